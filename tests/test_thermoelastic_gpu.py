"""Thermal expansion on the device: the right-hand side against the literal
reference's, the static solve and dR/dT against the CPU oracle, the free
expansion of a block on statically determinate supports (exact displacement,
zero stress), and a thermal + elastic job through the job host."""

import os
import tempfile

import numpy as np
import pytest

from helpers import GOLDEN, rel_norm
from misc_amd import astro
from misc_amd import job as J
from misc_amd.meshes import thermoelastic_block
from misc_amd.refformat import write_model
from oracle import design_oracle as DO
from thermoelastic_cases import CASES, case_model, free_block

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,seed", CASES)
def test_device_thermal_load_and_solve(dims, seed):
    ref = dict(np.load(os.path.join(GOLDEN, "ref_thermoelastic_%dx%dx%d.npz" % dims), allow_pickle=False))
    m, T = case_model(dims, seed)
    o = DO.DesignOracle(m)
    F = np.zeros(o.n)
    for nd, vals in zip(m.load_nodes, m.load_vals):
        for q in range(3):
            if o.nDof[nd - 1, q]:
                F[o.nDof[nd - 1, q] - 1] += vals[q]
    print("thermoelastic", dims, flush=True)
    ctx = astro.Context(m)
    ctx.setd("nodeTemp", T)
    ctx.elastic_soln_load()  # -R(0, T): the thermal load alone
    f = ctx.getd("elasticLoad")
    assert np.abs(f - (ref["elasticLoad"] - F)).max() <= 1e-13 * np.abs(ref["elasticLoad"]).max()
    ctx.static_solve()
    u = ctx.getd("nodeDisp")
    A, b = o._system("elastic")
    import scipy.sparse.linalg as spl

    ud = spl.spsolve(A, b + o.thermal_load(T))
    assert rel_norm(u, ud) <= 1e-10
    lam = np.random.default_rng(seed).standard_normal(o.n)
    np.testing.assert_allclose(ctx.dRdT_product(lam), o.dRdT_product(lam), rtol=1e-12,
                               atol=1e-12 * np.abs(o.dRdT_product(lam)).max())
    ctx.close()


def test_device_free_expansion():
    m = free_block()
    ctx = astro.Context(m, astro.Options(rtol=1e-12))
    T0 = 80.0
    ctx.setd("nodeTemp", np.full(m.num_nodes, T0))
    ctx.static_solve()
    u = ctx.getd("nodeDisp")
    nd = ctx.geti("nDofIndex").reshape(-1, 6)
    got = np.array([[u[nd[i, c] - 1] for c in range(3)] for i in range(m.num_nodes)])
    ue = T0 * 1.5e-5 * m.nodes
    assert np.abs(got - ue).max() <= 1e-9 * np.abs(ue).max()
    # the Green-Lagrange strain of u = a x is a + a^2 / 2 on the diagonal
    # (a = T alpha), so stress = C (strain - T alpha) is the stress of the
    # uniform strain a^2 / 2: E / (1 - 2 nu) a^2 / 2 on each normal, zero shear
    a = 1.5e-5 * T0
    s_exact = 70e9 / (1 - 2 * 0.3) * 0.5 * a * a
    for sp in ((0.0, 0.0, 0.0), (0.5, -0.5, 0.3)):
        st, sn = ctx.element_stress(sp)
        np.testing.assert_allclose(sn[:, :3], a + 0.5 * a * a, rtol=1e-9)
        assert np.abs(sn[:, 3:]).max() <= 1e-9 * a
        np.testing.assert_allclose(st[:, :3], s_exact, rtol=1e-6)
        assert np.abs(st[:, 3:]).max() <= 1e-6 * s_exact
    ctx.close()


SCRIPT = """*readModelInput:
    fileName: model.yaml
*solve:
    elastic: yes
    thermal: yes
*writeNodeResults:
    fileName: nodes.yaml
    fields:
        - displacement
        - temperature
"""


def test_thermoelastic_job():
    m = thermoelastic_block(3, 3, 2, alpha=1.2e-5, pressure=1e6)
    with tempfile.TemporaryDirectory() as wd:
        write_model(m, os.path.join(wd, "model.yaml"))
        with open(os.path.join(wd, "job.txt"), "w") as f:
            f.write(SCRIPT)
        print("job: thermoelastic 3x3x2", flush=True)
        job = J.Job(wd)
        try:
            job.run_script(os.path.join(wd, "job.txt"))
            T = job.state["temperature"]
            u = job.state["displacement"]
        finally:
            job.close()
        mj, info = J.read_model(os.path.join(wd, "model.yaml"))
    assert mj.therm_exp is not None and np.abs(mj.therm_exp - m.therm_exp).max() == 0.0
    o = DO.DesignOracle(J.model_at(mj, info, None, None, 0.0))
    Tr, ur = o.solve_thermoelastic()
    assert np.abs(T - Tr).max() <= 1e-9 * 100.0
    assert rel_norm(u, ur) <= 1e-9
