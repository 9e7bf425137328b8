"""Thermal expansion in the brick8 residual on the CPU oracle: its
right-hand side against the literal reference's (tests/golden/
make_thermoelastic.py: the reference's elastic step with given nodal
temperatures, stress = C (strain - T alpha)), the free expansion of a block on
statically determinate supports (u = T alpha x, zero stress, exactly), and
dR/dT against the linearity of the thermal load."""

import os

import numpy as np
import pytest

from helpers import GOLDEN
from oracle import design_oracle as DO
from thermoelastic_cases import CASES, case_model, free_block


def _nodal_rhs(o):
    b = np.zeros(o.n)
    for nd, vals in zip(o.model.load_nodes, o.model.load_vals):
        for q in range(6):
            d = o.nDof[nd - 1, q]
            if d and vals[q] != 0.0:
                b[d - 1] += vals[q]
    return b


@pytest.mark.parametrize("dims,seed", CASES)
def test_thermal_load_matches_the_reference(dims, seed):
    ref = dict(np.load(os.path.join(GOLDEN, "ref_thermoelastic_%dx%dx%d.npz" % dims), allow_pickle=False))
    m, T = case_model(dims, seed)
    np.testing.assert_array_equal(T, ref["temperatures"])
    o = DO.DesignOracle(m)
    np.testing.assert_array_equal(o.rank, ref["currentRank"])
    b = _nodal_rhs(o) + o.thermal_load(T)
    # the thermal load dominates the pressure here; both carry the same terms
    assert np.abs(b - ref["elasticLoad"]).max() <= 1e-13 * np.abs(ref["elasticLoad"]).max()


def test_free_expansion_is_exact():
    m = free_block()
    o = DO.DesignOracle(m)
    T0 = 80.0
    T = np.full(m.num_nodes, T0)
    A, b = o._system("elastic")
    import scipy.sparse.linalg as spl

    u = spl.spsolve(A, b + o.thermal_load(T))
    ue = np.array([[T0 * 1.5e-5 * x for x in m.nodes[i]] for i in range(m.num_nodes)])
    got = np.array([[u[o.nDof[i, c] - 1] for c in range(3)] for i in range(m.num_nodes)])
    assert np.abs(got - ue).max() <= 1e-9 * np.abs(ue).max()


def test_dRdT_is_minus_the_thermal_load_derivative():
    m, _ = case_model((3, 3, 3), 12)
    o = DO.DesignOracle(m)
    rng = np.random.default_rng(3)
    T1, T2, lam = rng.uniform(0, 100, m.num_nodes), rng.uniform(0, 100, m.num_nodes), rng.standard_normal(o.n)
    g = o.dRdT_product(lam)
    # R = K u - f_th(T) and f_th is linear in T
    lhs = lam @ (o.thermal_load(T2) - o.thermal_load(T1))
    assert abs(lhs + (T2 - T1) @ g) <= 1e-12 * abs(lhs)
