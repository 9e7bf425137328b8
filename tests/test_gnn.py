"""The learned block preconditioner on the device (graph network over the
node graph, hand-written HIP forward / backward / Adam kernels) against its
CPU oracle (oracle/gnn_oracle.py, pinned by finite differences on the CPU),
and what it gains: fewer GMRES(50) steps than the skyline LDL^T alone in
Newmark steps on a field it was not trained on, with the same displacements
(the plain LDL^T judges every solve and would continue it)."""

import numpy as np
import pytest

from helpers import csr_sparse, refined_solve, rel_norm
from misc_amd import astro as A
from misc_amd.meshes import half_sine, impulse_block
from oracle import gnn_oracle as GO

pytestmark = pytest.mark.gpu

DT = 1e-5


def dyn_ctx(seed, dims):
    m = impulse_block(*dims, seed=seed)
    ctx = A.Context(m, A.Options(numVecs=50, rtol=1e-8))
    ctx.dynamic_init(0.25, 0.5, DT)
    return m, ctx


def operator(ctx):
    """A = K + C^T C as the context holds it (after penalty scaling)."""
    n = ctx.scalar("elMatDim")
    K = csr_sparse(ctx.getd("elasticMat"), ctx.geti("elMatCols"), ctx.geti("elMatRange"), n)
    C = csr_sparse(ctx.getd("elMPCMat"), ctx.geti("elMPCMatCols"), ctx.geti("elMPCMatRange"), n)
    return K, C


def oracle_field(f, ctx):
    K, C = operator(ctx)
    nd = ctx.geti("nDofIndex").reshape(-1, 6)[:, :3] - 1
    return dict(range=f["range"], adj=f["adj"], X=f["X"].astype(np.float64), G=f["G"], A=(K + C.T @ C).tocsr(),
                node_dofs=nd, R1=f["R1"], eTr=f["eTr"])


def test_features_match_oracle():
    m, ctx = dyn_ctx(1, (6, 5, 3))
    print("gnn features 6x5x3", flush=True)
    N = m.num_nodes
    X = ctx.gnn_get("features")
    G = ctx.gnn_get("G").reshape(N, 3, 3)
    K, C = operator(ctx)
    nd = ctx.geti("nDofIndex").reshape(-1, 6)[:, :3] - 1
    constrained = np.zeros(N, bool)
    constrained[m.node_sets["bottom"] - 1] = True
    Xo, Go = GO.features(K, C, nd, constrained)
    cols = [0, 1, 2, 3, 4, 5, 7, 8]
    np.testing.assert_allclose(X[:, cols], Xo[:, cols], rtol=2e-6, atol=2e-7)
    mu = Xo[~constrained, 6].mean()
    lg = np.where(constrained, 0.0, np.clip(Xo[:, 6] - mu, -4, 4))
    np.testing.assert_allclose(X[:, 6], lg, rtol=0, atol=2e-5)
    np.testing.assert_allclose(G, Go, rtol=1e-10, atol=0)
    ctx.close()


@pytest.mark.parametrize("pairs", ["solution", "solve"])
def test_loss_gradient_and_blocks_match_oracle(pairs):
    gnn = A.Gnn(64, 3, seed=0)
    m, ctx = dyn_ctx(2, (6, 5, 3))
    print("gnn collect 6x5x3", pairs, flush=True)
    ctx.gnn_collect(gnn, 6, seed=1, rtol=1e-12, pairs=pairs)
    fo = oracle_field(gnn.field(0), ctx)
    ctx.close()
    # the pairs' error statistic: sum e1^T r1 with e1 = A^-1 r1 (exact for
    # solution pairs, to the GMRES tolerance for solved ones)
    eTr = sum(r @ refined_solve(fo["A"], r) for r in fo["R1"])
    assert abs(fo["eTr"] - eTr) <= 1e-8 * eTr
    rng = np.random.default_rng(3)
    p0 = gnn.get_params()
    for p in (p0, (p0 + 0.05 * rng.standard_normal(p0.size)).astype(np.float32)):
        gnn.set_params(p)
        l_dev, g_dev = gnn.loss_grad(0)
        l_o, g_o = GO.loss_grad(p.astype(np.float64), fo)
        print(f"loss device {l_dev:.8e} oracle {l_o:.8e}", flush=True)
        assert abs(l_dev - l_o) <= 1e-5 * abs(l_o)
        assert np.linalg.norm(g_dev - g_o) <= 1e-4 * np.linalg.norm(g_o)
        B_dev = gnn.field_blocks(0)
        B_o, _ = GO.forward(p.astype(np.float64), fo["X"], fo["range"], fo["adj"], fo["G"])
        scale = np.linalg.norm(B_o, axis=(1, 2))
        assert np.max(np.linalg.norm(B_dev - B_o, axis=(1, 2)) / scale) <= 1e-5
    gnn.close()


def test_training_lowers_the_loss_and_the_gmres_steps():
    dims = (12, 12, 6)
    gnn = A.Gnn(64, 3, seed=0)
    for seed in range(6):
        print("collect field", seed, flush=True)
        _, ctx = dyn_ctx(seed, dims)
        ctx.gnn_collect(gnn, 8, seed=100 + seed, rtol=1e-10)
        ctx.close()
    assert gnn.num_fields == 6
    scales = half_sine(20, DT, 5e-4)
    m, ctx = dyn_ctx(77, dims)  # a field the network is not trained on

    def newmark_steps(net):
        ctx.gnn_detach()
        ctx.dynamic_init(0.25, 0.5, DT)
        if net is not None:
            ctx.gnn_attach(net)
        ctx.dynamic_run(scales)
        return ctx.scalar("dynGmresSteps"), ctx.getd("nodeDisp")

    steps_plain, u_plain = newmark_steps(None)
    steps_init, _ = newmark_steps(gnn)  # untrained: close to block Jacobi
    hist = gnn.train(60, lr=2e-3, seed=0)
    print("loss per epoch", hist[0], hist[len(hist) // 2], hist[-1], flush=True)
    assert hist[-1] < hist[0]
    steps_gnn, u_gnn = newmark_steps(gnn)
    print(f"GMRES steps over 20 Newmark steps: plain LDL {steps_plain}, untrained network {steps_init}, "
          f"trained network {steps_gnn} (judge/fallback steps {ctx.scalar('gnnFallbackSteps')})", flush=True)
    assert steps_gnn < steps_plain
    assert steps_gnn <= steps_init
    assert rel_norm(u_gnn, u_plain) <= 1e-6
    ctx.close()
    gnn.close()
