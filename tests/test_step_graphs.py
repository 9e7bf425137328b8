"""Newmark steps replayed from HIP graphs (Options.stepGraphs) against the
same steps launched eagerly.  A replayed step runs the eager step's kernels
with the GMRES cycle's rotations and least squares on the device, so the two
agree to the solver's tolerance; the run is split into chunks (each starts
with an eager step, and the host reads the state between chunks), and every
chunk's end state is checked, so every replay is checked through the state
it hands on.  A leading zero load factor makes the first step need no
Arnoldi steps, so the first graph's cap is too small for the next step: that
step misses its target in the replay and is finished eagerly, the path a
replay takes when the cap does not suffice."""

import numpy as np
import pytest

from helpers import rel_norm
from misc_amd import astro as A
from misc_amd.meshes import half_sine, impulse_block

pytestmark = pytest.mark.gpu

DT = 1e-5
# both paths stop at rtol 1e-12 on |P^-1 r|; their iterates differ by the
# rotations' rounding and, after a miss, by the restart structure
TOL = 1e-9


def run(model, scales, graphs, chunk, opts, net=None):
    ctx = A.Context(model, A.Options(stepGraphs=graphs, **opts))
    ctx.dynamic_init(0.25, 0.5, DT)
    if net is not None:
        ctx.gnn_attach(net)
    states = []
    for k0 in range(0, len(scales), chunk):
        ctx.dynamic_run(scales[k0:k0 + chunk])
        states.append({f: ctx.getd(f) for f in ("nodeDisp", "nodeVel", "nodeAcc")})
    stats = {s: ctx.scalar(s) for s in ("dynSteps", "dynGmresSteps", "dynGraphSteps", "dynEagerSteps", "dynGraphs")}
    ctx.close()
    return states, stats


def compare(a, b):
    worst = 0.0
    for i, (sa, sb) in enumerate(zip(a, b)):
        for f in sa:
            e = rel_norm(sb[f], sa[f])
            worst = max(worst, e)
            assert e <= TOL, (i, f, e)
    return worst


@pytest.mark.parametrize("dims,chunk", [((6, 5, 3), 6), ((12, 10, 6), 10)])
def test_graph_steps_match_eager_steps(dims, chunk):
    model = impulse_block(*dims, seed=3)
    scales = np.concatenate([[0.0], half_sine(3 * chunk - 1, DT, 2e-4)])
    opts = dict(rtol=1e-12, numVecs=60)
    print("step graphs", dims, flush=True)
    eager, se = run(model, scales, False, chunk, opts)
    graph, sg = run(model, scales, True, chunk, opts)
    worst = compare(eager, graph)
    print(f"eager {se} graph {sg} worst rel diff {worst:.3e}", flush=True)
    assert sg["dynSteps"] == len(scales) == se["dynSteps"]
    # one eager step opens each chunk; the rest replayed or finished after a miss
    assert sg["dynGraphSteps"] + sg["dynEagerSteps"] == len(scales) - len(eager)
    assert sg["dynEagerSteps"] >= 1 and sg["dynGraphs"] >= 2
    assert sg["dynGraphSteps"] >= len(scales) // 2


def test_graph_steps_with_learned_blocks_match_eager():
    """The two-stage preconditioner in the replayed cycle and the plain
    LDL^T judge (an untrained network: blocks close to block Jacobi)."""
    model = impulse_block(8, 8, 4, seed=5)
    scales = np.concatenate([[0.0], half_sine(15, DT, 2e-4)])
    opts = dict(rtol=1e-10, numVecs=50)
    net = A.Gnn(64, 3, seed=0)
    print("step graphs with learned blocks", flush=True)
    eager, se = run(model, scales, False, 8, opts, net)
    graph, sg = run(model, scales, True, 8, opts, net)
    worst = max(rel_norm(g[f], e[f]) for e, g in zip(eager, graph) for f in e)
    print(f"eager {se} graph {sg} worst rel diff {worst:.3e}", flush=True)
    # rtol 1e-10 here: the tolerance scales with it
    assert worst <= 1e-8
    assert sg["dynGraphSteps"] >= 8
    net.close()
