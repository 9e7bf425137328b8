#!/usr/bin/env python3
"""Assembled+solved DOF/s of the linear-static brick8 hot path on one MI355X.

Workload (one "step"): the 1.0 x 1.0 x 0.5 m block of 64 x 64 x 32 brick8
elements (139,425 nodes, 418,275 DOF), E = 200 GPa, nu = 0.3, bottom face
clamped by penalty constraints, 1 MPa on the top face as consistent nodal
loads, solved the way the reference's solve() does it: element assembly into
the CSR, penalty scaling, profile extraction, skyline LDL^T of the profile,
then left-preconditioned GMRES(100) to |P^-1 r| <= 1e-12 |P^-1 r0|.  All FP64
and resident in HBM; the mesh upload and the symbolic analysis (node ordering,
DOF map, sparsity, profile) happen once before the timed region, as they are
per mesh, not per solve.

solverMaxBW: "int64" (default) evaluates the reference's rule
b^3 >= 27 N^2 without overflow (b = 8067, profile 1.25e9 entries, ~10 GB);
"literal" reproduces the reference's 32-bit wraparound (b = 957, profile
3.2e7 entries, ~30x more GMRES steps).

The JSON line carries the phase times, the roofline of the dominant kernel
(the triangular sweeps of the preconditioner, timed with HIP events on the
context stream over the timed region) and a CPU baseline: the literal
reference (its own Fortran compiled from source, oracle/_ref) run on a
bounded sample, a 16 x 16 x 8 block of the same problem, on one host core.
"""

import argparse
import json

import numpy as np
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3


def factor_flops(lt_range):
    """Exact LDL^T flop count of the profile: sum over rows i and stored
    columns j < i of 2 (j - max(s_i, s_j)), s the first stored column.
    When s never decreases down the rows (true for the level-sorted meshes
    under the 64-bit rule) the inner sum is L (L - 1) / 2 with L = i - s_i;
    otherwise the pairs are summed directly (profiles up to ~1e8 entries)."""
    import numpy as np
    R = np.asarray(lt_range, dtype=np.int64)
    rowlen = np.diff(R)
    s = np.arange(len(rowlen)) - rowlen + 1
    if len(s) <= 1 or not np.any(np.diff(s) < 0):
        L = (rowlen - 1).astype(np.float64)
        return float(np.sum(L * (L - 1.0)))
    if R[-1] > 2e8:
        return None
    total = 0.0
    n = len(rowlen)
    for a in range(0, n, 65536):  # rows in chunks, to bound memory
        rows = np.arange(a, min(a + 65536, n))
        cnt = rowlen[rows] - 1
        ri = np.repeat(rows, cnt)
        off = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        j = s[ri] + off
        total += 2.0 * float(np.sum(j - np.maximum(s[ri], s[j])))
    return total


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(sample=(16, 16, 8)):
    """Time the reference pipeline (assemble, scale, profile, factor, GMRES)
    on a sample block with one host core; returns the cpu_baseline object
    with the phase times and the sample's profile statistics (for the
    extrapolation in extrapolate_cpu)."""
    from misc_amd.meshes import cantilever_block
    from oracle import oracle as O

    model = cantilever_block(*sample)
    if O.ref_available():
        r = O.run_reference(model, "elastic")
        t = r["times"]
        phases = {k: t[k] for k in ("assemble", "scale", "profile", "factor", "solve")}
        n = int(r["scalars"][1])
        lt_range = r["elMatLTRange"]
        nnz = int(r["elMatRange"][-1])
        kind = "reference"
        what = "literal reference (flang -O2, its own GMRES stopping rule)"
    else:
        om = O.OracleModel(model).prep()
        t0 = time.perf_counter()
        om.static_elastic(mode=0)
        phases = {"total": time.perf_counter() - t0}
        n = om.scalar("elMatDim")
        lt_range = om.geti("elMatLTRange")
        nnz = int(om.geti("elMatRange")[-1])
        kind = "port"
        what = "C++ restatement of the reference algorithm (-O2)"
    total = sum(phases.values())
    return {"value": n / total, "unit": "DOF/s", "cores": 1, "kind": kind, "cpu_model": cpu_model(),
            "sample": f"brick8 block {sample[0]}x{sample[1]}x{sample[2]} ({n} DOF), same physics and "
                      f"solve sequence, {what}, {total:.1f} s",
            "phases_s": phases, "sample_dims": list(sample), "sample_dof": n, "sample_nnz": nnz,
            "sample_profile_entries": int(lt_range[-1]), "sample_factor_flops": factor_flops(lt_range)}


def extrapolate_cpu(base, steps_sample, target):
    """Flop-model extrapolation of the literal reference's time from its
    sample to the target workload (same rule, literal solverMaxBW): the factor
    by the exact profile flop ratio, GMRES by the work per step (two sweeps
    over the profile and one CSR product: 4 prof + 2 nnz) times the ratio of
    GMRES steps (both counted by the GPU solver under the same rule, as a
    proxy for the reference's own counts), the rest by the DOF ratio."""
    ph = base["phases_s"]
    if "factor" not in ph or not base["sample_factor_flops"] or not target["factor_flops"]:
        return None
    w_s = 4.0 * base["sample_profile_entries"] + 2.0 * base["sample_nnz"]
    w_t = 4.0 * target["profile_entries"] + 2.0 * target["nnz"]
    t_fact = ph["factor"] * target["factor_flops"] / base["sample_factor_flops"]
    t_solve = ph["solve"] * (w_t / w_s) * (target["gmres_steps"] / max(1.0, steps_sample))
    t_rest = (ph["assemble"] + ph["scale"] + ph["profile"]) * target["dof"] / base["sample_dof"]
    total = t_fact + t_solve + t_rest
    return {"seconds": total, "dof_per_s": target["dof"] / total,
            "factor_s": t_fact, "gmres_s": t_solve, "rest_s": t_rest,
            "model": "factor by exact profile flops; GMRES by (4 profile + 2 nnz) per step x steps "
                     "(the GPU solver's counts under the literal rule at both sizes); the rest by DOF; "
                     "one core, literal solverMaxBW rule"}


def load_traffic():
    """HBM bytes per sweep launch measured by rocprofv3 PMC runs (FETCH_SIZE and
    WRITE_SIZE, each calibrated against a copy kernel of known size), as
    recorded in profiles/k_sweep_traffic_brick8_64x64x32.md; None if absent."""
    f = os.path.join(ROOT, "profiles", "k_sweep_traffic_brick8_64x64x32.md")
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split("=")[1])
    return None


def bench_newmark(args):
    """One step = one Newmark run from rest of the impulse workload: set-up
    (tangent K + M/(beta dt^2), penalty scaling, profile, factor) and 200 time
    steps of dt = 1e-5 s, each a residual assembly plus GMRES(50) to 1e-8
    (relative, preconditioned) from the previous increment, replayed as one
    HIP graph per time step (the first step of a run is eager).  The value is
    DOF x time steps per second of the whole run; the same run with eager
    launches is timed once beside it."""
    from misc_amd import astro
    from misc_amd.meshes import half_sine, impulse_block

    nx, ny, nz = 48, 48, 24
    n_steps, dt, pulse = 200, 1e-5, 5e-4
    model = impulse_block(nx, ny, nz)
    scales = half_sine(n_steps, dt, pulse)
    print(f"bench: Newmark brick8 {nx}x{ny}x{nz}, {n_steps} steps", file=sys.stderr, flush=True)
    ctx = astro.Context(model, astro.Options(literalMaxBW=(args.maxbw == "literal"), numVecs=50, rtol=1e-8,
                                             stepGraphs=not args.eager_steps))
    n = ctx.scalar("elMatDim")
    for w in range(args.warmup):
        ctx.dynamic_init(0.25, 0.5, dt)
        ctx.dynamic_run(scales)
        print(f"bench: warmup {w} done", file=sys.stderr, flush=True)
    ctx.sync()
    init_s = run_s = 0.0
    gm_steps = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        a = time.perf_counter()
        ctx.dynamic_init(0.25, 0.5, dt)
        ctx.sync()
        b = time.perf_counter()
        ctx.dynamic_run(scales)
        ctx.sync()
        init_s += b - a
        run_s += time.perf_counter() - b
        gm_steps += ctx.scalar("dynGmresSteps")
        print(f"bench: step {k} done", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    u = ctx.getd("nodeDisp")
    graph_stats = {k: ctx.scalar(k) for k in ("dynGraphSteps", "dynEagerSteps", "dynGraphs", "dynGraphCap")}
    ctx.close()
    # the same run with one eager launch sequence per step, for comparison
    ctx = astro.Context(model, astro.Options(literalMaxBW=(args.maxbw == "literal"), numVecs=50, rtol=1e-8))
    ctx.dynamic_init(0.25, 0.5, dt)
    ctx.sync()
    a = time.perf_counter()
    ctx.dynamic_run(scales)
    ctx.sync()
    eager_s = time.perf_counter() - a
    u_eager = ctx.getd("nodeDisp")
    ctx.close()
    print("bench: eager run done", file=sys.stderr, flush=True)
    out = {
        "metric": "Newmark DOF x time steps per second (plain skyline-LDL-preconditioned GMRES)",
        "value": n * n_steps * args.steps / elapsed, "unit": "DOF*steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"brick8 block {nx}x{ny}x{nz} (1.0 x 1.0 x 0.5 m), log-normal E (median 70 GPa, "
                               f"log-sigma 0.5, correlation 8 elements, seed 0), nu 0.3, rho 2700, bottom clamped, "
                               f"1 kN half-sine ({pulse} s) on an 8x8-element top patch; Newmark beta 0.25 "
                               f"gamma 0.5, dt {dt}, {n_steps} steps; GMRES(50) to 1e-8",
                   "dof": n, "time_steps": n_steps, "profile_entries": ctx_profile_entries(model, args),
                   "gmres_steps_per_time_step": gm_steps / (args.steps * n_steps),
                   "max_abs_displacement_m": float(np.abs(u).max())},
        "phases_ms_per_step": {"setup_assemble_factor": 1e3 * init_s / args.steps,
                               "time_loop": 1e3 * run_s / args.steps},
        "step_graphs": dict(graph_stats, eager_time_loop_ms=1e3 * eager_s,
                            graph_vs_eager_time_loop=eager_s / (run_s / args.steps),
                            rel_diff_vs_eager=float(np.linalg.norm(u - u_eager) / np.linalg.norm(u_eager))),
    }
    print(json.dumps(out))


def bench_learned(args):
    """The learned preconditioner in the Newmark workload.  Training data:
    64 fields of the same block (log-normal E, seeds 1000..1063), each
    factored at the Newmark tangent, with 31 pairs (x standard normal,
    b = A x) and the residual the LDL^T stage leaves for each (1984
    solver-generated pairs); the graph network (3
    message-passing layers, 64 hidden, FP32) is trained on the GPU by the
    library's own forward/backward/Adam kernels.  One step = one Newmark run
    of 200 time steps on the field of seed 0 (not trained on) with the
    two-stage preconditioner (LDL^T, then the learned node blocks), plain
    LDL^T as judge and fallback, GMRES(50) to 1e-8, HIP-graph time steps;
    the same run with plain LDL^T is timed beside it.  Set-up (assembly,
    factor, features and blocks) is outside the timed region for both, as
    in the newmark workload's time loop."""
    from misc_amd import astro
    from misc_amd.meshes import half_sine, impulse_block

    nx, ny, nz = 48, 48, 24
    n_steps, dt, pulse = 200, 1e-5, 5e-4
    n_fields, n_rhs, epochs = args.fields, 31, args.epochs
    opts = astro.Options(literalMaxBW=(args.maxbw == "literal"), numVecs=50, rtol=1e-8,
                         stepGraphs=not args.eager_steps)
    gnn = astro.Gnn(64, 3, seed=0)
    t0 = time.perf_counter()
    for f in range(n_fields):
        m = impulse_block(nx, ny, nz, seed=1000 + f)
        c = astro.Context(m, opts)
        c.dynamic_init(0.25, 0.5, dt)
        c.gnn_collect(gnn, n_rhs, seed=5000 + f, rtol=1e-10)
        c.close()
        if f % 8 == 7:
            print(f"bench: collected {f + 1} fields", file=sys.stderr, flush=True)
    collect_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    hist = gnn.train(epochs, lr=2e-3, seed=0)
    train_s = time.perf_counter() - t0
    print(f"bench: trained {epochs} epochs, loss {hist[0]:.4f} -> {hist[-1]:.4f}", file=sys.stderr, flush=True)
    model = impulse_block(nx, ny, nz)
    scales = half_sine(n_steps, dt, pulse)
    ctx = astro.Context(model, opts)
    n = ctx.scalar("elMatDim")

    def newmark(net):
        ctx.gnn_detach()
        ctx.dynamic_init(0.25, 0.5, dt)
        if net is not None:
            ctx.gnn_attach(net)
        ctx.sync()
        a = time.perf_counter()
        ctx.dynamic_run(scales)
        ctx.sync()
        return time.perf_counter() - a, ctx.scalar("dynGmresSteps"), ctx.getd("nodeDisp")

    for w in range(args.warmup):
        newmark(None)
        newmark(gnn)
    t_plain = t_learned = 0.0
    s_plain = s_learned = 0
    for k in range(args.steps):
        tp, sp, u_plain = newmark(None)
        tl, sl, u_learned = newmark(gnn)
        t_plain += tp
        t_learned += tl
        s_plain += sp
        s_learned += sl
        print(f"bench: step {k} done", file=sys.stderr, flush=True)
    fallback = ctx.scalar("gnnFallbackSteps")
    ctx.close()
    gnn.close()
    out = {
        "metric": "Newmark DOF x time steps per second, learned two-stage vs plain skyline-LDL preconditioner",
        "value": n * n_steps * args.steps / t_learned, "unit": "DOF*steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t_learned / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 solve, f32 network", "data": "synthetic",
        "config": {"workload": f"brick8 block {nx}x{ny}x{nz} (1.0 x 1.0 x 0.5 m), log-normal E (median 70 GPa, "
                               f"log-sigma 0.5, correlation 8 elements), nu 0.3, rho 2700, bottom clamped, 1 kN "
                               f"half-sine ({pulse} s) on an 8x8-element top patch; Newmark beta 0.25 gamma 0.5, "
                               f"dt {dt}, {n_steps} steps; GMRES(50) to 1e-8; network 3 layers x 64 hidden FP32 "
                               f"trained on {n_fields * n_rhs} solves over {n_fields} seeded fields",
                   "dof": n, "time_steps": n_steps, "solverMaxBW_rule": args.maxbw},
        "learned": {"speedup_vs_plain": t_plain / t_learned,
                    "plain_value": n * n_steps * args.steps / t_plain,
                    "plain_ms_per_step": 1e3 * t_plain / args.steps,
                    "gmres_steps_per_time_step_plain": s_plain / (args.steps * n_steps),
                    "gmres_steps_per_time_step_learned": s_learned / (args.steps * n_steps),
                    "judge_fallback_steps_last_run": fallback,
                    "rel_diff_vs_plain": float(np.linalg.norm(u_learned - u_plain) / np.linalg.norm(u_plain)),
                    "collect_s": collect_s, "train_s": train_s, "epochs": epochs,
                    "loss_first_last": [float(hist[0]), float(hist[-1])]},
    }
    print(json.dumps(out))


def bench_sensitivity(args):
    """One step = the tip-compliance gradient of the cantilever shell plate
    (128 x 16 shell4, 13,158 DOF, thickness 2 mm x (1 + U(-0.1, 0.1)), 1 N on
    each tip node) with respect to its 64 two-column thickness strips by
    complex step (h = 1e-30): every variable's perturbed system assembled,
    factored (complex-symmetric LDL^T) and solved in complex128 in one
    persistent launch, then one double-double refinement step.  The value
    is design variables x DOF per second of that evaluation (tile shape
    from the library's autotuner).  Beside it: the adjoint form (one real
    static solve, complex-step element derivatives, u . dK/dd u) and, on
    the host, the oracle's per-variable loop of the gradient stage."""
    from misc_amd import astro
    from misc_amd.meshes import strip_plate

    model, dv = strip_plate(128, 16)
    ctx = astro.Context(model, astro.Options(literalMaxBW=(args.maxbw == "literal"), coupledProfile=True,
                                             rtol=1e-12, maxCycles=100000))
    n = ctx.scalar("elMatDim")
    ctx.set_design(dv, h=1e-30)
    n_dv = int(dv.max()) + 1
    print(f"bench: shell plate 128x16 ({n} DOF), {n_dv} design variables", file=sys.stderr, flush=True)
    tiles = ctx.cs_batch_autotune()
    for w in range(args.warmup):
        ctx.cs_batch_gradient()
    ctx.sync()
    t0 = time.perf_counter()
    kern_ms = 0.0
    for k in range(args.steps):
        g, comp = ctx.cs_batch_gradient()
        kern_ms += ctx.scalar_f("csBatchMs")
        print(f"bench: step {k} done", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    flops = ctx.scalar_f("csBatchFlops")
    # the adjoint form on the same context
    ctx.reset()
    a = time.perf_counter()
    ctx.static_solve()
    ctx.sync()
    b = time.perf_counter()
    g_adj = ctx.compliance_gradient()
    adj_grad_s = time.perf_counter() - b
    solve_s = b - a
    u = ctx.getd("nodeDisp")
    F = np.zeros(n)
    nd = ctx.geti("nDofIndex").reshape(-1, 6)
    for node, v in zip(model.load_nodes, model.load_vals):
        for c in range(6):
            if nd[node - 1, c] and v[c]:
                F[nd[node - 1, c] - 1] += v[c]
    out = {
        "metric": "complex-step design sensitivities: design variables x DOF per second (batched complex128 "
                  "assemble + factor + solve of every variable)",
        "value": n_dv * n * args.steps / elapsed, "unit": "DV*DOF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex128", "data": "synthetic",
        "config": {"workload": "shell4 cantilever plate 128x16 (1.0 x 0.1 m), E 70 GPa, nu 0.33, thickness "
                               "2 mm x (1 + U(-0.1, 0.1)) seed 0, clamped root, 1 N +z per tip node; 64 "
                               "two-column thickness strips; tip compliance; complex step 1e-30",
                   "dof": n, "design_variables": n_dv,
                   "instance_profile_entries": int(ctx.scalar_f("csBatchEntries")),
                   "instance_max_row": int(ctx.scalar_f("csBatchMaxRow")),
                   "compliance": float(F @ u), "gradient_l2": float(np.linalg.norm(g))},
        "roofline": {"bound": "latency (serial row-block chain per instance; one workgroup per instance)",
                     "achieved": flops / (kern_ms / args.steps * 1e-3) / 1e12,
                     "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": flops / (kern_ms / args.steps * 1e-3) / 1e12 / FP64_MFMA_PEAK_TFLOPS,
                     "traffic": None, "flop_model": flops},
        "batched": {"launch_ms": kern_ms / args.steps, "tiles_ms": tiles,
                    "max_rel_diff_vs_adjoint": float(np.max(np.abs(g - g_adj)) / np.max(np.abs(g_adj))),
                    "compliance_rel_diff": float(np.max(np.abs(comp - F @ u)) / abs(F @ u))},
        "adjoint": {"static_solve_ms": 1e3 * solve_s, "gradient_ms": 1e3 * adj_grad_s,
                    "gmres_steps": ctx.scalar("gmresSteps")},
    }
    ctx.close()
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        om = O.OracleModel(model).prep()
        om.static_elastic(stop_after="assemble")
        om.setd("nodeDisp", u)
        om.set_design(dv)
        t0 = time.perf_counter()
        g_cpu = om.compliance_gradient(h=1e-30)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n_dv * n / cpu_s, "unit": "DV*DOF/s", "cores": 1, "kind": "port",
                               "sample": f"the oracle's per-variable loop (complex-step element residuals of each "
                                         f"variable's elements, then the adjoint dot) on the whole plate given u, "
                                         f"{cpu_s:.2f} s; the one real solve it needs is not included",
                               "max_rel_diff_vs_batched": float(np.max(np.abs(g_cpu - g)) /
                                                                np.max(np.abs(g_cpu)))}
    print(json.dumps(out))


def timed_static(ctx, steps, warmup):
    """warmup + steps static solves from rest on ctx; returns (seconds per
    solve, phase ms per solve, GMRES steps per solve, sweep ms, sweep count)."""
    for _ in range(warmup):
        ctx.reset()
        ctx.static_solve()
    ctx.sync()
    ctx.set_timing(True)
    phases = {k: 0.0 for k in ("assemble", "scale", "profile", "factor", "solve")}
    gsteps = 0
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.reset()
        ctx.static_solve()
        for ph in phases:
            phases[ph] += ctx.phase_ms(ph)
        gsteps += ctx.scalar("gmresSteps")
        print(f"bench: step {k} done", file=sys.stderr, flush=True)
    ctx.sync()
    el = time.perf_counter() - t0
    sw_ms = ctx.kernel_stat("sweep_fwd_ms") + ctx.kernel_stat("sweep_bwd_ms")
    sw_n = ctx.kernel_stat("sweep_fwd_count") + ctx.kernel_stat("sweep_bwd_count")
    ctx.set_timing(False)
    return el / steps, {k: v / steps for k, v in phases.items()}, gsteps / steps, sw_ms, sw_n


def bench_precision(args):
    """Config 3, the precision form on the 64 x 64 x 32 brick block with E per
    element drawn uniform in [150, 250] GPa (seed 3): one step = one static
    solve from rest, timed for the plain FP64 path and for the mixed path
    (FP32-stored profile and factor, FP32 matrix-core trailing updates, FP64
    pivots / diagonal blocks / sweeps / operator, GMRES restarts forming the
    true residual in double-double).  The value is the mixed path's DOF/s;
    the line reports the speed-up over the plain path, the error of every
    mixed result against the plain FP64 result, and the steady thermal case
    (k = 50 W/m K, 100 K on top, 0 K at the bottom) on the same block in both
    precisions."""
    from misc_amd import astro
    from misc_amd.meshes import cantilever_block, thermal_block, uniform_e_field

    nx, ny, nz = args.nx, args.ny, args.nz
    model = cantilever_block(nx, ny, nz, e_field=uniform_e_field(nx * ny * nz, 150e9, 250e9, 3))
    lit = args.maxbw == "literal"
    print(f"bench: precision form, brick8 {nx}x{ny}x{nz}, E ~ U[150, 250] GPa", file=sys.stderr, flush=True)
    res = {}
    for name, prec in (("plain", 0), ("mixed", 1)):
        ctx = astro.Context(model, astro.Options(literalMaxBW=lit, precision=prec))
        n, prof = ctx.scalar("elMatDim"), ctx.scalar("elMatLTSize")
        sec, ph, gs, sw_ms, sw_n = timed_static(ctx, args.steps, args.warmup)
        u = ctx.getd("nodeDisp")
        bottom = model.node_sets["bottom"]
        react = ctx.reaction_force(bottom)
        res[name] = dict(sec=sec, phases=ph, gmres=gs, sweep_ms=sw_ms / max(1.0, sw_n), u=u, react=react,
                         rel_res=ctx.true_residual(), flops=factor_flops(ctx.geti("elMatLTRange")))
        ctx.close()
        print(f"bench: {name} {1e3 * sec:.1f} ms per solve, GMRES steps {gs}", file=sys.stderr, flush=True)
    th = {}
    tm = thermal_block(nx, ny, nz)
    for name, prec in (("plain", 0), ("mixed", 1)):
        ctx = astro.Context(tm, astro.Options(literalMaxBW=lit, thermLiteralD8=False, precision=prec))
        ctx.assemble_thermal()
        ctx.thermal_solve()  # warm-up
        ctx.reset()
        ctx.sync()
        t0 = time.perf_counter()
        ctx.assemble_thermal()
        ctx.thermal_solve()
        ctx.sync()
        th[name] = dict(sec=time.perf_counter() - t0, T=ctx.getd("nodeTemp"), rank=ctx.geti("currentRank"),
                        steps=ctx.scalar("thermGmresSteps"))
        ctx.close()
    z = tm.nodes[np.argsort(th["plain"]["rank"]), 2]
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    p, mx = res["plain"], res["mixed"]
    sweep_bytes = 4.0 * prof + 16.0 * n
    fac_s = mx["phases"]["factor"] * 1e-3
    out = {
        "metric": "assembled+solved DOF/s at <=1e-10 rel. err vs FP64; mixed vs plain speedup",
        "value": n / mx["sec"], "unit": "DOF/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * mx["sec"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 factor storage and trailing updates, f64 pivots / sweeps / operator / residuals",
        "data": "synthetic",
        "config": {"workload": f"brick8 block {nx}x{ny}x{nz} linear static, E per element U[150, 250] GPa (seed 3), "
                               f"nu 0.3, bottom clamped, 1 MPa top; skyline LDL^T-preconditioned GMRES(100) to "
                               f"1e-12 (preconditioned, relative), solverMaxBW rule {args.maxbw}; plain FP64 vs "
                               f"mixed (precision = 1)",
                   "dof": n, "profile_entries": prof,
                   "plain_ms_per_solve": 1e3 * p["sec"], "mixed_ms_per_solve": 1e3 * mx["sec"],
                   "speedup_mixed_vs_plain": p["sec"] / mx["sec"],
                   "gmres_steps_plain": p["gmres"], "gmres_steps_mixed": mx["gmres"],
                   "phases_ms_plain": p["phases"], "phases_ms_mixed": mx["phases"],
                   "err_vs_fp64": {"nodeDisp": rel(mx["u"], p["u"]),
                                   "reactionForce_bottom": rel(mx["react"], p["react"]),
                                   "true_rel_residual_plain": p["rel_res"], "true_rel_residual_mixed": mx["rel_res"]},
                   "thermal": {"dof": tm.num_nodes, "plain_ms": 1e3 * th["plain"]["sec"],
                               "mixed_ms": 1e3 * th["mixed"]["sec"],
                               "speedup_mixed_vs_plain": th["plain"]["sec"] / th["mixed"]["sec"],
                               "gmres_steps_plain": th["plain"]["steps"], "gmres_steps_mixed": th["mixed"]["steps"],
                               "err_vs_fp64_nodeTemp": rel(th["mixed"]["T"], th["plain"]["T"]),
                               "max_abs_T_minus_200z_K": float(np.abs(th["mixed"]["T"] - 200.0 * z).max())}},
        "roofline": {"kernel": "k_sweep_forward<float> / k_sweep_backward<float> (FP32-stored preconditioner "
                               "sweeps)", "bound": "hbm", "achieved": sweep_bytes / (mx["sweep_ms"] * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": sweep_bytes / (mx["sweep_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "avg_launch_ms": mx["sweep_ms"], "algorithmic_bytes_per_launch": sweep_bytes},
        "cpu_baseline": None,
    }
    if mx["flops"] is not None:
        out["roofline_factor"] = {"kernel": "k_ldl_lead<float> + k_ldl_rows<float> + k_sweep_blocks<float> "
                                            "(whole factor phase, FP32 matrix cores)", "bound": "mfma",
                                  "achieved": mx["flops"] / fac_s / 1e12, "peak": FP32_MFMA_PEAK_TFLOPS,
                                  "unit": "TFLOP/s", "frac": mx["flops"] / fac_s / 1e12 / FP32_MFMA_PEAK_TFLOPS,
                                  "flops_per_factor": mx["flops"]}
    print(json.dumps(out))


def ctx_profile_entries(model, args):
    from misc_amd import astro
    p = astro.Prep(model, astro.Options(literalMaxBW=(args.maxbw == "literal")))
    return int(p.geti("elMatLTRange")[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--nz", type=int, default=32)
    ap.add_argument("--maxbw", choices=["literal", "int64"], default="int64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-literal", action="store_true",
                    help="static: skip the second solve under the literal solverMaxBW rule (for profiling)")
    ap.add_argument("--workload", choices=["static", "precision", "newmark", "learned", "sensitivity"],
                    default="static",
                    help="static: the 64x64x32 linear-static block (default); newmark: the 48x48x24 "
                         "log-normal-E block, 200 Newmark steps of a 1 kN half-sine top-patch pulse; "
                         "sensitivity: 128x16 shell plate, compliance gradient of 64 thickness strips by "
                         "complex step; precision: the 64x64x32 block with E ~ U[150, 250] GPa in FP64 and "
                         "in the mixed form, plus the thermal case; learned: the newmark block with the "
                         "graph-network preconditioner trained on solver-generated pairs vs plain LDL^T")
    ap.add_argument("--fields", type=int, default=64, help="learned: training fields")
    ap.add_argument("--epochs", type=int, default=40, help="learned: training epochs")
    ap.add_argument("--eager-steps", action="store_true",
                    help="newmark / learned: launch every time step eagerly instead of replaying HIP graphs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported. "
              f"(The three displacement-component blocks of the preconditioner would shard across GPUs; "
              f"DESIGN.md describes the multi-GPU layout.)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "newmark":
        return bench_newmark(args)
    if args.workload == "sensitivity":
        return bench_sensitivity(args)
    if args.workload == "precision":
        return bench_precision(args)
    if args.workload == "learned":
        return bench_learned(args)

    from misc_amd import astro
    from misc_amd.meshes import cantilever_block

    model = cantilever_block(args.nx, args.ny, args.nz)
    print(f"bench: brick8 {args.nx}x{args.ny}x{args.nz}, solverMaxBW rule {args.maxbw}", file=sys.stderr,
          flush=True)
    ctx = astro.Context(model, astro.Options(literalMaxBW=(args.maxbw == "literal")))
    n = ctx.scalar("elMatDim")
    prof = ctx.scalar("elMatLTSize")
    nnz = ctx.scalar("elMatSize")
    for w in range(args.warmup):
        ctx.reset()
        ctx.static_solve()
        print(f"bench: warmup {w} done, GMRES steps {ctx.scalar('gmresSteps')}", file=sys.stderr, flush=True)
    ctx.sync()
    ctx.set_timing(True)
    phases = {k: 0.0 for k in ("assemble", "scale", "profile", "factor", "solve")}
    steps_total = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.reset()
        ctx.static_solve()
        for p in phases:
            phases[p] += ctx.phase_ms(p)
        steps_total += ctx.scalar("gmresSteps")
        print(f"bench: step {k} done", file=sys.stderr, flush=True)
    ctx.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    fwd_ms, bwd_ms = ctx.kernel_stat("sweep_fwd_ms"), ctx.kernel_stat("sweep_bwd_ms")
    fwd_n, bwd_n = ctx.kernel_stat("sweep_fwd_count"), ctx.kernel_stat("sweep_bwd_count")
    ctx.set_timing(False)
    rel_res = ctx.true_residual()
    flops = factor_flops(ctx.geti("elMatLTRange"))
    value = n * args.steps / elapsed
    # one sweep reads the strict lower profile (or, backward, all of it) and
    # reads b / writes x: 8 bytes per profile entry + 16 per row
    sweep_bytes = 8.0 * prof + 16.0 * n
    avg_sweep_ms = (fwd_ms + bwd_ms) / max(1.0, fwd_n + bwd_n)
    achieved = sweep_bytes / (avg_sweep_ms * 1e-3) / 1e9
    out = {
        "metric": "assembled+solved DOF/s at <=1e-10 rel. err vs FP64",
        "value": value,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"brick8 block {args.nx}x{args.ny}x{args.nz} linear static (E 200 GPa, nu 0.3, "
                               f"bottom clamped, 1 MPa top), skyline LDL^T-preconditioned GMRES(100), "
                               f"solverMaxBW rule {args.maxbw}",
                   "dof": n, "csr_nnz": nnz, "profile_entries": prof, "solverMaxBW": ctx.scalar("solverMaxBW"),
                   "gmres_steps_per_solve": steps_total / args.steps, "rtol_preconditioned": 1e-12,
                   "true_rel_residual": rel_res},
        "phases_ms_per_step": {k: v / args.steps for k, v in phases.items()},
        "roofline": {"kernel": "k_sweep_forward / k_sweep_backward (preconditioner triangular sweeps)",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic() if (args.nx, args.ny, args.nz, args.maxbw) == (64, 64, 32, "int64")
                     else None,
                     "avg_launch_ms": avg_sweep_ms,
                     "launches": int(fwd_n + bwd_n), "algorithmic_bytes_per_launch": sweep_bytes},
        "cpu_baseline": None,
    }
    if flops is not None:
        fac_s = phases["factor"] / args.steps * 1e-3
        out["roofline_factor"] = {"kernel": "k_ldl_lead + k_ldl_rows + k_sweep_blocks (whole factor phase)",
                                  "bound": "mfma", "achieved": flops / fac_s / 1e12, "peak": FP64_MFMA_PEAK_TFLOPS,
                                  "unit": "TFLOP/s", "frac": flops / fac_s / 1e12 / FP64_MFMA_PEAK_TFLOPS,
                                  "flops_per_factor": flops}
    # an error estimate beside the residual: a second solve from the
    # converged state (a restart of the refinement) moves u by about its error
    u1 = ctx.getd("nodeDisp")
    ctx.static_solve()
    u2 = ctx.getd("nodeDisp")
    out["config"]["error_estimate_rel"] = float(np.linalg.norm(u2 - u1) / np.linalg.norm(u2))
    ctx.close()
    # the same workload under the reference's literal (32-bit) solverMaxBW rule
    lit = None
    if args.maxbw == "int64" and not args.no_literal:
        lctx = astro.Context(model, astro.Options(literalMaxBW=True))
        lctx.static_solve()
        lctx.reset()
        a = time.perf_counter()
        lctx.static_solve()
        lctx.sync()
        lit_s = time.perf_counter() - a
        lit = {"value": n / lit_s, "unit": "DOF/s", "ms_per_step": 1e3 * lit_s,
               "solverMaxBW": lctx.scalar("solverMaxBW"), "profile_entries": lctx.scalar("elMatLTSize"),
               "gmres_steps": lctx.scalar("gmresSteps"), "true_rel_residual": lctx.true_residual(),
               "phases_ms": {p: lctx.phase_ms(p) for p in phases},
               "factor_flops": factor_flops(lctx.geti("elMatLTRange"))}
        lctx.close()
        out["literal_rule"] = {k: v for k, v in lit.items() if k != "factor_flops"}
        print("bench: literal-rule solve done", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        try:
            base = cpu_baseline()
            out["cpu_baseline"] = base
            if lit is not None:
                # the GPU solver's step count on the sample under the same rule
                sctx = astro.Context(cantilever_block(*base["sample_dims"]), astro.Options(literalMaxBW=True))
                sctx.static_solve()
                steps_s = sctx.scalar("gmresSteps")
                sctx.close()
                base["extrapolated_to_workload"] = extrapolate_cpu(
                    base, steps_s, {"dof": n, "nnz": nnz, "profile_entries": lit["profile_entries"],
                                    "factor_flops": lit["factor_flops"], "gmres_steps": lit["gmres_steps"]})
        except Exception as e:  # the baseline is a report, not the measurement
            print(f"bench: cpu baseline failed: {e}", file=sys.stderr)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
