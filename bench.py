"""Headline benchmark: batched mixed-precision BEM frequency sweep on one GPU.

Workload (a "step"): 256 wavenumbers log-spaced over 1-200 kHz, c = 1500 m/s,
on a fluid-filled prolate spheroid (semi-axes 0.02 x 0.01 m, g = 1.0357,
h = 1.0279), each frequency meshed as the smallest mapped icosphere with
longest edge <= lambda/6, clamped to N in [1280, 20480] triangles (2N Mueller
unknowns).  Each solve: FP32 regular-pair assembly + FP64 near/self panels,
complex-FP32 MFMA LU, FP64 iterative refinement to backward error 1e-13,
FP64 far field.  |dTS| is checked against the same sweep solved all-FP64.

Prints one JSON line (value = solves/s of the timed sweep).

The sweep runs six lanes (`--lanes`; two HIP streams each) under two
hardware queues per lane (GPU_MAX_HW_QUEUES, `--hw-queues`): with the
runtime's default of 4, the lanes' streams share queues and a lane's kernels
wait in order behind another lane's.  The variable is set in this process
before the library (and so the HIP runtime) is loaded.  Two solves with 16384
or more unknowns run at a time (`--heavy-lanes`): one heavy solve's panels,
triangular solves and refinement overlap the other's trailing updates, while
the other lanes take the smaller solves.

Time budget (`--budget-s`, default 400 s for the whole process): at most one
warm-up sweep runs; after `--min-steps` timed sweeps (3), further timed
sweeps run only while the rest of the run (the solo heavy solve, the FP64 reference sweep at
about 2.1x a mixed sweep, and the CPU leg) still fits.  `steps` and `warmup`
in the JSON are the counts actually run, beside the requested ones.

`--workload hmatrix` measures the hierarchical-matrix path: a rigid bent
capped cylinder (radius 2 mm, centerline 60 mm, bend radius 120 mm,
hemispherical caps) at 200k triangles, k = 1000 rad/m, cluster leaf 64,
eta = 2, ACA 1e-6 in complex FP32, GMRES(50) to an FP64 relative residual of
1e-8 for 64 incidences uniformly spaced in [0, pi], solved as one batch.

`--workload serve` measures the served form instead: seeded Poisson streams of
random bodies of revolution (rigid / pressure-release / fluid, N in [2000,
40000]) at 50, 200 and 800 requests/s through the server, reporting solves/s
and p50/p99 latency (submission to result) per rate.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "acousticscattering.jl_amd")]

C_SOUND = 1500.0
A_MAJOR, B_MINOR = 0.02, 0.01
G_RHO, H_C = 1.0357, 1.0279
MFMA_FP32_PEAK = 157.3e12
MFMA_FP64_PEAK = 78.6e12
MFMA_BF16_PEAK = 2.5e15  # dense
# matrix-core flops the bf16-split trailing update executes per algorithmic
# complex flop: 3 real products (3M) x 6 piece products x 2 M N K, over 8 M N K
BF16_SPLIT_FLOPS = 36.0 / 8.0
METRIC = "BEM sweep solves/s at |ΔTS|≤0.01 dB vs FP64 (N=20480); p50/p99 latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sweep_plan(bh, ctx, n_points, max_level):
    import workloads as W

    return W.sweep_plan(ctx, n_points, max_level)


TRAFFIC_PROFILE = os.path.join("profiles", "lu_gemm3m_traffic_fluid_spheroid_n20480.md")
TRAFFIC_PROFILE_BF16 = os.path.join("profiles", "lu_gemm_bf16_traffic_fluid_spheroid_n20480.md")
TRAFFIC_PROFILE_HM = os.path.join("profiles", "hmatrix_apply_traffic_bent_cylinder_200k.md")


def gemm_roofline(fp32, bf16, traffic_note):
    """Roofline of the trailing update that carried most of the time: the
    FP32 3M kernel (k_gemm3m, vs the FP32 MFMA peak) or the bf16-split one
    (k_gemm3m_bf, vs the dense bf16 MFMA peak in the matrix-core flops it
    executes, 4.5 per algorithmic flop; `algorithmic` = 8 M N K / time, which
    may exceed the FP32 peak)."""
    def rate(st):
        return st["flops"] / (st["ms"] * 1e-3) if st["ms"] > 0 else 0.0

    if bf16["ms"] > fp32["ms"]:
        g, r = bf16, rate(bf16)
        out = {"kernel": "k_gemm3m_bf (LU outer trailing update, 3M complex, bf16-split FP32 products)",
               "bound": "mfma", "achieved": r * BF16_SPLIT_FLOPS / 1e12, "peak": MFMA_BF16_PEAK / 1e12,
               "unit": "TFLOP/s", "frac": r * BF16_SPLIT_FLOPS / MFMA_BF16_PEAK,
               "algorithmic_tflops": r / 1e12, "algorithmic_vs_fp32_peak": r / MFMA_FP32_PEAK,
               "traffic": measured_traffic(TRAFFIC_PROFILE_BF16), "traffic_source": TRAFFIC_PROFILE_BF16}
    else:
        g, r = fp32, rate(fp32)
        out = {"kernel": "k_gemm3m<0, FULL> (LU trailing update, 3M complex)", "bound": "mfma",
               "achieved": r / 1e12, "peak": MFMA_FP32_PEAK / 1e12, "unit": "TFLOP/s",
               "frac": r / MFMA_FP32_PEAK, "traffic": measured_traffic(TRAFFIC_PROFILE),
               "traffic_source": TRAFFIC_PROFILE}
    out.update({"traffic_basis": traffic_note,
                "algorithmic_bytes_per_launch": g["bytes"] / max(1, g["launches"]),
                "launches": g["launches"], "avg_launch_ms": g["ms"] / max(1, g["launches"]),
                "other_update_kernel": {"launches": (fp32 if g is bf16 else bf16)["launches"],
                                        "ms": (fp32 if g is bf16 else bf16)["ms"]}})
    return out


def measured_traffic(path=TRAFFIC_PROFILE):
    """HBM bytes per launch of the FP32 trailing update, from the committed
    rocprofv3 --pmc summary of `bench.py --workload solve` (one N = 20480
    Mueller solve; the solves of that size carry most of the sweep's GEMM
    time): FETCH_SIZE and WRITE_SIZE passes, each scaled by the counter's
    reading for a plain copy of known size.  None when the summary is
    absent."""
    try:
        with open(os.path.join(ROOT, path)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


# Work of one fused kernel evaluation in the regular-pair assembly
# (k_assemble, all four operators S, D, K', H at one target, one source
# Gauss point and one medium), counting an FMA as 2 flops and the
# transcendentals apart: |r|^2 5, R and kR 2, the 1/(4 pi R) weight 2,
# e^{ikR} scaling 2, S 2, D 16, K' 16, H 18 = 63 flops, plus r = x - y
# (3 flops) shared by the two media of a Mueller pair (1.5 per medium),
# and 1 rsqrt + 1 sincos.
F_K_MULLER = 64.5
FP32_VECTOR_PEAK = 157.3e12  # published; packed FP32 (v_pk_*) is disabled in this build, which halves it


def assembly_roofline(sweep, solo):
    """Achieved FLOP rate of the regular-pair assembly (Mueller, FP32 in the
    mixed form) from the library's event timers: evaluations (N^2 x 7 points
    x 2 media per launch) x F_K_MULLER / event time."""
    def rate(st):
        return st["flops"] * F_K_MULLER / (st["ms"] * 1e-3) if st["ms"] > 0 else 0.0

    r_solo, r_sweep = rate(solo), rate(sweep)
    t_solo = solo["ms"] * 1e-3 / max(1, solo["launches"])
    return {"kernel": "k_assemble<float, MULLER> (regular panel pairs, both media)", "bound": "valu",
            "unit": "TFLOP/s", "achieved": r_solo / 1e12, "peak": FP32_VECTOR_PEAK / 1e12,
            "frac": r_solo / FP32_VECTOR_PEAK, "frac_of_nonpacked_peak": 2 * r_solo / FP32_VECTOR_PEAK,
            "basis": "one N=20480 launch alone (solo solve)",
            "sweep_achieved": r_sweep / 1e12, "sweep_launches": sweep["launches"],
            "flops_per_evaluation": F_K_MULLER, "transcendentals_per_evaluation": {"sincos": 1, "rsqrt": 1},
            "evaluations_per_launch": solo["flops"] / max(1, solo["launches"]),
            "bytes_written_per_launch": solo["bytes"] / max(1, solo["launches"]),
            "store_hbm_frac": (solo["bytes"] / max(1, solo["launches"]) / t_solo / 8e12) if t_solo > 0 else None,
            "avg_launch_ms": t_solo * 1e3}


def size_lanes(bh, ctx, plan, meshes, lanes, precision):
    """Allocate every lane's device arena at the sweep's largest footprint,
    untimed: one solve of the largest mesh and small ones on the other lanes.
    A context allocates its lane arenas once and reuses them for later sweeps
    of that footprint; allocating ~8 x 14 GB inside the timed sweep added
    0-3.4 s to it, depending on what the device was still clearing from the
    previous process."""
    fl = bh.FluidFilled(G_RHO, H_C)
    big, small = max(meshes), min(meshes)
    k_big = max(k for (_, k, L) in plan if L == big)
    ms = [meshes[big]] + [meshes[small]] * (lanes - 1)
    probs = [bh.make_problem(fl, k_big if i == 0 else plan[0][1], precision=precision) for i in range(lanes)]
    bh.batched_sweep(ms, probs, [(0.0, 0.0)] * lanes, ctx=ctx)


def run_sweep(bh, ctx, plan, meshes, precision, chunk=64, serial=False):
    """One pass over the sweep as one bem_sweep call (the library orders the
    solves, largest first, and runs the matrix-core-bound ones one at a time
    while its other lanes take the small ones); returns per-solve results in
    plan order and the wall time of the pass.  serial: one solve at a time on
    the calling thread (bem), for profilers that cannot follow the sweep's
    worker threads; the same solves and kernels, progress every `chunk`."""
    fl = bh.FluidFilled(G_RHO, H_C)
    order = sorted(range(len(plan)), key=lambda i: -plan[i][2])
    amps = np.empty(len(plan), dtype=np.complex128)
    lat = np.empty(len(plan))
    diags = [None] * len(plan)
    t0 = time.perf_counter()
    if serial:
        for n, i in enumerate(order):
            t1 = time.perf_counter()
            sol = bh.bem(meshes[plan[i][2]], fl, plan[i][1], incidence_angle=0.0, precision=precision)
            amps[i] = bh.scattering_amplitude(sol)
            lat[i] = (time.perf_counter() - t1) * 1e3
            diags[i] = sol.diagnostics
            if (n + 1) % chunk == 0 or n + 1 == len(order):
                log(f"  {precision}: {n + 1}/{len(plan)} solves, {time.perf_counter() - t0:.1f}s")
        return amps, lat, diags, time.perf_counter() - t0
    ms = [meshes[plan[i][2]] for i in order]
    probs = [bh.make_problem(fl, plan[i][1], precision=precision) for i in order]
    a, l, d = bh.batched_sweep(ms, probs, [(0.0, 0.0)] * len(order), ctx=ctx)
    for j, i in enumerate(order):
        amps[i], lat[i], diags[i] = a[j], l[j], d[j]
    log(f"  {precision}: {len(plan)}/{len(plan)} solves, {time.perf_counter() - t0:.1f}s")
    return amps, lat, diags, time.perf_counter() - t0


def cpu_baseline(plan, budget_s=60.0):
    """Oracle (C assembly with OpenMP + LAPACK LU through scipy) on a bounded
    sample of the sweep: one full solve per mesh class up to N = 5120
    (2N = 10240 unknowns), then one LU at 2N = 20480 of the N = 5120 system
    tiled as [[A, A/2], [A/2, A]] (nonsingular: its blocks diagonalise to
    (1 +- 1/2) A), when the 10240 LU's time x 8 fits the remaining budget.
    The sweep rate is extrapolated per class: assembly ~ N^2 from N = 5120,
    LU ~ n^3 from the largest LU timed (8x from 2N = 20480 to the heavy
    class's 2N = 40960)."""
    from oracle import meshes as om
    from oracle import oracle as O
    import scipy.linalg as sla

    threads = O.lib().orc_num_threads()
    classes = sorted(set(L for (_, _, L) in plan))
    counts = {L: sum(1 for p in plan if p[2] == L) for L in classes}
    per = {}
    sample = []
    t_start = time.perf_counter()
    A_keep = None
    for L in classes:
        if L > 4 or time.perf_counter() - t_start > budget_s:
            continue
        nodes, tri = om.spheroid(A_MAJOR, B_MINOR, L)
        m = O.Mesh(nodes, tri)
        k = next(p[1] for p in plan if p[2] == L)
        t0 = time.perf_counter()
        A = m.assemble(O.FLUID_MULLER, k, G_RHO, H_C)
        t1 = time.perf_counter()
        O.solve_fluid(m, G_RHO, H_C, k, A=A)
        t2 = time.perf_counter()
        log(f"  cpu baseline N={m.n}: assembly {t1 - t0:.2f}s, LU+solve {t2 - t1:.2f}s")
        per[L] = (t1 - t0, t2 - t1, m.n)
        sample.append(f"N={m.n} (2N={2 * m.n}): assembly {t1 - t0:.2f}s, LU+solve {t2 - t1:.2f}s")
        A_keep = A
    if not per:
        return None
    lmax = max(per)
    ta, tl, nref = per[lmax]
    lu_n, lu_t = 2 * nref, tl
    left = budget_s - (time.perf_counter() - t_start)
    if A_keep is not None and 9.0 * tl < left:
        n2 = 2 * A_keep.shape[0]
        big = np.empty((n2, n2), dtype=np.complex128, order="F")
        h = A_keep.shape[0]
        big[:h, :h] = A_keep
        big[h:, h:] = A_keep
        big[:h, h:] = 0.5 * A_keep
        big[h:, :h] = 0.5 * A_keep
        del A_keep
        t0 = time.perf_counter()
        sla.lu_factor(big, overwrite_a=True, check_finite=False)
        lu_t = time.perf_counter() - t0
        lu_n = n2
        del big
        log(f"  cpu baseline LU n={n2}: {lu_t:.2f}s")
        sample.append(f"complex LU n={n2} (tiled N={nref} system): {lu_t:.2f}s")
    total = 0.0
    for L in classes:
        n = 20 * 4 ** L
        if L in per:
            t = per[L][0] + per[L][1]
        else:
            t = ta * (n / nref) ** 2 + lu_t * (2 * n / lu_n) ** 3
        total += counts[L] * t
    return {"value": len(plan) / total, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": "FP64 oracle (OpenMP assembly, LAPACK LU) timed on " + "; ".join(sample) +
                      f"; sweep time extrapolated per class: assembly ~N^2 from N={nref}, "
                      f"LU ~n^3 from n={lu_n}",
            "seconds": time.perf_counter() - t_start}


def serve_rate(bh, ctx, rate, count, seed, lanes):
    """Replay one seeded stream through a server; latency = submission to result."""
    import workloads as W

    reqs = W.serving_stream(rate, count, seed)
    meshes = [bh.Mesh(r["nodes"], r["triangles"], ctx=ctx) for r in reqs]
    with bh.Server(ctx, lanes=lanes, keep_factors=True) as srv:
        t0 = time.perf_counter()
        tickets = []
        for r, m in zip(reqs, meshes):
            delay = r["t"] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(m, r["boundary"], r["k"]))
        res = [srv.wait(t) for t in tickets]
        wall = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array([x.total_ms for x in res])
    unknowns = [m.n * (2 if isinstance(r["boundary"], bh.FluidFilled) else 1) for r, m in zip(reqs, meshes)]
    return {"rate": rate, "requests": count, "seed": seed, "solves_per_s": count / wall, "wall_s": wall,
            "p50_latency_ms": float(np.percentile(lat, 50)), "p99_latency_ms": float(np.percentile(lat, 99)),
            "mean_queue_ms": float(np.mean([x.queue_ms for x in res])),
            "mean_run_ms": float(np.mean([x.run_ms for x in res])),
            "unknowns_min_max": [min(unknowns), max(unknowns)],
            "promoted_to_fp64": sum(1 for x in res if x.diagnostics["factor_precision"] == "double"),
            "failed": st["failed"]}


def serve_main(args):
    import bemhip as bh

    ctx = bh.Context(0)
    # warm-up: load every kernel once
    warm = bh.mesh(bh.Sphere(0.01), level=3, ctx=ctx)
    for bc in (bh.Rigid(), bh.PressureRelease(), bh.FluidFilled(1.03, 1.02)):
        with bh.Server(ctx, lanes=1) as srv:
            srv.wait(srv.submit(warm, bc, 100.0))
    rates = {}
    for i, rate in enumerate(args.rates):
        log(f"serving {args.requests} requests at {rate} req/s on {args.lanes} lanes ...")
        rates[str(rate)] = serve_rate(bh, ctx, rate, args.requests, 1000 + i, args.lanes)
        log(f"  {rates[str(rate)]['solves_per_s']:.1f} solves/s, p50 {rates[str(rate)]['p50_latency_ms']:.0f} ms, "
            f"p99 {rates[str(rate)]['p99_latency_ms']:.0f} ms")
    top = rates[str(args.rates[-1])]
    out = {
        "metric": "served BEM solves/s (Poisson stream of random bodies of revolution); p50/p99 latency",
        "value": top["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": top["wall_s"] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c32 LU / f64 refinement", "data": "synthetic",
        "config": {"workload": "served stream: bodies of revolution r=a sin(pi s)(1+eps cos 3 pi s), "
                               "a~U[5,50] mm, aspect~U[1,8], eps~U[0,0.2], rigid/soft/fluid, ka~U[0.5,15], "
                               f"N in [2000,40000], mixed precision, {args.lanes} lanes, factors resident, "
                               f"{os.environ['GPU_MAX_HW_QUEUES']} hardware queues",
                   "rates": args.rates, "requests_per_rate": args.requests},
        "p50_latency_ms": top["p50_latency_ms"], "p99_latency_ms": top["p99_latency_ms"],
        "per_rate": rates,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def solve_main(args):
    """One fluid Mueller solve of the 0.02 x 0.01 m spheroid at
    k = 500 rad/m, N = 20480 panels (2N = 40960 unknowns), mixed precision,
    on the calling thread (the workload the traffic counters are read on)."""
    import bemhip as bh

    ctx = bh.Context(0)
    fl = bh.FluidFilled(G_RHO, H_C)
    warm = bh.mesh(bh.Spheroid(A_MAJOR, B_MINOR), level=3, ctx=ctx)  # loads every kernel once
    bh.bem(warm, fl, 500.0, incidence_angle=0.0, precision="mixed")
    mesh = bh.mesh(bh.Spheroid(A_MAJOR, B_MINOR), level=5, ctx=ctx)
    ctx.set_lu_tuning(args.lu_tuning)
    if args.lu_tuning == "tuned":
        log(f"tuning the factorization shape for {2 * mesh.n} unknowns: {ctx.autotune(2 * mesh.n, 'mixed')}")
    times = []
    for s in range(args.warmup + args.steps):
        if s == args.warmup:
            ctx.reset_stats()
            ctx.set_profiling(True)
        t0 = time.perf_counter()
        sol = bh.bem(mesh, fl, 500.0, incidence_angle=0.0, precision="mixed")
        f = bh.scattering_amplitude(sol)
        if s >= args.warmup:
            times.append(time.perf_counter() - t0)
        log(f"solve {s + 1}: {time.perf_counter() - t0:.3f}s, TS {bh.target_strength(f):.4f} dB")
    ctx.set_profiling(False)
    g, gb = ctx.kernel_stats("lu_gemm"), ctx.kernel_stats("lu_gemm_bf16")
    t = float(np.mean(times))
    out = {
        "metric": "single fluid Mueller solve time (N=20480, 2N=40960 unknowns, mixed precision)",
        "value": t * 1e3, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "c32 LU / f64 refinement", "data": "synthetic",
        "config": {"workload": "fluid spheroid 0.02x0.01 m (g=1.0357, h=1.0279), k=500 rad/m, "
                               "icosphere L5 (N=20480), Mueller, mixed precision, one incidence"},
        "target_strength_db": bh.target_strength(f),
        "diagnostics": {k: v for k, v in sol.diagnostics.items() if isinstance(v, (int, float, str, bool))},
        "lu_tune": ctx.lu_tune(2 * mesh.n, "mixed"),
        "roofline": gemm_roofline(g, gb, "per launch of this solve's trailing updates"),
        "kernels": {k: ctx.kernel_stats(k) for k in ("lu_panel", "lu_gemm", "lu_gemm_bf16", "lu_split", "lu_trsm",
                                                        "lu_swap", "solve", "residual")},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


FP64_VALU_PEAK = 78.6e12


def nystrom_main(args):
    """The reference's own full-3D discretisation on the config-1 body: the
    38 kHz sphere (a = 1 cm, c = 1477.4 m/s, ka = 1.62) as a level-3 curved
    icosphere of 1280 quadratic (6-node) triangles with the order-4 rule
    (7680 quadrature nodes = unknowns), density-interpolation corrections.
    A step is four solves: rigid and pressure-release Burton-Miller and fluid
    Mueller (15360 unknowns) in FP64, and the fluid one in mixed precision."""
    import bemhip as bh
    from oracle import modal

    ctx = bh.Context(0)
    k = 2 * math.pi * 38000.0 / 1477.4
    g, h = 1028.9 / 1026.8, 1480.3 / 1477.4
    warm = bh.mesh(bh.Sphere(0.01), level=1, mesh_order=2, ctx=ctx)
    mesh = bh.mesh(bh.Sphere(0.01), level=3, mesh_order=2, ctx=ctx)
    cases = [("rigid", bh.Rigid(), "double"), ("soft", bh.PressureRelease(), "double"),
             ("fluid", bh.FluidFilled(g, h), "double"), ("fluid_mixed", bh.FluidFilled(g, h), "mixed")]
    for _, bc, prec in cases:  # loads every kernel once
        bh.bem(warm, bc, k, incidence_angle=0.0, precision=prec)
    times, per_case, amps = [], {c[0]: [] for c in cases}, {}
    for s in range(args.warmup + args.steps):
        if s == args.warmup:
            ctx.reset_stats()
            ctx.set_profiling(True)
        t0 = time.perf_counter()
        for name, bc, prec in cases:
            t1 = time.perf_counter()
            sol = bh.bem(mesh, bc, k, incidence_angle=0.0, precision=prec)
            amps[name] = bh.scattering_amplitude(sol)
            if s >= args.warmup:
                per_case[name].append(time.perf_counter() - t1)
        if s >= args.warmup:
            times.append(time.perf_counter() - t0)
        log(f"step {s + 1}: {time.perf_counter() - t0:.3f}s")
    ctx.set_profiling(False)
    stats = {c: ctx.kernel_stats(c) for c in ("nys_assemble", "nys_defect", "nys_fit", "lu_panel", "lu_gemm",
                                              "lu_gemm_fp64", "lu_gemm_bf16", "lu_trsm", "lu_swap", "solve",
                                              "residual", "far_field")}
    step = float(np.mean(times))
    errs = {}
    for name, bc, _ in cases:
        key = ("fluid", g, h) if name.startswith("fluid") else (name,)
        ref = modal.form_function(key, k, 0.01)
        errs[name] = abs(bh.target_strength(amps[name]) - modal.target_strength(ref))

    def line(st, bound, peak, unit):
        ms = st["ms"] / max(1, st["launches"])
        ach = st["flops"] / max(1, st["launches"]) / (ms * 1e-3) / 1e12
        return {"launches": st["launches"], "avg_launch_ms": ms, "achieved": ach, "peak": peak / 1e12,
                "unit": unit, "frac": ach * 1e12 / peak, "bound": bound}

    dfc = line(stats["nys_defect"], "valu", FP64_VALU_PEAK, "TFLOP/s")
    asm = line(stats["nys_assemble"], "valu", FP64_VALU_PEAK, "TFLOP/s")
    asm["bytes_per_launch"] = stats["nys_assemble"]["bytes"] / max(1, stats["nys_assemble"]["launches"])
    asm["store_GBps"] = asm["bytes_per_launch"] / (asm["avg_launch_ms"] * 1e-3) / 1e9
    roof = dict(dfc)
    roof.update({"kernel": "k_nys_defect (Green-identity defects of the density interpolation, FP64)",
                 "traffic": None,
                 "work": "2 complex multiply-adds (16 flops) per node pair, basis function (18 point sources) "
                         "and operator family; the kernel values themselves are counted under nys_assemble",
                 "assembly": dict(asm, kernel="k_nys_assemble (naive Nystrom entries, FP64, "
                                              "75 flops per pair and medium for S, D, K', H; 34 for S, D)")})
    # the reference's default for rigid bodies, compression=(method=:hmatrix,):
    # the same sphere on 20480 quadratic elements (122880 nodes), 64 incidences
    big = bh.mesh(bh.Sphere(0.01), level=5, mesh_order=2, ctx=ctx)
    angles = [(math.pi * c / 63, 0.0) for c in range(64)]
    t0 = time.perf_counter()
    H = bh.HMatrix(big, bh.Rigid(), k)
    t1 = time.perf_counter()
    hsol = H.solve(angles)
    t2 = time.perf_counter()
    ref_rigid = modal.target_strength(modal.form_function(("rigid",), k, 0.01))
    hts = [bh.target_strength(hsol.far_field(r)[0]) for r in range(len(angles))]
    hinfo = H.info()
    compressed = {
        "workload": "same sphere, curved icosphere L5 (20480 6-node triangles, 122880 nodes), rigid Burton-Miller, "
                    "H-matrix (leaf 64, eta 2, ACA 1e-6 c32) with the element corrections in the dense leaves, "
                    "GMRES to 1e-8, 64 incidences in [0, pi] in one batch",
        "build_s": t1 - t0, "solve_s": t2 - t1, "solves_per_s": len(angles) / (t2 - t0),
        "gmres_iterations": hsol.diagnostics(0)["iterations"],
        "max_abs_dTS_vs_modal_dB": max(abs(x - ref_rigid) for x in hts),
        "lowrank_blocks": hinfo["lowrank_blocks"], "dense_blocks": hinfo["dense_blocks"],
        "mean_rank": hinfo["mean_rank"],
    }
    log(f"compressed L5: build {t1 - t0:.2f}s solve {t2 - t1:.2f}s, max |dTS| "
        f"{compressed['max_abs_dTS_vs_modal_dB']:.2e} dB")
    cpu = None if args.no_cpu else nystrom_cpu_baseline(k, g, h)
    total = sum(v["ms"] for v in stats.values())
    out = {
        "metric": "Nystrom BEM solves/s (38 kHz sphere, 1280 quadratic elements, 7680 quadrature nodes)",
        "value": len(cases) / step, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": step * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (fluid also c32 LU / f64 refinement)", "data": "synthetic",
        "config": {"workload": "sphere a=0.01 m, f=38 kHz, c=1477.4 m/s (ka=1.62), curved icosphere L3 "
                               "(1280 6-node triangles), qorder 4 (7680 nodes), DIM corrections; rigid BM, "
                               "soft BM, fluid Mueller (g=1.0021, h=1.0020) FP64 and mixed"},
        "per_solve_ms": {n: 1e3 * float(np.mean(v)) for n, v in per_case.items()},
        "compressed": compressed,
        "abs_dTS_vs_modal_dB": errs,
        "roofline": roof,
        "kernel_share": {c: v["ms"] / total for c, v in stats.items() if v["launches"]},
        "kernels": stats,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def nystrom_cpu_baseline(k, g, h):
    """The numpy FP64 oracle (oracle/nystrom.py) on the level-2 mesh (1920
    nodes), fluid Mueller: assembly with corrections and LU+solve timed, then
    extrapolated to level 3 by n^2 and n^3."""
    from oracle import nystrom as NY
    from oracle import oracle as O
    import scipy.linalg as sla

    nodes, tri = NY.curved_icosphere(2, (0.01, 0.01, 0.01), 2)
    Q = NY.Quadrature(nodes, tri, 4)
    t0 = time.perf_counter()
    A = NY.assemble_fluid(Q, g, h, k)
    t1 = time.perf_counter()
    p_inc, q_inc = Q.incident(k, O.direction(0.0))
    rows, cols = O.equilibrate(A)
    lu = sla.lu_factor(A / rows[:, None] / cols[None, :], check_finite=False)
    sla.lu_solve(lu, np.concatenate([p_inc, q_inc]) / rows, check_finite=False)
    t2 = time.perf_counter()
    est = (t1 - t0) * 16 + (t2 - t1) * 64
    return {"value": 1.0 / est, "unit": "solves/s", "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())),
            "kind": "port", "seconds": t2 - t0,
            "sample": f"FP64 numpy oracle, fluid Mueller on the level-2 curved sphere (1920 nodes, 3840 unknowns): "
                      f"assembly+DIM {t1 - t0:.2f}s, LU+solve {t2 - t1:.2f}s; one level-3 fluid solve extrapolated "
                      f"by n^2 (assembly) and n^3 (LU): {est:.1f}s"}


def hmatrix_main(args):
    import bemhip as bh

    ctx = bh.Context(0)
    body = bh.Cylinder(0.002, 0.06, 0.12, 0.002)
    nodes, tri = bh.cylinder_arrays(body, args.hm_nphi)
    mesh = bh.Mesh(nodes, tri, ctx=ctx)
    k = 1000.0
    log(f"bent cylinder: {mesh.n} triangles, k = {k}")
    warm = bh.mesh(bh.Sphere(0.01), level=3, ctx=ctx)  # loads every kernel once
    bh.HMatrix(warm, bh.Rigid(), 100.0).solve([(0.0, 0.0)])
    angles = [(math.pi * c / (args.hm_rhs - 1), 0.0) for c in range(args.hm_rhs)]
    ctx.reset_stats()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    H = bh.HMatrix(mesh, bh.Rigid(), k, leaf_size=64, eta=2.0, aca_tolerance=1e-6, gmres_restart=50,
                   gmres_tolerance=1e-8)
    t1 = time.perf_counter()
    sol = H.solve(angles)
    t2 = time.perf_counter()
    ctx.set_profiling(False)
    info = H.info()
    diags = [sol.diagnostics(r) for r in range(len(angles))]
    ts = [bh.target_strength(sol.far_field(r)[0]) for r in range(len(angles))]
    stats = {name: ctx.kernel_stats(name) for name in ("h_apply", "h_proj", "h_rows", "h_aca", "h_ortho")}
    ap = stats["h_apply"]
    gbs = ap["bytes"] / (ap["ms"] * 1e-3) / 1e9 if ap["ms"] > 0 else 0.0
    tfs = ap["flops"] / (ap["ms"] * 1e-3) / 1e12 if ap["ms"] > 0 else 0.0
    log(f"build {t1 - t0:.2f}s solve {t2 - t1:.2f}s, {diags[0]['iterations']} iterations, "
        f"apply {ap['ms'] / max(1, ap['launches']):.1f} ms ({gbs:.0f} GB/s, {tfs:.1f} TF/s)")
    out = {
        "metric": "H-matrix batched solve: incidences solved per second (build + GMRES)",
        "value": len(angles) / (t2 - t0), "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": (t2 - t0) * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c32 blocks / f64 GMRES", "data": "synthetic",
        "config": {"workload": f"rigid bent capped cylinder a=2 mm, L=60 mm, R=120 mm, {mesh.n} triangles, "
                               f"k=1000, leaf 64, eta 2, ACA 1e-6 (c32), GMRES(50) to 1e-8, "
                               f"{len(angles)} incidences in [0, pi]"},
        "build_s": t1 - t0, "solve_s": t2 - t1, "gmres_iterations": diags[0]["iterations"],
        "max_relative_residual": max(d["relative_residual"] for d in diags),
        "converged": sum(d["converged"] for d in diags),
        "hmatrix": {key: info[key] for key in ("leaves", "dense_blocks", "lowrank_blocks", "dense_bytes",
                                                 "lowrank_bytes", "max_rank", "mean_rank", "unconverged_blocks",
                                                 "plan_ms", "assembly_ms", "aca_ms")},
        # 64 right-hand sides give the apply ~60 flops per byte moved (blocks
        # and vectors), six times FP64's machine balance (78.6 TFLOP/s over
        # 8 TB/s): the products, FP64 on the matrix cores, bound it
        "roofline": {"kernel": "compressed apply (64 right-hand sides)", "bound": "mfma", "achieved": tfs,
                     "peak": MFMA_FP64_PEAK / 1e12, "unit": "TFLOP/s", "frac": tfs * 1e12 / MFMA_FP64_PEAK,
                     "traffic": measured_traffic(TRAFFIC_PROFILE_HM), "traffic_source": TRAFFIC_PROFILE_HM,
                     "traffic_basis": "HBM bytes per 64-RHS apply (k_hproj, k_hproj_sum, k_hrows), FETCH_SIZE and WRITE_SIZE "
                                      "passes of this workload",
                     "hbm_gbs": gbs, "hbm_frac": gbs / 8000.0,
                     "flops_per_byte": ap["flops"] / ap["bytes"] if ap["bytes"] > 0 else None,
                     "launches": ap["launches"], "avg_launch_ms": ap["ms"] / max(1, ap["launches"]),
                     "kernels": {kn: {"avg_launch_ms": stats[kn]["ms"] / max(1, stats[kn]["launches"]),
                                      "tflops": (stats[kn]["flops"] / (stats[kn]["ms"] * 1e-3) / 1e12
                                                 if stats[kn]["ms"] > 0 else 0.0),
                                      "frac": (stats[kn]["flops"] / (stats[kn]["ms"] * 1e-3) / MFMA_FP64_PEAK
                                               if stats[kn]["ms"] > 0 else 0.0)}
                                 for kn in ("h_proj", "h_rows")}},
        "kernels": {kname: {"launches": v["launches"], "ms": round(v["ms"], 3)} for kname, v in stats.items()},
        "ts_db_min_max": [min(ts), max(ts)],
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


T_START = time.perf_counter()


def elapsed():
    return time.perf_counter() - T_START


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["sweep", "serve", "hmatrix", "solve", "nystrom"], default="sweep")
    ap.add_argument("--hm-nphi", type=int, default=138)
    ap.add_argument("--hm-rhs", type=int, default=64)
    ap.add_argument("--rates", type=float, nargs="+", default=[50.0, 200.0, 800.0])
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--lanes", type=int, default=None,
                    help="concurrent solves (sweep, default 6) or server lanes (serve, default 8)")
    ap.add_argument("--heavy-lanes", type=int, default=2,
                    help="sweep solves with >= 16384 unknowns allowed to run at once (default 2: one heavy "
                         "solve's panels, triangular solves and refinement overlap the other's trailing "
                         "updates; measured 14.5 solves/s against 12.9 with one and 13.8 with three)")
    ap.add_argument("--hw-queues", type=int, default=None, choices=range(1, 33), metavar="1..32",
                    help="HIP hardware queues (GPU_MAX_HW_QUEUES) the benchmark runs under "
                         "(default two per lane, at most 32)")
    ap.add_argument("--lu-tuning", choices=["tuned", "default", "split"], default="tuned",
                    help="factorization shapes: the tuner's per size class (default), the fixed default shape, "
                         "or the default with bf16-split outer updates (no tuning pass)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--min-steps", type=int, default=3,
                    help="timed sweeps run even when the budget estimate says they do not fit")
    ap.add_argument("--budget-s", type=float, default=400.0,
                    help="wall-clock budget of the whole sweep run (setup, warm-up, timed sweeps, FP64 "
                         "reference sweep, CPU leg): warm-up and timed sweeps beyond --min-steps stop "
                         "when the rest would not fit")
    ap.add_argument("--points", type=int, default=256)
    ap.add_argument("--max-level", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serial", action="store_true",
                    help="sweep one solve at a time on the main thread (for rocprofv3 --pmc)")
    ap.add_argument("--cpu-budget", type=float, default=60.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only (--gpus 1); across GPUs the frequency sweep would "
              "shard its frequencies, with no exchange but the amplitudes.", file=sys.stderr)
        sys.exit(2)
    if args.lanes is None:
        args.lanes = 8 if args.workload == "serve" else 6
    if args.hw_queues is None:
        args.hw_queues = min(32, 2 * args.lanes)
    # The HIP runtime reads GPU_MAX_HW_QUEUES when it initialises, which
    # happens when libbemhip is loaded below (nothing before it touches the
    # GPU), so setting it here takes effect in this process.
    os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    if args.workload == "serve":
        serve_main(args)
        return
    if args.workload == "hmatrix":
        hmatrix_main(args)
        return
    if args.workload == "solve":
        solve_main(args)
        return
    if args.workload == "nystrom":
        nystrom_main(args)
        return
    sweep_main(args)


def sweep_main(args):
    import bemhip as bh

    ctx = bh.Context(0)
    ctx.set_lanes(args.lanes)
    ctx.set_heavy_lanes(args.heavy_lanes)
    plan, meshes = sweep_plan(bh, ctx, args.points, args.max_level)
    sizes = {}
    for (_, _, L) in plan:
        sizes[meshes[L].n] = sizes.get(meshes[L].n, 0) + 1
    log(f"plan: {len(plan)} solves, panels per class {sizes}")

    # the tuner's pass over each size class and precision (untimed; a sweep
    # would otherwise run it on its first call)
    t_tune = time.perf_counter()
    tunes = {}
    ctx.set_lu_tuning(args.lu_tuning)
    for n_panels in sorted(sizes) if args.lu_tuning == "tuned" else []:
        for prec in ("mixed", "double"):
            r = ctx.autotune(2 * n_panels, prec)
            tunes[f"{2 * n_panels}/{prec}"] = r
            log(f"tuned {2 * n_panels} unknowns, {prec}: outer {r['outer']} group {r['group']} split {r['split']} "
                f"ahead {r['ahead']} graph {r['graph']}: {r['best_ms']:.1f} ms (default {r['default_ms']:.1f} ms)")
    tune_s = time.perf_counter() - t_tune
    # load every kernel once from one thread before the concurrent lanes start
    fl = bh.FluidFilled(G_RHO, H_C)
    small = meshes[min(meshes)]
    for prec in ("double", "mixed"):
        if args.serial:
            bh.bem(small, fl, plan[0][1], precision=prec)
        else:
            bh.batched_sweep([small], [bh.make_problem(fl, plan[0][1], precision=prec)], [(0.0, 0.0)], ctx=ctx)
    # The timed mixed sweep runs before the FP64 reference sweep: a mixed
    # sweep that follows an FP64 one in the same process (whose lane arenas
    # are twice as large and are freed for the mixed ones) ran its trailing
    # updates 18 % slower, while the FP64 sweep is not slowed by running second.
    if not args.serial:
        size_lanes(bh, ctx, plan, meshes, args.lanes, "mixed")
    log(f"setup done at {elapsed():.1f}s")

    # Budget: what must still run after the timed sweeps is the solo heavy
    # solve, the FP64 sweep (measured at about 2.1x the mixed one, with its
    # lane arenas) and the CPU leg.
    est = [30.0]  # one mixed sweep [s], until one is measured

    def reserve():
        cpu = 0.0 if args.no_cpu else args.cpu_budget + 15.0
        return 2.4 * est[0] + 20.0 + cpu

    def fits(n_sweeps):
        return elapsed() + n_sweeps * est[0] + reserve() <= args.budget_s

    # One warm-up sweep at most: more measured no difference (the arenas are
    # sized and every kernel loaded before it), and the budget goes to timed
    # sweeps instead.
    warm_done = 0
    for w in range(min(args.warmup, 1)):
        if not fits(1 + args.min_steps):
            break
        log(f"warmup {w + 1}/{args.warmup} ...")
        _, _, _, t = run_sweep(bh, ctx, plan, meshes, "mixed", serial=args.serial)
        est[0] = t
        warm_done += 1
    ctx.reset_stats()
    ctx.set_profiling(True)
    times, lats, amps, diags = [], [], None, None
    all_diags = []
    for s in range(args.steps):
        if s >= args.min_steps and not fits(1):
            log(f"budget: stopping after {s} timed sweeps ({elapsed():.0f}s of {args.budget_s:.0f}s used)")
            break
        amps, lat, diags, t = run_sweep(bh, ctx, plan, meshes, "mixed", serial=args.serial)
        times.append(t)
        est[0] = max(times)
        lats.extend(lat.tolist())
        all_diags.extend(diags)
        log(f"step {s + 1}/{args.steps}: {t:.2f}s ({len(plan) / t:.2f} solves/s), at {elapsed():.0f}s")
    ctx.set_profiling(False)
    stats = {name: ctx.kernel_stats(name) for name in
             ("assemble", "near", "lu_panel", "lu_gemm", "lu_gemm_bf16", "lu_split", "lu_graph", "lu_gemm_fp64",
              "lu_trsm", "lu_swap", "solve", "residual", "far_field")}
    # The same trailing-update kernel on one heavy solve with nothing else on
    # the GPU: under several lanes a launch's event time includes sharing the
    # CUs with the other lanes' kernels, which the sweep's figure cannot undo.
    heavy_mesh = meshes[max(meshes)]
    k_heavy = max(k for (_, k, L) in plan if meshes[L] is heavy_mesh)
    ctx.reset_stats()
    ctx.set_profiling(True)
    bh.bem(heavy_mesh, fl, k_heavy, incidence_angle=0.0, precision="mixed")
    ctx.set_profiling(False)
    ga = ctx.kernel_stats("assemble")
    solo = gemm_roofline(ctx.kernel_stats("lu_gemm"), ctx.kernel_stats("lu_gemm_bf16"),
                         "per launch of this solve's trailing updates")
    solo["workload"] = f"one mixed solve alone after the sweep: N={heavy_mesh.n} (2N unknowns), k={k_heavy:.1f}"
    log(f"FP64 reference sweep at {elapsed():.0f}s ...")
    if not args.serial:
        size_lanes(bh, ctx, plan, meshes, args.lanes, "double")
    a64, lat64, _, t64 = run_sweep(bh, ctx, plan, meshes, "double", serial=args.serial)
    log(f"FP64 sweep {t64:.2f}s ({len(plan) / t64:.2f} solves/s), at {elapsed():.0f}s")
    ts64 = 20 * np.log10(np.abs(a64))
    tsm = 20 * np.log10(np.abs(amps))
    dts = float(np.max(np.abs(tsm - ts64)))
    rel = float(np.max(np.abs(amps - a64) / np.abs(a64)))
    promoted = sum(1 for d in all_diags if d["factor_precision"] == "double")
    unconverged = sum(1 for d in all_diags if not d["converged"])
    worst_be = max(max(d["normwise_backward_error"], d["componentwise_backward_error"]) for d in all_diags)
    t_step = float(np.mean(times))
    value = len(plan) / t_step
    cpu_left = args.budget_s - elapsed() - 5.0
    cpu = None if args.no_cpu else cpu_baseline(plan, max(10.0, min(args.cpu_budget, cpu_left)))
    full = args.points == 256 and args.max_level == 5
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": len(times),
        "warmup": warm_done,
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "budget_s": args.budget_s,
        "ms_per_step": t_step * 1e3,
        "step_s": [round(t, 3) for t in times],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c32 LU / f64 refinement",
        "data": "synthetic",
        "config": {
            "workload": ("fluid spheroid 0.02x0.01 m (g=1.0357, h=1.0279) frequency sweep, "
                         f"{len(plan)} f log-spaced 1-200 kHz, c=1500, icosphere meshes with edge<=lambda/6 "
                         f"clamped to N<={20 * 4 ** args.max_level}, Mueller 2N unknowns, mixed precision"
                         + ("" if full else " [REDUCED]")),
            "panels_per_class": {str(k): v for k, v in sorted(sizes.items())},
            "lanes": "serial" if args.serial else args.lanes,
            "heavy_lanes": args.heavy_lanes,
            "hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"]),
            "global_batch": len(plan),
        },
        "max_abs_dTS_dB": dts,
        "max_rel_amplitude_vs_fp64": rel,
        "dTS_ok": dts <= 0.01,
        "promoted_to_fp64": promoted,
        "unconverged": unconverged,
        "max_backward_error": worst_be,
        "p50_latency_ms": float(np.percentile(lats, 50)),
        "p99_latency_ms": float(np.percentile(lats, 99)),
        "fp64_solves_per_s": len(plan) / t64,
        "speedup_vs_fp64": value / (len(plan) / t64),
        "roofline": dict(gemm_roofline(stats["lu_gemm"], stats["lu_gemm_bf16"],
                                       "per launch of an N=20480 solve's trailing updates (the launches of "
                                       "'solo'; compare with solo.algorithmic_bytes_per_launch): the counters "
                                       "are read one solve at a time"), solo=solo),
        "lu_tunes": tunes,
        "tune_s": tune_s,
        "assembly_roofline": assembly_roofline(stats["assemble"], ga),
        "kernels": {k: {"launches": v["launches"], "ms": round(v["ms"], 3)} for k, v in stats.items()},
        "cpu_baseline": cpu,
        "wall_s": elapsed(),
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
