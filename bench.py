"""Benchmark: forward+adjoint 4D-Var gradient evaluations per second.

Workload: 1.5-layer QG in a closed 512x512 basin, dx = 5 km, Ld = 20 km
(c = 2 m/s), dt = 900 s, 30-day window (2880 RK2 steps); 60 seeded nadir
tracks every 6 km plus one 120 km swath per day at 2 km pixels (sigma =
0.02 m); initial-state control with x_b = 0, sigma_b = 0.1 m; checkpoints
every 54 steps (about sqrt(2880)) plus every observation time.  One "step" of
this benchmark is one full ``cost_and_grad`` evaluation (forward sweep with
misfits, reverse sweep with the hand-written adjoint), replayed from a HIP
graph, with the control vectors already resident in HBM.  Every evaluation
is taken at X = 0.5 x the truth initial SSH of each member, a flowing state.

``value`` is the batched path: ``--batch`` independent problems of this
workload (own initial fields, tracks and noise) advanced in the same kernel
launches, with the whole trajectory tape kept in HBM.  Measured beside it in
the same run: the plain single-problem path, both paths with checkpointed
trajectories (segment start states kept, each segment recomputed in the
reverse sweep), the ratios, the split-product gradient against FP64-MFMA
over the whole window, the 48-problem uneven set, the stencil rooflines and
the CPU oracle.  ``--batch 1`` makes the plain path the headline.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: this build measures a
single GPU.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fwd+adjoint 4D-Var gradients/sec, QG 512² 30-day window, grad rel.err≤1e-5"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK = {"f64": 78.6, "f32": 157.3, "i8": 5000.0}  # dense TFLOP/s (TOP/s for int8), MI355X


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="float64", choices=["float64", "float32"])
    ap.add_argument("--batch", type=int, default=8,
                    help="independent 512x512 problems advanced in the same launches")
    ap.add_argument("--tape", default="auto", choices=["auto", "full", "checkpoint"],
                    help="trajectory storage of the batched path (auto: full if it fits in HBM)")
    ap.add_argument("--window-steps", type=int, default=None,
                    help="shorten the window (profiling only; the default is the full 2880)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the single-problem measurement")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-steps", type=int, default=96,
                    help="steps of the window the CPU oracle is timed on (one day by default)")
    ap.add_argument("--no-checkpoint", action="store_true",
                    help="skip the checkpointed (segment start states + recompute) legs")
    ap.add_argument("--no-problem-set", action="store_true",
                    help="skip the 48-problem uneven set in the default run")
    ap.add_argument("--minimize", action="store_true",
                    help="the 4D-Var minimisation instead: --iterations L-BFGS iterations with the "
                         "vectors in HBM, batched and plain paths")
    ap.add_argument("--iterations", type=int, default=30)
    ap.add_argument("--oracle-full", action="store_true",
                    help="one full-window FP64 oracle gradient (all host cores, minutes) against "
                         "the GPU gradients of both sine-GEMM modes, at X = 0.5 x truth")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernels eagerly instead of replaying a captured HIP graph")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the shape-based GEMM tile choice instead of timing the candidates")
    ap.add_argument("--no-stencil", action="store_true",
                    help="skip the batched stencil HBM measurement")
    ap.add_argument("--gemm-only", action="store_true",
                    help="only the inversion-GEMM roofline of the batched plan (no problem, no tape): "
                         "the command the GEMM counter runs profile")
    ap.add_argument("--stencil-only", action="store_true",
                    help="only the stencil measurement and a copy of known size (for counter runs)")
    ap.add_argument("--problem-set", action="store_true",
                    help="the 48-problem uneven-size workload instead (aggregate gradients/s)")
    ap.add_argument("--problem-set-size", type=int, default=48,
                    help="first N problems of the problem set (profiling)")
    ap.add_argument("--problem-set-split", type=int, default=6,
                    help="problem set: split-product digits of the ragged engine's sine GEMMs "
                         "(0 = FP64 MFMA)")
    ap.add_argument("--streams", type=int, default=4,
                    help="HIP streams the problem set is spread over")
    ap.add_argument("--dst-split", type=int, default=6,
                    help="batched path: sine-matrix GEMMs as split products on int8 MFMA with "
                         "this many digits per FP64 operand (0 = FP64 MFMA)")
    ap.add_argument("--plain-dst-split", type=int, default=0,
                    help="plain single-problem path: the same choice (FP64 MFMA by default, "
                         "the faster one for a single 512^2 problem)")
    ap.add_argument("--tiled", action="store_true",
                    help="the tiled QG model (QgmWithTiles) on the 512x512 grid instead: step and "
                         "adjoint of all tiles in shared launches (grid-point-steps/s)")
    ap.add_argument("--obs-build", action="store_true",
                    help="construction of the L3 observation operator for the 512x512 30-day "
                         "observations: the device neighbour search against the KDTree oracle")
    ap.add_argument("--merge", action="store_true",
                    help="time the merge of 9 overlapping subwindow analyses onto a 512x512 grid "
                         "over 64 dates")
    ap.add_argument("--basis", action="store_true",
                    help="the balanced-motion wavelet basis (Basis_bm) on a 512x512 lon/lat grid "
                         "instead: operg + operg_transpose over the window's basis times")
    ap.add_argument("--sw", action="store_true",
                    help="the 3-layer shallow-water model (256x256, dx 8 km, beta plane, dt 120 s) instead: "
                         "forward steps and the gradient of a final-state cost over the 10-day window")
    ap.add_argument("--sw-days", type=float, default=10.0, help="shallow-water window (days)")
    ap.add_argument("--sw4dvar", action="store_true",
                    help="the 3-layer shallow-water 4D-Var with the 3-D wavelet forcing control instead "
                         "(256x256, 10-day window, 30 nadir tracks): gradient evaluations per second, the "
                         "basis products' roofline, parity against the oracle cost on short windows")
    ap.add_argument("--tile-size", type=int, default=32)
    ap.add_argument("--tile-overlap", type=int, default=16)
    ap.add_argument("--tiled-steps", type=int, default=16,
                    help="tiled model: steps per step / step_adj call (one segment)")
    ap.add_argument("--check-f32", action="store_true",
                    help="also evaluate the FP32 path and report its gradient error vs FP64")
    return ap.parse_args()


def gemm_roofline(plan, precision, n_inv=100, reps=5):
    """Average duration of the inversion GEMMs on the GPU: n_inv inversions are
    captured in one HIP graph (so host launch cost is out of the picture) and
    the replays are timed with HIP events on the stream they run on.

    Algorithmic flops per inversion and batch member: 4 m_y m_x (m_y + m_x)
    (two 2-D sine transforms as dense products), a quarter of it per GEMM
    launch; ``achieved`` uses that count.  With folded sine matrices the
    kernels execute half of it: ``executed_tflops`` / ``executed_frac`` give
    the MFMA utilisation on the flops actually issued.  ``traffic`` (bytes per
    launch moved below L2) comes from the committed counter runs of this
    configuration (FETCH_SIZE / WRITE_SIZE, calibrated on a copy of known size),
    when there is one.

    Split-product plans (``dst_split`` digits) run the products on int8 MFMA:
    ``achieved`` is then the same algorithmic flops per second against the
    int8 peak, and ``executed_int8_tops`` the int8 work the kernels issue
    (digit pairs x 2 M N K over the padded product shapes)."""
    import torch
    q = torch.randn(plan.shape, dtype=plan.dtype, device=plan.device) * 1e-6
    hb = torch.zeros_like(q)
    out = torch.empty_like(q)
    for _ in range(3):
        plan.pv2h(q, hb, None, out)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(n_inv):
            plan.pv2h(q, hb, None, out)
    g.replay()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        g.replay()
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    my, mx = plan.ny - 2, plan.nx - 2
    alg_flops_launch = my * mx * (my + mx) * plan.nb
    exec_flops_launch = alg_flops_launch / 2 if plan.folded else alg_flops_launch
    per_launch_s = ms * 1e-3 / (4 * n_inv * reps)
    split = getattr(plan, "dst_split", 0)
    traffic, src = None, None
    rec = os.path.join(ROOT, "profiles", f"gemm_traffic_qg{plan.ny}_b{plan.nb}"
                                         f"{f'_split{split}' if split else ''}.json")
    if precision == "float64" and os.path.exists(rec):
        with open(rec) as f:
            traffic, src = json.load(f)["traffic_bytes_per_launch"], os.path.relpath(rec, ROOT)
    if split:
        # int8 MFMA work the split products issue: digit pairs x 2 M N K over the
        # zero-padded product shapes (half sizes when folded), averaged over the
        # four stages of an inversion.  ``achieved`` / ``frac`` count only the
        # algorithmic DST work (the same flops as the FP64 path) against the int8
        # MFMA peak the kernels run on; the executed int8 work has its own keys.
        def rup(v):
            return (v + 31) // 32 * 32
        pairs = split * (split + 1) // 2
        if plan.folded:
            HY, HX, MY, MX = rup(my // 2), rup(mx // 2), rup(my), rup(mx)
            macs, nz = [HY * MX * HY, MY * HX * HX] * 2, 2 * plan.nb
        else:
            MY, MX = rup(my), rup(mx)
            macs, nz = [MY * MX * MY, MY * MX * MX] * 2, plan.nb
        i8_ops = 2 * pairs * nz * sum(macs) / 4
        executed = i8_ops / per_launch_s / 1e12
        achieved = alg_flops_launch / per_launch_s / 1e12
        return {"bound": "mfma", "kernel": "k_gemm_split (sine-matrix GEMMs as int8 split products "
                                           "of FP64 operands, 4 launches per inversion)",
                "achieved": round(achieved, 3), "peak": MFMA_PEAK["i8"], "unit": "TFLOP/s",
                "frac": round(achieved / MFMA_PEAK["i8"], 4), "traffic": traffic,
                "traffic_source": src, "avg_launch_us": round(per_launch_s * 1e6, 3),
                "algorithmic_flops_per_launch": alg_flops_launch, "folded_dst": plan.folded,
                "note": "achieved = algorithmic DST flops per launch / launch time, against the "
                        "dense int8 MFMA peak the products run on",
                "executed_int8_ops_per_launch": i8_ops, "digits": split, "digit_pairs": pairs,
                "executed_int8_tops": round(executed, 3),
                "executed_int8_frac": round(executed / MFMA_PEAK["i8"], 4),
                "achieved_over_fp64_mfma_peak": round(achieved / MFMA_PEAK["f64"], 4)}
    peak = MFMA_PEAK["f64" if precision == "float64" else "f32"]
    achieved = alg_flops_launch / per_launch_s / 1e12
    executed = exec_flops_launch / per_launch_s / 1e12
    return {"bound": "mfma", "kernel": "k_gemm / k_gemm2 (sine-matrix GEMMs of the Helmholtz "
                                       "inversion, 4 launches per inversion)",
            "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": traffic, "traffic_source": src,
            "avg_launch_us": round(per_launch_s * 1e6, 3),
            "algorithmic_flops_per_launch": alg_flops_launch,
            "executed_flops_per_launch": exec_flops_launch, "folded_dst": plan.folded,
            "executed_tflops": round(executed, 3), "executed_frac": round(executed / peak, 4)}


def stencil_roofline(case, precision, nbatch=64, n_launch=20, reps=5, rounds_n=5):
    """HBM roofline of the fused RK stage stencil (``k_stage`` behind
    ``vd_qg_rhs``: out = base + dt * rhs(h, q)) on ``nbatch`` problems of the
    workload's grid at once, so that its working set (three input fields and
    one output per problem) is far larger than the 256 MB Infinity Cache.
    Algorithmic bytes per point: read h, q, base and write out (4 words) plus
    the shared int8 mask once per launch; base is a field of its own (as in
    the second RK2 stage), so all four streams go to memory.  The launch time is
    the median of ``rounds_n`` timed rounds of ``reps`` graph replays."""
    import torch
    from vardyn_amd.qg import QgPlan
    plan = QgPlan(case.ny, case.nx, case.dx, case.dx, case.dt, case.c, case.f0,
                  time_scheme=case.time_scheme, precision=precision, nbatch=nbatch)
    g = torch.Generator(device=plan.device).manual_seed(5)
    h = 0.1 * torch.randn(plan.shape, dtype=plan.dtype, device=plan.device, generator=g)
    q = plan.h2pv(h, torch.zeros_like(h))
    base = q + 1e-9 * torch.randn(plan.shape, dtype=plan.dtype, device=plan.device, generator=g)
    out = torch.empty_like(q)
    for _ in range(3):
        plan.rhs(h, q, base, case.dt, out)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        for _ in range(n_launch):
            plan.rhs(h, q, base, case.dt, out)
    graph.replay()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    rounds = []
    for _ in range(rounds_n):
        e0.record(s)
        for _ in range(reps):
            graph.replay()
        e1.record(s)
        torch.cuda.synchronize()
        rounds.append(e0.elapsed_time(e1) * 1e-3 / (n_launch * reps))
    per_launch_s = sorted(rounds)[len(rounds) // 2]  # median round
    npts = case.ny * case.nx
    nbytes = nbatch * npts * 4 * q.element_size() + npts
    gbs = nbytes / per_launch_s / 1e9
    del plan, h, q, base, out, graph
    torch.cuda.empty_cache()
    return {"bound": "hbm", "kernel": "k_stage (fused h2uv + upwind advection + RK stage update)",
            "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "batch": nbatch,
            "bytes_per_launch": nbytes, "avg_launch_us": round(per_launch_s * 1e6, 2),
            "rounds_us": [round(r * 1e6, 2) for r in rounds]}


def adj_stencil_roofline(case, precision, nbatch=64, n_launch=20, reps=5, rounds_n=5):
    """HBM roofline of the adjoint RK stage stencil (``k_adj_fused`` behind
    ``vd_qg_rhs_adj``: lq = acc + coef drhs/dq^T lam, lh = coef drhs/dh^T lam)
    on ``nbatch`` problems of the workload's grid at once (working set far above
    the 256 MB Infinity Cache).  Algorithmic bytes per point: read h, q, lam,
    acc and write lq, lh (6 words) plus the shared int8 mask once per launch.
    Launch time: median of ``rounds_n`` rounds of ``reps`` graph replays."""
    import torch
    from vardyn_amd.qg import QgPlan
    plan = QgPlan(case.ny, case.nx, case.dx, case.dx, case.dt, case.c, case.f0,
                  time_scheme=case.time_scheme, precision=precision, nbatch=nbatch)
    g = torch.Generator(device=plan.device).manual_seed(6)
    h = 0.1 * torch.randn(plan.shape, dtype=plan.dtype, device=plan.device, generator=g)
    q = plan.h2pv(h, torch.zeros_like(h))
    lam = torch.randn(plan.shape, dtype=plan.dtype, device=plan.device, generator=g)
    acc = torch.randn(plan.shape, dtype=plan.dtype, device=plan.device, generator=g)
    lq, lh = torch.empty_like(q), torch.empty_like(q)
    for _ in range(3):
        plan.rhs_adj(h, q, lam, case.dt, acc, lq, lh)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        for _ in range(n_launch):
            plan.rhs_adj(h, q, lam, case.dt, acc, lq, lh)
    graph.replay()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    rounds = []
    for _ in range(rounds_n):
        e0.record(s)
        for _ in range(reps):
            graph.replay()
        e1.record(s)
        torch.cuda.synchronize()
        rounds.append(e0.elapsed_time(e1) * 1e-3 / (n_launch * reps))
    per_launch_s = sorted(rounds)[len(rounds) // 2]
    npts = case.ny * case.nx
    nbytes = nbatch * npts * 6 * q.element_size() + npts
    gbs = nbytes / per_launch_s / 1e9
    del plan, h, q, lam, acc, lq, lh, graph
    torch.cuda.empty_cache()
    return {"bound": "hbm", "kernel": "k_adj_fused (adjoint RK stage: transposed upwind advection, "
                                      "beta / diffusion terms and h2uv)",
            "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "batch": nbatch,
            "bytes_per_launch": nbytes, "avg_launch_us": round(per_launch_s * 1e6, 2),
            "rounds_us": [round(r * 1e6, 2) for r in rounds]}


def copy_calibration(nbytes=256 << 20):
    """A plain device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE
    counters of a profiled run (bytes moved: nbytes read + nbytes written)."""
    import torch
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * nbytes / (e0.elapsed_time(e1) * 1e-3) / 1e9
    return {"bytes_read": nbytes, "bytes_written": nbytes, "achieved_GBs": round(gbs, 1)}


def control_point(truth):
    """The control the evaluations are timed and checked at: half the truth
    initial SSH of each member (a flowing state, away from the background
    x_b = 0 where the forward stays at rest and the tangent is time-invariant)."""
    return 0.5 * np.asarray(truth, dtype=np.float64).ravel()


def oracle_cost(pb, case):
    from oracle.fourdvar import BasisInitialState, CostOracle, Cov, ModelOracle
    from oracle.obsop import ObsopOracle
    from oracle.qg import QgmOracle
    qg = QgmOracle(dx=case.dx, dy=case.dx, dt=case.dt, c=case.c, f=case.f0,
                   time_scheme=case.time_scheme, shape=(case.ny, case.nx))
    H = ObsopOracle(pb.grid.lon, pb.grid.lat, pb.grid.DX, pb.grid.DY, pb.T, case.dt, pb.dict_obs,
                    Npix=4)
    return CostOracle(ModelOracle(qg), H, BasisInitialState((case.ny, case.nx), case.sigma_b),
                      pb.checkpoints, pb.T, dtbasis=case.checkpoint_steps, B=Cov(case.sigma_b),
                      Xb=np.zeros(case.ny * case.nx))


def timed_oracle(cost, X, threads, progress=False):
    """The oracle's cost_and_grad on ``threads`` host threads (BLAS and scipy.fft
    workers), optionally logging each model segment so that long runs report."""
    import scipy.fft
    from threadpoolctl import threadpool_limits
    M = cost.M
    step0, adj0 = M.step, M.step_adj
    t_start = time.perf_counter()
    if progress:
        count = {"f": 0, "a": 0}

        def step(var, params, nstep):
            count["f"] += 1
            if count["f"] % 12 == 0:
                log(f"oracle: forward segment {count['f']} ({time.perf_counter() - t_start:.0f} s)")
            return step0(var, params, nstep)

        def step_adj(advar, adparams, var, nstep):
            count["a"] += 1
            if count["a"] % 12 == 0:
                log(f"oracle: adjoint segment {count['a']} ({time.perf_counter() - t_start:.0f} s)")
            return adj0(advar, adparams, var, nstep)
        M.step, M.step_adj = step, step_adj
    try:
        with threadpool_limits(limits=threads), scipy.fft.set_workers(threads):
            t0 = time.perf_counter()
            J, G = cost.cost_and_grad(X)
            t = time.perf_counter() - t0
    finally:
        M.step, M.step_adj = step0, adj0
    return J, G, t


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def cpu_baseline(case, sample_steps, precision, dst_split=0):
    """The FP64 oracle's ``cost_and_grad`` over a bounded sample of the same
    workload (the first ``sample_steps`` steps of the window with their
    observations, at the control of the timed evaluations), on one host thread
    (the reference pins MKL/OMP to one thread, mapping/src/__init__.py:3-5) and
    on all host cores, each scaled to the full window.  The same sample is
    evaluated by the GPU engine (FP64-MFMA sine GEMMs, and with ``dst_split``
    digits of split products when given), and the relative errors of its
    gradient and cost against the oracle's are reported beside the timing."""
    import torch
    from vardyn_amd.problem import build_problem
    pb = build_problem(case, precision=precision, n_steps=sample_steps, control="ic",
                       tape="full")
    X = control_point(pb.truth0)
    J, G = pb.engine.cost_and_grad(X)
    Js = Gs = None
    if dst_split:
        pbs = build_problem(case, precision=precision, n_steps=sample_steps, control="ic",
                            tape="full", truth_plan=pb.plan, dst_split=dst_split)
        Js, Gs = pbs.engine.cost_and_grad(X)
        del pbs
    cost = oracle_cost(pb, case)
    ncpu = os.cpu_count() or 1
    Jr, Gr, t1 = timed_oracle(cost, X, 1)
    _, Ga, tn = timed_oracle(cost, X, ncpu)
    scale = case.n_steps / sample_steps
    del pb
    gc.collect()
    torch.cuda.empty_cache()
    extra = {}
    if Gs is not None:
        extra = {"gpu_split_grad_rel_err_vs_oracle": rel(Gs, Gr),
                 "gpu_split_J_rel_err_vs_oracle": float(abs(Js - Jr) / abs(Jr)),
                 "split_digits": dst_split}
    return {"value": round(1.0 / (tn * scale), 6), "unit": "gradients/s", "cores": ncpu,
            "kind": "port",
            "sample": f"FP64 NumPy/SciPy oracle cost_and_grad over the first {sample_steps} RK2 "
                      f"steps of the {case.ny}x{case.nx} window with their observations at "
                      f"X = 0.5 x truth ({tn:.1f} s on {ncpu} threads, {t1:.1f} s on 1), "
                      f"scaled by {case.n_steps}/{sample_steps}",
            "one_thread": {"value": round(1.0 / (t1 * scale), 6), "cores": 1,
                           "sample_s": round(t1, 2)},
            "all_cores": {"value": round(1.0 / (tn * scale), 6), "cores": ncpu,
                          "sample_s": round(tn, 2)},
            "host": platform.processor() or platform.machine(), "host_cpus": ncpu,
            "threads_agree": rel(Ga, Gr),
            "gpu_grad_rel_err_vs_oracle": rel(G, Gr),
            "gpu_J_rel_err_vs_oracle": float(abs(J - Jr) / abs(Jr)), **extra}


def measure(pb, args, label, X, tiles=None, steps=None, warmup=None):
    """Autotune (unless ``tiles`` are given), capture and time evaluations of
    ``pb.engine`` at the control ``X`` (host array of B x nctrl values)."""
    import torch
    eng = pb.engine
    nb = pb.plan.nb
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    t0 = time.time()
    if tiles is None:
        tiles = pb.plan.autotune(reps=5) if not args.no_autotune else None
    log(f"{label}: GEMM tile configurations {tiles} ({time.time() - t0:.2f}s); "
        f"{pb.H.n_total} obs at {len(pb.H.t_obs)} times per member, "
        f"{len(pb.checkpoints) - 1} segments, tape {eng.tape_mode}")
    X0 = torch.as_tensor(np.asarray(X, dtype=np.float64), device=pb.plan.device).reshape(-1)
    if not args.no_graph:
        t0 = time.time()
        eng.capture()
        log(f"{label}: captured one gradient evaluation in a HIP graph ({time.time() - t0:.1f}s)")
    for w in range(warmup):
        J, G = eng.evaluate(X0)
        torch.cuda.synchronize()
        log(f"{label}: warmup {w}: J[0]={J[0].item():.6e}")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        J, G = eng.evaluate(X0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t1
    ms = dt * 1e3 / steps
    r = {"value": nb * steps / dt, "ms_per_step": ms, "batch": nb, "tiles": tiles,
         "tape": eng.tape_mode, "J0": J[0].item(),
         "gnorm0": float(torch.linalg.vector_norm(G[0]).item()), "steps": steps}
    log(f"{label}: timed {steps} evaluations of {nb} problem(s): {ms:.1f} ms each; "
        f"J[0]={r['J0']:.10e} |G[0]|={r['gnorm0']:.10e}")
    return r


def retape(pb, tape):
    """Replace ``pb.engine`` by an engine of the same problem(s) with another
    trajectory storage ("full" or "checkpoint"); the old tape is freed first."""
    from vardyn_amd.fourdvar import FourDVar
    H = pb.Hs if getattr(pb, "Hs", None) is not None and pb.plan.nb > 1 else pb.H
    pb.engine = None
    release(None)
    pb.engine = FourDVar(pb.plan, H, pb.checkpoints, pb.T, control="ic",
                         sigma_b=pb.case.sigma_b, tape=tape, dtbasis=pb.case.checkpoint_steps)
    return pb


def split_check(case, pb_ref, precision, n_steps, digits, X):
    """Gradient and cost of the workload with split-product sine GEMMs
    (``digits`` per operand line) against the same problem -- same
    observations -- with FP64-MFMA GEMMs (``pb_ref``), over the whole window, at
    the control ``X`` of the timed evaluations."""
    import torch
    from vardyn_amd.problem import build_problem
    X0 = torch.as_tensor(X, device=pb_ref.plan.device)
    Jr, Gr = pb_ref.engine.run(X0)
    Jr, Gr = Jr.clone(), Gr.clone()
    pbs = build_problem(case, precision=precision, n_steps=n_steps, control="ic", tape="full",
                        truth_plan=pb_ref.plan, dst_split=digits)
    Js, Gs = pbs.engine.run(X0)
    torch.cuda.synchronize()
    r = {"grad_rel_err_vs_fp64_mfma": float((torch.linalg.vector_norm(Gs - Gr) /
                                             torch.linalg.vector_norm(Gr)).item()),
         "J_rel_err_vs_fp64_mfma": abs(Js[0].item() - Jr[0].item()) / abs(Jr[0].item()),
         "window_steps": n_steps, "digits": digits, "control": "0.5 x truth initial SSH"}
    # the metric's accuracy bar: gradient within 1e-5 (relative) of FP64
    r["within_1e-5"] = r["grad_rel_err_vs_fp64_mfma"] <= 1e-5
    if not r["within_1e-5"]:
        log(f"WARNING: split-product gradient error {r['grad_rel_err_vs_fp64_mfma']:.3e} exceeds "
            f"the 1e-5 bar; the batched figure does not meet the metric's accuracy")
    del pbs
    release(None)
    return r


def oracle_full(args):
    """One full-window gradient of the 512 x 512 workload by the FP64 oracle on
    all host cores, at X = 0.5 x truth, against the GPU engine's gradients with
    FP64-MFMA sine GEMMs and with split products of 5 and 6 digits (same
    problem, same observations).  Several minutes of host time: a run of its
    own, not part of the default benchmark."""
    import torch
    from vardyn_amd import synthetic as syn
    from vardyn_amd.problem import build_problem
    case = syn.case_basin512()
    n = args.window_steps or case.n_steps
    pb = build_problem(case, precision="float64", n_steps=n, control="ic", tape="full")
    X = control_point(pb.truth0)
    t0 = time.perf_counter()
    J0, G0 = pb.engine.cost_and_grad(X)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    gpu = {"fp64_mfma": (J0, G0)}
    for d in (5, 6):
        pbs = build_problem(case, precision="float64", n_steps=n, control="ic", tape="full",
                            truth_plan=pb.plan, dst_split=d)
        gpu[f"split{d}"] = pbs.engine.cost_and_grad(X)
        del pbs
        release(None)
    log(f"oracle: GPU gradients done ({gpu_s:.1f} s for the FP64-MFMA one); oracle over {n} "
        f"steps, {len(pb.checkpoints) - 1} segments")
    cost = oracle_cost(pb, case)
    ncpu = os.cpu_count() or 1
    Jr, Gr, t = timed_oracle(cost, X, ncpu, progress=True)
    out = {"workload": f"{case.name}: {case.ny}x{case.nx}, {n} RK2 steps, "
                       f"{pb.H.n_total} observations", "control": "X = 0.5 x truth initial SSH",
           "oracle": {"J": float(Jr), "grad_norm": float(np.linalg.norm(Gr)), "seconds": round(t, 1),
                      "threads": ncpu, "host": platform.processor() or platform.machine(),
                      "gradients_per_s": round(1.0 / t, 6)},
           "gpu": {}}
    for k, (J, G) in gpu.items():
        out["gpu"][k] = {"J": float(J), "J_rel_err_vs_oracle": float(abs(J - Jr) / abs(Jr)),
                         "grad_rel_err_vs_oracle": rel(G, Gr),
                         "grad_max_abs_err_over_max": float(np.max(np.abs(G - Gr)) /
                                                            np.max(np.abs(Gr))),
                         "within_1e-5": rel(G, Gr) <= 1e-5}
    log(f"oracle comparison: {json.dumps(out['gpu'])}")
    return out


def minimisation(args):
    """The 4D-Var minimisation of the configured workload: ``--iterations``
    L-BFGS iterations (30 by default) from the background X = 0 with the
    iterates, gradients and correction pairs in HBM (``inv.minimize_device``:
    FP64 device inner products, Moré-Thuente line search as L-BFGS-B's), for
    the batched path (``--batch`` problems, each its own minimisation in
    lockstep) and for the plain single-problem path.  Reports wall time per
    minimisation, evaluations, and the cost reduction of every member."""
    import torch
    from vardyn_amd import synthetic as syn
    from vardyn_amd.inv import minimize_device
    from vardyn_amd.problem import build_batch, build_problem
    case = syn.case_basin512()
    n_steps = args.window_steps or case.n_steps
    out = {"metric": f"4D-Var minimisation, {args.iterations} L-BFGS iterations, QG 512^2 "
                     f"{n_steps}-step window", "unit": "minimisations/s",
           "config": {"workload": case.name, "window_steps": n_steps, "iterations": args.iterations,
                      "start": "X = 0 (background)", "memory_pairs": 10}}
    legs = []
    if not args.no_plain:
        legs.append(("plain", 1, args.plain_dst_split))
    if args.batch > 1:
        legs.append(("batched", args.batch, args.dst_split))
    for label, nb, split in legs:
        t0 = time.time()
        if nb == 1:
            pb = build_problem(case, precision=args.precision, n_steps=n_steps, control="ic",
                               tape="full", dst_split=split)
        else:
            pb = build_batch(case, nb, precision=args.precision, n_steps=n_steps, tape="full",
                             dst_split=split)
        tiles = pb.plan.autotune(reps=5) if not args.no_autotune else None
        eng = pb.engine.capture()
        log(f"minimise {label}: {nb} problem(s) built and captured in {time.time() - t0:.1f}s")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        X, info = minimize_device(eng, np.zeros((nb, eng.nctrl)), maxiter=args.iterations)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        mem = info["members"]
        r = {"problems": nb, "dst_split": split, "gemm_tiles": tiles, "wall_s": round(wall, 2),
             "rounds": info["rounds"], "s_per_round": round(wall / info["rounds"], 4),
             "minimisations_per_s": round(nb / wall, 5),
             "gradients_per_s": round(sum(m["nfev"] for m in mem) / wall, 4),
             "iterations": [m["nit"] for m in mem], "evaluations": [m["nfev"] for m in mem],
             "status": sorted({m["status"] for m in mem}),
             "J_first_last": [[m["J"][0], m["J"][-1]] for m in mem]}
        log(f"minimise {label}: {r}")
        out[label] = r
        pb.engine = None
        del pb, eng, X
        release(None)
    if "batched" in out and "plain" in out:
        out["batched_over_plain"] = round(out["batched"]["minimisations_per_s"] /
                                          out["plain"]["minimisations_per_s"], 3)
    return out


def problem_set(args, steps=None, warmup=None, separate=True):
    """Independent problems of uneven grid sizes and windows (48 by default).

    Headline: the ragged engine -- one launch per RK stage / inversion / adjoint
    stage covers every problem still running, one HIP graph for the whole
    evaluation -- at the control X = 0.5 x truth of every problem.  Beside it
    (``separate``), the same problems as separate single-problem engines, one
    graph each, replayed concurrently on ``args.streams`` streams."""
    import torch
    from vardyn_amd import synthetic as syn
    from vardyn_amd.problem import build_problem
    from vardyn_amd.problemset import ProblemSet
    from vardyn_amd.ragged import RaggedFourDVar
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    cases = syn.config4_cases()[:args.problem_set_size]
    t0 = time.time()
    pbs = [build_problem(c, precision=args.precision, control="ic", with_engine=False,
                         dst_split=args.problem_set_split) for c in cases]
    eng = RaggedFourDVar(pbs, autotune=not args.no_autotune).capture()
    tuning = eng.tuning
    build_s = time.time() - t0
    log(f"problem set: {len(cases)} problems built and captured in {build_s:.1f}s, "
        f"tape {eng.tape_bytes / 2**30:.1f} GiB")
    X = torch.as_tensor(np.concatenate([control_point(pb.truth0) for pb in pbs]),
                        device=eng.dev)
    for _ in range(warmup):
        eng.evaluate(X)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        eng.evaluate(X)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t1
    ms = dt * 1e3 / steps
    gps = sum(c.ny * c.nx * c.n_steps for c in cases)
    value = len(cases) * steps / dt
    log(f"ragged engine: {ms:.1f} ms per evaluation of {len(cases)} problems")
    del eng, pbs
    release(None)
    out = {
        "metric": f"fwd+adjoint 4D-Var gradients/sec, {len(cases)} uneven regional QG problems",
        "value": round(value, 4), "unit": "gradients/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(ms, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "float64" else "f32",
        "data": "synthetic (seeded fields and nadir tracks per problem)",
        "config": {"workload": f"{len(cases)} problems, sides 96-384, 5-20 day windows, "
                               f"Ld 15-40 km, dt 900 s", "engine": "ragged (shared launches)",
                   "dst_split": args.problem_set_split, "control": "0.5 x truth initial SSH",
                   "grid_point_steps_per_evaluation": gps},
        "grid_point_steps_per_s": round(gps * steps / dt, 1),
        "per_problem_latency_ms": round(ms, 2), "build_s": round(build_s, 1),
        "group_tuning": tuning}
    if not separate:
        return out
    # the same problems as separate engines on several streams
    ps = ProblemSet(cases, precision=args.precision, tape="full", n_streams=args.streams,
                    autotune=False).capture()
    Xs = [torch.as_tensor(control_point(pb.truth0), device=ps.device) for pb in ps.problems]
    ps.evaluate(Xs)
    t2 = time.perf_counter()
    ps.evaluate(Xs)
    sep_ms = (time.perf_counter() - t2) * 1e3
    log(f"separate engines on {args.streams} streams: {sep_ms:.1f} ms per evaluation")
    del ps
    release(None)
    out["separate_engines"] = {"ms_per_step": round(sep_ms, 2), "streams": args.streams,
                               "value": round(len(cases) * 1e3 / sep_ms, 4)}
    out["ragged_over_separate"] = round(sep_ms / ms, 3)
    return out


def tiled(args):
    """The tiled QG model (``QgmWithTiles``, jqgm.py:1033-1258) on the 512 x 512
    grid of the headline workload with the reference's default tiling (32-point
    tiles overlapping by 16: 1024 tiles).  One unit is a ``step`` plus a
    ``step_adj`` call over a segment of ``--tiled-steps`` RK2 steps, each call
    one replayed HIP graph; ``value`` is domain grid points x steps per second
    of that pair.  Beside it: the same calls tile by tile (``engine="plans"``),
    the FP64 tiled oracle on one host thread, parity of both calls against the
    oracle, and the rooflines of the shared-launch inversion GEMMs and stage
    stencil (graph-timed with HIP events on the calls' stream)."""
    import numpy as np
    import torch
    from oracle.tiles import TiledQgmOracle
    from vardyn_amd import synthetic as syn
    from vardyn_amd.tiles import QgmWithTiles
    case = syn.case_basin512()
    ny, nx, ts, ov, n = case.ny, case.nx, args.tile_size, args.tile_overlap, args.tiled_steps
    kw = dict(dx=case.dx * np.ones((ny, nx)), dy=case.dx * np.ones((ny, nx)), dt=case.dt,
              SSH=np.zeros((ny, nx)), c=case.c, f=case.f0, time_scheme="rk2", tile_size=ts,
              tile_overlap=ov)
    t0 = time.time()
    m = QgmWithTiles(**kw, engine="group")
    tuning = m.autotune() if not args.no_autotune else None
    log(f"tiled: {m.ntile} tiles of {ts}x{ts} (overlap {ov}) built in {time.time() - t0:.1f}s, "
        f"group tuning {tuning}")
    h0n = syn.gaussian_random_field(ny, nx, case.dx, 100e3, 0.1, seed=11)
    rng = np.random.default_rng(12)
    hbn = 0.01 * rng.standard_normal((ny, nx))
    adn = rng.standard_normal((ny, nx))
    dev = m.device
    h0, hb, ad = (torch.as_tensor(a, device=dev) for a in (h0n, hbn, adn))

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            out = fn()
        g.replay()
        torch.cuda.synchronize()
        s_ = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s_)
        for _ in range(reps):
            g.replay()
        e1.record(s_)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps, out, g

    ms_f, _, gf = timed(lambda: m.step(h0, hb, n), max(3, args.steps))
    ms_a, _, ga = timed(lambda: m.step_adj(ad, h0, hb, n), max(3, args.steps))
    del gf, ga
    gps = ny * nx * n
    value = gps / ((ms_f + ms_a) * 1e-3)
    log(f"tiled, shared launches: step {ms_f:.2f} ms, step_adj {ms_a:.2f} ms per {n}-step call")
    # rooflines of the shared launches: inversion GEMMs (FP64 MFMA) and stage stencil
    grp, T = m.group, m.ntile
    F, H, Q, Q2 = (m._stack().normal_() for _ in range(4))
    qb = torch.zeros_like(F)
    reps = 50

    def inv():
        for _ in range(reps):
            grp.inversion(T, ptr_(F), ptr_(H), Fsub=ptr_(qb), hb=ptr_(qb))

    def stage():
        for _ in range(reps):
            grp.stage(T, ptr_(H), ptr_(Q), ptr_(Q), None, None, 0, 0, case.dt, ptr_(Q2))

    ms_inv = timed(inv, 3)[0] / reps
    ms_st = timed(stage, 3)[0] / reps
    my = ts - 2
    st_bytes = T * ts * ts * (4 * 8 + 1)
    st_gbs = st_bytes / (ms_st * 1e-3) / 1e9
    # tiles with interiors of at most 45 points invert in one fused launch
    # (k_inv_small: the four sine products of each tile in LDS, FP64 FMA);
    # larger ones in four GEMM launches (k_gemm_rg, FP64 MFMA)
    fused = my <= 45
    nlaunch = 1 if fused else 4
    flops_launch = T * 4 * my * my * (2 * my) // nlaunch
    inv_tf = flops_launch / (ms_inv / nlaunch * 1e-3) / 1e12
    roof = {"bound": "fp64" if fused else "mfma",
            "kernel": "k_inv_small (all four sine products of each tile in one workgroup, one "
                      "launch per inversion)" if fused else
                      "k_gemm_rg (sine-matrix GEMMs of all tiles, FP64 MFMA, 4 launches per "
                      "inversion)",
            "achieved": round(inv_tf, 3), "peak": MFMA_PEAK["f64"], "unit": "TFLOP/s",
            "frac": round(inv_tf / MFMA_PEAK["f64"], 4), "traffic": None,
            "avg_launch_us": round(ms_inv / nlaunch * 1e3, 3),
            "algorithmic_flops_per_launch": flops_launch}
    st_roof = {"bound": "hbm", "kernel": "k_stage_rg (stage stencil of all tiles)",
               "achieved": round(st_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(st_gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": st_bytes,
               "avg_launch_us": round(ms_st * 1e3, 3),
               "note": "working set of one launch (33 MB) is cache-resident"}
    del F, H, Q, Q2, qb
    # parity against the oracle over two steps (the oracle takes seconds per step)
    orc = TiledQgmOracle(case.dx, case.dx, case.dt, (ny, nx), c=case.c, f=case.f0,
                         SSH=np.zeros((ny, nx)), time_scheme="rk2", tile_size=ts, tile_overlap=ov)
    k = 2
    t1 = time.perf_counter()
    ref_h = orc.step(h0n, hbn, k)
    ref_a = orc.step_adj(adn, h0n, hbn, k)
    cpu_s = time.perf_counter() - t1
    gpu_h = m.step(h0n, hbn, k)
    gpu_a = m.step_adj(adn, h0n, hbn, k)
    err_h = float(np.linalg.norm(gpu_h - ref_h) / np.linalg.norm(ref_h))
    err_a = float(np.linalg.norm(gpu_a - ref_a) / np.linalg.norm(ref_a))
    log(f"tiled parity vs oracle ({k} steps): step {err_h:.2e}, step_adj {err_a:.2e}; "
        f"oracle {cpu_s:.1f} s")
    # the same calls tile by tile
    one = QgmWithTiles(**kw, engine="plans")
    k1 = 2
    one.step(h0, hb, k1)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    one.step(h0, hb, k1)
    one.step_adj(ad, h0, hb, k1)
    torch.cuda.synchronize()
    ms_one = (time.perf_counter() - t2) * 1e3
    del one
    release(None)
    return {
        "metric": f"tiled QG (QgmWithTiles) step + step_adj grid-point-steps/sec, {ny}x{nx}, "
                  f"{ts}-point tiles overlapping by {ov}",
        "value": round(value, 1), "unit": "grid-point-steps/s", "n_gpus": 1,
        "steps": max(3, args.steps), "warmup": 1, "ms_per_step": round(ms_f + ms_a, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Gaussian random SSH, random boundary field and adjoint input)",
        "config": {"workload": f"tiled_qg{ny}", "grid": [ny, nx], "dx_m": case.dx,
                   "dt_s": case.dt, "time_scheme": "rk2", "tile_size": ts, "tile_overlap": ov,
                   "tiles": m.ntile, "steps_per_call": n, "engine": "ragged group (shared launches)",
                   "hip_graph": True},
        "ms_step": round(ms_f, 3), "ms_step_adj": round(ms_a, 3), "group_tuning": tuning,
        "roofline": roof, "stencil_roofline": st_roof,
        "tile_by_tile": {"value": round(ny * nx * k1 / (ms_one * 1e-3), 1),
                         "ms_step_plus_adj": round(ms_one, 2), "steps_per_call": k1,
                         "launches": "eager"},
        "group_over_tile_by_tile": round(value / (ny * nx * k1 / (ms_one * 1e-3)), 2),
        "parity_vs_oracle": {"steps": k, "step_rel_err": err_h, "step_adj_rel_err": err_a},
        "cpu_baseline": {"value": round(ny * nx * k / cpu_s, 1), "unit": "grid-point-steps/s",
                         "cores": 1, "kind": "port",
                         "sample": f"FP64 tiled oracle, step + step_adj over {k} steps of the "
                                   f"same tiling ({cpu_s:.1f} s)"}}


def basis(args):
    """The balanced-motion wavelet basis (``Basis_bm``, basis.py:472-936) with
    the reference's default settings (BASIS_BM, config_default.py:804-860) on
    a regular 512 x 512 lon/lat grid of 0.045 degrees (about 5 km), at the
    54 basis times of a 30-day window with a basis time every 13.5 h (every
    checkpoint of the QG workload).  One unit is ``operg`` of all 54 times
    plus the summed ``operg_transpose`` of all 54 (the forward and reverse
    use of the reduced control in one gradient evaluation); ``value`` is grid
    points x basis times per second of that pair.  Beside it: the FP64 oracle
    (the reference's sparse-matrix algorithm) on one host thread over a
    sample of times, and parity of the GPU fields and coefficients against
    it."""
    import numpy as np
    import torch
    from types import SimpleNamespace
    from oracle.basis_bm import BasisBmOracle
    from vardyn_amd import _lib
    from vardyn_amd.basis_bm import Basis_bm
    from vardyn_amd.qg import current_stream_ptr, ptr
    ny = nx = 512
    d = 0.045
    lon, lat = np.meshgrid(-60 + d * np.arange(nx), 35 + d * np.arange(ny))
    times = np.arange(54) * 0.5625
    t0 = time.time()
    b = Basis_bm(SimpleNamespace(lon=lon, lat=lat, mask=None))
    _, Q = b.set_basis(times, return_q=True)
    setup_s = time.time() - t0
    log(f"basis: {b.nbasis} coefficients, {b.nf} frequencies x {b.ntheta} directions, "
        f"{b.nc} centres, {b.nnz_space:.3e} space-matrix entries; set-up {setup_s:.1f}s")
    T = times.size
    dev = b.device
    rng = np.random.default_rng(21)
    X = torch.as_tensor(rng.standard_normal(b.nbasis) * np.sqrt(Q), device=dev)
    gt = torch.as_tensor(b.time_factors(times), device=dev)
    ws = b._workspace(T)
    out = torch.empty((T, ny, nx), dtype=torch.float64, device=dev)
    ad = torch.as_tensor(rng.standard_normal((T, ny, nx)), device=dev)
    adX = torch.zeros(b.nbasis, dtype=torch.float64, device=dev)
    s_ = current_stream_ptr(dev)

    def fwd():
        _lib.call("vd_bm_operg", b._h, T, ptr(gt), ptr(X), ptr(out), ptr(ws), s_)

    def adj():
        _lib.call("vd_bm_operg_adj", b._h, T, ptr(gt), ptr(ad), ptr(adX), ptr(ws), s_)

    def clock(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    ms_f, ms_a = clock(fwd), clock(adj)
    value = ny * nx * T / ((ms_f + ms_a) * 1e-3)
    flops = 2 * b.nnz_space * T  # one multiply-add per space-matrix entry and time
    tf = 2 * flops / ((ms_f + ms_a) * 1e-3) / 1e12
    log(f"basis: operg {ms_f:.2f} ms, operg_transpose {ms_a:.2f} ms for {T} times")
    # oracle on a sample of times (its space matrix is built once, untimed)
    orc = BasisBmOracle(lon, lat, None)
    t1 = time.time()
    orc.set_basis(times)
    orc_setup = time.time() - t1
    sample = [0, T // 2]
    Xh = X.cpu().numpy()
    t2 = time.perf_counter()
    ref_f = [orc.operg(times[k], Xh) for k in sample]
    ref_a = sum(orc.operg_transpose(times[k], ad[k].cpu().numpy()) for k in sample)
    cpu_s = time.perf_counter() - t2
    got_f = b.operg_batch(times[sample], X).cpu().numpy()
    got_a = b.operg_transpose_batch(times[sample], ad[sample]).cpu().numpy()
    err_f = max(float(np.max(np.abs(got_f[i] - ref_f[i])) / np.max(np.abs(ref_f[i])))
                for i in range(len(sample)))
    err_a = float(np.max(np.abs(got_a - ref_a)) / np.max(np.abs(ref_a)))
    log(f"basis parity vs oracle ({len(sample)} times): operg {err_f:.2e}, "
        f"operg_transpose {err_a:.2e}; oracle {cpu_s:.1f} s (+{orc_setup:.1f} s set-up)")
    del orc
    return {
        "metric": "balanced-motion basis operg + operg_transpose grid-point-times/sec, 512x512, "
                  "54 basis times",
        "value": round(value, 1), "unit": "grid-point-times/s", "n_gpus": 1, "steps": 5,
        "warmup": 1, "ms_per_step": round(ms_f + ms_a, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded coefficients scaled by sqrt(Q), seeded adjoint fields)",
        "config": {"workload": "bm_basis512", "grid": [ny, nx], "dlon_dlat_deg": d,
                   "times": int(T), "window_days": 30, "lmin_km": 80.0, "lmax_km": 970.0,
                   "nbasis": int(b.nbasis), "frequencies": int(b.nf), "directions": int(b.ntheta),
                   "centres": int(b.nc), "space_matrix_entries": int(b.nnz_space)},
        "ms_operg": round(ms_f, 3), "ms_operg_transpose": round(ms_a, 3),
        "roofline": {"bound": "fp64 (matrix-free: space values from separable tables)",
                     "kernel": "k_bm_synth + k_bm_proj + k_bm_proj_sum", "achieved": round(tf, 3),
                     "peak": MFMA_PEAK["f64"], "unit": "TFLOP/s",
                     "frac": round(tf / MFMA_PEAK["f64"], 4), "traffic": None,
                     "algorithmic_flops": 2 * flops,
                     "note": "2 flops per space-matrix entry, time and direction of use; the "
                             "angle-addition and taper products that rebuild each entry "
                             "from the separable tables are not counted"},
        "parity_vs_oracle": {"times": len(sample), "operg_rel_err": err_f,
                             "operg_transpose_rel_err": err_a},
        "setup_s": round(setup_s, 2),
        "cpu_baseline": {"value": round(ny * nx * len(sample) / cpu_s, 1),
                         "unit": "grid-point-times/s", "cores": 1, "kind": "port",
                         "sample": f"FP64 oracle (sparse space matrix, as basis.py) operg + "
                                   f"operg_transpose at {len(sample)} of the {T} times "
                                   f"({cpu_s:.1f} s; matrix set-up {orc_setup:.1f} s not timed)"}}


def sw_case(nx=256, ny=256):
    """The 3-layer shallow-water configuration: 256 x 256 cells of 8 km, beta
    plane f = 1e-4 + 1.6e-11 y, layers 500 / 1000 / 3000 m, reduced gravities
    0.02 (surface) / 0.02 / 0.01 m/s^2, dt = 120 s, no-slip walls.  The
    initial state: a seeded smooth thickness anomaly in the top layer (RMS
    5 cm, 50 km Gaussian smoothing) and the geostrophic velocities of its
    pressure, zero in the layers below."""
    import numpy as np
    from scipy.ndimage import gaussian_filter
    dx = dy = 8e3
    y = (np.arange(ny + 1) - ny / 2) * dy
    f = np.broadcast_to(1e-4 + 1.6e-11 * y, (nx + 1, ny + 1)).copy()
    H = np.array([500.0, 1000.0, 3000.0])
    gp = np.array([0.02, 0.02, 0.01])
    param = dict(nx=nx, ny=ny, nl=3, dx=dx, dy=dy, H=H.reshape(3, 1, 1), g_prime=gp, f=f, dt=120.0,
                 slip_coef=0.0, bottom_drag_coef=0.0)
    rng = np.random.default_rng(2024)
    eta = gaussian_filter(rng.standard_normal((nx, ny)), 50e3 / dx)
    eta *= 0.05 / eta.std()
    h = np.zeros((3, nx, ny)); h[0] = eta
    # geostrophic velocities of the top-layer pressure g'_0 eta: f u = -g' deta/dy, f v = g' deta/dx
    f_u = 0.5 * (f[:, 1:] + f[:, :-1]); f_v = 0.5 * (f[1:, :] + f[:-1, :])
    u = np.zeros((3, nx + 1, ny)); v = np.zeros((3, nx, ny + 1))
    ey = np.gradient(eta, dy, axis=1); ex = np.gradient(eta, dx, axis=0)
    u[0, 1:nx, :] = -gp[0] * 0.5 * (ey[1:] + ey[:-1]) / f_u[1:nx]
    v[0, :, 1:ny] = gp[0] * 0.5 * (ex[:, 1:] + ex[:, :-1]) / f_v[:, 1:ny]
    return param, (u, v, h)


def sw_model(args):
    """The multilayer shallow-water model (sw.py:67-937) on the 256 x 256,
    3-layer configuration of ``sw_case``: forward RK3-SSP steps per second,
    and the gradient of the final-state cost J = 1/2 |x(T)|^2 over the window
    (``step`` then ``step_adj``: a forward sweep keeping sqrt(n) segment
    states, each segment recomputed, three adjoint RHS evaluations per step).
    Beside them the FP64 oracle (NumPy, one host thread) over a few steps of
    the same grid, and parity against it."""
    import numpy as np
    import torch
    from oracle.sw import SwOracle
    from vardyn_amd.sw import SW
    param, x0 = sw_case()
    model = SW(param)
    dev = torch.device("cuda")
    X = [torch.as_tensor(a, device=dev) for a in x0]
    n_steps = int(round(args.sw_days * 86400 / param["dt"]))
    st = torch.cuda.current_stream()
    ev = lambda: torch.cuda.Event(enable_timing=True)
    # forward
    model.step(*X, nstep=8)
    torch.cuda.synchronize()
    e0, e1 = ev(), ev()
    e0.record(st)
    xT = model.step(*X, nstep=n_steps)
    e1.record(st)
    torch.cuda.synchronize()
    fwd_s = e0.elapsed_time(e1) / 1e3
    # gradient of J = 1/2 |x(T)|^2 with respect to x0
    e2, e3 = ev(), ev()
    e2.record(st)
    xT = model.step(*X, nstep=n_steps)
    (g,) = model.step_adj(*X, *xT, nstep=n_steps)
    e3.record(st)
    torch.cuda.synchronize()
    grad_s = e2.elapsed_time(e3) / 1e3
    finite = all(bool(torch.isfinite(t).all()) for t in list(xT) + list(g[:3]))
    # oracle on the host: a few steps, timed and compared
    k = 6
    orc = SwOracle(param["nx"], param["ny"], param["dx"], param["dy"], param["H"].reshape(-1), param["g_prime"],
                   param["f"], param["dt"], slip_coef=param["slip_coef"], bottom_drag_coef=0.0)
    t0 = time.time()
    ref = orc.step(*x0, nstep=k)
    cpu_s = time.time() - t0
    got = model.step(*X, nstep=k)
    torch.cuda.synchronize()
    err = max(float(np.linalg.norm(a.cpu().numpy() - b) / np.linalg.norm(b)) for a, b in zip(got, ref))
    pts = param["nx"] * param["ny"] * 3
    return {
        "metric": "sw_gradients_per_sec",
        "value": round(1.0 / grad_s, 4),
        "unit": "gradients/s",
        "workload": f"3-layer shallow water {param['nx']}x{param['ny']}, dx 8 km, dt 120 s, "
                    f"{args.sw_days:g}-day window ({n_steps} steps), FP64",
        "gradient_s": round(grad_s, 3),
        "forward_steps_per_s": round(n_steps / fwd_s, 1),
        "forward_window_s": round(fwd_s, 3),
        "cell_steps_per_s": round(pts * n_steps / fwd_s, 0),
        "finite": finite,
        "oracle_rel_err": float(f"{err:.3e}"),
        "oracle_steps": k,
        "cpu_baseline": {"value": round(k / cpu_s, 3), "unit": "forward steps/s", "cores": 1, "kind": "port",
                         "sample": f"{k} forward steps of the same grid with the NumPy oracle"},
        "gpu_over_cpu_forward": round((n_steps / fwd_s) / (k / cpu_s), 1),
    }


def merge(args):
    """The merge of subwindow analyses onto the target grid
    (compute_weights_map / merge_output_date, run_assimilation.py:351-607):
    a 512 x 512 target grid of 0.045 degrees, 3 x 3 overlapping subwindows of
    200 x 200 points on grids of 0.05 degrees (offset from the target's, so every
    value is a true bilinear interpolation), 64 output dates merged per call.
    ``value`` is target grid points x dates per second; beside it the FP64
    oracle (scipy, as the reference) on one host thread over 2 dates, and
    parity against it."""
    import numpy as np
    import torch
    from oracle.merge import merge_date, weights_map
    from vardyn_amd.merge import SubwindowMerge
    ny = nx = 512
    d = 0.045
    lon, lat = np.meshgrid(-10.0 + d * np.arange(nx), 30.0 + d * np.arange(ny))
    subs = []
    for a in range(3):
        for c in range(3):
            lo0, la0 = -10.3 + 6.7 * c, 29.8 + 6.7 * a  # 9.95 degrees wide, 3.25 overlapping
            subs.append((lo0 + 0.05 * np.arange(200), la0 + 0.05 * np.arange(200)))
    T = 64
    rng = np.random.default_rng(11)
    dev = torch.device("cuda")
    t0 = time.time()
    m = SubwindowMerge(lon, lat, subs)
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    fields = [torch.as_tensor(rng.standard_normal((T, 200, 200)), device=dev) for _ in subs]
    packed, _ = m.pack(fields)
    out = torch.empty((T, ny, nx), dtype=torch.float64, device=dev)
    # the timed calls alternate between two input/output sets (636 MB together,
    # above the 256 MB Infinity Cache), so each call streams its data from HBM
    sets = [(packed, out), (packed.clone(), torch.empty_like(out))]
    st = torch.cuda.current_stream()
    for pk, o in sets:
        m.merge_packed(pk, T, o)
    torch.cuda.synchronize()
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for r in range(reps):
        pk, o = sets[r % 2]
        m.merge_packed(pk, T, o)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    launches = 1  # all dates in one launch, 8 per workgroup row
    w = m.weights_space.cpu().numpy()
    valid = int(np.isfinite(w).sum())
    np_ = ny * nx
    # per launch: each subwindow value read once and each output written once per
    # date, and per 8 dates the tables (wsum; cell per pair; fractions and weight
    # per valid pair)
    per_launch = 8 * T * (m.sub_points + np_) + (T // 8) * (8 * np_ + 8 * m.nsub * np_ + 24 * valid)
    gbs = per_launch / (ms * 1e-3 / launches) / 1e9
    traffic, tsrc = None, None
    rec = os.path.join(ROOT, "profiles", f"merge_traffic_merge{ny}_s{m.nsub}_t{T}.json")
    if os.path.exists(rec):
        with open(rec) as fh_:
            traffic = json.load(fh_)["traffic_bytes_per_launch"]
        tsrc = os.path.relpath(rec, ROOT)
    log(f"merge: {m.nsub} subwindows of 200x200 onto {ny}x{nx}, {T} dates: {ms:.3f} ms per call "
        f"({launches} launches); set-up {setup_s:.2f}s")
    wr, wsr = weights_map(lon, lat, subs)
    sample = [0, T - 1]
    fh = [f.cpu().numpy() for f in fields]
    t1 = time.perf_counter()
    refs = [merge_date(lon, lat, subs, [f[k] for f in fh], wr, wsr) for k in sample]
    cpu_s = time.perf_counter() - t1
    got = out.cpu().numpy()
    err = 0.0
    for k, ref in zip(sample, refs):
        ok = np.isfinite(ref)
        assert np.array_equal(ok, np.isfinite(got[k]))
        err = max(err, float(np.max(np.abs(got[k][ok] - ref[ok])) / np.max(np.abs(ref[ok]))))
    log(f"merge parity vs oracle ({len(sample)} dates): {err:.2e}; oracle {cpu_s:.2f} s")
    return {
        "metric": "subwindow merge target-grid-point-dates/sec, 9 subwindows onto 512x512, 64 dates",
        "value": round(np_ * T / (ms * 1e-3), 1), "unit": "grid-point-dates/s", "n_gpus": 1,
        "steps": reps, "warmup": 1, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded subwindow fields)",
        "config": {"workload": "merge512_s9_t64", "grid": [ny, nx], "subwindows": m.nsub,
                   "subwindow_grid": [200, 200], "dates": T, "sub_points": m.sub_points},
        "roofline": {"bound": "hbm", "kernel": "k_merge_apply", "achieved": round(gbs, 1),
                     "peak": 8000.0, "unit": "GB/s", "frac": round(gbs / 8000.0, 4), "traffic": traffic,
                     "traffic_source": tsrc,
                     "bytes_per_launch": per_launch, "avg_launch_us": round(ms * 1e3 / launches, 2),
                     "note": "algorithmic bytes: every subwindow value read once and every target "
                             "value written once per date, plus the tables once per 8 dates"},
        "parity_vs_oracle": {"dates": len(sample), "rel_err": err},
        "setup_s": round(setup_s, 3),
        "cpu_baseline": {"value": round(np_ * len(sample) / cpu_s, 1), "unit": "grid-point-dates/s",
                         "cores": 1, "kind": "port",
                         "sample": f"FP64 oracle (scipy RegularGridInterpolator, as the reference) "
                                   f"merge of {len(sample)} dates ({cpu_s:.2f} s; weights untimed)"},
    }


def obs_build(args):
    """Construction of the L3 observation operator (_sparse_op,
    obsop.py:417-460) for the observations of the 512 x 512 30-day
    configuration (60 nadir tracks at 6 km, one 120 km swath per day at 2 km
    pixels, 2880 steps): ``value`` is observations per second of the device
    neighbour search (``vd_obs_grid_query``, Npix = 4) over every observation
    of the window in one launch, the grid binned once (timed apart).  Beside it
    the end-to-end ``process_obs`` (device search plus the host's CSR
    assembly and upload) and the FP64 oracle, which builds a scipy KDTree per
    observation date and queries it on one host thread, as the reference
    does, over every date; tables compared on every date."""
    import numpy as np
    import torch
    from oracle.obsop import sparse_op
    from vardyn_amd import synthetic as syn
    from vardyn_amd.obsop import ObsGrid, Obsop_interp_l3, geo2cart, interp_weights
    case = syn.case_basin512()
    grid = syn.make_grid(case.ny, case.nx, case.dx)
    obs = syn.make_observations(case, grid)
    T = np.arange(case.n_steps + 1) * case.dt
    dict_obs = {float(T[s]): (e["lon"], e["lat"], np.zeros(len(e["lon"])), np.ones(len(e["lon"])))
                for s, e in obs.steps.items()}
    dates = list(dict_obs)
    n_obs = sum(len(v[0]) for v in dict_obs.values())
    npix = 4
    cc = geo2cart(np.column_stack((grid.lon.ravel(), grid.lat.ravel())))
    dmax = npix * np.mean(np.sqrt(grid.DX ** 2 + grid.DY ** 2)) * 1e-3 * np.sqrt(2) / 2
    dev = torch.device("cuda")
    t0 = time.perf_counter()
    og = ObsGrid(cc, dmax, dev)
    torch.cuda.synchronize()
    grid_s = time.perf_counter() - t0
    pts = np.concatenate([geo2cart(np.column_stack((v[0], v[1]))) for v in dict_obs.values()])
    xyz = torch.as_tensor(pts, device=dev)
    for _ in range(max(1, args.warmup)):
        og.query(xyz, npix)
    torch.cuda.synchronize()
    reps = max(3, args.steps)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        col, w, dist = og.query(xyz, npix)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    found = int((col >= 0).sum())
    t0 = time.perf_counter()
    H = Obsop_interp_l3(grid.lon, grid.lat, grid.DX, grid.DY, T, case.dt, dict_obs, Npix=npix,
                        device=dev).process_obs()
    torch.cuda.synchronize()
    process_s = time.perf_counter() - t0
    # oracle: a KDTree per date (the reference rebuilds it in _sparse_op), on
    # every date of the window (about 10 s on one core)
    sample = dates
    t0 = time.perf_counter()
    ref = [sparse_op(cc, dict_obs[d][0], dict_obs[d][1], npix, dmax) for d in sample]
    cpu_s = time.perf_counter() - t0
    n_sample = sum(len(dict_obs[d][0]) for d in sample)
    werr, same = 0.0, True
    for d, (w2, r2, c2) in zip(sample, ref):
        w1, r1, c1 = interp_weights(og, dict_obs[d][0], dict_obs[d][1], npix)
        same = same and np.array_equal(r1, r2) and np.array_equal(c1, c2)
        if w2.size:
            werr = max(werr, float(np.max(np.abs(w1 - w2) / np.abs(w2))))
    log(f"obs build: {n_obs} observations on {len(dates)} dates; device search {ms:.3f} ms "
        f"({found} neighbours), grid binning {grid_s:.2f} s, process_obs {process_s:.2f} s; "
        f"oracle {cpu_s:.2f} s for {len(sample)} dates ({n_sample} obs); indices equal {same}, "
        f"weights {werr:.1e}")
    per_launch = n_obs * (24 + npix * (4 + 8 + 8))
    gbs = per_launch / (ms * 1e-3) / 1e9
    return {
        "metric": "observation-operator construction: observations per second (Npix = 4)",
        "value": round(n_obs / (ms * 1e-3), 1), "unit": "obs/s", "n_gpus": 1, "steps": reps,
        "warmup": max(1, args.warmup), "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded nadir tracks and swaths of the 512x512 30-day configuration)",
        "config": {"workload": "obsop_build_qg512_30d", "grid": [case.ny, case.nx], "observations": n_obs,
                   "dates": len(dates), "npix": npix, "dmax_km": round(float(dmax), 4)},
        "roofline": {"bound": "hbm", "kernel": "k_obs_knn", "achieved": round(gbs, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(gbs / 8000.0, 5), "traffic": None,
                     "bytes_per_launch": per_launch, "avg_launch_us": round(ms * 1e3, 2),
                     "note": "algorithmic bytes: each observation's coordinates in and its npix (col, w, "
                             "dist) out; the search reads the binned grid from cache (binary searches "
                             "over the sorted cube keys, then the candidates), so it is latency-bound"},
        "parity_vs_oracle": {"dates": len(sample), "indices_equal": bool(same), "weights_rel_err": werr},
        "grid_binning_s": round(grid_s, 3), "process_obs_s": round(process_s, 3),
        "cpu_baseline": {"value": round(n_sample / cpu_s, 1), "unit": "obs/s", "cores": 1, "kind": "port",
                         "sample": f"FP64 oracle sparse_op (scipy KDTree built per date and queried, as "
                                   f"the reference) on {len(sample)} of the {len(dates)} dates "
                                   f"({n_sample} observations, {cpu_s:.2f} s)"},
    }


def ptr_(t):
    from vardyn_amd.qg import ptr
    return ptr(t)


def release(pb):
    import torch
    if pb is not None and pb.engine is not None:
        pb.engine._graph = None
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def sw_oracle_cost(pb, prec=False):
    """The FP64 oracle cost (oracle/sw4dvar.py) of a shallow-water problem."""
    from oracle.basis_wavelet3d import BasisWavelet3dOracle
    from oracle.obsop import ObsopOracle
    from oracle.sw import SwOracle
    from oracle.sw4dvar import SwCostOracle
    from vardyn_amd.sw_problem import basis_kwargs
    c, m = pb.case, pb.model
    sw = SwOracle(m.nx, m.ny, m.dx, m.dy, m.H, m.g_prime, m.f, m.dt)
    lon, lat = pb.grid.lon.T, pb.grid.lat.T
    H = ObsopOracle(lon, lat, np.full(lon.shape, c.dx), np.full(lon.shape, c.dx), pb.T, c.dt, pb.dict_obs, Npix=4)
    basis = BasisWavelet3dOracle(lon, lat, **basis_kwargs(c))
    basis.set_basis(pb.time_basis)
    return SwCostOracle(sw, H, basis, pb.checkpoints, pb.T, c.checkpoint_steps, pb.x0, sigma_b=pb.Q, prec=prec)


def sw_parity(pb, seed=3):
    """J and one directional derivative of the GPU engine against the oracle
    cost (complex step) at X = 0.5 x truth coefficients."""
    import torch
    orc = sw_oracle_cost(pb)
    X0 = 0.5 * pb.X_true
    J, G = pb.engine.run(torch.as_tensor(X0, device=pb.model.device))
    J, G = float(J[0]), G.cpu().numpy()
    t0 = time.time()
    Jo = orc.cost(X0)
    cost_s = time.time() - t0
    d = np.random.default_rng(seed).standard_normal(G.size)
    dd = orc.directional_derivative(X0, d)
    return {"steps": int(pb.T.size - 1), "grid": [pb.case.nx, pb.case.ny], "J_rel_err": abs(J - Jo) / abs(Jo),
            "dir_deriv_rel_err": abs(G @ d - dd) / abs(dd), "oracle_cost_s": round(cost_s, 2)}


def w3_roofline(basis, prep, X, reps=20):
    """Average durations of the basis synthesis (all prepared times) and of its
    transpose, HIP events on their stream.  Both stream the tile blocks of the
    space matrix once per pass of up to 16 times: algorithmic bytes = 256
    doubles per stored tile entry and pass (plus the coefficients, fields and
    partials, small beside it)."""
    import torch
    st = torch.cuda.current_stream()
    out = torch.empty((prep.T, basis.ny, basis.nx), dtype=torch.float64, device=basis.device)
    adX = torch.zeros(basis.nbasis, dtype=torch.float64, device=basis.device)
    res = {}
    for name, fn in (("synthesis", lambda: basis.operg_prepared(prep, X, out)),
                     ("transpose", lambda: basis.operg_transpose_prepared(prep, out, adX))):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        res[name + "_us"] = round(e0.elapsed_time(e1) / reps * 1e3, 2)
    passes = -(-prep.T // 16)
    bytes_ = basis.nentries * 256 * 8 * passes
    flops = 2.0 * 16 * basis.nentries * 256 * passes
    us = res["synthesis_us"]
    res.update({"bound": "hbm", "achieved": round(bytes_ / (us * 1e-6) / 1e9, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(bytes_ / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 3),
                "traffic": None, "block_bytes": bytes_, "times": prep.T,
                "mfma_f64_tflops_issued": round(flops / (us * 1e-6) / 1e12, 2),
                "space_matrix_nnz": basis.nnz_space, "stored_entries": basis.nentries * 256})
    return res


def sw4dvar(args):
    """The 3-layer shallow-water 4D-Var of the wavelet-forcing configuration:
    one evaluation = forward sweep with the misfits of 30 nadir tracks, reverse
    sweep with the hand-written adjoint (segment recomputation inside
    vd_sw_step_adj), the basis synthesis of every basis time before and its
    transpose after; at X = 0.5 x the truth's coefficients (a forced, flowing
    state).  Beside it: the basis products' roofline, parity of J and of a
    directional derivative against the oracle cost on two short windows (the
    40 x 36 test grid and the 256 x 256 grid over 30 steps), and the oracle's
    forward speed on one host thread."""
    import torch
    from vardyn_amd.sw_problem import SwCase, build_sw_problem, small_case
    case = SwCase(days=args.sw_days)
    t0 = time.time()
    pb = build_sw_problem(case)
    build_s = time.time() - t0
    log(f"sw4dvar: {case.nx}x{case.ny}x{len(case.H)}, {case.n_steps} steps, {pb.basis.nbasis} coefficients, "
        f"{pb.H.n_total} obs, {len(pb.checkpoints) - 1} segments, built in {build_s:.1f}s")
    eng = pb.engine
    X0 = torch.as_tensor(0.5 * pb.X_true, device=pb.model.device)
    for _ in range(max(1, args.warmup)):
        J, G = eng.run(X0)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    n = max(1, min(args.steps, 3))
    e0.record(st)
    for _ in range(n):
        J, G = eng.run(X0)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / n
    log(f"sw4dvar: {ms:.1f} ms per gradient")
    roof = w3_roofline(pb.basis, eng.bprep, eng.X)
    log(f"basis roofline: {roof}")
    # forward-only speed (the same window, the engine's model)
    xs = [torch.as_tensor(a, device=pb.model.device) for a in pb.x0]
    e0.record(st)
    pb.model.step(*xs, nstep=case.n_steps)
    e1.record(st)
    torch.cuda.synchronize()
    fwd_s = e0.elapsed_time(e1) / 1e3
    out = {
        "metric": "sw_4dvar_gradients_per_sec", "value": round(1e3 / ms, 4), "unit": "gradients/s", "n_gpus": 1,
        "steps": n, "warmup": args.warmup, "ms_per_step": round(ms, 1), "higher_is_better": True,
        "dtype": "f64", "data": "synthetic (seeded eddy initial state, forced truth run, nadir tracks)",
        "config": {"workload": f"{case.name}: 3-layer shallow water {case.nx}x{case.ny}, dx 8 km, beta plane, "
                               f"dt {case.dt:g} s, {case.days:g}-day window ({case.n_steps} steps), wavelet forcing "
                               f"control ({pb.basis.nbasis} coefficients, {pb.basis.nfs} scales "
                               f"{1 / pb.basis.fs.max():g}-{1 / pb.basis.fs.min():g} km, {pb.basis.ntheta} "
                               f"directions), {case.n_tracks} nadir tracks",
                   "obs": pb.H.n_total, "segments": len(pb.checkpoints) - 1, "basis_times": int(eng.bprep.T),
                   "evaluated_at": "X = 0.5 x truth coefficients", "J": float(J[0]),
                   "grad_norm": float(torch.linalg.vector_norm(G))},
        "cell_steps_per_s": round(case.nx * case.ny * len(case.H) * case.n_steps * 1e3 / ms, 0),
        "forward_window_s": round(fwd_s, 3), "build_s": round(build_s, 1),
        "roofline": roof,
    }
    del pb, eng
    release(None)
    small = build_sw_problem(small_case())
    out["parity_small"] = sw_parity(small)
    log(f"parity (40x36, 240 steps): {out['parity_small']}")
    del small
    short = build_sw_problem(SwCase(days=30 * 120.0 / 86400, n_tracks=6))
    par = sw_parity(short)
    out["parity_256_30steps"] = par
    log(f"parity (256x256, 30 steps): {par}")
    del short
    release(None)
    if not args.no_cpu_baseline:
        # the oracle's forward steps on one host thread (it has no adjoint: its
        # gradient is taken by complex steps, one direction at a time)
        steps_s = 30 / par["oracle_cost_s"]
        grad_est = steps_s / (case.n_steps * 3.0)  # forward + recomputation + adjoint ~ 3 forwards
        out["cpu_baseline"] = {"value": round(steps_s, 2), "unit": "forward steps/s (256x256x3)", "cores": 1,
                               "kind": "port", "sample": "oracle cost (forward sweep, misfits, basis) over the "
                                                         "30-step 256x256 window",
                               "gradients_per_s_if_adjoint_cost_3_forwards": float(f"{grad_est:.3e}")}
        out["gpu_over_cpu_forward"] = round((case.n_steps / fwd_s) / steps_s, 1)
    return out


def main():
    args = parse()
    if args.precision != "float64":
        args.dst_split = args.plain_dst_split = 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(independent problems shard across GPUs with no communication)", file=sys.stderr)
        sys.exit(2)
    import torch
    from vardyn_amd import synthetic as syn
    from vardyn_amd.problem import build_batch, build_problem

    case = syn.case_basin512()
    if args.problem_set:
        print(json.dumps(problem_set(args)), flush=True)
        return
    if args.tiled:
        print(json.dumps(tiled(args)), flush=True)
        return
    if args.oracle_full:
        print(json.dumps(oracle_full(args)), flush=True)
        return
    if args.minimize:
        print(json.dumps(minimisation(args)), flush=True)
        return
    if args.basis:
        print(json.dumps(basis(args)), flush=True)
        return
    if args.merge:
        print(json.dumps(merge(args)), flush=True)
        return
    if args.obs_build:
        print(json.dumps(obs_build(args)), flush=True)
        return
    if args.sw:
        print(json.dumps(sw_model(args)), flush=True)
        return
    if args.sw4dvar:
        print(json.dumps(sw4dvar(args)), flush=True)
        return
    if args.gemm_only:
        from vardyn_amd.qg import QgPlan
        plan = QgPlan(case.ny, case.nx, case.dx, case.dx, case.dt, case.c, case.f0, g=case.g,
                      time_scheme=case.time_scheme, precision=args.precision,
                      nbatch=max(1, args.batch), dst_split=args.dst_split)
        tiles = plan.autotune(reps=5) if not args.no_autotune else None
        print(json.dumps({"gemm_roofline": gemm_roofline(plan, args.precision), "gemm_tiles": tiles,
                          "batch": plan.nb, "dst_split": plan.dst_split}), flush=True)
        return
    if args.stencil_only:
        out = {"stencil_roofline": stencil_roofline(case, args.precision),
               "adjoint_stencil_roofline": adj_stencil_roofline(case, args.precision),
               "copy_calibration": copy_calibration()}
        print(json.dumps(out), flush=True)
        return
    n_steps = args.window_steps or case.n_steps
    nb = max(1, args.batch)
    extra_steps = max(2, min(args.steps, 10))  # timed evaluations of the side measurements
    legs = {}

    plain = split = None
    if nb == 1 or not args.no_plain:
        log(f"plain path: one {case.ny}x{case.nx} problem, {n_steps} steps, {args.precision}, "
            f"dst_split {args.plain_dst_split}")
        pb1 = build_problem(case, precision=args.precision, n_steps=n_steps, control="ic",
                            tape="full" if args.tape == "auto" else args.tape,
                            dst_split=args.plain_dst_split)
        X1 = control_point(pb1.truth0)
        plain = measure(pb1, args, "plain", X1)
        if args.dst_split and not args.plain_dst_split and nb > 1:
            split = split_check(case, pb1, args.precision, n_steps, args.dst_split, X1)
            log(f"split check: {split}")
        if nb > 1 and not args.no_checkpoint and plain["tape"] == "full":
            retape(pb1, "checkpoint")
            legs["plain_checkpoint"] = measure(pb1, args, "plain, checkpointed", X1,
                                               tiles=plain["tiles"], steps=extra_steps, warmup=1)
        if nb == 1:
            pb, head = pb1, plain
        else:
            del pb1
            release(None)
    if nb > 1:
        t0 = time.time()
        pb = build_batch(case, nb, precision=args.precision, n_steps=n_steps, tape=args.tape,
                         dst_split=args.dst_split)
        log(f"batched path: {nb} problems built in {time.time() - t0:.1f}s")
        Xb = control_point(pb.truth)
        head = measure(pb, args, "batched", Xb)
        if not args.no_checkpoint and head["tape"] == "full":
            retape(pb, "checkpoint")
            legs["batched_checkpoint"] = measure(pb, args, "batched, checkpointed", Xb,
                                                 tiles=head["tiles"], steps=extra_steps, warmup=1)

    roof = gemm_roofline(pb.plan, args.precision)
    log(f"roofline: {roof}")
    value, ms = head["value"], head["ms_per_step"]
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "gradients/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "float64" else "f32",
        "data": "synthetic (seeded random initial SSH, truth run, nadir tracks + swaths per member)",
        "config": {"workload": f"{case.name}" + (f" x{nb} batched" if nb > 1 else ""),
                   "grid": [case.ny, case.nx], "dx_m": case.dx, "dt_s": case.dt,
                   "window_steps": n_steps, "time_scheme": case.time_scheme,
                   "obs_per_member": pb.H.n_total, "segments": len(pb.checkpoints) - 1,
                   "checkpoint_steps": case.checkpoint_steps,
                   "control": "initial SSH", "evaluated_at": "X = 0.5 x truth initial SSH",
                   "tape": head["tape"], "batch": nb,
                   "hip_graph": not args.no_graph, "folded_dst": pb.plan.folded,
                   "dst_split": pb.plan.dst_split, "gemm_tiles": head["tiles"]},
        "grid_point_steps_per_s": round(case.ny * case.nx * n_steps * value, 1),
        "ms_per_gradient": round(ms / nb, 2),
        "J": head["J0"], "grad_norm": head["gnorm0"],
        "roofline": roof,
    }
    if split is not None:
        out["split_check"] = split
    if plain is not None and nb > 1:
        out["plain_path"] = {"value": round(plain["value"], 4), "unit": "gradients/s",
                             "ms_per_gradient": round(plain["ms_per_step"], 2),
                             "tape": plain["tape"], "gemm_tiles": plain["tiles"],
                             "dst_split": args.plain_dst_split}
        out["batched_over_plain"] = round(value / plain["value"], 3)
    for k, r in legs.items():
        out[k] = {"value": round(r["value"], 4), "unit": "gradients/s",
                  "ms_per_gradient": round(r["ms_per_step"] / r["batch"], 2),
                  "ms_per_step": round(r["ms_per_step"], 2), "batch": r["batch"],
                  "tape": r["tape"], "steps": r["steps"], "J": r["J0"]}
    if "batched_checkpoint" in legs and plain is not None:
        bc = legs["batched_checkpoint"]["value"]
        out["ratios"] = {
            "batched_full_over_plain_full": round(value / plain["value"], 3),
            "batched_checkpoint_over_plain_full": round(bc / plain["value"], 3)}
        if "plain_checkpoint" in legs:
            out["ratios"]["batched_checkpoint_over_plain_checkpoint"] = round(
                bc / legs["plain_checkpoint"]["value"], 3)
    if args.check_f32 and args.precision == "float64" and nb == 1:
        X0 = torch.as_tensor(X1, device=pb.plan.device)
        retape(pb, "full")
        J64, G64 = pb.engine.run(X0)
        G64 = G64.clone()
        pb32 = build_problem(case, precision="float32", n_steps=n_steps, control="ic",
                             tape="full", truth_plan=pb.plan)
        J32, G32 = pb32.engine.run(X0)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        pb32.engine.run(X0)
        torch.cuda.synchronize()
        out["f32_path"] = {
            "gradients_per_s": round(1.0 / (time.perf_counter() - t2), 4),
            "grad_rel_err_vs_f64": float((torch.linalg.vector_norm(G32 - G64) /
                                          torch.linalg.vector_norm(G64)).item()),
            "J_rel_err_vs_f64": abs(J32.item() - J64.item()) / abs(J64.item())}
        log(f"f32 path: {out['f32_path']}")
        del pb32
    pb.engine = None
    del pb
    release(None)
    if not args.no_problem_set:
        ps = problem_set(args, steps=extra_steps, warmup=1, separate=False)
        out["problem_set"] = {k: ps[k] for k in ("value", "unit", "ms_per_step", "steps",
                                                 "grid_point_steps_per_s", "build_s")}
        out["problem_set"]["workload"] = ps["config"]["workload"]
        out["problem_set"]["dst_split"] = ps["config"]["dst_split"]
        out["problem_set"]["group_tuning"] = ps["group_tuning"]
        log(f"problem set: {out['problem_set']}")
    if not args.no_stencil:
        out["stencil_roofline"] = stencil_roofline(case, args.precision)
        log(f"stencil roofline: {out['stencil_roofline']}")
        out["adjoint_stencil_roofline"] = adj_stencil_roofline(case, args.precision)
        log(f"adjoint stencil roofline: {out['adjoint_stencil_roofline']}")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_sample_steps, args.precision,
                                           args.dst_split)
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
