#!/usr/bin/env python3
"""Headline benchmark: one explicit ADS heat step (RHS assembly by element
quadrature + x/y/z banded mass-matrix sweeps) at 256^3 elements, p = 3
(259^3 DOFs).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Prints ONE JSON line.  ``value`` is FP64 DOF-updates/s of the plain step with
the state resident in HBM (timed with HIP events on the stream the library
runs on, bracketed by device synchronisation), after half a second of
untimed steps that bring the GPU to its steady clock and the W warm-up steps.  Alongside it:
  * ``fp32_ir``: the precision form as specified (FP32 element assembly + an
    FP32 segmented solve, then one correction: FP64 Kronecker residual + FP32
    segmented solve) on the same state, its speed-up and its error against
    the FP64 step;
  * ``fp32_ir_kron``: the precision form without element assembly (two passes
    of FP64 Kronecker residual + FP32 LU sweeps from a zero start; constant
    kappa);
  * ``fp32_ir_seg``: as ``fp32_ir_kron`` with each FP32 solve cut into line
    segments of at most 32 entries solved in parallel, each from zero state
    w entries outside it (the LU recurrences forget their state);
  * ``fp64_kron``: FP64 throughout, the right-hand side from the fused
    Kronecker operator of the 1-D matrices instead of the element assembly
    (exact for a constant kappa, where the Gauss sum over elements and points
    factorises per direction) + the same FP64 solve;
  every form carries its speed-up against the FP64 step and against
  ``fp64_kron`` (the fastest FP64 path for this constant-kappa workload), and
  whether it stays within 1e-10 of the FP64 step;
  * ``grad_fwd``: time of the checkpointed gradient (segment recomputation +
    adjoint steps) over the time of the forward rollout at 96^3 elements,
    p = 2, 50 steps, checkpoint 10;
  * ``config5``: the combined large case (384x384x256 elements, p = 3, a
    lognormal diffusivity field, 40 steps, checkpoint 8): the FP32 forward
    and adjoint against the FP64 ones (time per step, DOF/s, HBM fraction of
    its minimum traffic, gradient/forward, errors of the state and both
    gradients);
  * ``roofline``: the heat assembly kernel (the step's dominant kernel) timed
    alone with HIP events on the library stream (ads_time_kernel), the FP64
    flops it executes (position-first form) against MI355X's FP64 peak, with
    the equivalent rates of the element-operator and point-quadrature counts
    beside it; ``solves`` the three FP64/FP32 mass solves by bytes (16 B/DOF
    per direction in FP64); ``kron_residual`` the fused residual kernel of the
    fast precision form, by bytes;
  * ``shard``: one GPU's share of a z-slab decomposition over 8 GPUs
    (256x256x32 elements): the step and its kernels measured at that shape,
    with the all-to-all and halo traffic modelled (not measured);
  * ``oil``: the oil problem (problems/oil: kq(X) exp(mi u) permeability,
    pumps, drains) at 128^3 elements, p = 2: step time, its point-quadrature
    kernel against the FP64 peak, and the oracle's faithful step on a 48^3
    sample on the host cores;
  * ``cpu_baseline``: the oracle's reference-structured CPU path (per-point
    integrand loops, rebuilt 1-D factorisations, per-line solves) timed over
    one whole step of the same workload on the host cores, with the CPU
    model; the sum-factorised oracle step beside it, and a one-thread
    sum-factorised step at 128^3, p = 2.
Data are synthetic: u0 = prod sin(pi x_i) projected, plus U[-0.01, 0.01)
noise (seed 2) on the coefficients.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def heat_kernel_flops(E, P):
    """Flops of one launch of the position-first heat assembly (k_heat_pf)
    per element: the y and z contractions once per x position, 5 P^3 FMA
    (z: 2 P^3 for Mz c and Kz c; y: 3 P^3 for My, Ky of those); the x
    contractions over the element's P positions, 3 P^4 FMA (Mx GM, Kx GM,
    Mx GS); M c - dt kappa K c and c . M c, 2 P^3 FMA: 6 P^4 + 14 P^3
    flops (2,432 at p = 3).  The halo positions' share is not counted."""
    return E * (6 * P ** 4 + 14 * P ** 3)


def element_operator_flops(E, P):
    """Flops of the same operator applied element by element (x, z, y
    stages per element and x test index, 8 P^4 + 2 P^3 FMA): 16 P^4 + 4 P^3
    per element (4,352 at p = 3)."""
    return E * (16 * P ** 4 + 4 * P ** 3)


def residual_bytes(D):
    """Algorithmic HBM bytes of the first residual pass: read c (FP64) and
    write r (FP32) once: 12 B/DOF."""
    return 12 * D


OIL_PUMPS = [(0.3, 0.3, 0.5), (0.7, 0.7, 0.5)]
OIL_DRAINS = [(0.5, 0.5, 0.2), (0.5, 0.5, 0.8)]


def oil_kernel_flops(E, P):
    """Algorithmic flops of one oil right-hand-side launch per element: the
    sum-factorised FormUn (u and grad u at Q^3 = P^3 points: x 2P^4, y 3P^4,
    z 4P^4 FMA) and the transposed contraction of the four point weights
    (z 4P^4, y 3P^4, x 2P^4): 18 P^4 FMA = 36 P^4 flops (the per-point
    weights, exp and source sums not counted)."""
    return E * 36 * P ** 4


def oil_cpu_baseline(rhs, dt):
    """The oracle's faithful oil step (per-point integrand loops, rebuilt 1-D
    factorisations) at 48^3 elements, p = 2, same curves, pumps and drains."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    n, p = 49, 2
    S = O.Setup(n, p)
    mi, kmin, kmax, rad, ps, ds = rhs.params
    oil = O.Oil(rhs.cx, rhs.cy, rhs.cz, rhs.cN, rhs.cL, pumps=rhs.pumps, drains=rhs.drains, mi=mi, kq_min=kmin,
                kq_max=kmax, radius=rad, pump_strength=ps, drain_strength=ds)
    kq = S.oil_kq_cache(oil)
    c, _ = S.oil_project(oil)
    t0 = time.perf_counter()
    S.oil_step(oil, kq, c, dt)
    t = time.perf_counter() - t0
    return {"value": S.D / t, "unit": "DOF/s", "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
            "kind": "port", "sample": "oracle faithful oil step at 48^3 elements, p = 2 (%.2f s)" % t}


def cpu_model():
    """The host CPU's model name (lscpu's "Model name", else /proc/cpuinfo)."""
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


_ONE_THREAD_STEP = r"""
import sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
import oracle as O
S = O.Setup(129, 2)
c = np.random.default_rng(1).uniform(-1.0, 1.0, S.D)
S.step(c, 1e-6)
t = []
for _ in range(3):
    t0 = time.perf_counter(); S.step(c, 1e-6); t.append(time.perf_counter() - t0)
print(sorted(t)[1])
"""


def cpu_baseline(n, p, dt):
    """The oracle's faithful path (the reference's loop structure) timed over
    one whole step of the benchmark workload on all host threads, plus the
    sum-factorised step on the same threads and, in a child process, the
    sum-factorised 128^3 p = 2 step on one thread (median of 3)."""
    import subprocess

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    S = O.Setup(n, p)
    c = np.random.default_rng(2).uniform(-1.0, 1.0, S.D)
    t0 = time.perf_counter()
    S.step(c, dt, mode=O.MODE_FAITHFUL)
    t_step = time.perf_counter() - t0
    t0 = time.perf_counter()
    S.step(c, dt, mode=O.MODE_SUMFACT)
    t_sf = time.perf_counter() - t0
    del S
    one = None
    try:
        env = dict(os.environ, OMP_NUM_THREADS="1")
        res = subprocess.run([sys.executable, "-c", _ONE_THREAD_STEP, os.path.join(ROOT, "oracle")], env=env,
                             capture_output=True, text=True, timeout=120)
        if res.returncode == 0:
            t1 = float(res.stdout.strip().splitlines()[-1])
            one = {"workload": "128^3 elements, p=2, 130^3 DOFs", "mode": "sumfact", "cores": 1,
                   "s_per_step": t1, "value": 130 ** 3 / t1}
    except Exception:
        one = None
    return {
        "value": S_D(n) / t_step,
        "unit": "DOF/s",
        "cores": threads,
        "kind": "port",
        "cpu": cpu_model(),
        "sample": ("one whole step of the benchmark workload (256^3 elements, p=3) by the oracle's faithful "
                   "path (reference loop structure: FormUn fused per element, per-(q,a) integrand, "
                   "MKBBT+dgbtrf rebuilt per direction, per-line dgbtrs with reorder copies): %.1f s" % t_step),
        "sumfact": {"s_per_step": t_sf, "value": S_D(n) / t_sf, "cores": threads},
        "one_thread": one,
    }


def S_D(n):
    return (n + 1) ** 3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--clock-warmup-s", type=float, default=0.5,
                    help="seconds of untimed FP64 steps before the warm-up steps: the GPU reaches its "
                         "steady clock after some tens of ms of load")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-oil", action="store_true")
    ap.add_argument("--no-shard", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: no CPU baseline, no gradient, plus one cold 1 GiB device "
                         "copy (known bytes) to calibrate FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.profile:
        args.no_cpu_baseline = True
        args.no_grad = True
        args.no_shard = True
        args.steps = min(args.steps, 5)
        args.warmup = min(args.warmup, 1)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the heat step runs on a single MI355X here "
              "(multi-GPU sharding is a design note, not a measured path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from ads_mpi_amd import _lib, adss
    from ads_mpi_amd.autograd import HeatSolver, heat_rollout

    L = _lib.load()
    torch.cuda.init()
    stream = torch.cuda.current_stream()

    n, p, dt = 258, 3, 1e-7
    ads, data = adss.initialize(n, p)
    h = data.handle
    L.ads_set_stream(h, ctypes.c_void_p(stream.cuda_stream))
    D = data.D
    E = int(np.prod(ads.nelem))
    P = p + 1

    # synthetic state: projection of prod sin + noise (seed 2)
    rhs = adss.HeatRHS(Dt=dt, f0=adss.ADS_F0_SINPROD)
    adss.Step(0, rhs, ads, data)
    data.F = data.F + adss.uniform(2, D, -0.01, 0.01)

    def run(k, prec, refine=1):
        for _ in range(k):
            _lib.check(L.ads_step(h, dt, prec, refine, None), h)

    # ---- FP64 step (value) ------------------------------------------------
    # untimed load until the GPU runs at its steady clock (the first tens of
    # ms of a fresh process run up to 10% slower), then the warm-up steps
    t_clk = time.perf_counter()
    while time.perf_counter() - t_clk < args.clock_warmup_s:
        run(10, _lib.ADS_PREC_FP64)
        torch.cuda.synchronize()
    run(args.warmup, _lib.ADS_PREC_FP64)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record(stream)
    run(args.steps, _lib.ADS_PREC_FP64)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms64 = e0.elapsed_time(e1) / args.steps
    l2 = ctypes.c_double()
    _lib.check(L.ads_last_l2norm(h, ctypes.byref(l2)), h)

    # ---- precision forms on the same state -------------------------------
    cdev = torch.empty(D, dtype=torch.float64, device="cuda")
    ref = torch.empty_like(cdev)
    got = torch.empty_like(cdev)
    torch.cuda.synchronize()
    cdev.copy_(torch.from_numpy(data.F).to("cuda"))
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    _lib.check(L.ads_step_d(h, vp(cdev), vp(ref), None, dt, _lib.ADS_PREC_FP64, 0, None), h)
    forms = {}
    for key, prec, refine, what in (
            ("fp32_ir", _lib.ADS_PREC_FP32_IR, 1,
             "FP32 element assembly + FP32 segmented solve, then 1 correction: FP64 Kronecker residual + FP32 "
             "segmented solve"),
            ("fp32_ir_kron", _lib.ADS_PREC_FP32_IR_KRON, 2,
             "zero start, 2 passes of fused FP64 Kronecker residual + FP32 sweeps (no element assembly)"),
            ("fp32_ir_seg", _lib.ADS_PREC_FP32_IR_SEG, 2,
             "zero start, 2 passes of fused FP64 Kronecker residual + FP32 segmented solve (lines cut into "
             "segments of <= 32 solved in parallel from zero state 32 entries outside, truncation 1e-8)"),
            ("fp64_kron", _lib.ADS_PREC_FP64_KRON, 0,
             "FP64: fused Kronecker right-hand side (1-D Gauss matrices, constant kappa) + FP64 sweeps")):
        _lib.check(L.ads_step_d(h, vp(cdev), vp(got), None, dt, prec, refine, None), h)
        torch.cuda.synchronize()
        rel_l2 = float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref))
        max_err = float(torch.max(torch.abs(got - ref)))
        run(args.warmup, prec, refine)
        torch.cuda.synchronize()
        e0.record(stream)
        run(args.steps, prec, refine)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        forms[key] = {"value": D / (ms * 1e-3), "ms_per_step": ms, "refine": refine, "form": what,
                      "speedup_vs_fp64": ms64 / ms, "rel_l2_vs_fp64": rel_l2, "max_abs_err_vs_fp64": max_err,
                      "within_1e-10": bool(rel_l2 <= 1e-10)}
    # the FP64 Kronecker form is the fastest FP64 path for this constant-kappa
    # workload: every form's speed against it as well
    for key in forms:
        forms[key]["speedup_vs_fp64_kron"] = forms["fp64_kron"]["ms_per_step"] / forms[key]["ms_per_step"]
    bad = [k for k in forms if not forms[k]["within_1e-10"]]
    if bad:
        print("bench.py: forms beyond 1e-10 of the FP64 step: %s" % bad, file=sys.stderr)

    # ---- kernels alone (roofline): HIP events around the launches only ----
    ms_k = ctypes.c_double()
    _lib.check(L.ads_time_kernel(h, _lib.ADS_KERNEL_HEAT_FP64, dt, 2, ctypes.byref(ms_k)), h)
    _lib.check(L.ads_time_kernel(h, _lib.ADS_KERNEL_HEAT_FP64, dt, args.steps, ctypes.byref(ms_k)), h)
    ms_asm = ms_k.value
    _lib.check(L.ads_time_kernel(h, _lib.ADS_KERNEL_KRON_RESIDUAL, dt, 2, ctypes.byref(ms_k)), h)
    _lib.check(L.ads_time_kernel(h, _lib.ADS_KERNEL_KRON_RESIDUAL, dt, args.steps, ctypes.byref(ms_k)), h)
    ms_res = ms_k.value
    # the mass solves (x, y and z launches together), by bytes: each direction
    # reads and writes every value once (16 B/DOF in FP64, 8 in FP32)
    solves = {}
    for key, which, bpd, what in (
            ("fp64_partitioned", _lib.ADS_KERNEL_SOLVE_FP64, 48,
             "k_part_rows + 2 x k_part_lines<double,3> (no-interchange LU, lines in partitions of 24 "
             "solved in parallel, exact edge-state scans; persistent, the next group loaded meanwhile)"),
            ("fp64_lu", _lib.ADS_KERNEL_SOLVE_FP64_LU, 48,
             "k_sweep_rows + 2 x k_sweep_lines<double,3> (dgbtrf/dgbtrs form, one lane per line)"),
            ("fp32_segmented", _lib.ADS_KERNEL_SOLVE_SEG_FP32, 24,
             "k_seg_rows + 2 x k_seg_lines<3,32> (FP32, segments of <= 32 from zero state w = 32 outside)")):
        _lib.check(L.ads_time_kernel(h, which, dt, 2, ctypes.byref(ms_k)), h)
        _lib.check(L.ads_time_kernel(h, which, dt, args.steps, ctypes.byref(ms_k)), h)
        gbs = bpd * D / (ms_k.value * 1e-3) / 1e9
        solves[key] = {"bound": "hbm", "kernels": what, "ms_per_launch": ms_k.value,
                       "bytes_per_launch": bpd * D, "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                       "frac": gbs / PEAK_HBM_GBS}
    if args.profile:
        # calibration dispatch for the PMC byte counters: a copy of 1 GiB whose
        # source was evicted from the 256 MB MALL by 1 GiB written after it
        n_cal = 1 << 27
        a_cal = torch.ones(n_cal, dtype=torch.float64, device="cuda")
        b_cal = torch.ones(n_cal, dtype=torch.float64, device="cuda")
        c_cal = torch.empty_like(a_cal)
        torch.cuda.synchronize()
        c_cal.copy_(a_cal)
        torch.cuda.synchronize()
        del a_cal, b_cal, c_cal
    flops = heat_kernel_flops(E, P)
    achieved = flops / (ms_asm * 1e-3) / 1e12
    traffic = None
    tf = os.path.join(ROOT, "profiles", "heat_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    res_gbs = residual_bytes(D) / (ms_res * 1e-3) / 1e9

    adss.Cleanup(ads, data)

    # ---- grad / forward at 96^3, p = 2 --------------------------------------
    grad = None
    if not args.no_grad:
        n4, p4, dt4, steps4, ck4 = 97, 2, 1e-7, 50, 10
        solver = HeatSolver(n4, p4)
        c0 = torch.from_numpy(adss.uniform(3, solver.D, -1.0, 1.0)).to("cuda").requires_grad_(True)
        kap = torch.from_numpy(adss.lognormal(3, solver.E, 0.0, 0.5)).to("cuda").requires_grad_(True)
        cobs = torch.from_numpy(adss.uniform(4, solver.D, -1.0, 1.0)).to("cuda")
        for _ in range(2):
            cT = heat_rollout(c0, kap, solver, dt4, steps4, ck4)
            loss = 0.5 * ((cT - cobs) ** 2).sum()
            loss.backward()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cT = heat_rollout(c0, kap, solver, dt4, steps4, ck4)
        loss = 0.5 * ((cT - cobs) ** 2).sum()
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        t0 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t0
        grad = {"workload": "96^3 p=2, 50 steps, checkpoint every 10, loss 0.5||c_T-c_obs||^2",
                "definition": "time of loss.backward() (segment recomputation + adjoint steps) / time of the "
                              "forward rollout",
                "forward_ms": 1e3 * t_fwd, "gradient_ms": 1e3 * t_bwd, "ratio": t_bwd / t_fwd}
        solver.close()

    # ---- the combined large case: 384x384x256, p = 3, kappa field -----------
    cfg5 = None
    if not args.no_grad:
        n5c, p5c, dt5c, steps5c, ck5c = (386, 386, 258), 3, 5e-9, 40, 8
        solver5 = HeatSolver(n5c, p5c)
        D5c, E5c = solver5.D, solver5.E
        c05 = torch.from_numpy(adss.uniform(5, D5c, -1.0, 1.0)).to("cuda")
        kap5 = torch.from_numpy(adss.lognormal(5, E5c, 0.0, 0.5)).to("cuda")
        cobs5 = torch.from_numpy(adss.uniform(6, D5c, -1.0, 1.0)).to("cuda")
        runs = {}
        for name, kw in (("fp64", dict(prec=_lib.ADS_PREC_FP64)),
                         ("fp32", dict(prec=_lib.ADS_PREC_FP32_IR_SEG, refine=2,
                                       adjoint_prec=_lib.ADS_PREC_FP32_IR_SEG, adjoint_refine=2))):
            # warm-up: a short rollout and its gradient (first-use tables and scratch)
            cg = c05.clone().requires_grad_(True)
            kg = kap5.clone().requires_grad_(True)
            (0.5 * ((heat_rollout(cg, kg, solver5, dt5c, 2, 1, **kw) - cobs5) ** 2).sum()).backward()
            cg = c05.clone().requires_grad_(True)
            kg = kap5.clone().requires_grad_(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            cT = heat_rollout(cg, kg, solver5, dt5c, steps5c, ck5c, **kw)
            torch.cuda.synchronize()
            t_f = time.perf_counter() - t0
            loss = 0.5 * ((cT - cobs5) ** 2).sum()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            loss.backward()
            torch.cuda.synchronize()
            t_g = time.perf_counter() - t0
            runs[name] = (cT.detach(), cg.grad, kg.grad, t_f, t_g)
        relerr = lambda a, b: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))

        def leg(name, form):
            cT, gc, gk, t_f, t_g = runs[name]
            ms = 1e3 * t_f / steps5c
            bps = 24.0 * D5c / (ms * 1e-3)  # read c and kappa, write c': 24 B per DOF per step
            return {"form": form, "ms_per_step": ms, "value": D5c / (ms * 1e-3), "unit": "DOF/s",
                    "hbm_bytes_per_step": 24 * D5c, "hbm_frac": bps / (PEAK_HBM_GBS * 1e9),
                    "forward_ms": 1e3 * t_f, "gradient_ms": 1e3 * t_g, "grad_fwd": t_g / t_f}

        cfg5 = {"workload": "384x384x256 elements, p=3, 387x387x259 DOFs, lognormal(0, 0.5) kappa (seed 5), "
                            "dt=5e-9, 40 steps, checkpoint every 8, loss 0.5||c_T-c_obs||^2",
                "grad_fwd_definition": "time of loss.backward() (segment recomputation + adjoint steps) / time "
                                       "of the forward rollout",
                "fp32": leg("fp32", "forward ADS_PREC_FP32_IR_SEG refine 2 (b = A(kappa) c by the FP64 element "
                                    "assembly, FP32 segmented solves with FP64 mass residuals), adjoint FP32"),
                "fp64": leg("fp64", "forward and adjoint ADS_PREC_FP64"),
                "rel_err_vs_fp64": {"state": relerr(runs["fp32"][0], runs["fp64"][0]),
                                    "grad_c0": relerr(runs["fp32"][1], runs["fp64"][1]),
                                    "grad_kappa": relerr(runs["fp32"][2], runs["fp64"][2])}}
        del runs
        solver5.close()

    # ---- one GPU's share of a z-slab decomposition over 8 GPUs ------------
    # 256 x 256 x 32 elements (259 x 259 x 35 DOFs with its p halo planes):
    # the heat assembly and the x and y solves at their real shape; the z
    # solve here runs on 35-entry lines (after the all-to-all of a real
    # decomposition it would run on 259-entry lines of a y slab, the same
    # bytes per DOF).  Not a multi-GPU measurement: the communication is a
    # model (all-to-all over 7 xGMI links, one neighbour halo).
    shard = None
    if not args.no_shard:
        G = 8
        ads_s, data_s = adss.initialize((n, n, (n - p + 1) // G + p - 1), p)
        hs = data_s.handle
        L.ads_set_stream(hs, ctypes.c_void_p(stream.cuda_stream))
        data_s.F = adss.uniform(7, data_s.D, -1.0, 1.0)
        for _ in range(args.warmup):
            _lib.check(L.ads_step(hs, dt, _lib.ADS_PREC_FP64, 0, None), hs)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            _lib.check(L.ads_step(hs, dt, _lib.ADS_PREC_FP64, 0, None), hs)
        e1.record(stream)
        torch.cuda.synchronize()
        ms_s = e0.elapsed_time(e1) / args.steps
        _lib.check(L.ads_time_kernel(hs, _lib.ADS_KERNEL_HEAT_FP64, dt, 2, ctypes.byref(ms_k)), hs)
        _lib.check(L.ads_time_kernel(hs, _lib.ADS_KERNEL_HEAT_FP64, dt, args.steps, ctypes.byref(ms_k)), hs)
        ms_s_asm = ms_k.value
        _lib.check(L.ads_time_kernel(hs, _lib.ADS_KERNEL_SOLVE_FP64, dt, 2, ctypes.byref(ms_k)), hs)
        _lib.check(L.ads_time_kernel(hs, _lib.ADS_KERNEL_SOLVE_FP64, dt, args.steps, ctypes.byref(ms_k)), hs)
        ms_s_solve = ms_k.value
        # model: the z <-> y all-to-all before and after the z solve (each
        # GPU sends (G-1)/G of its D/G doubles over 7 links of 153 GB/s) and
        # the p-plane halo to one neighbour (c planes out, partial F planes
        # back) on one link, not overlapped
        a2a_ms = 2 * 8.0 * D / G * (G - 1) / G / (7 * 153e9) * 1e3
        halo_ms = 2 * 8.0 * p * (n + 1) ** 2 / 153e9 * 1e3
        t8 = ms_s + a2a_ms + halo_ms
        shard = {"workload": "one z slab of %d: 256x256x32 elements, p=3, %dx%dx%d DOFs (with its p halo planes)"
                             % (G, n + 1, n + 1, data_s.D // (n + 1) ** 2),
                 "gpus_modelled": G, "ms_per_step": ms_s, "heat_kernel_ms": ms_s_asm, "solve_fp64_ms": ms_s_solve,
                 "vs_full_step_over_G": ms_s / (ms64 / G),
                 "model": {"all_to_all_ms": a2a_ms, "halo_ms": halo_ms, "step_ms": t8,
                           "value": D / (t8 * 1e-3), "unit": "DOF/s",
                           "note": "compute measured on one GPU at the slab's shape; communication modelled, "
                                   "not measured (one GPU per box)"}}
        adss.Cleanup(ads_s, data_s)

    # ---- oil problem at 128^3, p = 2 ---------------------------------------
    oil = None
    if not args.no_oil:
        # explicit-step stability: dt Kqmax exp(mi u) lambda_max(M^-1 K) < 2 at h = 1/128
        n5, p5, dt5 = 129, 2, 1e-9
        ads5, data5 = adss.initialize(n5, p5)
        h5 = data5.handle
        L.ads_set_stream(h5, ctypes.c_void_p(stream.cuda_stream))
        rhs5 = adss.OilRHS.with_random_curves(dt5, 1, pumps=OIL_PUMPS, drains=OIL_DRAINS)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rhs5.setup(data5)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        adss.Step(0, rhs5, ads5, data5)
        for _ in range(args.warmup):
            _lib.check(L.ads_oil_step(h5, dt5, None, None), h5)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            _lib.check(L.ads_oil_step(h5, dt5, None, None), h5)
        e1.record(stream)
        torch.cuda.synchronize()
        ms_oil = e0.elapsed_time(e1) / args.steps
        l2o = ctypes.c_double()
        _lib.check(L.ads_last_l2norm(h5, ctypes.byref(l2o)), h5)
        _lib.check(L.ads_time_kernel(h5, _lib.ADS_KERNEL_OIL, dt5, 2, ctypes.byref(ms_k)), h5)
        _lib.check(L.ads_time_kernel(h5, _lib.ADS_KERNEL_OIL, dt5, args.steps, ctypes.byref(ms_k)), h5)
        E5 = 128 ** 3
        D5 = data5.D
        fl5 = oil_kernel_flops(E5, p5 + 1)
        ach5 = fl5 / (ms_k.value * 1e-3) / 1e12
        oil = {"workload": "oil problem, 128x128x128 elements, p=2, 130^3 DOFs, dt=1e-9, 30 curves x 16 points "
                           "(seed 1), 2 pumps, 2 drains, mi=10, Kq 1..1000, radius 0.15",
               "value": D5 / (ms_oil * 1e-3), "unit": "DOF/s", "ms_per_step": ms_oil, "l2norm": l2o.value,
               "kq_setup_ms": 1e3 * t_setup,
               "kernel": {"name": "k_oil<P=3> (FP64)", "bound": "fp64_valu", "ms_per_launch": ms_k.value,
                          "flops_per_launch": fl5, "achieved": ach5, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                          "frac": ach5 / PEAK_FP64_TFLOPS},
               "cpu_baseline": None if args.no_cpu_baseline else oil_cpu_baseline(rhs5, dt5)}
        adss.Cleanup(ads5, data5)

    line = {
        "metric": "ADS step DOF/s, 256^3 p=3, FP64 vs FP32+IR at <=1e-10 rel; grad/fwd time ratio",
        "value": D / (ms64 * 1e-3),
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "clock_warmup_s": args.clock_warmup_s,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "heat step, 256x256x256 elements, p=3, 259^3 DOFs, dt=1e-7, FP64 "
                               "(element-quadrature RHS assembly + x/y/z banded sweeps)",
                   "elements": [256, 256, 256], "p": 3, "dofs": D, "dt": dt},
        "wall_s_timed": wall,
        "l2norm": l2.value,
        "fp32_ir": forms["fp32_ir"],
        "fp32_ir_kron": forms["fp32_ir_kron"],
        "fp32_ir_seg": forms["fp32_ir_seg"],
        "fp64_kron": forms["fp64_kron"],
        "grad_fwd": grad,
        "config5": cfg5,
        "shard": shard,
        "oil": oil,
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "kernel": "k_heat_pf<double,P=4,plain> (FP64 VALU; MI355X's FP64 vector and matrix peaks are "
                               "equal)",
                     "basis": ("flops the kernel executes per element: the element operator F_e = (Mz(x)My(x)Mx - "
                               "dt kappa (Mz(x)My(x)Kx + Mz(x)Ky(x)Mx + Kz(x)My(x)Mx)) c_e with the 1-D element "
                               "matrices, its y and z contractions applied once per x position and shared by the "
                               "P elements there: 3P^4 + 7P^3 FMA = 6P^4 + 14P^3 flops per element (equal to the "
                               "Gauss-point sum: Q = p+1 points integrate the degree-2p integrands exactly)"),
                     "ms_per_launch": ms_asm, "flops_per_launch": flops,
                     "flops_per_element": 6 * P ** 4 + 14 * P ** 3, "elements_per_launch": E,
                     "element_operator_form": {
                         "flops_per_element": 16 * P ** 4 + 4 * P ** 3,
                         "note": ("the same operator applied element by element (8P^4 + 2P^3 FMA), which the "
                                  "element-first kernel k_heat executes; equivalent rate of this launch"),
                         "equivalent_achieved": element_operator_flops(E, P) / (ms_asm * 1e-3) / 1e12,
                         "equivalent_frac": element_operator_flops(E, P) / (ms_asm * 1e-3) / 1e12
                         / PEAK_FP64_TFLOPS},
                     "quadrature_form": {
                         "flops_per_element": 36 * P ** 4 + 12 * P ** 3,
                         "note": ("the sum-factorised point-quadrature count (u and grad u at the Q^3 points, "
                                  "the integrand, the transposed contraction), which this kernel does not "
                                  "execute: its equivalent rate exceeds the FP64 peak"),
                         "equivalent_achieved": E * (36 * P ** 4 + 12 * P ** 3) / (ms_asm * 1e-3) / 1e12,
                         "equivalent_frac": E * (36 * P ** 4 + 12 * P ** 3) / (ms_asm * 1e-3) / 1e12
                         / PEAK_FP64_TFLOPS}},
        "solves": solves,
        "kron_residual": {"bound": "hbm", "achieved": res_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": res_gbs / PEAK_HBM_GBS, "kernel": "k_kron_residual<P=4,x=0,r=FP32>",
                          "ms_per_launch": ms_res, "bytes_per_launch": residual_bytes(D)},
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(n, p, dt),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
