#!/usr/bin/env python
"""Benchmark of the MI355X analytic Kerr raytracer.

Default workload (one "step"): raytrace_ana over a 512x512 screen
(alpha, beta = meshgrid(linspace(-8, 8, 512))), a = 0.94, th_o = 17 deg,
r_o = 1e4, ngeo = 500 Mino-time samples per ray: per-pixel setup (kgeo_plan)
plus all 131,072,000 samples with r, theta, phi, t written to HBM
(kgeo_sample).  Inputs are resident in HBM before the timed region.

Prints ONE JSON line with the metric, the roofline of the dominant kernel
(kg_samples, FP64 vector ALU bound; timed with HIP events on the launch
stream) and the CPU baseline (the oracle, i.e. the reference algorithm in
NumPy/SciPy, on a bounded seeded sample of the same workload on this host).

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""

import argparse
import ctypes
import json
import multiprocessing as mp
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 geodesic samples/sec at 512², ngeo=500 (+% FP64 peak); grad-images/sec"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix), dense
HBM_PEAK_GBS = 8000.0

A_SPIN, TH_DEG, R_O, WINDOW, NPIX_SIDE, NGEO = 0.94, 17.0, 1.0e4, 8.0, 512, 500

# Algorithmic FLOP weights: +, -, *, /, sqrt count 1; each transcendental
# counts the FP64 VALU operations of its gfx950 (ROCm 7.2 ocml) expansion,
# FMA = 2, from the AMD assembly of this build (static count of the
# function body).
TRANS_WEIGHTS = dict(SIN=132, COS=132, TAN=163, ASIN=93, ACOS=83, ATAN=62, ATAN2=70, LOG=90, LOG1P=129, EXP=32,
                     CBRT=28, POW=196)
OP_NAMES = "ADD MUL DIV SQRT SIN COS TAN ASIN ACOS ATAN ATAN2 LOG LOG1P EXP CBRT POW".split()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def screen():
    al = np.linspace(-WINDOW, WINDOW, NPIX_SIDE)
    A, B = np.meshgrid(al, al)
    return A.flatten(), B.flatten()


def algorithmic_flops_per_sample(A, B, th, npix_sample=512, seed=0):
    """Counts the per-sample FP64 operations of the kernel's own algorithm on a
    seeded subset of the workload's pixels (CPU instantiation of the device
    source, kgeo_amd/libkgeo_opcount.so)."""
    path = os.path.join(ROOT, "kgeo_amd", "libkgeo_opcount.so")
    lib = ctypes.CDLL(path)
    n = lib.kgc_num_ops()
    idx = np.random.default_rng(seed).choice(len(A), npix_sample, replace=False)
    a_, b_ = np.ascontiguousarray(A[idx]), np.ascontiguousarray(B[idx])
    sc = np.array([A_SPIN, R_O, th, 1 - 1e-10, A_SPIN ** 2, np.sin(th), np.cos(th), np.cos(th) ** 2,
                   1 + np.sqrt(1 - A_SPIN ** 2), 1 - np.sqrt(1 - A_SPIN ** 2)])
    ps = np.zeros(n, np.uint64)
    su = np.zeros(n, np.uint64)
    P = lambda x, t=ctypes.c_double: x.ctypes.data_as(ctypes.POINTER(t))
    lib.kgc_count(P(sc), P(a_), P(b_), ctypes.c_int64(npix_sample), ctypes.c_int64(NGEO),
                  P(ps, ctypes.c_uint64), P(su, ctypes.c_uint64))
    per = {k: int(v) / (npix_sample * NGEO) for k, v in zip(OP_NAMES, ps)}
    setup = {k: int(v) / npix_sample for k, v in zip(OP_NAMES, su)}
    w = lambda d: sum(d[k] * (TRANS_WEIGHTS.get(k, 1)) for k in OP_NAMES)
    return w(per), w(setup), per


def _cpu_chunk(args):
    from oracle import kgeo_oracle as O
    a, r_o, th, A, B, G = args
    os.environ["OMP_NUM_THREADS"] = "1"
    t0 = time.perf_counter()
    O.raytrace_ana(a, r_o, th, A, B, G)
    return time.perf_counter() - t0


def cpu_baseline(A, B, th, px_per_core=768, seed=1):
    """The oracle (the reference algorithm restated in NumPy/SciPy) on a seeded
    sample of the workload's pixels, 1 process and then P processes."""
    from oracle import kgeo_oracle as O
    ncores = min(len(os.sched_getaffinity(0)), 16)
    rng = np.random.default_rng(seed)
    idx = rng.choice(len(A), px_per_core * ncores, replace=False)
    # warm-up + single process
    O.raytrace_ana(A_SPIN, R_O, th, A[idx[:32]], B[idx[:32]], NGEO)
    t1 = _cpu_chunk((A_SPIN, R_O, th, A[idx[:px_per_core]], B[idx[:px_per_core]], NGEO))
    rate1 = px_per_core * NGEO / t1
    log(f"cpu baseline, 1 process: {px_per_core} px x {NGEO} samples in {t1:.2f} s = {rate1:.3e} samples/s")
    chunks = [(A_SPIN, R_O, th, A[idx[i * px_per_core:(i + 1) * px_per_core]],
               B[idx[i * px_per_core:(i + 1) * px_per_core]], NGEO) for i in range(ncores)]
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(ncores) as pool:
        pool.map(_cpu_chunk, chunks)
    tp = time.perf_counter() - t0
    rateP = px_per_core * ncores * NGEO / tp
    log(f"cpu baseline, {ncores} processes: {px_per_core * ncores} px in {tp:.2f} s = {rateP:.3e} samples/s")
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    import scipy
    return {"value": rateP, "unit": "samples/s", "cores": ncores, "kind": "port",
            "sample": f"{px_per_core * ncores} seeded pixels of the 512x512 grid x ngeo=500 "
                      f"({px_per_core * ncores * NGEO} samples), {ncores} processes x 1 thread",
            "single_process_value": rate1, "cpu_model": model, "numpy": np.__version__,
            "scipy": scipy.__version__}


def read_traffic():
    """HBM bytes per kg_samples launch from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "kg_samples_512x512_ngeo500_pmc.txt")
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith("traffic_bytes_per_launch:"):
            return float(ln.split(":", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build measures one GPU only (every box has one MI355X); "
              "the raytracer shards over pixels with no collectives, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import torch
    from kgeo_amd import _lib
    from kgeo_amd.raytrace import _ptr, _stream_handle

    th = TH_DEG * np.pi / 180.0
    A, B = screen()
    npix = len(A)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    lib = _lib.load()
    al = torch.from_numpy(A).to(dev)
    be = torch.from_numpy(B).to(dev)
    f64 = dict(device=dev, dtype=torch.float64)
    out = {k: torch.empty((NGEO, npix), **f64) for k in ("t", "r", "th", "ph")}
    rclass = torch.empty(npix, device=dev, dtype=torch.int8)
    thclass = torch.empty(npix, device=dev, dtype=torch.int8)
    nmax = torch.empty(npix, device=dev, dtype=torch.int32)
    tau_tot = torch.empty(npix, **f64)
    status = torch.empty(npix, device=dev, dtype=torch.int32)
    wsb = _lib.workspace_bytes(npix)
    ws = torch.empty(wsb, device=dev, dtype=torch.uint8)
    p = _lib.make_params(A_SPIN, R_O, th, npix, NGEO, 1.0 - 1e-10, True)
    stream = _stream_handle()

    def step(ev=None):
        if ev is not None:
            ev[0].record()
        _lib.check(lib.kgeo_plan(p, _ptr(al), _ptr(be), _ptr(rclass), _ptr(thclass), _ptr(nmax), _ptr(tau_tot),
                                 _ptr(status), _ptr(ws), wsb, stream), "kgeo_plan")
        if ev is not None:
            ev[1].record()
        _lib.check(lib.kgeo_sample(p, None, NGEO, None, _ptr(out["t"]), _ptr(out["r"]), _ptr(out["th"]),
                                   _ptr(out["ph"]), None, _ptr(status), _ptr(ws), wsb, stream), "kgeo_sample")
        if ev is not None:
            ev[2].record()

    log(f"warmup {args.warmup} steps ({npix} px x {NGEO} samples)")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    log(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(evs[i])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    plan_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    samp_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    samples = npix * NGEO
    value = samples * args.steps / elapsed
    ms_step = elapsed / args.steps * 1e3
    log(f"step {ms_step:.2f} ms (plan {plan_ms:.3f} ms, samples {samp_ms:.3f} ms) -> {value:.3e} samples/s")

    st = status.cpu().numpy()
    nonfinite = int(np.sum(st & 1))

    flops_s, flops_setup, ops = algorithmic_flops_per_sample(A, B, th)
    achieved_tf = flops_s * samples / (samp_ms * 1e-3) / 1e12
    bytes_per_launch = 32.0 * samples
    hbm_gbs = bytes_per_launch / (samp_ms * 1e-3) / 1e9
    roofline = {"bound": "valu-fp64", "kernel": "kg_theta + kg_radial", "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved_tf / FP64_PEAK_TFLOPS, "traffic": read_traffic(),
                "flops_per_sample": flops_s, "algorithmic_bytes_per_launch": bytes_per_launch,
                "kernel_ms": samp_ms, "hbm_write_GBs": hbm_gbs, "hbm_frac": hbm_gbs / HBM_PEAK_GBS,
                "setup_kernel_ms": plan_ms, "setup_flops_per_pixel": flops_setup,
                "ops_per_sample": {k: round(v, 3) for k, v in ops.items()}}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(A, B, th)

    line = {"metric": METRIC, "value": value, "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "raytrace_ana 512x512 ngeo=500 (a=0.94, th_o=17 deg, r_o=1e4, "
                                   "alpha,beta in [-8,8]); writes r, theta, phi, t",
                       "npix": npix, "ngeo": NGEO, "parallelism": "single GPU, pixels x sample chunks"},
            "roofline": roofline, "cpu_baseline": cpu,
            "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
            "nonfinite_pixels": nonfinite, "host": platform.node()}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
