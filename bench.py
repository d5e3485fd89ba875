#!/usr/bin/env python
"""Benchmark of the MI355X analytic Kerr raytracer.

Default workload (one "step"): raytrace_ana over a 512x512 screen
(alpha, beta = meshgrid(linspace(-8, 8, 512))), a = 0.94, th_o = 17 deg,
r_o = 1e4, ngeo = 500 Mino-time samples per ray: per-pixel setup (kgeo_plan)
plus all 131,072,000 samples with r, theta, phi, t written to HBM
(kgeo_sample).  Inputs are resident in HBM before the timed region.

Prints ONE JSON line with the metric, the roofline of the sample passes
(the heads, far, walk and fixup kernels of theta and r, FP64 vector ALU;
timed with HIP events on the launch stream; algorithmic FLOPs op-counted on
the kernels' own code) and the CPU
baseline (the oracle, i.e. the reference algorithm in NumPy/SciPy, on a
bounded seeded sample of the same workload on this host).

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
  python bench.py --mode batch        256-image survey, one persistent graph launch
  python bench.py --mode equatorial   equatorial images/s and gradient images/s
  python bench.py --mode copy         plain 1 GiB copy (counter calibration)
"""

import argparse
import ctypes
import json
import multiprocessing as mp
import os
import platform
import subprocess
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


def algorithmic_flops_per_sample(A, B, th, chunks, npix_sample=512, seed=0):
    """Counts the per-sample FP64 operations of the kernels' own loops (the
    walk along the grid with its direct evaluations, split into `chunks`
    ranges per pixel as the launch does) on a seeded subset of the workload's
    pixels: CPU instantiation of the device source on an op-counting type,
    kgeo_amd/libkgeo_opcount.so."""
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
    lib.kgc_count(P(sc), P(a_), P(b_), ctypes.c_int64(npix_sample), ctypes.c_int64(NGEO), ctypes.c_int(chunks),
                  P(ps, ctypes.c_uint64), P(su, ctypes.c_uint64))
    per = {k: int(v) / (npix_sample * NGEO) for k, v in zip(OP_NAMES, ps)}
    setup = {k: int(v) / npix_sample for k, v in zip(OP_NAMES, su)}
    w = lambda d: sum(d[k] * (TRANS_WEIGHTS.get(k, 1)) for k in OP_NAMES)
    return w(per), w(setup), per


def _cpu_chunk(args):
    from oracle import kgeo_oracle as O
    a, r_o, th, A, B, G = args
    t0 = time.perf_counter()
    O.raytrace_ana(a, r_o, th, A, B, G)
    return time.perf_counter() - t0


def _cpu_pool_init():
    for k in ("OMP_NUM_THREADS", "MKL_NUM_THREADS", "OPENBLAS_NUM_THREADS"):
        os.environ[k] = "1"


def _cpu_time(A, B, th, G, procs, reps):
    """Median wall time of `reps` runs (after one warm-up) of the oracle over
    pixels A, B: one process, or `procs` processes over contiguous chunks."""
    def once(pool):
        if pool is None:
            return _cpu_chunk((A_SPIN, R_O, th, A, B, G))
        parts = np.array_split(np.arange(len(A)), procs)
        t0 = time.perf_counter()
        pool.map(_cpu_chunk, [(A_SPIN, R_O, th, A[ix], B[ix], G) for ix in parts])
        return time.perf_counter() - t0
    if procs == 1:
        once(None)
        return float(np.median([once(None) for _ in range(reps)]))
    with mp.get_context("fork").Pool(procs, initializer=_cpu_pool_init) as pool:
        once(pool)
        return float(np.median([once(pool) for _ in range(reps)]))


def cpu_baseline(th):
    """The oracle (the reference algorithm restated in NumPy/SciPy, with the
    same scipy.special calls) timed on this host: the 64x64 ngeo=100 image in
    full (409,600 samples) and the seeded 4096-pixel subset of the 512x512
    ngeo=500 image (2,048,000 samples, the pixel mix the parity tests use),
    each in one process and in P processes of one thread; one warm-up, then
    the median of 3 (one timed run for the single-process subset, about 6 s).
    P is this process's CPU affinity, capped at 16: a one-GPU box shares its
    host among GPUs and gives each GPU's job 16 cores, whatever nproc shows."""
    import scipy
    _cpu_pool_init()
    procs = max(1, min(len(os.sched_getaffinity(0)), 16))
    al = np.linspace(-WINDOW, WINDOW, 64)
    A1, B1 = np.meshgrid(al, al)
    A1, B1 = A1.flatten(), B1.flatten()
    A2, B2 = screen()
    sub = np.random.default_rng(0).choice(len(A2), 4096, replace=False)   # the parity tests' subset
    As, Bs = A2[sub], B2[sub]
    runs = {}
    for name, (A, B, G) in {"c1_full": (A1, B1, 100), "c2_subset": (As, Bs, NGEO)}.items():
        for P, reps in ((1, 3 if name == "c1_full" else 1), (procs, 3)):
            t = _cpu_time(A, B, th, G, P, reps)
            runs[f"{name}_{P}p"] = {"samples": len(A) * G, "processes": P, "seconds": t,
                                    "samples_per_s": len(A) * G / t, "reps": reps}
            log(f"cpu baseline {name}, {P} process(es): {len(A) * G} samples in {t:.2f} s "
                f"= {len(A) * G / t:.3e} samples/s (median of {reps})")
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    try:   # /proc/cpuinfo carries the full name where lscpu shortens it
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                full = ln.split(":", 1)[1].strip()
                model = full if len(full) > len(model) or model == "unknown" else model
                break
    except OSError:
        pass
    head = runs[f"c2_subset_{procs}p"]
    return {"value": head["samples_per_s"], "unit": "samples/s", "cores": procs, "kind": "port",
            "sample": "oracle on the seeded 4096-pixel subset of the 512x512 ngeo=500 image (2,048,000 "
                      f"samples), {procs} processes x 1 thread, median of 3 after a warm-up",
            "single_process_value": runs["c2_subset_1p"]["samples_per_s"], "runs": runs,
            "cpu_model": model, "affinity_cpus": len(os.sched_getaffinity(0)), "numpy": np.__version__,
            "scipy": scipy.__version__}


PMC_SUMMARY = os.path.join(ROOT, "profiles", "kg_sample_passes_512x512_ngeo500_pmc.txt")


IMAGE_SUMMARY = os.path.join(ROOT, "profiles", "kg_stokes_polarized_800x800_trace_pmc.md")


def read_profile_value(key, path=PMC_SUMMARY):
    """A figure from a committed rocprofv3 counter summary (by default the
    sample passes at the raytrace workload: FETCH_SIZE + WRITE_SIZE scaled by
    a plain copy's reading; hardware FP64 FLOP count), or None."""
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith(key + ":"):
            return float(ln.split(":", 1)[1])
    return None


def _events(n):
    import torch
    return [torch.cuda.Event(enable_timing=True) for _ in range(n)]


BATCH_CANDIDATES = [(64, 8), (128, 8), (256, 8), (128, 16)]


def batch_survey_figures(steps, warmup, retune=False, candidates=None):
    """Batched survey: 256 seeded images (spin, inclination, resolution,
    ngeo) in one persistent graph launch, against the same survey as 256
    sequential plain kgeo_raytrace_ana calls into the same arenas.  The
    launch shape comes from the autotuner (its cached choice, else a tuning
    run over every 8th image).  Returns the figures as a dict; the largest
    relative deviation between the two runs is taken over every 16th image
    (labels must be identical, NaN patterns equal)."""
    import torch
    from kgeo_amd import autotune
    from kgeo_amd.batch import KEYS, BatchSurvey, layout, run_sequential, survey_draw, survey_images, survey_size

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    draw = survey_draw(256, seed=0)
    images = survey_images(draw, window=WINDOW)
    pix, samples = survey_size(images)
    _, out_off = layout(images, r_o=R_O)
    log(f"batch survey: {len(images)} images, {pix} px, {samples} samples ({samples * 32 / 1e9:.1f} GB of r,th,ph,t)")
    arena = {k: torch.empty(int(out_off[-1]), device=dev, dtype=torch.float64) for k in KEYS}
    npix = np.array([len(im["alpha"]) for im in images])
    ngeo = np.array([im["ngeo"] for im in images])
    key = autotune.batch_key(npix, ngeo)
    choice = autotune.lookup(key)
    tune_ms = None
    if choice is None or retune:
        sub = images[::8]
        log(f"autotuning on {len(sub)} of the images (key {key})")
        t0 = time.perf_counter()
        choice, timings = autotune.tune_batch(sub, key=key, arena=arena, log=log, candidates=candidates)
        tune_ms = (time.perf_counter() - t0) * 1e3
        log(f"autotune choice {choice} in {tune_ms / 1e3:.1f} s")
    sv = BatchSurvey(images, r_o=R_O, tuning=choice, arena=arena)
    log(f"planned: {sv.items} work items, tuning {sv.tuning}")
    for _ in range(warmup):
        sv.run()
    sv.stream.synchronize()
    ev = _events(2 * steps)
    t0 = time.perf_counter()
    for i in range(steps):
        ev[2 * i].record(sv.stream)
        sv.run()
        ev[2 * i + 1].record(sv.stream)
    sv.stream.synchronize()
    wall_b = (time.perf_counter() - t0) / steps
    gpu_b = float(np.mean([ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(steps)])) / 1e3
    done, total = sv.progress()
    log(f"batched: {wall_b * 1e3:.1f} ms/survey (events {gpu_b * 1e3:.1f} ms) -> {samples / wall_b:.3e} samples/s")
    # the batched outputs of every 16th image, kept on the device for the comparison
    keep = list(range(0, len(images), 16))
    saved = {i: {k: sv.image(i)[k].clone() for k in KEYS} for i in keep}
    labels_b = {k: getattr(sv, k).clone() for k in ("rclass", "thclass", "nmax_eq", "status")}
    dev_labels = {k: torch.empty_like(getattr(sv, k)) for k in ("rclass", "thclass", "nmax_eq", "tau_tot", "status")}
    ws = run_sequential(images, arena, out_off, r_o=R_O, labels=dev_labels)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run_sequential(images, arena, out_off, r_o=R_O, workspace=ws, labels=dev_labels)
    torch.cuda.synchronize()
    wall_s = (time.perf_counter() - t0) / steps
    log(f"sequential plain launches: {wall_s * 1e3:.1f} ms/survey -> {samples / wall_s:.3e} samples/s")
    worst, nan_same = 0.0, True
    for i in keep:
        im = sv.image(i)
        for k in KEYS:
            b, p_ = saved[i][k], im[k]
            nan_same &= bool(torch.equal(torch.isnan(b), torch.isnan(p_)))
            inf = torch.isinf(p_)
            nan_same &= bool(torch.equal(b[inf], p_[inf]))   # +-inf at horizon ends must match exactly
            ok = torch.isfinite(p_)
            if ok.any():
                worst = max(worst, float(((b[ok] - p_[ok]).abs() / p_[ok].abs().clamp(min=1.0)).max()))
    labels_same = all(bool(torch.equal(labels_b[k], dev_labels[k])) for k in labels_b)
    log(f"batched vs sequential on {len(keep)} images: largest deviation {worst:.2e}, labels identical "
        f"{labels_same}, NaN and inf patterns equal {nan_same}")
    out = {"ms_per_survey": wall_b * 1e3, "samples_per_s": samples / wall_b, "event_ms": gpu_b * 1e3,
           "sequential_ms_per_survey": wall_s * 1e3, "sequential_samples_per_s": samples / wall_s,
           "speedup_vs_sequential": wall_s / wall_b, "largest_deviation_vs_sequential": worst,
           "compared_images": len(keep), "labels_identical": labels_same, "nonfinite_patterns_equal": nan_same,
           "all_items_done": bool(np.array_equal(done, total)), "images": len(images), "pixels": int(pix),
           "samples": int(samples), "tuning": {"samples_per_item": sv.tuning[0], "waves_per_cu": sv.tuning[1]},
           "work_items": sv.items, "autotune_key": key, "autotune_ms": tune_ms, "steps": steps}
    sv.close()
    del sv, saved, arena, ws
    torch.cuda.empty_cache()
    return out


BATCH_WORKLOAD = ("256-image survey: a~U(-0.99,0.99), th_o~U(1,89) deg, res from {64,128,256,512}^2, "
                  "ngeo~U{50..400}, default_rng(0); alpha,beta in [-8,8], r_o=1e4; writes r, theta, phi, t")


def bench_batch(args):
    f = batch_survey_figures(args.steps, args.warmup, retune=args.retune)
    line = {"metric": "batched survey geodesic samples/s (256 images, one persistent hipGraph launch)",
            "value": f["samples_per_s"], "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": f["ms_per_survey"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": BATCH_WORKLOAD}, "batch": f}
    print(json.dumps(line), flush=True)


def bench_equatorial(args):
    """Equatorial-emission image 1024^2 (a = first draw of default_rng(0)
    over [-0.99, 0.99], th_o = 60 deg, mbar 0..2, p = 2), and the gradient
    image (vector-Jacobian product over (a, th_o, p)) at 256^2, a = 0.7,
    th_o = 40 deg, against the plain forward image at that size."""
    import torch
    from kgeo_amd import _lib
    from kgeo_amd.equatorial import equatorial_image_device, equatorial_image_vjp_device

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    res = {}
    for name, (a, thd, n, w) in {"image_1024": (0.27118414089647946, 60.0, 1024, 12.0),
                                 "image_256": (0.7, 40.0, 256, 12.0)}.items():
        al = np.linspace(-w, w, n)
        A, B = np.meshgrid(al, al)
        A = torch.from_numpy(A.flatten()).to(dev)
        B = torch.from_numpy(B.flatten()).to(dev)
        th = thd * np.pi / 180.0
        for _ in range(args.warmup):
            equatorial_image_device(a, R_O, th, A, B, want_parts=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            equatorial_image_device(a, R_O, th, A, B, want_parts=False)
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / args.steps
        log(f"{name}: {res[name] * 1e3:.3f} ms/image")
        if n == 256:
            for _ in range(args.warmup):
                equatorial_image_vjp_device(a, R_O, th, A, B)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                equatorial_image_vjp_device(a, R_O, th, A, B)
            torch.cuda.synchronize()
            res["grad_256"] = (time.perf_counter() - t0) / args.steps
            log(f"grad_256: {res['grad_256'] * 1e3:.3f} ms/grad-image")
    line = {"metric": "equatorial images/s at 1024^2; grad-images/s at 256^2",
            "value": 1.0 / res["grad_256"], "unit": "grad-images/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["grad_256"] * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "d image / d (a, th_o, p) at 256^2, a=0.7, th_o=40 deg, p=2, mbar 0..2, "
                                   "alpha,beta in [-12,12]; forward images at 1024^2 (a=0.2712, th_o=60 deg) and "
                                   "256^2"},
            "images_per_s_1024": 1.0 / res["image_1024"], "images_per_s_256": 1.0 / res["image_256"],
            "grad_over_forward_time_256": res["grad_256"] / res["image_256"]}
    print(json.dumps(line), flush=True)


def grad_image_rate(steps=20, warmup=3):
    """The metric's second half: d(image)/d(a, th_o, p) at 256^2 (a = 0.7,
    th_o = 40 deg, p = 2, mbar 0..2, alpha, beta in [-12, 12]), timed on the
    host clock around `steps` calls after `warmup`."""
    import torch
    from kgeo_amd.equatorial import equatorial_image_vjp_device

    dev = torch.device("cuda", 0)
    al = np.linspace(-12.0, 12.0, 256)
    A, B = np.meshgrid(al, al)
    A = torch.from_numpy(A.flatten()).to(dev)
    B = torch.from_numpy(B.flatten()).to(dev)
    th = 40.0 * np.pi / 180.0
    for _ in range(warmup):
        equatorial_image_vjp_device(0.7, R_O, th, A, B)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        equatorial_image_vjp_device(0.7, R_O, th, A, B)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return {"value": 1.0 / dt, "unit": "grad-images/s", "ms_per_grad_image": dt * 1e3, "steps": steps,
            "workload": "d image / d (a, th_o, p) at 256^2, a=0.7, th_o=40 deg, p=2, mbar 0..2, "
                        "alpha,beta in [-12,12] (bench.py --mode equatorial for the 1024^2 image)"}


IMAGE_WORKLOAD = ("polarized make_image 800x800, subrings 0..1 (a=0.7, th_o=45 deg, r_o=inf, alpha,beta in "
                  "[-10,10]; general velocity fac_subkep=0.6 beta_r=0.7 beta_phi=0.8, bz_guess field, ring "
                  "emissivity, specind 1): the reference's tests/test_images.py case")


def polarized_image_figures(steps=5, warmup=2, cpu_pixels=2000):
    """kgeo_equatorial_stokes on the reference's image test (IMAGE_WORKLOAD):
    kernel time by HIP events on the launching stream, pixels x subrings per
    second, and the oracle's rate on a seeded pixel sample (one core)."""
    import torch
    from kgeo_amd import images
    from kgeo_amd import models as M

    a, th, nx = 0.7, 45.0 * np.pi / 180.0, 800
    A, B = images.pixel_grid(-10., 10., -10., 10., 20. / nx)
    dev = torch.device("cuda", 0)
    al, be = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    vel = M.Velocity("general", fac_subkep=0.6, beta_r=0.7, beta_phi=0.8)
    bf, em = M.Bfield("bz_guess", C=1.), M.Emissivity("ring")

    import ctypes
    from kgeo_amd import _lib
    from kgeo_amd.raytrace import _ptr, _stream_handle
    lib = _lib.load()
    model = M.emission_model(vel, bf, em, polarization=True, specind=1.)
    prm = _lib.make_params(a, np.inf, th, len(A), 1, 1.0)
    outs = [torch.empty((len(A), 2), device=dev, dtype=torch.float64) for _ in range(3)]
    stream = _stream_handle()

    def run():   # one kgeo_equatorial_stokes launch, arguments prepared once
        _lib.check(lib.kgeo_equatorial_stokes(prm, _ptr(al), _ptr(be), 1, ctypes.byref(model),
                                              *[_ptr(o) for o in outs], *([None] * 8), stream),
                   "kgeo_equatorial_stokes")
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for e0, e1 in evs:
        e0.record()
        run()
        e1.record()
    torch.cuda.synchronize()
    ref = images.stokes_device(a, np.inf, th, 1, al, be, emissivity=em, velocity=vel, bfield=bf, polarization=True,
                               specind=1., want=("I",))
    torch.cuda.synchronize()
    assert torch.equal(ref["I"], outs[0])
    ms = float(np.median([e0.elapsed_time(e1) for e0, e1 in evs]))
    units = len(A) * 2
    out = {"workload": IMAGE_WORKLOAD, "ms_per_image": ms, "pixel_subrings_per_s": units / (ms * 1e-3),
           "steps": steps}
    hw = read_profile_value("hardware_fp64_flops_per_launch", IMAGE_SUMMARY)
    if hw:
        tf = hw / (ms * 1e-3) / 1e12
        out["roofline"] = {"bound": "valu-fp64", "kernel": "kg_stokes", "achieved": tf, "peak": FP64_PEAK_TFLOPS,
                           "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFLOPS, "flops": "hardware FP64 count",
                           "hardware_fp64_flops_per_launch": hw, "traffic": 64.0 * len(A)}
    try:
        from oracle import kgeo_oracle_emission as E
        rng = np.random.default_rng(0)
        idx = rng.choice(len(A), cpu_pixels, replace=False)
        ev = E.Velocity("general", fac_subkep=0.6, beta_r=0.7, beta_phi=0.8)
        t0 = time.perf_counter()
        E.make_image_pixels(a, np.inf, th, 1, A[idx], B[idx], emissivity=E.Emissivity("ring"), velocity=ev,
                            bfield=E.Bfield("bz_guess", C=1.), polarization=True, specind=1.)
        dt = time.perf_counter() - t0
        out["cpu_oracle"] = {"pixel_subrings_per_s": cpu_pixels * 2 / dt, "cores": 1, "kind": "port",
                             "sample": f"{cpu_pixels} seeded pixels of the same image (numpy, vectorised)"}
    except ImportError:
        out["cpu_oracle"] = None
    log(f"polarized image: {ms:.3f} ms/image ({units / (ms * 1e-3):.3e} pixel-subrings/s)")
    return out


def bench_image(args):
    f = polarized_image_figures(args.steps, args.warmup)
    print(json.dumps({"metric": "polarized image pixel-subrings/s", "value": f["pixel_subrings_per_s"],
                      "unit": "pixel-subrings/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": f["ms_per_image"], "higher_is_better": True, "scaling": "weak",
                      "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                      "config": {"workload": IMAGE_WORKLOAD}, "image": f}), flush=True)


def bench_lensing(args):
    """The lensed-curve library (equatorial_lensing.py generate_library) on
    the GPU: the curves mbar = 0..5 of r = r_h for 8 spins x 89 inclinations
    (4,272 curves x 360 screen angles = 1,537,920 Brent root finds) in one
    kgeo_rho_of_req launch, timed by HIP events; the oracle's brentq (scipy,
    one core) on 360 of them for comparison."""
    import torch
    from kgeo_amd import lensing
    spins = np.linspace(0.01, 0.99, 8)
    incs = np.linspace(1, 89, 89)
    S, I = np.meshgrid(spins, incs, indexing="ij")
    sp = np.repeat(S.ravel(), 6)
    th = np.repeat(I.ravel() * np.pi / 180, 6)
    req = 1 + np.sqrt(1 - sp ** 2)
    mb = np.tile(np.arange(6, dtype=np.int32), S.size)
    vp = lensing.default_varphis()
    torch.cuda.set_device(0)
    for _ in range(args.warmup):
        lensing.rho_of_req_device(sp, th, req, mb, vp)
    torch.cuda.synchronize()
    ms = []
    for _ in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rho, status = lensing.rho_of_req_device(sp, th, req, mb, vp)
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    dt = float(np.median(ms))
    roots = len(sp) * len(vp)
    found = float((status == 0).float().mean())
    cpu = None
    try:
        from oracle import kgeo_oracle as O
        t0 = time.perf_counter()
        O.rho_of_req(float(sp[6 * 40 + 1]), float(th[6 * 40 + 1]), float(req[6 * 40 + 1]), 1, vp)
        cpu = {"roots_per_s": len(vp) / (time.perf_counter() - t0), "cores": 1, "kind": "port",
               "sample": "one curve (360 angles) of the same library, scipy brentq"}
    except (ImportError, ValueError, RuntimeError) as e:
        log(f"lensing cpu sample skipped: {e}")
    log(f"lensing: {roots} roots in {dt:.2f} ms -> {roots / (dt * 1e-3):.3e} roots/s ({found:.4f} found)")
    print(json.dumps({"metric": "lensed-curve root finds/s", "value": roots / (dt * 1e-3), "unit": "roots/s",
                      "ms_per_step": dt, "steps": args.steps, "warmup": args.warmup, "dtype": "f64",
                      "fraction_found": found, "cpu_baseline": cpu,
                      "config": {"workload": "r = r_h curves, mbar 0..5, 8 spins x 89 inclinations x 360 angles"}}),
          flush=True)


def bench_dropin(args):
    """The Python drop-in raytrace_ana at the raytrace workload, host arrays
    in and out: the kernels plus the PCIe transfers of alpha, beta and the six
    (ngeo, npix) outputs (tau, t, r, theta, phi, sigma: 6.3 GB) and the labels."""
    import torch
    from kgeo_amd.raytrace import raytrace_ana
    th = TH_DEG * np.pi / 180.0
    A, B = screen()
    torch.cuda.set_device(0)
    times = []
    for i in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        g = raytrace_ana(A_SPIN, [0, R_O, th, 0], [A, B], ngeo=NGEO, plotdata=False)
        dt = time.perf_counter() - t0
        if i >= args.warmup:
            times.append(dt)
        del g
    dt = float(np.median(times))
    samples = len(A) * NGEO
    log(f"drop-in raytrace_ana with host arrays: {dt * 1e3:.1f} ms -> {samples / dt:.3e} samples/s")
    print(json.dumps({"metric": "drop-in raytrace_ana samples/s, host arrays in and out (PCIe-inclusive)",
                      "value": samples / dt, "unit": "samples/s", "ms_per_call": dt * 1e3, "steps": args.steps,
                      "warmup": args.warmup, "bytes_to_host": samples * 6 * 8,
                      "config": {"workload": "raytrace_ana 512x512 ngeo=500 (a=0.94, th_o=17 deg, r_o=1e4)"}}),
          flush=True)


def bench_copy(args):
    """A plain device-to-device copy of known size (1 GiB read + 1 GiB
    written per launch): the reference point that scales the FETCH_SIZE and
    WRITE_SIZE counters into bytes for the traffic figure."""
    import torch
    dev = torch.device("cuda", 0)
    n = 1 << 27  # float64 elements = 1 GiB
    x = torch.ones(n, device=dev, dtype=torch.float64)
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"metric": "device copy GB/s (counter calibration)", "value": 2 * n * 8 / dt / 1e9,
                      "unit": "GB/s", "bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8,
                      "steps": args.steps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="leave the batched-survey figures out of the line")
    ap.add_argument("--mode", choices=("raytrace", "batch", "equatorial", "image", "dropin", "lensing", "copy"),
                    default="raytrace")
    ap.add_argument("--retune", action="store_true", help="batch mode: re-run the autotuner")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: measures one GPU only (--gpus 1); a multi-GPU run would shard the pixels, "
              "with no collectives on the path, and is not built", file=sys.stderr)
        sys.exit(2)
    if args.mode == "batch":
        return bench_batch(args)
    if args.mode == "equatorial":
        return bench_equatorial(args)
    if args.mode == "copy":
        return bench_copy(args)
    if args.mode == "image":
        return bench_image(args)
    if args.mode == "dropin":
        return bench_dropin(args)
    if args.mode == "lensing":
        return bench_lensing(args)

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
    wsb = _lib.workspace_bytes(npix, NGEO)
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

    chunks = int(lib.kgeo_sample_chunks(npix, NGEO))
    flops_s, flops_setup, ops = algorithmic_flops_per_sample(A, B, th, chunks)
    achieved_tf = flops_s * samples / (samp_ms * 1e-3) / 1e12
    bytes_per_launch = 32.0 * samples
    hbm_gbs = bytes_per_launch / (samp_ms * 1e-3) / 1e9
    roofline = {"bound": "valu-fp64", "kernel": "sample passes (kg_theta_* + kg_radial_*: heads, far, repair, walk, fixup)", "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved_tf / FP64_PEAK_TFLOPS,
                "traffic": read_profile_value("traffic_bytes_per_launch"),
                "flops_per_sample": flops_s, "algorithmic_bytes_per_launch": bytes_per_launch,
                "kernel_ms": samp_ms, "hbm_write_GBs": hbm_gbs, "hbm_frac": hbm_gbs / HBM_PEAK_GBS,
                "setup_kernel_ms": plan_ms, "setup_flops_per_pixel": flops_setup,
                "hardware_fp64_flops_per_launch": read_profile_value("hardware_fp64_flops_per_launch"),
                "ops_per_sample": {k: round(v, 3) for k, v in ops.items()}}

    grad = grad_image_rate()
    log(f"grad-images: {grad['ms_per_grad_image']:.3f} ms -> {grad['value']:.0f} grad-images/s")
    image = polarized_image_figures()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(th)

    batch = None
    if not args.no_batch:
        del out, ws
        torch.cuda.empty_cache()
        batch = batch_survey_figures(2, 1, candidates=BATCH_CANDIDATES)
        batch["workload"] = BATCH_WORKLOAD

    line = {"metric": METRIC, "value": value, "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "raytrace_ana 512x512 ngeo=500 (a=0.94, th_o=17 deg, r_o=1e4, "
                                   "alpha,beta in [-8,8]); writes r, theta, phi, t",
                       "npix": npix, "ngeo": NGEO, "parallelism": "single GPU, pixels x sample chunks"},
            "roofline": roofline, "cpu_baseline": cpu,
            "gpu_vs_cpu": (value / cpu["value"]) if cpu else None, "grad_images": grad, "polarized_image": image,
            "batch": batch,
            "nonfinite_pixels": nonfinite, "host": platform.node()}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
