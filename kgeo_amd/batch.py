"""Batched survey: many ``raytrace_ana`` images in one graph launch.

The reference traces one image per ``raytrace_ana`` call
(kerr_raytracing_ana.py:46-122); a survey over spins, inclinations,
resolutions and ngeo is a Python loop of such calls.  ``BatchSurvey`` runs
the whole survey as one HIP graph: a setup kernel over every pixel of every
image, then for theta and r the plain path's passes (direct heads, long
steps, walk, fixups) as persistent kernels whose waves pull (image, 64-pixel
tile, unit) work items from one atomic queue per pass (kgeo_batch.hip).
Each image's results equal a plain ``raytrace_ana_device`` call on that
image to the last bits.

Outputs live in one arena per coordinate (t, r, th, ph; tau and sigma are
not stored): image k is the (ngeo_k, npix_k) block at ``out_off[k]``.
"""

import ctypes

import numpy as np
import torch

from . import _lib, autotune
from .raytrace import MAXTAUFRAC, _as_device_f64, _device, _ptr, ctypes_void

KEYS = ("t", "r", "th", "ph")


def survey_draw(n=256, seed=0):
    """The seeded survey: spin a ~ U(-0.99, 0.99), inclination ~ U(1, 89) deg,
    resolution from {64, 128, 256, 512} squared, ngeo ~ U{50..400}, drawn in
    that order from numpy's default_rng(seed)."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(-0.99, 0.99, n)
    th = rng.uniform(1.0, 89.0, n)
    res = rng.choice([64, 128, 256, 512], n)
    ngeo = rng.integers(50, 401, n)
    return [dict(a=float(a[k]), th_o=float(th[k] * np.pi / 180.0), res=int(res[k]), ngeo=int(ngeo[k]))
            for k in range(n)]


def square_grid(window, n):
    """alpha, beta = meshgrid(linspace(-w, w, n)) flattened, alpha fastest."""
    al = np.linspace(-window, window, n)
    A, B = np.meshgrid(al, al)
    return A.flatten(), B.flatten()


def survey_images(draw, window=8.0):
    """Image dicts (a, th_o, ngeo, alpha, beta) for a survey draw."""
    out = []
    for d in draw:
        A, B = square_grid(window, d["res"])
        out.append(dict(a=d["a"], th_o=d["th_o"], ngeo=d["ngeo"], alpha=A, beta=B))
    return out


def survey_size(images):
    """(total pixels, total samples) of a list of image dicts."""
    pix = sum(len(im["alpha"]) for im in images)
    smp = sum(len(im["alpha"]) * int(im["ngeo"]) for im in images)
    return pix, smp


def _check_phi_t(images, default=True):
    """Every image samples t, phi and sigma (as raytrace_ana does,
    kerr_raytracing_ana.py:97); a survey or image asking otherwise is refused."""
    for k, im in enumerate(images):
        if not bool(im.get("do_phi_and_t", default)):
            raise _lib.KgeoError(f"image {k}: do_phi_and_t=False is not supported by the batched survey "
                                 "(t, phi and sigma are always sampled)")


def _params_array(images, r_o, maxtaufrac):
    arr = (_lib.RayParams * len(images))()
    for k, im in enumerate(images):
        arr[k] = _lib.make_params(im["a"], r_o, im["th_o"], len(im["alpha"]), int(im["ngeo"]), maxtaufrac)
    return arr


def layout(images, r_o=1.0e4, maxtaufrac=MAXTAUFRAC):
    """(pix_off, out_off) int64 arrays of nimg + 1 entries (kgeo_batch_layout)."""
    lib = _lib.load()
    n = len(images)
    params = _params_array(images, r_o, maxtaufrac)
    pix = np.zeros(n + 1, dtype=np.int64)
    out = np.zeros(n + 1, dtype=np.int64)
    rc = lib.kgeo_batch_layout(params, n, pix.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
    _lib.check(rc, "kgeo_batch_layout")
    return pix, out


class BatchSurvey:
    """A planned survey: device arenas, workspace and the captured graph.

    images: sequence of dicts with a, th_o, ngeo, alpha, beta (arrays or
    tensors).  Every image samples t, phi and sigma; do_phi_and_t=False (for
    the survey or in an image's dict) raises KgeoError.
    tuning: (samples_per_item, waves_per_cu), or None to use the
    autotuner's cached choice for this survey's shape (default (64, 8) when
    nothing is cached).  ``arena`` may pass preallocated flat float64 tensors
    for t, r, th, ph (at least out_off[-1] elements each), e.g. to reuse one
    allocation across plans.
    """

    def __init__(self, images, r_o=1.0e4, maxtaufrac=MAXTAUFRAC, do_phi_and_t=True, tuning=None, arena=None):
        self.lib = _lib.load()
        self.dev = _device()
        self.images = list(images)
        n = len(self.images)
        if n == 0:
            raise ValueError("empty survey")
        self.r_o = float(r_o)
        _check_phi_t(self.images, do_phi_and_t)
        self.params = _params_array(self.images, self.r_o, maxtaufrac)
        self.pix_off, self.out_off = layout(self.images, self.r_o, maxtaufrac)
        self.npix = np.array([len(im["alpha"]) for im in self.images], dtype=np.int64)
        self.ngeo = np.array([int(im["ngeo"]) for im in self.images], dtype=np.int64)
        dev = self.dev
        total_pix = int(self.pix_off[-1])
        total_out = int(self.out_off[-1])
        self.alpha = torch.cat([_as_device_f64(im["alpha"], dev).reshape(-1) for im in self.images])
        self.beta = torch.cat([_as_device_f64(im["beta"], dev).reshape(-1) for im in self.images])
        arena = dict(arena or {})
        self.arena = {}
        for k in KEYS:
            a = arena.get(k)
            if a is None:
                a = torch.empty(total_out, device=dev, dtype=torch.float64)
            elif a.numel() < total_out or a.dtype != torch.float64 or not a.is_contiguous():
                raise ValueError(f"arena[{k!r}] must be contiguous float64 with >= {total_out} elements")
            self.arena[k] = a
        self.rclass = torch.empty(total_pix, device=dev, dtype=torch.int8)
        self.thclass = torch.empty(total_pix, device=dev, dtype=torch.int8)
        self.nmax_eq = torch.empty(total_pix, device=dev, dtype=torch.int32)
        self.tau_tot = torch.empty(total_pix, device=dev, dtype=torch.float64)
        self.status = torch.empty(total_pix, device=dev, dtype=torch.int32)
        wsb = int(self.lib.kgeo_batch_workspace_bytes(self.params, n))
        if wsb == 0:
            _lib.check(-1, "kgeo_batch_workspace_bytes")
        self.workspace = torch.empty(wsb, device=dev, dtype=torch.uint8)
        if tuning is None:
            tuning = autotune.lookup(self.shape_key()) or autotune.DEFAULT_BATCH
        self.tuning = tuple(int(x) for x in tuning)
        tn = _lib.BatchTuning(self.tuning[0], self.tuning[1], (ctypes.c_int32 * 2)(0, 0))
        # graph capture needs a non-default stream; default priority
        self.stream = torch.cuda.Stream(device=dev)
        handle = ctypes.c_void_p()
        torch.cuda.current_stream().synchronize()
        rc = self.lib.kgeo_batch_plan(self.params, n, ctypes.byref(tn), _ptr(self.alpha), _ptr(self.beta),
                                      _ptr(self.arena["t"]), _ptr(self.arena["r"]), _ptr(self.arena["th"]),
                                      _ptr(self.arena["ph"]), _ptr(self.rclass), _ptr(self.thclass),
                                      _ptr(self.nmax_eq), _ptr(self.tau_tot), _ptr(self.status),
                                      _ptr(self.workspace), wsb, ctypes_void(self.stream.cuda_stream),
                                      ctypes.byref(handle))
        _lib.check(rc, "kgeo_batch_plan")
        self.handle = handle

    def shape_key(self):
        return autotune.batch_key(self.npix, self.ngeo)

    @property
    def items(self):
        return int(self.lib.kgeo_batch_items(self.handle))

    @property
    def samples(self):
        return int(self.out_off[-1])

    def run(self, stream=None):
        """Replay the survey graph on `stream` (default: the survey's own
        stream), ordered after the work already queued on torch's current
        stream (writes to alpha, beta or the arenas) and before any work queued
        on it afterwards (reads of the arenas through image()).  Returns the
        stream the graph ran on."""
        s = stream if stream is not None else self.stream
        cur = torch.cuda.current_stream(self.dev)
        if s != cur:
            s.wait_stream(cur)
        rc = self.lib.kgeo_batch_run(self.handle, ctypes_void(s.cuda_stream))
        _lib.check(rc, "kgeo_batch_run")
        if s != cur:
            # the caching allocator must not hand these to other work before the replay ends
            for t in [self.workspace, self.alpha, self.beta] + [a for a in self.arena.values() if a is not None]:
                t.record_stream(s)
            cur.wait_stream(s)
        return s

    def progress(self):
        """(done, total) work items per image of the last run (synchronise first)."""
        n = len(self.images)
        done = np.zeros(n, dtype=np.uint32)
        total = np.zeros(n, dtype=np.uint32)
        rc = self.lib.kgeo_batch_progress(self.handle, done.ctypes.data_as(ctypes.c_void_p),
                                          total.ctypes.data_as(ctypes.c_void_p))
        _lib.check(rc, "kgeo_batch_progress")
        return done, total

    def image(self, k):
        """Device views of image k: t, r, th, ph as (ngeo, npix) and its per-pixel labels."""
        o0, o1 = int(self.out_off[k]), int(self.out_off[k + 1])
        p0, p1 = int(self.pix_off[k]), int(self.pix_off[k + 1])
        G, P = int(self.ngeo[k]), int(self.npix[k])
        out = {}
        for key in KEYS:
            a = self.arena[key]
            out[key] = None if a is None else a[o0:o1].view(G, P)
        for key in ("rclass", "thclass", "nmax_eq", "tau_tot", "status"):
            out[key] = getattr(self, key)[p0:p1]
        return out

    def close(self):
        if getattr(self, "handle", None) is not None:
            self.lib.kgeo_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_sequential(images, arena, out_off, r_o=1.0e4, maxtaufrac=MAXTAUFRAC, workspace=None, labels=None):
    """The same survey as len(images) plain kgeo_raytrace_ana calls, one per
    image, writing into the same arena layout (tau and sigma not stored).
    Asynchronous on torch's current stream."""
    lib = _lib.load()
    dev = _device()
    maxpix = max(len(im["alpha"]) for im in images)
    wsb = _lib.workspace_bytes(maxpix, max(int(im["ngeo"]) for im in images))
    if workspace is None or workspace.numel() < wsb:
        workspace = torch.empty(wsb, device=dev, dtype=torch.uint8)
    stream = ctypes_void(torch.cuda.current_stream().cuda_stream)
    dev_ab = [(_as_device_f64(im["alpha"], dev).reshape(-1), _as_device_f64(im["beta"], dev).reshape(-1))
              for im in images]
    labels = labels or {}
    pix0 = 0
    for k, im in enumerate(images):
        P, G = len(im["alpha"]), int(im["ngeo"])
        p = _lib.make_params(im["a"], r_o, im["th_o"], P, G, maxtaufrac)
        o0 = int(out_off[k])

        def view(key):
            a = arena.get(key)
            return None if a is None else a[o0:o0 + P * G]

        def lab(key):
            a = labels.get(key)
            return None if a is None else a[pix0:pix0 + P]

        al, be = dev_ab[k]
        rc = lib.kgeo_raytrace_ana(p, _ptr(al), _ptr(be), None, _ptr(view("t")), _ptr(view("r")), _ptr(view("th")),
                                   _ptr(view("ph")), None, _ptr(lab("rclass")), _ptr(lab("thclass")),
                                   _ptr(lab("nmax_eq")), _ptr(lab("tau_tot")), _ptr(lab("status")),
                                   _ptr(workspace), wsb, stream)
        _lib.check(rc, "kgeo_raytrace_ana")
        pix0 += P
    return workspace
