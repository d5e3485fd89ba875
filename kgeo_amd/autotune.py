"""Autotuner for the batched survey's launch shape.

Two knobs (kgeo_batch_tuning in include/kgeo.h):
  samples_per_item  Mino-time samples per walk work item; a walk item is
                    one 64-pixel tile x this many samples, so it sets how
                    many lanes share one pixel's ngeo range
                    (ceil(ngeo / this)): small images need several to fill
                    256 CUs, large ones want long runs per lane to amortise
                    the per-lane set-up (plan slots, step record).
  waves_per_cu      size of each pass's persistent grid (CUs x this waves).

Choices are cached in a JSON file keyed by (gfx arch, log2 pixel-count
bucket, log2 ngeo bucket) of the survey.  The cache lives at
$KGEO_AUTOTUNE_CACHE, else $XDG_CACHE_HOME/kgeo_amd/autotune.json, else
~/.cache/kgeo_amd/autotune.json; a cache that cannot be read or written is
ignored.
"""

import json
import math
import os

import numpy as np

DEFAULT_BATCH = (64, 8)
CANDIDATES = [(s, w) for s in (16, 32, 64, 128, 256) for w in (4, 8, 16)]


def cache_path():
    p = os.environ.get("KGEO_AUTOTUNE_CACHE")
    if p:
        return p
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "kgeo_amd", "autotune.json")


def _arch():
    try:
        import torch
        if torch.cuda.is_available():
            name = getattr(torch.cuda.get_device_properties(0), "gcnArchName", "") or "gpu"
            return name.split(":")[0]
    except Exception:
        pass
    return "cpu"


def _bucket(x):
    return int(math.floor(math.log2(max(float(x), 1.0))))


def batch_key(npix, ngeo, arch=None):
    """Cache key of a survey: arch, log2 bucket of the total pixel count, of
    the mean pixels per image and of the pixel-weighted mean ngeo."""
    npix = np.asarray(npix, dtype=np.float64)
    ngeo = np.asarray(ngeo, dtype=np.float64)
    tot = npix.sum()
    mean_ngeo = float((npix * ngeo).sum() / max(tot, 1.0))
    return f"{arch or _arch()}|batch|pix{_bucket(tot)}|img{_bucket(tot / max(len(npix), 1))}|ngeo{_bucket(mean_ngeo)}"


def _load():
    try:
        with open(cache_path()) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except Exception:
        return {}


def lookup(key):
    e = _load().get(key)
    if isinstance(e, dict) and "choice" in e:
        return tuple(int(x) for x in e["choice"])
    return None


def store(key, choice, timings):
    d = _load()
    d[key] = {"choice": [int(x) for x in choice],
              "timings_ms": {f"{s}x{w}": round(float(t), 4) for (s, w), t in timings.items()}}
    try:
        path = cache_path()
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def tune_batch(images, candidates=None, reps=2, r_o=1.0e4, key=None, arena=None, log=None):
    """Time each candidate (samples_per_item, waves_per_cu) on `images` (one
    warm-up replay, then the best of `reps` replays), store the fastest under
    `key` (default: the key of `images`) and return (choice, timings_ms).
    `images` may be a representative subset of a larger survey; pass that
    survey's key to file the result under it."""
    import torch

    from .batch import BatchSurvey

    cands = list(candidates or CANDIDATES)
    timings = {}
    key_used = key
    for c in cands:
        sv = BatchSurvey(images, r_o=r_o, tuning=c, arena=arena)
        arena = sv.arena
        sv.run()
        sv.stream.synchronize()
        best = float("inf")
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(sv.stream)
            sv.run()
            e1.record(sv.stream)
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        timings[c] = best
        if log:
            log(f"autotune {c}: {best:.3f} ms")
        key_used = key_used or sv.shape_key()
        sv.close()
        del sv
    choice = min(timings, key=timings.get)
    store(key_used, choice, timings)
    return choice, timings
