"""Host-side pieces of the batched survey (no GPU calls): the seeded survey
draw and its size, the C-ABI layout and workspace arithmetic, argument
errors that the library rejects before touching the GPU, and the
autotuner's cache."""

import ctypes

import numpy as np
import pytest

from kgeo_amd import _lib, autotune
from kgeo_amd.batch import layout, survey_draw, survey_images, survey_size


def test_survey_draw_is_the_seeded_configuration():
    d = survey_draw(256, seed=0)
    assert len(d) == 256
    a = np.array([x["a"] for x in d])
    th = np.rad2deg([x["th_o"] for x in d])
    res = np.array([x["res"] for x in d])
    ngeo = np.array([x["ngeo"] for x in d])
    assert a.min() >= -0.99 and a.max() < 0.99
    assert th.min() >= 1.0 and th.max() < 89.0
    assert set(res) <= {64, 128, 256, 512} and ngeo.min() >= 50 and ngeo.max() <= 400
    vals, counts = np.unique(res, return_counts=True)
    assert dict(zip(vals.tolist(), counts.tolist())) == {64: 62, 128: 70, 256: 68, 512: 56}
    pix = int((res.astype(np.int64) ** 2).sum())
    smp = int((res.astype(np.int64) ** 2 * ngeo).sum())
    assert pix == 20537344
    assert 5.0e9 < smp < 5.05e9
    # the first draw equals the equatorial-image spin drawn from the same seed
    assert a[0] == 0.27118414089647946


def test_layout_offsets_and_workspace():
    images = survey_images(survey_draw(6, seed=3), window=5.0)
    pix, out = layout(images)
    npix = [len(im["alpha"]) for im in images]
    ngeo = [im["ngeo"] for im in images]
    assert pix.tolist() == np.concatenate([[0], np.cumsum(npix)]).tolist()
    assert out.tolist() == np.concatenate([[0], np.cumsum(np.multiply(npix, ngeo))]).tolist()
    assert survey_size(images) == (pix[-1], out[-1])
    lib = _lib.load()
    ragged = [dict(a=0.5, th_o=1.0, ngeo=10, alpha=np.zeros(n), beta=np.zeros(n)) for n in (1, 63, 64, 65, 1000)]
    params = (_lib.RayParams * len(ragged))(*[_lib.make_params(0.5, 1e4, 1.0, len(im["alpha"]), 10, 1.0)
                                              for im in ragged])
    slots = sum((n + 63) // 64 * 64 for n in (1, 63, 64, 65, 1000))
    plan_slots = (_lib.workspace_bytes(1000) - 256) // (1000 * 8)
    # per pixel slot at ngeo = 10: the plan, theta records (ceil(10/32) + 3 of 6
    # fields), radial records (ceil(10/16) + 3 of 13 fields), two queue items
    # (the ceil(10/8) ranges)
    core = slots * 8 * (plan_slots + 4 * 6 + 4 * 13 + 2)
    ws = lib.kgeo_batch_workspace_bytes(params, len(ragged))
    assert ws >= core and ws - core < 4096


def test_batch_argument_errors_without_gpu():
    lib = _lib.load()
    assert lib.kgeo_batch_workspace_bytes(None, 0) == 0
    bad = (_lib.RayParams * 1)(_lib.make_params(0.5, 1e4, 1.0, 0, 10, 1.0))
    assert lib.kgeo_batch_workspace_bytes(bad, 1) == 0
    assert lib.kgeo_batch_layout(bad, 1, None, None) == -1
    assert b"npix" in lib.kgeo_batch_last_error()
    good = (_lib.RayParams * 1)(_lib.make_params(0.5, 1e4, 1.0, 64, 10, 1.0))
    h = ctypes.c_void_p()
    tn = _lib.BatchTuning(64, 8, (ctypes.c_int32 * 2)(0, 0))
    fake = ctypes.c_void_p(256)
    # NULL stream: graph capture needs a real stream; rejected before any HIP call
    rc = lib.kgeo_batch_plan(good, 1, ctypes.byref(tn), fake, fake, fake, fake, fake, fake, None, None, None, None,
                             None, fake, 1 << 30, None, ctypes.byref(h))
    assert rc == -1 and not h.value
    assert lib.kgeo_batch_run(None, None) == -1
    assert lib.kgeo_batch_items(None) == -1
    assert lib.kgeo_batch_destroy(None) == 0
    with pytest.raises(_lib.KgeoError):
        _lib.check(lib.kgeo_batch_run(None, None), "kgeo_batch_run")


def test_autotune_cache_roundtrip(tmp_path, monkeypatch):
    monkeypatch.setenv("KGEO_AUTOTUNE_CACHE", str(tmp_path / "sub" / "t.json"))
    key = autotune.batch_key([4096, 16384, 262144], [100, 300, 50], arch="gfx950")
    assert key.startswith("gfx950|batch|")
    assert autotune.lookup(key) is None
    autotune.store(key, (32, 8), {(32, 8): 1.5, (64, 8): 2.0})
    assert autotune.lookup(key) == (32, 8)
    # buckets: same key for nearby shapes, different for a 4x larger survey
    assert autotune.batch_key([4100, 16384, 262144], [100, 300, 50], arch="gfx950") == key
    assert autotune.batch_key([4 * 4096, 4 * 16384, 4 * 262144], [100, 300, 50], arch="gfx950") != key
    (tmp_path / "sub" / "t.json").write_text("not json")
    assert autotune.lookup(key) is None
