"""GPU tests of the batched survey (persistent kernel in one HIP graph).

Each image of a batch must equal a plain kgeo_raytrace_ana call on that
image alone: integer labels bit-exact and every coordinate equal to the
walk's rounding (the batch kernel runs the same per-pixel code; only the
work partition and the compiled instruction selection differ), and the plain path in turn matches the oracle (checked
again here on three images with the same 1e-10 bar as the single-image
tests).  Replays and every work partition of the batch itself are
bit-identical.
"""

import numpy as np
import pytest

from oracle import kgeo_oracle as O

from .kgeo_testlib import conditioning_mask, grid, summarize

pytestmark = pytest.mark.gpu

R_O = 1.0e4


def _line(n, w):
    al = np.linspace(-w, w, n)
    return al, 0.37 * al[::-1] + 0.11


def _images():
    """Mixed survey: ragged pixel counts (not multiples of 64), ngeo 1 and 2,
    a = 0 and |a| < 1e-4, retrograde spin, near-polar and near-equatorial
    observers, one southern-hemisphere observer."""
    ims = []

    def add(a, th_deg, AB, ngeo):
        ims.append(dict(a=a, th_o=float(np.deg2rad(th_deg)), alpha=AB[0], beta=AB[1], ngeo=ngeo))

    add(0.94, 17.0, grid(8.0, 32), 60)
    add(-0.6, 70.0, grid(7.0, 30), 37)
    add(0.0, 45.0, grid(6.0, 20), 1)
    add(5.0e-5, 30.0, grid(9.0, 16), 2)
    add(0.99, 88.0, grid(8.0, 24), 130)
    add(-0.3, 3.0, grid(10.0, 40), 7)
    add(0.5, 120.0, _line(1000, 7.0), 45)
    add(0.8, 60.0, _line(5, 4.0), 3)
    return ims


def _host(t):
    return None if t is None else t.cpu().numpy()


def _batch(images, tuning=(64, 8)):
    import torch
    from kgeo_amd.batch import BatchSurvey
    sv = BatchSurvey(images, r_o=R_O, tuning=tuning)
    sv.run()
    sv.stream.synchronize()
    torch.cuda.synchronize()
    return sv


def _plain(images, out_off):
    import torch
    from kgeo_amd.batch import KEYS, run_sequential
    dev = torch.device("cuda", torch.cuda.current_device())
    arena = {k: torch.full((int(out_off[-1]),), float("nan"), device=dev, dtype=torch.float64) for k in KEYS}
    tot = sum(len(im["alpha"]) for im in images)
    labels = {"rclass": torch.empty(tot, device=dev, dtype=torch.int8),
              "thclass": torch.empty(tot, device=dev, dtype=torch.int8),
              "nmax_eq": torch.empty(tot, device=dev, dtype=torch.int32),
              "tau_tot": torch.empty(tot, device=dev, dtype=torch.float64),
              "status": torch.empty(tot, device=dev, dtype=torch.int32)}
    run_sequential(images, arena, out_off, r_o=R_O, labels=labels)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in arena.items()}, {k: v.cpu().numpy() for k, v in labels.items()}


def _same(x, y):
    """bit-identical, NaN payloads aside"""
    return np.array_equal(x, y, equal_nan=True)


def test_batch_equals_plain_launches():
    """Same labels bit for bit; coordinates equal to the walk's rounding (the
    two kernels inline the same per-pixel code, but the compiler may contract
    it into FMAs differently, and a step on the edge of the walk's
    conditioning test may then be walked in one kernel and evaluated directly
    in the other: both within 2e-11 of the direct evaluation), and every
    image's work items completed."""
    images = _images()
    sv = _batch(images)
    arena_p, labels_p = _plain(images, sv.out_off)
    arena_b = {k: _host(v) for k, v in sv.arena.items()}
    for k in ("rclass", "thclass", "nmax_eq", "tau_tot", "status"):
        assert _same(_host(getattr(sv, k)), labels_p[k]), k
    nbit = 0
    for k in arena_b:
        assert np.array_equal(np.isnan(arena_b[k]), np.isnan(arena_p[k])), k
        inf = np.isinf(arena_p[k])
        assert np.array_equal(arena_b[k][inf], arena_p[k][inf]), k
        ok = np.isfinite(arena_p[k])
        rel = np.abs(arena_b[k][ok] - arena_p[k][ok]) / np.maximum(np.abs(arena_p[k][ok]), 1.0)
        assert rel.max(initial=0.0) < 4e-11, (k, float(rel.max()))
        nbit += int((arena_b[k][ok] != arena_p[k][ok]).sum())
    done, total = sv.progress()
    assert np.array_equal(done, total) and total.sum() == sv.items
    print(f"\n[batch vs plain] {len(images)} images, {sv.samples} samples, {sv.items} work items: "
          f"{nbit} values differ in the last bits")
    sv.close()


def test_batch_refuses_surveys_without_phi_t():
    """Every image samples t, phi and sigma, as raytrace_ana does
    (kerr_raytracing_ana.py:97): a survey, or one image, asking for
    do_phi_and_t=False is refused before anything is allocated or launched."""
    from kgeo_amd import _lib
    from kgeo_amd.batch import BatchSurvey
    images = _images()
    with pytest.raises(_lib.KgeoError, match="do_phi_and_t=False"):
        BatchSurvey(images, r_o=R_O, do_phi_and_t=False)
    images[3]["do_phi_and_t"] = False
    with pytest.raises(_lib.KgeoError, match="image 3"):
        BatchSurvey(images, r_o=R_O)


@pytest.mark.skip(reason="surveys without phi and t are refused (KgeoError, see "
                         "test_batch_refuses_surveys_without_phi_t): a batch kernel variant with per-survey "
                         "phi/t instantiations ended a GPU run in a fault whose cause was not shown, so the "
                         "no-phi/t kernels were taken out instead of being run")
def test_batch_without_phi_t_and_mixed_surveys():
    """Surveys where no image, or only some images, want phi and t: r and
    theta of every image, and t and phi of the images that want them, equal
    plain calls with the same per-image flag to the walk's rounding."""
    for mixed in (False, True):
        images = _images()
        for k, im in enumerate(images):
            im["do_phi_and_t"] = mixed and k % 2 == 0
        sv = _batch(images)
        arena_p, labels_p = _plain(images, sv.out_off)
        for k in ("rclass", "thclass", "nmax_eq", "status"):
            assert _same(_host(getattr(sv, k)), labels_p[k]), (mixed, k)
        for key, a_b in sv.arena.items():
            a_b = _host(a_b)
            for k, im in enumerate(images):
                if key in ("t", "ph") and not im["do_phi_and_t"]:
                    continue
                o0, o1 = int(sv.out_off[k]), int(sv.out_off[k + 1])
                x, y = a_b[o0:o1], arena_p[key][o0:o1]
                assert np.array_equal(np.isnan(x), np.isnan(y)), (mixed, key, k)
                ok = ~np.isnan(y)
                rel = np.abs(x[ok] - y[ok]) / np.maximum(np.abs(y[ok]), 1.0)
                assert rel.max(initial=0.0) < 4e-11, (mixed, key, k, float(rel.max()))
        sv.close()


@pytest.mark.parametrize("k", [0, 1, 6])
def test_batch_images_match_oracle(k):
    images = _images()
    sv = _batch(images)
    im = images[k]
    got = {key: _host(v) for key, v in sv.image(k).items()}
    a, th, A, B, G = im["a"], im["th_o"], im["alpha"], im["beta"], im["ngeo"]
    ref = O.raytrace_ana(a, R_O, th, A, B, G)
    assert np.array_equal(got["rclass"].astype(int), ref["rclass"])
    assert np.array_equal(got["thclass"].astype(int), ref["thclass"])
    assert np.array_equal(got["nmax_eq"].astype(float), O.n_equatorial_crossings(a, th, A, B, ref["tau"][-1]))
    ill = conditioning_mask(a, R_O, th, A, B, ref["tau"], ref)
    for key in ("r", "th", "ph", "t"):
        s = summarize(got[key], ref[key], key, ill)
        assert s["nan_mismatch"] == 0, (k, key, s)
        assert s["max_well"] < 1e-10, (k, key, s)
        print(f"\n[batch image {k}] {key:3s} max {s['max']:.2e} p99 {s['p99']:.2e} max(well) {s['max_well']:.2e}")
    sv.close()


def test_batch_replays_and_tunings_agree():
    """Graph replays with eager work between them, and every work partition
    (samples_per_item, waves_per_cu), give bit-identical outputs."""
    import torch
    images = _images()
    sv = _batch(images)
    first = {k: _host(v) for k, v in sv.arena.items()}
    for rep in range(3):
        for v in sv.arena.values():
            v.fill_(float("nan"))  # eager work between replays; every replay must rewrite everything
        torch.cuda.synchronize()
        sv.run()
        sv.stream.synchronize()
        for k, v in sv.arena.items():
            assert _same(_host(v), first[k]), (rep, k)
    sv.close()
    for tuning in ((1, 1), (16, 4), (1000, 64), (7, 3)):
        sv2 = _batch(images, tuning)
        for k, v in sv2.arena.items():
            assert _same(_host(v), first[k]), (tuning, k)
        sv2.close()


def test_batch_autotuner_picks_and_caches(tmp_path, monkeypatch):
    from kgeo_amd import autotune
    monkeypatch.setenv("KGEO_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    images = _images()
    choice, timings = autotune.tune_batch(images, candidates=[(16, 4), (64, 8)], reps=1)
    assert choice in timings and len(timings) == 2
    from kgeo_amd.batch import BatchSurvey
    sv = BatchSurvey(images, r_o=R_O)
    assert autotune.lookup(sv.shape_key()) == choice
    assert sv.tuning == choice
    sv.close()
