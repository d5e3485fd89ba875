"""The batched survey at its workload: the seeded 256-image survey
(a ~ U(-0.99, 0.99), th_o ~ U(1, 89) deg, 64^2..512^2 pixels, ngeo 50..400;
20,537,344 pixels, about 5.0e9 samples) in one graph launch.

* every 16th image against a plain kgeo_raytrace_ana call on that image
  alone, into separate buffers (two full arenas would not fit in one GPU's
  HBM): labels bit-exact, NaN patterns equal, every coordinate within 4e-11
  (the two kernels inline the same per-pixel code; the compiler may contract
  it into FMAs differently, and a step at the edge of the walk's
  conditioning test may be walked in one and evaluated directly in the
  other);
* three images on 512 seeded pixels each against the oracle: labels
  bit-exact, coordinates within 1e-10 on the samples the oracle resolves at
  that level;
* every image's work items completed.
The reference has no batched form: a survey is a loop of raytrace_ana calls
(kerr_raytracing_ana.py:46-122)."""

import numpy as np
import pytest

from oracle import kgeo_oracle as O

from .kgeo_testlib import conditioning_mask, summarize

pytestmark = pytest.mark.gpu

R_O = 1.0e4


@pytest.fixture(scope="module")
def survey():
    import torch
    from kgeo_amd.batch import BatchSurvey, survey_draw, survey_images
    images = survey_images(survey_draw(256, 0))
    sv = BatchSurvey(images, r_o=R_O)
    sv.run()
    sv.stream.synchronize()
    torch.cuda.synchronize()
    print(f"\n[survey256] {len(images)} images, {int(sv.pix_off[-1])} pixels, {sv.samples} samples, "
          f"tuning {sv.tuning}", flush=True)
    yield images, sv
    sv.close()
    del sv
    torch.cuda.empty_cache()


def test_survey256_completes(survey):
    images, sv = survey
    done, total = sv.progress()
    assert np.array_equal(done, total) and int(total.sum()) == sv.items
    assert sv.samples > 4.9e9


def test_survey256_every_16th_image_equals_plain_call(survey):
    import torch
    from kgeo_amd.raytrace import raytrace_ana_device
    images, sv = survey
    worst = 0.0
    nbit = 0
    for k in range(0, len(images), 16):
        im = images[k]
        res = raytrace_ana_device(im["a"], R_O, im["th_o"], im["alpha"], im["beta"], im["ngeo"],
                                  want=("t", "r", "th", "ph"))
        torch.cuda.synchronize()
        got = sv.image(k)
        for key in ("rclass", "thclass", "nmax_eq", "status"):
            assert torch.equal(got[key], getattr(res, key).to(got[key].dtype)), (k, key)
        assert torch.equal(got["tau_tot"], res.tau_tot), k
        for key in ("t", "r", "th", "ph"):
            b, p = got[key], getattr(res, key)
            assert torch.equal(torch.isnan(b), torch.isnan(p)), (k, key)
            # +-inf (t, phi at the horizon end where r - r_+ rounds to 0) must match exactly
            inf = torch.isinf(p)
            assert torch.equal(b[inf], p[inf]), (k, key)
            ok = torch.isfinite(p)
            rel = (b[ok] - p[ok]).abs() / p[ok].abs().clamp(min=1.0)
            m = float(rel.max()) if rel.numel() else 0.0
            assert m < 4e-11, (k, key, m)
            worst = max(worst, m)
            nbit += int((b[ok] != p[ok]).sum())
        del res
        torch.cuda.empty_cache()
    print(f"\n[survey256 vs plain] 16 images: largest deviation {worst:.2e}, {nbit} values differ in the last bits")


@pytest.mark.parametrize("k", [5, 100, 200])
def test_survey256_images_match_oracle(survey, k):
    images, sv = survey
    im = images[k]
    got = {key: v.cpu().numpy() for key, v in sv.image(k).items()}
    idx = np.sort(np.random.default_rng(k).choice(len(im["alpha"]), 512, replace=False))
    a, th, A, B, G = im["a"], im["th_o"], im["alpha"][idx], im["beta"][idx], im["ngeo"]
    ref = O.raytrace_ana(a, R_O, th, A, B, G)
    assert np.array_equal(got["rclass"][idx].astype(int), ref["rclass"])
    assert np.array_equal(got["thclass"][idx].astype(int), ref["thclass"])
    assert np.array_equal(got["nmax_eq"][idx].astype(float), O.n_equatorial_crossings(a, th, A, B, ref["tau"][-1]))
    ill = conditioning_mask(a, R_O, th, A, B, ref["tau"], ref)
    for key in ("r", "th", "ph", "t"):
        s = summarize(got[key][:, idx], ref[key], key, ill)
        assert s["nan_mismatch"] == 0, (k, key, s)
        assert s["max_well"] < 1e-10, (k, key, s)
        print(f"\n[survey256 image {k}: a={a:.3f} th={np.rad2deg(th):.1f} {len(im['alpha'])} px ngeo={G}] "
              f"{key:3s} max {s['max']:.2e} p99 {s['p99']:.2e} max(well) {s['max_well']:.2e}")
