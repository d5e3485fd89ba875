"""Geodesics files (kerr_raytracing_utils.py:201-241): savegeos writes, and
loadgeos reads, HDF5 files with the reference's datasets.  h5py is not
installed here, so both go through kgeo_amd.h5lite; the HDF5 C library
(through ctypes) is the independent check in both directions: it reads what
h5lite writes, and h5lite reads what the library writes with its default
layout (the layout h5py's create_dataset produces), including group B-trees
split over several symbol-table nodes and a level-1 B-tree.  plotgeos is run
headless on oracle rays and its line colours checked against the reference's
labelling rule."""

import os

import numpy as np
import pytest

from kgeo_amd import h5lite
from kgeo_amd.geodesics import Geodesics, loadgeos
from oracle import kgeo_oracle as O

from . import h5capi

needs_lib = pytest.mark.skipif(h5capi.find() is None, reason="no libhdf5 to check against")

KEYS = ("spin", "inc", "r_o", "alpha", "beta", "t", "r", "theta", "phi", "affine", "mino")


def _sample(seed=0, ngeo=9, npix=13):
    rng = np.random.default_rng(seed)
    d = {"spin": 0.94, "inc": 17 * np.pi / 180, "r_o": np.inf, "alpha": rng.normal(size=npix),
         "beta": rng.normal(size=npix)}
    for k in ("t", "r", "theta", "phi", "affine", "mino"):
        d[k] = rng.normal(size=(ngeo, npix))
    d["t"][-1, :3] = [np.inf, -np.inf, np.nan]   # horizon ends and undefined samples survive
    return d


def _same(x, y):
    x, y = np.asarray(x), np.asarray(y)
    return x.shape == y.shape and np.array_equal(x, y, equal_nan=True)


@needs_lib
def test_library_reads_h5lite_files(tmp_path):
    d = _sample()
    path = str(tmp_path / "w.h5")
    h5lite.write(path, d)
    got = h5capi.read(path)
    assert sorted(got) == sorted(d)
    for k in d:
        assert _same(got[k], d[k]), k
    assert got["spin"].shape == ()   # scalars are rank 0, as h5py writes them


@needs_lib
@pytest.mark.parametrize("ndatasets", [11, 40, 300])
def test_h5lite_reads_library_files(tmp_path, ndatasets):
    rng = np.random.default_rng(ndatasets)
    d = dict(_sample(ndatasets))
    for i in range(ndatasets - len(d)):
        d[f"extra{i:03d}"] = rng.normal(size=(i % 4 + 1, 3)) if i % 3 else rng.normal()
    path = str(tmp_path / "lib.h5")
    h5capi.write(path, d)
    got = h5lite.read(path)
    assert sorted(got) == sorted(d)
    for k in d:
        assert _same(got[k], d[k]), k


def test_h5lite_round_trip_and_limits(tmp_path):
    d = _sample(3)
    d["empty"] = np.zeros((0, 4))
    path = str(tmp_path / "rt.h5")
    h5lite.write(path, d)
    got = h5lite.read(path)
    for k in d:
        assert _same(got[k], d[k]), k
    with pytest.raises(ValueError):
        h5lite.write(str(tmp_path / "big.h5"), {f"k{i}": 0.0 for i in range(2 * h5lite.LEAF_K + 1)})
    bad = tmp_path / "bad.h5"
    bad.write_bytes(b"not hdf5" * 20)
    with pytest.raises(ValueError):
        h5lite.read(str(bad))


def _geos(d):
    return Geodesics(d["spin"], [0, d["r_o"], d["inc"], 0], [d["alpha"], d["beta"]], d["mino"], d["affine"],
                     [d["t"], d["r"], d["theta"], d["phi"]])


def test_savegeos_default_name_and_loadgeos(tmp_path):
    d = _sample(5)
    g = _geos(d)
    fname = g.savegeos(path=str(tmp_path) + "/")
    assert os.path.basename(fname) == "a%0.2f_th%0.2f_geo.h5" % (0.94, 17.0)
    h = loadgeos(fname)
    assert h.a == g.a and h.r_o == g.r_o and h.th_o == g.th_o
    for name in ("alpha", "beta", "t_s", "r_s", "th_s", "ph_s", "sig_s", "tausteps"):
        assert _same(getattr(h, name), getattr(g, name)), name
    other = str(tmp_path / "named.h5")
    assert g.savegeos(outfile=other) == other and os.path.isfile(other)


@needs_lib
def test_savegeos_file_has_reference_datasets(tmp_path):
    g = _geos(_sample(6))
    got = h5capi.read(g.savegeos(outfile=str(tmp_path / "g.h5")))
    assert sorted(got) == sorted(KEYS)
    assert _same(got["mino"], g.tausteps) and _same(got["affine"], g.sig_s) and _same(got["theta"], g.th_s)


def test_plotgeos_colours_follow_crossing_counts():
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.colors as mcolors
    import matplotlib.pyplot as plt

    a, th_o = 0.9, 60 * np.pi / 180
    alpha = np.repeat(np.linspace(-7, 7, 6), 6)
    beta = np.tile(np.linspace(-7, 7, 6), 6)
    res = O.raytrace_ana(a, np.inf, th_o, alpha, beta, 64)
    g = Geodesics(a, [0, np.inf, th_o, 0], [alpha, beta], res["tau"], res["sig"],
                  [res["t"], res["r"], res["th"], res["ph"]])
    nmax = O.n_equatorial_crossings(a, th_o, alpha, beta, res["tau_tot"])
    g._device_labels = {"nmax_eq": nmax}
    fig = g.plotgeos(ngeoplot=None)
    colors = ["dimgrey", "b", "g", "orange", "r", "m", "c", "y"]
    drawn = [ln for ln in fig.axes[0].get_lines() if ln.get_linestyle() != "--"]
    label = nmax + 1
    expect = []
    for j in range(min(int(np.nanmax(label)) + 1, len(colors))):
        mask = ((label == 0) | (label == -1)) if j == 0 else (label == j)
        expect += [mcolors.to_rgba(colors[j])] * int(mask.sum())
    assert len(drawn) == len(expect) > 0
    assert [mcolors.to_rgba(ln.get_color()) for ln in drawn] == expect
    plt.close(fig)
