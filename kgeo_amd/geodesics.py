"""Result container with the reference's attribute names
(kgeo.kerr_raytracing_utils.Geodesics, kerr_raytracing_utils.py:15-93)."""

import numpy as np


class Geodesics(object):
    """Geodesic samples of one raytrace.

    mino_times (= tausteps), affine_times (= sig_s) and geo_coords = [t, r, th, ph]
    are (ngeo, npix) float64 arrays, exactly as the reference returns them.
    """

    def __init__(self, a, observer_coords, image_coords, mino_times, affine_times, geo_coords):
        self.a = a
        self.observer_coords = observer_coords
        self.image_coords = image_coords
        self.mino_times = mino_times
        self.affine_times = affine_times
        self.geo_coords = geo_coords
        self._device_labels = None

    # observer (kerr_raytracing_utils.py:29-41)
    @property
    def t_o(self):
        return self.observer_coords[0]

    @property
    def r_o(self):
        return self.observer_coords[1]

    @property
    def th_o(self):
        return self.observer_coords[2]

    @property
    def ph_o(self):
        return self.observer_coords[3]

    # image (:43-58)
    @property
    def alpha(self):
        return self.image_coords[0]

    @property
    def beta(self):
        return self.image_coords[1]

    @property
    def npix(self):
        return len(self.alpha)

    @property
    def lam(self):
        return -self.alpha * np.sin(self.th_o)

    @property
    def eta(self):
        return (self.alpha ** 2 - self.a ** 2) * np.cos(self.th_o) ** 2 + self.beta ** 2

    @property
    def nmax_eq(self):
        """Number of equatorial crossings at the last Mino time (:63-67); computed
        on the GPU alongside the rays."""
        if self._device_labels is not None:
            return self._device_labels["nmax_eq"].astype(float)
        from .counts import n_equatorial_crossings
        return n_equatorial_crossings(self.a, self.th_o, self.alpha, self.beta, self.tautot)

    @property
    def n_poloidal(self):
        """Fractional number of poloidal orbits at every sample (:59-62)."""
        from .counts import n_poloidal_orbits
        return n_poloidal_orbits(self.a, self.th_o, self.alpha, self.beta, self.tausteps)

    def turning_counts(self):
        """Turning points passed at every sample, (ngeo, npix) arrays computed on
        the GPU: (theta, r).  theta = floor(n_angular_turnings) (int32,
        kerr_raytracing_utils.py:686-792; needs th_o <= pi/2 as the reference
        does); r = 1 once a ray outside the critical curve has passed its
        turning point at half its total Mino time (int8)."""
        from .counts import turnings_device
        import torch
        out = turnings_device(self.a, self.th_o, self.alpha, self.beta, self.tausteps, r_o=self.r_o,
                              want=("turns_theta", "turns_r"))
        torch.cuda.current_stream().synchronize()
        return out["turns_theta"].cpu().numpy(), out["turns_r"].cpu().numpy()

    # labels computed with the rays (no reference attribute; rclass/thclass are
    # the reference's radial_roots / angular_turning classes)
    @property
    def rclass(self):
        return None if self._device_labels is None else self._device_labels["rclass"].astype(int)

    @property
    def thclass(self):
        return None if self._device_labels is None else self._device_labels["thclass"].astype(int)

    # geodesics (:69-93)
    @property
    def tausteps(self):
        return self.mino_times

    @property
    def tautot(self):
        return self.mino_times[-1]

    @property
    def affinesteps(self):
        return self.affine_times

    @property
    def t_s(self):
        return self.geo_coords[0]

    @property
    def r_s(self):
        return self.geo_coords[1]

    @property
    def th_s(self):
        return self.geo_coords[2]

    @property
    def ph_s(self):
        return self.geo_coords[3]

    @property
    def sig_s(self):
        return self.affine_times

    def plotgeos(self, xlim=12, nplot=None, ngeoplot=50):
        """Minimal x-z projection of up to ngeoplot rays (the reference's
        plotgeos, :95-199, draws the same projection with more decoration)."""
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
        r, th, ph = self.r_s, self.th_s, self.ph_s
        idx = np.linspace(0, self.npix - 1, min(ngeoplot, self.npix)).astype(int)
        fig = plt.figure()
        ax = fig.add_subplot(111)
        for i in idx:
            x = r[:, i] * np.sin(th[:, i]) * np.cos(ph[:, i])
            z = r[:, i] * np.cos(th[:, i])
            ax.plot(x, z, lw=0.5)
        ax.set_xlim(-xlim, xlim)
        ax.set_ylim(-xlim, xlim)
        ax.set_aspect("equal")
        return fig

    def savegeos(self, path="./"):
        """HDF5 output with the reference's keys (:201-222); needs h5py."""
        import h5py  # noqa: F401  (raises ImportError where h5py is absent)
        fname = path + "a%0.2f_th%0.2f_geo.h5" % (self.a, self.th_o * 180 / np.pi)
        with h5py.File(fname, "w") as f:
            f.create_dataset("spin", data=self.a)
            f.create_dataset("inc", data=self.th_o)
            f.create_dataset("r_o", data=self.r_o)
            f.create_dataset("alpha", data=self.alpha)
            f.create_dataset("beta", data=self.beta)
            f.create_dataset("t", data=self.t_s)
            f.create_dataset("r", data=self.r_s)
            f.create_dataset("theta", data=self.th_s)
            f.create_dataset("phi", data=self.ph_s)
            f.create_dataset("affine", data=self.sig_s)
            f.create_dataset("mino", data=self.mino_times)
        return fname
