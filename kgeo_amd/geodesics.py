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

    def plotgeos(self, xlim=12, rmax=15, nplot=None, ngeoplot=50, plot_disk=True, plot_inside_cc=True,
                 plot_outside_cc=True, n_label="equatorial"):
        """3D plot of the rays coloured by their number of equatorial crossings
        (n_label='poloidal': by completed poloidal orbits), over the horizon
        sphere and the equatorial disk, as kerr_raytracing_utils.py:95-199
        draws it: rays with label j use colour j of (dimgrey, b, g, orange, r,
        m, c, y); at most ngeoplot rays per label for labels below 4, each
        drawn where r < rmax or before half its Mino time, and within 3 rmax.
        Returns the figure."""
        import matplotlib.pyplot as plt
        a, th_o = self.a, self.th_o
        r_s, th_s, ph_s, tausteps = self.r_s, self.th_s, self.ph_s, self.tausteps
        if n_label == "poloidal":
            nmax_eq = np.floor(self.n_poloidal[-1])
        else:
            nmax_eq = np.array(self.nmax_eq, dtype=float)
        nmax_eq = nmax_eq + 1
        rplus = 1 + np.sqrt(1 - a ** 2)
        x_s = r_s * np.cos(ph_s) * np.sin(th_s)
        y_s = r_s * np.sin(ph_s) * np.sin(th_s)
        z_s = r_s * np.cos(th_s)

        fig = plt.figure()
        ax = fig.add_subplot(projection="3d")
        u, v = np.mgrid[0:2 * np.pi:30j, 0:np.pi:20j]
        ax.plot_surface(rplus * np.cos(u) * np.sin(v), rplus * np.sin(u) * np.sin(v), rplus * np.cos(v), color="black")
        if plot_disk:
            rr, pp = np.mgrid[0:xlim, 0:2 * np.pi:20j]
            ax.plot_surface(rr * np.cos(pp), rr * np.sin(pp), np.zeros(rr.shape), alpha=0.5)
        ax.set_xlim(-xlim, xlim)
        ax.set_ylim(-xlim, xlim)
        ax.set_zlim(-xlim, xlim)
        ax.auto_scale_xyz([-xlim, xlim], [-xlim, xlim], [-xlim, xlim])
        ax.set_axis_off()
        rmax2 = 3 * rmax
        ax.plot3D([0, rmax2 * np.sin(th_o)], [0, 0], [0, rmax2 * np.cos(th_o)], "black", ls="dashed")

        colors = ["dimgrey", "b", "g", "orange", "r", "m", "c", "y"]
        maxwraps = int(np.nanmax(nmax_eq))
        print("maxwraps ", maxwraps)
        if nplot is None:
            nplot = range(min(maxwraps + 1, len(colors)))
        else:
            nplot = np.array([nplot]).flatten()
        for j in nplot:
            mask = ((nmax_eq == 0) | (nmax_eq == -1)) if j == 0 else (nmax_eq == j)
            if not plot_inside_cc:
                mask = mask & (r_s[-1] > 10)
            if not plot_outside_cc:
                mask = mask & (r_s[-1] < 10)
            if not np.any(mask):
                continue
            xs, ys, zs = x_s[:, mask], y_s[:, mask], z_s[:, mask]
            rs, tau = r_s[:, mask], tausteps[:, mask]
            trim = max(int(xs.shape[-1] / ngeoplot), 1) if ngeoplot is not None else 0
            if xs.shape[-1] < 5 or j >= 4 or trim == 0:
                geos = range(xs.shape[-1])
            else:
                geos = range(0, xs.shape[-1], trim)
            for i in geos:
                keep = ((rs[:, i] < rmax) | (tau[:, i] < .5 * tau[-1, i])) & (rs[:, i] < rmax2)
                ax.plot3D(xs[keep, i], ys[keep, i], zs[keep, i], colors[j])
        return fig

    def savegeos(self, outfile=None, path="./"):
        """HDF5 file with the reference's datasets (kerr_raytracing_utils.py:
        201-222): scalars spin, inc, r_o; alpha, beta (npix); t, r, theta,
        phi, affine, mino (ngeo, npix).  Default name
        path + 'a%0.2f_th%0.2f_geo.h5' % (a, th_o in degrees).  Written with
        h5py when it is installed, else by h5lite (the same datasets in the
        HDF5 format's original layout).  Returns the file name."""
        fname = outfile if outfile is not None else path + "a%0.2f_th%0.2f_geo.h5" % (self.a, self.th_o * 180 / np.pi)
        data = {"spin": self.a, "inc": self.th_o, "r_o": self.r_o, "alpha": self.alpha, "beta": self.beta,
                "t": self.t_s, "r": self.r_s, "theta": self.th_s, "phi": self.ph_s, "affine": self.sig_s,
                "mino": self.tausteps}
        try:
            import h5py
        except ImportError:
            h5py = None
        if h5py is None:
            from . import h5lite
            h5lite.write(fname, data)
        else:
            with h5py.File(fname, "w") as f:
                for k, v in data.items():
                    f.create_dataset(k, data=v)
        return fname


def loadgeos(infile):
    """Geodesics from a savegeos file (kerr_raytracing_utils.py:224-241),
    read with h5py when it is installed, else by h5lite."""
    keys = ("spin", "r_o", "inc", "alpha", "beta", "mino", "affine", "t", "r", "theta", "phi")
    try:
        import h5py
    except ImportError:
        h5py = None
    if h5py is None:
        from . import h5lite
        hf = h5lite.read(infile)
        d = {k: hf[k] for k in keys}
    else:
        with h5py.File(infile, "r") as hf:
            d = {k: hf[k][()] for k in keys}
    return Geodesics(d["spin"], [0, d["r_o"], d["inc"], 0], [d["alpha"], d["beta"]], d["mino"], d["affine"],
                     [d["t"], d["r"], d["theta"], d["phi"]])
