"""Shared helpers for the test suite: workload definitions, the error
metrics used for parity, the conditioning probe, and a ctypes view of the
CPU instantiation of the device headers (libkgeo_hostcheck.so)."""

import ctypes
import os

import numpy as np

from oracle import kgeo_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOSTCHECK = os.path.join(ROOT, "kgeo_amd", "libkgeo_hostcheck.so")

# Workloads (seeds and unstated choices fixed here)
C1 = dict(a=0.94, th_deg=17.0, r_o=1.0e4, window=8.0, n=64, ngeo=100)
C2 = dict(a=0.94, th_deg=17.0, r_o=1.0e4, window=8.0, n=512, ngeo=500, subset=4096, seed=0)


def grid(window, n):
    """alpha, beta = meshgrid(linspace(-w, w, n)) flattened, alpha fastest."""
    al = np.linspace(-window, window, n)
    A, B = np.meshgrid(al, al)
    return A.flatten(), B.flatten()


def c2_subset():
    A, B = grid(C2["window"], C2["n"])
    idx = np.random.default_rng(C2["seed"]).choice(len(A), C2["subset"], replace=False)
    return A[idx], B[idx], idx


# --- error metrics ---------------------------------------------------------
# r: pointwise relative.  theta: |d theta| / max(|theta|, 1) (radians for
# small angles).  phi, t, sigma start at exactly 0 on every ray and
# accumulate; their pointwise relative error is undefined near tau = 0, so
# they are measured against the ray's own magnitude: |dx| / max(max_j |x_j|, 1).

def errors(x, y, key):
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    if key == "r":
        return np.abs(x - y) / np.abs(y)
    if key == "th":
        return np.abs(x - y) / np.maximum(np.abs(y), 1.0)
    scale = np.maximum(np.nanmax(np.abs(y), axis=0, keepdims=True), 1.0)
    return np.abs(x - y) / scale


def conditioning_mask(a, r_o, th_o, alpha, beta, tau, ref, rel=1.0e-14, ulps=2, thresh=1.0e-11):
    """True where the oracle itself is ill-conditioned at the 1e-10 level,
    i.e. where the reference's own rounding noise can reach ~1e-10:

    * the last sample of horizon-bound rays (rclass 2, 3, 4): tau =
      (1 - 1e-10) tau_tot leaves r - r_+ known to ~1e-6 relative, and phi, t,
      sigma contain log(r - r_+);
    * any sample whose oracle value moves by more than `thresh` (same
      metric) when tau is scaled by (1 -+ rel), or when the intermediates
      the samples inherit are moved by -+`ulps` ulp: u_+ and u_- (for
      near-polar rays u_+ = dth + sqrt(dth^2 + eta/a^2) cancels to ~1e-7 and
      phi varies as 1/sin^2(theta) near the pole) and the radial roots (near
      the critical curve r3 ~ r4).  This also catches the removable log
      singularities of the reference's R1/S1 split.
    """
    bad = np.zeros(tau.shape, dtype=bool)
    horizon = np.isin(ref["rclass"], (2, 3, 4)) if "rclass" in ref else np.zeros(tau.shape[1], bool)
    bad[-1, horizon] = True

    def flag(pert):
        nonlocal bad
        for k in pert:
            bad |= errors(pert[k], ref[k], k) > thresh

    for f in (1.0 - rel, 1.0 + rel):
        sig2, (t2, r2, th2, ph2) = O.coords_at_tau(a, r_o, th_o, alpha, beta, tau * f)
        flag(dict(t=t2, r=r2, th=th2, ph=ph2, sig=sig2))

    lam, eta = O.conserved_fixed(a, th_o, alpha, beta)
    s_o = O.sign_ge0(beta)
    up, um, _, _ = O.uplus_uminus(a, th_o, lam, eta)
    r1, r2, r3, r4, _ = O.radial_roots(a, lam, eta)
    rs, Iph, It, Isig = O.r_integrate(a, r_o, lam, eta, r1, r2, r3, r4, tau)
    thb, Gph, Gt = O.th_integrate(a, th_o, s_o, lam, eta, up, um, tau)
    eps = ulps * 2.0 ** -52
    with np.errstate(all="ignore"):
        for f in (1.0 - eps, 1.0 + eps):
            for upp, ump in ((up * f, um), (up, um * f)):
                th2, Gph2, Gt2 = O.th_integrate(a, th_o, s_o, lam, eta, upp, ump, tau)
                flag(dict(th=th2, ph=Iph + lam * Gph2, t=It + a ** 2 * Gt2, sig=Isig + a ** 2 * Gt2))
            rs2, Iph2, It2, Isig2 = O.r_integrate(a, r_o, lam, eta, r1 * f, r2 * f, r3 * f, r4 * f, tau)
            flag(dict(r=rs2, ph=Iph2 + lam * Gph, t=It2 + a ** 2 * Gt, sig=Isig2 + a ** 2 * Gt))
    return bad


def summarize(x, y, key, ill=None):
    e = errors(x, y, key)
    d = dict(max=float(np.nanmax(e)), p99=float(np.nanpercentile(e, 99)))
    if ill is not None:
        ew = np.where(ill, np.nan, e)
        d["max_well"] = float(np.nanmax(ew))
        d["max_ill"] = float(np.nanmax(np.where(ill, e, np.nan))) if ill.any() else 0.0
        d["n_ill"] = int(ill.sum())
    d["nan_mismatch"] = int(np.sum(np.isnan(np.asarray(x, float)) != np.isnan(np.asarray(y, float))))
    return d


# --- host instantiation of the device headers ------------------------------
class Host:
    def __init__(self):
        if not os.path.exists(HOSTCHECK):
            raise RuntimeError("libkgeo_hostcheck.so not built: run python -m kgeo_amd.build")
        self.lib = ctypes.CDLL(HOSTCHECK)

    @staticmethod
    def _p(a, t=ctypes.c_double):
        return a.ctypes.data_as(ctypes.POINTER(t))

    def _call(self, name, ins, nout):
        ins = [np.ascontiguousarray(a, dtype=float) for a in ins]
        n = len(ins[0])
        outs = [np.empty(n) for _ in range(nout)]
        getattr(self.lib, name)(*[self._p(a) for a in ins], ctypes.c_int64(n), *[self._p(o) for o in outs])
        return outs

    def rf(self, x, y, z):
        return self._call("kgh_rf", [x, y, z], 1)[0]

    def rd(self, x, y, z):
        return self._call("kgh_rd", [x, y, z], 1)[0]

    def rj(self, x, y, z, p):
        return self._call("kgh_rj", [x, y, z, p], 1)[0]

    def rc(self, x, y):
        return self._call("kgh_rc", [x, y], 1)[0]

    def ellipj(self, u, m):
        return self._call("kgh_ellipj", [u, m], 4)

    def legendre(self, n, phi, m):
        return self._call("kgh_legendre", [n, phi, m], 3)

    def stokes(self, a, r_o, th_o, model, alpha, beta, nsub):
        """kg_emission.hpp's stokes_pixel on the CPU: dict of (n, nsub) arrays
        I, Q, U, g, r_s, sin_thb, Ir, lp, Ie and (n,) nmax_eq, n_poloidal."""
        alpha = np.ascontiguousarray(alpha, float)
        beta = np.ascontiguousarray(beta, float)
        n = len(alpha)
        names = ("I", "Q", "U", "g", "r_s", "sin_thb", "Ir", "nmax_eq", "n_poloidal", "lp", "Ie")
        out = {k: np.full((n,) if k in ("nmax_eq", "n_poloidal") else (n, nsub), np.nan) for k in names}
        self.lib.kgh_stokes(self._p(self.scalars(a, r_o, th_o)), ctypes.byref(model), self._p(alpha), self._p(beta),
                            ctypes.c_int64(n), ctypes.c_int(nsub), *[self._p(out[k]) for k in names])
        return out

    def model_eval(self, model, a, r, th=None, g=None, sinthb=None, bmag=None):
        """kg_emission.hpp's model_point on the CPU: dict u (4, n), tetrad
        (16, n), B (3, n), omega, bsq (n,) and, when the emissivity is
        evaluable, j (n,)."""
        r = np.ascontiguousarray(r, float)
        n = len(r)
        out = {"u": np.empty((4, n)), "tetrad": np.empty((16, n)), "B": np.empty((3, n)), "omega": np.empty(n),
               "bsq": np.empty(n), "j": np.empty(n)}
        P = lambda x: None if x is None else self._p(np.ascontiguousarray(x, float))
        self.lib.kgh_model_eval(ctypes.byref(model), ctypes.c_double(a), self._p(r), P(th), ctypes.c_int64(n), P(g),
                                P(sinthb), P(bmag), *[self._p(out[k]) for k in ("u", "tetrad", "B", "omega", "bsq",
                                                                                "j")])
        return out

    def special(self, which, x):
        x = np.ascontiguousarray(x, float)
        out, out2 = np.empty(len(x)), np.empty(len(x))
        self.lib.kgh_special(ctypes.c_int(which), self._p(x), ctypes.c_int64(len(x)), self._p(out), self._p(out2))
        return out, out2

    @staticmethod
    def scalars(a, r_o, th_o, maxtaufrac=O.MAXTAUFRAC):
        return np.array([a, r_o, th_o, maxtaufrac, a ** 2, np.sin(th_o), np.cos(th_o), np.cos(th_o) ** 2,
                         1 + np.sqrt(1 - a ** 2), 1 - np.sqrt(1 - a ** 2)])

    def setup(self, a, r_o, th_o, alpha, beta):
        n = len(alpha)
        I = ctypes.c_int32
        rc = np.empty(n, np.int32); tc = np.empty(n, np.int32); cs = np.empty(n, np.int32)
        st = np.empty(n, np.uint32); tt = np.empty(n); r8 = np.empty(8 * n); uu = np.empty(2 * n)
        self.lib.kgh_setup(self._p(self.scalars(a, r_o, th_o)), self._p(np.ascontiguousarray(alpha, float)),
                           self._p(np.ascontiguousarray(beta, float)), ctypes.c_int64(n), self._p(rc, I),
                           self._p(tc, I), self._p(cs, I), self._p(st, ctypes.c_uint32), self._p(tt), self._p(r8),
                           self._p(uu))
        return dict(rclass=rc, thclass=tc, rcase=cs, status=st, tau_tot=tt, roots=r8.reshape(n, 8),
                    upum=uu.reshape(n, 2))

    def counts(self, a, r_o, th_o, alpha, beta, tau):
        """The count kernels' formulas (kg_counts.hpp) on the CPU at Mino times
        tau (ntau, npix): n_equatorial, n_poloidal, n_angular, turns_theta, turns_r."""
        alpha = np.ascontiguousarray(alpha, float)
        beta = np.ascontiguousarray(beta, float)
        tau = np.ascontiguousarray(tau, float).reshape(-1, len(alpha))
        ntau, n = tau.shape
        out = dict(n_equatorial=np.empty((ntau, n)), n_poloidal=np.empty((ntau, n)), n_angular=np.empty((ntau, n)),
                   turns_theta=np.empty((ntau, n), np.int32), turns_r=np.empty((ntau, n), np.int8))
        self.lib.kgh_counts(self._p(self.scalars(a, r_o, th_o)), self._p(alpha), self._p(beta), ctypes.c_int64(n),
                            self._p(tau), ctypes.c_int64(ntau), self._p(out["n_equatorial"]),
                            self._p(out["n_poloidal"]), self._p(out["n_angular"]),
                            self._p(out["turns_theta"], ctypes.c_int32), self._p(out["turns_r"], ctypes.c_int8))
        return out

    def raytrace(self, a, r_o, th_o, alpha, beta, ngeo, tau_in=None):
        n = len(alpha)
        outs = {k: np.empty((ngeo, n)) for k in ("tau", "t", "r", "th", "ph", "sig")}
        nm = np.empty(n, np.int32)
        ti = None if tau_in is None else self._p(np.ascontiguousarray(tau_in, float))
        self.lib.kgh_raytrace(self._p(self.scalars(a, r_o, th_o)), self._p(np.ascontiguousarray(alpha, float)),
                              self._p(np.ascontiguousarray(beta, float)), ctypes.c_int64(n), ctypes.c_int64(ngeo),
                              ctypes.c_int(1), ti, *[self._p(outs[k]) for k in ("tau", "t", "r", "th", "ph", "sig")],
                              self._p(nm, ctypes.c_int32))
        outs["nmax_eq"] = nm
        return outs

    def raytrace_loops(self, a, r_o, th_o, alpha, beta, ngeo, tau_in=None, chunks=1, passes=3, do_phi_t=1):
        """The kernels' per-pixel sample code on the CPU, with ngeo split into
        `chunks` ranges as the kernels' grid does; adds the status bits.
        passes=3: the anchor / walk / fixup passes (the kernels' path on the
        linspace grid; adds the theta and radial fixup-queue lengths and the
        radial repair items as "nfix"); passes=1: one
        combined loop per coordinate.  do_phi_t=0 runs the passes without t
        and phi (NULL outputs), which only the CPU build keeps."""
        n = len(alpha)
        outs = {k: np.empty((ngeo, n)) for k in ("tau", "t", "r", "th", "ph", "sig")}
        nm = np.empty(n, np.int32)
        st = np.empty(n, np.uint32)
        nfix = np.zeros(3, np.int64)
        ti = None if tau_in is None else self._p(np.ascontiguousarray(tau_in, float))
        self.lib.kgh_raytrace_passes(self._p(self.scalars(a, r_o, th_o)), self._p(np.ascontiguousarray(alpha, float)),
                                     self._p(np.ascontiguousarray(beta, float)), ctypes.c_int64(n),
                                     ctypes.c_int64(ngeo), ctypes.c_int(do_phi_t), ti, ctypes.c_int(chunks),
                                     ctypes.c_int(passes),
                                     *[self._p(outs[k]) for k in ("tau", "t", "r", "th", "ph", "sig")],
                                     self._p(nm, ctypes.c_int32), self._p(st, ctypes.c_uint32),
                                     self._p(nfix, ctypes.c_int64))
        outs["nmax_eq"] = nm
        outs["status"] = st
        outs["nfix"] = nfix
        # the kernels' fixup queue holds npix * ceil(ngeo / 8) items per pass
        assert np.all(nfix >= 0), "fixup queue over the kernels' capacity"
        return outs


# --- equatorial image workloads ---------------------------------------------
C3 = dict(a=0.27118414089647946,  # np.random.default_rng(0).uniform(-0.99, 0.99)
          th_deg=60.0, r_o=1.0e4, window=12.0, n=1024, mbar_max=2, p=2.0)
C5 = dict(a=0.7, th_deg=40.0, p=2.0, target=(0.5, 60.0, 2.5), r_o=1.0e4, window=12.0, n=256, mbar_max=2,
          fd_step=1.0e-6, crit_dist=1.0e-3)


def host_equatorial(host, a, r_o, th, A, B, mbar_max=2, p=2.0):
    n = len(A)
    M = mbar_max + 1
    out = dict(image=np.empty(n), I=np.empty(M * n), g=np.empty(M * n), r_s=np.empty(M * n), Ir=np.empty(M * n),
               Imax=np.empty(n), Nmax=np.empty(n, np.int32), zmask=np.empty(n, np.int32))
    P = Host._p
    host.lib.kgh_equatorial(P(Host.scalars(a, r_o, th)), P(np.ascontiguousarray(A, float)),
                            P(np.ascontiguousarray(B, float)), ctypes.c_int64(n), ctypes.c_int(mbar_max),
                            ctypes.c_double(p), P(out["image"]), P(out["I"]), P(out["g"]), P(out["r_s"]),
                            P(out["Ir"]), P(out["Imax"]), P(out["Nmax"], ctypes.c_int32),
                            P(out["zmask"], ctypes.c_int32))
    for k in ("I", "g", "r_s", "Ir"):
        out[k] = out[k].reshape(M, n)
    return out


def host_equatorial_jac(host, a, th, r_o, p, A, B, mbar_max=2):
    n = len(A)
    img = np.empty(n)
    jac = np.empty(3 * n)
    P = Host._p
    host.lib.kgh_equatorial_jac(ctypes.c_double(a), ctypes.c_double(th), ctypes.c_double(r_o), ctypes.c_double(p),
                                P(np.ascontiguousarray(A, float)), P(np.ascontiguousarray(B, float)),
                                ctypes.c_int64(n), ctypes.c_int(mbar_max), P(img), P(jac))
    return img, jac.reshape(3, n)


def image_errors(x, y):
    """Relative error of image values; exact zeros (masked pixels) compared absolutely."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    return np.where(y == 0, np.abs(x), np.abs(x - y) / np.maximum(np.abs(y), 1e-300))


def fd_loss_gradient(theta, target, keep, A, B, r_o, mbar_max, h):
    """Central finite differences (step h) of ||image(theta) - target||_2 over
    `keep`, with the oracle's image.  Also returns the per-parameter integer
    fingerprints (N_max, masks) used to detect discontinuities."""
    L = lambda im: np.sqrt(np.sum(((im - target) * keep) ** 2))
    grad = np.zeros(3)
    for k in range(3):
        vals = []
        for s in (1.0, -1.0):
            q = np.array(theta, float)
            q[k] += s * h
            im, _ = O.equatorial_image(q[0], r_o, q[1], A, B, mbar_max, q[2])
            vals.append(L(im))
        grad[k] = (vals[0] - vals[1]) / (2 * h)
    return grad


def discontinuity_mask(theta, A, B, r_o, mbar_max, h):
    """Pixels whose N_max or subring masks change within the FD stencil."""
    _, base = O.equatorial_image(theta[0], r_o, theta[1], A, B, mbar_max, theta[2])
    disc = np.zeros(len(A), bool)
    for k in range(3):
        for s in (1.0, -1.0):
            q = np.array(theta, float)
            q[k] += s * h
            _, parts = O.equatorial_image(q[0], r_o, q[1], A, B, mbar_max, q[2])
            for m in range(mbar_max + 1):
                disc |= (parts[m]["Nmax"] != base[m]["Nmax"]) | (parts[m]["zeromask"] != base[m]["zeromask"])
    return disc
