"""Emission models of the equatorial disk: drop-ins for the reference's
Velocity (kgeo/velocities.py:23-176), Bfield (kgeo/bfields.py:33-427) and
Emissivity (kgeo/emissivities.py:37-118).

The constructors take the reference's arguments, with its defaults and its
error messages; the objects describe a model (kgeo_velocity_model,
kgeo_bfield_model, kgeo_emissivity_model of include/kgeo.h) and evaluate it
on the GPU through kgeo_model_eval: u_lab, u_lab_cov, tetrades, bfield_lab,
omega_field, bfield_fluid, bsq and jrest return numpy arrays as the
reference's methods do.  Tables of the 'fromfile' models live in device
memory for as long as the object does.

Not supported (NotImplementedError): driftframe with nu_parallel='FF' or
gammamax (the force-free stagnation-surface solver of ff_boost.py and the
image-wide gamma rescaling are outside the equatorial image).
"""

import ctypes

import numpy as np

from . import _lib

_allowed_velocity_models = ["zamo", "infall", "kep", "cunningham", "subkep", "cunningham_subkep", "general",
                            "gelles", "simfit", "fromfile", "driftframe"]

# model constants of the reference (velocities.py:6-12, emissivities.py:17-23)
ELLISCO, VRISCO, P1, P2, DD = 1., 2, 6., 2., 0.2
BETA, CHI = 0.3, -150 * np.pi / 180.
P1E_230, P2E_230 = -2.0, -0.5
GAMMAOFF, SIGMA_GLM, R_RING, SIGMA_RING = -1.5, 0.5, 4.5, 0.3
OBSFREQ = 230.e9
SPECIND = 0.0

_VEL_KIND = {"zamo": 0, "infall": 1, "kep": 2, "cunningham": 2, "subkep": 3, "cunningham_subkep": 3, "general": 4,
             "gelles": 5, "simfit": 6, "fromfile": 7, "driftframe": 8}
_BF_KIND = {"rad": 0, "vert": 0, "tor": 0, "simple": 0, "simple_rm1": 1, "monopoleA": 2, "bz_monopole": 3,
            "bz_guess": 4, "bz_para": 5, "power": 6, "gen_power": 7, "fromfile": 8}
_EMIS_KIND = {"constant": 0, "bpl": 1, "glm": 2, "ring": 2, "thermal": 3}


def load_table(filename, names):
    """A '# r, ...' CSV table (velocities.py:626-655, bfields.py:749-787) as
    a (4, n) array: r and the named columns (zeros where absent)."""
    with open(filename, "r") as f:
        header = f.readline().strip()
    if header[0] != "#":
        raise Exception("file %s does not have a header!" % filename)
    header = [x.strip() for x in header[1:].split(",")]
    data = np.loadtxt(filename, delimiter=",", skiprows=1, ndmin=2)
    radii = data[:, header.index("r")]
    cols = [data[:, header.index(n)] if n in header else np.zeros_like(radii) for n in names]
    return np.ascontiguousarray(np.vstack([radii] + cols), dtype=np.float64)


class _Table:
    """A host table and, on first use, its device copy."""

    def __init__(self, host):
        self.host = host
        self._dev = None

    def pointer(self, device):
        if not device:
            return self.host.ctypes.data, self.host.shape[1]
        if self._dev is None:
            from .raytrace import _device
            import torch
            self._dev = torch.from_numpy(self.host).to(_device())
        return self._dev.data_ptr(), self.host.shape[1]


def _r_array(r):
    return np.ascontiguousarray(np.array([r]).flatten() if not isinstance(r, np.ndarray) else r.reshape(-1),
                                dtype=np.float64)


def _check_spin(a):
    if not (isinstance(a, float) and (0 <= np.abs(a) < 1)):
        raise Exception("|a| should be a float in range [0,1)")


class Bfield(object):
    """bfields.py:33-91 (constructor semantics and messages)."""

    def __init__(self, fieldtype="rad", **kwargs):
        self.fieldtype = fieldtype
        if fieldtype in ["rad", "vert", "tor", "simple", "simple_rm1", "monopoleA", "bz_monopole", "bz_guess",
                         "bz_para", "power", "fromfile", "gen_power"]:
            self.fieldframe = "lab"
        elif fieldtype in ["const_comoving"]:
            self.fieldframe = "comoving"
        else:
            raise Exception("fieldtype %s not recognized in Bfield!" % fieldtype)
        self.kwargs = kwargs
        self._table = None
        if fieldtype in ["bz_monopole", "bz_guess", "bz_para"]:
            self.secondorder_only = kwargs.get("secondorder_only", False)
            self.C = kwargs.get("C", 1)
        elif fieldtype in ["monopoleA"]:
            self.omegafac = kwargs.get("omegafac", 0.5)
            self.C = kwargs.get("C", 1)
        elif fieldtype == "power":
            self.secondorder_only = kwargs.get("secondorder_only", False)
            self.C = kwargs.get("C", 1)
            self.pval = kwargs.get("p", 1)
            self.usemono = kwargs.get("usemono", False)
        elif fieldtype == "rad":
            self.Cr, self.Cvert, self.Cph = 1, 0, 0
        elif fieldtype == "vert":
            self.Cr, self.Cvert, self.Cph = 0, 1, 0
        elif fieldtype == "tor":
            self.Cr, self.Cvert, self.Cph = 0, 0, 1
        elif fieldtype == "fromfile":
            self.filename = kwargs.get("file", None)
            self._table = _Table(load_table(self.filename, ["Br", "Btheta", "Bphi"]))
        elif fieldtype == "gen_power":
            self.n_I = kwargs.get("n_I", 0)
            self.p_val = kwargs.get("p_val", 1)
            self.isAbove = kwargs.get("isAbove", True)
        else:
            self.Cr = kwargs.get("Cr", 0)
            self.Cvert = kwargs.get("Cvert", 0)
            self.Cph = kwargs.get("Cph", 0)
            if self.Cr == self.Cvert == self.Cph == 0.:
                raise Exception("all field coefficients are 0!")

    def to_c(self, device=True):
        """The kgeo_bfield_model of this field (table pointer on the device,
        or on the host with device=False)."""
        if self.fieldframe != "lab":
            raise Exception("Bfield.bfield_lab only supported for Bfield.fieldtype==lab")
        m = _lib.BfieldModel()
        m.kind = _BF_KIND[self.fieldtype]
        m.secondorder_only = int(bool(getattr(self, "secondorder_only", False)))
        m.usemono = int(bool(getattr(self, "usemono", False)))
        m.is_above = int(bool(getattr(self, "isAbove", True)))
        m.C = float(getattr(self, "C", 1))
        m.Cr, m.Cvert, m.Cph = (float(getattr(self, k, 0)) for k in ("Cr", "Cvert", "Cph"))
        m.omegafac = float(getattr(self, "omegafac", 0.5))
        m.p = float(getattr(self, "pval", 1))
        m.n_I = float(getattr(self, "n_I", 0))
        m.p_val = float(getattr(self, "p_val", 1))
        if self._table is not None:
            ptr, n = self._table.pointer(device)
            m.table, m.table_n = ptr, n
        return m

    def bfield_lab(self, a, r, th=np.pi / 2):
        """lab-frame B^i (bfields.py:93-131)."""
        if self.fieldframe != "lab":
            raise Exception("Bfield.bfield_lab only supported for Bfield.fieldtype==lab")
        out = model_eval(a, r, th, bfield=self, want=("B",))
        return tuple(out["B"])

    def bfield_comoving(self, a, r):
        """bfields.py:159-172."""
        r = _r_array(r)
        if self.fieldframe != "comoving":
            raise Exception("Bfield.bfield_comoving only supported for Bfield.fieldtype==comoving")
        return (self.Cr * np.ones(r.shape), -1 * self.Cvert * np.ones(r.shape), self.Cph * np.ones(r.shape))

    def omega_field(self, a, r, th=np.pi / 2):
        """field-line angular velocity (bfields.py:175-196)."""
        if self.fieldtype not in ("bz_monopole", "monopoleA", "bz_guess", "bz_para", "power", "gen_power"):
            raise Exception("self.omega_field not implemented for fieldtype %s'!" % self.fieldtype)
        return model_eval(a, r, th, bfield=self, want=("omega",))["omega"]

    def bfield_fluid(self, a, r, velocity, th=np.pi / 2.):
        """b^mu in the fluid frame, ideal MHD (bfields.py:330-362)."""
        out = model_eval(a, r, th, bfield=self, velocity=velocity, want=("u", "B"))
        u0, u1, u2, u3 = out["u"]
        B1, B2, B3 = out["B"]
        th = np.full(len(u0), th) if np.ndim(th) == 0 else th
        r = _r_array(r)
        a2, r2 = a ** 2, r ** 2
        cth2, sth2 = np.cos(th) ** 2, np.sin(th) ** 2
        Sigma = r2 + a2 * cth2
        g11, g22 = Sigma / (r2 - 2 * r + a2), Sigma
        g33 = (r2 + a2 + 2 * r * a2 * sth2 / Sigma) * sth2
        g03 = -2 * r * a * sth2 / Sigma
        b0 = B1 * g11 * u1 + B2 * g22 * u2 + B3 * (g33 * u3 + g03 * u0)
        return (b0, (B1 + b0 * u1) / u0, (B2 + b0 * u2) / u0, (B3 + b0 * u3) / u0)

    def bsq(self, a, r, velocity, th=np.pi / 2.):
        """b^2 in the fluid frame (bfields.py:365-395)."""
        return model_eval(a, r, th, bfield=self, velocity=velocity, want=("bsq",))["bsq"]


BFIELD_DEFAULT = Bfield("bz_para")


class Velocity(object):
    """velocities.py:23-68 (constructor semantics and messages)."""

    def __init__(self, veltype="kep", **kwargs):
        self.veltype = veltype
        self.kwargs = kwargs
        self._table = None
        if veltype in ("zamo", "infall"):
            pass
        elif veltype in ("kep", "cunningham"):
            self.retrograde = kwargs.get("retrograde", False)
        elif veltype in ("subkep", "cunningham_subkep"):
            self.retrograde = kwargs.get("retrograde", False)
            self.fac_subkep = kwargs.get("fac_subkep", 1)
        elif veltype == "general":
            self.retrograde = kwargs.get("retrograde", False)
            self.fac_subkep = kwargs.get("fac_subkep", 1)
            self.beta_phi = kwargs.get("beta_phi", 1)
            self.beta_r = kwargs.get("beta_r", 1)
        elif veltype == "gelles":
            self.gelles_beta = kwargs.get("gelles_beta", BETA)
            self.gelles_chi = kwargs.get("gelles_chi", CHI)
        elif veltype == "simfit":
            self.ell_isco = kwargs.get("ell_isco", ELLISCO)
            self.vr_isco = kwargs.get("vr_isco", VRISCO)
            self.p1 = kwargs.get("p1", P1)
            self.p2 = kwargs.get("p2", P2)
            self.dd = kwargs.get("dd", DD)
        elif veltype == "fromfile":
            self.filename = kwargs.get("file", None)
            self._table = _Table(load_table(self.filename, ["U1", "U2", "U3"]))
        elif veltype == "driftframe":
            self.bfield = kwargs.get("bfield", BFIELD_DEFAULT)
            self.nu_parallel = kwargs.get("nu_parallel", 0)
            self.gammamax = kwargs.get("gammamax", None)
        else:
            raise Exception("veltype %s not recognized in Velocity!" % veltype)

    def to_c(self, device=True):
        """The kgeo_velocity_model of this velocity."""
        m = _lib.VelocityModel()
        m.kind = _VEL_KIND[self.veltype]
        m.retrograde = int(bool(getattr(self, "retrograde", False)))
        m.fac_subkep = float(getattr(self, "fac_subkep", 1))
        m.beta_phi = float(getattr(self, "beta_phi", 1))
        m.beta_r = float(getattr(self, "beta_r", 1))
        m.gelles_beta = float(getattr(self, "gelles_beta", BETA))
        m.gelles_chi = float(getattr(self, "gelles_chi", CHI))
        m.ell_isco = float(getattr(self, "ell_isco", ELLISCO))
        m.vr_isco = float(getattr(self, "vr_isco", VRISCO))
        m.p1 = float(getattr(self, "p1", P1))
        m.p2 = float(getattr(self, "p2", P2))
        m.dd = float(getattr(self, "dd", DD))
        if self.veltype == "driftframe":
            if isinstance(self.nu_parallel, str) or self.gammamax:
                raise NotImplementedError("driftframe with nu_parallel='FF' or gammamax is not supported")
            m.nu_parallel = float(self.nu_parallel)
            m.drift_field = self.bfield.to_c(device)
        if self._table is not None:
            ptr, n = self._table.pointer(device)
            m.table, m.table_n = ptr, n
        return m

    def u_lab(self, a, r, th=np.pi / 2., retqty=False):
        """contravariant four-velocity (velocities.py:70-96)."""
        if retqty:
            raise NotImplementedError("u_lab(retqty=True) is not supported")
        return tuple(model_eval(a, r, th, velocity=self, want=("u",))["u"])

    def u_lab_cov(self, a, r, th=np.pi / 2.):
        """covariant four-velocity (velocities.py:98-122)."""
        u0, u1, u2, u3 = self.u_lab(a, r, th=th)
        r = _r_array(r)
        a2, r2 = a ** 2, r ** 2
        cth2, sth2 = np.cos(th) ** 2, np.sin(th) ** 2
        Delta = r2 - 2 * r + a2
        Sigma = r2 + a2 * cth2
        g00, g11, g22 = -(1 - 2 * r / Sigma), Sigma / Delta, Sigma
        g33 = (r2 + a2 + 2 * r * a2 * sth2 / Sigma) * sth2
        g03 = -2 * r * a * sth2 / Sigma
        return (g00 * u0 + g03 * u3, g11 * u1, g22 * u2, g33 * u3 + g03 * u0)

    def tetrades(self, a, r, th=np.pi / 2):
        """((e_t), (e_x), (e_y), (e_z)) of the comoving frame
        (velocities.py:124-176)."""
        t = model_eval(a, r, th, velocity=self, want=("tetrad",))["tetrad"]
        return tuple(tuple(t[4 * row + k] for k in range(4)) for row in range(4))


class Emissivity(object):
    """emissivities.py:37-74 (constructor semantics and messages)."""

    def __init__(self, emistype="bpl", **kwargs):
        self.emistype = emistype
        self.kwargs = kwargs
        self.emiscut_in = kwargs.get("emiscut_in", 0)
        self.emiscut_out = kwargs.get("emiscut_out", 1.e10)
        if emistype == "thermal":
            self.Rb = float(kwargs.get("Rb", 5.0))
            self.ne0 = float(kwargs["ne0"])
            self.Te0 = float(kwargs["Te0"])
            self.B0 = float(kwargs["B0"])
            self.alpha_n = float(kwargs.get("alpha_n", 0.7))
            self.alpha_T = float(kwargs.get("alpha_T", 1.0))
            self.alpha_B = float(kwargs.get("alpha_B", 1.5))
            self.bfield_model = bool(kwargs.get("bfield_model", False))
        elif emistype == "bpl":
            self.p1 = kwargs.get("p1", P1E_230)
            self.p2 = kwargs.get("p2", P2E_230)
        elif emistype == "ring":
            self.mu_ring = kwargs.get("r_ring", R_RING)
            self.gamma_off = kwargs.get("gamma_off", 0)
            self.sigma = kwargs.get("sigma", SIGMA_RING)
        elif emistype == "glm":
            self.mu_ring = False
            self.gamma_off = kwargs.get("gamma_off", GAMMAOFF)
            self.sigma = kwargs.get("sigma", SIGMA_GLM)
        elif emistype == "constant":
            pass
        else:
            raise Exception("emistype %s not recognized in Emissivity!" % emistype)

    def to_c(self):
        m = _lib.EmissivityModel()
        m.kind = _EMIS_KIND[self.emistype]
        m.bfield_model = int(bool(getattr(self, "bfield_model", False)))
        m.emiscut_in = float(self.emiscut_in)
        m.emiscut_out = float(self.emiscut_out)
        m.p1 = float(getattr(self, "p1", P1E_230))
        m.p2 = float(getattr(self, "p2", P2E_230))
        mu = getattr(self, "mu_ring", False)
        m.mu_ring = float(mu) if mu else 0.0
        m.gamma_off = float(getattr(self, "gamma_off", GAMMAOFF))
        m.sigma = float(getattr(self, "sigma", SIGMA_GLM))
        for k in ("Rb", "ne0", "Te0", "B0", "alpha_n", "alpha_T", "alpha_B"):
            setattr(m, k, float(getattr(self, k, 0.0)))
        return m

    def jrest(self, a, r, g=None, sinthetab=None, Bmag=None, nu_obs=OBSFREQ, specind=SPECIND):
        """rest-frame emissivity (emissivities.py:83-118)."""
        r = np.asarray(r, dtype=float)
        n = r.size
        extra = {}
        for name, v in (("g", g), ("sinthb", sinthetab), ("bmag", Bmag)):
            if v is not None:
                extra[name] = np.broadcast_to(np.asarray(v, dtype=float), r.shape).reshape(n)
        out = model_eval(a, r.reshape(n), np.pi / 2, emissivity=self, want=("j",), nu_obs=nu_obs, specind=specind,
                         **extra)
        return out["j"].reshape(r.shape)


def emission_model(velocity=None, bfield=None, emissivity=None, polarization=False, pathlength=False, specind=1.,
                   diskangle=0.1, nu_obs=OBSFREQ, device=True):
    """The kgeo_emission_model for Iobs with these models and flags."""
    m = _lib.EmissionModel()
    m.velocity = (velocity or Velocity("zamo")).to_c(device)
    m.bfield = (bfield or Bfield("rad")).to_c(device)
    m.emissivity = (emissivity or Emissivity("bpl")).to_c()
    m.polarization = int(bool(polarization))
    m.pathlength = int(bool(pathlength))
    m.specind = float(specind)
    m.diskangle = float(diskangle)
    m.nu_obs = float(nu_obs)
    return m


def model_eval(a, r, th=np.pi / 2, velocity=None, bfield=None, emissivity=None, want=("u",), g=None, sinthb=None,
               bmag=None, nu_obs=OBSFREQ, specind=SPECIND):
    """kgeo_model_eval on the GPU: the wanted outputs ('u' (4, n), 'tetrad'
    (16, n), 'B' (3, n), 'omega', 'bsq', 'j' (n,)) as numpy arrays."""
    import torch
    from .raytrace import _device, _ptr, _stream_handle
    _check_spin(a)
    lib = _lib.load()
    dev = _device()
    rr = _r_array(r)
    n = len(rr)
    rt = torch.from_numpy(rr).to(dev)
    tht = None
    if not (np.ndim(th) == 0 and float(th) == np.pi / 2):
        tht = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(np.asarray(th, float), (n,)))).to(dev)
    model = emission_model(velocity, bfield, emissivity, nu_obs=nu_obs, specind=specind)
    shapes = {"u": 4, "tetrad": 16, "B": 3, "omega": 1, "bsq": 1, "j": 1}
    outs = {k: torch.empty((shapes[k], n), dtype=torch.float64, device=dev) for k in want}
    ins = {k: None if v is None else torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(dev)
           for k, v in (("g", g), ("sinthb", sinthb), ("bmag", bmag))}
    _lib.check(lib.kgeo_model_eval(ctypes.byref(model), float(a), _ptr(rt), _ptr(tht), n, _ptr(ins["g"]),
                                   _ptr(ins["sinthb"]), _ptr(ins["bmag"]), _ptr(outs.get("u")),
                                   _ptr(outs.get("tetrad")), _ptr(outs.get("B")), _ptr(outs.get("omega")),
                                   _ptr(outs.get("bsq")), _ptr(outs.get("j")), _stream_handle()), "kgeo_model_eval")
    torch.cuda.current_stream().synchronize()
    res = {}
    for k, t in outs.items():
        h = t.cpu().numpy()
        res[k] = h[0] if shapes[k] == 1 else h
    return res
