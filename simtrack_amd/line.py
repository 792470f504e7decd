"""Element / Line construction API mirroring SimTrack's C++ classes.

The reference builds a lattice from Element subclasses (simtrack.h:373-1059)
held in a Line (simtrack.h:1067-1100) and edits them with SetP/GetP.  This
module keeps those names, constructor signatures and defaults (Nint from the
slice lengths, Norder from the highest non-zero multipole) so a lattice
written for the reference can be written here unchanged, and lowers the Line
to the st_element records of the C ABI.  Unknown parameter names raise
KeyError instead of exiting the process.
"""
import math

import numpy as np

from ._capi import ELEMENT_DTYPE, TYPES, Globals

# slice lengths used by the constructors (simtrack.cpp:45-47)
BLslice = 0.25
QLslice = 0.05
GLslice = 0.05


# ----------------------------------------------------------------- globals

def globals_default():
    """GlobalVariables() defaults, simtrack.h:31-58."""
    g = Globals()
    g.A, g.Q = 1, 1
    g.m0 = g.m = 1.6726216e-27
    g.q = 1.6021765e-19
    g.gamma, g.beta = 106., 0.999955
    g.energy0, g.energy = 938.27201, 99456.8
    g.p = g.cp = 1.
    g.brho = 331.738
    g.G = 1.7928474
    g.H_expand = g.radiate = g.quad_fringe = 0
    g.rng_seed = 0
    return g


def set_ref_part_energy(g, gamma, E0, Q, A):
    """Set_RefPartEnergy, simtrack.cpp:87-108 (same operations, same order)."""
    g.Q = int(Q)
    g.A = int(A)
    g.m0 = E0 * 1.0e6 * 1.6021765e-19 / 299792458. / 299792458.
    g.m = gamma * g.m0
    g.q = Q * 1.6021765e-19
    g.gamma = gamma
    g.beta = math.sqrt(1.0 - 1.0 / gamma / gamma)
    g.energy0 = E0
    g.energy = gamma * g.energy0
    g.cp = g.energy * g.beta
    g.p = g.energy * g.beta * 1.0e6 * 1.6021765e-19 / 299792458.
    g.brho = 3.33564095198 * g.beta * (g.energy / 1000) / g.Q
    return g


def make_globals(gamma, E0, Q, A, **flags):
    g = set_ref_part_energy(globals_default(), gamma, E0, Q, A)
    for k, v in flags.items():
        setattr(g, k, v)
    return g


# ---------------------------------------------------------------- elements

def _norder(knl, knsl):
    """Highest non-zero order at construction (simtrack.cpp:8944-8950)."""
    for i in range(10):
        if knl[10 - i] != 0. or knsl[10 - i] != 0.:
            return 10 - i
    return 1


class Element:
    """Common fields of the reference's Element (simtrack.h:384-395)."""
    TYPE = None
    _params = {}            # SetP/GetP name -> attribute (or (attr, index))

    def __init__(self, name):
        self.NAME = name
        self.S = 0.
        self.L = 0.
        self.DX = self.DY = self.DS = 0.
        self.DTHETA = self.DPHI = self.DPSI = 0.
        self.APx = self.APy = 1.

    def SetP(self, name, value):
        key = self._params.get(name)
        if key is None:
            raise KeyError(f"{self.TYPE} does not have a parameter of {name}")
        if isinstance(key, tuple):
            getattr(self, key[0])[key[1]] = float(value)
        elif key in ("Nint", "Norder", "NSLICE", "TREATMENT"):
            setattr(self, key, int(value))
        else:
            setattr(self, key, float(value))

    def GetP(self, name):
        key = self._params.get(name)
        if key is None:
            raise KeyError(f"{self.TYPE} does not have a parameter of {name}")
        if isinstance(key, tuple):
            return getattr(self, key[0])[key[1]]
        return getattr(self, key)

    def record(self):
        r = np.zeros((), ELEMENT_DTYPE)
        r["type"] = TYPES[self.TYPE]
        r["l"] = self.L
        r["dx"], r["dy"], r["ds"] = self.DX, self.DY, self.DS
        r["dtheta"], r["dphi"], r["dpsi"] = self.DTHETA, self.DPHI, self.DPSI
        r["apx"], r["apy"] = self.APx, self.APy
        self._fill(r)
        return r

    def _fill(self, r):
        pass

    def __repr__(self):
        return f"{self.TYPE}({self.NAME!r}, L={self.L})"


def _check_len(l, strict=False):
    if (l <= 0) if strict else (l < 0):
        raise ValueError("element length must be %s" % ("positive" if strict else "non-negative"))


class DRIFT(Element):
    TYPE = "DRIFT"

    def __init__(self, name, l):
        super().__init__(name)
        _check_len(l)
        self.L = float(l)


class MARKER(DRIFT):
    TYPE = "MARKER"


class BPM(DRIFT):
    TYPE = "BPM"


class HBPM(DRIFT):
    TYPE = "HBPM"


class VBPM(DRIFT):
    TYPE = "VBPM"


class SBEND(Element):
    TYPE = "SBEND"
    _params = {"ANGLE": "ANGLE", "E1": "E1", "E2": "E2", "Nint": "Nint"}

    def __init__(self, name, l, angle, e1=0., e2=0.):
        super().__init__(name)
        _check_len(l, strict=True)
        self.L, self.ANGLE, self.E1, self.E2 = float(l), float(angle), float(e1), float(e2)
        self.Nint = int(self.L / BLslice) + 1

    def _fill(self, r):
        r["angle"], r["e1"], r["e2"], r["nint"] = self.ANGLE, self.E1, self.E2, self.Nint


class GSBEND(SBEND):
    TYPE = "GSBEND"
    _params = {"ANGLE": "ANGLE", "K0L": "K0L", "E1": "E1", "E2": "E2", "Nint": "Nint"}

    def __init__(self, name, l, angle, k0l, e1=0., e2=0.):
        super().__init__(name, l, angle, e1, e2)
        self.K0L = float(k0l)

    def _fill(self, r):
        super()._fill(r)
        r["k0l"] = self.K0L


class _Multipole(Element):
    ORDER = 1

    def __init__(self, name, l, kn, ks):
        super().__init__(name)
        _check_len(l)
        self.L = float(l)
        self.KN, self.KS = float(kn), float(ks)
        self.Nint = int(self.L / QLslice) + 1
        self.Norder = self.ORDER

    def _fill(self, r):
        r["nint"], r["norder"] = self.Nint, self.Norder
        r["knl"][self.ORDER] = self.KN
        r["knsl"][self.ORDER] = self.KS


class QUAD(_Multipole):
    TYPE, ORDER = "QUAD", 1
    _params = {"K1L": "KN", "K1SL": "KS", "Nint": "Nint"}


class SKEWQ(_Multipole):
    TYPE, ORDER = "SKEWQ", 1
    _params = {"K1SL": "KS", "Nint": "Nint"}

    def __init__(self, name, l, k1sl):
        super().__init__(name, l, 0., k1sl)


class SEXT(_Multipole):
    TYPE, ORDER = "SEXT", 2
    _params = {"K2L": "KN", "K2SL": "KS", "Nint": "Nint"}


class OCT(_Multipole):
    TYPE, ORDER = "OCT", 3
    _params = {"K3L": "KN", "K3SL": "KS", "Nint": "Nint"}


def _kn_params(extra):
    """KnL/KnSL names after the class's own: a name the class already uses
    keeps its meaning (GSBENDMULT's "K0L" is the bend's K0L, simtrack.cpp:9575)."""
    p = dict(extra)
    for i in range(11):
        p.setdefault(f"K{i}L", ("KNL", i))
        p.setdefault(f"K{i}SL", ("KNSL", i))
    return p


class MULT(Element):
    TYPE = "MULT"
    SLICE = QLslice
    _params = _kn_params({"Nint": "Nint", "Norder": "Norder"})

    def __init__(self, name, l, knl, knsl):
        super().__init__(name)
        _check_len(l)
        self.L = float(l)
        self.KNL = [float(v) for v in knl] + [0.] * (11 - len(knl))
        self.KNSL = [float(v) for v in knsl] + [0.] * (11 - len(knsl))
        self.Nint = int(self.L / self.SLICE) + 1
        self.Norder = _norder(self.KNL, self.KNSL)

    def _fill(self, r):
        r["nint"], r["norder"] = self.Nint, self.Norder
        r["knl"][:] = self.KNL
        r["knsl"][:] = self.KNSL


class GMULT(MULT):
    TYPE = "GMULT"
    SLICE = BLslice
    _params = _kn_params({"ANGLE": "ANGLE", "E1": "E1", "E2": "E2", "Nint": "Nint"})

    def __init__(self, name, l, angle, e1, e2, knl, knsl):
        super().__init__(name, l, knl, knsl)
        self.ANGLE, self.E1, self.E2 = float(angle), float(e1), float(e2)

    def GetP(self, name):
        if name == "Norder":
            return self.Norder
        return super().GetP(name)

    def _fill(self, r):
        super()._fill(r)
        r["angle"], r["e1"], r["e2"] = self.ANGLE, self.E1, self.E2


class SBENDMULT(GMULT):
    TYPE = "SBENDMULT"
    _params = _kn_params({"ANGLE": "ANGLE", "E1": "E1", "E2": "E2", "Nint": "Nint", "Norder": "Norder"})


class SMULT(SBENDMULT):
    TYPE = "SMULT"
    SLICE = GLslice


class GSBENDMULT(SBENDMULT):
    TYPE = "GSBENDMULT"
    _params = _kn_params({"ANGLE": "ANGLE", "K0L": "K0L", "E1": "E1", "E2": "E2", "Nint": "Nint",
                          "Norder": "Norder"})

    def __init__(self, name, l, angle, k0l, e1, e2, knl, knsl):
        super().__init__(name, l, angle, e1, e2, knl, knsl)
        self.K0L = float(k0l)

    def _fill(self, r):
        super()._fill(r)
        r["k0l"] = self.K0L


class KICKER(Element):
    TYPE = "KICKER"
    _params = {"HKICK": "HKICK", "VKICK": "VKICK"}

    def __init__(self, name, l, hkick=0., vkick=0.):
        super().__init__(name)
        _check_len(l)
        self.L, self.HKICK, self.VKICK = float(l), float(hkick), float(vkick)

    def _fill(self, r):
        r["hkick"], r["vkick"] = self.HKICK, self.VKICK


class HKICKER(KICKER):
    TYPE = "HKICKER"
    _params = {"HKICK": "HKICK"}

    def __init__(self, name, l, hkick):
        super().__init__(name, l, hkick, 0.)


class VKICKER(KICKER):
    TYPE = "VKICKER"
    _params = {"VKICK": "VKICK"}

    def __init__(self, name, l, vkick):
        super().__init__(name, l, 0., vkick)


class RFCAV(Element):
    TYPE = "RFCAV"
    _params = {"VRF": "VRF", "FRF": "FRF", "PHASE0": "PHASE0"}

    def __init__(self, name, l, vrf, frf, phase0):
        super().__init__(name)
        _check_len(l)
        self.L, self.VRF, self.FRF, self.PHASE0 = float(l), float(vrf), float(frf), float(phase0)

    def _fill(self, r):
        r["vrf"], r["frf"], r["phase0"] = self.VRF, self.FRF, self.PHASE0


class CRABRF(RFCAV):
    TYPE = "CRABRF"


class CCMULT(RFCAV):
    """Crab-cavity multipoles; B1..A4 feed CCMULT_Pass's B2..A5 (simtrack.cpp:10943-10947)."""
    TYPE = "CCMULT"
    _params = {"VRF": "VRF", "FRF": "FRF", "PHASE0": "PHASE0", **{k: ("CC", i) for i, k in enumerate(
        ["B1", "A1", "B2", "A2", "B3", "A3", "B4", "A4"])}}

    def __init__(self, name, l, vrf, frf, phase0, b1, a1, b2, a2, b3, a3, b4, a4):
        super().__init__(name, l, vrf, frf, phase0)
        self.CC = [float(v) for v in (b1, a1, b2, a2, b3, a3, b4, a4)]

    def _fill(self, r):
        super()._fill(r)
        r["cc"][:] = self.CC


class LBT(Element):
    TYPE = "LBT"
    _params = {"THETA": "THETA"}

    def __init__(self, name, theta):
        super().__init__(name)
        self.THETA = float(theta)

    def _fill(self, r):
        r["theta"] = self.THETA


class ILBT(LBT):
    TYPE = "ILBT"


_BB_PARAMS = {"NP": "NP", "BBSCALE": "BBSCALE", "SIGMAL": "SIGMAL", "NSLICE": "NSLICE", "EMITX": "EMITX",
              "BETAX": "BETAX", "ALFAX": "ALFAX", "EMITY": "EMITY", "BETAY": "BETAY", "ALFAY": "ALFAY"}


class BEAMBEAM(Element):
    TYPE = "BEAMBEAM"
    _params = {"TREATMENT": "TREATMENT", **_BB_PARAMS}

    def __init__(self, name, treatment, np_, bbscale, sigmal, nslice, emitx, betax, alfax, emity, betay, alfay):
        super().__init__(name)
        self.TREATMENT, self.NP, self.BBSCALE, self.SIGMAL = int(treatment), float(np_), float(bbscale), float(sigmal)
        self.NSLICE = int(nslice)
        self.EMITX, self.BETAX, self.ALFAX = float(emitx), float(betax), float(alfax)
        self.EMITY, self.BETAY, self.ALFAY = float(emity), float(betay), float(alfay)

    def SetP(self, name, value):
        if name == "TREATMENT" and int(value) not in (4, 6):
            raise ValueError("beambeam treatment only can be 4 or 6")
        super().SetP(name, value)

    def _fill(self, r):
        r["treatment"], r["np"], r["bbscale"], r["sigmal"] = self.TREATMENT, self.NP, self.BBSCALE, self.SIGMAL
        r["nslice"] = self.NSLICE
        r["emitx"], r["betax"], r["alfax"] = self.EMITX, self.BETAX, self.ALFAX
        r["emity"], r["betay"], r["alfay"] = self.EMITY, self.BETAY, self.ALFAY


class CRBB(BEAMBEAM):
    TYPE = "CRBB"
    _params = {"THETA": "THETA", **_BB_PARAMS}

    def __init__(self, name, np_, theta, bbscale, sigmal, nslice, emitx, betax, alfax, emity, betay, alfay):
        super().__init__(name, 6, np_, bbscale, sigmal, nslice, emitx, betax, alfax, emity, betay, alfay)
        self.THETA = float(theta)

    def _fill(self, r):
        super()._fill(r)
        r["theta"] = self.THETA


class MATRIX(Element):
    TYPE = "MATRIX"

    def __init__(self, name, l, xco_in, xco_out, m66):
        super().__init__(name)
        self.L = float(l)
        self.XCO_IN, self.XCO_OUT = list(map(float, xco_in)), list(map(float, xco_out))
        self.M66 = list(map(float, np.ravel(m66)))
        self._params = {f"M{i + 1}{j + 1}": ("M66", 6 * i + j) for i in range(6) for j in range(6)}

    def _fill(self, r):
        r["m66"][:] = self.M66
        r["xco_in"][:] = self.XCO_IN
        r["xco_out"][:] = self.XCO_OUT


class ROTAT(Element):
    TYPE = "ROTAT"

    def __init__(self, name, l, n, angle):
        super().__init__(name)
        self.L, self.N, self.ANGLE = float(l), list(map(float, n)), float(angle)

    def _fill(self, r):
        r["rot_n"][:] = self.N
        r["rot_angle"] = self.ANGLE


class SNAKE(ROTAT):
    TYPE = "SNAKE"


class PATCH(Element):
    """PATCH (simtrack.cpp:12262-12334): a frame translation by DX, DY, DS
    and a rotation about y by THETA.  As in the reference the offsets live in
    the element's DX, DY, DS (which are not misalignments here) and L = DS."""
    TYPE = "PATCH"
    _params = {"DX": "DX", "DY": "DY", "DS": "DS", "THETA": "THETA"}

    def __init__(self, name, dx, dy, ds, theta):
        super().__init__(name)
        self.L = float(ds)
        self.DX, self.DY, self.DS, self.THETA = float(dx), float(dy), float(ds), float(theta)

    def _fill(self, r):
        r["theta"] = self.THETA


class SOLEN(Element):
    """SOLEN (simtrack.cpp:9645-9704), KS = Bz / brho, Nint = 1: a hard-edge
    solenoid (SOLEN_Pass cpp:5804, SOLEN_sPass cpp:8118); KS = 0 is a drift."""
    TYPE = "SOLEN"
    _params = {"KS": "KS"}

    def __init__(self, name, l, ks):
        super().__init__(name)
        _check_len(l, strict=True)
        self.L, self.KS = float(l), float(ks)
        self.Nint = 1

    def _fill(self, r):
        r["ks"] = self.KS
        r["nint"] = self.Nint


class TRANS(Element):
    """TRANS (simtrack.cpp:12056-12118): a frame translation by DX, DY, DS
    (the element's own DX, DY, DS, not misalignments; L = DS)."""
    TYPE = "TRANS"
    _params = {"DX": "DX", "DY": "DY", "DS": "DS"}

    def __init__(self, name, dx, dy, ds):
        super().__init__(name)
        self.L = float(ds)
        self.DX, self.DY, self.DS = float(dx), float(dy), float(ds)


class SROTAT(Element):
    """SROTAT (simtrack.cpp:12121-12165): frame rotation about s by PSI."""
    TYPE = "SROTAT"
    _params = {"PSI": "PSI"}

    def __init__(self, name, psi):
        super().__init__(name)
        self.PSI = float(psi)

    def _fill(self, r):
        r["psi"] = self.PSI


class YROTAT(Element):
    """YROTAT (simtrack.cpp:12168-12212): frame rotation about y by THETA."""
    TYPE = "YROTAT"
    _params = {"THETA": "THETA"}

    def __init__(self, name, theta):
        super().__init__(name)
        self.THETA = float(theta)

    def _fill(self, r):
        r["theta"] = self.THETA


class XROTAT(Element):
    """XROTAT (simtrack.cpp:12215-12259): frame rotation about x by PHI."""
    TYPE = "XROTAT"
    _params = {"PHI": "PHI"}

    def __init__(self, name, phi):
        super().__init__(name)
        self.PHI = float(phi)

    def _fill(self, r):
        r["phi"] = self.PHI


class COOLING(Element):
    """COOLING (simtrack.cpp:12009-12053): x, px, y, py scaled by 1 - ALPHA."""
    TYPE = "COOLING"
    _params = {"ALPHA": "ALPHA"}

    def __init__(self, name, alpha):
        super().__init__(name)
        self.ALPHA = float(alpha)

    def _fill(self, r):
        r["alpha"] = self.ALPHA


class LRBB(Element):
    """LRBB (simtrack.cpp:11220-11313): a long-range 4-D beam-beam kick of a
    beam separated by (SEPX, SEPY)."""
    TYPE = "LRBB"
    _params = {"NP": "NP", "BBSCALE": "BBSCALE", "SEPX": "SEPX", "SEPY": "SEPY", "SIGMAX": "SIGMAX",
               "SIGMAY": "SIGMAY"}

    def __init__(self, name, np_, bbscale, sepx, sepy, sigmax, sigmay):
        super().__init__(name)
        self.NP, self.BBSCALE = float(np_), float(bbscale)
        self.SEPX, self.SEPY, self.SIGMAX, self.SIGMAY = float(sepx), float(sepy), float(sigmax), float(sigmay)

    def _fill(self, r):
        r["np"], r["bbscale"] = self.NP, self.BBSCALE
        r["sepx"], r["sepy"], r["sigmax"], r["sigmay"] = self.SEPX, self.SEPY, self.SIGMAX, self.SIGMAY


class ELENS(Element):
    """ELENS (simtrack.cpp:11562-11675): NSLICE Gaussian electron-lens kicks
    between half drifts."""
    TYPE = "ELENS"
    _params = {"NE": "NE", "BBSCALE": "BBSCALE", "NSLICE": "NSLICE", "BETAE": "BETAE", "SIGMAX": "SIGMAX",
               "SIGMAY": "SIGMAY"}

    def __init__(self, name, l, ne, bbscale, nslice, betae, sigmax, sigmay):
        super().__init__(name)
        _check_len(l)
        self.L, self.NE, self.BBSCALE, self.NSLICE = float(l), float(ne), float(bbscale), int(nslice)
        self.BETAE, self.SIGMAX, self.SIGMAY = float(betae), float(sigmax), float(sigmay)

    def _fill(self, r):
        r["ne"], r["bbscale"], r["nslice"], r["betae"] = self.NE, self.BBSCALE, self.NSLICE, self.BETAE
        r["sigmax"], r["sigmay"] = self.SIGMAX, self.SIGMAY


class HELENS(Element):
    """HELENS (simtrack.cpp:11678-11781): a hollow electron lens, NSLICE
    radial kicks between half drifts."""
    TYPE = "HELENS"
    _params = {"NE": "NE", "BBSCALE": "BBSCALE", "NSLICE": "NSLICE", "BETAE": "BETAE", "ROUT": "ROUT",
               "RIN": "RIN"}

    def __init__(self, name, l, ne, bbscale, nslice, betae, rout, rin):
        super().__init__(name)
        _check_len(l)
        self.L, self.NE, self.BBSCALE, self.NSLICE = float(l), float(ne), float(bbscale), int(nslice)
        self.BETAE, self.ROUT, self.RIN = float(betae), float(rout), float(rin)

    def _fill(self, r):
        r["ne"], r["bbscale"], r["nslice"], r["betae"] = self.NE, self.BBSCALE, self.NSLICE, self.BETAE
        r["rout"], r["rin"] = self.ROUT, self.RIN


class HACDIP(Element):
    """HACDIP (simtrack.cpp:10278-10371): horizontal AC dipole, kick
    HKICKMAX sin(2 pi NUD (turn - TURNS) + PHID), ramped up between turns
    TURNS and TURNE."""
    TYPE = "HACDIP"
    _params = {"HKICKMAX": "HKICKMAX", "NUD": "NUD", "PHID": "PHID", "TURNS": "TURNS", "TURNE": "TURNE"}
    _kick = "HKICKMAX"

    def __init__(self, name, l, kickmax, nud, phid, turns, turne):
        super().__init__(name)
        _check_len(l)
        self.L = float(l)
        setattr(self, self._kick, float(kickmax))
        self.NUD, self.PHID, self.TURNS, self.TURNE = float(nud), float(phid), int(turns), int(turne)

    def SetP(self, name, value):
        if name in ("TURNS", "TURNE"):
            if name not in self._params:
                raise KeyError(f"{self.TYPE} does not have a parameter of {name}")
            setattr(self, name, int(value))
        else:
            super().SetP(name, value)

    def _fill(self, r):
        r["hkick" if self._kick == "HKICKMAX" else "vkick"] = getattr(self, self._kick)
        r["nud"], r["phid"], r["turns"], r["turne"] = self.NUD, self.PHID, self.TURNS, self.TURNE


class VACDIP(HACDIP):
    """VACDIP (simtrack.cpp:10374-10467): the vertical AC dipole."""
    TYPE = "VACDIP"
    _params = {"VKICKMAX": "VKICKMAX", "NUD": "NUD", "PHID": "PHID", "TURNS": "TURNS", "TURNE": "TURNE"}
    _kick = "VKICKMAX"


class HACMULT(Element):
    """HACMULT (simtrack.cpp:10104-10187): one normal multipole order Norder
    (0 = dipole) of integrated strength KLMAX modulated by sin(2 pi turn /
    TTURNS + PHI0), between half drifts."""
    TYPE = "HACMULT"
    _params = {"Norder": "Norder", "KLMAX": "KLMAX", "TTURNS": "TTURNS", "PHI0": "PHI0"}
    _k = "KLMAX"

    def __init__(self, name, l, norder, klmax, tturns, phi0):
        super().__init__(name)
        _check_len(l)
        self.L, self.Norder = float(l), int(norder)
        setattr(self, self._k, float(klmax))
        self.TTURNS, self.PHI0 = int(tturns), float(phi0)

    def SetP(self, name, value):
        if name == "TTURNS":
            self.TTURNS = int(value)
        else:
            super().SetP(name, value)

    def _fill(self, r):
        r["norder"], r["klmax"] = self.Norder, getattr(self, self._k)
        r["tturns"], r["phi0"] = self.TTURNS, self.PHI0


class VACMULT(HACMULT):
    """VACMULT (simtrack.cpp:10191-10274): the skew counterpart, KSLMAX."""
    TYPE = "VACMULT"
    _params = {"Norder": "Norder", "KSLMAX": "KSLMAX", "TTURNS": "TTURNS", "PHI0": "PHI0"}
    _k = "KSLMAX"


class WIGGLER(Element):
    """WIGGLER (simtrack.cpp:9707-9771): a sinusoidal wiggler field of peak
    B0 [T], wave numbers KX, KZ and phase PHIZ0; Pass integrates with
    Forest's implicit step (cpp:4400), sPass with Wu's splitting (cpp:4516)
    and leaves the spin unrotated."""
    TYPE = "WIGGLER"
    _params = {"Nint": "Nint", "B0": "B0", "KX": "KX", "KZ": "KZ", "PHIZ0": "PHIZ0"}

    def __init__(self, name, l, nint, b0, kx, kz, phiz0):
        super().__init__(name)
        if not l > 0:
            raise ValueError("Wiggler: L must be non-zero.")
        self.L, self.Nint = float(l), int(nint)
        self.B0, self.KX, self.KZ, self.PHIZ0 = float(b0), float(kx), float(kz), float(phiz0)

    def _fill(self, r):
        r["nint"] = self.Nint
        r["b0"], r["kx"], r["kz"], r["phiz0"] = self.B0, self.KX, self.KZ, self.PHIZ0


# -------------------------------------------------------------------- line

class Line:
    """Ordered list of elements (simtrack.cpp:12341-12431)."""

    def __init__(self, cells=()):
        self.Cell = list(cells)
        self.Update()

    def Update(self):
        s = 0.
        for c in self.Cell:
            s = c.L + s
            c.S = s
        self.Ncell = len(self.Cell)
        self.Length = s

    def Append(self, el):
        self.Cell.append(el)
        self.Update()

    def Delete(self, i):
        if 0 <= i < len(self.Cell):
            del self.Cell[i]
            self.Update()

    def Insert(self, i, el):
        if 0 <= i < len(self.Cell):
            self.Cell.insert(i, el)
            self.Update()

    def Replace(self, i, el):
        if 0 <= i < len(self.Cell):
            self.Cell[i] = el
            self.Update()

    def records(self):
        out = np.zeros(len(self.Cell), ELEMENT_DTYPE)
        for i, c in enumerate(self.Cell):
            out[i] = c.record()
        return out

    def __len__(self):
        return len(self.Cell)
