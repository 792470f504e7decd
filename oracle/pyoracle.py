"""ctypes loader of the CPU oracle (test infrastructure only).

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU baseline;
never by the simtrack_amd package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}


def build():
    subprocess.run(["sh", os.path.join(HERE, "build.sh")], check=True)


def load(glibc=False):
    """The oracle library: glibc=False the project libm, True the C library
    for every function, or one of "trig", "asin", "exp", "pow" for the C
    library in that function family only."""
    key = glibc if isinstance(glibc, str) else bool(glibc)
    if key in _libs:
        return _libs[key]
    if isinstance(glibc, str):
        name = f"liboracle_glibc_{glibc}.so"
    else:
        name = "liboracle_glibc.so" if glibc else "liboracle.so"
    path = os.path.join(HERE, "build", name)
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.orc_set_globals.argtypes = [P]
    lib.orc_get_yoshida.argtypes = [P]
    lib.orc_element_pass.argtypes = [P, P, I, ctypes.c_uint32]
    lib.orc_element_pass.restype = I
    lib.orc_track.argtypes = [P, I, P, I, I, P, P, P, I, P, I, I]
    lib.orc_track.restype = I
    lib.orc_radial_x04.argtypes = [P, I, P]
    lib.orc_draws_per_turn.argtypes = [P, I, I]
    lib.orc_draws_per_turn.restype = ctypes.c_uint64
    lib.orc_cerrf.argtypes = [D, D, P, P]
    lib.orc_bb4d.argtypes = [D, D, D, D, D, D, P, P]
    lib.orc_bb_soft_gaussian_4d.argtypes = [D, I, P, D, I, P, D, D, D, D]
    lib.orc_bb_soft_gaussian_6d.argtypes = [D, I, P, D, I, P, D, D, D, D, I, I, P, P]
    lib.orc_bb_soft_gaussian_6d_zb.argtypes = [D, I, P, D, I, P, D, D, D, D, I, I, P, P]
    lib.orc_zboundary_slices_gaussian.argtypes = [P, I, D]
    lib.orc_bb_pic_4d.argtypes = [D, I, P, D, I, P, D, D, D, D]
    lib.orc_bb_pic_4d_nonoffset.argtypes = [D, I, P, D, I, P, D, D, D, D, D, D]
    lib.orc_bb_pic_4d_fixed_gridfield.argtypes = [D, I, P, D, I, P, D, D, D, D, D, D]
    lib.orc_bb_pic_6d.argtypes = [D, I, P, D, I, P, D, D, D, D, I, I]
    lib.orc_fft2d_128.argtypes = [P, P, I]
    lib.orc_bb_slice_centers.argtypes = [I, D, P]
    lib.orc_fine_tune.argtypes = [I, P, D, D, P]
    lib.orc_closed_orbit.argtypes = [P, I, D, P, P, ctypes.POINTER(I)]
    lib.orc_spin_orbit.argtypes = [P, I, D, P, P, ctypes.POINTER(I)]
    lib.orc_track_dz.argtypes = [P, I, P, P, I, I, P, P, P]
    lib.orc_spin_tune.argtypes = [P, I, D, P, P]
    lib.orc_set_initial_spin.argtypes = [P, I, I, P, P]
    lib.orc_one_turn_matrix.argtypes = [P, I, P, D, P]
    lib.orc_element_maps.argtypes = [P, I, P, D, D, P, P]
    lib.orc_fma_scan.argtypes = [P, I, D, D, D, P, D, D, I, D, D, P, I]
    lib.orc_lyapunov_scan.argtypes = [P, I, D, D, D, I, D, D, P, I]
    lib.orc_action_diff.argtypes = [P, I, D, D, D, D, D, D, D, D, I, I, I, P, P]
    lib.orc_fine_tune_xpx.argtypes = [I, P, P, D, D, P]
    for f in ("orc_sin", "orc_cos", "orc_tan", "orc_asin", "orc_exp", "orc_cosh", "orc_sinh", "orc_log", "orc_atan",
              "orc_ugaussian_pinv"):
        getattr(lib, f).argtypes = [D]
        getattr(lib, f).restype = D
    lib.orc_powi.argtypes = [D, I]
    lib.orc_powi.restype = D
    if glibc is False:
        lib.orc_asin_small.restype = D
        lib.orc_asin_small.argtypes = [D]
        lib.orc_exp_phase.restype = D
        lib.orc_exp_phase.argtypes = [D, I]
    lib.orc_rng_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64]
    lib.orc_rng_normal.restype = D
    lib.orc_philox.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32, P]
    lib.orc_rad2_sample.argtypes = [ctypes.c_uint64, I, ctypes.c_uint64, D, P, P]
    lib.orc_rad2_energy.argtypes = [D]
    lib.orc_rad2_energy.restype = D
    _libs[key] = lib
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Oracle:
    """Reference restatement on the CPU: Track/Track_spin/Track_tbt/Track_Radial."""

    def __init__(self, gp, records, glibc=False, nthreads=None):
        self.lib = load(glibc)
        self.gp = gp
        self.records = np.ascontiguousarray(records)
        self.nthreads = nthreads or min(16, os.cpu_count() or 1)

    def _set(self):
        self.lib.orc_set_globals(ctypes.byref(self.gp))

    def track(self, x, nturn, spin=False, radial=False, stable=None, lost_turn=None, lost_post=None,
              tbt=False, bpm_index=-1, nthreads=None):
        self._set()
        ncol = 9 if spin else 6
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        n = x.shape[0]
        stable = np.ones(n, np.int32) if stable is None else np.ascontiguousarray(stable, np.int32).copy()
        lt = np.zeros(n, np.int32) if lost_turn is None else np.ascontiguousarray(lost_turn, np.int32).copy()
        lp = np.zeros(n, np.int32) if lost_post is None else np.ascontiguousarray(lost_post, np.int32).copy()
        xt = np.zeros((n, nturn, ncol)) if tbt else None
        mode = (1 if spin else 0) | (2 if radial else 0)
        rc = self.lib.orc_track(_p(self.records), len(self.records), _p(x), n, int(nturn), _p(stable), _p(lt),
                                _p(lp), mode, _p(xt), int(bpm_index), int(nthreads or self.nthreads))
        if rc != 0:
            raise RuntimeError(f"oracle track failed ({rc})")
        if tbt:
            return x, xt, stable, lt, lp
        return x, stable, lt, lp

    def element_pass(self, index, x, spin=False):
        self._set()
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        rec = self.records[index:index + 1]
        for p in range(x.shape[0]):
            row = x[p]
            rc = self.lib.orc_element_pass(_p(rec), row.ctypes.data_as(ctypes.c_void_p), 1 if spin else 0, p)
            if rc != 0:
                raise RuntimeError(f"oracle pass failed ({rc})")
        return x

    def closed_orbit(self, deltap):
        """Cal_Orbit_Num (simtrack.cpp:14947-15030): (x_co, x_end, iterations)."""
        self._set()
        xc, xe, it = np.zeros(6), np.zeros(6), ctypes.c_int()
        rc = self.lib.orc_closed_orbit(_p(self.records), len(self.records), deltap, _p(xc), _p(xe), ctypes.byref(it))
        if rc != 0:
            raise RuntimeError(f"oracle closed orbit failed ({rc})")
        return xc, xe, it.value

    def track_dz(self, x, nturn, dz0):
        """Track_wrt_OnMomentumPart (simtrack.cpp:20942-20966)."""
        self._set()
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        np_ = x.shape[0]
        s, lt, lp = np.ones(np_, np.int32), np.zeros(np_, np.int32), np.zeros(np_, np.int32)
        dz0 = np.ascontiguousarray(dz0, dtype=np.float64)
        rc = self.lib.orc_track_dz(_p(self.records), len(self.records), _p(dz0), _p(x), np_, int(nturn), _p(s),
                                   _p(lt), _p(lp))
        if rc != 0:
            raise RuntimeError(f"oracle track failed ({rc})")
        return x, s, lt, lp

    def spin_orbit(self, deltap):
        """Cal_Spin_Orbit (simtrack.cpp:20510-20572): (x0[9], n0[n, 3], iterations)."""
        self._set()
        x0, n0, it = np.zeros(9), np.zeros((len(self.records), 3)), ctypes.c_int()
        rc = self.lib.orc_spin_orbit(_p(self.records), len(self.records), deltap, _p(x0), _p(n0), ctypes.byref(it))
        if rc != 0:
            raise RuntimeError(f"oracle spin closed orbit failed ({rc})")
        return x0, n0, it.value

    def spin_tune(self, deltap):
        """Cal_Spin_Tune (simtrack.cpp:20574-20601): (tune, n0 of the last element)."""
        self._set()
        tune, n0 = np.zeros(1), np.zeros(3)
        rc = self.lib.orc_spin_tune(_p(self.records), len(self.records), deltap, _p(tune), _p(n0))
        if rc != 0:
            raise RuntimeError(f"oracle spin tune failed ({rc})")
        return float(tune[0]), n0

    def set_initial_spin(self, nturn, n0, x):
        """Set_Initial_Spin (simtrack.cpp:20603-20635): x (9) with its spin replaced."""
        self._set()
        x = np.array(x, dtype=np.float64).reshape(9).copy()
        n0 = np.ascontiguousarray(n0, dtype=np.float64).reshape(3)
        rc = self.lib.orc_set_initial_spin(_p(self.records), len(self.records), int(nturn), _p(n0), _p(x))
        if rc != 0:
            raise RuntimeError(f"oracle Set_Initial_Spin failed ({rc})")
        return x

    def one_turn_matrix(self, x_co, h=1e-6):
        """The 6x6 one-turn matrix of Cal_OneTurnMap (simtrack.cpp:15122-15137)
        by symmetric differences of step h about x_co, RFCAV voltages off."""
        self._set()
        m = np.zeros((6, 6))
        xc = np.ascontiguousarray(np.asarray(x_co, dtype=np.float64)[:6])
        rc = self.lib.orc_one_turn_matrix(_p(self.records), len(self.records), _p(xc), float(h), _p(m))
        if rc != 0:
            raise RuntimeError(f"oracle one-turn matrix failed ({rc})")
        return m

    def element_maps(self, x_co, x5, h=1e-6):
        """(T, X): every element's 6x6 map and the orbit after it (st_element_maps)."""
        self._set()
        n = len(self.records)
        T, X = np.zeros((n, 6, 6)), np.zeros((n, 6))
        xc = np.ascontiguousarray(np.asarray(x_co, dtype=np.float64)[:6])
        rc = self.lib.orc_element_maps(_p(self.records), n, _p(xc), float(x5), float(h), _p(T), _p(X))
        if rc:
            raise RuntimeError(f"orc_element_maps: {rc}")
        return T, X

    def fma_scan(self, deltap0, sigmax0, sigmay0, x_end, tune_low, tune_high, nturn, sigma_step, sigma_max):
        """Track_tbt_FMA (simtrack.cpp:21179-21241) one particle at a time: rows of 8."""
        self._set()
        cap = (int(sigma_max / sigma_step) + 2) ** 2
        rows = np.zeros((cap, 8))
        xe = np.ascontiguousarray(x_end[:4], dtype=np.float64)
        m = self.lib.orc_fma_scan(_p(self.records), len(self.records), deltap0, sigmax0, sigmay0, _p(xe), tune_low,
                                  tune_high, nturn, sigma_step, sigma_max, _p(rows), cap)
        if m < 0:
            raise RuntimeError(f"oracle fma scan failed ({m})")
        return rows[:m]

    def lyapunov_scan(self, deltap0, sigmax0, sigmay0, nturn, sigma_step, sigma_max):
        """Track_tbt_Lyapunov (simtrack.cpp:21243-21303) one pair at a time: rows of 3."""
        self._set()
        cap = (int(sigma_max / sigma_step) + 2) ** 2
        rows = np.zeros((cap, 3))
        m = self.lib.orc_lyapunov_scan(_p(self.records), len(self.records), deltap0, sigmax0, sigmay0, nturn,
                                       sigma_step, sigma_max, _p(rows), cap)
        if m < 0:
            raise RuntimeError(f"oracle lyapunov scan failed ({m})")
        return rows[:m]

    def action_diff(self, betx, alfx, bety, alfy, nsigma, phase, emitx0, emity0, nturn, nstep, npart, seed=1234.):
        """Track_tbt_ActionDiff (simtrack.cpp:21305-21383): (rows, seed after)."""
        self._set()
        rows = np.zeros((nstep, 6))
        sd = np.array([seed], dtype=np.float64)
        rc = self.lib.orc_action_diff(_p(self.records), len(self.records), betx, alfx, bety, alfy, nsigma, phase,
                                      emitx0, emity0, nturn, nstep, npart, _p(sd), _p(rows))
        if rc != 0:
            raise RuntimeError(f"oracle action diffusion failed ({rc})")
        return rows, float(sd[0])


# ------------------------------------------- strong-strong beam-beam

def fine_tune(x, tune_lower, tune_upper):
    """FineTuneFinder (simtrack.cpp:2288-2319) of the samples x."""
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = ctypes.c_double()
    lib.orc_fine_tune(len(x), _p(x), float(tune_lower), float(tune_upper), ctypes.byref(out))
    return out.value


def _beam(x):
    return np.ascontiguousarray(np.array(x, dtype=np.float64, copy=True).reshape(-1, 6))


def bb_soft_gaussian_4d(Np1, x1, Np2, x2, gamma1, gamma2, bbscale1, bbscale2):
    """BeamBeam4D_SoftGaussian (simtrack.cpp:22846): new (x1, x2)."""
    lib = load()
    a, b = _beam(x1), _beam(x2)
    lib.orc_bb_soft_gaussian_4d(Np1, len(a), _p(a), Np2, len(b), _p(b), gamma1, gamma2, bbscale1, bbscale2)
    return a, b


def bb_soft_gaussian_6d(Np1, x1, Np2, x2, gamma1, gamma2, bbscale1, bbscale2, nslice1, nslice2):
    """BeamBeam6D_SoftGaussian (simtrack.cpp:22874): new (x1, x2) sorted by z
    descending, and the slice centres (Z1, Z2)."""
    lib = load()
    a, b = _beam(x1), _beam(x2)
    z1, z2 = np.zeros(nslice1), np.zeros(nslice2)
    lib.orc_bb_soft_gaussian_6d(Np1, len(a), _p(a), Np2, len(b), _p(b), gamma1, gamma2, bbscale1, bbscale2, nslice1,
                                nslice2, _p(z1), _p(z2))
    return a, b, z1, z2


def bb_soft_gaussian_6d_zb(Np1, x1, Np2, x2, gamma1, gamma2, bbscale1, bbscale2, zb1, zb2):
    """BeamBeam6D_SoftGaussian with slice boundaries (simtrack.cpp:23134): new (x1, x2)."""
    lib = load()
    a, b = _beam(x1), _beam(x2)
    zb1 = np.ascontiguousarray(zb1, dtype=np.float64)
    zb2 = np.ascontiguousarray(zb2, dtype=np.float64)
    lib.orc_bb_soft_gaussian_6d_zb(Np1, len(a), _p(a), Np2, len(b), _p(b), gamma1, gamma2, bbscale1, bbscale2,
                                   len(zb1) - 1, len(zb2) - 1, _p(zb1), _p(zb2))
    return a, b


def zboundary_slices_gaussian(nslice, sigmal):
    """Zboundary_slices_Gaussian (simtrack.cpp:3157): nslice + 1 boundaries, descending."""
    zb = np.zeros(nslice + 1)
    load().orc_zboundary_slices_gaussian(_p(zb), nslice, sigmal)
    return zb


def bb4d(x, y, gamma, N, sigmax, sigmay):
    """BB4D (simtrack.cpp:6325-6382): (Dpx, Dpy)."""
    a, b = ctypes.c_double(), ctypes.c_double()
    load().orc_bb4d(x, y, gamma, N, sigmax, sigmay, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def bb_pic_4d_nonoffset(Np1, x1, Np2, x2, gamma1, gamma2, bbscale1, bbscale2, hx, hy):
    """BeamBeam4D_PIC_NonOffset (simtrack.cpp:24226) on the 64 x 64 grid: new (x1, x2)."""
    a, b = _beam(x1), _beam(x2)
    load().orc_bb_pic_4d_nonoffset(Np1, len(a), _p(a), Np2, len(b), _p(b), gamma1, gamma2, bbscale1, bbscale2,
                                   hx, hy)
    return a, b


def bb_pic_4d_fixed_gridfield(Np1, x1, Np2, x2, gamma1, gamma2, bbscale1, bbscale2, hx, hy):
    """The fixed-grid collision by the full-grid field path (a check of bb_pic_4d_nonoffset)."""
    a, b = _beam(x1), _beam(x2)
    load().orc_bb_pic_4d_fixed_gridfield(Np1, len(a), _p(a), Np2, len(b), _p(b), gamma1, gamma2, bbscale1,
                                         bbscale2, hx, hy)
    return a, b


def bb_pic_4d(Np1, x1, Np2, x2, gamma1, gamma2, bbscale1, bbscale2):
    """BeamBeam4D_PIC (simtrack.cpp:23286): new (x1, x2)."""
    a, b = _beam(x1), _beam(x2)
    load().orc_bb_pic_4d(Np1, len(a), _p(a), Np2, len(b), _p(b), gamma1, gamma2, bbscale1, bbscale2)
    return a, b


def bb_pic_6d(Np1, x1, Np2, x2, gamma1, gamma2, bbscale1, bbscale2, nslice1, nslice2):
    """BeamBeam6D_PIC (simtrack.cpp:23380): new (x1, x2), sorted by z descending."""
    a, b = _beam(x1), _beam(x2)
    load().orc_bb_pic_6d(Np1, len(a), _p(a), Np2, len(b), _p(b), gamma1, gamma2, bbscale1, bbscale2, nslice1,
                         nslice2)
    return a, b


def fft2d_128(re, im, direction):
    """FFT2D (simtrack.cpp:2649) of a 128 x 128 complex grid: new (re, im)."""
    x = np.ascontiguousarray(np.array(re, dtype=np.float64, copy=True).reshape(128, 128))
    y = np.ascontiguousarray(np.array(im, dtype=np.float64, copy=True).reshape(128, 128))
    load().orc_fft2d_128(_p(x), _p(y), int(direction))
    return x, y


def rad2_sample(seed, n, draw, nr):
    """radiate2's photon sampler for particles 0..n-1 at one draw number:
    (photon counts, summed energies in units of the critical energy)."""
    cnt = np.zeros(n, dtype=np.int32)
    et = np.zeros(n)
    load().orc_rad2_sample(int(seed), int(n), int(draw), float(nr), _p(cnt), _p(et))
    return cnt, et


def rad2_energy(r):
    """Photon energy / critical energy from a uniform r (radiate2's inverse CDF)."""
    return load().orc_rad2_energy(float(r))


# --------------------------------------------------- operation counts

_opc = None


def opcount(records, index, gp, x, spin=False):
    """Algorithmic operations of element `index`'s map on particle state x
    (oracle/opcount.cpp): (operations, pt-derived operations, pt-cache
    updates)."""
    global _opc
    if _opc is None:
        path = os.path.join(HERE, "build", "liboracle_opcount.so")
        if not os.path.exists(path):
            build()
        _opc = ctypes.CDLL(path)
        _opc.opc_element_ops.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p]
        _opc.opc_element_ops.restype = ctypes.c_int
    rec = np.ascontiguousarray(records[index:index + 1])
    xx = np.ascontiguousarray(x, dtype=np.float64)
    c = np.zeros(3, np.int64)
    rc = _opc.opc_element_ops(_p(rec), ctypes.byref(gp), _p(xx), 1 if spin else 0, _p(c))
    if rc != 0:
        raise RuntimeError(f"opcount failed ({rc})")
    return int(c[0]), int(c[1]), int(c[2])


# ------------------------------------------- first-order (dual) maps

_dual = None


def _dual_lib():
    global _dual
    if _dual is None:
        path = os.path.join(HERE, "build", "liboracle_dual.so")
        if not os.path.exists(path):
            build()
        _dual = ctypes.CDLL(path)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        _dual.dual_one_turn_matrix.argtypes = [P, I, P, P, P]
        _dual.dual_element_maps.argtypes = [P, I, P, P, D, P]
    return _dual


def dual_one_turn_matrix(records, gp, x_co):
    """The exact first-order one-turn matrix about x_co, RF voltages off
    (oracle/dualmap.cpp)."""
    rec = np.ascontiguousarray(records)
    m = np.zeros((6, 6))
    xc = np.ascontiguousarray(np.asarray(x_co, dtype=np.float64)[:6])
    rc = _dual_lib().dual_one_turn_matrix(_p(rec), len(rec), ctypes.byref(gp), _p(xc), _p(m))
    if rc:
        raise RuntimeError(f"dual_one_turn_matrix: {rc}")
    return m


def dual_element_maps(records, gp, X, x5):
    """Every element's exact first-order matrix about the orbit X (n x 6,
    after each element), z = 0 and x[5] = x5, RF voltages off."""
    rec = np.ascontiguousarray(records)
    Xc = np.ascontiguousarray(X, dtype=np.float64)
    T = np.zeros((len(rec), 6, 6))
    rc = _dual_lib().dual_element_maps(_p(rec), len(rec), ctypes.byref(gp), _p(Xc), float(x5), _p(T))
    if rc:
        raise RuntimeError(f"dual_element_maps: {rc}")
    return T
