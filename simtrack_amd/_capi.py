"""ctypes binding of the C ABI in include/simtrack_amd.h.

Python is glue here: the tracking runs in the HIP library.  Loading fails
loudly when the library is missing; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import build as _build

ELEMENT_DTYPE = np.dtype([
    ("type", "<i4"), ("nint", "<i4"), ("norder", "<i4"), ("nslice", "<i4"),
    ("treatment", "<i4"), ("reserved", "<i4", (3,)),
    ("l", "<f8"), ("angle", "<f8"), ("e1", "<f8"), ("e2", "<f8"), ("k0l", "<f8"),
    ("knl", "<f8", (11,)), ("knsl", "<f8", (11,)),
    ("dx", "<f8"), ("dy", "<f8"), ("ds", "<f8"),
    ("dtheta", "<f8"), ("dphi", "<f8"), ("dpsi", "<f8"),
    ("apx", "<f8"), ("apy", "<f8"),
    ("vrf", "<f8"), ("frf", "<f8"), ("phase0", "<f8"),
    ("hkick", "<f8"), ("vkick", "<f8"),
    ("theta", "<f8"),
    ("np", "<f8"), ("bbscale", "<f8"), ("sigmal", "<f8"),
    ("emitx", "<f8"), ("betax", "<f8"), ("alfax", "<f8"),
    ("emity", "<f8"), ("betay", "<f8"), ("alfay", "<f8"),
    ("cc", "<f8", (8,)),
    ("rot_n", "<f8", (3,)), ("rot_angle", "<f8"),
    ("m66", "<f8", (36,)),
    ("xco_in", "<f8", (6,)), ("xco_out", "<f8", (6,)),
    ("ks", "<f8"),
    ("psi", "<f8"), ("phi", "<f8"), ("alpha", "<f8"),
    ("sepx", "<f8"), ("sepy", "<f8"), ("sigmax", "<f8"), ("sigmay", "<f8"),
    ("ne", "<f8"), ("betae", "<f8"), ("rout", "<f8"), ("rin", "<f8"),
    ("nud", "<f8"), ("phid", "<f8"), ("turns", "<f8"), ("turne", "<f8"),
    ("klmax", "<f8"), ("tturns", "<f8"), ("phi0", "<f8"),
    ("b0", "<f8"), ("kx", "<f8"), ("kz", "<f8"), ("phiz0", "<f8"),
])


class Globals(ctypes.Structure):
    """st_globals: the GP fields the element maps read (simtrack.h:28-68)."""
    _fields_ = [
        ("gamma", ctypes.c_double), ("beta", ctypes.c_double),
        ("energy0", ctypes.c_double), ("energy", ctypes.c_double),
        ("cp", ctypes.c_double), ("p", ctypes.c_double), ("m0", ctypes.c_double),
        ("m", ctypes.c_double), ("q", ctypes.c_double), ("brho", ctypes.c_double),
        ("G", ctypes.c_double),
        ("Q", ctypes.c_int32), ("A", ctypes.c_int32), ("H_expand", ctypes.c_int32),
        ("radiate", ctypes.c_int32), ("quad_fringe", ctypes.c_int32),
        ("reserved", ctypes.c_int32), ("rng_seed", ctypes.c_uint64),
    ]


# element type codes (enum st_element_type)
TYPES = {
    "DRIFT": 1, "SBEND": 2, "GSBEND": 3, "QUAD": 4, "SKEWQ": 5, "SEXT": 6, "OCT": 7,
    "MULT": 8, "GMULT": 9, "SBENDMULT": 10, "SMULT": 11, "GSBENDMULT": 12,
    "KICKER": 13, "HKICKER": 14, "VKICKER": 15, "MARKER": 16, "BPM": 17, "HBPM": 18,
    "VBPM": 19, "RFCAV": 20, "CRABRF": 21, "CCMULT": 22, "LBT": 23, "ILBT": 24,
    "BEAMBEAM": 25, "CRBB": 26, "MATRIX": 27, "ROTAT": 28, "SNAKE": 29, "PATCH": 30, "SOLEN": 31,
    "TRANS": 32, "SROTAT": 33, "YROTAT": 34, "XROTAT": 35, "COOLING": 36, "LRBB": 37, "ELENS": 38,
    "HELENS": 39, "HACDIP": 40, "VACDIP": 41, "HACMULT": 42, "VACMULT": 43, "WIGGLER": 44,
}
TRACK_SPIN = 1
TRACK_RADIAL = 2
TRACK_BARE = 4
TRACK_PERF = 8    # perf-mode kernels: contraction and reciprocal division, not bit-equal
# limits of st_track_device_noalloc's fixed scratch (include/simtrack_amd.h)
NOALLOC_MAX_PARTICLES = 1 << 20
NOALLOC_MAX_ELEMENTS = 16384

# every symbol declared in include/simtrack_amd.h
EXPORTS = [
    "st_globals_default", "st_globals_set_energy", "st_abi_version", "st_struct_sizes",
    "st_last_error", "st_ctx_create", "st_ctx_destroy", "st_table_build", "st_table_destroy",
    "st_table_size", "st_track", "st_track_spin", "st_track_tbt", "st_track_radial",
    "st_element_pass", "st_track_device", "st_track_device_noalloc", "st_ctx_sync", "st_ctx_set_ss_lanes", "st_last_gpu_seconds",
    "st_ctx_set_timing", "st_ctx_kernel_stats", "st_da_angles", "st_da_bisect", "st_da_uniform",
    "st_fine_tune", "st_fine_tune_device", "st_batch_create", "st_batch_destroy", "st_batch_default_config",
    "st_batch_track", "st_batch_track_device", "st_batch_autotune", "st_mlp_create", "st_mlp_destroy",
    "st_mlp_num_params", "st_mlp_get_params", "st_mlp_set_params", "st_mlp_forward", "st_mlp_gradients",
    "st_mlp_train_step", "st_batch_track_tbt_device", "st_surrogate_features", "st_closed_orbit", "st_fma_scan",
    "st_one_turn_matrix", "st_element_maps", "st_lyapunov_scan", "st_fine_tune_window_device", "st_action_diff",
    "st_bb_soft_gaussian_4d", "st_bb_soft_gaussian_6d", "st_bb_soft_gaussian_6d_zb", "st_bb_soft_gaussian_4d_host",
    "st_bb_soft_gaussian_6d_host", "st_bb_soft_gaussian_6d_zb_host", "st_zboundary_slices_gaussian",
    "st_bb_pic_4d", "st_bb_pic_6d", "st_bb_pic_4d_host", "st_bb_pic_6d_host", "st_bb_pic_4d_nonoffset",
    "st_bb_pic_4d_nonoffset_host", "st_spin_orbit", "st_spin_tune",
    "st_set_initial_spin", "st_track_dz",
]


class BatchConfig(ctypes.Structure):
    """st_batch_config: launch shape of the persistent multi-table kernel."""
    _fields_ = [("wg_waves", ctypes.c_int), ("waves_per_simd", ctypes.c_int), ("turn_block", ctypes.c_int)]

    def __repr__(self):
        return f"BatchConfig(wg_waves={self.wg_waves}, waves_per_simd={self.waves_per_simd}, " \
               f"turn_block={self.turn_block})"
ST_DA_TRACK_DA, ST_DA_TRACK_DA1, ST_DA_TRACK_DA_RADIAL, ST_DA_TRACK_DA2, ST_DA_TRACK_DA3 = 0, 1, 2, 3, 4
ST_DA_RADIAL, ST_DA_SINGLE = 1, 2

_lib = None


class STError(RuntimeError):
    pass


def _load_torch_first():
    """PyTorch ships its own HIP runtime (libamdhip64, libhsa-runtime64 under
    torch/lib).  When this library, linked against the system ROCm, is loaded
    into a process before torch, torch's device initialisation later fails
    with "No HIP GPUs are available"; loading torch first avoids that.  Torch
    is optional: without it nothing is loaded."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load(path=None):
    """Load libsimtrack_amd.so (raises if it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or _build.library_path()
    if not os.path.exists(path):
        raise STError(f"HIP library not built: {path} (run simtrack_amd/build.py)")
    _load_torch_first()
    lib = ctypes.CDLL(path)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.st_last_error.restype = ctypes.c_char_p
    lib.st_abi_version.restype = I
    lib.st_struct_sizes.argtypes = [ctypes.POINTER(I), ctypes.POINTER(I)]
    lib.st_globals_default.argtypes = [ctypes.POINTER(Globals)]
    lib.st_globals_set_energy.argtypes = [ctypes.POINTER(Globals), D, D, D, D]
    lib.st_ctx_create.argtypes = [I, ctypes.POINTER(P)]
    lib.st_ctx_destroy.argtypes = [P]
    lib.st_ctx_sync.argtypes = [P]
    lib.st_table_build.argtypes = [P, ctypes.POINTER(Globals), P, I, ctypes.POINTER(P)]
    lib.st_table_destroy.argtypes = [P]
    lib.st_table_size.argtypes = [P, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_int64)]
    for f in ("st_track", "st_track_spin", "st_track_radial"):
        getattr(lib, f).argtypes = [P, P, I, I, P, P, P]
    lib.st_track_tbt.argtypes = [P, P, I, I, P, I, I, P, P, P]
    lib.st_track_dz.argtypes = [P, P, P, I, I, P, P, P]
    lib.st_element_pass.argtypes = [P, I, P, I, I]
    lib.st_track_device.argtypes = [P, P, I, I, I, P, P, I]
    lib.st_track_device_noalloc.argtypes = [P, P, I, I, I, P, P, I]
    lib.st_last_gpu_seconds.argtypes = [P, ctypes.POINTER(D)]
    lib.st_ctx_set_timing.argtypes = [P, I]
    lib.st_ctx_set_ss_lanes.argtypes = [P, I]
    lib.st_ctx_kernel_stats.argtypes = [P, ctypes.POINTER(D), ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_int64)]
    lib.st_da_angles.argtypes = [I, P, ctypes.POINTER(I), ctypes.POINTER(D)]
    lib.st_da_bisect.argtypes = [P, I, D, D, D, P, I, D, I, P]
    lib.st_da_uniform.argtypes = [P, I, D, D, D, D, D, D, I, I, P, P, ctypes.POINTER(I)]
    lib.st_fine_tune.argtypes = [P, P, P, I, I, D, D, P]
    lib.st_fine_tune_device.argtypes = [P, P, P, ctypes.c_int64, ctypes.c_int64, I, I, D, D, P]
    lib.st_fine_tune_window_device.argtypes = [P, P, P, ctypes.c_int64, ctypes.c_int64, I, I, D, D, P, P]
    BC = ctypes.POINTER(BatchConfig)
    lib.st_batch_create.argtypes = [P, P, I, ctypes.POINTER(P)]
    lib.st_batch_destroy.argtypes = [P]
    lib.st_batch_default_config.argtypes = [BC]
    lib.st_batch_track.argtypes = [P, P, P, I, P, P, P, I, BC]
    lib.st_batch_track_device.argtypes = [P, P, P, I, I, P, P, I, BC]
    lib.st_batch_autotune.argtypes = [P, P, P, I, I, BC, P, ctypes.POINTER(I)]
    lib.st_batch_track_tbt_device.argtypes = [P, P, P, I, I, P, P, P, I, BC]
    lib.st_surrogate_features.argtypes = [P, P, I, P, I, D, D, D, D, P]
    lib.st_closed_orbit.argtypes = [P, ctypes.POINTER(Globals), P, I, D, P, P, ctypes.POINTER(I)]
    lib.st_one_turn_matrix.argtypes = [P, ctypes.POINTER(Globals), P, I, P, D, P]
    lib.st_element_maps.argtypes = [P, ctypes.POINTER(Globals), P, I, P, D, D, P, P]
    lib.st_spin_orbit.argtypes = [P, ctypes.POINTER(Globals), P, I, D, P, P, ctypes.POINTER(I)]
    lib.st_spin_tune.argtypes = [P, ctypes.POINTER(Globals), P, I, D, P, P]
    lib.st_set_initial_spin.argtypes = [P, ctypes.POINTER(Globals), P, I, I, P, P]
    for f in ("st_bb_soft_gaussian_4d", "st_bb_soft_gaussian_4d_host", "st_bb_pic_4d", "st_bb_pic_4d_host"):
        getattr(lib, f).argtypes = [P, D, I, P, D, I, P, D, D, D, D]
    for f in ("st_bb_soft_gaussian_6d", "st_bb_soft_gaussian_6d_host", "st_bb_soft_gaussian_6d_zb",
              "st_bb_soft_gaussian_6d_zb_host", "st_bb_pic_6d", "st_bb_pic_6d_host"):
        getattr(lib, f).argtypes = [P, D, I, P, D, I, P, D, D, D, D, I, I, P, P]
    for f in ("st_bb_pic_4d_nonoffset", "st_bb_pic_4d_nonoffset_host"):
        getattr(lib, f).argtypes = [P, D, I, P, D, I, P, D, D, D, D, I, I, D, D]
    lib.st_zboundary_slices_gaussian.argtypes = [P, I, D]
    lib.st_fma_scan.argtypes = [P, D, D, D, P, D, D, I, D, D, P, I, ctypes.POINTER(I)]
    lib.st_lyapunov_scan.argtypes = [P, D, D, D, I, D, D, P, I, ctypes.POINTER(I)]
    lib.st_action_diff.argtypes = [P, D, D, D, D, D, D, D, D, I, I, I, P, P]
    F = ctypes.c_float
    lib.st_mlp_create.argtypes = [P, I, I, ctypes.c_uint64, ctypes.POINTER(P)]
    lib.st_mlp_destroy.argtypes = [P]
    lib.st_mlp_num_params.argtypes = [P, ctypes.POINTER(ctypes.c_int64)]
    lib.st_mlp_get_params.argtypes = [P, P, ctypes.c_int64]
    lib.st_mlp_set_params.argtypes = [P, P, ctypes.c_int64]
    lib.st_mlp_forward.argtypes = [P, P, I, P]
    lib.st_mlp_gradients.argtypes = [P, P, P, I, P, ctypes.POINTER(F)]
    lib.st_mlp_train_step.argtypes = [P, P, P, I, F, ctypes.POINTER(F)]
    es, gs = I(), I()
    lib.st_struct_sizes(ctypes.byref(es), ctypes.byref(gs))
    if es.value != ELEMENT_DTYPE.itemsize or gs.value != ctypes.sizeof(Globals):
        raise STError(f"ABI layout mismatch: element {es.value} vs {ELEMENT_DTYPE.itemsize}, "
                      f"globals {gs.value} vs {ctypes.sizeof(Globals)}")
    _lib = lib
    return lib


def check(rc, what=""):
    if rc != 0:
        msg = _lib.st_last_error().decode() if _lib is not None else ""
        raise STError(f"{what} failed ({rc}): {msg}")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Context:
    """A device context (HIP stream and work buffers) on one GPU."""

    def __init__(self, device=0):
        self.lib = load()
        self.h = ctypes.c_void_p()
        check(self.lib.st_ctx_create(device, ctypes.byref(self.h)), "st_ctx_create")

    def close(self):
        if self.h:
            self.lib.st_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def gpu_seconds(self):
        v = ctypes.c_double()
        check(self.lib.st_last_gpu_seconds(self.h, ctypes.byref(v)), "st_last_gpu_seconds")
        return v.value

    def sync(self):
        check(self.lib.st_ctx_sync(self.h), "st_ctx_sync")

    def set_ss_lanes(self, lanes):
        """At most `lanes` strong-strong slice pairs in flight (1..8)."""
        check(self.lib.st_ctx_set_ss_lanes(self.h, int(lanes)), "st_ctx_set_ss_lanes")

    def set_timing(self, on=True):
        check(self.lib.st_ctx_set_timing(self.h, 1 if on else 0), "st_ctx_set_timing")

    def fine_tune_device(self, d_x, stride_sig, stride, n, nsig, lower, upper, d_tune, d_px=None, d_window=None):
        if d_window:
            check(self.lib.st_fine_tune_window_device(
                self.h, ctypes.c_void_p(d_x), ctypes.c_void_p(d_px) if d_px else None, int(stride_sig), int(stride),
                int(n), int(nsig), float(lower), float(upper), ctypes.c_void_p(d_window), ctypes.c_void_p(d_tune)),
                "st_fine_tune_window_device")
            return
        check(self.lib.st_fine_tune_device(self.h, ctypes.c_void_p(d_x), ctypes.c_void_p(d_px) if d_px else None,
                                           int(stride_sig), int(stride), int(n), int(nsig), float(lower), float(upper),
                                           ctypes.c_void_p(d_tune)), "st_fine_tune_device")

    def surrogate_features(self, d_x0, ncol, d_tunes, n, beta_x, alpha_x, beta_y, alpha_y, d_feat):
        check(self.lib.st_surrogate_features(self.h, ctypes.c_void_p(d_x0), int(ncol), ctypes.c_void_p(d_tunes), int(n),
                                             float(beta_x), float(alpha_x), float(beta_y), float(alpha_y),
                                             ctypes.c_void_p(d_feat)), "st_surrogate_features")

    def fine_tune(self, x, lower, upper, px=None):
        """FineTuneFinder (simtrack.cpp:2288, x/px form cpp:2321) of each row of x."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
        if px is not None:
            px = np.ascontiguousarray(np.atleast_2d(px), dtype=np.float64)
            assert px.shape == x.shape
        out = np.zeros(x.shape[0])
        check(self.lib.st_fine_tune(self.h, _ptr(x), _ptr(px), x.shape[1], x.shape[0], float(lower), float(upper),
                                    _ptr(out)), "st_fine_tune")
        return out

    def kernel_stats(self):
        """(summed kernel seconds, launches, launched particle-turns) since the last call."""
        s, n, w = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        check(self.lib.st_ctx_kernel_stats(self.h, ctypes.byref(s), ctypes.byref(n), ctypes.byref(w)),
              "st_ctx_kernel_stats")
        return s.value, n.value, w.value


class Table:
    """Packed device table of a line (st_table_build)."""

    def __init__(self, ctx, gp, records):
        self.ctx = ctx
        self.lib = ctx.lib
        self.records = np.ascontiguousarray(records, dtype=ELEMENT_DTYPE)
        self.n = len(self.records)
        self.h = ctypes.c_void_p()
        check(self.lib.st_table_build(ctx.h, ctypes.byref(gp), _ptr(self.records), self.n,
                                      ctypes.byref(self.h)), "st_table_build")

    def close(self):
        if self.h:
            self.lib.st_table_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def track(self, x, nturn, stable=None, lost_turn=None, lost_post=None, spin=False, radial=False):
        """Track (spin: Track_spin, radial: Track_Radial).  Arrays are updated in place."""
        ncol = 9 if spin else 6
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.ndim == 2 and x.shape[1] == ncol
        npart = x.shape[0]
        stable = np.ones(npart, np.int32) if stable is None else np.ascontiguousarray(stable, np.int32)
        lost_turn = np.zeros(npart, np.int32) if lost_turn is None else np.ascontiguousarray(lost_turn, np.int32)
        lost_post = np.zeros(npart, np.int32) if lost_post is None else np.ascontiguousarray(lost_post, np.int32)
        f = self.lib.st_track_spin if spin else (self.lib.st_track_radial if radial else self.lib.st_track)
        assert not (spin and radial)
        check(f(self.h, _ptr(x), npart, int(nturn), _ptr(stable), _ptr(lost_turn), _ptr(lost_post)), "st_track")
        return x, stable, lost_turn, lost_post

    def track_dz(self, x, nturn, dz0, stable=None, lost_turn=None, lost_post=None):
        """Track_wrt_OnMomentumPart: Track with dz0[j] subtracted from z after
        element j.  Arrays are updated in place."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.ndim == 2 and x.shape[1] == 6
        dz0 = np.ascontiguousarray(dz0, dtype=np.float64)
        npart = x.shape[0]
        stable = np.ones(npart, np.int32) if stable is None else np.ascontiguousarray(stable, np.int32)
        lost_turn = np.zeros(npart, np.int32) if lost_turn is None else np.ascontiguousarray(lost_turn, np.int32)
        lost_post = np.zeros(npart, np.int32) if lost_post is None else np.ascontiguousarray(lost_post, np.int32)
        check(self.lib.st_track_dz(self.h, _ptr(dz0), _ptr(x), npart, int(nturn), _ptr(stable), _ptr(lost_turn),
                                   _ptr(lost_post)), "st_track_dz")
        return x, stable, lost_turn, lost_post

    def track_tbt(self, x, nturn, bpm_index=-1, spin=False, x_tbt=None):
        ncol = 9 if spin else 6
        x = np.ascontiguousarray(x, dtype=np.float64)
        npart = x.shape[0]
        stable = np.ones(npart, np.int32)
        lost_turn = np.zeros(npart, np.int32)
        lost_post = np.zeros(npart, np.int32)
        if x_tbt is None:
            x_tbt = np.zeros((npart, nturn, ncol))
        check(self.lib.st_track_tbt(self.h, _ptr(x), npart, int(nturn), _ptr(x_tbt), int(bpm_index),
                                    1 if spin else 0, _ptr(stable), _ptr(lost_turn), _ptr(lost_post)),
              "st_track_tbt")
        return x, x_tbt, stable, lost_turn, lost_post

    def element_pass(self, index, x, spin=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        check(self.lib.st_element_pass(self.h, int(index), _ptr(x), x.shape[0], 1 if spin else 0),
              "st_element_pass")
        return x

    def track_device(self, d_x, npart, turn0, nturn, d_state, d_lost, flags=0):
        """Device-resident tracking; d_* are integer device addresses (e.g. torch data_ptr())."""
        check(self.lib.st_track_device(self.h, ctypes.c_void_p(d_x), int(npart), int(turn0), int(nturn),
                                       ctypes.c_void_p(d_state), ctypes.c_void_p(d_lost), int(flags)),
              "st_track_device")

    def track_device_noalloc(self, d_x, npart, turn0, nturn, d_state, d_lost, flags=0):
        """As track_device, with the work buffers in the library's fixed
        per-GPU scratch (st_track_device_noalloc): no device allocation."""
        check(self.lib.st_track_device_noalloc(self.h, ctypes.c_void_p(d_x), int(npart), int(turn0), int(nturn),
                                               ctypes.c_void_p(d_state), ctypes.c_void_p(d_lost), int(flags)),
              "st_track_device_noalloc")


class Batch:
    """Tables of several lattices (seeds) tracked together by one persistent
    launch per turn block (st_batch_*)."""

    def __init__(self, ctx, tables):
        self.ctx = ctx
        self.lib = ctx.lib
        self.tables = list(tables)          # keeps the tables alive
        arr = (ctypes.c_void_p * len(self.tables))(*[t.h.value for t in self.tables])
        self.h = ctypes.c_void_p()
        check(self.lib.st_batch_create(ctx.h, arr, len(self.tables), ctypes.byref(self.h)), "st_batch_create")

    def close(self):
        if self.h:
            self.lib.st_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def default_config():
        cfg = BatchConfig()
        load().st_batch_default_config(ctypes.byref(cfg))
        return cfg

    def track(self, x, offsets, nturn, spin=False, cfg=None, stable=None, lost_turn=None, lost_post=None):
        """Track rows offsets[k]..offsets[k+1]-1 of x on table k; arrays updated in place."""
        ncol = 9 if spin else 6
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.ndim == 2 and x.shape[1] == ncol
        offs = np.ascontiguousarray(offsets, dtype=np.int32)
        assert len(offs) == len(self.tables) + 1 and offs[-1] == x.shape[0]
        n = x.shape[0]
        stable = np.ones(n, np.int32) if stable is None else np.ascontiguousarray(stable, np.int32)
        lost_turn = np.zeros(n, np.int32) if lost_turn is None else np.ascontiguousarray(lost_turn, np.int32)
        lost_post = np.zeros(n, np.int32) if lost_post is None else np.ascontiguousarray(lost_post, np.int32)
        check(self.lib.st_batch_track(self.h, _ptr(x), _ptr(offs), int(nturn), _ptr(stable), _ptr(lost_turn),
                                      _ptr(lost_post), TRACK_SPIN if spin else 0,
                                      ctypes.byref(cfg) if cfg is not None else None), "st_batch_track")
        return x, stable, lost_turn, lost_post

    def track_device(self, d_x, offsets, turn0, nturn, d_state, d_lost, flags=0, cfg=None):
        offs = np.ascontiguousarray(offsets, dtype=np.int32)
        check(self.lib.st_batch_track_device(self.h, ctypes.c_void_p(d_x), _ptr(offs), int(turn0), int(nturn),
                                             ctypes.c_void_p(d_state), ctypes.c_void_p(d_lost), int(flags),
                                             ctypes.byref(cfg) if cfg is not None else None),
              "st_batch_track_device")

    def track_tbt_device(self, d_x, offsets, turn0, nturn, d_tbt, d_state, d_lost, flags=0, cfg=None):
        offs = np.ascontiguousarray(offsets, dtype=np.int32)
        check(self.lib.st_batch_track_tbt_device(self.h, ctypes.c_void_p(d_x), _ptr(offs), int(turn0), int(nturn),
                                                 ctypes.c_void_p(d_tbt), ctypes.c_void_p(d_state),
                                                 ctypes.c_void_p(d_lost), int(flags),
                                                 ctypes.byref(cfg) if cfg is not None else None),
              "st_batch_track_tbt_device")

    def autotune(self, x, offsets, calib_turns, spin=False):
        """(best BatchConfig, [(config, seconds), ...]) over the candidate shapes."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        offs = np.ascontiguousarray(offsets, dtype=np.int32)
        best = BatchConfig()
        secs = np.zeros(32)
        n = ctypes.c_int()
        check(self.lib.st_batch_autotune(self.h, _ptr(x), _ptr(offs), int(calib_turns), TRACK_SPIN if spin else 0,
                                         ctypes.byref(best), _ptr(secs), ctypes.byref(n)), "st_batch_autotune")
        cands = [(w, o, t) for w in (1, 2, 4) for o in (1, 2) for t in (16, 64, 256)]
        return best, list(zip(cands, secs[:n.value]))


class Mlp:
    """The survival surrogate network (st_mlp_*): n_in -> H -> H -> H -> 1.
    Device arguments are integer addresses (e.g. torch tensor data_ptr())."""

    def __init__(self, ctx, n_in, n_hidden=256, seed=1):
        self.ctx = ctx
        self.lib = ctx.lib
        self.n_in, self.n_hidden = n_in, n_hidden
        self.h = ctypes.c_void_p()
        check(self.lib.st_mlp_create(ctx.h, n_in, n_hidden, seed, ctypes.byref(self.h)), "st_mlp_create")
        n = ctypes.c_int64()
        check(self.lib.st_mlp_num_params(self.h, ctypes.byref(n)), "st_mlp_num_params")
        self.nparam = n.value

    def close(self):
        if self.h:
            self.lib.st_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.nparam, np.float32)
        check(self.lib.st_mlp_get_params(self.h, _ptr(p), self.nparam), "st_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparam
        check(self.lib.st_mlp_set_params(self.h, _ptr(p), self.nparam), "st_mlp_set_params")

    def unpack(self, p):
        """[(W, b)] per layer from a flat parameter vector."""
        H, n = self.n_hidden, self.n_in
        shapes = [(H, n), (H, H), (H, H), (1, H)]
        out, o = [], 0
        Ws = []
        for sh in shapes:
            Ws.append(p[o:o + sh[0] * sh[1]].reshape(sh))
            o += sh[0] * sh[1]
        for sh, W in zip(shapes, Ws):
            out.append((W, p[o:o + sh[0]]))
            o += sh[0]
        return out

    def forward(self, d_x, n, d_prob):
        check(self.lib.st_mlp_forward(self.h, ctypes.c_void_p(d_x), int(n), ctypes.c_void_p(d_prob)), "st_mlp_forward")

    def gradients(self, d_x, d_y, n):
        g = np.zeros(self.nparam, np.float32)
        loss = ctypes.c_float()
        check(self.lib.st_mlp_gradients(self.h, ctypes.c_void_p(d_x), ctypes.c_void_p(d_y), int(n), _ptr(g),
                                        ctypes.byref(loss)), "st_mlp_gradients")
        return g, loss.value

    def train_step(self, d_x, d_y, n, lr):
        loss = ctypes.c_float()
        check(self.lib.st_mlp_train_step(self.h, ctypes.c_void_p(d_x), ctypes.c_void_p(d_y), int(n), float(lr),
                                         ctypes.byref(loss)), "st_mlp_train_step")
        return loss.value
