"""Strong-strong beam-beam between two macroparticle beams on the GPU, with
the reference's function names: soft-Gaussian (simtrack.cpp:22846-23284)
and particle-in-cell (cpp:23286-23615).

Beams are (Mp, 6) arrays of rows (x, px, y, py, z, pt).  numpy arrays are
copied to the device and back (the reference's in-place update is kept:
the arrays passed in are overwritten); torch tensors on the GPU are updated
in place without copies.  Moments are summed in a fixed order, so results
repeat bit for bit and agree with the reference's sequential sums to
rounding.
"""
import ctypes

import numpy as np

from ._capi import STError, check, load


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _is_device(x):
    return hasattr(x, "data_ptr") and getattr(getattr(x, "device", None), "type", "cpu") == "cuda"


def _rows(x, what):
    if _is_device(x):
        import torch
        if x.dtype != torch.float64 or x.dim() != 2 or x.shape[1] != 6 or not x.is_contiguous():
            raise STError(f"{what}: a device beam must be a contiguous (Mp, 6) float64 tensor")
        return x, ctypes.c_void_p(x.data_ptr()), x.shape[0]
    if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.ndim == 2 and x.shape[1] == 6
            and x.flags.c_contiguous and x.flags.writeable):
        raise STError(f"{what}: a host beam must be a writable C-contiguous (Mp, 6) float64 array")
    return x, _ptr(x), x.shape[0]


def _pair(x1, x2, what):
    a, pa, n1 = _rows(x1, what)
    b, pb, n2 = _rows(x2, what)
    if _is_device(a) != _is_device(b):
        raise STError(f"{what}: both beams must be on the host or both on the device")
    if _is_device(a):
        # the library works on its own stream: let torch's pending writes of
        # the beams finish first
        import torch
        torch.cuda.current_stream(a.device).synchronize()
    return pa, n1, pb, n2, _is_device(a)


def BeamBeam4D_SoftGaussian(ctx, Np1, x1, Np2, x2, gamma1, gamma2, bbscale1, bbscale2):
    """BeamBeam4D_SoftGaussian (simtrack.cpp:22846): kicks x1 and x2 in place."""
    lib = load()
    pa, n1, pb, n2, dev = _pair(x1, x2, "BeamBeam4D_SoftGaussian")
    f = lib.st_bb_soft_gaussian_4d if dev else lib.st_bb_soft_gaussian_4d_host
    check(f(ctx.h, Np1, n1, pa, Np2, n2, pb, gamma1, gamma2, bbscale1, bbscale2), "BeamBeam4D_SoftGaussian")


def BeamBeam6D_SoftGaussian(ctx, Np1, x1, Np2, x2, gamma1, gamma2, bbscale1, bbscale2, Nslice1, Nslice2,
                            zb1=None, zb2=None):
    """BeamBeam6D_SoftGaussian: without zb (simtrack.cpp:22874) both beams are
    sorted by z in place and cut into equal-population slices, and the slice
    centres (Z1, Z2) are returned; with zb1/zb2 (cpp:23134) the slices are
    the z windows zb[i+1] <= z < zb[i] and None is returned."""
    lib = load()
    pa, n1, pb, n2, dev = _pair(x1, x2, "BeamBeam6D_SoftGaussian")
    if zb1 is None and zb2 is None:
        z1, z2 = np.zeros(int(Nslice1)), np.zeros(int(Nslice2))
        f = lib.st_bb_soft_gaussian_6d if dev else lib.st_bb_soft_gaussian_6d_host
        check(f(ctx.h, Np1, n1, pa, Np2, n2, pb, gamma1, gamma2, bbscale1, bbscale2, int(Nslice1), int(Nslice2),
                _ptr(z1), _ptr(z2)), "BeamBeam6D_SoftGaussian")
        return z1, z2
    zb1 = np.ascontiguousarray(zb1, dtype=np.float64)
    zb2 = np.ascontiguousarray(zb2, dtype=np.float64)
    if len(zb1) != Nslice1 + 1 or len(zb2) != Nslice2 + 1:
        raise STError("BeamBeam6D_SoftGaussian: zb must hold Nslice + 1 boundaries")
    f = lib.st_bb_soft_gaussian_6d_zb if dev else lib.st_bb_soft_gaussian_6d_zb_host
    check(f(ctx.h, Np1, n1, pa, Np2, n2, pb, gamma1, gamma2, bbscale1, bbscale2, int(Nslice1), int(Nslice2),
            _ptr(zb1), _ptr(zb2)), "BeamBeam6D_SoftGaussian")
    return None


def BeamBeam4D_PIC(ctx, Np1, x1, Np2, x2, gamma1, gamma2, bbscale1, bbscale2):
    """BeamBeam4D_PIC (simtrack.cpp:23286): particle-in-cell kicks, in place."""
    lib = load()
    pa, n1, pb, n2, dev = _pair(x1, x2, "BeamBeam4D_PIC")
    f = lib.st_bb_pic_4d if dev else lib.st_bb_pic_4d_host
    check(f(ctx.h, Np1, n1, pa, Np2, n2, pb, gamma1, gamma2, bbscale1, bbscale2), "BeamBeam4D_PIC")


def BeamBeam4D_PIC_NonOffset(ctx, Np1, x1, Np2, x2, gamma1, gamma2, bbscale1, bbscale2, Nx, Ny, hx, hy):
    """BeamBeam4D_PIC_NonOffset (simtrack.cpp:24226): particle-in-cell kicks
    on the given Nx x Ny grid of hx x hy cells centred on the origin, in
    place.  The reference takes the grid's Green function from its caller;
    it is computed here from hx, hy.  Nx = Ny = 64."""
    lib = load()
    pa, n1, pb, n2, dev = _pair(x1, x2, "BeamBeam4D_PIC_NonOffset")
    f = lib.st_bb_pic_4d_nonoffset if dev else lib.st_bb_pic_4d_nonoffset_host
    check(f(ctx.h, Np1, n1, pa, Np2, n2, pb, gamma1, gamma2, bbscale1, bbscale2, int(Nx), int(Ny), float(hx),
            float(hy)), "BeamBeam4D_PIC_NonOffset")


def BeamBeam6D_PIC(ctx, Np1, x1, Np2, x2, gamma1, gamma2, bbscale1, bbscale2, Nslice1, Nslice2):
    """BeamBeam6D_PIC (simtrack.cpp:23380): beams sorted by z in place, slice
    pairs collided by BeamBeam4D_PIC; returns the slice centres (Z1, Z2)."""
    lib = load()
    pa, n1, pb, n2, dev = _pair(x1, x2, "BeamBeam6D_PIC")
    z1, z2 = np.zeros(int(Nslice1)), np.zeros(int(Nslice2))
    f = lib.st_bb_pic_6d if dev else lib.st_bb_pic_6d_host
    check(f(ctx.h, Np1, n1, pa, Np2, n2, pb, gamma1, gamma2, bbscale1, bbscale2, int(Nslice1), int(Nslice2),
            _ptr(z1), _ptr(z2)), "BeamBeam6D_PIC")
    return z1, z2


def Zboundary_slices_Gaussian(Nslice, sigmal):
    """Zboundary_slices_Gaussian (simtrack.cpp:3157): Nslice + 1 boundaries, decreasing."""
    zb = np.zeros(int(Nslice) + 1)
    check(load().st_zboundary_slices_gaussian(_ptr(zb), int(Nslice), float(sigmal)), "Zboundary_slices_Gaussian")
    return zb


__all__ = ["BeamBeam4D_SoftGaussian", "BeamBeam6D_SoftGaussian", "BeamBeam4D_PIC", "BeamBeam6D_PIC",
           "Zboundary_slices_Gaussian"]
