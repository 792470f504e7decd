"""Algorithmic FP64 operation count per particle-turn (F_turn).

Counts the per-particle arithmetic of each element map: every IEEE add,
subtract, multiply, divide and square root is one operation and every
elementary-function call (sin, cos, asin, exp, pow) is one.  Not counted:
operations whose operands are all element or reference-particle constants
(evaluated once when the table is built), the pt-derived delta1/gamma1/beta1
except after a map that changes pt (the kernels cache them), zero-length
drifts, and comparisons.  The counts follow the operation sequences of the
maps (simtrack.cpp:3743-8348) as the kernels execute them.

roofline.achieved = F_turn * particle-turns/s; the bound is the FP64 vector
peak of MI355X (78.6 TFLOP/s, FMA = 2).
"""
import math

import numpy as np

from ._capi import TYPES

PEAK_FP64_VECTOR = 78.6e12

PT_UPDATE = 15        # delta1 (7), gamma1 (2), beta1 (4), 1+delta1, (1+delta1)^2
DRIFT = 16            # DRIFT_Pass exact, cpp:3753-3765 with delta1/gamma1/beta1 cached
DRIFT_HEXP = 14       # DRIFT_Pass H_expand branch, cpp:3745-3750 (1 + pt formed twice)
QUAD_KICK, SEXT_KICK, OCT_KICK = 8, 15, 22
SBEND_EXACT = 56      # sbend_exact_pass, cpp:3833-3858 (2 asin counted as 1 each)
GSBEND_EXACT = 57
EDGE = 4
SPIN = 112            # spin(), cpp:4574-4630 (sin and cos counted)
SPIN_UNIFORM = 110    # spin() with a field given by the element (bends, kickers)
# spin field of QUAD/SEXT/OCT_sPass at the slice midpoint, cpp:7538-7660:
# zero-length elements, and thick slices (which also divide by Nint)
SPIN_FIELD = {1: 6, 2: 13, 3: 22}
SPIN_FIELD_SLICE = {1: 8, 2: 15, 3: 22}
RADIATE = 48 + PT_UPDATE
RADIATE1 = 63 + PT_UPDATE
BENDKICK = 8
GBENDKICK = 9
QUAD_FRINGE = 72      # one end of QUAD_Pass's fringe map, cpp:4895-4911
BB4D_ROUND = 17       # BB4D round branch, cpp:6336-6342
# cerrf (cpp:6258-6323) is data-dependent: ~100 operations far from the core,
# up to ~600 inside it (Taylor/continued-fraction lengths 7 + 23 q, 10 + 21 q);
# 550 is its mean over Gaussian beam-core arguments (instrumented)
CERRF = 550
BB4D_ELLIPTIC = 24 + 2 * CERRF   # two cerrf calls, cpp:6343-6380
BB6D_SLICE_ROUND = 78            # one slice of BB6D (cpp:6622-6676) with the round BB4D
BB6D_SLICE_ELLIPTIC = 113 + 2 * CERRF
CRBB_BOOSTS = 60                 # BB6D_Angle's boost and inverse boost, cpp:6958-7078
# strong-strong slice pair, per row (cpp:23035-23110): moments of the drifted
# row (drift 4, sums 6), then drift 4, BB4D, kick 4, drift back 4
SS_MOMENTS = 10
SS_KICK_ELLIPTIC = 12 + BB4D_ELLIPTIC
SS_BYTES = 96         # x, px, y, py read by both passes and written by the kick
# particle-in-cell pair, per row: x, y read for the moments (16) and by the
# deposition (16), x, y, px, py read and px, py written by the kick (48),
# 9 contributions of (cell, row, weight) written, sorted and read (~9 x 48)
PIC_BYTES = 16 + 16 + 48 + 9 * 48


def mult_kick(norder, mask, bend=False):
    n = 0
    for i in range(1, norder + 1):
        n += 6
        if (mask >> i) & 1:
            n += 10
    return n + 2 + (8 if bend else 0)


def _mask(kn, ks, norder):
    m = 0
    for i in range(1, norder + 1):
        if kn[i] != 0. or ks[i] != 0.:
            m |= 1 << i
    return m


def _misalignment(r):
    """GtoL + LtoG (simtrack.cpp:3701-3741): shifts (2 each way) and the
    tilt rotation (12 each way) of GtoL1/LtoG1, or the pitched frame change
    of GtoL2/LtoG2 (W-matrix product and path-length terms); none of them
    changes pt."""
    a = [abs(float(r[k])) for k in ("dx", "dy", "ds", "dtheta", "dphi", "dpsi")]
    if sum(a) < 1e-9:
        return 0
    if a[3] + a[4] < 1e-9:
        return 2 * ((2 if a[0] + a[1] > 1e-10 else 0) + (12 if a[5] > 1e-10 else 0))
    return 2 * (72 + 5)


def _acdip_kick(r, turn):
    """ACDIP_Pass's kick amplitudes at a turn (simtrack.cpp:5930-5950):
    None before TURNS (a plain drift), else whether any kick is non-zero."""
    ts, te = float(r["turns"]), float(r["turne"])
    if turn < ts:
        return None
    amp = (turn - ts) / (te - ts) if ts <= turn <= te and te != ts else 1.
    hk = float(r["hkick"]) if int(r["type"]) == TYPES["HACDIP"] else 0.
    vk = float(r["vkick"]) if int(r["type"]) == TYPES["VACDIP"] else 0.
    s = math.sin(2.0 * 3.14159265 * float(r["nud"]) * (turn - ts) + float(r["phid"]))
    return (hk * amp * s != 0.) or (vk * amp * s != 0.)


def element_ops(r, gp, spin=False, turn=0):
    """Counted operations of one element map; turn is the reference's
    GP.turn, which only the AC elements read."""
    t = int(r["type"])
    L = float(r["l"])
    nint = max(int(r["nint"]), 1)
    hexp = bool(gp.H_expand)
    rad = bool(gp.radiate)
    d = DRIFT_HEXP if hexp else DRIFT
    drift_l = (lambda length: 0 if length == 0. else d)
    name = {v: k for k, v in TYPES.items()}[t]
    mis = _misalignment(r)
    if name in ("DRIFT", "MARKER", "BPM", "HBPM", "VBPM"):
        return drift_l(L)
    if name in ("ROTAT", "SNAKE"):
        # Pass is one drift; sPass two half drifts around the rotation
        return 2 * drift_l(L / 2) + 18 if spin else drift_l(L)
    if name in ("QUAD", "SKEWQ", "SEXT", "OCT"):
        order = {"QUAD": 1, "SKEWQ": 1, "SEXT": 2, "OCT": 3}[name]
        kick = {1: QUAD_KICK, 2: SEXT_KICK, 3: OCT_KICK}[order]
        kn = 0. if name == "SKEWQ" else r["knl"][order]
        if kn == 0. and r["knsl"][order] == 0.:
            return drift_l(L) + mis
        if L == 0.:
            return kick + (SPIN + 4 + SPIN_FIELD[order] if spin else 0) + mis
        per = 4 * d + 3 * kick + (SPIN + 4 + SPIN_FIELD_SLICE[order] if spin else 0)
        # QUAD_Pass's hard-edge fringe maps at both ends (cpp:4895-4957)
        fringe = 2 * QUAD_FRINGE if gp.quad_fringe and order == 1 and not spin else 0
        return nint * per + fringe + mis
    if name == "MULT" or (name in ("GMULT", "SBENDMULT", "SMULT", "GSBENDMULT") and L == 0.):
        n = int(r["norder"])
        m0 = _mask(r["knl"], r["knsl"], n)
        if L == 0.:
            # the bend families' zero-length sPass forms the midpoint of x and y too
            extra = 1 if name == "MULT" else 5
            return mult_kick(n, m0) + (SPIN + mult_kick(n, m0) + extra if spin else 0) + mis
        k1 = _mask(r["knl"] * 1.0, r["knsl"] * 1.0, n)
        per = 4 * d + 3 * mult_kick(n, k1) + (SPIN + 4 + mult_kick(n, m0) + 1 if spin else 0)
        return nint * per + mis
    if name in ("SBEND", "GSBEND"):
        if r["angle"] == 0.:
            return drift_l(L) + mis
        if name == "GSBEND" and not spin:
            body = GSBEND_EXACT if not hexp else nint * (4 * d + 3 * GBENDKICK)
            return 2 * EDGE + body + mis
        if not spin and not hexp and not rad:
            return 2 * EDGE + SBEND_EXACT + mis
        step = SBEND_EXACT if not hexp else 4 * d + 3 * BENDKICK
        extra = (SPIN_UNIFORM if spin else 0) + ((RADIATE if spin else RADIATE1) if rad else 0)
        return 2 * EDGE + nint * (step + extra) + mis
    if name in ("GMULT", "SBENDMULT", "SMULT", "GSBENDMULT"):
        n = int(r["norder"])
        m = _mask(r["knl"], r["knsl"], n)
        gen = name == "GSBENDMULT" and not spin
        if hexp:
            # general_bend_mult_kick_pass (GSBENDMULT) has the K0L term besides the geometric one
            step = 4 * d + 3 * (mult_kick(n, m, bend=True) + (1 if gen else 0))
        else:
            step = 4 * (GSBEND_EXACT if gen else SBEND_EXACT) + 3 * mult_kick(n, m)
        extra = 0
        if spin or rad:
            extra += 4 + mult_kick(n, m) + (2 if rad else 1)
        if spin:
            extra += SPIN
        if rad:
            extra += RADIATE
        return 2 * EDGE + nint * (step + extra) + mis
    if name in ("KICKER", "HKICKER", "VKICKER"):
        sp = SPIN_UNIFORM if spin and (r["hkick"] != 0. or r["vkick"] != 0.) else 0
        return 2 * drift_l(L / 2) + 2 + sp + mis
    if name == "RFCAV":
        if r["vrf"] == 0.:
            return drift_l(L) + mis
        return 2 * drift_l(L / 2) + 9 + PT_UPDATE + mis
    if name == "CRABRF":
        if r["vrf"] == 0. or spin:
            return drift_l(L) + mis
        return 2 * drift_l(L / 2) + 18 + PT_UPDATE + mis
    if name == "CCMULT":
        if r["vrf"] == 0.:
            return drift_l(L) + mis
        return 2 * drift_l(L / 2) + 147 + PT_UPDATE + mis
    if name in ("LBT", "ILBT"):
        return (41 if name == "LBT" else 34) + PT_UPDATE + (0 if spin else mis)
    if name == "BEAMBEAM":
        if r["np"] == 0.:
            return mis
        if int(r["treatment"]) == 6:
            sx, sy = r["emitx"] * r["betax"], r["emity"] * r["betay"]
            round_ = abs(np.sqrt(sx) - np.sqrt(sy)) / np.sqrt(sx) < 1e-6
            per = BB6D_SLICE_ROUND if round_ else BB6D_SLICE_ELLIPTIC
            return int(r["nslice"]) * per + PT_UPDATE + mis
        sx, sy = r["emitx"] * r["betax"], r["emity"] * r["betay"]
        return (BB4D_ROUND if abs(np.sqrt(sx) - np.sqrt(sy)) / np.sqrt(sx) < 1e-6 else BB4D_ELLIPTIC) + mis
    if name == "CRBB":
        return CRBB_BOOSTS + int(r["nslice"]) * BB6D_SLICE_ELLIPTIC + PT_UPDATE + mis
    if name == "MATRIX":
        m = np.asarray(r["m66"], dtype=float).reshape(6, 6)
        keeps_pt = (np.array_equal(m[5], [0., 0., 0., 0., 0., 1.]) and r["xco_in"][5] == 0. and
                    r["xco_out"][5] == 0.)
        return 6 + 72 + 6 + (0 if keeps_pt else PT_UPDATE)
    if name == "PATCH":
        # TRANS_Pass 3; YROTAT_Pass: pz 5 (sqrt), temp 3, x 2, px 3, y 5, z 7
        return 3 + (0 if abs(r["theta"]) < 1e-9 else 24)
    if name == "WIGGLER":
        # per step: Forest's implicit step 85 (Pass) or Wu's six field
        # evaluations 116 (112 when KX = 0: x itself replaces sin(KX x)) (sPass);
        # with radiation the step's field and radiate()
        if spin:
            per = (116 if r["kx"] != 0. else 112) + (71 + PT_UPDATE if rad else 0)
        else:
            per = 85 + (60 + PT_UPDATE if rad else 0)
        return nint * per + mis
    if name == "SOLEN":
        if r["ks"] == 0.:
            return drift_l(L) + mis
        # x', y' and vz 46; per body step the 4x4 product 32, and in the spin
        # pass the step's px, py 8 and spin() in a constant longitudinal field 108
        return 46 + nint * (32 + (116 if spin else 0)) + mis
    # frame changes and cooling: no GtoL/LtoG
    if name == "TRANS":
        return 3
    if name == "SROTAT":
        return 0 if r["psi"] == 0. else 12
    if name in ("YROTAT", "XROTAT"):
        # pz 5 (sqrt), temp 3, the rotated pair 5, the other coordinate 5, z 7
        a = r["theta"] if name == "YROTAT" else r["phi"]
        return 0 if abs(a) < 1e-9 else 24
    if name == "COOLING":
        return 4
    if name == "LRBB":
        # the separation added and removed around the 4-D kick
        sx, sy = float(r["sigmax"]), float(r["sigmay"])
        return (BB4D_ROUND if abs(sx - sy) / sx < 1e-6 else BB4D_ELLIPTIC) + 4 + mis
    if name == "ELENS":
        if r["ne"] == 0.:
            return drift_l(L) + mis
        sx, sy = float(r["sigmax"]), float(r["sigmay"])
        bb = BB4D_ROUND if abs(sx - sy) / sx < 1e-6 else BB4D_ELLIPTIC
        # each kick scaled twice, (1 + beta_e) then BBSCALE
        return int(r["nslice"]) * (2 * drift_l(L / int(r["nslice"]) / 2) + bb + 2) + mis
    if name == "HELENS":
        # r 4, scale inside the hollow band 5, two kicks of 4 each
        return int(r["nslice"]) * (2 * drift_l(L / int(r["nslice"]) / 2) + 17) + mis
    if name in ("HACDIP", "VACDIP"):
        k = _acdip_kick(r, turn)
        if k is None:
            return drift_l(L) + mis
        return 2 * drift_l(L / 2) + 2 + (SPIN_UNIFORM if spin and k else 0) + mis
    if name in ("HACMULT", "VACMULT"):
        n = int(r["norder"])
        if n == 0:
            # a constant field: two kicks, the spin rotation with constant B
            sp = SPIN_UNIFORM if spin and r["klmax"] != 0. else 0
            return 2 * drift_l(L / 2) + 2 + sp + mis
        # (x + iy)^n 6 per order, By and Bx 5 each, two kicks 2 each; the
        # spin field's two products and spin()
        return 2 * drift_l(L / 2) + 6 * n + 10 + 4 + (2 + SPIN if spin else 0) + mis
    return 0


def flops_per_turn(records, gp, spin=False):
    """F_turn of a line: sum of element_ops over one turn."""
    return int(sum(element_ops(r, gp, spin) for r in records))


def breakdown(records, gp, spin=False):
    from collections import Counter
    c = Counter()
    name = {v: k for k, v in TYPES.items()}
    for r in records:
        c[name[int(r["type"])]] += element_ops(r, gp, spin)
    return dict(c)
