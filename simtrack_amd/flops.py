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
import numpy as np

from ._capi import TYPES

PEAK_FP64_VECTOR = 78.6e12

PT_UPDATE = 15        # delta1 (7), gamma1 (2), beta1 (4), 1+delta1, (1+delta1)^2
DRIFT = 16            # DRIFT_Pass exact, cpp:3753-3765 with delta1/gamma1/beta1 cached
DRIFT_HEXP = 11       # DRIFT_Pass H_expand branch, cpp:3745-3750
QUAD_KICK, SEXT_KICK, OCT_KICK = 8, 15, 22
SBEND_EXACT = 51      # sbend_exact_pass, cpp:3833-3858 (2 asin counted as 1 each)
GSBEND_EXACT = 56
EDGE = 4
SPIN = 114            # spin(), cpp:4574-4630 (sin and cos counted)
RADIATE = 44 + PT_UPDATE
RADIATE1 = 60 + PT_UPDATE
BENDKICK = 7
GBENDKICK = 9
BB4D_ROUND = 40       # BB4D round branch, cpp:6336-6342
BB4D_ELLIPTIC = 40 + 2 * 150   # two cerrf calls, cpp:6343-6380
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


def element_ops(r, gp, spin=False):
    t = int(r["type"])
    L = float(r["l"])
    nint = max(int(r["nint"]), 1)
    hexp = bool(gp.H_expand)
    rad = bool(gp.radiate)
    d = DRIFT_HEXP if hexp else DRIFT
    drift_l = (lambda length: 0 if length == 0. else d)
    name = {v: k for k, v in TYPES.items()}[t]
    mis = 0
    if abs(r["dx"]) + abs(r["dy"]) + abs(r["ds"]) + abs(r["dtheta"]) + abs(r["dphi"]) + abs(r["dpsi"]) >= 1e-9:
        if abs(r["dtheta"]) + abs(r["dphi"]) < 1e-9:
            mis = 2 * (4 + 12)
        else:
            mis = 2 * (72 + 5)
    if name in ("DRIFT", "MARKER", "BPM", "HBPM", "VBPM"):
        return drift_l(L)
    if name in ("ROTAT", "SNAKE"):
        return 2 * drift_l(L / 2) + (18 if spin else 0)
    if name in ("QUAD", "SKEWQ", "SEXT", "OCT"):
        order = {"QUAD": 1, "SKEWQ": 1, "SEXT": 2, "OCT": 3}[name]
        kick = {1: QUAD_KICK, 2: SEXT_KICK, 3: OCT_KICK}[order]
        kn = 0. if name == "SKEWQ" else r["knl"][order]
        if kn == 0. and r["knsl"][order] == 0.:
            return drift_l(L) + mis
        if L == 0.:
            return kick + (SPIN + 4 + kick // 2 + 2 if spin else 0) + mis
        per = 4 * d + 3 * kick + (SPIN + 4 + kick + 2 if spin else 0)
        return nint * per + mis
    if name == "MULT" or (name in ("GMULT", "SBENDMULT", "SMULT", "GSBENDMULT") and L == 0.):
        n = int(r["norder"])
        m0 = _mask(r["knl"], r["knsl"], n)
        if L == 0.:
            return mult_kick(n, m0) + (SPIN + mult_kick(n, m0) if spin else 0) + mis
        k1 = _mask(r["knl"] * 1.0, r["knsl"] * 1.0, n)
        per = 4 * d + 3 * mult_kick(n, k1) + (SPIN + 4 + mult_kick(n, m0) + 3 if spin else 0)
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
        extra = (SPIN if spin else 0) + ((RADIATE if spin else RADIATE1) if rad else 0)
        return 2 * EDGE + nint * (step + extra) + mis
    if name in ("GMULT", "SBENDMULT", "SMULT", "GSBENDMULT"):
        n = int(r["norder"])
        m = _mask(r["knl"], r["knsl"], n)
        gen = name == "GSBENDMULT" and not spin
        if hexp:
            step = 4 * d + 3 * mult_kick(n, m, bend=True)
        else:
            step = 4 * (GSBEND_EXACT if gen else SBEND_EXACT) + 3 * mult_kick(n, m)
        extra = 0
        if spin or rad:
            extra += 4 + mult_kick(n, m) + 3
        if spin:
            extra += SPIN
        if rad:
            extra += RADIATE
        return 2 * EDGE + nint * (step + extra) + mis
    if name in ("KICKER", "HKICKER", "VKICKER"):
        sp = SPIN if spin and (r["hkick"] != 0. or r["vkick"] != 0.) else 0
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
        return 2 * drift_l(L / 2) + 140 + PT_UPDATE + mis
    if name in ("LBT", "ILBT"):
        return (30 if name == "LBT" else 28) + PT_UPDATE + (0 if spin else mis)
    if name == "BEAMBEAM":
        if r["np"] == 0.:
            return mis
        if int(r["treatment"]) == 6:
            sx, sy = r["emitx"] * r["betax"], r["emity"] * r["betay"]
            per = 97 if abs(np.sqrt(sx) - np.sqrt(sy)) / np.sqrt(sx) < 1e-6 else 97 + 2 * 150 + 20
            return int(r["nslice"]) * per + PT_UPDATE + mis
        return 40 + 2 * 150 + mis
    if name == "CRBB":
        return 60 + int(r["nslice"]) * (97 + 2 * 150 + 20) + PT_UPDATE + mis
    if name == "MATRIX":
        return 6 + 72 + 6 + PT_UPDATE
    if name == "PATCH":
        # TRANS_Pass 3; YROTAT_Pass: pz 5 (sqrt), temp 3, x 2, px 3, y 5, z 7
        return 3 + (0 if abs(r["theta"]) < 1e-9 else 25)
    if name == "SOLEN":
        return drift_l(L) + mis
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
