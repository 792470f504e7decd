"""Synthetic lattices and particle sets for tests and benchmarks.

Everything is generated from explicit seeds with numpy's PCG64, so the CPU
oracle and the GPU engine receive identical inputs.
"""
import math

import numpy as np

from . import line as L

PROTON_E0 = 938.27201


def proton_globals(total_energy_mev=100000.0, **flags):
    """Set_RefPartEnergy(E/E0, 938.27201, 1, 1): a proton of total energy E."""
    return L.make_globals(total_energy_mev / PROTON_E0, PROTON_E0, 1, 1, **flags)


def electron_globals(total_energy_mev=18000.0, **flags):
    e0 = 0.51099895
    return L.make_globals(total_energy_mev / e0, e0, -1, 1, **flags)


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def all_types_line(seed=7, misalign=True, with_bb=True, with_rf=True, with_gsbend=True):
    """One of every element type the engine tracks, moderate random strengths,
    thick and thin variants, offset/roll and pitched misalignments."""
    r = _rng(seed)
    u = lambda a, b: float(r.uniform(a, b))  # noqa: E731
    knl = [0.] * 11
    knsl = [0.] * 11
    for j in range(1, 9):
        knl[j] = u(-1, 1) * 10.0 ** (-1 - 0.5 * j) * math.factorial(j)
        knsl[j] = u(-1, 1) * 10.0 ** (-2 - 0.5 * j) * math.factorial(j)
    knl[4] = 0.  # exercise the zero-order mask
    knsl[4] = 0.
    els = [
        L.DRIFT("D1", 1.3), L.MARKER("M0", 0.), L.BPM("B0", 0.),
        L.QUAD("QF", 1.1, 0.05, 0.002), L.QUAD("QT", 0., 0.03, -0.001), L.SKEWQ("SQ", 0.4, 0.01),
        L.SEXT("SF", 0.5, 0.3, 0.02), L.SEXT("ST", 0., -0.2, 0.01),
        L.OCT("OF", 0.3, 2.0, 0.5), L.OCT("OT", 0., -1.0, 0.2),
        L.MULT("MU", 0.9, knl, knsl), L.MULT("MT", 0., knl, knsl),
        L.SBEND("BE", 3.0, 0.02, 0.003, -0.002), L.SBEND("BN", 2.0, -0.015, 0., 0.),
        L.SBENDMULT("SBM", 2.5, 0.018, 0.001, 0.001, knl, knsl),
        L.GMULT("GM", 1.5, -0.012, 0., 0., knl, knsl), L.SMULT("SM", 1.2, 0.01, 0., 0., knl, knsl),
        L.GSBENDMULT("GBM", 1.9, 0.014, 0.0139, 0., 0., knl, knsl),
        L.SBENDMULT("SBT", 0., 0.01, 0., 0., knl, knsl),
        L.KICKER("K", 0.4, 2e-5, -1e-5), L.HKICKER("HK", 0., 3e-5), L.VKICKER("VK", 0.2, -2e-5),
        L.ROTAT("ROT", 0.5, [0.3, 0.9, 0.1], 0.2), L.SNAKE("SNK", 0.1, [1.0, 0., 0.2], math.pi / 2),
        L.MATRIX("MAT", 0., np.zeros(6), np.zeros(6), np.eye(6) + 1e-3 * np.triu(np.ones((6, 6)), 1)),
    ]
    if with_gsbend:
        els.insert(17, L.GSBEND("GB", 2.2, 0.016, 0.0161, 0.001, 0.0))
    if with_rf:
        els += [L.RFCAV("RF", 0.6, 3.0, 28.15e6, 0.05), L.RFCAV("RF0", 0.3, 0., 28e6, 0.),
                L.CRABRF("CR", 0.4, 2.0, 197e6, 0.1),
                L.CCMULT("CCM", 0., 25.8, 197e6, 0.0, 2.8e-3, -9.8e-7, 2.2e-3, -3.9e-5, 0.59, 2.3e-3, 0.1, 0.01)]
    if with_bb:
        eps = 2.5e-9
        els += [L.LBT("LBT", 0.0125),
                L.BEAMBEAM("BB6", 6, 1.72e11, -1.0, 0.02, 5, 20e-9, 0.45, 0., 1.3e-9, 0.056, 0.),
                L.ILBT("ILBT", 0.0125),
                L.BEAMBEAM("BB6R", 6, 2e11, 1.0, 0.3, 5, eps, 0.5, 0., eps, 0.5, 0.),
                L.BEAMBEAM("BB6E", 6, 2e11, 1.0, 0.1, 11, 3 * eps, 0.6, 0.1, eps, 0.3, -0.05),
                L.BEAMBEAM("BB4", 4, 1.5e11, 1.0, 0.1, 1, 4 * eps, 0.5, 0., eps, 0.5, 0.),
                L.CRBB("CRB", 1.5e11, 0.01, 1.0, 0.05, 5, 2 * eps, 0.4, 0.02, eps, 0.2, 0.)]
    els += [L.HBPM("HB", 0.2), L.VBPM("VB", 0.)]
    els += [L.PATCH("PT", -0.012, 0.001, 3e-4, 0.021), L.PATCH("PT0", 0.003, 0., -1e-4, 0.),
            L.SOLEN("SOL", 1.1, 0.), L.SOLEN("SOLK", 1.2, 0.05)]
    els += [L.TRANS("TR", 2e-4, -1e-4, 0.05), L.SROTAT("SR", 0.013), L.YROTAT("YR", -0.007),
            L.XROTAT("XR", 0.009), L.COOLING("CO", 1e-4), L.SROTAT("SR0", 0.), L.XROTAT("XR0", 1e-10)]
    if with_bb:
        els += [L.LRBB("LR", 1.5e11, 1.0, 8e-4, -3e-4, 2e-4, 2e-4),
                L.ELENS("EL", 2.0, 1e11, 1.0, 3, 0.5, 3e-4, 3e-4), L.ELENS("EL0", 1.0, 0., 1.0, 2, 0.5, 3e-4, 3e-4),
                L.HELENS("HE", 2.0, 5e11, 1.0, 2, 0.5, 1e-2, 1e-6)]
    # AC elements: at turn 0 HAC is inside its ramp and VAC past it
    els += [L.HACDIP("HAC", 0.3, 2e-5, 0.31, 0.2, -3, 5), L.VACDIP("VAC", 0.0, 1e-5, 0.27, 0.0, -10, -2),
            L.HACMULT("HAM", 0.2, 2, 0.05, 100, 0.3), L.VACMULT("VAM", 0.0, 0, 1e-5, 50, 0.1),
            L.WIGGLER("WIG", 0.8, 8, 0.5, 2.0, 2 * math.pi / 0.2, 0.1),
            L.WIGGLER("WIP", 0.6, 6, 0.3, 0.0, 2 * math.pi / 0.15, 0.0)]
    for e in els:
        if hasattr(e, "Nint") and e.L > 0:
            e.Nint = max(1, min(e.Nint, 5))
    if misalign:
        for i, e in enumerate(els):
            if e.TYPE in ("DRIFT", "MARKER", "BPM", "HBPM", "VBPM", "MATRIX", "ROTAT", "SNAKE", "PATCH", "TRANS",
                          "SROTAT", "YROTAT", "XROTAT", "COOLING"):
                continue
            k = i % 4
            if k == 1:
                e.DX, e.DY = u(-1e-4, 1e-4), u(-1e-4, 1e-4)
            elif k == 2:
                e.DX, e.DPSI = u(-1e-4, 1e-4), u(-1e-3, 1e-3)
            elif k == 3:
                e.DX, e.DY, e.DS = u(-1e-4, 1e-4), u(-1e-4, 1e-4), u(-1e-3, 1e-3)
                e.DTHETA, e.DPHI, e.DPSI = u(-1e-4, 1e-4), u(-1e-4, 1e-4), u(-1e-3, 1e-3)
    return L.Line(els)


def random_particles(n, seed=1, scale=1e-3, spin=False, pt_scale=1e-4):
    r = _rng(seed)
    x = np.zeros((n, 9 if spin else 6))
    x[:, 0] = r.normal(0, scale, n)
    x[:, 1] = r.normal(0, scale * 0.05, n)
    x[:, 2] = r.normal(0, scale, n)
    x[:, 3] = r.normal(0, scale * 0.05, n)
    x[:, 4] = r.normal(0, 0.05, n)
    x[:, 5] = r.normal(0, pt_scale, n)
    if spin:
        v = r.normal(size=(n, 3))
        v[:, 1] += 3.0
        v /= np.linalg.norm(v, axis=1)[:, None]
        x[:, 6:9] = v
    return x


# ------------------------------------------------------- DA ring (config 0)
#
# 48 FODO arc cells of 76 m (two 32 m sector dipoles, thick quadrupoles and
# sextupoles as MULT with b3..b10 errors, a corrector, BPMs), two 91 m low-beta
# straights with a matched quadruplet on each side of the IP (beta* = 1 m),
# one 28 MHz RF cavity at h = 360.  999 elements, C = 3830 m.

ARC_KQ = 0.03355          # arc quadrupole K1L (QF +, QD -), ~90 deg cells
ARC_ANGLE = 2 * math.pi / 96
IR_LQ = (2.0, 3.0, 3.0, 2.0)
IR_GAP = (1.5, 1.5, 1.5, 1.0)
IR_DRIFT = 30.0
BETA_STAR = 1.0
EMIT_N = 2.5e-6           # normalised rms emittance [m]
R_ARC, R_IR = 0.04, 0.06  # multipole reference radii [m]


def _m_quad(L, k):
    def m(kk):
        if kk > 0:
            s = math.sqrt(kk)
            return np.array([[math.cos(s * L), math.sin(s * L) / s], [-s * math.sin(s * L), math.cos(s * L)]])
        s = math.sqrt(-kk)
        return np.array([[math.cosh(s * L), math.sinh(s * L) / s], [s * math.sinh(s * L), math.cosh(s * L)]])
    return m(k), m(-k)


def _m_drift(L):
    m = np.array([[1., L], [0., 1.]])
    return m, m


def _m_bend(L, th):
    h = th / L
    return np.array([[math.cos(th), math.sin(th) / h], [-h * math.sin(th), math.cos(th)]]), _m_drift(L)[1]


def _arc_cell_linear():
    d = _m_drift
    return [_m_quad(0.75, ARC_KQ / 1.5), d(0.75), d(0.5), d(0.75), _m_bend(32., ARC_ANGLE), d(0.75), d(0.5),
            d(0.75), _m_quad(1.5, -ARC_KQ / 1.5), d(0.75), d(0.5), d(0.75), _m_bend(32., ARC_ANGLE), d(0.75),
            d(0.5), d(0.75), _m_quad(0.75, ARC_KQ / 1.5)]


def _ir_half_linear(k):
    seq = [_m_drift(IR_DRIFT)]
    for i in range(4):
        seq += [_m_quad(IR_LQ[i], k[i] / IR_LQ[i]), _m_drift(IR_GAP[i])]
    return seq


def _product(seq):
    mx, my = np.eye(2), np.eye(2)
    for a, b in seq:
        mx, my = a @ mx, b @ my
    return mx, my


def _periodic(m):
    mu = math.acos((m[0, 0] + m[1, 1]) / 2)
    if m[0, 1] < 0:
        mu = 2 * math.pi - mu
    return m[0, 1] / math.sin(mu), (m[0, 0] - m[1, 1]) / (2 * math.sin(mu)), mu / (2 * math.pi)


def _transport(m, b, a):
    c, s, cp, sp = m[0, 0], m[0, 1], m[1, 0], m[1, 1]
    g = (1 + a * a) / b
    return c * c * b - 2 * c * s * a + s * s * g, -c * cp * b + (c * sp + s * cp) * a - s * sp * g


def match_ir():
    """Quadruplet K1L matching the IP (beta*, alpha = 0) onto the arc cell's
    periodic Twiss in both planes (linear thick-element model)."""
    from scipy.optimize import least_squares
    mx, my = _product(_arc_cell_linear())
    bx0, ax0, _ = _periodic(mx)
    by0, ay0, _ = _periodic(my)

    def res(k):
        mx, my = _product(_ir_half_linear(k))
        bx, ax = _transport(mx, BETA_STAR, 0.)
        by, ay = _transport(my, BETA_STAR, 0.)
        return [math.log(bx / bx0) if bx > 0 else 10., ax - ax0, math.log(by / by0) if by > 0 else 10., ay - ay0]

    sol = least_squares(res, [-0.21, 0.27, -0.05, -0.42], xtol=1e-15, ftol=1e-15, gtol=1e-15)
    return [float(v) for v in sol.x]


# result of match_ir() (checked by the test suite)
IR_K1L = (-0.21269946112305568, 0.27000198780176277, -0.049269261431258135, -0.4178277891509443)


def ring_linear_tunes(k=IR_K1L):
    h = _ir_half_linear(k)
    full = h + _arc_cell_linear() * 24 + h[::-1] + h + _arc_cell_linear() * 24 + h[::-1]
    mx, my = _product(full)
    return _periodic(mx), _periodic(my)


def _errors(r, main, order_main, R, orders):
    """b/a errors of relative size 1e-4 of the main field at radius R:
    KNL[j] R^j / j! = 1e-4 * (KNL[n] R^n / n!) * N(0,1)."""
    kn = [0.] * 11
    ks = [0.] * 11
    kn[order_main] = main
    ref = abs(main) * R ** order_main / math.factorial(order_main)
    for j in orders:
        kn[j] = 1e-4 * ref * float(r.normal()) * math.factorial(j) / R ** j
        ks[j] = 1e-4 * ref * float(r.normal()) * math.factorial(j) / R ** j
    return kn, ks


def da_ring(seed=2026, k2f=0.10, k2d=-0.18, vrf=3.0, ip_elements=None, nint=4, aperture=0.5):
    """The ~1000-element DA ring.  ip_elements: two element lists placed at
    IP1 (start of the line) and IP2; default: zero-length markers."""
    r = _rng(seed)

    def mult(name, L, main, order, R, orders):
        kn, ks = _errors(r, main, order, R, orders)
        m = L_.MULT(name, L, kn, ks)
        m.Nint = nint
        return m

    L_ = L
    quad_err = range(2, 10)
    sext_err = range(3, 10)

    def cell(i):
        return [
            mult(f"QFH{i}a", 0.75, ARC_KQ / 2, 1, R_ARC, quad_err), L.DRIFT("D", 0.75),
            mult(f"SF{i}", 0.5, k2f, 2, R_ARC, sext_err), L.DRIFT("D", 0.75), L.SBEND(f"B{i}a", 32., ARC_ANGLE),
            L.DRIFT("D", 0.75), L.KICKER(f"C{i}", 0.5, 0., 0.), L.DRIFT("D", 0.75), L.BPM(f"BPM{i}a", 0.),
            mult(f"QD{i}", 1.5, -ARC_KQ, 1, R_ARC, quad_err), L.DRIFT("D", 0.75),
            mult(f"SD{i}", 0.5, k2d, 2, R_ARC, sext_err), L.DRIFT("D", 0.75), L.SBEND(f"B{i}b", 32., ARC_ANGLE),
            L.DRIFT("D", 0.75), L.BPM(f"BPM{i}b", 0.), L.DRIFT("D", 0.5), L.MARKER(f"MK{i}", 0.),
            L.DRIFT("D", 0.75), mult(f"QFH{i}b", 0.75, ARC_KQ / 2, 1, R_ARC, quad_err)]

    def ir_out(tag):
        seq = [L.DRIFT("DIR", IR_DRIFT)]
        for q in range(4):
            seq += [mult(f"Q{q + 1}{tag}", IR_LQ[q], IR_K1L[q], 1, R_IR, quad_err), L.DRIFT("DQ", IR_GAP[q])]
        return seq

    def ir_in(tag):
        seq = []
        for q in reversed(range(4)):
            seq += [L.DRIFT("DQ", IR_GAP[q]), mult(f"Q{q + 1}{tag}", IR_LQ[q], IR_K1L[q], 1, R_IR, quad_err)]
        return seq + [L.DRIFT("DIR", IR_DRIFT)]

    ip1, ip2 = ip_elements or ([L.MARKER("IP1", 0.)], [L.MARKER("IP2", 0.)])
    cells = []
    cells += list(ip1) + ir_out("1R")
    for i in range(24):
        cells += cell(i)
    cells += ir_in("2L") + list(ip2) + ir_out("2R")
    for i in range(24, 48):
        cells += cell(i)
    cells += ir_in("1L")
    # one RF cavity at h = 360 in the first arc cell
    length = sum(e.L for e in cells)
    g_beta = None
    ln = L.Line(cells)
    ln.info = dict(length=length)
    for e in ln.Cell:
        e.APx = e.APy = aperture
    k = [i for i, e in enumerate(ln.Cell) if e.NAME == "MK0"][0]
    ln.Cell[k] = L.RFCAV("RF28", 0., vrf, 0., 0.)     # FRF set by set_rf_frequency
    ln.Cell[k].APx = ln.Cell[k].APy = aperture
    ln.Update()
    del g_beta
    return ln


def set_rf_frequency(ln, gp, harmonic=360):
    """FRF = h * beta c / C for every RFCAV (Ex1:43 sets FRF = frev * h)."""
    frev = 299792458. * gp.beta / ln.Length
    for e in ln.Cell:
        if e.TYPE == "RFCAV":
            e.FRF = harmonic * frev


def beam_sizes(gp, emit_n=EMIT_N, beta=BETA_STAR):
    emit = emit_n / (gp.beta * gp.gamma)
    return emit, math.sqrt(emit * beta)


def da_grid(gp, n_amp=205, angles_deg=(15., 30., 45., 60., 75.), amax_sigma=20., pt=3e-4):
    """DA scan grid (Track_DA1 angles, cpp:21455): x = (A cos t sx, 0, A sin t sy, 0, 0, pt),
    A_k = (k+1) * amax/n_amp sigma.  x[5] carries deltap0 unconverted, as the
    reference's DA drivers do (cpp:21415)."""
    _, sig = beam_sizes(gp)
    amps = (np.arange(n_amp) + 1) * amax_sigma / n_amp
    rows = []
    for t in angles_deg:
        th = math.radians(t)
        for a in amps:
            rows.append((a * math.cos(th) * sig, 0., a * math.sin(th) * sig, 0., 0., pt))
    return np.array(rows), amps, np.array(angles_deg)


def config0(seed=2026):
    """Config 0: the DA ring, 100 GeV protons, 5 x 205 grid over 0-20 sigma."""
    gp = proton_globals(100000.0)
    ln = da_ring(seed=seed)
    set_rf_frequency(ln, gp)
    x, amps, angles = da_grid(gp)
    return gp, ln, x


def bb_intensity(gp, xi=0.012, emit_n=EMIT_N):
    """Bunch intensity giving head-on beam-beam parameter xi for round beams:
    xi = N r_p / (4 pi gamma eps)."""
    emit, _ = beam_sizes(gp, emit_n)
    return 4 * math.pi * gp.gamma * emit * xi / 1.534698e-18


def config1(n_particles=65536, seed=2027, lattice_seed=2026, nslice=5, sigma_z=0.3, sigma_pt=2.4e-4):
    """Config 1: config-0 ring with 6-D weak-strong beam-beam at both IPs
    (round strong beam sigma = 0.15 mm, sigma_z = 0.3 m, 5 slices, xi = 0.012
    per IP), spin on; Gaussian particles truncated at 6 sigma."""
    gp = proton_globals(100000.0)
    emit, sig = beam_sizes(gp)
    npb = bb_intensity(gp)

    def bb(name):
        return [L.BEAMBEAM(name, 6, npb, 1.0, sigma_z, nslice, emit, BETA_STAR, 0., emit, BETA_STAR, 0.)]

    ln = da_ring(seed=lattice_seed, ip_elements=(bb("BB1"), bb("BB2")))
    set_rf_frequency(ln, gp)
    r = _rng(seed)

    def tg(n, s):
        v = r.normal(0., 1., n)
        bad = np.abs(v) > 6.
        while bad.any():
            v[bad] = r.normal(0., 1., bad.sum())
            bad = np.abs(v) > 6.
        return v * s

    x = np.zeros((n_particles, 9))
    x[:, 0] = tg(n_particles, sig)
    x[:, 1] = tg(n_particles, sig / BETA_STAR)
    x[:, 2] = tg(n_particles, sig)
    x[:, 3] = tg(n_particles, sig / BETA_STAR)
    x[:, 4] = tg(n_particles, sigma_z)
    x[:, 5] = tg(n_particles, sigma_pt)
    x[:, 7] = 1.0          # spin along the vertical guide field
    return gp, ln, x


# ------------------------------------------------ config 2: many seeds

def cc_voltage(theta, freq_cc, beta_star, beta_cc, energy_mev):
    """Cal_CC_Voltage, simtrack.cpp:13570-13577: crab voltage [MV] for half
    crossing angle theta with ideal 90-degree phase advance."""
    c = 2.99792458e8
    energy = energy_mev * 1e6
    return c * energy * (2 * theta) / math.sqrt(beta_star * beta_cc) / 4 / 3.14159265358979323 / freq_cc / 1e6


def arc_beta():
    """Periodic horizontal beta at the arc-cell boundary (linear model)."""
    mx, _ = _product(_arc_cell_linear())
    return _periodic(mx)[0]


def crab_ring(lattice_seed, n_elements, gp, theta=0.0125, f_cc=197e6, nslice=5, sigma_z=0.3, rng_seed=None):
    """One config-2 lattice: the DA ring (errors from lattice_seed) with a
    crossing angle at IP1 (LBT, 6-D beam-beam, ILBT, as Example2/DA_cal.cpp:
    777-805), crab cavities at 197 MHz on both sides of IP1 at the arc ends
    of its insertion, head-on 6-D beam-beam at IP2, and zero-length extra
    correctors (KICKER) and thin multipole-error slices (MULT, b3..b6)
    inserted at random positions until the line has n_elements."""
    emit, sig = beam_sizes(gp)
    npb = bb_intensity(gp)
    ip1 = [L.LBT("LBT1", theta),
           L.BEAMBEAM("BB1", 6, npb, 1.0, sigma_z, nslice, emit, BETA_STAR, 0., emit, BETA_STAR, 0.),
           L.ILBT("ILBT1", theta)]
    ip2 = [L.BEAMBEAM("BB2", 6, npb, 1.0, sigma_z, nslice, emit, BETA_STAR, 0., emit, BETA_STAR, 0.)]
    ln = da_ring(seed=lattice_seed, ip_elements=(ip1, ip2))
    set_rf_frequency(ln, gp)
    v = cc_voltage(theta, f_cc, BETA_STAR, arc_beta(), gp.energy)
    cells = list(ln.Cell)
    # the first arc cell after IP1's insertion, and the last before it
    i_out = [i for i, e in enumerate(cells) if e.NAME == "QFH0a"][0]
    i_in = [i for i, e in enumerate(cells) if e.NAME == "QFH47b"][0] + 1
    # the insertion turns the phase by ~pi between the two cavities, so equal
    # voltages close the crab bump (opposite ones double it and destroy the DA)
    cells.insert(i_in, L.CRABRF("CC1L", 0., v, f_cc, 0.))
    cells.insert(i_out, L.CRABRF("CC1R", 0., v, f_cc, 0.))
    r = _rng(lattice_seed if rng_seed is None else rng_seed)
    n_extra = max(0, n_elements - len(cells))
    lo = i_out + 1
    for k in range(n_extra):
        pos = int(r.integers(lo, len(cells) - 60))
        if k % 2 == 0:
            e = L.KICKER(f"XC{k}", 0., float(r.normal(0, 2e-7)), float(r.normal(0, 2e-7)))
        else:
            kn = [0.] * 11
            ks = [0.] * 11
            for j in range(2, 6):
                kn[j] = float(r.normal(0, 1e-4)) * math.factorial(j) / 0.04 ** (j - 1) * 1e-2
                ks[j] = float(r.normal(0, 1e-4)) * math.factorial(j) / 0.04 ** (j - 1) * 1e-2
            e = L.MULT(f"XM{k}", 0., kn, ks)
        cells.insert(pos, e)
    for e in cells:
        e.APx = e.APy = 0.5
    return L.Line(cells)


def config2(n_seeds=64, n_per_seed=4096, seed=2028, min_elements=1000, max_elements=1800, n_angles=8,
            amax_sigma=20., pt=3e-4):
    """Config 2: n_seeds lattices of uneven length (uniform in [min, max]
    elements), each with its own n_per_seed DA-grid particles (n_angles
    angles in (0, 90) degrees x amplitudes up to amax_sigma).  Returns
    (globals, [Line], x with all seeds' particles stacked, offsets)."""
    gp = proton_globals(100000.0)
    r = _rng(seed)
    sizes = r.integers(min_elements, max_elements + 1, n_seeds)
    lines = [crab_ring(2026 + s, int(sizes[s]), gp) for s in range(n_seeds)]
    _, sig = beam_sizes(gp)
    n_amp = n_per_seed // n_angles
    amps = (np.arange(n_amp) + 1) * amax_sigma / n_amp
    angles = (np.arange(n_angles) + 1) * 90. / (n_angles + 1)
    one = np.zeros((n_angles * n_amp, 6))
    k = 0
    for t in angles:
        th = math.radians(t)
        for a in amps:
            one[k] = (a * math.cos(th) * sig, 0., a * math.sin(th) * sig, 0., 0., pt)
            k += 1
    x = np.concatenate([one] * n_seeds)
    offsets = np.arange(n_seeds + 1, dtype=np.int32) * len(one)
    return gp, lines, x, offsets


# ------------------------------------------------- electron ring (config 4)
#
# FODO arc cells of two sector dipoles, thick quadrupoles and sextupoles,
# with horizontal and vertical BPMs, and two dispersion-carrying straights of
# bend-free cells that hold the RF cavities and the IP markers.  Every bend
# radiates (GP.radiate) and the maps use the expanded Hamiltonian
# (GP.H_expand), the only bend branch with quantum excitation besides the
# exact one (SBEND_Pass, simtrack.cpp:4720-4746).

ELECTRON_G = 0.00115965218128      # anomalous magnetic moment a_e (CODATA 2018)
E_LQ, E_LS = 0.6, 0.25             # quadrupole and sextupole lengths [m]


def _e_half_drifts(lb):
    """Drifts of a half cell: quad-sext, sext-bend, bend-BPM, BPM-quad."""
    return 0.3, 0.3, 0.25, 0.3


def _e_cell(tag, lb, angle, kq, k2f, k2d, quad_nint, sext_nint, rf=None):
    """One arc cell (17 elements).  rf: an RFCAV that replaces the VBPM."""
    d1, d2, d3, d4 = _e_half_drifts(lb)

    def q(name, length, k1):
        e = L.QUAD(name, length, k1 * length, 0.)
        e.Nint = quad_nint
        return e

    def s(name, k2l):
        e = L.SEXT(name, E_LS, k2l, 0.)
        e.Nint = sext_nint
        return e
    return [q(f"QFA{tag}", E_LQ / 2, kq), L.DRIFT("D1", d1), s(f"SF{tag}", k2f), L.DRIFT("D2", d2),
            L.SBEND(f"BA{tag}", lb, angle), L.DRIFT("D3", d3), L.HBPM(f"HBPM{tag}", 0.), L.DRIFT("D4", d4),
            q(f"QD{tag}", E_LQ, -kq), L.DRIFT("D1", d1), s(f"SD{tag}", k2d), L.DRIFT("D2", d2),
            L.SBEND(f"BB{tag}", lb, angle), L.DRIFT("D3", d3),
            rf if rf is not None else L.VBPM(f"VBPM{tag}", 0.), L.DRIFT("D4", d4), q(f"QFB{tag}", E_LQ / 2, kq)]


def _e_straight_cell(tag, lb, kq, quad_nint, centre, kq_scale=1.0):
    """A bend-free cell of the arc cell's length (9 elements); centre is
    the element placed in its middle (an RFCAV or a MARKER)."""
    d1, d2, d3, d4 = _e_half_drifts(lb)
    half = d1 + E_LS + d2 + lb + d3 + d4

    def q(name, length, k1):
        e = L.QUAD(name, length, k1 * length, 0.)
        e.Nint = quad_nint
        return e
    kq = kq * kq_scale
    return [q(f"QFA{tag}", E_LQ / 2, kq), L.DRIFT("DS", half / 2), centre, L.DRIFT("DS", half / 2),
            q(f"QD{tag}", E_LQ, -kq), L.DRIFT("DS", half / 2), L.MARKER(f"MS{tag}", 0.), L.DRIFT("DS", half / 2),
            q(f"QFB{tag}", E_LQ / 2, kq)]


def _e_cell_kq(lb, angle, mu_deg):
    """Quadrupole gradient giving the arc cell a horizontal phase advance of mu_deg."""
    from scipy.optimize import brentq
    from . import optics as O

    def f(kq):
        cell = L.Line(_e_cell("X", lb, angle, kq, 0., 0., 4, 2))
        m = np.eye(3)
        for e in cell.Cell:
            m = O.element_matrices(e)[0] @ m
        return math.degrees(math.acos(max(-1., min(1., (m[0, 0] + m[1, 1]) / 2)))) - mu_deg
    return brentq(f, 0.05, 3.0, xtol=1e-14)


def _e_sextupoles(lb, angle, kq, xi_target):
    """Arc sextupole K2L (SF, SD) that set both chromaticities of an arc
    cell to xi_target / n_cells each (thin-lens sums over the cell)."""
    from . import optics as O
    cell = L.Line(_e_cell("X", lb, angle, kq, 0., 0., 4, 2))
    opt = O.Optics(cell)
    tx = np.array([[opt.beta_x, -opt.alpha_x], [-opt.alpha_x, (1 + opt.alpha_x ** 2) / opt.beta_x]])
    ty = np.array([[opt.beta_y, -opt.alpha_y], [-opt.alpha_y, (1 + opt.alpha_y ** 2) / opt.beta_y]])
    eta = np.array([opt.D, opt.Dp, 1.])
    xi_x = xi_y = 0.
    at = {}
    for e in cell.Cell:
        mx, my = O.element_matrices(e)
        b0x, b0y, d0 = tx[0, 0], ty[0, 0], eta[0]
        eta = mx @ eta
        tx = mx[:2, :2] @ tx @ mx[:2, :2].T
        ty = my @ ty @ my.T
        bx, by, d = (b0x + tx[0, 0]) / 2, (b0y + ty[0, 0]) / 2, (d0 + eta[0]) / 2
        if e.TYPE == "QUAD":
            xi_x -= bx * e.KN / (4 * math.pi)
            xi_y += by * e.KN / (4 * math.pi)
        if e.TYPE == "SEXT":
            at[e.NAME[:2]] = (bx, by, d)
    (bfx, bfy, df), (bdx, bdy, dd) = at["SF"], at["SD"]
    # xi_x + (bfx df k2f + bdx dd k2d)/4pi = target, xi_y - (bfy df k2f + bdy dd k2d)/4pi = target
    a = np.array([[bfx * df, bdx * dd], [-bfy * df, -bdy * dd]]) / (4 * math.pi)
    return np.linalg.solve(a, [xi_target - xi_x, xi_target - xi_y])


def electron_globals_ring(total_energy_mev=18000.0, spin=False):
    """18 GeV electrons with radiation on and the expanded Hamiltonian."""
    gp = electron_globals(total_energy_mev, radiate=1, H_expand=1)
    gp.G = ELECTRON_G
    return gp


def electron_ring(gp, n_arc_cells=200, n_straight_cells=6, bend_length=3.8, mu_deg=86., chromaticity=1.0,
                  harmonic=4800, overvoltage=1.6, rf_in_arcs=False, quad_nint=4, sext_nint=2, aperture=0.1,
                  straight_kq_scale=1.0):
    """An electron ring of 17-element FODO arc cells (two SBENDs each) split
    into two arcs by straights of n_straight_cells bend-free cells.  The RF
    cavities sit in the straights (one per straight cell), or, with
    rf_in_arcs, one per arc cell in place of its VBPM, which keeps the
    radiation sawtooth small.  Total voltage = overvoltage x U0 (the loss of
    Cal_SRLoss_U0rad, simtrack.cpp:20691-20725); PHASE0 = 0 puts the stable
    phase where the RF restores U0.  chromaticity: target of both planes
    (None switches the sextupoles off)."""
    from . import optics as O
    angle = 2 * math.pi / (2 * n_arc_cells)
    kq = _e_cell_kq(bend_length, angle, mu_deg)
    if chromaticity is None:
        k2f = k2d = 0.
    else:
        k2f, k2d = _e_sextupoles(bend_length, angle, kq, chromaticity / (n_arc_cells + n_straight_cells * 0.5))
    rho = bend_length / angle
    u0 = O.CR * (gp.energy / 1000.) ** 4 / rho * 1000.        # MeV per turn

    def cav(tag):
        return L.RFCAV(f"RF{tag}", 0., 0., 0., 0.)             # VRF set below
    half = n_arc_cells // 2
    ns = max(n_straight_cells // 2, 1)
    cells = []
    for arc in range(2):
        for j in range(ns):
            centre = L.MARKER(f"IP{arc + 1}", 0.) if j == 0 else (L.MARKER(f"MK{arc}{j}", 0.) if rf_in_arcs
                                                                  else cav(f"S{arc}{j}"))
            cells += _e_straight_cell(f"S{arc}{j}", bend_length, kq, quad_nint, centre, straight_kq_scale)
        for i in range(half):
            k = arc * half + i
            cells += _e_cell(k, bend_length, angle, kq, k2f, k2d, quad_nint, sext_nint,
                             rf=cav(f"A{k}") if rf_in_arcs else None)
    if not rf_in_arcs and ns == 1:
        cells[1:1] = [cav("S0")]
    ln = L.Line(cells)
    cavities = [e for e in ln.Cell if e.TYPE == "RFCAV"]
    vcav = overvoltage * u0 / len(cavities)
    for e in cavities:
        e.VRF = vcav
    for e in ln.Cell:
        e.APx = e.APy = aperture
    ln.Update()
    set_rf_frequency(ln, gp, harmonic)
    ln.info = dict(kq=kq, k2f=k2f, k2d=k2d, rho=rho, u0_mev=u0, n_cav=len(cavities), vcav_mv=vcav)
    return ln


def config4(n_particles=1048576, seed=2029, spin=True):
    """Config 4: the 18 GeV electron ring (~3,500 elements, radiation, spin)
    with n_particles Gaussian electrons matched to its radiation
    equilibrium, spins vertical.  Returns (globals, line, x)."""
    from . import optics as O
    gp = electron_globals_ring(18000.0)
    ln = electron_ring(gp)
    opt = O.Optics(ln)
    eq = opt.equilibrium(gp)
    sz = _bunch_length(opt, gp, ln, eq["sigma_delta"])
    zs, _ = rf_synchronous(opt, gp, ln)
    x = opt.matched_beam(n_particles, eq["eps_x"], eq["eps_x"] * 0.1, eq["sigma_delta"], sz, seed=seed, spin=spin)
    x[:, 4] += zs
    return gp, ln, x


def rf_synchronous(opt, gp, ln):
    """(z_s, Qs): the synchronous z where the RF restores U0, and the
    linear synchrotron tune.  RFCAV_Pass (simtrack.cpp:6002-6012) adds
    |VRF Q| sin(-k z + pi - PHASE0) / beta0 / E0 to pt with k = 2 pi FRF / c;
    with PHASE0 = 0 that is VRF sin(k z), so sin(k z_s) = U0 / V."""
    u0 = opt.equilibrium(gp)["U0_MeV"]
    cav = [e for e in ln.Cell if e.TYPE == "RFCAV"]
    v = sum(e.VRF for e in cav)
    k = 2 * math.pi * cav[0].FRF / 299792458.
    zs = math.asin(u0 / v) / k
    a = v * k * math.cos(k * zs) / gp.energy
    b = opt.alpha_c * opt.length
    qs = math.acos(1 - a * b / 2) / (2 * math.pi)
    return zs, qs


def _bunch_length(opt, gp, ln, sigma_delta):
    """Equilibrium rms bunch length of the linearised RF: alpha_c C sigma_delta / (2 pi Qs)."""
    _, qs = rf_synchronous(opt, gp, ln)
    return opt.alpha_c * opt.length * sigma_delta / (2 * math.pi * qs)


# ----------------------------------------- the reference's example lattices

def rhic_example(path, rf_volts_mv=4.0):
    """The RHIC set-up of the reference's first example driver: Au ions at
    gamma = 100 (Set_RefPartEnergy(100, 938.27201*197, 79, 197)), the MAD-X
    table read by Read_MADX_Lattice, and the first RFCAV set to VRF = 4 MV,
    FRF = 78250.42279 * harm with GP.harm = 360.  Returns (globals, line)."""
    from . import lattice_io
    gp = L.make_globals(100., 938.27201 * 197, 79, 197)
    ln = lattice_io.read_madx_twiss(path)
    rf = [e for e in ln.Cell if e.TYPE == "RFCAV"][0]
    rf.VRF, rf.FRF = rf_volts_mv, 78250.42279 * 360
    return gp, ln


def hsr_example(path):
    """The EIC hadron ring of the reference's second example driver:
    275 GeV protons, the Bmad table read by Read_BMAD_Lattice, and its two
    cavities set as the driver's (6 MV at h = 2520 and 20 MV at h = 7560 of
    c / length).  Returns (globals, line)."""
    from . import lattice_io
    gp = proton_globals(275000.)
    ln = lattice_io.read_bmad_table(path)
    cav = [e for e in ln.Cell if e.TYPE == "RFCAV"]
    for e, (v, h) in zip(cav, ((6., 2520), (20., 7560))):
        e.VRF, e.FRF = v, h * 299792458. / ln.Length
    return gp, ln
