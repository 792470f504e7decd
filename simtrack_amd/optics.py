"""Host linear optics of a Line: uncoupled Twiss and dispersion, synchrotron
radiation integrals, and the electron equilibrium they imply.

The transfer matrices are the linear parts of the engine's maps in the
expanded-Hamiltonian (H_expand) frame, where x[5] is delta:
- DRIFT and every zero-field element of length L: drift;
- QUAD, and the normal quadrupole term of MULT: thick quadrupole with
  k = K1L / L (a zero-length MULT is a thin lens);
- SBEND: sector bend of curvature h = ANGLE / L with edge focusing
  tan(E) h (SBEND_Pass, simtrack.cpp:4693-4752; the edge kicks are
  cpp:4705-4707 and cpp:4748-4750);
- higher multipoles, kickers, RF and beam-position monitors: no linear
  transverse effect beyond their length.

The radiation integrals are the sums that the tracked electron actually
sees.  SBEND_Pass radiates once after each of its Nint slices (radiate1,
cpp:4730-4746) with the step Lint = L / Nint, so

    I2 = sum h^2 Lint,   I3 = sum |h|^3 Lint,
    I4 = sum D h^3 Lint,   I5 = sum H |h|^3 Lint

over the slice ends, with D and H = gamma D^2 + 2 alpha D D' + beta D'^2 at
each slice end.  I4 has no gradient term, because the engine's bends carry
no gradient, and no edge term, because every bend this module builds has
E1 = E2 = 0.  The classical loss per turn is the reference's
Cal_SRLoss_U0rad (cpp:20691-20725).  The equilibrium follows from the
textbook relations

    eps_x = Cq gamma^2 I5 / (Jx I2),   sigma_delta^2 = Cq gamma^2 I3 / (Jz I2),
    Jx = 1 - I4 / I2,   Jz = 2 + I4 / I2,

with Cq = 55 hbar c / (32 sqrt(3) m_e c^2) = 3.8319e-13 m.
"""
import math

import numpy as np

# CODATA-2018 electron values
HBAR_C_MEV_M = 197.3269804e-15
ELECTRON_MEV = 0.51099895
CQ = 55.0 * HBAR_C_MEV_M / (32.0 * math.sqrt(3.0) * ELECTRON_MEV)
CR = 8.84627632682e-5      # m / GeV^3, Cr of simtrack.cpp:25

_ZERO_FIELD = {"DRIFT", "MARKER", "BPM", "HBPM", "VBPM", "KICKER", "HKICKER", "VKICKER", "RFCAV", "CRABRF",
               "SEXT", "OCT", "SKEWQ"}


def _drift(L):
    mx = np.array([[1., L, 0.], [0., 1., 0.], [0., 0., 1.]])
    my = np.array([[1., L], [0., 1.]])
    return mx, my


def _quad(L, k1l):
    if L == 0.:
        mx = np.array([[1., 0., 0.], [-k1l, 1., 0.], [0., 0., 1.]])
        my = np.array([[1., 0.], [k1l, 1.]])
        return mx, my
    k = k1l / L

    def m2(kk):
        if kk > 0:
            s = math.sqrt(kk)
            return np.array([[math.cos(s * L), math.sin(s * L) / s], [-s * math.sin(s * L), math.cos(s * L)]])
        if kk < 0:
            s = math.sqrt(-kk)
            return np.array([[math.cosh(s * L), math.sinh(s * L) / s], [s * math.sinh(s * L), math.cosh(s * L)]])
        return np.array([[1., L], [0., 1.]])
    mx = np.eye(3)
    mx[:2, :2] = m2(k)
    return mx, m2(-k)


def _sector(L, angle):
    """Sector bend body (no edges) of length L and bending angle `angle`."""
    h = angle / L
    c, s = math.cos(angle), math.sin(angle)
    mx = np.array([[c, s / h, (1. - c) / h], [-h * s, c, s], [0., 0., 1.]])
    my = np.array([[1., L], [0., 1.]])
    return mx, my


def _edge(h, e):
    t = math.tan(e) * h
    mx = np.array([[1., 0., 0.], [t, 1., 0.], [0., 0., 1.]])
    my = np.array([[1., 0.], [-t, 1.]])
    return mx, my


def element_matrices(e):
    """(3x3 horizontal (x, px, delta), 2x2 vertical) linear transfer matrices."""
    t = e.TYPE
    if t == "SBEND":
        if e.ANGLE == 0.:
            return _drift(e.L)
        h = e.ANGLE / e.L
        a1, b1 = _edge(h, e.E1)
        a2, b2 = _sector(e.L, e.ANGLE)
        a3, b3 = _edge(h, e.E2)
        return a3 @ a2 @ a1, b3 @ b2 @ b1
    if t == "QUAD":
        return _quad(e.L, e.KN)
    if t == "MULT":
        return _quad(e.L, e.KNL[1])
    if t in _ZERO_FIELD:
        return _drift(e.L)
    raise ValueError(f"linear optics: element type {t} is not handled")


def _periodic_x(m):
    """Periodic beta, alpha, mu and dispersion (D, D') of a 3x3 one-turn matrix."""
    cosmu = (m[0, 0] + m[1, 1]) / 2
    if abs(cosmu) >= 1:
        raise ValueError("horizontal motion is unstable")
    mu = math.acos(cosmu)
    if m[0, 1] < 0:
        mu = 2 * math.pi - mu
    beta = m[0, 1] / math.sin(mu)
    alpha = (m[0, 0] - m[1, 1]) / (2 * math.sin(mu))
    d = np.linalg.solve(np.eye(2) - m[:2, :2], m[:2, 2])
    return beta, alpha, mu, d[0], d[1]


def _periodic_y(m):
    cosmu = (m[0, 0] + m[1, 1]) / 2
    if abs(cosmu) >= 1:
        raise ValueError("vertical motion is unstable")
    mu = math.acos(cosmu)
    if m[0, 1] < 0:
        mu = 2 * math.pi - mu
    return m[0, 1] / math.sin(mu), (m[0, 0] - m[1, 1]) / (2 * math.sin(mu)), mu


def _twiss_matrix(beta, alpha):
    return np.array([[beta, -alpha], [-alpha, (1 + alpha * alpha) / beta]])


class Optics:
    """Periodic uncoupled optics of a ring Line at its start, propagated to
    every element end and to every radiation point of the bends."""

    def __init__(self, line):
        self.line = line
        mats = [element_matrices(e) for e in line.Cell]
        mx, my = np.eye(3), np.eye(2)
        for a, b in mats:
            mx, my = a @ mx, b @ my
        self.mx, self.my = mx, my
        self.beta_x, self.alpha_x, mux, self.D, self.Dp = _periodic_x(mx)
        self.beta_y, self.alpha_y, muy = _periodic_y(my)
        self.tune_x, self.tune_y = mux / (2 * math.pi), muy / (2 * math.pi)
        self.length = sum(e.L for e in line.Cell)
        self._radiation_integrals()

    def _radiation_integrals(self):
        tx = _twiss_matrix(self.beta_x, self.alpha_x)
        eta = np.array([self.D, self.Dp, 1.])
        I1 = I2 = I3 = I4 = I5 = 0.
        for e in self.line.Cell:
            if e.TYPE == "SBEND" and e.ANGLE != 0.:
                h = e.ANGLE / e.L
                lint = e.L / e.Nint
                a, _ = _edge(h, e.E1)
                eta = a @ eta
                tx = a[:2, :2] @ tx @ a[:2, :2].T
                ms, _ = _sector(lint, e.ANGLE / e.Nint)
                for _k in range(e.Nint):
                    d_mid = eta[0]
                    eta = ms @ eta
                    tx = ms[:2, :2] @ tx @ ms[:2, :2].T
                    # radiation point: the end of the slice
                    beta, alpha = tx[0, 0], -tx[0, 1]
                    gamma = tx[1, 1]
                    H = gamma * eta[0] ** 2 + 2 * alpha * eta[0] * eta[1] + beta * eta[1] ** 2
                    I1 += 0.5 * (d_mid + eta[0]) * h * lint
                    I2 += h * h * lint
                    I3 += abs(h) ** 3 * lint
                    I4 += eta[0] * h ** 3 * lint
                    I5 += H * abs(h) ** 3 * lint
                a, _ = _edge(h, e.E2)
                eta = a @ eta
                tx = a[:2, :2] @ tx @ a[:2, :2].T
            else:
                a, _ = element_matrices(e)
                eta = a @ eta
                tx = a[:2, :2] @ tx @ a[:2, :2].T
        self.I1, self.I2, self.I3, self.I4, self.I5 = I1, I2, I3, I4, I5
        self.alpha_c = I1 / self.length

    def equilibrium(self, gp):
        """Electron equilibrium (dict): U0 [MeV], damping partition numbers,
        damping times in turns, eps_x [m], sigma_delta."""
        g = gp.gamma
        e_gev = gp.energy / 1000.
        u0 = CR * e_gev ** 4 * self.I2 / (2 * math.pi) * 1000.
        jx = 1. - self.I4 / self.I2
        jz = 2. + self.I4 / self.I2
        u0_over_e = u0 / gp.energy
        return dict(U0_MeV=u0, Jx=jx, Jy=1.0, Jz=jz,
                    tau_x_turns=2. / (jx * u0_over_e), tau_y_turns=2. / u0_over_e,
                    tau_z_turns=2. / (jz * u0_over_e),
                    eps_x=CQ * g * g * self.I5 / (jx * self.I2),
                    sigma_delta=math.sqrt(CQ * g * g * self.I3 / (jz * self.I2)))

    def matched_beam(self, n, eps_x, eps_y, sigma_delta, sigma_z, seed=1, spin=False):
        """Gaussian particles matched to the periodic Twiss and dispersion at
        the start (x[5] = delta)."""
        r = np.random.Generator(np.random.PCG64(seed))
        x = np.zeros((n, 9 if spin else 6))
        u = r.standard_normal((n, 6))
        d = sigma_delta * u[:, 5]
        bx, ax, by, ay = self.beta_x, self.alpha_x, self.beta_y, self.alpha_y
        x[:, 0] = math.sqrt(eps_x * bx) * u[:, 0] + self.D * d
        x[:, 1] = math.sqrt(eps_x / bx) * (u[:, 1] - ax * u[:, 0]) + self.Dp * d
        x[:, 2] = math.sqrt(eps_y * by) * u[:, 2]
        x[:, 3] = math.sqrt(eps_y / by) * (u[:, 3] - ay * u[:, 2])
        x[:, 4] = sigma_z * u[:, 4]
        x[:, 5] = d
        if spin:
            x[:, 7] = 1.0
        return x


def eigen_emittances(x):
    """Normal-mode emittances of a particle set (rows of x[:, :6]): the
    moduli of the eigenvalues of Sigma S, S the symplectic unit, sorted
    descending.  They are invariant under any linear symplectic change of
    coordinates, so dispersion needs no model."""
    z = np.asarray(x[:, :6], dtype=np.float64)
    z = z - z.mean(axis=0)
    sig = z.T @ z / len(z)
    s = np.zeros((6, 6))
    for k in range(3):
        s[2 * k, 2 * k + 1] = 1.
        s[2 * k + 1, 2 * k] = -1.
    ev = np.linalg.eigvals(sig @ s)
    em = np.sort(np.abs(ev.imag))[::-1]
    return em[::2]
