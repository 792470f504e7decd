/*
 * st_lower.cpp -- lowering of st_element records (the reference's Element
 * objects) to the packed device table: per-element constants evaluated once
 * on the host with the reference's own expression trees; st_table_build.
 */
#include "st_capi_internal.h"

/* ---------------------------------------------------------- lowering */

namespace {

struct Builder {
    const st_globals &gp;
    bool spin;
    dev_table &t;
    std::string err;

    Builder(const st_globals &g, bool s, dev_table &tab) : gp(g), spin(s), t(tab) {}

    int put(double v)
    {
        t.par.push_back(v);
        return (int)t.par.size() - 1;
    }

    st_op new_op(int kind)
    {
        st_op op;
        std::memset(&op, 0, sizeof op);
        op.kind = kind;
        op.po = (int)t.par.size();
        op.pm = -1;
        return op;
    }

    /* Cal_W_Matrix, simtrack.cpp:3485-3503 */
    static void w_matrix(double theta, double phi, double psi, double W[9])
    {
        double costhe = st_cos(theta), cosphi = st_cos(phi), cospsi = st_cos(psi);
        double sinthe = st_sin(theta), sinphi = st_sin(phi), sinpsi = st_sin(psi);
        W[0] = +costhe * cospsi - sinthe * sinphi * sinpsi;
        W[1] = -costhe * sinpsi - sinthe * sinphi * cospsi;
        W[2] = sinthe * cosphi;
        W[3] = cosphi * sinpsi;
        W[4] = cosphi * cospsi;
        W[5] = sinphi;
        W[6] = -sinthe * cospsi - costhe * sinphi * sinpsi;
        W[7] = +sinthe * sinpsi - costhe * sinphi * cospsi;
        W[8] = costhe * cosphi;
    }

    /* mat_mult, simtrack.cpp:1351-1362 */
    static void mat_mult(const double *a, const double *b, double *c, int m, int n, int k)
    {
        for (int i = 0; i <= m - 1; i++)
            for (int j = 0; j <= k - 1; j++) {
                int u = i * k + j;
                c[u] = 0.0;
                for (int l = 0; l <= n - 1; l++) c[u] = c[u] + a[i * n + l] * b[l * k + j];
            }
    }

    /*
     * Misalignment block of GtoL/LtoG (simtrack.cpp:3701-3741) for an element of
     * length L and bend angle `angle`: offset/roll fast path (GtoL1/LtoG1) or
     * the general MAD-X frame change (GtoL2/LtoG2) with its element-constant
     * matrices evaluated here.
     */
    void misalign(st_op &op, const st_element &e, double L, double angle)
    {
        double dx = e.dx, dy = e.dy, ds = e.ds, theta = e.dtheta, phi = e.dphi, psi = e.dpsi;
        double temp = std::fabs(dx) + std::fabs(dy) + std::fabs(ds) + std::fabs(theta) + std::fabs(phi) +
                      std::fabs(psi);
        if (temp < 1.0e-9) return;
        if (std::fabs(theta) + std::fabs(phi) < 1.0e-9) {
            op.flags |= ST_F_MIS1;
            if (std::fabs(dx) + std::fabs(dy) > 1.0e-10) op.flags |= ST_F_MIS1_OFS;
            if (std::fabs(psi) > 1.0e-10) op.flags |= ST_F_MIS1_ROT;
            op.pm = put(dx);
            put(dy);
            put(st_cos(psi));
            put(st_sin(psi));
            return;
        }
        op.flags |= ST_F_MIS2;
        const double beta = gp.beta, gamma = gp.gamma;
        double W[9], RM[36];
        /* GtoL2, cpp:3567-3626 */
        w_matrix(theta, phi, psi, W);
        double s2 = (W[0 * 3 + 2] * dx + W[1 * 3 + 2] * dy + W[2 * 3 + 2] * ds) / W[2 * 3 + 2];
        for (int i = 0; i < 36; i++) RM[i] = 0.;
        for (int i = 0; i < 6; i++) RM[i * 6 + i] = 1.0;
        RM[1 * 6 + 1] = W[0 * 3 + 0];
        RM[1 * 6 + 3] = W[1 * 3 + 0];
        RM[1 * 6 + 5] = W[2 * 3 + 0] / beta;
        RM[3 * 6 + 1] = W[0 * 3 + 1];
        RM[3 * 6 + 3] = W[1 * 3 + 1];
        RM[3 * 6 + 5] = W[2 * 3 + 1] / beta;
        RM[0 * 6 + 0] = W[1 * 3 + 1] / W[2 * 3 + 2];
        RM[0 * 6 + 1] = RM[0 * 6 + 0] * s2;
        RM[0 * 6 + 2] = -W[0 * 3 + 1] / W[2 * 3 + 2];
        RM[0 * 6 + 3] = RM[0 * 6 + 2] * s2;
        RM[2 * 6 + 0] = -W[1 * 3 + 0] / W[2 * 3 + 2];
        RM[2 * 6 + 1] = RM[2 * 6 + 0] * s2;
        RM[2 * 6 + 2] = W[0 * 3 + 0] / W[2 * 3 + 2];
        RM[2 * 6 + 3] = RM[2 * 6 + 2] * s2;
        RM[4 * 6 + 0] = W[0 * 3 + 2] / (W[2 * 3 + 2] * beta);
        RM[4 * 6 + 1] = RM[4 * 6 + 0] * s2;
        RM[4 * 6 + 2] = W[1 * 3 + 2] / (W[2 * 3 + 2] * beta);
        RM[4 * 6 + 3] = RM[4 * 6 + 2] * s2;
        RM[4 * 6 + 5] = -s2 / (beta * gamma * beta * gamma);
        op.pm = (int)t.par.size();
        for (int i = 0; i < 36; i++) put(RM[i]);
        put(-((W[1 * 3 + 1] * dx - W[0 * 3 + 1] * dy) / W[2 * 3 + 2]));
        put(W[2 * 3 + 0]);
        put(-((W[0 * 3 + 0] * dy - W[1 * 3 + 0] * dx) / W[2 * 3 + 2]));
        put(W[2 * 3 + 1]);
        put(-(s2 / beta));
        /* LtoG2, cpp:3628-3699, with Cal_VE_WE_Matrx (cpp:3505-3565); an
           angle of exactly +-1e-12 (left uninitialised by the reference) is
           treated as straight */
        double VE[3], WE[9], V[3], V1[3], M1[9], M2[9];
        w_matrix(theta, phi, psi, W);
        if (!(std::fabs(angle) > 1.0e-12)) {
            VE[0] = VE[1] = 0.;
            VE[2] = L;
            for (int i = 0; i < 9; i++) WE[i] = 0.;
            for (int i = 0; i < 3; i++) WE[i * 3 + i] = 1.0;
        } else {
            double tilt = 0.;
            double cospsi = st_cos(tilt), sinpsi = st_sin(tilt);
            double costhe = st_cos(angle), sinthe = st_sin(angle);
            double tt = L * (costhe - 1.0) / angle;
            VE[0] = tt * cospsi;
            VE[1] = tt * sinpsi;
            VE[2] = L * st_sin(angle) / angle;
            WE[0 * 3 + 0] = costhe * cospsi * cospsi + sinpsi * sinpsi;
            WE[1 * 3 + 0] = (costhe - 1.0) * cospsi * sinpsi;
            WE[2 * 3 + 0] = sinthe * cospsi;
            WE[0 * 3 + 1] = WE[1 * 3 + 0];
            WE[1 * 3 + 1] = costhe * sinpsi * sinpsi + cospsi * cospsi;
            WE[2 * 3 + 1] = sinthe * sinpsi;
            WE[0 * 3 + 2] = -WE[2 * 3 + 0];
            WE[1 * 3 + 2] = -WE[2 * 3 + 1];
            WE[2 * 3 + 2] = costhe;
        }
        V[0] = dx + W[0 * 3 + 0] * VE[0] + W[0 * 3 + 1] * VE[1] + W[0 * 3 + 2] * VE[2] - VE[0];
        V[1] = dy + W[1 * 3 + 0] * VE[0] + W[1 * 3 + 1] * VE[1] + W[1 * 3 + 2] * VE[2] - VE[1];
        V[2] = ds + W[2 * 3 + 0] * VE[0] + W[2 * 3 + 1] * VE[1] + W[2 * 3 + 2] * VE[2] - VE[2];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) M1[i * 3 + j] = WE[j * 3 + i];
        for (int i = 0; i < 3; i++) {
            V1[i] = 0;
            for (int j = 0; j < 3; j++) V1[i] = V1[i] + M1[i * 3 + j] * V[j];
        }
        for (int i = 0; i < 3; i++) V[i] = V1[i];
        mat_mult(M1, W, M2, 3, 3, 3);
        mat_mult(M2, WE, W, 3, 3, 3);
        s2 = -(W[0 * 3 + 2] * V[0] + W[1 * 3 + 2] * V[1] + W[2 * 3 + 2] * V[2]) / W[2 * 3 + 2];
        for (int i = 0; i < 36; i++) RM[i] = 0.;
        for (int i = 0; i < 6; i++) RM[i * 6 + i] = 1.0;
        RM[0 * 6 + 0] = W[0 * 3 + 0];
        RM[2 * 6 + 0] = W[1 * 3 + 0];
        RM[4 * 6 + 0] = W[2 * 3 + 0] / beta;
        RM[0 * 6 + 2] = W[0 * 3 + 1];
        RM[2 * 6 + 2] = W[1 * 3 + 1];
        RM[4 * 6 + 2] = W[2 * 3 + 1] / beta;
        RM[1 * 6 + 1] = W[1 * 3 + 1] / W[2 * 3 + 2];
        RM[0 * 6 + 1] = RM[1 * 6 + 1] * s2;
        RM[3 * 6 + 1] = -W[0 * 3 + 1] / W[2 * 3 + 2];
        RM[2 * 6 + 1] = RM[3 * 6 + 1] * s2;
        RM[1 * 6 + 3] = -W[1 * 3 + 0] / W[2 * 3 + 2];
        RM[0 * 6 + 3] = RM[1 * 6 + 3] * s2;
        RM[3 * 6 + 3] = W[0 * 3 + 0] / W[2 * 3 + 2];
        RM[2 * 6 + 3] = RM[3 * 6 + 3] * s2;
        RM[1 * 6 + 5] = W[0 * 3 + 2] / (W[2 * 3 + 2] * beta);
        RM[0 * 6 + 5] = RM[1 * 6 + 5] * s2;
        RM[3 * 6 + 5] = W[1 * 3 + 2] / (W[2 * 3 + 2] * beta);
        RM[2 * 6 + 5] = RM[3 * 6 + 5] * s2;
        RM[4 * 6 + 5] = -s2 / (beta * gamma * beta * gamma);
        for (int i = 0; i < 36; i++) put(RM[i]);
        put((W[1 * 3 + 1] * V[0] - W[0 * 3 + 1] * V[1]) / W[2 * 3 + 2]);
        put(-W[2 * 3 + 0]);
        put((W[0 * 3 + 0] * V[1] - W[1 * 3 + 0] * V[0]) / W[2 * 3 + 2]);
        put(-W[2 * 3 + 1]);
        put(-(s2 / beta));
    }

    int drift_op(double L, const st_element *mis_e = nullptr, double mis_L = 0., double mis_angle = 0.)
    {
        st_op op = new_op(L == 0.0 && !std::signbit(L) ? K_DRIFT0 : K_DRIFT);
        if (op.kind == K_DRIFT) {
            put(L);
            put(L / gp.beta);
        }
        if (mis_e) misalign(op, *mis_e, mis_L, mis_angle);
        t.ops.push_back(op);
        return 0;
    }

    static int order_mask(const double *kn, const double *ks, int norder)
    {
        int m = 0;
        for (int i = 1; i < norder + 1; i++)
            if (kn[i] != 0. || ks[i] != 0.) m |= 1 << i;
        return m;
    }

    /* QUAD_Pass (cpp:4882) / SEXT_Pass (cpp:5004) / OCT_Pass (cpp:5077) and sPass twins */
    int multipole_op(const st_element &e, int thick_kind, int thin_kind, double kn, double ks)
    {
        double L = e.l;
        int Nint = e.nint;
        if (kn == 0. && ks == 0.) return drift_op(L, &e, L, 0.);
        if (L == 0.) {
            st_op op = new_op(thin_kind);
            put(kn);
            put(ks);
            put((double)Nint);
            misalign(op, e, L, 0.);
            t.ops.push_back(op);
            return 0;
        }
        if (Nint < 1) return -1;
        st_op op = new_op(thick_kind);
        op.nint = Nint;
        double Lint = L / Nint;
        put(Fdrift1 * Lint);
        put((Fdrift1 * Lint) / gp.beta);
        put(Fdrift2 * Lint);
        put((Fdrift2 * Lint) / gp.beta);
        put(Fkick1 * kn / Nint);
        put(Fkick1 * ks / Nint);
        put(Fkick2 * kn / Nint);
        put(Fkick2 * ks / Nint);
        put(kn);
        put(ks);
        put((double)Nint);
        put(kn / L);
        if (!spin && thick_kind == K_QUAD && gp.quad_fringe) op.flags |= ST_F_FRINGE;
        misalign(op, e, L, 0.);
        t.ops.push_back(op);
        return 0;
    }

    void kick_arrays(const double *KNL, const double *KNSL, int Nint, int norder, st_op &op)
    {
        double k1[11], s1[11], k2[11], s2[11];
        for (int i = 0; i < 11; i++) {
            k1[i] = Fkick1 * KNL[i] / Nint;
            s1[i] = Fkick1 * KNSL[i] / Nint;
            k2[i] = Fkick2 * KNL[i] / Nint;
            s2[i] = Fkick2 * KNSL[i] / Nint;
        }
        for (int i = 0; i <= norder; i++) put(k1[i]);
        for (int i = 0; i <= norder; i++) put(s1[i]);
        for (int i = 0; i <= norder; i++) put(k2[i]);
        for (int i = 0; i <= norder; i++) put(s2[i]);
        op.mask1 = order_mask(k1, s1, norder);
        op.mask2 = order_mask(k2, s2, norder);
    }

    void plain_arrays(const double *KNL, const double *KNSL, int norder)
    {
        for (int i = 0; i <= norder; i++) put(KNL[i]);
        for (int i = 0; i <= norder; i++) put(KNSL[i]);
    }

    /* MULT_Pass (cpp:5152) / MULT_sPass (cpp:7702).  mis_angle: the ANGLE the
       owning element passes to GtoL/LtoG (0 for MULT, ANGLE for the bend
       multipoles, whose L = 0 branch is a plain mult_kick_pass). */
    int mult_op(const st_element &e, double mis_angle = 0.)
    {
        double L = e.l;
        int Nint = e.nint, norder = e.norder;
        if (L == 0.) {
            st_op op = new_op(K_MULT_THIN);
            op.norder = norder;
            op.mask1 = order_mask(e.knl, e.knsl, norder);
            plain_arrays(e.knl, e.knsl, norder);
            misalign(op, e, L, mis_angle);
            t.ops.push_back(op);
            return 0;
        }
        if (Nint < 1) return -1;
        st_op op = new_op(K_MULT);
        op.nint = Nint;
        op.norder = norder;
        double Lint = L / Nint;
        put(Fdrift1 * Lint);
        put((Fdrift1 * Lint) / gp.beta);
        put(Fdrift2 * Lint);
        put((Fdrift2 * Lint) / gp.beta);
        put((double)Nint);
        put((double)order_mask(e.knl, e.knsl, norder));
        kick_arrays(e.knl, e.knsl, Nint, norder, op);
        plain_arrays(e.knl, e.knsl, norder);
        misalign(op, e, L, 0.);
        t.ops.push_back(op);
        return 0;
    }

    /* sbend_exact_pass piece (L, href, 1/href, cos, sin, L/beta) */
    void exact_piece(double L, double A, double c, double s)
    {
        double href = A / L;
        put(L);
        put(href);
        put(1.0 / href);
        put(c);
        put(s);
        put(L / gp.beta);
    }

    /* general_sbend_exact_pass piece (L, rhoc, b0, b0*rhoc, cos, sin, L/beta) */
    void general_piece(double L, double A, double K0L, double c, double s)
    {
        double rhoc = L / A, b0 = K0L / L;
        put(L);
        put(rhoc);
        put(b0);
        put(b0 * rhoc);
        put(c);
        put(s);
        put(L / gp.beta);
    }

    void yoshida_pieces(double Lint)
    {
        put(Fdrift1 * Lint);
        put((Fdrift1 * Lint) / gp.beta);
        put(Fdrift2 * Lint);
        put((Fdrift2 * Lint) / gp.beta);
        put(Fkick1 * Lint);
        put(Fkick2 * Lint);
    }

    /* SBEND_Pass (cpp:4762) / SBEND_sPass (cpp:7486) */
    int sbend_op(const st_element &e)
    {
        double L = e.l, Angle = e.angle;
        int Nint = e.nint;
        if (Angle == 0.) return drift_op(L, &e, L, Angle);
        double href = Angle / L;
        bool hexp = gp.H_expand != 0, rad = gp.radiate != 0;
        if (!spin && !hexp && !rad) {
            st_op op = new_op(K_SBEND_EXACT);
            put(st_tan(e.e1) * href);
            put(st_tan(e.e2) * href);
            put(L);
            put(href);
            put(1.0 / href);
            put(st_cos(Angle));
            put(st_sin(Angle));
            put(L / gp.beta);
            misalign(op, e, L, Angle);
            t.ops.push_back(op);
            return 0;
        }
        if (Nint < 1) return -1;
        st_op op = new_op(K_SBEND_SLICED);
        op.nint = Nint;
        double Lint = L / Nint, An = Angle / Nint;
        put(st_tan(e.e1) * href);
        put(st_tan(e.e2) * href);
        put(href);
        put(Lint);
        put(An);
        put((double)Nint);
        if (hexp) {
            yoshida_pieces(Lint);
        } else {
            double hs = An / Lint;
            put(hs);
            put(1.0 / hs);
            put(st_cos(An));
            put(st_sin(An));
            put(Lint / gp.beta);
        }
        if (rad) op.flags |= ST_F_RAD;
        misalign(op, e, L, Angle);
        t.ops.push_back(op);
        return 0;
    }

    /* GSBEND_Pass (cpp:4818); with GP.radiate, radiate2 after every slice */
    int gsbend_op(const st_element &e)
    {
        double L = e.l, Angle = e.angle, K0L = e.k0l;
        int Nint = e.nint;
        if (Angle == 0.) return drift_op(L, &e, L, Angle);
        double href = Angle / L, hreal = K0L / L;
        bool rad = gp.radiate != 0;
        if ((gp.H_expand || rad) && Nint < 1) return -1;
        if (gp.H_expand) {
            st_op op = new_op(K_GSBEND_YOSHIDA);
            op.nint = Nint;
            double Lint = L / Nint;
            put(st_tan(e.e1) * hreal);
            put(st_tan(e.e2) * hreal);
            put(Fdrift1 * Lint);
            put((Fdrift1 * Lint) / gp.beta);
            put(Fdrift2 * Lint);
            put((Fdrift2 * Lint) / gp.beta);
            put(Fkick1 * Lint);
            put(Fkick2 * Lint);
            put(href);
            put(hreal);
            if (rad) {
                put(K0L / Nint);
                put(Lint);
                op.flags |= ST_F_RAD;
            }
            misalign(op, e, L, Angle);
            t.ops.push_back(op);
            return 0;
        }
        st_op op = new_op(K_GSBEND_EXACT);
        put(st_tan(e.e1) * hreal);
        put(st_tan(e.e2) * hreal);
        if (rad) {
            double Lint = L / Nint, An = Angle / Nint;
            op.nint = Nint;
            general_piece(Lint, An, K0L / Nint, st_cos(An), st_sin(An));
            put(href);
            put(K0L / Nint);
            put(Lint);
            op.flags |= ST_F_RAD;
        } else {
            general_piece(L, Angle, K0L, st_cos(Angle), st_sin(Angle));
        }
        misalign(op, e, L, Angle);
        t.ops.push_back(op);
        return 0;
    }

    /* GMULT/SBENDMULT/SMULT_Pass (cpp:5223-5595) and their sPass twins */
    int bendmult_op(const st_element &e)
    {
        double L = e.l, Angle = e.angle;
        int Nint = e.nint, norder = e.norder;
        if (L == 0.) {
            if (!spin) return mult_op(e, Angle);
            st_op op = new_op(K_BENDMULT_THIN);
            op.norder = norder;
            op.mask1 = order_mask(e.knl, e.knsl, norder);
            put(Angle);
            plain_arrays(e.knl, e.knsl, norder);
            misalign(op, e, L, Angle);
            t.ops.push_back(op);
            return 0;
        }
        if (Nint < 1) return -1;
        st_op op = new_op(K_BENDMULT);
        op.nint = Nint;
        op.norder = norder;
        double href = Angle / L, Lint = L / Nint;
        put(st_tan(e.e1) * href);
        put(st_tan(e.e2) * href);
        put(href);
        put(Lint);
        put((double)Nint);
        put(Angle);
        put(Angle / Nint);
        put((double)order_mask(e.knl, e.knsl, norder));
        if (gp.H_expand) {
            yoshida_pieces(Lint);
            for (int i = 0; i < 6; i++) put(0.);
        } else {
            double A1 = Fdrift1 * Angle / Nint, A2 = Fdrift2 * Angle / Nint;
            exact_piece(Fdrift1 * Lint, A1, st_cos(A1), st_sin(A1));
            exact_piece(Fdrift2 * Lint, A2, st_cos(A2), st_sin(A2));
        }
        kick_arrays(e.knl, e.knsl, Nint, norder, op);
        plain_arrays(e.knl, e.knsl, norder);
        if (gp.radiate) op.flags |= ST_F_RAD;
        misalign(op, e, L, Angle);
        t.ops.push_back(op);
        return 0;
    }

    /* GSBENDMULT_Pass (cpp:5597) */
    int gsbendmult_op(const st_element &e)
    {
        double L = e.l, Angle = e.angle, K0L = e.k0l;
        int Nint = e.nint, norder = e.norder;
        if (spin) return bendmult_op(e); /* GSBENDMULT_sPass == SBENDMULT_sPass (cpp:8024) */
        if (L == 0.) return mult_op(e, Angle);
        if (Nint < 1) return -1;
        st_op op = new_op(K_GSBENDMULT);
        op.nint = Nint;
        op.norder = norder;
        double href = Angle / L, hreal = K0L / L, Lint = L / Nint;
        put(st_tan(e.e1) * hreal);
        put(st_tan(e.e2) * hreal);
        put(href);
        put(hreal);
        put(Lint);
        put((double)Nint);
        put(K0L / Nint);
        put(0.);
        if (gp.H_expand) {
            yoshida_pieces(Lint);
            for (int i = 0; i < 8; i++) put(0.);
        } else {
            double A1 = Fdrift1 * Angle / Nint, A2 = Fdrift2 * Angle / Nint;
            general_piece(Fdrift1 * Lint, A1, Fdrift1 * K0L / Nint, st_cos(A1), st_sin(A1));
            general_piece(Fdrift2 * Lint, A2, Fdrift2 * K0L / Nint, st_cos(A2), st_sin(A2));
        }
        kick_arrays(e.knl, e.knsl, Nint, norder, op);
        if (gp.radiate) op.flags |= ST_F_RAD;
        misalign(op, e, L, Angle);
        t.ops.push_back(op);
        return 0;
    }

    /* slice centroids of BB6D / BB6D_Angle (cpp:6594-6620, 6937-6962) */
    static void slice_centers(int N_slice, double sigma_l, double *z)
    {
        if (N_slice == 11) {
            static const double tab[11] = {1.7996765362153655,   1.1049618149811977,   0.75072324235797372,
                                           0.47407703558819131,  0.23041113693387344,  0.,
                                           -0.23041113693387344, -0.47407703558819131, -0.75072324235797372,
                                           -1.1049618149811977,  -1.7996765362153655};
            for (int i = 0; i < 11; i++) z[i] = tab[i] * sigma_l;
            return;
        }
        double y[64];
        for (int i = 0; i < N_slice - 1; i++) y[i] = st_ugaussian_pinv((i + 1) * 1.0 / N_slice);
        z[0] = N_slice * (st_exp(-y[0] * y[0] / 2)) / std::sqrt(2.0 * 3.14159265);
        z[N_slice - 1] = N_slice * (0 - st_exp(-y[0] * y[0] / 2)) / std::sqrt(2.0 * 3.14159265);
        for (int i = 1; i < N_slice - 1; i++)
            z[i] = N_slice * (st_exp(-y[i] * y[i] / 2) - st_exp(-y[i - 1] * y[i - 1] / 2)) / std::sqrt(2.0 * 3.14159265);
        for (int i = 0; i < N_slice; i++) z[i] = z[i] * sigma_l;
    }

    int lower(const st_element &e)
    {
        const double PI = 3.14159265358979323;
        switch (e.type) {
        case ST_DRIFT: case ST_MARKER: case ST_BPM: case ST_HBPM: case ST_VBPM:
            return drift_op(e.l);
        case ST_ROTAT: case ST_SNAKE: {
            if (!spin) return drift_op(e.l);
            /* ROTAT_sPass / SNAKE_sPass, cpp:8286-8348 */
            st_op op = new_op(K_ROTAT);
            double costheta = st_cos(-e.rot_angle), sintheta = st_sin(-e.rot_angle);
            const double *n = e.rot_n;
            double F = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            double uF[3] = {n[0] / F, n[1] / F, n[2] / F};
            double rot[9];
            rot[0] = uF[0] * uF[0] * (1 - costheta) + costheta;
            rot[1] = uF[0] * uF[1] * (1 - costheta) - uF[2] * sintheta;
            rot[2] = uF[0] * uF[2] * (1 - costheta) + uF[1] * sintheta;
            rot[3] = uF[0] * uF[1] * (1 - costheta) + uF[2] * sintheta;
            rot[4] = uF[1] * uF[1] * (1 - costheta) + costheta;
            rot[5] = uF[1] * uF[2] * (1 - costheta) - uF[0] * sintheta;
            rot[6] = uF[0] * uF[2] * (1 - costheta) - uF[1] * sintheta;
            rot[7] = uF[1] * uF[2] * (1 - costheta) + uF[0] * sintheta;
            rot[8] = uF[2] * uF[2] * (1 - costheta) + costheta;
            put(e.l / 2.);
            put((e.l / 2.) / gp.beta);
            for (int i = 0; i < 9; i++) put(rot[i]);
            t.ops.push_back(op);
            return 0;
        }
        case ST_SBEND:
            return sbend_op(e);
        case ST_GSBEND:
            if (spin) return sbend_op(e); /* GSBEND::sPass calls SBEND_sPass, cpp:8619-8623 */
            return gsbend_op(e);
        case ST_QUAD:
            return multipole_op(e, K_QUAD, K_QUAD_THIN, e.knl[1], e.knsl[1]);
        case ST_SKEWQ:
            return multipole_op(e, K_QUAD, K_QUAD_THIN, 0., e.knsl[1]);
        case ST_SEXT:
            return multipole_op(e, K_SEXT, K_SEXT_THIN, e.knl[2], e.knsl[2]);
        case ST_OCT:
            return multipole_op(e, K_OCT, K_OCT_THIN, e.knl[3], e.knsl[3]);
        case ST_MULT:
            return mult_op(e);
        case ST_GMULT: case ST_SBENDMULT: case ST_SMULT:
            return bendmult_op(e);
        case ST_GSBENDMULT:
            return gsbendmult_op(e);
        case ST_KICKER: case ST_HKICKER: case ST_VKICKER: {
            /* KICK_Pass cpp:5887, KICK_sPass cpp:8193 */
            st_op op = new_op(K_KICK);
            double hk = e.type == ST_VKICKER ? 0. : e.hkick;
            double vk = e.type == ST_HKICKER ? 0. : e.vkick;
            put(e.l / 2.0);
            put((e.l / 2.0) / gp.beta);
            put(hk);
            put(vk);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_RFCAV: {
            /* RFCAV_Pass cpp:6002 (sPass identical, cpp:10712) */
            if (e.vrf == 0.) return drift_op(e.l, &e, e.l, 0.);
            st_op op = new_op(K_RFCAV);
            put(e.l / 2.);
            put((e.l / 2.) / gp.beta);
            put(std::fabs(e.vrf * gp.Q));
            put(-2.0 * PI * e.frf);
            put(e.phase0);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_CRABRF: {
            /* CRABRF_Pass cpp:6014; CRABRF::sPass is a drift (cpp:10808-10812) */
            if (e.vrf == 0. || spin) return drift_op(e.l, &e, e.l, 0.);
            st_op op = new_op(K_CRABRF);
            put(e.l / 2.);
            put((e.l / 2.) / gp.beta);
            put(e.vrf * gp.Q);
            put(2.0 * PI * e.frf);
            put(e.phase0);
            put(2.0 * PI * e.frf / 2.99792458e8);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_CCMULT: {
            /* CCMULT_Pass cpp:6027 (B1..A4 of the element feed B2..A5, cpp:10943-10947) */
            if (e.vrf == 0.) return drift_op(e.l, &e, e.l, 0.);
            st_op op = new_op(K_CCMULT);
            double c = 2.99792458e8;
            put(e.l / 2.);
            put((e.l / 2.) / gp.beta);
            put(2.0 * PI * e.frf);
            put(e.phase0);
            put(2 * PI * e.frf / c / gp.brho);
            put(gp.brho);
            for (int i = 0; i < 8; i++) put(e.cc[i]);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_LBT: case ST_ILBT: {
            /* LBT_Pass cpp:6069 / ILBT_Pass cpp:6095; sPass without misalignment */
            st_op op = new_op(e.type == ST_LBT ? K_LBT : K_ILBT);
            double sn = st_sin(e.theta), cs = st_cos(e.theta);
            put(sn);
            put(cs);
            if (e.type == ST_LBT) {
                put(sn / cs);
                put(st_cos(e.theta));
            }
            if (!spin) misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_BEAMBEAM: {
            /* BEAMBEAM::Pass cpp:11191 -> BEAMBEAM_Pass cpp:7343 */
            st_op op = new_op(K_NOP);
            if (e.np != 0.) {
                if (e.treatment == 6) {
                    if (e.nslice < 2 || e.nslice > 64) {
                        err = "BEAMBEAM NSLICE must be in [2, 64]";
                        return -3;
                    }
                    op.kind = K_BB6D;
                    op.norder = e.nslice;
                    double Nps = 1.0 * e.np / e.nslice;
                    put(2 * Nps * 1.53469826769e-18 / gp.gamma);
                    put(2. * Nps * 1.534698e-18 / gp.gamma);
                    put(e.bbscale);
                    put((1.0 + e.alfax * e.alfax) / e.betax);
                    put((1.0 + e.alfay * e.alfay) / e.betay);
                    put(e.betax);
                    put(e.alfax);
                    put(e.betay);
                    put(e.alfay);
                    put(e.emitx);
                    put(e.emity);
                    put(2.0 * e.emitx);
                    put(2.0 * e.emity);
                    double z[64];
                    slice_centers(e.nslice, e.sigmal, z);
                    for (int i = 0; i < e.nslice; i++) put(z[i]);
                } else {
                    op.kind = K_BB4D;
                    put(2 * e.np * 1.53469826769e-18 / gp.gamma);
                    put(std::sqrt(e.emitx * e.betax));
                    put(std::sqrt(e.emity * e.betay));
                    put(e.bbscale);
                }
            }
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_CRBB: {
            /* CRBB::Pass cpp:11442 -> BB6D_Angle cpp:6917 */
            if (e.nslice < 2 || e.nslice > 64) {
                err = "CRBB NSLICE must be in [2, 64]";
                return -3;
            }
            st_op op = new_op(K_BB6DA);
            op.norder = e.nslice;
            double sn = st_sin(e.theta), cs = st_cos(e.theta), tn = sn / cs;
            double Nps = 1.0 * e.np / e.nslice;
            put(2 * Nps * 1.53469826769e-18 / gp.gamma);
            put(2. * Nps * 1.534698e-18 / gp.gamma);
            put(e.bbscale);
            put(sn);
            put(cs);
            put(tn);
            put(st_cos(e.theta));
            double gx = (1.0 + e.alfax * e.alfax) / e.betax;
            double gy = (1.0 + e.alfay * e.alfay) / e.betay;
            put(gx / cs);
            put(gy / cs);
            put(e.betax * cs);
            put(e.alfax);
            put(e.betay * cs);
            put(e.alfay);
            put(e.emitx / cs);
            put(e.emity / cs);
            double z[64];
            slice_centers(e.nslice, e.sigmal, z);
            for (int i = 0; i < e.nslice; i++) put(z[i] / cs);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_PATCH: {
            /* PATCH::Pass / sPass = PATCH_Pass (cpp:12323-12333, 6248-6256):
               TRANS_Pass by DX, DY, DS, then YROTAT_Pass by THETA (skipped
               for |THETA| < 1e-9, cpp:6195); no GtoL/LtoG */
            st_op op = new_op(K_PATCH);
            put(e.dx);
            put(e.dy);
            put(e.ds);
            put(std::fabs(e.theta) < 1.0e-9 ? 0. : 1.);
            put(st_cos(e.theta));
            put(st_sin(e.theta));
            put(st_tan(e.theta));
            t.ops.push_back(op);
            return 0;
        }
        case ST_SOLEN: {
            /* SOLEN::Pass / sPass (cpp:9688-9704): GtoL, SOLEN_Pass (cpp:5804) or
               SOLEN_sPass (cpp:8118), LtoG; with KS = 0 both are DRIFT_Pass.
               The spin pass takes Bz = KS*brho/L where the orbit pass takes
               KS*brho (the reference's two bodies differ so) */
            if (e.ks == 0.) return drift_op(e.l, &e, e.l, 0.);
            if (e.nint < 1) {
                err = "SOLEN Nint must be at least 1";
                return -3;
            }
            st_op op = new_op(K_SOLEN);
            op.nint = e.nint;
            double Bz = spin ? e.ks * gp.brho / e.l : e.ks * gp.brho;
            double Lint = e.l / e.nint;
            put(e.l);
            put(Lint);
            put(gp.q * Bz / 2.0);
            put(gp.m0);
            put(gp.p);
            put(e.l / gp.beta / 2.99792458e8);
            put(Bz * Lint / gp.brho);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_TRANS: {
            /* TRANS::Pass / sPass = TRANS_Pass (cpp:12108-12118, 6140-6151):
               PATCH_Pass without its rotation; no GtoL/LtoG */
            st_op op = new_op(K_PATCH);
            put(e.dx);
            put(e.dy);
            put(e.ds);
            put(0.);
            put(1.);
            put(0.);
            put(0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_YROTAT: {
            /* YROTAT::Pass / sPass = YROTAT_Pass (cpp:12202-12212, 6192-6218):
               PATCH_Pass's rotation; the zero translation is exact (x - 0 = x,
               also for -0 and NaN) */
            st_op op = new_op(K_PATCH);
            put(0.);
            put(0.);
            put(0.);
            put(std::fabs(e.theta) < 1.0e-9 ? 0. : 1.);
            put(st_cos(e.theta));
            put(st_sin(e.theta));
            put(st_tan(e.theta));
            t.ops.push_back(op);
            return 0;
        }
        case ST_SROTAT: {
            /* SROTAT::Pass / sPass = SROTAT_Pass (cpp:12155-12165, 6153-6170),
               a no-op for PSI == 0 */
            st_op op = new_op(e.psi == 0. ? K_NOP : K_SROTAT);
            if (e.psi != 0.) {
                put(st_cos(e.psi));
                put(st_sin(e.psi));
            }
            t.ops.push_back(op);
            return 0;
        }
        case ST_XROTAT: {
            /* XROTAT::Pass / sPass = XROTAT_Pass (cpp:12249-12259, 6220-6245),
               a no-op for |PHI| < 1e-9 */
            st_op op = new_op(std::fabs(e.phi) < 1.0e-9 ? K_NOP : K_XROTAT);
            if (op.kind == K_XROTAT) {
                put(st_cos(e.phi));
                put(st_sin(e.phi));
                put(st_tan(e.phi));
            }
            t.ops.push_back(op);
            return 0;
        }
        case ST_COOLING: {
            /* COOLING::Pass / sPass = COOLING_Pass (cpp:12043-12053, 6132-6136) */
            st_op op = new_op(K_COOLING);
            put(1. - e.alpha);
            t.ops.push_back(op);
            return 0;
        }
        case ST_LRBB: {
            /* LRBB::Pass / sPass (cpp:11299-11313): GtoL, LRBB_Pass (cpp:7358-7363:
               the 6-argument BB4D, cpp:6384-6447, at the separated position), LtoG */
            st_op op = new_op(K_LRBB);
            put(e.sepx);
            put(e.sepy);
            put(2 * e.np * 1.53469826769e-18 / gp.gamma);
            put(e.sigmax);
            put(e.sigmay);
            put(e.bbscale);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_ELENS: {
            /* ELENS::Pass / sPass (cpp:11649-11675): GtoL, ELENS_Pass (cpp:7370-7390),
               LtoG; NE == 0 is DRIFT_Pass(x, Le) */
            if (e.nslice < 1) {
                err = "ELENS NSLICE must be at least 1";
                return -3;
            }
            if (e.ne == 0.) return drift_op(e.l, &e, e.l, 0.);
            st_op op = new_op(K_ELENS);
            op.nint = e.nslice;
            double Le_slice = e.l * 1.0 / e.nslice, Ne_slice = e.ne * 1.0 / e.nslice;
            put(Le_slice / 2.);
            put((Le_slice / 2.) / gp.beta);
            put(2 * Ne_slice * 1.53469826769e-18 / gp.gamma);
            put(e.sigmax);
            put(e.sigmay);
            put(1. + e.betae);
            put(e.bbscale);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_HELENS: {
            /* HELENS::Pass / sPass (cpp:11765-11781): GtoL, HELENS_Pass
               (cpp:7392-7431), LtoG; speed_light = 299792458. (cpp:24) */
            if (e.nslice < 1) {
                err = "HELENS NSLICE must be at least 1";
                return -3;
            }
            st_op op = new_op(K_HELENS);
            op.nint = e.nslice;
            double Le_slice = e.l * 1.0 / e.nslice;
            double I_e = 1.6021765e-19 * e.ne * e.betae * 299792458. / e.l;
            double Theta_max = -2e-7 * Le_slice * I_e * (1 + e.betae * gp.beta) / e.betae / gp.beta / e.rout / gp.brho;
            put(Le_slice / 2.);
            put((Le_slice / 2.) / gp.beta);
            put(Theta_max);
            put(e.rin);
            put(e.rout);
            put(e.rin * e.rin);
            put(e.rout * e.rout - e.rin * e.rin);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_HACDIP: case ST_VACDIP: {
            /* HACDIP/VACDIP::Pass (cpp:10357, 10453): GtoL, ACDIP_Pass (cpp:5930-5950)
               with the other plane's maximum 0, LtoG; sPass: ACDIP_sPass (cpp:8251-8284) */
            st_op op = new_op(K_ACDIP);
            put(e.l);
            put(e.l / gp.beta);
            put(e.l / 2.);
            put((e.l / 2.) / gp.beta);
            put(e.type == ST_HACDIP ? e.hkick : 0.);
            put(e.type == ST_VACDIP ? e.vkick : 0.);
            put(2.0 * 3.14159265 * e.nud);
            put(e.phid);
            put(e.turns);
            put(e.turne);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_HACMULT: case ST_VACMULT: {
            /* HACMULT/VACMULT::Pass (cpp:10173, 10260): GtoL, ACMULT_Pass
               (cpp:5895-5928) with KL = KLMAX, KSL = 0 or KL = 0, KSL = KSLMAX, LtoG;
               sPass: ACMULT_sPass (cpp:8210-8249) */
            st_op op = new_op(K_ACMULT);
            op.norder = e.norder;
            put(e.l / 2.0);
            put((e.l / 2.0) / gp.beta);
            put(e.type == ST_HACMULT ? e.klmax : 0.);
            put(e.type == ST_VACMULT ? e.klmax : 0.);
            put(2.0 * PI);
            put(e.tturns);
            put(e.phi0);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_WIGGLER: {
            /* WIGGLER::Pass (cpp:9755): GtoL, Wiggler_Pass_Forest (cpp:4400), LtoG;
               sPass (cpp:9767): Wiggler_Pass_Wu (cpp:4516), no spin rotation.
               Coefficients as get_Axy_wiggler (cpp:4371), get_Ax_Wu/get_Ay_Wu
               (cpp:4456-4494) and get_wiggler_B (cpp:4497) group them */
            if (e.nint < 1) {
                err = "WIGGLER Nint must be at least 1";
                return -3;
            }
            st_op op = new_op(K_WIGGLER);
            op.nint = e.nint;
            const double b0 = e.b0, kx = e.kx, kz = e.kz, ky = std::sqrt(kx * kx + kz * kz);
            put(e.l / e.nint);
            put(b0 / kz);
            put(b0 * kx / ky / kz);
            put(-(b0 * kx / kz));
            put(b0 * kx * kx / ky / kz);
            put(b0 * ky / kz);
            put(b0 * kx / kz);
            put(b0);
            put(b0 * kx / ky);
            put(kx);
            put(ky);
            put(kz);
            put(e.phiz0);
            put(gp.brho);
            put(kx == 0. ? b0 * ky / kz : b0 * ky / kx / kz);
            put(b0 * kx * kx / ky / ky / kz);
            put((kx / ky) * b0);
            put(-(kz * b0 / ky));
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_MATRIX: {
            st_op op = new_op(K_MATRIX);
            for (int i = 0; i < 6; i++) put(e.xco_in[i]);
            for (int i = 0; i < 36; i++) put(e.m66[i]);
            for (int i = 0; i < 6; i++) put(e.xco_out[i]);
            t.ops.push_back(op);
            return 0;
        }
        default:
            return -2;
        }
    }
};

} // namespace

/* ------------------------------------------------------------- table */

static int upload(dev_table &t)
{
    size_t no = t.ops.size() ? t.ops.size() : 1, np = t.par.size() ? t.par.size() : 1;
    HIPCHK(hipMalloc(&t.d_ops, no * sizeof(st_op)));
    HIPCHK(hipMalloc(&t.d_par, np * sizeof(double)));
    if (!t.ops.empty()) HIPCHK(hipMemcpy(t.d_ops, t.ops.data(), t.ops.size() * sizeof(st_op), hipMemcpyHostToDevice));
    if (!t.par.empty()) HIPCHK(hipMemcpy(t.d_par, t.par.data(), t.par.size() * sizeof(double), hipMemcpyHostToDevice));
    return ST_OK;
}

extern "C" int st_table_destroy(st_table *t)
{
    if (!t) return ST_OK;
    if (t->ctx) hipSetDevice(t->ctx->device);
    hipFree(t->orbit.d_ops);
    hipFree(t->orbit.d_par);
    hipFree(t->spin.d_ops);
    hipFree(t->spin.d_par);
    hipFree(t->d_ap);
    hipFree(t->d_kg);
    delete t;
    return ST_OK;
}

static int validate(const st_element &e, int j)
{
    switch (e.type) {
    case ST_MULT: case ST_GMULT: case ST_SBENDMULT: case ST_SMULT: case ST_GSBENDMULT:
    case ST_HACMULT: case ST_VACMULT:
        if (e.norder < 0 || e.norder > 10)
            return fail(ST_ERR_PARAM, "element %d: Norder %d outside [0, 10] (KNL has 11 entries)", j, e.norder);
        break;
    default:
        break;
    }
    /* a PATCH's or TRANS's L is its DS, which may be negative (cpp:12267,
       12062); no map uses it */
    if (e.l < 0. && e.type != ST_PATCH && e.type != ST_TRANS)
        return fail(ST_ERR_PARAM, "element %d: negative length", j);
    return ST_OK;
}

extern "C" int st_table_build(st_ctx *ctx, const st_globals *gp, const st_element *elems, int n, st_table **out)
{
    if (!ctx || !gp || !out || (n > 0 && !elems) || n < 0) return fail(ST_ERR_ARG, "st_table_build: bad argument");
    *out = nullptr;
    HIPCHK(hipSetDevice(ctx->device));
    init_yoshida();
    st_table *t = new st_table();
    t->ctx = ctx;
    t->gp = *gp;
    t->n = n;
    st_kglobals &k = t->kg;
    std::memset(&k, 0, sizeof k);
    k.beta = gp->beta;
    k.gamma = gp->gamma;
    k.sg = std::sqrt(gp->gamma * gp->gamma - 1.);
    k.G = gp->G;
    k.energy = gp->energy;
    k.brho = gp->brho;
    k.bg2 = gp->beta * gp->gamma * gp->beta * gp->gamma;
    k.e3 = st_powi(gp->energy / 1000., 3);
    k.Cr = 8.84627632682e-5;
    k.Cf = (55. / 24. / std::sqrt(3.)) * 2.8179403e-15 * 1.054572e-34 / 9.1093822e-31;
    k.rng_seed = gp->rng_seed;
    k.H_expand = gp->H_expand;
    k.radiate = gp->radiate;
    std::vector<double> ap(2 * (size_t)(n > 0 ? n : 1));
    for (int pass = 0; pass < 2; ++pass) {
        Builder b(*gp, pass == 1, pass == 1 ? t->spin : t->orbit);
        for (int j = 0; j < n; ++j) {
            int rc = validate(elems[j], j);
            if (rc) {
                st_table_destroy(t);
                return rc;
            }
            rc = b.lower(elems[j]);
            if (rc) {
                std::string msg = b.err.empty() ? "unsupported element type or parameters" : b.err;
                st_table_destroy(t);
                return fail(rc == -2 ? ST_ERR_TYPE : ST_ERR_PARAM, "element %d (type %d): %s", j, elems[j].type,
                            msg.c_str());
            }
        }
    }
    for (int j = 0; j < n; ++j) {
        ap[2 * j] = elems[j].apx;
        ap[2 * j + 1] = elems[j].apy;
        if (gp->radiate && (elems[j].type == ST_SBEND || elems[j].type == ST_GSBEND) && elems[j].angle != 0.)
            t->draws_per_turn += elems[j].nint > 0 ? elems[j].nint : 0;
    }
    int rc = upload(t->orbit);
    if (!rc) rc = upload(t->spin);
    if (!rc) {
        hipError_t e = hipMalloc(&t->d_ap, ap.size() * sizeof(double));
        if (e == hipSuccess) e = hipMemcpy(t->d_ap, ap.data(), ap.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&t->d_kg, sizeof(st_kglobals));
        if (e == hipSuccess) e = hipMemcpy(t->d_kg, &t->kg, sizeof(st_kglobals), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(ST_ERR_DEVICE, "table upload: %s", hipGetErrorString(e));
    }
    if (rc) {
        st_table_destroy(t);
        return rc;
    }
    *out = t;
    return ST_OK;
}

extern "C" int st_table_size(const st_table *t, int *n_elems, int64_t *bytes)
{
    if (!t) return fail(ST_ERR_ARG, "st_table_size: null table");
    if (n_elems) *n_elems = t->n;
    if (bytes)
        *bytes = (int64_t)((t->orbit.ops.size() + t->spin.ops.size()) * sizeof(st_op) +
                           (t->orbit.par.size() + t->spin.par.size()) * sizeof(double) + 2 * t->n * sizeof(double));
    return ST_OK;
}
