/*
 * simtrack_amd.hpp -- SimTrack's C++ tracking API on the MI355X engine.
 *
 * Header-only C++17 layer over the C ABI (simtrack_amd.h) that keeps the
 * reference's names and signatures, so that a SimTrack driver (e.g.
 * Example2/DA_cal.cpp) switches engines by changing its include:
 *
 *   GlobalVariables GP, Set_RefPartEnergy, Set_Spin_G   simtrack.h:28-72, simtrack.cpp:87-115
 *   Element and the element classes below              simtrack.h:373-959
 *   Line                                               simtrack.h:1067-1100, simtrack.cpp:12341-12431
 *   Track, Track_spin, Track_Spin, Track_tbt(_spin),   simtrack.h:1360-1390
 *   Track_wrt_OnMomentumPart                           simtrack.cpp:20942
 *   Track_Radial, Track_DA, Track_DA1, Track_DA2, Track_DA3, simtrack.cpp:21389-21790
 *   Track_DA_Radial, Track_DA_Uniform, Track_DA_Radial_Uniform
 *   Read_MADX_Lattice, Read_BMAD_Lattice               simtrack.cpp:13591, 14054
 *   Cal_Orbit_Num, Track_tbt_FMA, Track_tbt_Lyapunov   simtrack.cpp:14947, 21179, 21243
 *   Track_tbt_tune_footprint                           simtrack.cpp:21100
 *   Cal_Spin_Orbit, Cal_Spin_Tune, Set_Initial_Spin    simtrack.cpp:20510, 20574, 20603
 *   Cal_Tunes_Track (both forms), Cal_Chrom_Track      simtrack.cpp:20971, 20992, 21028
 *   BeamBeam4D_SoftGaussian, BeamBeam6D_SoftGaussian   simtrack.cpp:22846, 22874, 23134
 *   (both forms), Zboundary_slices_Gaussian             simtrack.cpp:3157
 *   BeamBeam4D_PIC, BeamBeam6D_PIC                     simtrack.cpp:23286, 23380
 *   BeamBeam4D_PIC_NonOffset                           simtrack.cpp:24226
 *   Cal_OneTurnMap (matrix only), Track_tbt_ActionDiff simtrack.cpp:15122, 21309
 *   Cal_Twiss, Cal_Dispersion, Cal_Chrom, Print_Twiss   simtrack.cpp:15540, 17049, 16964, 19846
 *
 * Element classes: DRIFT, MARKER, BPM, HBPM, VBPM, SBEND, GSBEND, QUAD,
 * SKEWQ, SEXT, OCT, MULT, GMULT, SBENDMULT, SMULT, GSBENDMULT, KICKER,
 * HKICKER, VKICKER, RFCAV, CRABRF, CCMULT, LBT, ILBT, BEAMBEAM, CRBB,
 * MATRIX, ROTAT, SNAKE, PATCH, SOLEN, TRANS, SROTAT, YROTAT, XROTAT, COOLING,
 * LRBB, ELENS, HELENS, HACDIP, VACDIP, HACMULT, VACMULT, WIGGLER, with the reference's
 * constructor arguments, Nint and Norder defaults (simtrack.cpp:8451, 8635,
 * 8943-8950) and SetP/GetP names.
 *
 * Differences, all deliberate:
 *  - errors throw simtrack_amd::Error instead of printing and calling exit();
 *  - the per-particle Track* functions lower the Line to a device table on
 *    first use and reuse it while the lowered records are unchanged;
 *  - every Track* also has a batched overload taking np particles;
 *  - Track_DA* run the whole scan as one batch (st_da_bisect/st_da_uniform)
 *    and write the reference's files; the uniform scans return where the
 *    reference calls exit(0).
 *
 * records_from_line() lowers any Line-like object whose Cell[] elements have
 * TYPE, L, DX..DPSI, APx, APy and GetP() -- this header's Line or the
 * reference's own.
 */
#ifndef SIMTRACK_AMD_HPP
#define SIMTRACK_AMD_HPP

#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "simtrack_amd.h"

namespace simtrack_amd {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(int rc, const char *what)
{
    if (rc != 0) throw Error(std::string(what) + ": " + st_last_error());
}

/* slice lengths, simtrack.cpp:45-47 */
constexpr double BLslice = 0.25, QLslice = 0.05, GLslice = 0.05;

/* ------------------------------------------------------------- globals */

/* GlobalVariables GP (simtrack.h:28-68): the fields the maps read */
struct GlobalVariables : st_globals {
    GlobalVariables() { st_globals_default(this); }
    int harm = 360;             /* GP.harm (simtrack.h:51), read by drivers         */
    double circumference = 0.;  /* GP.circumference, set by the lattice readers     */
};
inline GlobalVariables GP;

inline void Set_RefPartEnergy(double gamma, double E0, double Q, double A)
{
    st_globals_set_energy(&GP, gamma, E0, Q, A);
}

/* Set_Spin_G, simtrack.cpp:110-115 (GP.Gr is not used by the maps) */
inline void Set_Spin_G(double g) { GP.G = g; }

/* ------------------------------------------------------------- elements */

class Element {
public:
    explicit Element(std::string name) : NAME(std::move(name)) {}
    virtual ~Element() = default;

    virtual void SetP(const char *name, double value)
    {
        const Param *p = find(name);
        if (!p) throw Error(TYPE + " does not have a parameter of " + name);
        if (p->is_int)
            *reinterpret_cast<int32_t *>(reinterpret_cast<char *>(&rec) + p->off) = (int32_t)value;
        else
            *reinterpret_cast<double *>(reinterpret_cast<char *>(&rec) + p->off) = value;
        after_setp(name, value);
    }
    virtual double GetP(const char *name)
    {
        const Param *p = find(name);
        if (!p) throw Error(TYPE + " does not have a parameter of " + name);
        if (p->is_int) return *reinterpret_cast<const int32_t *>(reinterpret_cast<const char *>(&rec) + p->off);
        return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(&rec) + p->off);
    }

    /* the C-ABI record of this element */
    st_element record() const
    {
        st_element r = rec;
        r.l = L;
        r.dx = DX;
        r.dy = DY;
        r.ds = DS;
        r.dtheta = DTHETA;
        r.dphi = DPHI;
        r.dpsi = DPSI;
        r.apx = APx;
        r.apy = APy;
        return r;
    }

    std::string NAME, TYPE, GROUP;
    double S = 0., L = 0.;
    double DX = 0., DY = 0., DS = 0.;
    double DTHETA = 0., DPHI = 0., DPSI = 0.;
    double APx = 1., APy = 1.; /* Element::Element, simtrack.cpp:8394-8395 */
    double X[6] = {0., 0., 0., 0., 0., 0.}; /* closed orbit after the element (Cal_Orbit_Num) */
    double M[36] = {};                       /* one-turn matrix, last element only (Cal_OneTurnMap) */
    double Beta1 = 0., Alfa1 = 0., Beta2 = 0., Alfa2 = 0.; /* Twiss after the element (Cal_Twiss) */
    double Mu1 = 0., Mu2 = 0.;                               /* phase advance in turns (Trace_A)  */
    double r = 0., c11 = 0., c12 = 0., c21 = 0., c22 = 0.;   /* coupling (Cal_Twiss)              */
    double Etax = 0., Etaxp = 0., Etay = 0., Etayp = 0.;     /* dispersion (Cal_Dispersion)       */
    double T[36] = {};                                       /* element map (Cal_ElementMap)      */
    double n0[3] = {0., 0., 0.};                             /* periodic spin (Cal_Spin_Orbit)    */
    double A[36] = {};                                       /* normalising matrix (Trace_A)      */

protected:
    struct Param {
        std::string name;
        size_t off;
        bool is_int;
    };
    st_element rec{};
    std::vector<Param> params;

    void init(const char *type, int code, double l)
    {
        TYPE = type;
        rec.type = code;
        if (l < 0.) throw Error(TYPE + " length should be non-negative");
        L = l;
    }
    void bind(const char *name, double &field) { params.push_back({name, off(&field), false}); }
    void bind_int(const char *name, int32_t &field) { params.push_back({name, off(&field), true}); }
    void bind_kn()
    {
        for (int i = 0; i < 11; ++i) {
            bind(("K" + std::to_string(i) + "L").c_str(), rec.knl[i]);
            bind(("K" + std::to_string(i) + "SL").c_str(), rec.knsl[i]);
        }
    }
    virtual void after_setp(const char *, double) {}

    /* Norder at construction: highest order with KNL or KNSL != 0, else 1
       (simtrack.cpp:8943-8950); SetP("KnL") does not refresh it */
    void init_norder()
    {
        rec.norder = 1;
        for (int i = 0; i < 10; ++i)
            if (rec.knl[10 - i] != 0. || rec.knsl[10 - i] != 0.) {
                rec.norder = 10 - i;
                break;
            }
    }

private:
    size_t off(const void *field) const
    {
        return (size_t)(reinterpret_cast<const char *>(field) - reinterpret_cast<const char *>(&rec));
    }
    const Param *find(const char *name) const
    {
        for (const Param &p : params)
            if (p.name == name) return &p;
        return nullptr;
    }
};

#define SIMTRACK_AMD_SIMPLE(CLS, CODE)                                                                              \
    class CLS : public Element {                                                                                    \
    public:                                                                                                         \
        CLS(std::string name, double l) : Element(std::move(name)) { init(#CLS, CODE, l); }                         \
    };
SIMTRACK_AMD_SIMPLE(DRIFT, ST_DRIFT)
SIMTRACK_AMD_SIMPLE(MARKER, ST_MARKER)
SIMTRACK_AMD_SIMPLE(BPM, ST_BPM)
SIMTRACK_AMD_SIMPLE(HBPM, ST_HBPM)
SIMTRACK_AMD_SIMPLE(VBPM, ST_VBPM)
#undef SIMTRACK_AMD_SIMPLE

class SBEND : public Element {
public:
    SBEND(std::string name, double l, double angle, double e1, double e2) : Element(std::move(name))
    {
        init("SBEND", ST_SBEND, l);
        if (l <= 0.) throw Error("SBEND length should be positive");
        rec.angle = angle;
        rec.e1 = e1;
        rec.e2 = e2;
        rec.nint = int(l / BLslice) + 1; /* simtrack.cpp:8451 */
        bind("ANGLE", rec.angle);
        bind("E1", rec.e1);
        bind("E2", rec.e2);
        bind_int("Nint", rec.nint);
    }
};

class GSBEND : public Element {
public:
    GSBEND(std::string name, double l, double angle, double k0l, double e1, double e2) : Element(std::move(name))
    {
        init("GSBEND", ST_GSBEND, l);
        if (l <= 0.) throw Error("GSBEND length should be positive");
        rec.angle = angle;
        rec.k0l = k0l;
        rec.e1 = e1;
        rec.e2 = e2;
        rec.nint = int(l / BLslice) + 1; /* simtrack.cpp:8540 */
        bind("ANGLE", rec.angle);
        bind("K0L", rec.k0l);
        bind("E1", rec.e1);
        bind("E2", rec.e2);
        bind_int("Nint", rec.nint);
    }
};

/* QUAD / SKEWQ / SEXT / OCT: one order, Nint = int(L/QLslice)+1 */
class Multipole1 : public Element {
protected:
    Multipole1(std::string name, const char *type, int code, int order, double l, double kn, double ks,
               const char *kn_name, const char *ks_name)
        : Element(std::move(name))
    {
        init(type, code, l);
        rec.knl[order] = kn;
        rec.knsl[order] = ks;
        rec.nint = int(l / QLslice) + 1;
        rec.norder = order;
        if (kn_name) bind(kn_name, rec.knl[order]);
        bind(ks_name, rec.knsl[order]);
        bind_int("Nint", rec.nint);
    }
};

class QUAD : public Multipole1 {
public:
    QUAD(std::string name, double l, double k1l, double k1sl)
        : Multipole1(std::move(name), "QUAD", ST_QUAD, 1, l, k1l, k1sl, "K1L", "K1SL") {}
};
class SKEWQ : public Multipole1 {
public:
    SKEWQ(std::string name, double l, double k1sl)
        : Multipole1(std::move(name), "SKEWQ", ST_SKEWQ, 1, l, 0., k1sl, nullptr, "K1SL") {}
};
class SEXT : public Multipole1 {
public:
    SEXT(std::string name, double l, double k2l, double k2sl)
        : Multipole1(std::move(name), "SEXT", ST_SEXT, 2, l, k2l, k2sl, "K2L", "K2SL") {}
};
class OCT : public Multipole1 {
public:
    OCT(std::string name, double l, double k3l, double k3sl)
        : Multipole1(std::move(name), "OCT", ST_OCT, 3, l, k3l, k3sl, "K3L", "K3SL") {}
};

class MULT : public Element {
public:
    MULT(std::string name, double l, const double knl[11], const double knsl[11]) : Element(std::move(name))
    {
        init("MULT", ST_MULT, l);
        setup(knl, knsl, QLslice);
    }

protected:
    MULT(std::string name, const char *type, int code, double l) : Element(std::move(name)) { init(type, code, l); }
    void setup(const double knl[11], const double knsl[11], double slice)
    {
        for (int i = 0; i < 11; ++i) {
            rec.knl[i] = knl[i];
            rec.knsl[i] = knsl[i];
        }
        rec.nint = int(L / slice) + 1;
        init_norder();
        bind_kn();
        bind_int("Norder", rec.norder);
        bind_int("Nint", rec.nint);
    }
};

/* GMULT / SBENDMULT / SMULT: bend body with multipoles */
class BendMult : public MULT {
protected:
    BendMult(std::string name, const char *type, int code, double l, double angle, double e1, double e2,
             const double knl[11], const double knsl[11], double slice, const double *k0l = nullptr)
        : MULT(std::move(name), type, code, l)
    {
        rec.angle = angle;
        rec.e1 = e1;
        rec.e2 = e2;
        bind("ANGLE", rec.angle);
        if (k0l) {
            /* GSBENDMULT: "K0L" names the bend's K0L before KNL[0], as in
               its SetP/GetP chain (simtrack.cpp:9519, 9575) */
            rec.k0l = *k0l;
            bind("K0L", rec.k0l);
        }
        bind("E1", rec.e1);
        bind("E2", rec.e2);
        setup(knl, knsl, slice);
    }
};

class GMULT : public BendMult {
public:
    GMULT(std::string name, double l, double angle, double e1, double e2, const double knl[11], const double knsl[11])
        : BendMult(std::move(name), "GMULT", ST_GMULT, l, angle, e1, e2, knl, knsl, BLslice) {}
};
class SBENDMULT : public BendMult {
public:
    SBENDMULT(std::string name, double l, double angle, double e1, double e2, const double knl[11],
              const double knsl[11])
        : BendMult(std::move(name), "SBENDMULT", ST_SBENDMULT, l, angle, e1, e2, knl, knsl, BLslice) {}
};
class SMULT : public BendMult {
public:
    SMULT(std::string name, double l, double angle, double e1, double e2, const double knl[11], const double knsl[11])
        : BendMult(std::move(name), "SMULT", ST_SMULT, l, angle, e1, e2, knl, knsl, GLslice) {}
};
class GSBENDMULT : public BendMult {
public:
    GSBENDMULT(std::string name, double l, double angle, double k0l, double e1, double e2, const double knl[11],
               const double knsl[11])
        : BendMult(std::move(name), "GSBENDMULT", ST_GSBENDMULT, l, angle, e1, e2, knl, knsl, BLslice, &k0l)
    {
    }
};

class KICKER : public Element {
public:
    KICKER(std::string name, double l, double hkick, double vkick) : Element(std::move(name))
    {
        init("KICKER", ST_KICKER, l);
        set(hkick, vkick, true, true);
    }

protected:
    KICKER(std::string name, const char *type, int code, double l) : Element(std::move(name)) { init(type, code, l); }
    void set(double h, double v, bool bh, bool bv)
    {
        rec.hkick = h;
        rec.vkick = v;
        if (bh) bind("HKICK", rec.hkick);
        if (bv) bind("VKICK", rec.vkick);
    }
};
class HKICKER : public KICKER {
public:
    HKICKER(std::string name, double l, double hkick) : KICKER(std::move(name), "HKICKER", ST_HKICKER, l)
    {
        set(hkick, 0., true, false);
    }
};
class VKICKER : public KICKER {
public:
    VKICKER(std::string name, double l, double vkick) : KICKER(std::move(name), "VKICKER", ST_VKICKER, l)
    {
        set(0., vkick, false, true);
    }
};

class RFCAV : public Element {
public:
    RFCAV(std::string name, double l, double vrf, double frf, double phase0) : Element(std::move(name))
    {
        init("RFCAV", ST_RFCAV, l);
        set(vrf, frf, phase0);
    }

protected:
    RFCAV(std::string name, const char *type, int code, double l) : Element(std::move(name)) { init(type, code, l); }
    void set(double vrf, double frf, double phase0)
    {
        rec.vrf = vrf;
        rec.frf = frf;
        rec.phase0 = phase0;
        bind("VRF", rec.vrf);
        bind("FRF", rec.frf);
        bind("PHASE0", rec.phase0);
    }
};
class CRABRF : public RFCAV {
public:
    CRABRF(std::string name, double l, double vrf, double frf, double phase0)
        : RFCAV(std::move(name), "CRABRF", ST_CRABRF, l)
    {
        set(vrf, frf, phase0);
    }
};
/* B1..A4 feed CCMULT_Pass's B2..A5 (simtrack.cpp:10943-10947) */
class CCMULT : public RFCAV {
public:
    CCMULT(std::string name, double l, double vrf, double frf, double phase0, double b1, double a1, double b2,
           double a2, double b3, double a3, double b4, double a4)
        : RFCAV(std::move(name), "CCMULT", ST_CCMULT, l)
    {
        set(vrf, frf, phase0);
        const double v[8] = {b1, a1, b2, a2, b3, a3, b4, a4};
        const char *n[8] = {"B1", "A1", "B2", "A2", "B3", "A3", "B4", "A4"};
        for (int i = 0; i < 8; ++i) {
            rec.cc[i] = v[i];
            bind(n[i], rec.cc[i]);
        }
    }
};

class LBT : public Element {
public:
    LBT(std::string name, double theta) : Element(std::move(name))
    {
        init("LBT", ST_LBT, 0.);
        rec.theta = theta;
        bind("THETA", rec.theta);
    }

protected:
    LBT(std::string name, const char *type, int code, double theta) : Element(std::move(name))
    {
        init(type, code, 0.);
        rec.theta = theta;
        bind("THETA", rec.theta);
    }
};
class ILBT : public LBT {
public:
    ILBT(std::string name, double theta) : LBT(std::move(name), "ILBT", ST_ILBT, theta) {}
};

class BEAMBEAM : public Element {
public:
    BEAMBEAM(std::string name, int treatment, double np, double bbscale, double sigmal, int nslice, double emitx,
             double betax, double alfax, double emity, double betay, double alfay)
        : Element(std::move(name))
    {
        init("BEAMBEAM", ST_BEAMBEAM, 0.);
        if (treatment != 4 && treatment != 6) throw Error("beambeam treatment only can be 4 or 6");
        rec.treatment = treatment;
        bind_int("TREATMENT", rec.treatment);
        set(np, bbscale, sigmal, nslice, emitx, betax, alfax, emity, betay, alfay);
    }

protected:
    BEAMBEAM(std::string name, const char *type, int code) : Element(std::move(name)) { init(type, code, 0.); }
    void set(double np, double bbscale, double sigmal, int nslice, double emitx, double betax, double alfax,
             double emity, double betay, double alfay)
    {
        rec.np = np;
        rec.bbscale = bbscale;
        rec.sigmal = sigmal;
        rec.nslice = nslice;
        rec.emitx = emitx;
        rec.betax = betax;
        rec.alfax = alfax;
        rec.emity = emity;
        rec.betay = betay;
        rec.alfay = alfay;
        bind("NP", rec.np);
        bind("BBSCALE", rec.bbscale);
        bind("SIGMAL", rec.sigmal);
        bind_int("NSLICE", rec.nslice);
        bind("EMITX", rec.emitx);
        bind("EMITY", rec.emity);
        bind("BETAX", rec.betax);
        bind("ALFAX", rec.alfax);
        bind("BETAY", rec.betay);
        bind("ALFAY", rec.alfay);
    }
    void after_setp(const char *name, double value) override
    {
        if (std::strcmp(name, "TREATMENT") == 0 && (int)value != 4 && (int)value != 6)
            throw Error("beambeam treatment only can be 4 or 6");
    }
};
class CRBB : public BEAMBEAM {
public:
    CRBB(std::string name, double np, double theta, double bbscale, double sigmal, int nslice, double emitx,
         double betax, double alfax, double emity, double betay, double alfay)
        : BEAMBEAM(std::move(name), "CRBB", ST_CRBB)
    {
        rec.treatment = 6;
        rec.theta = theta;
        bind("THETA", rec.theta);
        set(np, bbscale, sigmal, nslice, emitx, betax, alfax, emity, betay, alfay);
    }
};

class MATRIX : public Element {
public:
    MATRIX(std::string name, double l, const double xco_in[6], const double xco_out[6], const double m66[36])
        : Element(std::move(name))
    {
        init("MATRIX", ST_MATRIX, l);
        for (int i = 0; i < 6; ++i) {
            rec.xco_in[i] = xco_in[i];
            rec.xco_out[i] = xco_out[i];
        }
        for (int i = 0; i < 36; ++i) rec.m66[i] = m66[i];
    }
};

class ROTAT : public Element {
public:
    ROTAT(std::string name, double l, const double n[3], double angle) : Element(std::move(name))
    {
        init("ROTAT", ST_ROTAT, l);
        set(n, angle);
    }

protected:
    ROTAT(std::string name, const char *type, int code, double l) : Element(std::move(name)) { init(type, code, l); }
    void set(const double n[3], double angle)
    {
        for (int i = 0; i < 3; ++i) rec.rot_n[i] = n[i];
        rec.rot_angle = angle;
        bind("NX", rec.rot_n[0]);
        bind("NY", rec.rot_n[1]);
        bind("NS", rec.rot_n[2]);
        bind("ANGLE", rec.rot_angle);
    }
};
class SNAKE : public ROTAT {
public:
    SNAKE(std::string name, double l, const double n[3], double angle)
        : ROTAT(std::move(name), "SNAKE", ST_SNAKE, l)
    {
        set(n, angle);
    }
};

/* PATCH (simtrack.cpp:12262-12334): TRANS_Pass by DX, DY, DS (held in the
   Element's own DX, DY, DS, which are not misalignments here, L = DS), then
   a frame rotation about y by THETA */
class PATCH : public Element {
public:
    PATCH(std::string name, double dx, double dy, double ds, double theta) : Element(std::move(name))
    {
        TYPE = "PATCH";
        rec.type = ST_PATCH;
        L = ds;
        DX = dx;
        DY = dy;
        DS = ds;
        rec.theta = theta;
        bind("THETA", rec.theta);
    }
    void SetP(const char *name, double value) override
    {
        if (!std::strcmp(name, "DX")) DX = value;
        else if (!std::strcmp(name, "DY")) DY = value;
        else if (!std::strcmp(name, "DS")) DS = value;
        else Element::SetP(name, value);
    }
    double GetP(const char *name) override
    {
        if (!std::strcmp(name, "DX")) return DX;
        if (!std::strcmp(name, "DY")) return DY;
        if (!std::strcmp(name, "DS")) return DS;
        return Element::GetP(name);
    }
};

/* SOLEN (simtrack.cpp:9645-9704), KS = Bz / brho, Nint = 1: a hard-edge
   solenoid (SOLEN_Pass cpp:5804, SOLEN_sPass cpp:8118); KS = 0 is a drift */
class SOLEN : public Element {
public:
    SOLEN(std::string name, double l, double ks) : Element(std::move(name))
    {
        if (!(l > 0.)) throw Error("Solenoid: L must be positive.");
        init("SOLEN", ST_SOLEN, l);
        rec.ks = ks;
        rec.nint = 1;
        bind("KS", rec.ks);
    }
};

/* TRANS (simtrack.cpp:12056-12118): frame translation by the element's own
   DX, DY, DS (not misalignments), L = DS */
class TRANS : public Element {
public:
    TRANS(std::string name, double dx, double dy, double ds) : Element(std::move(name))
    {
        TYPE = "TRANS";
        rec.type = ST_TRANS;
        L = ds;
        DX = dx;
        DY = dy;
        DS = ds;
    }
    void SetP(const char *name, double value) override
    {
        if (!std::strcmp(name, "DX")) DX = value;
        else if (!std::strcmp(name, "DY")) DY = value;
        else if (!std::strcmp(name, "DS")) DS = value;
        else Element::SetP(name, value);
    }
    double GetP(const char *name) override
    {
        if (!std::strcmp(name, "DX")) return DX;
        if (!std::strcmp(name, "DY")) return DY;
        if (!std::strcmp(name, "DS")) return DS;
        return Element::GetP(name);
    }
};

/* SROTAT (cpp:12121), YROTAT (cpp:12168), XROTAT (cpp:12215): frame
   rotations about s, y and x; COOLING (cpp:12009): 1 - ALPHA scaling */
class SROTAT : public Element {
public:
    SROTAT(std::string name, double psi) : Element(std::move(name))
    {
        init("SROTAT", ST_SROTAT, 0.);
        rec.psi = psi;
        bind("PSI", rec.psi);
    }
};
class YROTAT : public Element {
public:
    YROTAT(std::string name, double theta) : Element(std::move(name))
    {
        init("YROTAT", ST_YROTAT, 0.);
        rec.theta = theta;
        bind("THETA", rec.theta);
    }
};
class XROTAT : public Element {
public:
    XROTAT(std::string name, double phi) : Element(std::move(name))
    {
        init("XROTAT", ST_XROTAT, 0.);
        rec.phi = phi;
        bind("PHI", rec.phi);
    }
};
class COOLING : public Element {
public:
    COOLING(std::string name, double alpha) : Element(std::move(name))
    {
        init("COOLING", ST_COOLING, 0.);
        rec.alpha = alpha;
        bind("ALPHA", rec.alpha);
    }
};

/* LRBB (simtrack.cpp:11220-11313): long-range 4-D beam-beam at a separation */
class LRBB : public Element {
public:
    LRBB(std::string name, double np, double bbscale, double sepx, double sepy, double sigmax, double sigmay)
        : Element(std::move(name))
    {
        init("LRBB", ST_LRBB, 0.);
        rec.np = np;
        rec.bbscale = bbscale;
        rec.sepx = sepx;
        rec.sepy = sepy;
        rec.sigmax = sigmax;
        rec.sigmay = sigmay;
        bind("NP", rec.np);
        bind("BBSCALE", rec.bbscale);
        bind("SEPX", rec.sepx);
        bind("SEPY", rec.sepy);
        bind("SIGMAX", rec.sigmax);
        bind("SIGMAY", rec.sigmay);
    }
};

/* ELENS (simtrack.cpp:11562-11675) and HELENS (cpp:11678-11781): Gaussian
   and hollow electron lenses, NSLICE kicks between half drifts */
class ELENS : public Element {
public:
    ELENS(std::string name, double l, double ne, double bbscale, int nslice, double betae, double sigmax,
          double sigmay)
        : Element(std::move(name))
    {
        if (l < 0.) throw Error("Error: check signs of L ELENS ");
        init("ELENS", ST_ELENS, l);
        rec.ne = ne;
        rec.bbscale = bbscale;
        rec.nslice = nslice;
        rec.betae = betae;
        rec.sigmax = sigmax;
        rec.sigmay = sigmay;
        bind("NE", rec.ne);
        bind("BBSCALE", rec.bbscale);
        bind_int("NSLICE", rec.nslice);
        bind("BETAE", rec.betae);
        bind("SIGMAX", rec.sigmax);
        bind("SIGMAY", rec.sigmay);
    }
};
class HELENS : public Element {
public:
    HELENS(std::string name, double l, double ne, double bbscale, int nslice, double betae, double rout, double rin)
        : Element(std::move(name))
    {
        init("HELENS", ST_HELENS, l);
        rec.ne = ne;
        rec.bbscale = bbscale;
        rec.nslice = nslice;
        rec.betae = betae;
        rec.rout = rout;
        rec.rin = rin;
        bind("NE", rec.ne);
        bind("BBSCALE", rec.bbscale);
        bind_int("NSLICE", rec.nslice);
        bind("BETAE", rec.betae);
        bind("ROUT", rec.rout);
        bind("RIN", rec.rin);
    }
};

/* HACDIP (cpp:10278-10371) / VACDIP (cpp:10374-10467): AC dipoles, kick
   KICKMAX sin(2 pi NUD (turn - TURNS) + PHID), ramped from TURNS to TURNE */
class HACDIP : public Element {
public:
    HACDIP(std::string name, double l, double hkickmax, double nud, double phid, int turns, int turne)
        : Element(std::move(name))
    {
        init("HACDIP", ST_HACDIP, l);
        rec.hkick = hkickmax;
        bind("HKICKMAX", rec.hkick);
        set(nud, phid, turns, turne);
    }

protected:
    HACDIP(std::string name, const char *type, int code, double l) : Element(std::move(name)) { init(type, code, l); }
    void set(double nud, double phid, int turns, int turne)
    {
        rec.nud = nud;
        rec.phid = phid;
        rec.turns = turns;
        rec.turne = turne;
        bind("NUD", rec.nud);
        bind("PHID", rec.phid);
        bind("TURNS", rec.turns);
        bind("TURNE", rec.turne);
    }
};
class VACDIP : public HACDIP {
public:
    VACDIP(std::string name, double l, double vkickmax, double nud, double phid, int turns, int turne)
        : HACDIP(std::move(name), "VACDIP", ST_VACDIP, l)
    {
        rec.vkick = vkickmax;
        bind("VKICKMAX", rec.vkick);
        set(nud, phid, turns, turne);
    }
};

/* HACMULT (cpp:10104-10187) / VACMULT (cpp:10191-10274): one multipole
   order modulated by sin(2 pi turn / TTURNS + PHI0) */
class HACMULT : public Element {
public:
    HACMULT(std::string name, double l, int norder, double klmax, int tturns, double phi0) : Element(std::move(name))
    {
        init("HACMULT", ST_HACMULT, l);
        bind("KLMAX", rec.klmax);
        set(norder, klmax, tturns, phi0);
    }

protected:
    HACMULT(std::string name, const char *type, int code, double l) : Element(std::move(name)) { init(type, code, l); }
    void set(int norder, double k, int tturns, double phi0)
    {
        rec.norder = norder;
        rec.klmax = k;
        rec.tturns = tturns;
        rec.phi0 = phi0;
        bind_int("Norder", rec.norder);
        bind("TTURNS", rec.tturns);
        bind("PHI0", rec.phi0);
    }
};
class VACMULT : public HACMULT {
public:
    VACMULT(std::string name, double l, int norder, double kslmax, int tturns, double phi0)
        : HACMULT(std::move(name), "VACMULT", ST_VACMULT, l)
    {
        bind("KSLMAX", rec.klmax);
        set(norder, kslmax, tturns, phi0);
    }
};

/* WIGGLER (simtrack.cpp:9707-9771): B0 [T], KX, KZ, PHIZ0; Pass by Forest's
   implicit step (cpp:4400), sPass by Wu's splitting (cpp:4516), no spin */
class WIGGLER : public Element {
public:
    WIGGLER(std::string name, double l, int nint, double b0, double kx, double kz, double phiz0)
        : Element(std::move(name))
    {
        if (!(l > 0.)) throw Error("Wiggler: L must be non-zero.");
        init("WIGGLER", ST_WIGGLER, l);
        rec.nint = nint;
        rec.b0 = b0;
        rec.kx = kx;
        rec.kz = kz;
        rec.phiz0 = phiz0;
        bind_int("Nint", rec.nint);
        bind("B0", rec.b0);
        bind("KX", rec.kx);
        bind("KZ", rec.kz);
        bind("PHIZ0", rec.phiz0);
    }
};

/* ----------------------------------------------------------------- line */

/* Line (simtrack.h:1067-1100): owns its elements, as the reference's does
   (Delete/Replace/Empty delete them, simtrack.cpp:12381-12431) */
class Line {
public:
    double Tune1 = 0., Tune2 = 0.;                                 /* Trace_A                  */
    double Chromx1 = 0., Chromy1 = 0., Chromx2 = 0., Chromy2 = 0.; /* Cal_Chrom                */
    double SpinTune = 0.;                                          /* Cal_Spin_Tune            */
    Line() = default;
    Line(const Line &) = delete;
    Line &operator=(const Line &) = delete;
    ~Line() { Empty(); }

    void Update()
    {
        double s = 0.;
        for (Element *e : Cell) {
            s = e->L + s;
            e->S = s;
        }
        Ncell = (long)Cell.size();
        Length = s;
        frev0 = (299792458. * GP.beta) / Length;
    }
    void Append(Element *x)
    {
        Cell.push_back(x);
        Update();
    }
    void Delete(int i)
    {
        if (i < 0 || i >= (int)Cell.size()) return;
        delete Cell[i];
        Cell.erase(Cell.begin() + i);
        Update();
    }
    void Insert(int i, Element *x)
    {
        if (i < 0 || i >= (int)Cell.size()) return;
        Cell.insert(Cell.begin() + i, x);
        Update();
    }
    void Replace(int i, Element *x)
    {
        if (i < 0 || i >= (int)Cell.size()) return;
        delete Cell[i];
        Cell[i] = x;
        Update();
    }
    void Empty()
    {
        for (Element *e : Cell) delete e;
        Cell.clear();
        Ncell = 0;
        Length = 0.;
    }

    std::vector<Element *> Cell;
    double Length = 0.;
    long Ncell = 0;
    double frev0 = 0.;
};

/* ------------------------------------------------------- lattice files */

namespace detail {
inline bool read_row(const std::string &line, std::string &name, std::string &type, double *v, int n)
{
    std::istringstream ss(line);
    if (!(ss >> name >> type)) return false;
    for (int i = 0; i < n; ++i)
        if (!(ss >> v[i])) return false;
    return true;
}
} // namespace detail

/* Read_MADX_Lattice (simtrack.cpp:13591-13796): a MAD-X TWISS table with the
   columns NAME KEYWORD S L ANGLE E1 E2 TILT K0L K0SL ... K10L K10SL KS,
   keyword by keyword as the reference maps them (kickers get zero kicks;
   SOLENOID with KS = 0 or L = 0, and unknown keywords, become DRIFTs). */
inline void Read_MADX_Lattice(Line &line, const char *filename)
{
    std::ifstream f(filename);
    if (!f) throw Error(std::string("error in opening the file: ") + filename);
    std::vector<std::string> rows;
    for (std::string r; std::getline(f, r);) rows.push_back(r);
    size_t nskip = rows.size();
    for (size_t i = 0; i < rows.size() && i < 100; ++i)
        if (rows[i].find("KEYWORD") != std::string::npos) {
            nskip = i + 2;
            break;
        }
    if (nskip >= rows.size()) throw Error(std::string(filename) + ": no KEYWORD header");
    for (size_t i = nskip; i < rows.size(); ++i) {
        std::string name, type;
        double v[29];
        if (rows[i].find_first_not_of(" \t\r") == std::string::npos) continue;
        if (!detail::read_row(rows[i], name, type, v, 29)) throw Error(std::string(filename) + ": short row");
        name = name.substr(1, name.size() - 2);
        type = type.substr(1, type.size() - 2);
        double l = v[1], angle = v[2], e1 = v[3], e2 = v[4], tilt = v[5], knl[11], knsl[11], sk = v[28];
        for (int k = 0; k < 11; ++k) {
            knl[k] = v[6 + 2 * k];
            knsl[k] = v[7 + 2 * k];
        }
        Element *e;
        if (type == "DRIFT") e = new DRIFT(name, l);
        else if (type == "SBEND") {
            if (knl[0] != 0. && knl[2] == 0.) e = new GSBEND(name, l, angle, knl[0], e1, e2);
            else if (knl[1] == 0. && knl[2] == 0.) e = new SBEND(name, l, angle, e1, e2);
            else e = new GMULT(name, l, angle, e1, e2, knl, knsl);
            e->DPSI = tilt;
        } else if (type == "RBEND") {
            if (knl[1] == 0. && knl[2] == 0.) e = new SBEND(name, l, angle, angle / 2, angle / 2);
            else e = new GMULT(name, l, angle, e1, e2, knl, knsl);
        } else if (type == "QUADRUPOLE") {
            e = new QUAD(name, l, knl[1], knsl[1]);
            e->DPSI = tilt;
        } else if (type == "SEXTUPOLE") {
            e = new SEXT(name, l, knl[2], knsl[2]);
            e->DPSI = tilt;
        } else if (type == "OCTUPOLE") {
            e = new OCT(name, l, knl[3], knsl[3]);
            e->DPSI = tilt;
        } else if (type == "HKICKER") e = new HKICKER(name, l, 0.);
        else if (type == "VKICKER") e = new VKICKER(name, l, 0.);
        else if (type == "TKICKER") e = new KICKER(name, l, 0., 0.);
        else if (type == "MULTIPOLE") {
            e = new MULT(name, l, knl, knsl);
            e->DPSI = tilt;
        } else if (type == "VMONITOR") e = new VBPM(name, l);
        else if (type == "MONITOR") e = new BPM(name, l);
        else if (type == "MARKER") e = new MARKER(name, l);
        else if (type == "SOLENOID") {
            if (sk != 0. && l != 0.) e = new SOLEN(name, l, sk);
            else e = new DRIFT(name, l);
        } else if (type == "RFCAVITY") e = new RFCAV(name, l, 0., 0., 0.);
        else if (type == "BEAMBEAM") {
            if (l != 0.) throw Error("Beam-Beam element should have zero length!");
            e = new BEAMBEAM(name, 6, 0., 1, 0.4545, 11, 2.5e-06, 2.5e-06, 0.53, 0.0, 0.53, 0.0);
        } else if (type == "ELENS") e = new ELENS(name, l, 0.0, 1, 0.14, 4, .31e-03, .31e-03);
        else e = new DRIFT(name, l);
        line.Cell.push_back(e);
    }
    line.Update();
    GP.circumference = line.Cell.empty() ? 0. : line.Cell.back()->S;
}

/* Read_BMAD_Lattice (simtrack.cpp:14054-14237): a Bmad element table with
   the columns name key s l angle actual_angle e1 e2 tilt x_offset z_offset
   x_pitch k0l k0sl ... k10l k10sl ks hkick vkick.  Only the first line is
   skipped, as in the reference: the column-name and units lines become
   zero-length DRIFTs (the reference gives them the length of an unset
   variable).  Keys are matched with the reference's capitalisation. */
inline void Read_BMAD_Lattice(Line &line, const char *filename)
{
    std::ifstream f(filename);
    if (!f) throw Error(std::string("error in opening the file: ") + filename);
    std::string row;
    std::getline(f, row);
    while (std::getline(f, row)) {
        std::string name, type;
        double v[35];
        if (row.find_first_not_of(" \t\r") == std::string::npos) continue;
        if (!detail::read_row(row, name, type, v, 35)) {
            std::istringstream ss(row);
            ss >> name;
            line.Cell.push_back(new DRIFT(name, 0.));
            continue;
        }
        double l = v[1], angle = v[2], e1 = v[4], e2 = v[5], xoff = v[7], zoff = v[8], xpitch = v[9];
        double knl[11], knsl[11], sk = v[32], hkick = v[33], vkick = v[34];
        for (int k = 0; k < 11; ++k) {
            knl[k] = v[10 + 2 * k];
            knsl[k] = v[11 + 2 * k];
        }
        Element *e;
        if (type == "Drift") e = new DRIFT(name, l);
        else if (type == "Sbend") {
            if (std::fabs(angle) < 1.0e-9) e = new MULT(name, l, knl, knsl);
            else if (std::fabs(angle - knl[0]) > 1.0e-6) e = new GSBEND(name, l, angle, knl[0], e1, e2);
            else {
                knl[0] = 0.;
                if (knl[2] == 0. && knl[1] == 0.) e = new SBEND(name, l, angle, e1, e2);
                else e = new SBENDMULT(name, l, angle, e1, e2, knl, knsl);
            }
        } else if (type == "RBEND") {
            if (knl[1] == 0. && knl[2] == 0.) e = new SBEND(name, l, angle, angle / 2., angle / 2.);
            else e = new SBENDMULT(name, l, angle, angle / 2., angle / 2., knl, knsl);
        } else if (type == "Quadrupole") e = new QUAD(name, l, knl[1], knsl[1]);
        else if (type == "Sextupole") e = new SEXT(name, l, knl[2], knsl[2]);
        else if (type == "OCTUPOLE") e = new OCT(name, l, knl[3], knsl[3]);
        else if (type == "HKICKER") e = new HKICKER(name, l, hkick);
        else if (type == "VKICKER") e = new VKICKER(name, l, vkick);
        else if (type == "Kicker") e = new KICKER(name, l, hkick, vkick);
        else if (type == "Multipole") {
            if (knl[0] == 0) e = new MULT(name, l, knl, knsl);
            else e = new SBEND(name + "_tb", 1.0e-6, knl[0], 0, 0);
        } else if (type == "HMONITOR") e = new HBPM(name, l);
        else if (type == "VMONITOR") e = new VBPM(name, l);
        else if (type == "Monitor") e = new BPM(name, l);
        else if (type == "MARKER") e = new MARKER(name, l);
        else if (type == "Solenoid") {
            if (l != 0.) e = new SOLEN(name, l, sk);
            else e = new DRIFT(name, l);
        } else if (type == "RFcavity") e = new RFCAV(name, l, 0., 0., 0.);
        else if (type == "BEAMBEAM") {
            if (l != 0.) throw Error("Beam-Beam element should have zero length!");
            e = new BEAMBEAM(name, 6, 0., 1, 0.4545, 11, 2.5e-06, 2.5e-06, 0.53, 0.0, 0.53, 0.0);
        } else if (type == "ELENS") e = new ELENS(name, l, 0.0, 1, 0.14, 4, .31e-03, .31e-03);
        else if (type == "Patch") e = new PATCH(name, xoff, 0.0, zoff, -xpitch);
        else e = new DRIFT(name, l);
        line.Cell.push_back(e);
    }
    line.Update();
    GP.circumference = line.Cell.empty() ? 0. : line.Cell.back()->S;
}

/* --------------------------------------------------------- lowering */

/* st_element code of a reference TYPE string, -1 if the engine lacks it */
inline int type_code(const std::string &t)
{
    static const char *names[] = {"", "DRIFT", "SBEND", "GSBEND", "QUAD", "SKEWQ", "SEXT", "OCT", "MULT", "GMULT",
                                  "SBENDMULT", "SMULT", "GSBENDMULT", "KICKER", "HKICKER", "VKICKER", "MARKER",
                                  "BPM", "HBPM", "VBPM", "RFCAV", "CRABRF", "CCMULT", "LBT", "ILBT", "BEAMBEAM",
                                  "CRBB", "MATRIX", "ROTAT", "SNAKE", "PATCH", "SOLEN", "TRANS", "SROTAT",
                                  "YROTAT", "XROTAT", "COOLING", "LRBB", "ELENS", "HELENS", "HACDIP", "VACDIP",
                                  "HACMULT", "VACMULT", "WIGGLER"};
    for (int i = 1; i < (int)(sizeof(names) / sizeof(names[0])); ++i)
        if (t == names[i]) return i;
    return -1;
}

/* The record of one element of any Line-like type, read through the public
   fields and GetP() only (simtrack.h:379-395 and each class's GetP names):
   what a reference-side driver can see.  MATRIX keeps M66/XCO private in
   the reference and has no GetP, so its matrix is read only from this
   header's own MATRIX. */
template <class E>
st_element record_of(E &e)
{
    st_element r{};
    if constexpr (std::is_base_of_v<Element, E>) {
        r = e.record();
        return r;
    } else {
        int code = type_code(e.TYPE);
        if (code < 0) throw Error("element " + e.NAME + ": type " + e.TYPE + " is not supported by the engine");
        r.type = code;
        r.l = e.L;
        r.dx = e.DX;
        r.dy = e.DY;
        r.ds = e.DS;
        r.dtheta = e.DTHETA;
        r.dphi = e.DPHI;
        r.dpsi = e.DPSI;
        r.apx = e.APx;
        r.apy = e.APy;
        auto P = [&](const char *n) { return e.GetP(n); };
        auto kn = [&]() {
            for (int i = 0; i < 11; ++i) {
                r.knl[i] = P(("K" + std::to_string(i) + "L").c_str());
                r.knsl[i] = P(("K" + std::to_string(i) + "SL").c_str());
            }
        };
        switch (code) {
        case ST_SBEND:
        case ST_GSBEND:
            r.angle = P("ANGLE");
            r.e1 = P("E1");
            r.e2 = P("E2");
            r.nint = (int)P("Nint");
            if (code == ST_GSBEND) r.k0l = P("K0L");
            break;
        case ST_QUAD:
            r.knl[1] = P("K1L");
            r.knsl[1] = P("K1SL");
            r.nint = (int)P("Nint");
            r.norder = 1;
            break;
        case ST_SKEWQ:
            r.knsl[1] = P("K1SL");
            r.nint = (int)P("Nint");
            r.norder = 1;
            break;
        case ST_SEXT:
            r.knl[2] = P("K2L");
            r.knsl[2] = P("K2SL");
            r.nint = (int)P("Nint");
            r.norder = 2;
            break;
        case ST_OCT:
            r.knl[3] = P("K3L");
            r.knsl[3] = P("K3SL");
            r.nint = (int)P("Nint");
            r.norder = 3;
            break;
        case ST_MULT:
        case ST_GMULT:
        case ST_SBENDMULT:
        case ST_SMULT:
        case ST_GSBENDMULT:
            kn();
            /* GSBENDMULT's "K0L" is the bend's K0L; its KNL[0] has no name */
            if (code == ST_GSBENDMULT) r.knl[0] = 0.;
            r.norder = (int)P("Norder"); /* the constructor's, not recomputed */
            r.nint = (int)P("Nint");
            if (code != ST_MULT) {
                r.angle = P("ANGLE");
                r.e1 = P("E1");
                r.e2 = P("E2");
            }
            if (code == ST_GSBENDMULT) r.k0l = P("K0L");
            break;
        case ST_KICKER:
            r.hkick = P("HKICK");
            r.vkick = P("VKICK");
            break;
        case ST_HKICKER:
            r.hkick = P("HKICK");
            break;
        case ST_VKICKER:
            r.vkick = P("VKICK");
            break;
        case ST_RFCAV:
        case ST_CRABRF:
        case ST_CCMULT:
            r.vrf = P("VRF");
            r.frf = P("FRF");
            r.phase0 = P("PHASE0");
            if (code == ST_CCMULT) {
                const char *n[8] = {"B1", "A1", "B2", "A2", "B3", "A3", "B4", "A4"};
                for (int i = 0; i < 8; ++i) r.cc[i] = P(n[i]);
            }
            break;
        case ST_LBT:
        case ST_ILBT:
            r.theta = P("THETA");
            break;
        case ST_BEAMBEAM:
        case ST_CRBB:
            r.treatment = code == ST_CRBB ? 6 : (int)P("TREATMENT");
            if (code == ST_CRBB) r.theta = P("THETA");
            r.np = P("NP");
            r.bbscale = P("BBSCALE");
            r.sigmal = P("SIGMAL");
            r.nslice = (int)P("NSLICE");
            r.emitx = P("EMITX");
            r.emity = P("EMITY");
            r.betax = P("BETAX");
            r.alfax = P("ALFAX");
            r.betay = P("BETAY");
            r.alfay = P("ALFAY");
            break;
        case ST_ROTAT:
        case ST_SNAKE:
            r.rot_n[0] = P("NX");
            r.rot_n[1] = P("NY");
            r.rot_n[2] = P("NS");
            r.rot_angle = P("ANGLE");
            break;
        case ST_PATCH:
            r.theta = P("THETA");   /* DX, DY, DS are the public fields read above */
            break;
        case ST_SOLEN:
            r.ks = P("KS");
            r.nint = 1;
            break;
        case ST_TRANS:
            break;                  /* DX, DY, DS are the public fields read above */
        case ST_SROTAT:
            r.psi = P("PSI");
            break;
        case ST_YROTAT:
            r.theta = P("THETA");
            break;
        case ST_XROTAT:
            r.phi = P("PHI");
            break;
        case ST_COOLING:
            r.alpha = P("ALPHA");
            break;
        case ST_LRBB:
            r.np = P("NP");
            r.bbscale = P("BBSCALE");
            r.sepx = P("SEPX");
            r.sepy = P("SEPY");
            r.sigmax = P("SIGMAX");
            r.sigmay = P("SIGMAY");
            break;
        case ST_ELENS:
        case ST_HELENS:
            r.ne = P("NE");
            r.bbscale = P("BBSCALE");
            r.nslice = (int)P("NSLICE");
            r.betae = P("BETAE");
            if (code == ST_ELENS) {
                r.sigmax = P("SIGMAX");
                r.sigmay = P("SIGMAY");
            } else {
                r.rout = P("ROUT");
                r.rin = P("RIN");
            }
            break;
        case ST_HACDIP:
        case ST_VACDIP:
            if (code == ST_HACDIP) r.hkick = P("HKICKMAX");
            else r.vkick = P("VKICKMAX");
            r.nud = P("NUD");
            r.phid = P("PHID");
            r.turns = P("TURNS");
            r.turne = P("TURNE");
            break;
        case ST_WIGGLER:
            r.nint = (int)P("Nint");
            r.b0 = P("B0");
            r.kx = P("KX");
            r.kz = P("KZ");
            r.phiz0 = P("PHIZ0");
            break;
        case ST_HACMULT:
        case ST_VACMULT:
            r.norder = (int)P("Norder");
            r.klmax = P(code == ST_HACMULT ? "KLMAX" : "KSLMAX");
            r.tturns = P("TTURNS");
            r.phi0 = P("PHI0");
            break;
        case ST_MATRIX:
            throw Error("element " + e.NAME + ": the reference's MATRIX exposes no parameters to read");
        default:
            break;
        }
        return r;
    }
}

template <class LineT>
std::vector<st_element> records_from_line(LineT &line)
{
    std::vector<st_element> out;
    out.reserve(line.Cell.size());
    for (auto *e : line.Cell) out.push_back(record_of(*e));
    return out;
}

/* -------------------------------------------------------------- engine */

/* the process's device context on HIP device 0 (one per process, as the
   reference's GP is) */
inline st_ctx *default_context()
{
    struct Holder {
        st_ctx *c = nullptr;
        Holder() { check(st_ctx_create(0, &c), "st_ctx_create"); }
        ~Holder()
        {
            if (c) st_ctx_destroy(c);
        }
    };
    static Holder h;
    return h.c;
}

/* A Line lowered to a device table, for repeated tracking */
class Tracker {
public:
    template <class LineT>
    explicit Tracker(LineT &line, const st_globals &gp = GP, st_ctx *ctx = default_context())
        : ctx_(ctx), recs_(records_from_line(line)), gp_(gp)
    {
        check(st_table_build(ctx_, &gp_, recs_.data(), (int)recs_.size(), &t_), "st_table_build");
    }
    Tracker(const Tracker &) = delete;
    Tracker &operator=(const Tracker &) = delete;
    ~Tracker()
    {
        if (t_) st_table_destroy(t_);
    }
    st_table *table() const { return t_; }
    bool same(const std::vector<st_element> &r, const st_globals &gp) const
    {
        return r.size() == recs_.size() && std::memcmp(r.data(), recs_.data(), r.size() * sizeof(st_element)) == 0 &&
               std::memcmp(&gp, &gp_, sizeof(st_globals)) == 0;
    }

    void Track(double *x, int np, int nturn, int *stable, int *lost_turn, int *lost_post)
    {
        check(st_track(t_, x, np, nturn, stable, lost_turn, lost_post), "Track");
    }
    void Track_spin(double *x, int np, int nturn, int *stable, int *lost_turn, int *lost_post)
    {
        check(st_track_spin(t_, x, np, nturn, stable, lost_turn, lost_post), "Track_spin");
    }
    void Track_Radial(double *x, int np, int nturn, int *stable, int *lost_turn, int *lost_post)
    {
        check(st_track_radial(t_, x, np, nturn, stable, lost_turn, lost_post), "Track_Radial");
    }
    void Track_wrt_OnMomentumPart(double *x, const double *dz0, int np, int nturn, int *stable, int *lost_turn,
                                  int *lost_post)
    {
        check(st_track_dz(t_, dz0, x, np, nturn, stable, lost_turn, lost_post), "Track_wrt_OnMomentumPart");
    }
    void Track_tbt(double *x, int np, int nturn, double *x_tbt, int bpm_index, bool spin, int *stable,
                   int *lost_turn, int *lost_post)
    {
        check(st_track_tbt(t_, x, np, nturn, x_tbt, bpm_index, spin ? 1 : 0, stable, lost_turn, lost_post),
              "Track_tbt");
    }

private:
    st_ctx *ctx_;
    std::vector<st_element> recs_;
    st_globals gp_;
    st_table *t_ = nullptr;
};

/* the table of a Line for the per-call API: rebuilt only when the lowered
   records or GP change */
template <class LineT>
Tracker &tracker_for(LineT &line)
{
    static std::unique_ptr<Tracker> cached;
    static const void *owner = nullptr;
    std::vector<st_element> r = records_from_line(line);
    if (!cached || owner != (const void *)&line || !cached->same(r, GP)) {
        cached.reset();
        cached = std::make_unique<Tracker>(line);
        owner = (const void *)&line;
    }
    return *cached;
}

/* ------------------------------------------------- Track* (reference API) */

inline void Track(Line &l, double x[], int nturn, int &stable, int &lost_turn, int &lost_post)
{
    tracker_for(l).Track(x, 1, nturn, &stable, &lost_turn, &lost_post);
}
inline void Track_spin(Line &l, double x[], int nturn, int &stable, int &lost_turn, int &lost_post)
{
    tracker_for(l).Track_spin(x, 1, nturn, &stable, &lost_turn, &lost_post);
}
inline void Track_Spin(Line &l, double x[], int nturn, int &stable, int &lost_turn, int &lost_post)
{
    Track_spin(l, x, nturn, stable, lost_turn, lost_post);
}
inline void Track_Radial(Line &l, double x[], int nturn, int &stable, int &lost_turn, int &lost_post)
{
    tracker_for(l).Track_Radial(x, 1, nturn, &stable, &lost_turn, &lost_post);
}
/* Track_wrt_OnMomentumPart (simtrack.cpp:20942-20966) */
inline void Track_wrt_OnMomentumPart(Line &l, double x[], double dz0[], int nturn, int &stable, int &lost_turn,
                                     int &lost_post)
{
    tracker_for(l).Track_wrt_OnMomentumPart(x, dz0, 1, nturn, &stable, &lost_turn, &lost_post);
}
inline void Track_tbt(Line &l, double x[], int nturn, double x_tbt[], int &stable, int &lost_turn, int &lost_post)
{
    tracker_for(l).Track_tbt(x, 1, nturn, x_tbt, -1, false, &stable, &lost_turn, &lost_post);
}
inline void Track_tbt(Line &l, double x[], int nturn, double x_tbt[], int &bpm_index, int &stable, int &lost_turn,
                      int &lost_post)
{
    tracker_for(l).Track_tbt(x, 1, nturn, x_tbt, bpm_index, false, &stable, &lost_turn, &lost_post);
}
inline void Track_tbt_spin(Line &l, double x[], int nturn, double x_tbt[], int &stable, int &lost_turn,
                           int &lost_post)
{
    tracker_for(l).Track_tbt(x, 1, nturn, x_tbt, -1, true, &stable, &lost_turn, &lost_post);
}
inline void Track_tbt_spin(Line &l, double x[], int nturn, double x_tbt[], int &bpm_index, int &stable,
                           int &lost_turn, int &lost_post)
{
    tracker_for(l).Track_tbt(x, 1, nturn, x_tbt, bpm_index, true, &stable, &lost_turn, &lost_post);
}

/* batched forms: np particles, x is np x 6 (np x 9 with spin) */
inline void Track(Line &l, double *x, int np, int nturn, int *stable, int *lost_turn, int *lost_post)
{
    tracker_for(l).Track(x, np, nturn, stable, lost_turn, lost_post);
}
inline void Track_spin(Line &l, double *x, int np, int nturn, int *stable, int *lost_turn, int *lost_post)
{
    tracker_for(l).Track_spin(x, np, nturn, stable, lost_turn, lost_post);
}
inline void Track_Radial(Line &l, double *x, int np, int nturn, int *stable, int *lost_turn, int *lost_post)
{
    tracker_for(l).Track_Radial(x, np, nturn, stable, lost_turn, lost_post);
}

/* --------------------------------------------------------- DA drivers */

namespace detail {
inline void da_bisect_file(Line &l, int variant, int nturn, double deltap0, double sigmax0, double sigmay0,
                           const char *filename)
{
    double angles[10], smax;
    int n;
    check(st_da_angles(variant, angles, &n, &smax), "st_da_angles");
    std::vector<double> sig((size_t)n);
    const int flags = (variant == ST_DA_TRACK_DA_RADIAL ? ST_DA_RADIAL : 0) |
                      (variant == ST_DA_TRACK_DA2 || variant == ST_DA_TRACK_DA3 ? ST_DA_SINGLE : 0);
    check(st_da_bisect(tracker_for(l).table(), nturn, deltap0, sigmax0, sigmay0, angles, n, smax, flags, sig.data()),
          "st_da_bisect");
    FILE *f = std::fopen(filename, "w");
    if (!f) throw Error(std::string("cannot open ") + filename);
    for (int i = 0; i < n; ++i) {
        double deg = (variant == ST_DA_TRACK_DA1 || variant == ST_DA_TRACK_DA3) ? 15. * (i + 1) : (90 / 11.) * (i + 1);
        std::fprintf(f, "%f %f \n", deg, sig[i]);
    }
    std::fclose(f);
}

inline void da_uniform_file(Line &l, int nturn, double deltap0, double angle, double sigmax0, double sigmay0,
                            const char *filename, double sigma0, int n, bool radial)
{
    std::vector<double> sig((size_t)n);
    std::vector<int> st((size_t)n);
    int m = 0;
    check(st_da_uniform(tracker_for(l).table(), nturn, deltap0, angle, sigmax0, sigmay0, sigma0, 0.2, n,
                        radial ? 1 : 0, sig.data(), st.data(), &m),
          "st_da_uniform");
    FILE *f = std::fopen(filename, "w");
    if (!f) throw Error(std::string("cannot open ") + filename);
    for (int i = 0; i < m; ++i) std::fprintf(f, "%f %f %f \n", angle, sig[i], 1.0 * st[i]);
    std::fclose(f);
}
} // namespace detail

inline void Track_DA(Line &l, int nturn, double deltap0, double sigmax0, double sigmay0, const char *filename)
{
    detail::da_bisect_file(l, ST_DA_TRACK_DA, nturn, deltap0, sigmax0, sigmay0, filename);
}
inline void Track_DA1(Line &l, int nturn, double deltap0, double sigmax0, double sigmay0, const char *filename)
{
    detail::da_bisect_file(l, ST_DA_TRACK_DA1, nturn, deltap0, sigmax0, sigmay0, filename);
}
inline void Track_DA_Radial(Line &l, int nturn, double deltap0, double sigmax0, double sigmay0, const char *filename)
{
    detail::da_bisect_file(l, ST_DA_TRACK_DA_RADIAL, nturn, deltap0, sigmax0, sigmay0, filename);
}
/* Track_DA2 (simtrack.cpp:21497-21533): one particle per bisection midpoint */
inline void Track_DA2(Line &l, int nturn, double deltap0, double sigmax0, double sigmay0, const char *filename)
{
    detail::da_bisect_file(l, ST_DA_TRACK_DA2, nturn, deltap0, sigmax0, sigmay0, filename);
}
inline void Cal_Orbit_Num(Line &l, double deltap);
/* Track_DA3 (simtrack.cpp:21535-21572): Cal_Orbit_Num(l, 0) (filling every
   element's X, as the reference does), then one particle per midpoint */
inline void Track_DA3(Line &l, int nturn, double deltap0, double sigmax0, double sigmay0, const char *filename)
{
    Cal_Orbit_Num(l, 0);
    detail::da_bisect_file(l, ST_DA_TRACK_DA3, nturn, deltap0, sigmax0, sigmay0, filename);
}
inline void Track_DA_Uniform(Line &l, int nturn, double deltap0, double angle, double sigmax0, double sigmay0,
                             const char *filename)
{
    detail::da_uniform_file(l, nturn, deltap0, angle, sigmax0, sigmay0, filename, 1., 150, false);
}
inline void Track_DA_Radial_Uniform(Line &l, int nturn, double deltap0, double angle, double sigmax0,
                                    double sigmay0, const char *filename)
{
    detail::da_uniform_file(l, nturn, deltap0, angle, sigmax0, sigmay0, filename, 1.0, 150, true);
}

/* ----------------------------------------- closed orbit, TBT diagnostics */

/* the reference's global tune search window, simtrack.cpp:27 */
inline double tune_low = 0.660, tune_high = 0.70;

/* Cal_Orbit_Num (simtrack.cpp:14947-15030): the closed orbit at deltap with
   RF off; fills every element's X with the orbit after it, as the
   reference does (cpp:15015-15018). */
inline void Cal_Orbit_Num(Line &l, double deltap)
{
    std::vector<st_element> rec = records_from_line(l);
    double xc[6], xe[6];
    int it = 0;
    st_ctx *c = default_context();
    check(st_closed_orbit(c, &GP, rec.data(), (int)rec.size(), deltap, xc, xe, &it), "Cal_Orbit_Num");
    for (auto &r : rec)
        if (r.type == ST_RFCAV || r.type == ST_CRABRF || r.type == ST_CCMULT) r.vrf = 0.;
    st_table *t = nullptr;
    check(st_table_build(c, &GP, rec.data(), (int)rec.size(), &t), "st_table_build");
    double x[6];
    for (int k = 0; k < 6; ++k) x[k] = xc[k];
    for (size_t j = 0; j < rec.size(); ++j) {
        int rc = st_element_pass(t, (int)j, x, 1, 0);
        if (rc) {
            st_table_destroy(t);
            check(rc, "Cal_Orbit_Num");
        }
        for (int k = 0; k < 6; ++k) l.Cell[j]->X[k] = x[k];
    }
    st_table_destroy(t);
}

/* Cal_Spin_Orbit (simtrack.cpp:20510-20572): Cal_Orbit_Num (every element's
   X), then the periodic spin by Newton iteration on one-turn sPass maps;
   fills every element's n0 with the normalised spin after it. */
inline void Cal_Spin_Orbit(Line &l, double deltap)
{
    Cal_Orbit_Num(l, deltap);
    std::vector<st_element> rec = records_from_line(l);
    std::vector<double> n0(3 * rec.size());
    double x0[9];
    int it = 0;
    check(st_spin_orbit(default_context(), &GP, rec.data(), (int)rec.size(), deltap, x0, n0.data(), &it),
          "Cal_Spin_Orbit");
    for (size_t j = 0; j < rec.size(); ++j)
        for (int i = 0; i < 3; ++i) l.Cell[j]->n0[i] = n0[3 * j + i];
}

/* Cal_Spin_Tune (simtrack.cpp:20574-20601): Cal_Spin_Orbit, then the spin
   tune of 500 tracked turns (FineTuneFinder on [0.001, 0.499]) into
   l.SpinTune. */
inline void Cal_Spin_Tune(Line &l, double deltap)
{
    Cal_Spin_Orbit(l, deltap);
    std::vector<st_element> rec = records_from_line(l);
    double tune = 0., n0[3];
    check(st_spin_tune(default_context(), &GP, rec.data(), (int)rec.size(), deltap, &tune, n0), "Cal_Spin_Tune");
    l.SpinTune = tune;
}

/* Set_Initial_Spin (simtrack.cpp:20603-20635): x[6..8] becomes the
   normalised projection on n0 of the three unit spins, summed over Nturn
   tracked turns from x's orbit. */
inline void Set_Initial_Spin(Line &l, int Nturn, double n0[3], double x[9])
{
    std::vector<st_element> rec = records_from_line(l);
    check(st_set_initial_spin(default_context(), &GP, rec.data(), (int)rec.size(), Nturn, n0, x),
          "Set_Initial_Spin");
}

/* Cal_Tunes_Track (simtrack.cpp:20971-20990): x[0] and x[2] below 1e-9 are
   raised by 1e-9 (in the caller's array, as the reference does), Track_tbt
   for Nturn turns, and FineTuneFinder of the x and y samples in the global
   window [tune_low, tune_high]; the tunes are left untouched when the
   particle is lost (the reference leaves them unset). */
inline void Cal_Tunes_Track(Line &l, double x[], int Nturn, double &tune1, double &tune2)
{
    int stable = 1, lost_turn = 0, lost_post = 0;
    std::vector<double> x_tbt((size_t)Nturn * 6), xtemp((size_t)Nturn);
    if (x[0] < 1.0e-9) x[0] = x[0] + 1.0e-9;
    if (x[2] < 1.0e-9) x[2] = x[2] + 1.0e-9;
    Track_tbt(l, x, Nturn, x_tbt.data(), stable, lost_turn, lost_post);
    if (stable == 1) {
        st_ctx *c = default_context();
        for (int k = 0; k < Nturn; k++) xtemp[k] = x_tbt[(size_t)k * 6 + 0];
        check(st_fine_tune(c, xtemp.data(), nullptr, Nturn, 1, tune_low, tune_high, &tune1), "FineTuneFinder");
        for (int k = 0; k < Nturn; k++) xtemp[k] = x_tbt[(size_t)k * 6 + 2];
        check(st_fine_tune(c, xtemp.data(), nullptr, Nturn, 1, tune_low, tune_high, &tune2), "FineTuneFinder");
    }
}

/* Cal_Tunes_Track (simtrack.cpp:20992-21026): Cal_Orbit_Num at deltap0,
   then the tunes of 1,024 tracked turns from the closed orbit (raised by
   1e-9 where below it) into l.Tune1, l.Tune2 */
inline void Cal_Tunes_Track(Line &l, double deltap0)
{
    Cal_Orbit_Num(l, deltap0);
    const Element *last = l.Cell[l.Ncell - 1];
    double x[6] = {last->X[0], last->X[1], last->X[2], last->X[3], 0., deltap0};
    double tunex1 = l.Tune1, tuney1 = l.Tune2;
    Cal_Tunes_Track(l, x, 1024, tunex1, tuney1);
    l.Tune1 = tunex1;
    l.Tune2 = tuney1;
}

/* Cal_Chrom_Track (simtrack.cpp:21028-21051): tracked tunes at deltap
   +-0.0003 and 0; first and second order chromaticities by differences */
inline void Cal_Chrom_Track(Line &l)
{
    double deltap;
    double qx0, qy0, qxp, qyp, qxm, qym;
    deltap = 0.0003;
    Cal_Tunes_Track(l, deltap);
    qxp = l.Tune1;
    qyp = l.Tune2;
    deltap = -0.0003;
    Cal_Tunes_Track(l, deltap);
    qxm = l.Tune1;
    qym = l.Tune2;
    deltap = 0.00;
    Cal_Tunes_Track(l, deltap);
    qx0 = l.Tune1;
    qy0 = l.Tune2;
    l.Chromx1 = (qxp - qxm) / 2. / 0.0003;
    l.Chromy1 = (qyp - qym) / 2. / 0.0003;
    l.Chromx2 = (qxp + qxm - 2 * qx0) / 2. / 0.0003 / 0.0003;
    l.Chromy2 = (qyp + qym - 2 * qy0) / 2. / 0.0003 / 0.0003;
}

namespace detail {
/* the reference's fstream keeps scientific/precision(18) after the first
   line it writes, so only that line has the default format (%g) */
inline std::string amp(double v, bool first)
{
    char b[64];
    std::snprintf(b, sizeof b, first ? "%g" : "%.18e", v);
    return b;
}
inline std::string sci25(double v)
{
    char b[64];
    std::snprintf(b, sizeof b, "%25.18e", v);
    return b;
}
} // namespace detail

/* Track_tbt_FMA (simtrack.cpp:21179-21241): Cal_Orbit_Num, then the tunes of
   two 1024-turn windows of every stable particle of the 0-6 sigma grid,
   written to ./FMA-output.dat. */
inline void Track_tbt_FMA(Line &l, double deltap0, double sigmax0, double sigmay0)
{
    Cal_Orbit_Num(l, deltap0);
    const double *xe = l.Cell.back()->X;
    Tracker &tr = tracker_for(l);
    int n = 0;
    check(st_fma_scan(tr.table(), deltap0, sigmax0, sigmay0, xe, tune_low, tune_high, 2048, 0.1, 6., nullptr, 0, &n),
          "Track_tbt_FMA");
    std::vector<double> rows((size_t)(n > 0 ? n : 1) * 8);
    check(st_fma_scan(tr.table(), deltap0, sigmax0, sigmay0, xe, tune_low, tune_high, 2048, 0.1, 6., rows.data(), n,
                      &n),
          "Track_tbt_FMA");
    FILE *f = std::fopen("./FMA-output.dat", "w");
    if (!f) throw Error("cannot open ./FMA-output.dat");
    for (int m = 0; m < n; ++m) {
        const double *r = &rows[(size_t)m * 8];
        std::fprintf(f, "%s %s %s %s %s %s  %s %s\n", detail::amp(r[0], m == 0).c_str(),
                     detail::amp(r[1], m == 0).c_str(), detail::sci25(r[2]).c_str(), detail::sci25(r[3]).c_str(),
                     detail::sci25(r[4]).c_str(), detail::sci25(r[5]).c_str(), detail::sci25(r[6]).c_str(),
                     detail::sci25(r[7]).c_str());
    }
    std::fclose(f);
}

/* Track_tbt_Lyapunov (simtrack.cpp:21243-21303): ./Lmap-output.dat */
inline void Track_tbt_Lyapunov(Line &l, double deltap0, double sigmax0, double sigmay0)
{
    Tracker &tr = tracker_for(l);
    int n = 0;
    check(st_lyapunov_scan(tr.table(), deltap0, sigmax0, sigmay0, 2048, 0.05, 6., nullptr, 0, &n),
          "Track_tbt_Lyapunov");
    std::vector<double> rows((size_t)(n > 0 ? n : 1) * 3);
    check(st_lyapunov_scan(tr.table(), deltap0, sigmax0, sigmay0, 2048, 0.05, 6., rows.data(), n, &n),
          "Track_tbt_Lyapunov");
    FILE *f = std::fopen("./Lmap-output.dat", "w");
    if (!f) throw Error("cannot open ./Lmap-output.dat");
    for (int m = 0; m < n; ++m) {
        const double *r = &rows[(size_t)m * 3];
        std::fprintf(f, "%s  %s%s\n", detail::amp(r[0], m == 0).c_str(), detail::amp(r[1], m == 0).c_str(),
                     detail::sci25(r[2]).c_str());
    }
    std::fclose(f);
}

inline void Cal_Twiss(Line &l, double deltap);

/* Track_tbt_tune_footprint (simtrack.cpp:21100-21176): Cal_Twiss and
   Cal_Orbit_Num at deltap0; on 5 angles x 6 amplitudes (0.1, 1..5 sigma of
   the unnormalised emittances emitx, emity, with the Twiss at the start)
   about the closed orbit, 2,048 turns of Track_tbt (one batch) and
   FineTuneFinder of turns [0, 1024) and [1023, 2047) in [tune_low,
   tune_high]: ./footprint-output.dat */
inline void Track_tbt_tune_footprint(Line &l, double deltap0, double emitx, double emity)
{
    const double PI_ = 3.14159265358979323; /* PI, simtrack.cpp:23 */
    const int Nturn = 2048;
    Cal_Twiss(l, deltap0);
    const Element *last = l.Cell[l.Ncell - 1];
    const double betax0 = last->Beta1, betay0 = last->Beta2, alfax0 = last->Alfa1, alfay0 = last->Alfa2;
    Cal_Orbit_Num(l, deltap0);
    std::vector<double> x(30 * 6);
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 6; j++) {
            double *p = &x[(size_t)(i * 6 + j) * 6];
            double nsigma = j;
            if (nsigma == 0) nsigma = 0.1;
            double nsigmax = nsigma * std::cos((i + 1) * 15 * PI_ / 180);
            double nsigmay = nsigma * std::sin((i + 1) * 15 * PI_ / 180);
            double xn = std::sqrt(emitx) * nsigmax, pxn = 0.;
            double yn = std::sqrt(emity) * nsigmay, pyn = 0.;
            p[0] = std::sqrt(betax0) * xn;
            p[1] = -alfax0 * xn / std::sqrt(betax0) + 1.0 * pxn / std::sqrt(betax0);
            p[2] = std::sqrt(betay0) * yn;
            p[3] = -alfay0 * yn / std::sqrt(betay0) + 1.0 * pyn / std::sqrt(betay0);
            p[4] = 0.;
            p[5] = deltap0;
            for (int k = 0; k < 4; k++) p[k] = p[k] + last->X[k];
            if (p[0] == 0.) p[0] = p[0] + 1.0e-09;
            if (p[2] == 0.) p[2] = p[2] + 1.0e-09;
        }
    std::vector<double> tbt((size_t)30 * Nturn * 6);
    std::vector<int> st(30, 1), lt(30, 0), lp(30, 0);
    tracker_for(l).Track_tbt(x.data(), 30, Nturn, tbt.data(), -1, false, st.data(), lt.data(), lp.data());
    std::vector<double> sig;
    std::vector<int> good;
    for (int q = 0; q < 30; ++q) {
        if (st[q] != 1) continue;
        good.push_back(q);
        const double *b = &tbt[(size_t)q * Nturn * 6];
        for (int w = 0; w < 4; ++w) {
            const int first = (w & 1) ? 1023 : 0, col = (w & 2) ? 2 : 0;
            for (int k = 0; k < 1024; k++) sig.push_back(b[(size_t)(k + first) * 6 + col]);
        }
    }
    std::vector<double> tunes(good.size() * 4);
    if (!good.empty())
        check(st_fine_tune(default_context(), sig.data(), nullptr, 1024, (int)tunes.size(), tune_low, tune_high,
                           tunes.data()),
              "FineTuneFinder");
    FILE *f = std::fopen("./footprint-output.dat", "w");
    if (!f) throw Error("cannot open ./footprint-output.dat");
    for (size_t g = 0; g < good.size(); ++g) {
        const double *t = &tunes[g * 4]; /* x1, x2, y1, y2 */
        std::fprintf(f, "%d %d %s %s %s %s  %s %s\n", (good[g] / 6 + 1) * 15, good[g] % 6, detail::sci25(t[0]).c_str(),
                     detail::sci25(t[1]).c_str(), detail::sci25(t[2]).c_str(), detail::sci25(t[3]).c_str(),
                     detail::sci25(t[0] - t[1]).c_str(), detail::sci25(t[2] - t[3]).c_str());
    }
    std::fclose(f);
}

/* the reference's global random seed of rnd(), simtrack.cpp:51 */
inline double seed = 1234.;

/* Cal_OneTurnMap (simtrack.cpp:15122-15179), its linear part: the 6x6
   one-turn matrix about the last element's X (set by Cal_Orbit_Num) with
   z = 0 and delta = deltap, RFCAV voltages off, by symmetric differences of
   1e-6 of GPU-tracked turns; stored row-major in the last element's M. */
inline void Cal_OneTurnMap(Line &l, double deltap)
{
    if (l.Cell.empty()) throw Error("Cal_OneTurnMap: empty line");
    std::vector<st_element> rec = records_from_line(l);
    Element *last = l.Cell.back();
    double x0[6];
    for (int k = 0; k < 6; ++k) x0[k] = last->X[k];
    x0[4] = 0.;
    x0[5] = deltap;
    check(st_one_turn_matrix(default_context(), &GP, rec.data(), (int)rec.size(), x0, 1e-6, last->M),
          "Cal_OneTurnMap");
}

/* The uncoupled Twiss at the end of the line from the last element's M:
   what Track_tbt_ActionDiff reads of Cal_Twiss (simtrack.cpp:21313-21317)
   for a ring whose planes do not couple. */
inline void Cal_Twiss_End(Line &l)
{
    Element *e = l.Cell.back();
    const double *m = e->M;
    for (int p = 0; p < 2; ++p) {
        const double a = m[(2 * p) * 6 + 2 * p], b = m[(2 * p) * 6 + 2 * p + 1], d = m[(2 * p + 1) * 6 + 2 * p + 1];
        const double cosmu = (a + d) / 2;
        if (!(std::fabs(cosmu) < 1.)) throw Error("Cal_Twiss: one-turn map unstable");
        const double sinmu = std::copysign(std::sqrt(1 - cosmu * cosmu), b);
        (p == 0 ? e->Beta1 : e->Beta2) = b / sinmu;
        (p == 0 ? e->Alfa1 : e->Alfa2) = (a - d) / (2 * sinmu);
    }
}

/* GP.step_deltap (simtrack.h:48): the momentum step of Cal_Chrom */
inline double step_deltap = 0.0001;

/* DeltaToPt, simtrack.cpp:125-129 */
inline double DeltaToPt(double delta)
{
    return (-2.0 / GP.beta + std::sqrt(4.0 / GP.beta / GP.beta - 4.0 * (1.0 - (1 + delta) * (1 + delta)))) / 2.;
}

namespace detail {
/* in-place inverse of a row-major 4x4 by Gauss-Jordan elimination with
   partial pivoting; 0 if singular */
inline int mat_inv4(double *a)
{
    double b[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) b[i][j] = j < 4 ? a[4 * i + j] : (j - 4 == i ? 1. : 0.);
    for (int k = 0; k < 4; ++k) {
        int p = k;
        for (int i = k + 1; i < 4; ++i)
            if (std::fabs(b[i][k]) > std::fabs(b[p][k])) p = i;
        if (b[p][k] == 0.) return 0;
        for (int j = 0; j < 8; ++j) std::swap(b[k][j], b[p][j]);
        const double d = b[k][k];
        for (int j = 0; j < 8; ++j) b[k][j] /= d;
        for (int i = 0; i < 4; ++i)
            if (i != k) {
                const double f = b[i][k];
                for (int j = 0; j < 8; ++j) b[i][j] -= f * b[k][j];
            }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) a[4 * i + j] = b[i][j + 4];
    return 1;
}

/* determinant of a row-major 4x4 (overwritten) by elimination with
   partial pivoting */
inline double det4(double *a)
{
    double d = 1.;
    for (int k = 0; k < 4; ++k) {
        int p = k;
        for (int i = k + 1; i < 4; ++i)
            if (std::fabs(a[4 * i + k]) > std::fabs(a[4 * p + k])) p = i;
        if (a[4 * p + k] == 0.) return 0.;
        if (p != k) {
            for (int j = 0; j < 4; ++j) std::swap(a[4 * k + j], a[4 * p + j]);
            d = -d;
        }
        d *= a[4 * k + k];
        for (int i = k + 1; i < 4; ++i) {
            const double f = a[4 * i + k] / a[4 * k + k];
            for (int j = k; j < 4; ++j) a[4 * i + j] -= f * a[4 * k + j];
        }
    }
    return d;
}

/* the null vector of the complex 4x4 matrix b (rank 3) by Gaussian
   elimination with full pivoting, scaled to unit Euclidean norm */
inline void null_vector4(std::complex<double> b[4][4], std::complex<double> v[4])
{
    int col[4] = {0, 1, 2, 3};
    for (int k = 0; k < 3; ++k) {
        int pi = k, pj = k;
        for (int i = k; i < 4; ++i)
            for (int j = k; j < 4; ++j)
                if (std::abs(b[i][j]) > std::abs(b[pi][pj])) {
                    pi = i;
                    pj = j;
                }
        for (int j = 0; j < 4; ++j) std::swap(b[k][j], b[pi][j]);
        for (int i = 0; i < 4; ++i) std::swap(b[i][k], b[i][pj]);
        std::swap(col[k], col[pj]);
        if (std::abs(b[k][k]) == 0.) throw Error("Cal_Twiss: degenerate one-turn map");
        for (int i = k + 1; i < 4; ++i) {
            std::complex<double> f = b[i][k] / b[k][k];
            for (int j = k; j < 4; ++j) b[i][j] -= f * b[k][j];
        }
    }
    std::complex<double> y[4];
    y[3] = 1.;
    for (int k = 2; k >= 0; --k) {
        std::complex<double> acc = 0.;
        for (int j = k + 1; j < 4; ++j) acc += b[k][j] * y[j];
        y[k] = -acc / b[k][k];
    }
    double nrm = 0.;
    for (int k = 0; k < 4; ++k) nrm += std::norm(y[k]);
    nrm = std::sqrt(nrm);
    for (int k = 0; k < 4; ++k) v[col[k]] = y[k] / nrm;
}

/* Cal_A (simtrack.cpp:15299-15477) on the transverse block of a 6x6
   one-turn matrix (row-major): A[i][j].  The eigenvalues of a symplectic
   4x4 come as lambda, 1/lambda: u = lambda + 1/lambda solves
   u^2 - tr(M) u + (c2 - 2) = 0, c2 = (tr(M)^2 - tr(M^2)) / 2, and a stable
   mode has |u| < 2, lambda ~ u/2 + i sqrt(1 - u^2/4).  The eigenvectors are
   null vectors of M - lambda I.  Then the reference's conventions: mode 1
   is the pair with the larger x, px content, each vector is turned so that
   A[0][1] = A[2][3] = 0, the second vector of a mode is flipped to make
   A[1][1], A[3][3] positive, and each pair is scaled to unit symplectic
   product. */
inline void cal_a_4d(const double *m66, double A[4][4])
{
    typedef std::complex<double> cd;
    double m[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) m[i][j] = m66[6 * i + j];
    double tr = m[0][0] + m[1][1] + m[2][2] + m[3][3], tr2 = 0.;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) tr2 += m[i][j] * m[j][i];
    const double c2 = (tr * tr - tr2) / 2, disc = tr * tr - 4 * (c2 - 2);
    if (!(disc >= 0.)) throw Error("Cal_Twiss: one-turn map unstable");
    const double us[2] = {(tr + std::sqrt(disc)) / 2, (tr - std::sqrt(disc)) / 2};
    /* a one-turn matrix from finite differences is symplectic only to its
       accuracy, so the eigenvalues sit slightly off the unit circle: the
       estimates are refined by Newton's method on the characteristic
       polynomial of M itself (Faddeev-LeVerrier coefficients), as an
       eigen-solver of the given matrix would find them */
    double m2[4][4], m3[4][4], t3 = 0.;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            m2[i][j] = 0.;
            for (int q = 0; q < 4; ++q) m2[i][j] += m[i][q] * m[q][j];
        }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            m3[i][j] = 0.;
            for (int q = 0; q < 4; ++q) m3[i][j] += m2[i][q] * m[q][j];
        }
    for (int i = 0; i < 4; ++i) t3 += m3[i][i];
    double lu[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) lu[4 * i + j] = m[i][j];
    const double k1 = tr, k2 = c2, k3 = (tr * tr * tr - 3 * tr * tr2 + 2 * t3) / 6, k4 = det4(lu);
    cd vec[2][4];
    double content[2];
    for (int q = 0; q < 2; ++q) {
        if (!(std::fabs(us[q]) < 2.)) throw Error("Cal_Twiss: one-turn map unstable");
        cd lam(us[q] / 2, std::sqrt(1 - us[q] * us[q] / 4));
        for (int it = 0; it < 8; ++it) {
            const cd pl = (((lam - k1) * lam + k2) * lam - k3) * lam + k4;
            const cd dp = ((4. * lam - 3. * k1) * lam + 2. * k2) * lam - k3;
            if (std::abs(dp) == 0.) break;
            const cd step = pl / dp;
            lam -= step;
            if (std::abs(step) < 1e-16) break;
        }
        cd b[4][4];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) b[i][j] = cd(m[i][j], 0.) - (i == j ? lam : cd(0., 0.));
        null_vector4(b, vec[q]);
        content[q] = std::fabs(vec[q][0].real()) + std::fabs(vec[q][0].imag()) + std::fabs(vec[q][1].real()) +
                     std::fabs(vec[q][1].imag());
    }
    const int mode[2] = {content[0] >= content[1] ? 0 : 1, content[0] >= content[1] ? 1 : 0};
    double a[4][4];
    for (int p = 0; p < 2; ++p) {
        cd *v = vec[mode[p]];
        const int comp = 2 * p;
        const double th = std::atan2(v[comp].imag(), v[comp].real());
        const cd rot(std::cos(-M_PI / 2 - th), std::sin(-M_PI / 2 - th));
        for (int i = 0; i < 4; ++i) {
            cd w = v[i] * rot;
            a[2 * p][i] = -2 * w.imag();
            a[2 * p + 1][i] = 2 * w.real();
        }
    }
    if (a[1][1] < 0.)
        for (int i = 0; i < 4; ++i) a[1][i] = -a[1][i];
    if (a[3][3] < 0.)
        for (int i = 0; i < 4; ++i) a[3][i] = -a[3][i];
    const double t1 = -a[0][1] * a[1][0] + a[0][0] * a[1][1] - a[0][3] * a[1][2] + a[0][2] * a[1][3];
    const double t2 = -a[2][1] * a[3][0] + a[2][0] * a[3][1] - a[2][3] * a[3][2] + a[2][2] * a[3][3];
    if (!(t1 > 0. && t2 > 0.)) throw Error("Cal_Twiss: normal modes not normalisable");
    for (int i = 0; i < 4; ++i) {
        a[0][i] /= std::sqrt(t1);
        a[1][i] /= std::sqrt(t1);
        a[2][i] /= std::sqrt(t2);
        a[3][i] /= std::sqrt(t2);
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) A[i][j] = a[j][i];
}
} // namespace detail

/* Cal_Twiss (simtrack.cpp:15540-15603): the closed orbit and element maps
   (Cal_Orbit_Num, Cal_ElementMap: every element's X and T from
   st_element_maps), the one-turn matrix (Cal_OneTurnMap, the last
   element's M), Cal_A and Trace_A (cpp:15480-15538), then every element's
   Beta1, Alfa1, Beta2, Alfa2, Mu1, Mu2, r, c11..c22 and the Line's Tune1,
   Tune2.  deltap is dp/p; without H_expand it enters x[5] as pt. */
inline void Cal_Twiss(Line &l, double deltap)
{
    if (l.Cell.empty()) throw Error("Cal_Twiss: empty line");
    const double x5 = GP.H_expand ? deltap : DeltaToPt(deltap);
    std::vector<st_element> rec = records_from_line(l);
    const int n = (int)rec.size();
    st_ctx *c = default_context();
    double xc[6], xe[6];
    int it = 0;
    check(st_closed_orbit(c, &GP, rec.data(), n, x5, xc, xe, &it), "Cal_Twiss");
    std::vector<double> T((size_t)n * 36), X((size_t)n * 6);
    check(st_element_maps(c, &GP, rec.data(), n, xc, x5, 1e-6, T.data(), X.data()), "Cal_Twiss");
    for (int k = 0; k < n; ++k) {
        for (int j = 0; j < 6; ++j) l.Cell[k]->X[j] = X[6 * (size_t)k + j];
        for (int j = 0; j < 36; ++j) l.Cell[k]->T[j] = T[36 * (size_t)k + j];
    }
    Element *last = l.Cell.back();
    double x0[6];
    for (int k = 0; k < 6; ++k) x0[k] = last->X[k];
    x0[4] = 0.;
    x0[5] = x5;
    check(st_one_turn_matrix(c, &GP, rec.data(), n, x0, 1e-6, last->M), "Cal_Twiss");
    double A[4][4], G[4][4];
    detail::cal_a_4d(last->M, A);
    double mu1 = 0., mu2 = 0.;
    for (int k = 0; k < n; ++k) {
        Element *e = l.Cell[k];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                double acc = 0.;
                for (int q = 0; q < 4; ++q) acc += e->T[6 * i + q] * A[q][j];
                G[i][j] = acc;
            }
        const double d1 = std::atan2(G[0][1], G[0][0]), d2 = std::atan2(G[2][3], G[2][2]);
        const double s1 = std::sqrt(G[0][0] * G[0][0] + G[0][1] * G[0][1]);
        const double s2 = std::sqrt(G[2][2] * G[2][2] + G[2][3] * G[2][3]);
        for (int i = 0; i < 4; ++i) {
            A[i][0] = (G[i][0] * G[0][0] + G[i][1] * G[0][1]) / s1;
            A[i][1] = (-G[i][0] * G[0][1] + G[i][1] * G[0][0]) / s1;
            A[i][2] = (G[i][2] * G[2][2] + G[i][3] * G[2][3]) / s2;
            A[i][3] = (-G[i][2] * G[2][3] + G[i][3] * G[2][2]) / s2;
        }
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) e->A[6 * i + j] = (i < 4 && j < 4) ? A[i][j] : 0.;
        mu1 = mu1 + d1 / 2. / M_PI;
        mu2 = mu2 + d2 / 2. / M_PI;
        e->Mu1 = mu1;
        e->Mu2 = mu2;
        e->Beta1 = A[0][0] / A[1][1];
        e->Alfa1 = -A[1][0] / A[1][1];
        e->Beta2 = A[2][2] / A[3][3];
        e->Alfa2 = -A[3][2] / A[3][3];
        const double p = A[2][2] * A[3][3];
        e->r = std::sqrt(A[2][2] * A[3][3] - A[2][3] * A[3][2]);
        const double inv[2][2] = {{A[3][3] / p, -A[2][3] / p}, {-A[3][2] / p, A[2][2] / p}};
        double cm[2][2];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) cm[i][j] = (A[i][2] * inv[0][j] + A[i][3] * inv[1][j]) * e->r;
        e->c11 = cm[0][0];
        e->c12 = cm[0][1];
        e->c21 = cm[1][0];
        e->c22 = cm[1][1];
    }
    l.Tune1 = mu1;
    l.Tune2 = mu2;
}

/* Cal_Dispersion (simtrack.cpp:17049-17105) on the maps of the last
   Cal_Twiss: the periodic dispersion -(M44 - I)^-1 M[0:4][5] carried
   through every element's T, times beta without H_expand. */
inline void Cal_Dispersion(Line &l)
{
    if (l.Cell.empty()) throw Error("Cal_Dispersion: empty line");
    const double *m = l.Cell.back()->M;
    double r44[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r44[4 * i + j] = m[6 * i + j] - (i == j ? 1.0 : 0.);
    if (detail::mat_inv4(r44) == 0) throw Error("Cal_Dispersion: singular matrix");
    double x[4], t[4];
    for (int i = 0; i < 4; ++i) {
        x[i] = 0.;
        for (int j = 0; j < 4; ++j) x[i] += -r44[4 * i + j] * m[6 * j + 5];
    }
    const double scale = GP.H_expand ? 1.0 : GP.beta;
    for (Element *e : l.Cell) {
        for (int i = 0; i < 4; ++i) {
            t[i] = e->T[6 * i + 5];
            for (int j = 0; j < 4; ++j) t[i] += e->T[6 * i + j] * x[j];
        }
        for (int i = 0; i < 4; ++i) x[i] = t[i];
        e->Etax = x[0] * scale;
        e->Etaxp = x[1] * scale;
        e->Etay = x[2] * scale;
        e->Etayp = x[3] * scale;
    }
}

/* Cal_Chrom (simtrack.cpp:16964-16992): chromaticities from the tunes of
   Cal_Twiss at +-step_deltap and 0 (converted to pt first without H_expand,
   as the reference does before Cal_Twiss converts again); the optics of
   deltap = 0 stay in the line. */
inline void Cal_Chrom(Line &l)
{
    const double d = GP.H_expand ? step_deltap : DeltaToPt(step_deltap);
    Cal_Twiss(l, d);
    const double qxp = l.Tune1, qyp = l.Tune2;
    Cal_Twiss(l, -d);
    const double qxm = l.Tune1, qym = l.Tune2;
    Cal_Twiss(l, 0.);
    const double qx0 = l.Tune1, qy0 = l.Tune2;
    l.Chromx1 = (qxp - qxm) / 2. / step_deltap;
    l.Chromy1 = (qyp - qym) / 2. / step_deltap;
    l.Chromx2 = (qxp + qxm - 2 * qx0) / 2. / step_deltap / step_deltap;
    l.Chromy2 = (qyp + qym - 2 * qy0) / 2. / step_deltap / step_deltap;
}

/* Print_Twiss (simtrack.cpp:19846-19862): one line per element, fields of
   width 15 in the stream's default format */
inline void Print_Twiss(Line &l, const char *filename)
{
    std::ofstream f(filename);
    if (!f) throw Error(std::string("cannot open ") + filename);
    const char *h[] = {"NAME", "TYPE", "S", "BETA1", "ALFA1", "MU1", "BETA2", "ALFA2", "MU2",
                       "ETAX", "ETAXP", "ETAY", "ETAYP"};
    for (const char *c : h) f << std::setw(15) << c;
    f << "\n";
    for (const Element *e : l.Cell)
        f << std::setw(15) << e->NAME << std::setw(15) << e->TYPE << std::setw(15) << e->S << std::setw(15) << e->Beta1
          << std::setw(15) << e->Alfa1 << std::setw(15) << e->Mu1 << std::setw(15) << e->Beta2 << std::setw(15)
          << e->Alfa2 << std::setw(15) << e->Mu2 << std::setw(15) << e->Etax << std::setw(15) << e->Etaxp
          << std::setw(15) << e->Etay << std::setw(15) << e->Etayp << "\n";
}

/* Track_tbt_ActionDiff (simtrack.cpp:21309-21383): npart particles at
   nsigma (split by phase) with rnd(seed) betatron phases, nstep blocks of
   nturn turns; per block the surviving count and the mean and rms of the
   2Jx, 2Jy fitted to each survivor's turns, appended to ./Diff-output.dat.
   The Twiss come from Cal_Orbit_Num + Cal_OneTurnMap at deltap = 0. */
inline void Track_tbt_ActionDiff(Line &l, double nsigma, double phase, double emitx0, double emity0, int nturn,
                                 int nstep, int npart)
{
    Cal_Orbit_Num(l, 0.);
    Cal_OneTurnMap(l, 0.);
    Cal_Twiss_End(l);
    const Element *e = l.Cell.back();
    Tracker &tr = tracker_for(l);
    std::vector<double> rows((size_t)(nstep > 0 ? nstep : 1) * 6);
    check(st_action_diff(tr.table(), e->Beta1, e->Alfa1, e->Beta2, e->Alfa2, nsigma, phase, emitx0, emity0, nturn,
                         nstep, npart, &seed, rows.data()),
          "Track_tbt_ActionDiff");
    FILE *f = std::fopen("./Diff-output.dat", "w");
    if (!f) throw Error("cannot open ./Diff-output.dat");
    for (int j = 0; j < nstep; ++j) {
        const double *r = &rows[(size_t)j * 6];
        std::fprintf(f, "%d   %d    %g  %g     %g   %g\n", (int)r[0], (int)r[1], r[2], r[3], r[4], r[5]);
    }
    std::fclose(f);
}

/* ------------------------------------------ strong-strong beam-beam */

/* BeamBeam4D_SoftGaussian (simtrack.cpp:22846-22872): both beams (rows of
   6) kicked in place on the GPU */
inline void BeamBeam4D_SoftGaussian(double Np1, int Npart_1, double PartDist_1[], double Np2, int Npart_2,
                                    double PartDist_2[], double gamma1, double gamma2, double bbscale1,
                                    double bbscale2)
{
    check(st_bb_soft_gaussian_4d_host(default_context(), Np1, Npart_1, PartDist_1, Np2, Npart_2, PartDist_2, gamma1,
                                      gamma2, bbscale1, bbscale2),
          "BeamBeam4D_SoftGaussian");
}

/* BeamBeam6D_SoftGaussian (simtrack.cpp:22874-23132): beams sorted by z in
   place and collided slice by slice; writes ./bunch1_Z.dat and
   ./bunch2_Z.dat (slice number and centre) as the reference does */
inline void BeamBeam6D_SoftGaussian(double Np1, int Mp1, double x1[], double Np2, int Mp2, double x2[],
                                    double gamma1, double gamma2, double bbscale1, double bbscale2, int Nslice1,
                                    int Nslice2)
{
    std::vector<double> Z1((size_t)(Nslice1 > 0 ? Nslice1 : 1)), Z2((size_t)(Nslice2 > 0 ? Nslice2 : 1));
    check(st_bb_soft_gaussian_6d_host(default_context(), Np1, Mp1, x1, Np2, Mp2, x2, gamma1, gamma2, bbscale1,
                                      bbscale2, Nslice1, Nslice2, Z1.data(), Z2.data()),
          "BeamBeam6D_SoftGaussian");
    const char *names[2] = {"./bunch1_Z.dat", "./bunch2_Z.dat"};
    const std::vector<double> *Z[2] = {&Z1, &Z2};
    const int ns[2] = {Nslice1, Nslice2};
    for (int b = 0; b < 2; ++b) {
        FILE *f = std::fopen(names[b], "w");
        if (!f) throw Error(std::string("cannot open ") + names[b]);
        for (int j = 0; j < ns[b]; ++j) std::fprintf(f, "%d  %g\n", j, (*Z[b])[(size_t)j]);
        std::fclose(f);
    }
}

/* BeamBeam6D_SoftGaussian with predefined slice boundaries
   (simtrack.cpp:23134-23284): zb1/zb2 hold Nslice + 1 decreasing values */
inline void BeamBeam6D_SoftGaussian(double Np1, int Mp1, double x1[], double Np2, int Mp2, double x2[],
                                    double gamma1, double gamma2, double bbscale1, double bbscale2, int Nslice1,
                                    int Nslice2, double zb1[], double zb2[])
{
    check(st_bb_soft_gaussian_6d_zb_host(default_context(), Np1, Mp1, x1, Np2, Mp2, x2, gamma1, gamma2, bbscale1,
                                         bbscale2, Nslice1, Nslice2, zb1, zb2),
          "BeamBeam6D_SoftGaussian");
}

/* BeamBeam4D_PIC (simtrack.cpp:23286-23378): particle-in-cell kicks in place */
inline void BeamBeam4D_PIC(double Np1, int Mp1, double xc1[], double Np2, int Mp2, double xc2[], double gamma1,
                           double gamma2, double bbscale1, double bbscale2)
{
    check(st_bb_pic_4d_host(default_context(), Np1, Mp1, xc1, Np2, Mp2, xc2, gamma1, gamma2, bbscale1, bbscale2),
          "BeamBeam4D_PIC");
}

/* BeamBeam4D_PIC_NonOffset (simtrack.cpp:24226-24280): particle-in-cell
   kicks on the given grid centred on the origin, in place.  The grid's Green
   function is computed from hx, hy (Cal_Effective_Green_Function with zero
   offsets, what the reference's caller passes in grn_c_r, grn_c_i); the two
   arrays are not read.  Nx = Ny = 64. */
inline void BeamBeam4D_PIC_NonOffset(double Np1, int Mp1, double xc1[], double Np2, int Mp2, double xc2[],
                                     double gamma1, double gamma2, double bbscale1, double bbscale2, int Nx, int Ny,
                                     double hx, double hy, const double *grn_c_r = nullptr,
                                     const double *grn_c_i = nullptr)
{
    (void)grn_c_r;
    (void)grn_c_i;
    check(st_bb_pic_4d_nonoffset_host(default_context(), Np1, Mp1, xc1, Np2, Mp2, xc2, gamma1, gamma2, bbscale1,
                                      bbscale2, Nx, Ny, hx, hy),
          "BeamBeam4D_PIC_NonOffset");
}

/* BeamBeam6D_PIC (simtrack.cpp:23380-23615): beams sorted by z in place,
   slice pairs by BeamBeam4D_PIC; writes ./bunch1_Z.dat and ./bunch2_Z.dat */
inline void BeamBeam6D_PIC(double Np1, int Mp1, double x1[], double Np2, int Mp2, double x2[], double gamma1,
                           double gamma2, double bbscale1, double bbscale2, int Nslice1, int Nslice2)
{
    std::vector<double> Z1((size_t)(Nslice1 > 0 ? Nslice1 : 1)), Z2((size_t)(Nslice2 > 0 ? Nslice2 : 1));
    check(st_bb_pic_6d_host(default_context(), Np1, Mp1, x1, Np2, Mp2, x2, gamma1, gamma2, bbscale1, bbscale2,
                            Nslice1, Nslice2, Z1.data(), Z2.data()),
          "BeamBeam6D_PIC");
    const char *names[2] = {"./bunch1_Z.dat", "./bunch2_Z.dat"};
    const std::vector<double> *Z[2] = {&Z1, &Z2};
    const int ns[2] = {Nslice1, Nslice2};
    for (int b = 0; b < 2; ++b) {
        FILE *f = std::fopen(names[b], "w");
        if (!f) throw Error(std::string("cannot open ") + names[b]);
        for (int j = 0; j < ns[b]; ++j) std::fprintf(f, "%d  %g\n", j, (*Z[b])[(size_t)j]);
        std::fclose(f);
    }
}

/* Zboundary_slices_Gaussian (simtrack.cpp:3157-3168) */
inline void Zboundary_slices_Gaussian(double zb[], int Nslice, double sigmal)
{
    check(st_zboundary_slices_gaussian(zb, Nslice, sigmal), "Zboundary_slices_Gaussian");
}

} // namespace simtrack_amd

#endif /* SIMTRACK_AMD_HPP */
