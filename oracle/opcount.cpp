/*
 * opcount.cpp -- op-counting build of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (loaded by tests/test_flops.py): it measures the
 * algorithmic FP64 work of each element map, the figure the roofline's
 * F_turn (simtrack_amd/flops.py) models by hand.
 *
 * st_oracle.c is compiled here as C++ with every `double` of its maps
 * replaced by Tracked, a double carrying two flags:
 *   dep    the value depends on the particle's coordinates;
 *   ptonly it is one of the reference's delta1 = sqrt(1 + 2 pt/beta + pt^2)
 *          - 1, gamma1, beta1 (marked by ORC_PT_CACHED in st_oracle.c),
 *          which the GPU kernels keep per particle and recompute only after
 *          a map that changes pt, or a function of them and constants alone.
 * An add, subtract, multiply, divide or square root whose result depends on
 * the particle counts as one operation, in `ops` or, when pt-only, in
 * `ptops`; each elementary-function call on such a value (sin, cos, tan,
 * asin, exp, pow) and each random draw counts one in the same way.
 * Operations on element or reference constants count nothing: the table
 * builder evaluates them once.  Comparisons and sign manipulations are free.
 * The element records (st_element) and the globals keep real doubles: they
 * are included before the substitution.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <cmath>
#include <cstddef>

#include "simtrack_amd.h"

#define ST_FN static inline
#define ST_CONST_TABLE static const
#define ST_FMA(a, b, c) fma((a), (b), (c))
#define ST_RINT(x) nearbyint(x)
#define ST_SQRT(x) sqrt(x)
#include "st_libm.h"
#include "st_rng.h"

#include <vector>

/* Every counted operation is a node of the evaluation graph; a node counts
   only if an output of the map (the particle's coordinates and spin)
   depends on it, so work the reference does and then discards (e.g. the
   midpoint of all nine coordinates where spin() reads two) is left out. */
enum { OPC_OP = 0, OPC_PT = 1, OPC_FREE = 2 };
struct opc_node {
    int a, b;        /* operand nodes, -1 for inputs and constants */
    int cat;         /* OPC_OP, OPC_PT (pt-derived cache) or OPC_FREE */
};
static std::vector<opc_node> g_nodes;
static int g_cat_override = -1;   /* >= 0 inside an uncounted scope */

struct Tracked {
    double v;
    bool dep, ptonly;
    bool input = false;  /* the particle's own input coordinate, not yet operated on */
    int node = -1;
    Tracked() : v(0.), dep(false), ptonly(true) {}
    Tracked(double d) : v(d), dep(false), ptonly(true) {}
    Tracked(double d, bool dp, bool pt) : v(d), dep(dp), ptonly(pt) {}
    operator double() const { return v; }
};

static inline Tracked counted(double v, const Tracked &a, const Tracked &b)
{
    Tracked r(v, a.dep || b.dep, a.ptonly && b.ptonly);
    if (r.dep) {
        r.node = (int)g_nodes.size();
        g_nodes.push_back({a.node, b.node, g_cat_override >= 0 ? g_cat_override : (r.ptonly ? OPC_PT : OPC_OP)});
    }
    return r;
}
static inline Tracked counted1(double v, const Tracked &a) { return counted(v, a, Tracked(0.)); }

#define ST_BINOP(op)                                                                                    \
    static inline Tracked operator op(const Tracked &a, const Tracked &b) { return counted(a.v op b.v, a, b); } \
    static inline Tracked operator op(const Tracked &a, double b) { return counted(a.v op b, a, Tracked(b)); }   \
    static inline Tracked operator op(double a, const Tracked &b) { return counted(a op b.v, Tracked(a), b); }   \
    static inline Tracked operator op(const Tracked &a, int b) { return a op (double)b; }                       \
    static inline Tracked operator op(int a, const Tracked &b) { return (double)a op b; }
ST_BINOP(+)
ST_BINOP(-)
ST_BINOP(*)
ST_BINOP(/)
#undef ST_BINOP

#define ST_CMP(op)                                                                                      \
    static inline bool operator op(const Tracked &a, const Tracked &b) { return a.v op b.v; }            \
    static inline bool operator op(const Tracked &a, double b) { return a.v op b; }                      \
    static inline bool operator op(double a, const Tracked &b) { return a op b.v; }                      \
    static inline bool operator op(const Tracked &a, int b) { return a.v op b; }                         \
    static inline bool operator op(int a, const Tracked &b) { return a op b.v; }
ST_CMP(<)
ST_CMP(>)
ST_CMP(<=)
ST_CMP(>=)
ST_CMP(==)
ST_CMP(!=)
#undef ST_CMP

static inline Tracked operator-(const Tracked &a) { return Tracked(-a.v, a.dep, a.ptonly); }  /* a new value */
static inline Tracked &operator+=(Tracked &a, const Tracked &b) { return a = a + b; }
static inline Tracked &operator-=(Tracked &a, const Tracked &b) { return a = a - b; }
static inline Tracked &operator*=(Tracked &a, const Tracked &b) { return a = a * b; }
static inline Tracked &operator/=(Tracked &a, const Tracked &b) { return a = a / b; }

/* functions: arithmetic ones count, sign manipulations and tests do not */
static inline Tracked sqrt(const Tracked &a) { return counted1(std::sqrt(a.v), a); }
static inline Tracked fabs(const Tracked &a) { return Tracked(std::fabs(a.v), a.dep, a.ptonly); }
static inline bool isnan(const Tracked &a) { return std::isnan(a.v); }
static inline Tracked st_sin(const Tracked &a) { return counted1(st_sin(a.v), a); }
static inline Tracked st_cos(const Tracked &a) { return counted1(st_cos(a.v), a); }
static inline Tracked st_tan(const Tracked &a) { return counted1(st_tan(a.v), a); }
static inline Tracked st_asin(const Tracked &a) { return counted1(st_asin(a.v), a); }
static inline Tracked st_exp(const Tracked &a) { return counted1(st_exp(a.v), a); }
static inline Tracked st_powi(const Tracked &a, int n) { return counted1(st_powi(a.v, n), a); }
static inline Tracked pow(const Tracked &a, const Tracked &b) { return counted(std::pow(a.v, b.v), a, b); }
/* a random draw is one operation of the particle (keyed by its index) */
static inline Tracked opc_draw(double v) { return counted1(v, Tracked(v, true, false)); }

/* st_element with Tracked members, field for field (st_oracle.c reads the
   element through this type in this build) */
struct opc_element {
    int32_t type, nint, norder, nslice, treatment, reserved[3];
    Tracked l, angle, e1, e2, k0l, knl[11], knsl[11], dx, dy, ds, dtheta, dphi, dpsi, apx, apy, vrf, frf, phase0,
        hkick, vkick, theta, np, bbscale, sigmal, emitx, betax, alfax, emity, betay, alfay, cc[8], rot_n[3],
        rot_angle, m66[36], xco_in[6], xco_out[6], ks, psi, phi, alpha, sepx, sepy, sigmax, sigmay, ne, betae,
        rout, rin, nud, phid, turns, turne, klmax, tturns, phi0, b0, kx, kz, phiz0;
};
static const int OPC_NDOUBLE = (int)((offsetof(st_element, phiz0) - offsetof(st_element, l)) / sizeof(double)) + 1;
static_assert((offsetof(opc_element, phiz0) - offsetof(opc_element, l)) / sizeof(Tracked) + 1 ==
                  (offsetof(st_element, phiz0) - offsetof(st_element, l)) / sizeof(double) + 1,
              "opc_element must mirror st_element");

/* operations in the scope of a zero-length drift are not counted (the F_turn
   definition leaves L = 0 drifts out: the kernels reduce them to two adds) */
struct opc_uncounted_scope {
    int saved;
    bool on;
    explicit opc_uncounted_scope(bool c) : saved(g_cat_override), on(c)
    {
        if (on) g_cat_override = OPC_FREE;
    }
    ~opc_uncounted_scope()
    {
        if (on) g_cat_override = saved;
    }
};
#define ORC_UNCOUNTED_IF(cond) opc_uncounted_scope opc_scope_(cond)

/* The reference's delta1, gamma1, beta1: functions of pt that the kernels
   keep per particle and recompute only after pt changes.  Their operations
   (and any later operation on them and constants alone) are pt-derived and
   not counted; instead each change of pt that a later map reads, or that
   the map leaves behind, is one cache update (counts[2], costed by the
   caller as the kernels' update).  g_cache_pt holds the pt the cache was
   last computed from. */
static double g_cache_pt;
static int64_t g_updates;
template <class F>
static inline Tracked opc_derived(F f)
{
    size_t first = g_nodes.size();
    Tracked r = f();
    for (size_t i = first; i < g_nodes.size(); i++)
        if (g_nodes[i].cat == OPC_OP) g_nodes[i].cat = OPC_PT;
    r.ptonly = true;
    return r;
}
template <class F>
static inline Tracked opc_cached(const Tracked &pt, F f)
{
    if (memcmp(&pt.v, &g_cache_pt, sizeof(double)) != 0) {
        g_updates += 1;
        g_cache_pt = pt.v;
    }
    return opc_derived(f);
}
#define ORC_PT_CACHED(pt, v) opc_cached((pt), [&]() -> Tracked { return (v); })
#define ORC_PT_DERIVED(v) opc_derived([&]() -> Tracked { return (v); })

#define ORC_VALUE(f) ((f).v)
#define ORC_SCALAR_BUILD 1   /* st_oracle.c: leave out the exports that take plain double pointers */
#define st_element opc_element
#define double Tracked
#define st_rng_normal(s, p, c) opc_draw(st_rng_normal((s), (p), (c)))
#define st_rad2_photons(s, p, c, n) opc_draw(st_rad2_photons((s), (p), (c), (n)))
#include "st_oracle.c"
#undef double
#undef st_element

/*
 * Counts of one element map (Element::Pass, or sPass with spin) on one
 * particle state x (6 values, 9 with spin): counts[0] = operations that
 * depend on the particle other than through pt alone, counts[1] = pt-only
 * operations, counts[2] = cache updates (changes of pt that a later
 * operation reads or that the map leaves behind).  Returns the map's rc.
 */
extern "C" __attribute__((visibility("default"))) int opc_element_ops(const st_element *e, const st_globals *gp,
                                                                         const double *x, int spin, int64_t *counts)
{
    orc_set_globals(gp);
    opc_element te;
    memcpy(&te.type, &e->type, 8 * sizeof(int32_t));
    const double *src = &e->l;
    Tracked *dst = &te.l;
    for (int i = 0; i < OPC_NDOUBLE; i++) dst[i] = Tracked(src[i]);
    Tracked t[9];
    for (int k = 0; k < (spin ? 9 : 6); k++) {
        t[k] = Tracked(x[k], true, false);
        t[k].input = true;
    }
    orc_rng rng = {gp->rng_seed, 0u, 0u};
    g_nodes.clear();
    g_cat_override = -1;
    g_cache_pt = x[5];
    g_updates = 0;
    int rc = spin ? elem_spass(&te, t) : elem_pass(&te, t, &rng);
    /* live nodes: those the outputs depend on */
    std::vector<char> live(g_nodes.size(), 0);
    std::vector<int> stack;
    for (int k = 0; k < (spin ? 9 : 6); k++)
        if (t[k].node >= 0) stack.push_back(t[k].node);
    while (!stack.empty()) {
        int n = stack.back();
        stack.pop_back();
        if (n < 0 || live[n]) continue;
        live[n] = 1;
        stack.push_back(g_nodes[n].a);
        stack.push_back(g_nodes[n].b);
    }
    counts[0] = counts[1] = 0;
    for (size_t i = 0; i < g_nodes.size(); i++)
        if (live[i] && g_nodes[i].cat != OPC_FREE) counts[g_nodes[i].cat] += 1;
    if (memcmp(&t[5].v, &g_cache_pt, sizeof(double)) != 0) g_updates += 1;   /* for the next map */
    counts[2] = g_updates;
    return rc;
}
