/*
 * simtrack_amd.h -- C ABI of the MI355X tracking engine.
 *
 * This is the drop-in boundary for SimTrack's tracking hot path.  The
 * reference has no FFI: its "plugin" interface is the abstract C++ class
 * Element (simtrack.h:373-396, virtual Pass/sPass) and its operator is the
 * free function Track(Line&, double x[], nturn, stable, lost_turn,
 * lost_post) (simtrack.h:1360, simtrack.cpp:20791).  Here a Line is
 * flattened into an array of st_element records (one per Element, same
 * parameter names as the reference's SetP/GetP), uploaded once as a packed
 * device table, and whole particle sets are tracked per call.
 *
 * Conventions (all identical to the reference):
 *   - particle state x[6] = (x, px, y, py, z = -c dt, pt = dE/(p0 c))
 *     (simtrack.cpp:11-21); spin runs use x[9] with (Sx, Sy, Ss) in x[6..8];
 *     arrays are particle-major (AoS), np x 6 or np x 9 doubles.
 *   - loss: before turn 0, |x| > 0.1 or |y| > 0.1 or stable == 0 marks the
 *     particle lost at turn 0, element 0 (simtrack.cpp:20795-20796); after
 *     every element, |x| > APx or |y| > APy or isnan(x) or isnan(y) stops it
 *     with lost_turn = turn, lost_post = element index, x left at the loss
 *     point (simtrack.cpp:20802-20803).  Surviving particles keep their
 *     input lost_turn/lost_post values, as the reference leaves them.
 *   - errors are negative return codes with a message from st_last_error();
 *     nothing here exits the process (the reference calls exit()).
 *
 * Threading: calls on different contexts may run concurrently; calls on one
 * context are serialised by the caller.  All calls block until the GPU work
 * they issue has finished, except the *_device entry points, which are
 * ordered on the context's stream and return after launching.
 *
 * Stream ordering: the context's stream is a non-blocking stream, not
 * ordered after any other.  A caller whose device inputs were written on
 * another stream (a framework's current stream, a copy stream) makes that
 * work finish first (synchronise the stream, or make it wait on an event),
 * and reads device outputs only after st_ctx_sync or a blocking call on the
 * same context.
 */
#ifndef SIMTRACK_AMD_H
#define SIMTRACK_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ST_ABI_VERSION 3

/* ------------------------------------------------------------ status */
enum st_status {
    ST_OK = 0,
    ST_ERR_ARG = -1,          /* bad argument (null pointer, size, index)       */
    ST_ERR_TYPE = -2,         /* unknown or unsupported element type            */
    ST_ERR_PARAM = -3,        /* element parameter out of the supported range   */
    ST_ERR_DEVICE = -4,       /* HIP runtime error                              */
    ST_ERR_NOMEM = -5,        /* host or device allocation failed               */
    ST_ERR_STATE = -6         /* call not valid in the current state            */
};

/* ----------------------------------------------------- element types */
/* Element classes of simtrack.h:397-1059 that the engine tracks. */
enum st_element_type {
    ST_DRIFT = 1,      /* DRIFT      cpp:8399  */
    ST_SBEND = 2,      /* SBEND      cpp:8441  */
    ST_GSBEND = 3,     /* GSBEND     cpp:8529  */
    ST_QUAD = 4,       /* QUAD       cpp:8626  knl[1]=K1L knsl[1]=K1SL */
    ST_SKEWQ = 5,      /* SKEWQ      cpp:8706  knsl[1]=K1SL */
    ST_SEXT = 6,       /* SEXT       cpp:8773  knl[2]=K2L knsl[2]=K2SL */
    ST_OCT = 7,        /* OCT        cpp:8853  knl[3]=K3L knsl[3]=K3SL */
    ST_MULT = 8,       /* MULT       cpp:8933  */
    ST_GMULT = 9,      /* GMULT      cpp:9046  */
    ST_SBENDMULT = 10, /* SBENDMULT  cpp:9192  */
    ST_SMULT = 11,     /* SMULT      cpp:9330  */
    ST_GSBENDMULT = 12,/* GSBENDMULT cpp:9483  */
    ST_KICKER = 13,    /* KICKER     cpp:9912  */
    ST_HKICKER = 14,   /* HKICKER    cpp:9982  */
    ST_VKICKER = 15,   /* VKICKER    cpp:10043 */
    ST_MARKER = 16,    /* MARKER     cpp:10590 */
    ST_BPM = 17,       /* BPM        cpp:10470 */
    ST_HBPM = 18,      /* HBPM       cpp:10510 */
    ST_VBPM = 19,      /* VBPM       cpp:10550 */
    ST_RFCAV = 20,     /* RFCAV      cpp:10630 */
    ST_CRABRF = 21,    /* CRABRF     cpp:10719 */
    ST_CCMULT = 22,    /* CCMULT     cpp:10815 cc[0..7]=B1,A1,B2,A2,B3,A3,B4,A4 */
    ST_LBT = 23,       /* LBT        cpp:10960 */
    ST_ILBT = 24,      /* ILBT       cpp:11009 */
    ST_BEAMBEAM = 25,  /* BEAMBEAM   cpp:11058 */
    ST_CRBB = 26,      /* CRBB       cpp:11316 */
    ST_MATRIX = 27,    /* MATRIX     cpp:9774  */
    ST_ROTAT = 28,     /* ROTAT      cpp:11784 (spin rotator; orbit: drift)   */
    ST_SNAKE = 29,     /* SNAKE      cpp:11864 (spin rotator; orbit: drift)   */
    ST_PATCH = 30,     /* PATCH      cpp:12262 dx,dy,ds = DX,DY,DS, theta = THETA, L = DS */
    ST_SOLEN = 31,     /* SOLEN      cpp:9645  ks = KS, nint = Nint (1)                    */
    ST_TRANS = 32,     /* TRANS      cpp:12056 dx,dy,ds = DX,DY,DS (not misalignments), L = DS */
    ST_SROTAT = 33,    /* SROTAT     cpp:12121 psi = PSI   */
    ST_YROTAT = 34,    /* YROTAT     cpp:12168 theta = THETA */
    ST_XROTAT = 35,    /* XROTAT     cpp:12215 phi = PHI   */
    ST_COOLING = 36,   /* COOLING    cpp:12009 alpha = ALPHA */
    ST_LRBB = 37,      /* LRBB       cpp:11220 np, bbscale, sepx, sepy, sigmax, sigmay */
    ST_ELENS = 38,     /* ELENS      cpp:11562 l, ne, bbscale, nslice, betae, sigmax, sigmay */
    ST_HELENS = 39,    /* HELENS     cpp:11678 l, ne, bbscale, nslice, betae, rout, rin */
    ST_HACDIP = 40,    /* HACDIP     cpp:10278 l, hkick = HKICKMAX, nud, phid, turns, turne */
    ST_VACDIP = 41,    /* VACDIP     cpp:10374 l, vkick = VKICKMAX, nud, phid, turns, turne */
    ST_HACMULT = 42,   /* HACMULT    cpp:10104 l, norder, klmax = KLMAX, tturns, phi0 */
    ST_VACMULT = 43,   /* VACMULT    cpp:10191 l, norder, klmax = KSLMAX, tturns, phi0 */
    ST_WIGGLER = 44    /* WIGGLER    cpp:9707  l, nint, b0, kx, kz, phiz0 */
};

/*
 * One Element.  Field names follow the reference's member and SetP names;
 * unused fields must be zero.  nint/norder carry the values GetP("Nint") and
 * GetP("Norder") return (constructor defaults: Nint = int(L/slice)+1,
 * simtrack.cpp:8451, 8635; Norder = highest non-zero order at construction,
 * cpp:8944-8950, and SetP("KnL") does not refresh it).
 */
typedef struct st_element {
    int32_t type;           /* enum st_element_type                          */
    int32_t nint;           /* Nint                                          */
    int32_t norder;         /* Norder (multipole families, 1..10)            */
    int32_t nslice;         /* NSLICE (BEAMBEAM, CRBB)                       */
    int32_t treatment;      /* TREATMENT (BEAMBEAM: 4 or 6)                  */
    int32_t reserved[3];
    double l;               /* L                                             */
    double angle;           /* ANGLE                                         */
    double e1, e2;          /* E1, E2                                        */
    double k0l;             /* K0L (GSBEND, GSBENDMULT)                      */
    double knl[11];         /* KNL[0..10]  (QUAD K1L in knl[1], ...)         */
    double knsl[11];        /* KNSL[0..10]                                   */
    double dx, dy, ds;      /* DX, DY, DS misalignment                       */
    double dtheta, dphi, dpsi; /* DTHETA, DPHI, DPSI                         */
    double apx, apy;        /* APx, APy (reference default 1.0)              */
    double vrf, frf, phase0;/* RFCAV/CRABRF/CCMULT: VRF [MV], FRF [Hz], PHASE0 */
    double hkick, vkick;    /* KICKER family                                 */
    double theta;           /* LBT/ILBT/CRBB THETA                           */
    double np, bbscale, sigmal;             /* beam-beam NP, BBSCALE, SIGMAL  */
    double emitx, betax, alfax;             /* EMITX, BETAX, ALFAX            */
    double emity, betay, alfay;             /* EMITY, BETAY, ALFAY            */
    double cc[8];           /* CCMULT B1,A1,B2,A2,B3,A3,B4,A4                */
    double rot_n[3];        /* ROTAT/SNAKE rotation axis n[3]                */
    double rot_angle;       /* ROTAT/SNAKE angle                             */
    double m66[36];         /* MATRIX M66 (row major)                        */
    double xco_in[6], xco_out[6]; /* MATRIX XCO_IN, XCO_OUT                  */
    double ks;              /* SOLEN KS                                      */
    /* ABI 3: the frame, cooling, e-lens, long-range beam-beam and AC elements */
    double psi, phi, alpha; /* SROTAT PSI, XROTAT PHI, COOLING ALPHA         */
    double sepx, sepy;      /* LRBB SEPX, SEPY                               */
    double sigmax, sigmay;  /* LRBB, ELENS SIGMAX, SIGMAY                    */
    double ne, betae;       /* ELENS, HELENS NE, BETAE                       */
    double rout, rin;       /* HELENS ROUT, RIN                              */
    double nud, phid;       /* HACDIP/VACDIP NUD, PHID                       */
    double turns, turne;    /* HACDIP/VACDIP TURNS, TURNE (integers)         */
    double klmax;           /* HACMULT KLMAX, VACMULT KSLMAX                 */
    double tturns, phi0;    /* HACMULT/VACMULT TTURNS (integer), PHI0        */
    double b0, kx, kz, phiz0; /* WIGGLER B0, KX, KZ, PHIZ0                   */
} st_element;

/*
 * Reference-particle state: the fields of GlobalVariables GP
 * (simtrack.h:28-68) that the element maps read.  st_globals_set_energy
 * fills them exactly as Set_RefPartEnergy (simtrack.cpp:87-108) does.
 */
typedef struct st_globals {
    double gamma, beta;     /* GP.gamma, GP.beta                             */
    double energy0, energy; /* GP.energy0, GP.energy [MeV]                   */
    double cp, p, m0, m, q; /* GP.cp, GP.p, GP.m0, GP.m, GP.q                */
    double brho;            /* GP.brho                                       */
    double G;               /* GP.G (spin)                                   */
    int32_t Q, A;           /* GP.Q, GP.A (int, as in the reference)         */
    int32_t H_expand;       /* GP.H_expand                                   */
    int32_t radiate;        /* GP.radiate                                    */
    int32_t quad_fringe;    /* GP.quad_fringe                                */
    int32_t reserved;
    /* radiation runs (GP.radiate): seed of the per-particle counter-based
       normal stream that replaces gsl_ran_gaussian(gsl_r2, 1.0) */
    uint64_t rng_seed;
} st_globals;

/* GP defaults (simtrack.h:31-58) followed by Set_RefPartEnergy(gamma, E0, Q, A). */
void st_globals_default(st_globals *gp);
void st_globals_set_energy(st_globals *gp, double gamma, double E0, double Q, double A);

/* --------------------------------------------------- library/context */
typedef struct st_ctx st_ctx;
typedef struct st_table st_table;

int         st_abi_version(void);
/* sizeof(st_element), sizeof(st_globals): lets bindings check their layout */
void        st_struct_sizes(int *element_bytes, int *globals_bytes);
/* Message of the last failing call on this thread ("" if none). */
const char *st_last_error(void);

/* Create a context on HIP device `device` (own stream and buffers). */
int st_ctx_create(int device, st_ctx **out);
int st_ctx_destroy(st_ctx *ctx);

/*
 * Build the packed device table of a Line (replaces walking
 * linename.Cell[j] in Track, simtrack.cpp:20800).  The globals are copied:
 * later changes to *gp do not affect the table.  Element constants that the
 * reference recomputes per call (Fkick*KNL/Nint, slice lengths, cos/sin of
 * bend angles, beam-beam slice centroids, ...) are evaluated here once with
 * the reference's own expressions, so results are unchanged bit for bit.
 */
int st_table_build(st_ctx *ctx, const st_globals *gp, const st_element *elems,
                   int n_elems, st_table **out);
int st_table_destroy(st_table *t);
int st_table_size(const st_table *t, int *n_elems, int64_t *bytes);

/*
 * Track (simtrack.cpp:20791-20814) for np particles at once: x is np x 6
 * (in/out), stable/lost_turn/lost_post are np ints (in/out).  Host arrays.
 */
int st_track(st_table *t, double *x, int np, int nturn,
             int *stable, int *lost_turn, int *lost_post);

/* Track_spin / Track_Spin (simtrack.cpp:20816, 21766): x is np x 9. */
int st_track_spin(st_table *t, double *x, int np, int nturn,
                  int *stable, int *lost_turn, int *lost_post);

/*
 * Track_tbt (simtrack.cpp:20841): also stores x at the start of every turn,
 * x_tbt[(p*nturn + turn)*6 + k] (np x nturn x 6, caller-allocated; rows of
 * turns after a loss are left untouched).  With bpm_index >= 0 the overload
 * of cpp:20885 is followed instead: x is stored after element bpm_index.
 * spin != 0 gives Track_tbt_spin (cpp:20863, 20909) with 9 columns.
 */
int st_track_tbt(st_table *t, double *x, int np, int nturn, double *x_tbt,
                 int bpm_index, int spin,
                 int *stable, int *lost_turn, int *lost_post);

/*
 * Track_Radial (simtrack.cpp:21641) and Example2's Track_Radial_Test
 * (Example2/DA_cal.cpp:19-57): the on-momentum path length z after every
 * element of one pass from x = 0 is subtracted from z after that element.
 */
int st_track_radial(st_table *t, double *x, int np, int nturn,
                    int *stable, int *lost_turn, int *lost_post);

/*
 * Element::Pass / sPass of one element (simtrack.h:380-382) applied to np
 * particles (x np x 6, or np x 9 with spin != 0).  turn is GP.turn.
 */
int st_element_pass(st_table *t, int elem_index, double *x, int np, int spin);

/*
 * Track_wrt_OnMomentumPart (simtrack.cpp:20942-20966): Track with dz0[j]
 * (one value per element, the on-momentum particle's z after element j)
 * subtracted from z after element j, the caller's version of
 * st_track_radial's own x04.
 */
int st_track_dz(st_table *t, const double *dz0, double *x, int np, int nturn, int *stable, int *lost_turn,
                int *lost_post);

/*
 * Device-resident tracking of a turn range, for callers that keep particle
 * state in HBM (device pointers from hipMalloc or a framework allocator):
 * d_x is np x 6 (x 9 with spin), d_state is np ints holding stable, d_lost
 * np x 2 ints (lost_turn, lost_post).  Turns turn0 .. turn0+nturn-1 are
 * tracked; the 0.1 m pre-check is applied only when turn0 == 0, so a run
 * split into ranges equals one call.  flags: ST_TRACK_SPIN, ST_TRACK_RADIAL,
 * ST_TRACK_BARE (no pre-check and no loss test: Element::Pass turn after
 * turn, as Cal_Orbit_Num, simtrack.cpp:14970-14982, passes particles),
 * ST_TRACK_PERF (the perf-mode kernels: the same maps with FMA contraction
 * and division by a reciprocal multiplication, so results differ from the
 * reference's operations by rounding; never the default).
 * Returns after launching on the context stream; st_ctx_sync waits.
 */
#define ST_TRACK_SPIN 1
#define ST_TRACK_RADIAL 2
#define ST_TRACK_BARE 4
#define ST_TRACK_PERF 8
int st_track_device(st_table *t, double *d_x, int np, int turn0, int nturn,
                    int *d_state, int *d_lost, int flags);

/*
 * st_track_device without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as st_track_device, except that the survivor lists
 * and compaction counts live in a fixed per-GPU scratch array held until the
 * process ends (about 8.5 MB), and per-launch timing is not recorded.
 * np <= ST_NOALLOC_MAX_PARTICLES, and with ST_TRACK_RADIAL the table has at
 * most ST_NOALLOC_MAX_ELEMENTS elements; larger inputs return ST_ERR_ARG.
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one
 * GPU (from several contexts or threads) race.
 */
#define ST_NOALLOC_MAX_PARTICLES (1 << 20)
#define ST_NOALLOC_MAX_ELEMENTS 16384
int st_track_device_noalloc(st_table *t, double *d_x, int np, int turn0, int nturn,
                            int *d_state, int *d_lost, int flags);
int st_ctx_sync(st_ctx *ctx);

/* Seconds spent in the last blocking call's GPU work (HIP events). */
int st_last_gpu_seconds(st_ctx *ctx, double *seconds);

/*
 * Per-launch timing of the tracking kernel: when enabled, every tracking
 * launch on this context is bracketed by HIP events on the context stream,
 * and so is every slice pair of a strong-strong collision (its two
 * launches).  st_ctx_kernel_stats returns (after st_ctx_sync) the summed
 * kernel time, the number of launches and the sum over launches of the work
 * (tracking: launched particles x turns; slice pairs: rows of both slices),
 * then clears the counters.
 */
int st_ctx_set_timing(st_ctx *ctx, int enable);
int st_ctx_kernel_stats(st_ctx *ctx, double *kernel_seconds, int64_t *launches, int64_t *launched_particle_turns);

/* ------------------------------------------------- dynamic aperture */

/*
 * Angle sets and bisection ranges of the reference's bisection drivers:
 * ST_DA_TRACK_DA     Track_DA        (simtrack.cpp:21389): 10 angles
 *                    (90/11)(i+1) deg, sigma in [0, 50];
 * ST_DA_TRACK_DA1    Track_DA1       (cpp:21443): 5 angles 15(i+1) deg, [0, 50];
 * ST_DA_TRACK_DA_RADIAL Track_DA_Radial (cpp:21682): 10 angles, [0, 15];
 * ST_DA_TRACK_DA2    Track_DA2       (cpp:21497): 10 angles, [0, 50];
 * ST_DA_TRACK_DA3    Track_DA3       (cpp:21535): 5 angles 15(i+1) deg, [0, 30].
 * angles (radians) needs room for 10.
 */
#define ST_DA_TRACK_DA 0
#define ST_DA_TRACK_DA1 1
#define ST_DA_TRACK_DA_RADIAL 2
#define ST_DA_TRACK_DA2 3
#define ST_DA_TRACK_DA3 4
int st_da_angles(int variant, double *angles, int *nangle, double *sigma_max);

/*
 * The bisection of Track_DA / Track_DA1 / Track_DA_Radial (cpp:21389-21441,
 * 21443-21495, 21682-21734) and Track_DA2 / Track_DA3 (cpp:21497-21572)
 * for nangle angles at once: particles x = (s cos(a) sigmax0, 0,
 * s sin(a) sigmay0, 0, 0, deltap0); a midpoint s is stable when its twins
 * at s and s + 0.05 both survive nturn turns, or with ST_DA_SINGLE in
 * `radial` when the one particle at s does (Track_DA2/DA3); Track, or
 * Track_Radial when `radial` has ST_DA_RADIAL (any odd value).  The
 * interval [0, sigma_max] is halved until it is <= 0.1.  sigma_out[i] is
 * the last midpoint tested -- the number the reference writes as "angle
 * sigma".  Replaces the per-particle Track calls (cpp:21417, 21426, 21522)
 * with one batched launch over the whole bisection tree.
 */
#define ST_DA_RADIAL 1
#define ST_DA_SINGLE 2
int st_da_bisect(st_table *t, int nturn, double deltap0, double sigmax0, double sigmay0,
                 const double *angles, int nangle, double sigma_max, int radial, double *sigma_out);

/*
 * Track_DA_Uniform (cpp:21574; sigma0 = 1, dsigma = 0.2, n = 150),
 * Track_DA_Radial_Uniform (cpp:21736; radial) and
 * Track_DA_Radial_Uniform_Test (Example2/DA_cal.cpp:61; sigma0 = 3, n = 100,
 * radial): one particle at sigma_i = sigma0 + i*dsigma on the given angle.
 * sigma_out/stable_out get all n results; *n_scanned is the number of lines
 * the reference writes (up to and including the first loss, after which it
 * calls exit(0); this library returns instead).
 */
int st_da_uniform(st_table *t, int nturn, double deltap0, double angle, double sigmax0, double sigmay0,
                  double sigma0, double dsigma, int n, int radial, double *sigma_out, int *stable_out,
                  int *n_scanned);

/* ------------------------------------------ many lattices, one launch */

/*
 * A batch of tables (lattice seeds of uneven element counts) tracked by one
 * persistent launch per turn block: work items of up to 64 particles of one
 * table are claimed from a global counter, longest table first; survivors
 * are regrouped between turn blocks.  Results equal st_track on each table
 * separately.  All tables must share the context and the globals.  The
 * reference has no counterpart: its drivers call Track per particle and
 * per seed (simtrack.cpp:20791, Example2/DA_cal.cpp:81).
 */
typedef struct st_batch st_batch;
typedef struct st_batch_config {
    int wg_waves;        /* wavefronts per workgroup: 1, 2 or 4                 */
    int waves_per_simd;  /* resident wavefronts per SIMD of the persistent grid */
    int turn_block;      /* turns between regroupings of the survivors          */
} st_batch_config;

int st_batch_create(st_ctx *ctx, st_table *const *tables, int n_tables, st_batch **out);
int st_batch_destroy(st_batch *b);
int st_batch_default_config(st_batch_config *cfg);
/* particles of table k are x rows offsets[k] .. offsets[k+1]-1 (n_tables+1
   offsets, offsets[0] = 0); flags: ST_TRACK_SPIN; cfg null: defaults */
int st_batch_track(st_batch *b, double *x, const int *offsets, int nturn, int *stable, int *lost_turn,
                   int *lost_post, int flags, const st_batch_config *cfg);
int st_batch_track_device(st_batch *b, double *d_x, const int *offsets, int turn0, int nturn, int *d_state,
                          int *d_lost, int flags, const st_batch_config *cfg);
/* as st_batch_track_device, also storing x at the start of every turn:
   d_tbt[(p*nturn + (turn - turn0))*ncol + k] (rows after a loss untouched) */
int st_batch_track_tbt_device(st_batch *b, double *d_x, const int *offsets, int turn0, int nturn, double *d_tbt,
                              int *d_state, int *d_lost, int flags, const st_batch_config *cfg);
/* times every candidate configuration on calib_turns turns of x (host,
   fresh particles) and returns the fastest; seconds_out (room for 32) and
   n_candidates may be null */
int st_batch_autotune(st_batch *b, const double *x, const int *offsets, int calib_turns, int flags,
                      st_batch_config *best, double *seconds_out, int *n_candidates);

/* ------------------------------------------------ survival surrogate */

/*
 * A multilayer perceptron n_in -> H -> H -> H -> 1 (ReLU hidden layers,
 * sigmoid output) predicting long-term survival from features the tracker
 * generates.  BF16 weights and activations on the matrix cores with FP32
 * accumulation; FP32 master weights, binary cross-entropy loss and Adam
 * (beta1 0.9, beta2 0.999, eps 1e-8).  No counterpart in the reference.
 * Parameters are flat FP32: W0 (H x n_in), W1, W2 (H x H), W3 (1 x H), then
 * b0, b1, b2 (H), b3 (1); weights row-major (out x in).  d_* arguments are
 * device pointers; calls are ordered on the context stream.
 */
typedef struct st_mlp st_mlp;
int st_mlp_create(st_ctx *ctx, int n_in, int n_hidden, uint64_t seed, st_mlp **out);
int st_mlp_destroy(st_mlp *m);
int st_mlp_num_params(const st_mlp *m, int64_t *n);
int st_mlp_get_params(st_mlp *m, float *params, int64_t n);
int st_mlp_set_params(st_mlp *m, const float *params, int64_t n);
/* probabilities of survival for n feature rows (n x n_in FP32) */
int st_mlp_forward(st_mlp *m, const float *d_x, int n, float *d_prob);
/* loss and gradients (host, same layout as the parameters) of a batch with
   labels d_y (1 = survived), without updating */
int st_mlp_gradients(st_mlp *m, const float *d_x, const float *d_y, int n, float *grads, float *loss);
/* one Adam step on a batch; *loss (may be null) gets the batch's mean loss */
int st_mlp_train_step(st_mlp *m, const float *d_x, const float *d_y, int n, float lr, float *loss);

/*
 * The surrogate's 12 FP32 features per particle: Courant-Snyder actions
 * Jx, Jy and angle cos/sin in both planes of the initial coordinates d_x0
 * (n rows of ncol doubles) with the Twiss parameters of the start point,
 * the four tunes d_tunes[n][4] = (nux, nuy of turn window 1, nux, nuy of
 * window 2) and log10 |tune change| per plane.
 */
int st_surrogate_features(st_ctx *ctx, const double *d_x0, int ncol, const double *d_tunes, int n, double beta_x,
                          double alpha_x, double beta_y, double alpha_y, float *d_feat);

/* ------------------------------------------------------ tune finder */

/*
 * FineTuneFinder (simtrack.cpp:2288-2319) for nsig signals of n samples:
 * the peak of sum_i x[i] exp(-2 pi i tune i) over [tune_lower, tune_upper),
 * scanned in steps of 1e-3 and refined by factors of 10 down to 1e-12.
 * With px != null the x/px form (cpp:2321-2352) is used:
 * sum_i (x[i] + i px[i]) exp(-2 pi i tune i).  Host arrays are nsig x n
 * (signal-major).  The range may span at most 4096 steps of 1e-3, and the
 * samples (x and px) plus one level's candidates must fit in 60 KiB.
 */
int st_fine_tune(st_ctx *ctx, const double *x, const double *px, int n, int nsig, double tune_lower,
                 double tune_upper, double *tune);

/*
 * Device form on strided data, e.g. turn-by-turn records: sample i of
 * signal j is d_x[j*stride_sig + i*stride] (and d_px likewise, or null).
 * Launches on the context stream; st_ctx_sync waits.
 */
int st_fine_tune_device(st_ctx *ctx, const double *d_x, const double *d_px, int64_t stride_sig, int64_t stride,
                        int n, int nsig, double tune_lower, double tune_upper, double *d_tune);
/* st_fine_tune_device on w[i] * x[i] (and w[i] * px[i]): d_window holds n
   device weights, e.g. a Hann window for NAFF-like tunes.  No counterpart
   in the reference (its FineTuneFinder is unwindowed). */
int st_fine_tune_window_device(st_ctx *ctx, const double *d_x, const double *d_px, int64_t stride_sig, int64_t stride,
                               int n, int nsig, double tune_lower, double tune_upper, const double *d_window,
                               double *d_tune);

/* ------------------------------------------- closed orbit and TBT scans */

/*
 * Cal_Orbit_Num (simtrack.cpp:14947-15030): the 4-D closed orbit at
 * deltap (x[5]) of the line `rec` with its RF, crab and CCMULT voltages set
 * to zero, by the reference's Newton iteration (finite differences 1e-9,
 * tolerance 1e-12, at most 20 iterations, its mat_inv).  x_co: the fixed
 * point (x[4] = 0, x[5] = deltap); x_end (may be null): x_co passed once
 * more through the line, the Cell[Ncell-1]->X the drivers read.
 * iterations (may be null): iterations used.  ST_ERR_PARAM when the
 * reference would report failure.
 */
int st_closed_orbit(st_ctx *ctx, const st_globals *gp, const st_element *rec, int n, double deltap,
                    double x_co[6], double x_end[6], int *iterations);

/*
 * Cal_Spin_Orbit (simtrack.cpp:20510-20572): Cal_Orbit_Num at deltap, then
 * the periodic spin n0 by the reference's Newton iteration on (Sx, Ss) with
 * one-turn sPass maps of the line as given (RF on, no loss test; spin
 * difference 1e-6, tolerance 1e-12, at most 20 iterations, Sy = sqrt(1 -
 * Sx^2 - Ss^2) after each step).  x0: the orbit Cell[Ncell-1]->X and the
 * converged spin at the start; n0 (3*n doubles, may be null): the
 * normalised spin after every element of one more pass, the reference's
 * Cell[j]->n0.  ST_ERR_PARAM when the reference would report failure.
 */
int st_spin_orbit(st_ctx *ctx, const st_globals *gp, const st_element *rec, int n, double deltap, double x0[9],
                  double *n0, int *iterations);

/*
 * Cal_Spin_Tune (simtrack.cpp:20574-20601): st_spin_orbit, then 500 turns
 * of sPass of the closed orbit with a spin perpendicular to the last
 * element's n0, and FineTuneFinder (on [0.001, 0.499]) of its Sx after
 * each turn: the reference's Line::SpinTune.  n0_end (may be null): the
 * last element's n0.
 */
int st_spin_tune(st_ctx *ctx, const st_globals *gp, const st_element *rec, int n, double deltap, double *tune,
                 double n0_end[3]);

/*
 * Set_Initial_Spin (simtrack.cpp:20603-20635): x's orbit with unit spins
 * along x, y and s, each tracked nturn turns of sPass (no loss test); the
 * projections on n0 after every turn are summed, and x[6..8] becomes the
 * normalised sums.
 */
int st_set_initial_spin(st_ctx *ctx, const st_globals *gp, const st_element *rec, int n, int nturn,
                        const double n0[3], double x[9]);

/*
 * The 6x6 one-turn matrix about x_co (row-major, m[6*i+k] = dx_i/dx_k):
 * symmetric differences of step h of bare passes (no loss test) with the
 * RFCAV voltages off (crab cavities and CCMULT on) -- the linear part of
 * the map that Cal_OneTurnMap (simtrack.cpp:15122-15137) expands in TPS
 * and Cal_Twiss (cpp:15540) analyses.  h = 1e-6 suits proton and ion rings.
 */
int st_one_turn_matrix(st_ctx *ctx, const st_globals *gp, const st_element *rec, int n, const double x_co[6],
                       double h, double m[36]);

/*
 * The closed orbit after every element and every element's linear map, the
 * inputs of Trace_A (simtrack.cpp:15480) that Cal_Orbit_Num (cpp:14947,
 * its last pass) and Cal_ElementMap (cpp:15205-15253) compute: X[6*k..] is
 * x_co passed through elements 0..k with every RF, crab and CCMULT voltage
 * off; T[36*k..] (row-major) is element k's map about X of element k-1 (of
 * the last element for k = 0) with z = 0 and x[5] = x5, RFCAV voltages off,
 * as symmetric differences of step h.  T: n*36 doubles, X: n*6.
 */
int st_element_maps(st_ctx *ctx, const st_globals *gp, const st_element *rec, int n, const double x_co[6],
                    double x5, double h, double *T, double *X);

/*
 * Track_tbt_FMA (simtrack.cpp:21179-21241) with its grid and turns as
 * arguments (the reference: nturn 2048, sigma_step 0.1, sigma_max 6):
 * rows of 8 doubles (i*step, j*step, tunex1, tunex2, tuney1, tuney2,
 * tunex1-tunex2, tuney1-tuney2) for the stable grid particles, in the
 * reference's order; x_end[0..3] is the closed orbit added to the grid.
 * *nrows receives the row count (rows may be null to ask for it).
 */
int st_fma_scan(st_table *t, double deltap0, double sigmax0, double sigmay0, const double x_end[4],
                double tune_low, double tune_high, int nturn, double sigma_step, double sigma_max,
                double *rows, int max_rows, int *nrows);

/*
 * Track_tbt_Lyapunov (simtrack.cpp:21243-21303) with its grid and turns
 * as arguments (the reference: nturn 2048, sigma_step 0.05, sigma_max 6):
 * rows of 3 doubles (i*step, j*step, mean of the last 25 log(d/1e-6)/(k+1))
 * for the grid points whose twin particles both survive.
 */
int st_lyapunov_scan(st_table *t, double deltap0, double sigmax0, double sigmay0, int nturn, double sigma_step,
                     double sigma_max, double *rows, int max_rows, int *nrows);

/*
 * Track_tbt_ActionDiff (simtrack.cpp:21305-21383) with the start Twiss
 * (betx, alfx, bety, alfy) as arguments: rows (nstep x 6) of (turns, count,
 * mean 2Jx, rms 2Jx, mean 2Jy, rms 2Jy); *seed is the reference's LCG state
 * (its global `seed`, initially 1234), advanced as the reference does.
 */
int st_action_diff(st_table *t, double betx, double alfx, double bety, double alfy, double nsigma, double phase,
                   double emitx0, double emity0, int nturn, int nstep, int npart, double *seed, double *rows);

/* ------------------------------------ strong-strong beam-beam */

/*
 * Soft-Gaussian strong-strong beam-beam between two macroparticle beams of
 * Mp1 and Mp2 rows of 6 doubles (the reference's layout), charges Np1 and
 * Np2, Lorentz factors gamma1 and gamma2.  Each beam is kicked by BB4D
 * about the other beam's (or slice's) centroid with its rms sizes, times
 * its bbscale.  Moments are summed in a fixed order (repeatable bit for
 * bit), not the reference's sequential order, so they agree with it to
 * rounding.  Blocking; d_x1/d_x2 are device arrays, the _host forms take
 * host arrays and copy through the context.  The work runs on the context's
 * own stream, which is not ordered after other streams: a caller that
 * writes d_x1/d_x2 on another stream (a framework's current stream) waits
 * for that write before the call (simtrack_amd.strong_strong does).  The
 * same holds for every *_device entry point.
 *
 * Slice pairs that touch disjoint slices go out together on up to 8 streams
 * (the result is the sequential one bit for bit); st_ctx_set_ss_lanes
 * bounds their number (1: one pair at a time).
 *
 * st_bb_soft_gaussian_4d: BeamBeam4D_SoftGaussian (simtrack.cpp:22846).
 * st_bb_soft_gaussian_6d: BeamBeam6D_SoftGaussian (cpp:22874): both beams
 *   sorted by z (descending; the rows are left sorted; equal z keep row
 *   order), Mp/Nslice rows per slice (the last Mp % Nslice collide with
 *   nothing), slice pairs collided in order of their collision point with
 *   drifts of +-S.  Z1/Z2 (host, may be null) receive the slice centres the
 *   reference writes to bunch1_Z.dat / bunch2_Z.dat.
 * st_bb_soft_gaussian_6d_zb: the slice-boundary form (cpp:23134): slice i
 *   holds the rows with zb[i+1] <= z < zb[i] (zb: Nslice + 1 host doubles,
 *   strictly decreasing); rows stay in place; pair charges Np * m / Mp.
 * st_zboundary_slices_gaussian: Zboundary_slices_Gaussian (cpp:3157).
 */
int st_ctx_set_ss_lanes(st_ctx *ctx, int lanes);
int st_bb_soft_gaussian_4d(st_ctx *ctx, double Np1, int Mp1, double *d_x1, double Np2, int Mp2, double *d_x2,
                           double gamma1, double gamma2, double bbscale1, double bbscale2);
int st_bb_soft_gaussian_6d(st_ctx *ctx, double Np1, int Mp1, double *d_x1, double Np2, int Mp2, double *d_x2,
                           double gamma1, double gamma2, double bbscale1, double bbscale2, int Nslice1, int Nslice2,
                           double *Z1, double *Z2);
int st_bb_soft_gaussian_6d_zb(st_ctx *ctx, double Np1, int Mp1, double *d_x1, double Np2, int Mp2, double *d_x2,
                              double gamma1, double gamma2, double bbscale1, double bbscale2, int Nslice1,
                              int Nslice2, const double *zb1, const double *zb2);
int st_bb_soft_gaussian_4d_host(st_ctx *ctx, double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2,
                                double gamma1, double gamma2, double bbscale1, double bbscale2);
int st_bb_soft_gaussian_6d_host(st_ctx *ctx, double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2,
                                double gamma1, double gamma2, double bbscale1, double bbscale2, int Nslice1,
                                int Nslice2, double *Z1, double *Z2);
int st_bb_soft_gaussian_6d_zb_host(st_ctx *ctx, double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2,
                                   double gamma1, double gamma2, double bbscale1, double bbscale2, int Nslice1,
                                   int Nslice2, const double *zb1, const double *zb2);
int st_zboundary_slices_gaussian(double *zb, int Nslice, double sigmal);

/*
 * Particle-in-cell strong-strong beam-beam (the same beams and arguments):
 * st_bb_pic_4d: BeamBeam4D_PIC (simtrack.cpp:23286): a 64 x 64 grid of
 *   15 times the smaller rms size centred on the source beam, 3 x 3 charge
 *   weights, the integrated Green function convolved by the reference's
 *   radix-2 FFT on the doubled grid, the field by its 4th-order stencil,
 *   interpolated back; beam 2 is kicked by beam 1's field, then beam 1 by
 *   beam 2's.  Only particles within 29 cells of their beam's centre are
 *   deposited or kicked.  Every step after the moments follows the
 *   reference's operations in its order (deposition sums each cell in row
 *   order); the moments are fixed-order sums as above.
 * st_bb_pic_6d: BeamBeam6D_PIC (cpp:23380): the z-sorted slicing and the
 *   collision order of st_bb_soft_gaussian_6d, st_bb_pic_4d per slice pair
 *   with charges Np/Nslice.
 * st_bb_pic_4d_nonoffset: BeamBeam4D_PIC_NonOffset (cpp:24226): the grid is
 *   given, Nx x Ny cells of hx x hy centred on the origin (no moments, no
 *   centroid offsets); the kick takes the field from the potential by the
 *   same stencil (Interpolate_Electric_Field_1Potential, cpp:3047); the
 *   window is 29 cells around the origin.  The reference's caller passes
 *   the Green function of that grid (grn_c_r, grn_c_i); here it is computed
 *   from hx, hy (Cal_Effective_Green_Function with zero offsets, cpp:2881).
 *   Nx = Ny = 64 only (else ST_ERR_PARAM), as the solver's FFT is.
 */
int st_bb_pic_4d(st_ctx *ctx, double Np1, int Mp1, double *d_x1, double Np2, int Mp2, double *d_x2, double gamma1,
                 double gamma2, double bbscale1, double bbscale2);
int st_bb_pic_6d(st_ctx *ctx, double Np1, int Mp1, double *d_x1, double Np2, int Mp2, double *d_x2, double gamma1,
                 double gamma2, double bbscale1, double bbscale2, int Nslice1, int Nslice2, double *Z1, double *Z2);
int st_bb_pic_4d_host(st_ctx *ctx, double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2, double gamma1,
                      double gamma2, double bbscale1, double bbscale2);
int st_bb_pic_4d_nonoffset(st_ctx *ctx, double Np1, int Mp1, double *d_x1, double Np2, int Mp2, double *d_x2,
                           double gamma1, double gamma2, double bbscale1, double bbscale2, int Nx, int Ny, double hx,
                           double hy);
int st_bb_pic_4d_nonoffset_host(st_ctx *ctx, double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2,
                                double gamma1, double gamma2, double bbscale1, double bbscale2, int Nx, int Ny,
                                double hx, double hy);
int st_bb_pic_6d_host(st_ctx *ctx, double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2, double gamma1,
                      double gamma2, double bbscale1, double bbscale2, int Nslice1, int Nslice2, double *Z1,
                      double *Z2);

#ifdef __cplusplus
}
#endif

#endif /* SIMTRACK_AMD_H */
