/*
 * oracle.h -- CPU oracle for the 2-D electrostatic PIC timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/ (as the checker), by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 *
 * Two families of functions live here:
 *
 *  orc_ref_*  FP64, array-of-structs restatement of the reference's algorithm
 *             (the 48-byte particleVec of particleVec.h:26-35, absolute
 *             positions, the reference's truncation-based cell tests).  These
 *             follow pic.cl, particle.cpp, fields.cpp, mglin.cpp and
 *             boundary.cpp statement by statement; each function cites the
 *             lines it follows.  They are also the CPU baseline.
 *
 *  orc32_*    the same algorithm restated on the GPU library's data layout:
 *             SoA fp32 state, position = packed int16 cell (jx, jy) + fp32
 *             in-cell offset, 64-bit fixed-point charge.  Every arithmetic
 *             operation is spelled out in the order the HIP kernels use, so
 *             that with -ffp-contract=off on both sides the integer outputs
 *             (cells, fixed-point charge, loss counts) and the fp32/fp64
 *             state are bit-identical to the GPU's.
 *
 * How each function is pinned (independently of its own formula) is listed in
 * the test files under tests/ (test_oracle_*.py); functions with no
 * independent pin say "parity unpinned" below.
 */
#ifndef PIC2D_ORACLE_H
#define PIC2D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 counter-based RNG (Salmon et al., SC'11, "Parallel random   */
/* numbers: as easy as 1, 2, 3").  Pinned by the published Random123 KAT     */
/* vectors in tests/test_oracle_philox.py.                                   */
/* ------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Philox2x32-10 (the same paper): counter (c0, c1), one key word */
void orc_philox2x32_10(const uint32_t ctr[2], uint32_t key, uint32_t out[2]);
/* uniform in the open interval (0,1): ((x >> 9) + 0.5) * 2^-23, exact in fp32 */
float orc_u01(uint32_t x);
/* 53-bit uniform in (0,1) from two words (the loader's and the neutral table's) */
double orc_u01d(uint32_t hi, uint32_t lo);

/* ------------------------------------------------------------------------ */
/* FP64 reference semantics (particleVec AoS)                               */
/* ------------------------------------------------------------------------ */
typedef struct {
    double x, y, vx, vy, vz, weight;   /* particleVec.h:26-35 */
} orc_particle;

/* pic.cl:56-92 (get_particle_electric_field).  Ex/Ey use the reference's
 * (N+1)^2 layout with row stride N (fields.cpp:65-76). */
void orc_ref_gather(const orc_particle *p, long n, const double *Ex, const double *Ey,
                    double dx, double dy, int nx, int ny, double *Exp, double *Eyp);

/* pic.cl:95-149 (move_particles); constants as set in particle.cpp:1103-1112 */
void orc_ref_push(orc_particle *p, long n, const double *Exp, const double *Eyp,
                  double dx, double dy, int nx, int ny, double coef, double cos_a,
                  double sin_a, double dt, double cos_w, double sin_w, double bfield);

/* pic.cl:22-52 (collectCharge) with serial (race-free) accumulation into q,
 * which the caller zeroes (particle.cpp:681-685). */
void orc_ref_deposit(const orc_particle *p, long n, double qc_dxdy, double dx, double dy,
                     int nx, int ny, double *q);

/* particle.cpp:830-845: edges x2, then corners x2 again (corners end x8). */
void orc_ref_charge_postscale(double *q, int nx, int ny);

/* fields.cpp:100-143 (one ion species): rho = q_ion; rho += q_e */
void orc_ref_grid_rho(const double *q_ion, const double *q_e, int n, double *rho);

/* mglin.cpp:21-196: full multigrid, ncycle V(1,1) cycles per level, red-black
 * Gauss-Seidel, half-weighting restriction, bilinear prolongation.  u (n*n,
 * row-major, u[i*n+j] = NR's u[i][j]) holds the rhs on entry and the solution
 * on exit.  Returns 0, or -1 if n-1 is not a power of two (mglin.cpp:25-27). */
int orc_mglin(double *u, int n, int ncycle, double length);

/* fields.cpp:169-220: rhs = -rho/EPSILON, Dirichlet folding, Mglin, boundary
 * overwrite.  eps is the permittivity (constants.h:28). */
int orc_ref_grid_potential(const double *rho, int n, double dx, double dy, double eps,
                           double phi0X, double phiLX, double phi0Y, double phiLY,
                           double timestep, double phi_period, int ncycle, double length,
                           double *phi);

/* fields.cpp:230-268.  Ex/Ey are (n+1)^2 arrays with row stride n that the
 * caller zero-initialises once and keeps (entries the reference never writes
 * stay 0, see the never-written edge rows). */
void orc_ref_grid_efield(const double *phi, int n, double dx, double *Ex, double *Ey);

/* boundary.cpp:581-677 (e_boundary) and 765-824 (ion_boundary) with all
 * surface yields zero: every particle with (int)(x/dx) or (int)(y/dy)
 * outside [0, N) is removed by swap-with-last + pop.  id (may be NULL) is
 * permuted alongside.  Returns the new count. */
long orc_ref_absorb(orc_particle *p, uint32_t *id, long n, double dx, double dy, int nx, int ny);

/* ------------------------------------------------------------------------ */
/* Build-layout (fp32 SoA, cell + offset) restatement                        */
/* ------------------------------------------------------------------------ */
#define ORC_MODE_REF 0        /* reference wall / frozen-particle semantics  */
#define ORC_MODE_PHYS 1       /* particles leave at x<0 or x>=L              */
#define ORC_MODE_REF_STEADY 2 /* REF with steady-state walls: nobody leaves;
                                 outside electrons and ions are paired and
                                 re-injected (boundary.cpp:890-1097)          */

#define ORC_ST_KEPT 0
#define ORC_ST_LOST_LEFT 1
#define ORC_ST_LOST_RIGHT 2
#define ORC_ST_LOST_BOTTOM 3
#define ORC_ST_LOST_TOP 4

static inline uint32_t orc_cell_pack(int jx, int jy) {
    return (uint32_t)(uint16_t)(int16_t)jx | ((uint32_t)(uint16_t)(int16_t)jy << 16);
}
static inline int orc_cell_jx(uint32_t c) { return (int)(int16_t)(c & 0xFFFFu); }
static inline int orc_cell_jy(uint32_t c) { return (int)(int16_t)(c >> 16); }

typedef struct {
    int N;                     /* nodes per side (2^k + 1)                    */
    int mode;                  /* ORC_MODE_REF / ORC_MODE_PHYS                */
    float cE;                  /* fp32(0.5 * dt * q/m)                        */
    float dtdx, dtdy;          /* fp32(dt/dx), fp32(dt/dy)                    */
    int bfield;                /* magnetic rotation on/off                    */
    float cos_a, sin_a, cos_w, sin_w;
} orc32_push_params;

/* Gather + push + reposition + wall classification of one species
 * (pic.cl:56-149 on the cell/offset layout; wall test boundary.cpp:595-599).
 * E32 holds (Ex, Ey) fp32 pairs in the reference's padded stride-N layout,
 * (N+1)*N + 1 pairs.  status[i] receives ORC_ST_*. */
void orc32_push(long n, uint32_t *cell, float *ox, float *oy, float *vx, float *vy, float *vz,
                const float *E32, const orc32_push_params *pp, uint8_t *status);

/* Fixed-point CIC deposit (pic.cl:22-52 on the cell/offset layout): weights
 * w = rint(offset * 2^20), corner products summed in int64, q[N*N]. */
void orc32_deposit(long n, const uint32_t *cell, const float *ox, const float *oy, int N,
                   int mode, int64_t *qfix);

/* Fixed-point charge -> rho -> Poisson rhs (particle.cpp:830-845 post-scale,
 * fields.cpp:100-143 sum order, fields.cpp:184-204 rhs and Dirichlet folding).
 * scale_s = qc_s/(dx*dy) * weight_s * 2^-48.  corner8: corners x8 (reference)
 * else x4.  rho and rhs are n*n. */
void orc_fixed_to_rhs(const int64_t *q_ion, const int64_t *q_e, int n, double scale_i,
                      double scale_e, int corner8, double eps, double phi0X, double phiR,
                      double phi0Y, double phiLY, double dxdy_inv, double *rho, double *rhs);

/* fields.cpp:209-219 boundary overwrite + fields.cpp:230-268 E stencil,
 * computed per node from phi (boundary already overwritten by the caller),
 * and packed into the fp32 (Ex, Ey) gather layout.  Ex64/Ey64 may be NULL. */
void orc_efield_pack(const double *phi, int n, double dx, double *Ex64, double *Ey64, float *E32);

/* Residual-controlled FP64 multigrid (the TOL mode): FMG with ncycle cycles,
 * then V(1,1) cycles until ||f - A u||_2 <= tol * ||f||_2 over interior nodes
 * or max_cycles.  Returns the number of extra V-cycles, -1 on a bad n. */
int orc_mg_tol(double *u_rhs, int n, double length, double tol, int max_cycles, double *rel_res_out);

/* relative residual ||f - A u|| / ||f|| over interior nodes (A as mglin.cpp:133-144) */
double orc_mg_relres(const double *u, const double *f, int n, double length);


/* Batch diagnostics of one species for one step (graphs.cpp:459-691, kernels
 * pic.cl:151-258) on the build layout: velocity sums along the centre row
 * (vx, cells with (int)(y/dy) == N/2 and x > 0) and column (vy) in 64-bit
 * fixed point (2^-20 m/s) with counts; histograms of v/vt * divide_bins_by
 * (truncated) at the two locations +- spread: dist = 4 x no_bins (first X,
 * first Y, second X, second Y).  Accumulates into the outputs. */
void orc32_diag_particles(long n, const uint32_t *cell, const float *ox, const float *oy,
                          const float *vx, const float *vy, int N, double vt, int no_bins,
                          int divide_bins_by, int spread, const int loc[4], int64_t *vsx,
                          int64_t *vsy, uint32_t *cntx, uint32_t *cnty, uint64_t *dist);
/* end of a step: vp += mean velocity of every cell holding a particle
 * (graphs.cpp:683-691); the step sums are cleared */
void orc_diag_vfinal(int N, int64_t *vsx, int64_t *vsy, uint32_t *cntx, uint32_t *cnty,
                     double *vpx, double *vpy);

/* ------------------------------------------------------------------------ */
/* Whole timesteps (pic::pic_run order, pic.cpp:1069-1162, collisionless)     */
/* ------------------------------------------------------------------------ */
typedef struct {
    int N, ncycle, dt_dti, pad;
    double L, dx, dy, eps;
    double coef_e, coef_i, dt_e, dt_i, qc_dxdy_e, qc_dxdy_i;
    double cos_a, sin_a, cos_we, sin_we, cos_wi, sin_wi, bfield;
    double phi_left, phi_right, phi_bottom, phi_top, phi_period;
    long start_step;
    /* 0: the reference's FMG with ncycle V-cycles (fields.cpp:174); > 0: the
     * same folded right-hand side solved to ||f - Au|| <= mg_tol ||f||
     * (FMG start, then FP64 V(1,1) cycles): a converged FP64 solve */
    double mg_tol;
} orc_ref_config;

/* fields::grid_potential (fields.cpp:169-220) with the solve selected by
 * c->mg_tol (see orc_ref_config) */
int orc_ref_potential_cfg(const orc_ref_config *c, const double *rho, double timestep,
                          double *phi);

/* FP64 reference step loop (also the CPU baseline; OpenMP-parallel gather,
 * push and deposit when more than one thread is set).  Ex/Ey: (N+1)^2. */
int orc_ref_run(const orc_ref_config *c, int nsteps, long *ne, orc_particle *e, uint32_t *eid,
                long *ni, orc_particle *ip, uint32_t *iid, double *Ex, double *Ey, double *qe,
                double *qi, double *rho, double *phi);
/* OpenMP thread count for orc_ref_run; returns the count in effect */
int orc_set_threads(int n);

typedef struct {
    int N, mode, corner8, mg_mode; /* mg_mode 0: FMG ncycle fixed; 1: residual tolerance */
    int mg_ncycle, mg_max_cycles, dt_dti, preloop;
    double L, dx, eps, mg_tol;
    double phi_left, phi_right, phi_bottom, phi_top, phi_period;
    double scale_e, scale_i;
    long start_step;
    orc32_push_params pe, pi;
    uint32_t key0, key1;       /* Philox key (seed) of the re-injection draws */
    float steady_se, steady_si; /* re-injection velocity scales (VTE, VTI)   */
} orc32_config;

/* Build-layout step loop: push + wall pass (stable compaction, per-wall loss
 * counts into lost8[0..3] electrons, [4..7] ions; in ORC_MODE_REF_STEADY the
 * pairing pass of orc32_steady instead), fixed-point deposit, rho/rhs,
 * multigrid, phi boundary overwrite, packed fp32 E. */
int orc32_run(const orc32_config *c, int nsteps, long *ne, uint32_t *ecell, float *eox, float *eoy,
              float *evx, float *evy, float *evz, uint32_t *eid, long *ni, uint32_t *icell,
              float *iox, float *ioy, float *ivx, float *ivy, float *ivz, uint32_t *iid,
              float *E32, int64_t *qe, int64_t *qi, double *rho, double *phi, long *lost8,
              int *mg_cycles_out);

/* Reference steady-state electron wall (boundary.cpp:890-1097) on the build
 * layout: per side (left x<0, right (int)(x/dx)>=N, then bottom, top among
 * the particles the x pass left outside) the m-th electron outside, in id
 * order, pairs with the m-th ion outside; both move to one point uniform in
 * [dx, L-dx]^2 with velocities v = vt cos(2 pi R2) sqrt|ln R1| (boundary.cpp:
 * 833-878), draws keyed by the ion's id (counter (id, step, 16+3..5,
 * 0x4D434331)).  Returns the number of pairs. */
long orc32_steady(long ne, uint32_t *ecell, float *eox, float *eoy, float *evx, float *evy,
                  float *evz, const uint32_t *eid, long ni, uint32_t *icell, float *iox,
                  float *ioy, float *ivx, float *ivy, float *ivz, const uint32_t *iid, int N,
                  uint32_t step, uint32_t key0, uint32_t key1, float se, float si);

/* ------------------------------------------------------------------------ */
/* Collision cross sections [1e-20 m^2], energies in eV (orc_collisions.c)    */
/* ------------------------------------------------------------------------ */
#define ORC_SIG_AR_EL_REF 0   /* e_collisions.cpp:205-224, branches as written     */
#define ORC_SIG_AR_EL_PHYS 1  /* same fit, closed breakpoints, clamped >= 0        */
#define ORC_SIG_AR_IZ_REF 2   /* e_collisions.cpp:235-256                          */
#define ORC_SIG_AR_IZ_PHYS 3  /* same fit above threshold thr, clamped >= 0        */
#define ORC_SIG_AR_EXC 4      /* physical-mode excitation, threshold thr           */
#define ORC_SIG_H_EL 5        /* e_collisions.cpp:267-283                          */
#define ORC_SIG_H_IZ 6        /* e_collisions.cpp:293-308 (bracket / (I E))        */
#define ORC_SIG_AR_CX 7       /* ion_collisions.cpp:170-176                        */
#define ORC_SIG_AR_ISO 8      /* Phelps (1994) Ar+ - Ar isotropic, physical mode   */
#define ORC_SIG_H_CX 9        /* ion_collisions.cpp:185-199, E in keV/amu          */
double orc_sigma(int kind, double E, double thr);

/* ------------------------------------------------------------------------ */
/* Physical-mode Monte-Carlo collisions in FP64 (orc_mcc.c).  Null-collision */
/* method of Vahedi & Surendra (Comput. Phys. Commun. 87 (1995) 179) with the */
/* cross sections above; draws = Philox4x32-10 with counter (id, step,        */
/* species*16 + stream, 0x4D434331) and key (key0, key1), the GPU library's   */
/* streams.  Pinned by tests/test_oracle_collisions.py (rates against         */
/* sigma*g*n, energy and momentum bookkeeping of every channel).             */
/* ------------------------------------------------------------------------ */
#define ORC_MCC_BINS 64
typedef struct {
    int gas;          /* 0 argon, 1 hydrogen                                  */
    int electron;     /* 1 electrons, 0 ions                                  */
    double ng;        /* neutral density [m^-3]                               */
    double dt;        /* species time step [s]                                */
    double mass, M;   /* species and neutral mass [kg]                        */
    double q;         /* elementary charge [C]                                */
    double vtn;       /* neutral sigma per component [m/s]                    */
    double e_exc, e_iz, opal_B, kev_amu_div;
    double nu_max, p_null; /* set by orc_mcc_setup                            */
    uint32_t key0, key1;
    /* electrons (set by orc_mcc_setup): null-collision bound per energy bin
     * of the particle, two bins per octave of E = m v^2 / 2q from 2^-12 eV
     * (bin 0 takes everything below, bin 63 everything above); bin_nu[b] =
     * 1.02 max of the total frequency over the relative speeds the bin's
     * particles reach against neutrals up to 10 vtn fast (257 speeds),
     * bin_p[b] = 1 - exp(-bin_nu[b] dt).  binned = 0: the global bound. */
    int binned, pad_;
    double bin_p[ORC_MCC_BINS], bin_nu[ORC_MCC_BINS];
} orc_mcc_params;

/* nu' = 1.02 * max of the total collision frequency over 1e-6 .. 3e4 eV
 * (electrons) or 1e-6 .. 1e3 eV (ions); p_null = 1 - exp(-nu' dt); the
 * electrons' per-bin bounds */
void orc_mcc_setup(orc_mcc_params *c);
/* the bin of a particle of speed^2 v2 (electrons; see orc_mcc_params) */
int orc_energy_bin(const orc_mcc_params *c, double v2);
/* returns 0 none, 1 elastic, 2 excitation, 3 ionisation (child = ejected
 * electron v[0..2], new ion v[3..5]) */
int orc_collide_electron(const orc_mcc_params *c, double *vx, double *vy, double *vz, uint32_t id,
                         uint32_t step, double child[6]);
/* returns 0 none, 4 charge exchange, 5 isotropic elastic */
int orc_collide_ion(const orc_mcc_params *c, double *vx, double *vy, double *vz, uint32_t id,
                    uint32_t step);
/* the two above over arrays: event[k] per particle, child6 (may be NULL) 6
 * doubles per particle (zeros unless ionised) */
void orc_mcc_apply(const orc_mcc_params *c, long n, double *vx, double *vy, double *vz,
                   const uint32_t *id, uint32_t step, int32_t *event, double *child6);

/* Reference-mode ion charge exchange (ion_collisions.cpp:255-308) on the
 * build layout, fp32 in the GPU's operation order: eligible ions ((int)(x/dx)
 * and (int)(y/dy) in [0, N)) visited in (column, id) order add
 * P = 1 - exp(-u dt 1e-20 n_g sigma) to pcol[column]; above 1 the ion takes
 * the velocity of neutral-table entry (int)(u01 * 9999) (Philox counter (id,
 * step, 16+9, 0x4D434331)) and pcol restarts at 0.  Returns the events. */
typedef struct {
    int gas;         /* 0 argon, 1 hydrogen                                   */
    int sig_energy;  /* 1: energy-dependent sigma, 0: sig_const               */
    float sig_const; /* [1e-20 m^2]                                           */
    float q, mass, dt, ng;
    uint32_t key0, key1;
} orc_refcx_params;
float orc32_ref_cx_prob(const orc_refcx_params *c, float vx, float vy, float vz);
long orc32_ref_cx(long n, const uint32_t *cell, const float *ox, const float *oy, float *vx,
                  float *vy, float *vz, const uint32_t *id, int N, const orc_refcx_params *c,
                  float *pcol, const float *ntab, int ntab_n, uint32_t step);
/* the reference's neutral velocity table (collisions.cpp:66-89) drawn from
 * Philox: 3 x 10000 floats (vx | vy | vz) */
void orc_neutral_table(uint64_t seed, double vtn, float *t);

/* Collisional FP64 step loop in the reference's phase order (pic.cpp:1069-
 * 1162): electron MCC (children appended in creator order), electron push,
 * ion MCC + push every dt_dti steps, walls x<0 or x>=L with, when plop is
 * set, an electron-ion pair re-injected per lost ion (boundary.cpp:854-878
 * placement, Maxwellian sigma vte / vti), deposit, fields.  events8 counts
 * collision events by kind.  next_id[2] is advanced by the ids handed out. */
int orc_ref_run_mcc(const orc_ref_config *c, const orc_mcc_params *me, const orc_mcc_params *mi,
                    int plop, double vte, double vti, int nsteps, long *ne, long cap_e,
                    orc_particle *e, uint32_t *eid, long *ni, long cap_i, orc_particle *ip,
                    uint32_t *iid, uint32_t *next_id, double *Ex, double *Ey, double *qe,
                    double *qi, double *rho, double *phi, long *events8);

/* ------------------------------------------------------------------------ */
/* dataDump text I/O (particle.cpp:359-401 reader, 1361-1382 writer)          */
/* ------------------------------------------------------------------------ */
/* precision = significant digits (6 reproduces the reference's ostream default) */
int orc_dump_write(const char *path, const orc_particle *p, long n, int precision);
/* returns count read, or -1 if the file cannot be opened; reads at most cap */
long orc_dump_read(const char *path, orc_particle *p, long cap);

/* ------------------------------------------------------------------------ */
/* Particle-count control (orc_control.c): source, reduction, weighted       */
/* deposit.  Pinned by tests/test_oracle_control.py (distribution of the     */
/* source against its closed form; quotas, group weights and conserved       */
/* weight against a literal restatement of reduce_population; weighted      */
/* deposit against the unweighted one).                                      */
/* ------------------------------------------------------------------------ */
void orc_inject(long n, uint32_t species, uint32_t step, uint32_t key0, uint32_t key1, double dx,
                double L, double vt, uint32_t *cell, float *ox, float *oy, float *vx, float *vy,
                float *vz);
typedef struct {
    int ng;
    uint32_t size_check, species, key0, key1;
    double vlim[8];
    uint32_t glimit[8], gnew[8];
} orc_reduce_params;
long orc_reduce(const orc_reduce_params *c, uint32_t step, long n, const float *vx,
                const float *vy, const uint32_t *id, float *w, uint8_t *keep, uint32_t *ngroup,
                uint32_t *nrem);
void orc32_deposit_w(long n, const uint32_t *cell, const float *ox, const float *oy,
                     const float *w, int N, int mode, int64_t *qfix);

/* ------------------------------------------------------------------------ */
/* Surface physics of the non-steady walls (orc_wall.c).  Pinned by          */
/* tests/test_oracle_wall.py (yields at their closed-form points, branch     */
/* frequencies, emission energy and cosine-law angle distributions, the     */
/* reference-mode emission positions).                                       */
/* ------------------------------------------------------------------------ */
typedef struct {
    int mode, N;
    double dx, Lw, dt, me, mi, q, max_yield, E0, Emax0, thr, wf;
    int see_e, refl, back, see_i;
    uint32_t key0, key1;
} orc_wall_params;
typedef struct {
    uint32_t *cell;
    float *ox, *oy, *vx, *vy, *vz, *w;
    uint64_t *key;
} orc_wall_out;
double orc_vaughan(double E, double theta, double max_yield, double E0, double Emax0);
double orc_ar_ion_yield(double E);
/* species 0: electrons reaching a wall; 1: ions.  Returns the number of
 * particles returned or emitted (written up to cap); counts[6] = reached,
 * absorbed, reflected, backscattered, emitting, secondaries created. */
long orc_wall(const orc_wall_params *c, int species, uint32_t step, long n, const uint32_t *cell,
              const float *ox, const float *oy, const float *vx, const float *vy, const float *vz,
              const float *w, const uint32_t *id, orc_wall_out *out, long cap, int64_t counts[6]);

#ifdef __cplusplus
}
#endif
#endif
