/*
 * pic2d.h -- C ABI of the MI355X-native 2-D electrostatic PIC timestep.
 *
 * Drop-in seam for the reference's pic / particle / fields / Mglin /
 * collisions / boundary classes.  Every entry point returns an int status
 * (PIC_OK or a negative PIC_ERR_*); none exits the process or throws.
 * Plain C types only: caller-owned host buffers are copied in or out and
 * never retained.  A handle is single-thread affine; calls are ordered on the
 * handle's HIP stream and the getters synchronise it.
 *
 * Species index s: 0 = electrons, 1 = ions (NO_ION_SPECIES is 1 in every
 * reference configuration, parameters.h:283).
 */
#ifndef PIC2D_H
#define PIC2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define PIC_OK 0
#define PIC_ERR_ARG (-1)         /* bad argument / handle                    */
#define PIC_ERR_GRID (-2)        /* nodes-1 not a power of two (mglin.cpp:25-27) */
#define PIC_ERR_FILE (-3)        /* file cannot be opened                    */
#define PIC_ERR_FORMAT (-4)      /* garbled parameters.h or dataDump         */
#define PIC_ERR_DEVICE (-5)      /* HIP runtime error or no GPU              */
#define PIC_ERR_CAPACITY (-6)    /* particle store full                      */
#define PIC_ERR_UNSUPPORTED (-7) /* option outside this build's scope        */
#define PIC_ERR_ALLOC (-8)       /* device or host allocation failed         */
#define PIC_ERR_NUMERIC (-9)     /* the potential or field became non-finite */
#define PIC_ERR_TIMEOUT (-10)    /* pic_serve_wait: the request is not done  */

/* --------------------------------------------------------------- options */
#define PIC_MODE_REF 0      /* the reference's wall tests, frozen particles,
                               corner x8, collision decision rules           */
#define PIC_MODE_PHYSICAL 1 /* x<0 or x>=L leaves, null-collision MCC        */

#define PIC_MG_FIXED 0 /* FP64 FMG, ncycle V(1,1) per level (fields.cpp:174) */
#define PIC_MG_TOL 1   /* FP32 smoothing + FP64 residual refinement to tol  */

#define PIC_LOAD_RANDOM 0  /* Philox uniform positions, Maxwellian velocities */
#define PIC_LOAD_LATTICE 1 /* particle.cpp:410-443 lattice, both species co-located */

#define PIC_WALL_ABSORB 0 /* boundary.cpp:581-677, 765-824 with zero yields  */
#define PIC_WALL_STEADY 1 /* boundary.cpp:854-1103 electron-ion "plop"       */

#define PIC_GAS_ARGON 0
#define PIC_GAS_HYDROGEN 1

#define PIC_SIGMA_CONSTANT 0
#define PIC_SIGMA_ENERGY 1

/*
 * Runtime mirror of parameters.h (parameters.h:41-366, constants.h:15-36).
 * Derived quantities are stored already evaluated, as the reference's macros
 * evaluate them.  pic_params_default() fills the reference defaults;
 * pic_params_load() evaluates a parameters.h-style file.
 */
typedef struct pic_params {
    int32_t nodes;            /* X_NO_OF_CELLS == Y_NO_OF_CELLS (node count, 2^k+1) */
    int32_t init_particle_no; /* INIT_PARTICLE_NO, per species                      */
    int32_t mode;             /* PIC_MODE_*                                         */
    int32_t loader;           /* PIC_LOAD_*                                         */
    int32_t wall;             /* PIC_WALL_* (STEADYSTATE)                           */
    int32_t have_collisions;  /* HAVE_COLLISIONS                                    */
    int32_t gas;              /* SPECIES                                            */
    int32_t sigma_cx;         /* SIGMA            (PIC_SIGMA_*)                     */
    int32_t sigma_elastic;    /* SIGMA_ELASTIC    (PIC_SIGMA_*)                     */
    int32_t sigma_ionization; /* SIGMA_IONIZATION (PIC_SIGMA_*)                     */
    int32_t mg_mode;          /* PIC_MG_*                                           */
    int32_t mg_ncycle;        /* V-cycles per FMG level (fields.cpp:174 uses 2)     */
    int32_t mg_max_iter;      /* PIC_MG_TOL: refinement sweeps cap                  */
    int32_t dt_dti;           /* DT_DTI                                             */
    int32_t run_time;         /* TOTAL_RUN_TIME                                     */
    int32_t number_to_inject; /* NUMBER_TO_INJECT: particles of each species created
                                 per step (particle.cpp:491-510)                    */
    int32_t no_ion_species;   /* NO_ION_SPECIES (only 1 is supported)               */
    int32_t corner8;          /* charge post-scale at corners: 1 -> x8 (reference)  */
    uint64_t seed;            /* RAND_SEED                                          */
    double length;            /* XLENGTH == YLENGTH [m]                             */
    double dt, dti;           /* DT, DTI [s]                                        */
    double electron_mass, ion_mass;     /* ELECTRON_MASS, ION_MASS [kg]             */
    double electron_charge, ion_charge; /* ELECTRON_CHARGE (<0), ION_CHARGE [C]     */
    double qce, qci;          /* QCE, QCI: macro-particle charges [C]               */
    double vte, vti, vtn;     /* VTE, VTI, VTN thermal speeds [m/s]                 */
    double sigma_load_e, sigma_load_i; /* Maxwellian sigma used by the loader       */
    double epsilon;           /* EPSILON                                            */
    double neutral_density;   /* NEUTRAL_DENSITY [m^-3]                             */
    double threshold;         /* THRESHOLD [eV]                                     */
    double sigma_value, elastic_sigma_value, ionization_sigma_value; /* constants   */
    double magnetic_field, angle;       /* MAGNETIC_FIELD [T], ANGLE [rad]          */
    double phi_left, phi_right, phi_bottom, phi_top, phi_period;
    double mg_tol;            /* PIC_MG_TOL relative residual target                */
    int32_t sort_interval;    /* particle passes per counting sort (1 = every step;
                                 in between, particles are updated in place);
                                 0 = chosen per species from its thermal
                                 displacement per step                              */
    /* batch-mode diagnostics (parameters.h:225-313; graphs.cpp) */
    int32_t graph_start;      /* START_GRAPH_TIME: first step of the averaging window */
    int32_t graph_steps;      /* NO_GRAPH_STEPS (0: no graph averages)                */
    int32_t time_data_freq;   /* TIME_DATA_FREQUENCY (0: no timeData rows)            */
    int32_t no_bins;          /* NO_BINS velocity bins (odd)                          */
    int32_t divide_bins_by;   /* DIVIDE_BINS_BY: bins per thermal speed               */
    int32_t spread;           /* SPREAD: cells about the distribution locations       */
    int32_t dist_loc[4];      /* DIST_LOCATIONX, DIST_LOCATIONY, DIST_LOCATION2X,
                                 DIST_LOCATION2Y                                      */
    int32_t probe[6];         /* POINT1X, POINT1Y, POINT2X, POINT2Y, POINT3X, POINT3Y */
    int32_t write_graphs;     /* pic::writeGraphDataToFiles (pic.cpp:58)              */
    /* particle-count control: contain_particle_growth (particle.cpp:520-653,
       parameters.h:320-366, set up at pic.cpp:257-327).  Index [0] electrons,
       [1] ions; group g holds the particles with |(vx,vy)| <= v_limit[s][g]
       not in a lower group. */
    int32_t reduce;              /* 1: run it every step (the reference does when
                                    STEADYSTATE is "NO", pic.cpp:1132-1141)       */
    int32_t no_groups;           /* NO_GROUPS (1..8)                              */
    int32_t size_to_check[2];    /* SIZE_TO_CHECK_E, SIZE_TO_CHECK                */
    int32_t group_limit[2][8];   /* UPPER_GROUP_LIMIT, UPPER_GROUP_LIMIT2         */
    int32_t new_group_size[2][8];/* NEW_GROUP_SIZE, NEW_GROUP_SIZE2               */
    double v_limit[2][8];        /* VE1, VE2, UPPER_V_LIMIT / VI1, VI2,
                                    UPPER_V_LIMIT_I [m/s]                         */
    /* surface physics of the non-steady walls (parameters.h:136-152, set up at
       pic.cpp:455-505; boundary.cpp:261-529, 586-824) */
    int32_t e_impact_see;        /* E_IMPACT_SEE: Vaughan electron-impact emission */
    int32_t ion_impact_see;      /* ION_IMPACT_SEE: Ar+ impact emission (argon)    */
    int32_t reflection;          /* REFLECTION                                    */
    int32_t backscatter;         /* BACKSCATTER                                   */
    double see_max_yield;        /* max_yield at normal incidence (1.4)           */
    double see_E0;               /* emission threshold E0 [eV] (4.32)             */
    double see_Emax0;            /* energy of the maximum yield [eV] (650)        */
    double work_function;        /* WORK_FUNCTION [eV]                            */
    int32_t sort_key;            /* counting-sort granularity: key = cell key >> k,
                                    k = 0..10 (0: per cell; 6: per 8x8-cell tile;
                                    7..10: 2..16 tiles along x); -1 = chosen by
                                    the library                                    */
    int32_t mg_sweeps;           /* PIC_MG_TOL: red-black sweeps before and after
                                    the coarse correction on each level, 1..3
                                    (V(1,1)..V(3,3)); -1 = chosen by the library   */
    int32_t null_bound;          /* physical-mode electron collisions: the null-
                                    collision bound per energy bin of the electron
                                    (PIC_NULL_BOUND_BINNED, the default) or one
                                    bound for all (PIC_NULL_BOUND_GLOBAL, Vahedi &
                                    Surendra 1995); both are exact                */
    int32_t graphs;              /* 1: repeated launch sequences (the tolerance
                                    solve's batches of V-cycles) are captured once
                                    as hipGraphs and replayed; 0 (default):
                                    launched kernel by kernel.  Results are
                                    identical; on MI355X the replay measured
                                    slower (C2 field solve 0.235 vs 0.229 ms).     */
} pic_params;

#define PIC_NULL_BOUND_BINNED 0
#define PIC_NULL_BOUND_GLOBAL 1

/* particleVec.h:26-35: the 48-byte AoS record of the reference */
typedef struct pic_particle {
    double x, y, vx, vy, vz, weight;
} pic_particle;

typedef struct pic_handle pic_handle;

/* ------------------------------------------------------------ setup, run */

/* Reference defaults of parameters.h/constants.h (parameters.h:41-366). */
int pic_params_default(pic_params *p);

/* Evaluates the #define lines of a parameters.h-style file (parameters.h:41-366)
 * on top of the defaults; replaces the compile-time configuration step. */
int pic_params_load(const char *path, pic_params *p);

/* pic::pic (pic.cpp:55-221).  restart != 0 loads dump_dir/dataDump_e.txt and
 * dump_dir/dataDump_i1.txt (particle.cpp:359-401); otherwise particles are
 * created by p->loader (particle.cpp:321-482). */
int pic_create(const pic_params *p, int restart, const char *dump_dir, pic_handle **out);
int pic_destroy(pic_handle *h);

/* pic::setRuntime (pic.h) */
int pic_set_runtime(pic_handle *h, int32_t run_time);

/* pic::pic_run (pic.cpp:1040-1279): pre-loop wall pass, then run_time steps of
 * the fused timestep (collisions, gather+push, walls, deposit, rho, phi, E).
 * Returns when the run is complete.  PIC_ERR_CAPACITY: particles were dropped
 * because the particle store or a staging buffer was full (the run went on
 * without them; pic_get_overflow counts them). */
int pic_run(pic_handle *h);

/* The timestep loop of pic_run without its pre-loop wall pass: queues nsteps
 * steps on the handle's stream (no reference counterpart; used by drivers
 * that step a run repeatedly, e.g. a benchmark or a serving loop).  Returns
 * PIC_ERR_CAPACITY when particles dropped by an earlier call are seen (once
 * per increase of the count); pic_sync reports those of the last call. */
int pic_advance(pic_handle *h, int32_t nsteps);

/* pic::dump_data (pic.cpp:1285-1295) -> particle::dump_data
 * (particle.cpp:1361-1382): dir/dataDump_e.txt, dir/dataDump_i1.txt.
 * precision = significant digits (6 = the reference's ostream default). */
int pic_dump_data(pic_handle *h, const char *dir, int32_t precision);

/* Exact-resume checkpoint (no reference counterpart: the reference restarts
 * from the dataDump text above, which rounds every value).  The file holds
 * the parameters, the step, both species' stores as they lie, their
 * counters and charge, rho, phi, E, the tolerance solve's iterate and
 * previous solution, the diagnostics' sums and rows, and a checksum; a run
 * continued from it takes exactly the steps the writer would have taken.
 * pic_checkpoint_params reads the parameters so that the caller can create
 * the handle (pic_create, restart 0) that pic_checkpoint_read fills; that
 * handle must have the writer's parameters apart from run_time,
 * write_graphs and graphs (PIC_ERR_ARG otherwise).  PIC_ERR_FORMAT: not a
 * checkpoint, truncated or corrupt; after a failed read the handle's state
 * is undefined (destroy it). */
int pic_checkpoint_write(pic_handle *h, const char *path);
int pic_checkpoint_params(const char *path, pic_params *out);
int pic_checkpoint_read(pic_handle *h, const char *path);

/* Launch-shape choices of a live run (no reference counterpart; used by the
 * autotuner, autotune.py): particle passes per counting sort of electrons
 * and ions (0: leave) and red-black sweeps per leg of the tolerance solve
 * (0: leave, else 1..3).  Results do not depend on them.
 * pic_get_tuning: {sort interval e, sort interval i, sort key, sweeps}. */
int pic_set_tuning(pic_handle *h, int32_t sort_interval_e, int32_t sort_interval_i,
                   int32_t mg_sweeps);
int pic_get_tuning(pic_handle *h, int32_t *out4);

/* ------------------------------------------------ per-phase entry points */
/* particle::get_particle_electric_field (particle.cpp:870-1085, pic.cl:56-92) */
int pic_species_gather(pic_handle *h, int32_t s);
/* particle::move_particles (particle.cpp:1101-1341, pic.cl:95-149) */
int pic_species_move(pic_handle *h, int32_t s);
/* particle::collect_charge (particle.cpp:663-858, pic.cl:22-52) */
int pic_species_collect_charge(pic_handle *h, int32_t s);
/* fields::grid_rho (fields.cpp:100-143) */
int pic_fields_grid_rho(pic_handle *h);
/* fields::grid_potential (fields.cpp:169-220) */
int pic_fields_grid_potential(pic_handle *h, double phi0X, double phiLX, double phi0Y,
                              double phiLY, double timestep, double phi_period);
/* fields::grid_electric_field (fields.cpp:230-268) */
int pic_fields_grid_electric_field(pic_handle *h);
/* e_collisions::collision (e_collisions.cpp:480-597) */
int pic_e_collisions(pic_handle *h);
/* ion_collisions::collision (ion_collisions.cpp:255-308) */
int pic_ion_collisions(pic_handle *h);
/* (sec_ions.*implement_i_boundary) / (sec_e.*implement_e_boundary)
 * (boundary.h:372-383; pic.cpp:1114-1128): which = 0 ion wall, 1 electron wall */
int pic_boundary(pic_handle *h, int32_t which);
/* particle::inject_more_particles (particle.cpp:491-510): NUMBER_TO_INJECT
 * particles of species s appended unmoved, uniform in [dx, L-dx]^2, each
 * velocity component vt cos(2 pi R2) sqrt|ln R1|; weight 0 in reference mode
 * (particleVec() value-initialises it), 1 in physical mode.  The fused
 * timestep does the same after the electron collisions (pic.cpp:1079-1082). */
int pic_inject_particles(pic_handle *h, int32_t s);
/* particle::contain_particle_growth (particle.cpp:640-653; reduce_population
 * 538-632): above size_to_check particles, every velocity group above its
 * limit drops to its new size by a uniform random choice and its survivors'
 * weights are multiplied by Ngroup / (Ngroup - removed).  Needs
 * pic_params.reduce; the fused timestep runs it every step. */
int pic_contain_particle_growth(pic_handle *h, int32_t s);

/* Accuracy of the last field solve against an FP64 multigrid (no reference
 * counterpart: the error report of the split-precision PIC_MG_TOL solver,
 * which replaces Mglin's fixed-cycle FP64 FMG).  The last solve's folded
 * right-hand side (fields.cpp:169-220) is solved again by the FP64 multigrid
 * alone -- the full multigrid of pic_mglin (mglin.cpp:21-55), then FP64
 * V(1,1) cycles (mglin.cpp:176-196) -- until ||f - Au||_2 <= tol ||f||_2, and
 * compared node by node: out5 = relative L2 error of phi, of Ex, of Ey (the
 * nodes*nodes values), the reference solve's relative residual, its V-cycles
 * after the full multigrid.  Synchronises the handle's stream. */
int pic_field_error(pic_handle *h, double tol, double *out5);
/* *finite = 1 when every value of phi, Ex and Ey is finite, else 0
 * (synchronises the handle's stream; no reference counterpart) */
int pic_field_finite(pic_handle *h, int32_t *finite);
/* pic_field_error(h, tol) after every `every`-th step of pic_run and
 * pic_advance (0: off), one row of 6 doubles per report: step, the solver's
 * relative residual of that step, error of phi, Ex, Ey, the reference
 * solve's relative residual.  pic_get_accuracy_report copies at most
 * cap_rows rows and returns the count in *rows. */
int pic_set_accuracy_report(pic_handle *h, int32_t every, double tol);
int pic_get_accuracy_report(pic_handle *h, double *out, int32_t cap_rows, int32_t *rows);

/* Cross sections the collision kernels use, evaluated on the host in FP64
 * (e_collisions::sigma_Ar_elastic / sigma_Ar_ionize / sigma_H_elastic /
 * sigma_H_ionize, e_collisions.h:192-195; ion_collisions::sigma_Ar_cx /
 * sigma_H_cx, ion_collisions.cpp:170-199).  channel: PIC_XS_*; mode:
 * PIC_MODE_REF (the fits as the reference writes them) or PIC_MODE_PHYSICAL
 * (breakpoints closed, clamped >= 0; excitation and Ar+ isotropic elastic
 * exist only here).  energy in eV (H+ charge exchange: eV as well);
 * *sigma in 1e-20 m^2.  Needs no GPU. */
#define PIC_XS_ELASTIC 0    /* e-neutral elastic                              */
#define PIC_XS_EXCITATION 1 /* e-Ar lumped electronic excitation (physical)   */
#define PIC_XS_IONIZATION 2 /* e-neutral ionisation                           */
#define PIC_XS_CX 3         /* ion-neutral charge exchange                    */
#define PIC_XS_ION_ISO 4    /* Ar+-Ar isotropic elastic (physical)            */
int pic_cross_section(int32_t gas, int32_t channel, int32_t mode, double energy, double *sigma);

/* Step counter used by the per-phase entry points (collision random streams
 * and the ion sub-cycling test); pic_fields_grid_electric_field, the last
 * phase of a step, advances it by one.  No reference counterpart. */
int pic_set_step(pic_handle *h, int64_t step);

/* Standalone Mglin (mglin.cpp:21-55) on the GPU, FP64: u is n*n row-major,
 * holds the rhs on entry and the solution on exit. */
int pic_mglin(double *u, int32_t n, int32_t ncycle, double length);

/* pic_mglin without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  u is a
 * DEVICE pointer to n*n doubles (the rhs on entry, the solution on exit, as
 * pic_mglin); the work is queued on `stream` (a hipStream_t of the current
 * device; NULL = its default stream) and not waited for.  The multigrid
 * levels live in a library-owned static device array of 19.7 MB per GPU
 * (grids up to n = 1025; larger n returns PIC_ERR_CAPACITY), so no call
 * allocates or frees.  Calls that overlap on one GPU share that array and
 * race: order them on one stream.  Results are pic_mglin's, bit for bit. */
int pic_mglin_noalloc(double *u, int32_t n, int32_t ncycle, double length, void *stream);

/* Residual-controlled Poisson solve (no reference counterpart; the
 * PIC_MG_TOL solver of the timestep): A u = f with the 5-point Laplacian of
 * mglin.cpp:133-144, zero walls; u (n*n) is the initial guess on entry and
 * the solution on exit.  FP32 V(1,1) cycles on the correction equation, FP64
 * residual and accumulation, until ||f-Au||_2 <= tol ||f||_2 (interior). */
int pic_poisson_tol(const double *f, double *u, int32_t n, double length, double tol,
                    int32_t max_iter, int32_t *cycles, double *relres);

/* --------------------------------------------------------------- getters */
/* multigrid work of the last solve: V-cycles (tolerance mode) or ncycle,
 * cumulative V-cycles, and the relative residual reached (tolerance mode).
 * Grids up to 65^2 are solved without a host wait during the step: this
 * call then waits for the last solve's verdict if the device has not
 * finished it. */
int pic_get_mg_stats(pic_handle *h, int32_t *last_cycles, int64_t *total_cycles, double *relres);
int pic_particle_size(pic_handle *h, int32_t s, int64_t *n);        /* particle::particle_size */
int pic_get_particles(pic_handle *h, int32_t s, pic_particle *out, int64_t cap); /* 48-B AoS copy */
int pic_get_phi(pic_handle *h, double *out);            /* nodes*nodes (fields::get_phi)      */
int pic_get_rho(pic_handle *h, double *out);            /* nodes*nodes (fields::get_rho)      */
int pic_get_E(pic_handle *h, double *ex, double *ey);   /* (nodes+1)^2 each, stride nodes
                                                           (fields::get_elec_fieldX/Y)        */
int pic_get_q(pic_handle *h, int32_t s, double *out);   /* post-scaled charge (particle::get_q) */
int pic_get_step(pic_handle *h, int64_t *step);
/* particles removed: out5 = left, right, bottom, top wall, removed by a
 * collision (the reference's ionisation rule deletes the parent).  With
 * surface physics on, the wall columns count every particle that reached the
 * wall, the reflected and backscattered ones included (pic_get_wall_counts) */
int pic_get_loss_counts(pic_handle *h, int32_t s, int64_t *out5);
/* collision events of species s: out8[k] counts events of kind k =
 * 1 elastic, 2 excitation, 3 ionisation, 4 charge exchange, 5 isotropic ion
 * elastic, 6 candidates whose total collision frequency exceeded the
 * null-collision bound (probability clipped; 0 in a well-posed run) */
int pic_get_collision_counts(pic_handle *h, int32_t s, int64_t *out8);
/* particles dropped because a store or staging buffer was full (0 normally) */
int pic_get_overflow(pic_handle *h, int64_t *dropped);
int pic_get_nodes(pic_handle *h, int32_t *nodes);
/* message of the handle's last error; NULL handle: the last error of a call
 * that had no handle (or failed before it had one), kept per calling thread */
const char *pic_last_error(pic_handle *h);

/* Native-layout state access (SoA: packed int16 cell (jx | jy<<16), fp32
 * offsets in [0,1), fp32 velocities, uint32 id) for tests and tools. */
int pic_get_state(pic_handle *h, int32_t s, uint32_t *cell, float *ox, float *oy, float *vx,
                  float *vy, float *vz, uint32_t *id, int64_t cap);
int pic_set_state(pic_handle *h, int32_t s, const uint32_t *cell, const float *ox,
                  const float *oy, const float *vx, const float *vy, const float *vz,
                  const uint32_t *id, int64_t n);
int pic_get_qfixed(pic_handle *h, int32_t s, int64_t *out); /* nodes*nodes, 2^-48 units   */
/* per-particle weights (particleVec.h: weight) in pic_get_state order; the
 * charge of a particle is weight * qc (pic.cl:46-49).  Setting them switches
 * the species to per-particle weights. */
int pic_get_weights(pic_handle *h, int32_t s, float *w, int64_t cap);
int pic_set_weights(pic_handle *h, int32_t s, const float *w, int64_t n);
/* out3 = population reductions performed (particle.cpp:648 count_reduce),
 * particles they removed, particles created by the source */
int pic_get_population_stats(pic_handle *h, int32_t s, int64_t *out3);
/* surface events of species s since creation (boundary.cpp:586-824, the
 * reference's count_absorb/count_ref/count_scat/count_sec): out6 = reached a
 * wall, absorbed, reflected, backscattered, emitting true secondaries,
 * secondary electrons created */
int pic_get_wall_counts(pic_handle *h, int32_t s, int64_t *out6);
/* live particles of species s without compacting the store (no reference
 * counterpart; pic_particle_size compacts first) */
int pic_particle_count(pic_handle *h, int32_t s, int64_t *n);
/* device memory the handle allocated (bytes) */
int pic_get_device_bytes(pic_handle *h, int64_t *bytes);
/* upper estimate of the device memory pic_create(p, restart=0) allocates */
int pic_estimate_bytes(const pic_params *p, int64_t *bytes);

/* ------------------------------------------- ensembles and serving */
/* No reference counterpart (the reference runs one configuration per
 * process).  Runs are independent handles, each on its own HIP stream; they
 * share the GPU by being stepped from their own host threads, so one run's
 * kernels fill the device while another waits on its stream.  A run's results
 * are bit-identical alone and in a batch. */

/* pic_advance(runs[k], steps) for every k at once, then each synchronised */
int pic_batch_advance(pic_handle *const *runs, int32_t n, int32_t steps);

typedef struct pic_serve_request {
    pic_params params;   /* the run (created without a restart file)         */
    int32_t steps;       /* fused steps after the pre-loop wall pass         */
    double arrival_s;    /* arrival time, seconds after pic_serve starts     */
} pic_serve_request;

typedef struct pic_serve_result {
    double admit_s, finish_s;   /* admission and completion times [s]         */
    int64_t particle_steps;     /* sum over its steps of the live particles   */
    double step_p50_ms, step_p99_ms;  /* per-step latency (issue to completion) */
    int32_t status;             /* PIC_OK or the run's error                   */
    double phi_error, e_error;  /* last step against the FP64 multigrid
                                   (pic_serve_status), -1 if the run failed    */
} pic_serve_result;

typedef struct pic_serve_stats {
    double wall_s;              /* first arrival window to last completion    */
    int64_t particle_steps;
    double throughput;          /* particle-steps/s over wall_s               */
    double step_p50_ms, step_p99_ms;  /* over every step of every request      */
    int64_t peak_bytes;         /* largest admitted memory estimate in use    */
    int32_t max_concurrent;     /* most runs active at once                   */
} pic_serve_stats;

/* Continuous batching: requests are admitted in arrival order as soon as they
 * have arrived, the estimated memory of the active runs plus theirs fits
 * hbm_budget (pic_estimate_bytes) and fewer than max_active runs are active;
 * each runs to completion on its own thread.  Returns the first run error
 * (every result carries its own status). */
int pic_serve(const pic_serve_request *reqs, int32_t n, int64_t hbm_budget, int32_t max_active,
              pic_serve_result *res, pic_serve_stats *stats);

/* A serving loop as an object (no reference counterpart).  Requests are
 * submitted at any time and queue in FIFO order; a scheduler thread admits
 * the head of the queue whenever the estimated memory of the active runs plus
 * its own fits hbm_budget and fewer than max_active runs are active, so
 * requests that arrive while others run join them as memory and slots free
 * up (continuous batching); each admitted run steps on its own stream and
 * thread, and its results are those of the same run alone, bit for bit.
 * Between steps a run obeys pic_serve_cancel; a run whose potential or field
 * turns non-finite stops with PIC_ERR_NUMERIC, the others go on.  Thread
 * safe; pic_server_destroy cancels what is queued or running and waits. */
typedef struct pic_server pic_server;
#define PIC_REQ_QUEUED 0
#define PIC_REQ_RUNNING 1
#define PIC_REQ_DONE 2
#define PIC_REQ_CANCELLED 3
#define PIC_REQ_FAILED 4
typedef struct pic_serve_status {
    int32_t state;              /* PIC_REQ_*                                        */
    int32_t error;              /* PIC_OK, or the PIC_ERR_* that failed it          */
    int32_t steps_done, steps;  /* steps completed / requested                      */
    double submit_s, admit_s, finish_s; /* seconds since pic_server_create          */
    double step_p50_ms, step_p99_ms;    /* per-step latency, issue to completion    */
    int64_t particle_steps;     /* sum over its steps of the live particles         */
    double phi_error, e_error;  /* DONE: relative L2 error of the last step's phi
                                   and E against the FP64 multigrid converged to
                                   1e-12 (pic_field_error); -1 otherwise            */
    uint64_t fingerprint;       /* DONE: FNV-1a of the final fixed-point charge of
                                   both species and of phi (bit patterns)           */
    int64_t device_bytes;       /* device memory the run allocated                  */
} pic_serve_status;
int pic_server_create(int64_t hbm_budget, int32_t max_active, pic_server **out);
int pic_server_destroy(pic_server *srv);
/* queues a run of `steps` fused steps (after pic_run's pre-loop wall pass) */
int pic_serve_submit(pic_server *srv, const pic_params *p, int32_t steps, int64_t *id);
int pic_serve_poll(pic_server *srv, int64_t id, pic_serve_status *out);
/* a queued request never starts; a running one stops before its next step */
int pic_serve_cancel(pic_server *srv, int64_t id);
/* blocks until the request is DONE, CANCELLED or FAILED (PIC_OK) or timeout_s
 * passes (PIC_ERR_TIMEOUT; timeout_s < 0 waits without limit); out may be NULL */
int pic_serve_wait(pic_server *srv, int64_t id, double timeout_s, pic_serve_status *out);
int pic_get_E32(pic_handle *h, float *out);  /* fp32 (Ex,Ey) pairs, (nodes+1)*nodes+1      */
int pic_set_E32(pic_handle *h, const float *in);

/* ------------------------------------------------------ diagnostics */
/* Batch-mode graph averages and time data (graphs.cpp:412-1369, pic.cpp:
 * 1167-1274), collected on the device inside pic_run / pic_advance over the
 * window [graph_start, graph_start + graph_steps) and every time_data_freq
 * steps (pic_params).  Velocity sums use 64-bit fixed point and histograms
 * integer counts, so the averages are deterministic. */

/* directory where pic_run writes the graph files at the end of the window
 * (if write_graphs) and timeData.txt at the end of the run; NULL or "" (the
 * default) writes nothing.  The reference writes into its working directory. */
int pic_set_output_dir(pic_handle *h, const char *dir);
/* graphs::average_fields: rho, phi, Ex, Ey window sums / NO_GRAPH_STEPS, each
 * nodes*nodes (any may be NULL) */
int pic_get_field_averages(pic_handle *h, double *rho, double *phi, double *ex, double *ey);
/* graphs::average_density: sum of get_q / particle charge over the window /
 * NO_GRAPH_STEPS, nodes*nodes */
int pic_get_density_average(pic_handle *h, int32_t s, double *out);
/* graphs::average_velocities: per-cell mean vx along the centre row and vy
 * along the centre column, / (VTE * NO_GRAPH_STEPS) for both species
 * (graphs.cpp:1114-1121), nodes entries each (either may be NULL) */
int pic_get_velocity_slices(pic_handle *h, int32_t s, double *vx, double *vy);
/* graphs::average_dist1/2: 4 * no_bins values -- first X, first Y, second
 * X, second Y histograms / NO_GRAPH_STEPS */
int pic_get_distributions(pic_handle *h, int32_t s, double *out);
/* timeData rows recorded so far: 8 doubles per row (step, n_e, n_i,
 * phi(POINT1), phi(POINT2), phi(POINT3), q_e(POINT3), q_i(POINT3)) */
int pic_get_time_data(pic_handle *h, double *out, int32_t cap_rows, int32_t *rows);
/* graphs::write_*_to_file (graphs.cpp:1167-1369) into dir: rho_phi_EgridX.txt,
 * rho_phi_EgridY.txt, rho2d.txt, phi2d.txt, egrid2dX.txt, egrid2dY.txt and per
 * species density2d_<t>.txt, av_velocity_<t>.txt, first/second_v_distX/Y_<t>.txt */
int pic_write_graphs(pic_handle *h, const char *dir);
/* timeData.txt (pic.cpp:1262-1274) into dir */
int pic_write_time_data(pic_handle *h, const char *dir);
/* graphs::getPhiMax, getExMax, getEyMax, getQMax of both species
 * (graphs.cpp:1000-1112; pic.cpp:924-975): out5 = max averaged phi, Ex, Ey,
 * electron density, ion density */
int pic_get_graph_max(pic_handle *h, double *out5);

/* -------------------------------------------------- snapshot export */
/* pic::getElectrons / getIons (pic.cpp:901-910) thinned on the device: the
 * particles whose id is a multiple of stride (stride 1: all), as 48-byte
 * records; *n receives the subset size (PIC_ERR_CAPACITY if above cap) */
int pic_get_particle_subset(pic_handle *h, int32_t s, int32_t stride, pic_particle *out,
                            int64_t cap, int64_t *n);
/* dir/snapshot_<step>_phi.txt and dir/snapshot_<step>_<e|i1>.txt (the subset
 * of pic_get_particle_subset, "x y vx vy vz weight" per line): frames for
 * offline plotting in place of the OpenGL views */
int pic_write_snapshot(pic_handle *h, const char *dir, int32_t stride);

/* ------------------------------------------------------------ profiling */
/* Event timing of the push pipeline (collide/gather/push/sort/deposit) and
 * of the field solve, on the handle's stream. */
int pic_profile_enable(pic_handle *h, int32_t on);
/* sums since the last read, n entries written (at most 12): out[0] step ms,
 * [1] push-phase ms (collisions, gather, push, sort, deposit, commits),
 * [2] k_push_scatter ms, [3] field solve ms, [4] k_push_count ms, [5] bin-scan
 * ms, [6] k_push_inplace ms, [7] k_push_scatter launches, [8] k_push_inplace
 * launches (species passes: both species' collisionless passes in one launch,
 * k_push_inplace_pair, count as two), [9..11] 0 */
int pic_profile_read(pic_handle *h, double *out, int32_t n);
/* per-step durations (ms) of the profiled steps since the last read, oldest
 * first; *n receives the count written (at most cap) */
int pic_profile_steps(pic_handle *h, double *out, int32_t cap, int32_t *n);
int pic_sync(pic_handle *h);

#ifdef __cplusplus
}
#endif
#endif
