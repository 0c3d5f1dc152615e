// pic_internal.h -- host/device structures shared by the library's sources.
#pragma once

#include <functional>
#include <vector>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mcc.h"
#include "pic_dev.h"

namespace pic {

// SoA view of one particle buffer (kernel argument, by value)
struct SpView {
    uint32_t *cell;
    float *ox, *oy, *vx, *vy, *vz;
    uint32_t *id;
    float *w;  // per-particle weight factor; null: every particle has weight 1
};

__device__ __forceinline__ float load_w(const SpView &v, uint32_t i) { return v.w ? v.w[i] : 1.0f; }

// Device counters of one species.  n[b] is the particle count of buffer b.
// Counters that every block flushes (losses, holes, collision events) are
// spread over CNT_SLOTS copies, block b adding into copy b mod CNT_SLOTS:
// atomics of all blocks on one address serialise at the memory side.  The
// host sums the copies.
#define CNT_SLOTS 64
struct SpCounters {
    uint32_t n[2];
    uint32_t next_id;        // id of the next particle created
    uint32_t overflow;       // particles dropped for lack of capacity
    uint32_t stage_count;    // entries in the staging buffer
    uint32_t pad_;
    // left, right, bottom, top, removed by collision, -
    uint32_t lost[CNT_SLOTS][6];
    uint32_t holes[CNT_SLOTS];  // empty slots left by in-place steps (current buffer)
    unsigned long long events[CNT_SLOTS][8];  // collision events by mcc::Event
    uint32_t reductions;     // population reductions performed
    uint32_t wall_count;     // entries in the wall queue
    // surface events: reached a wall, absorbed, reflected, backscattered,
    // emitting true secondaries, secondary electrons created
    uint32_t wall[CNT_SLOTS][6];
};

__device__ __forceinline__ uint32_t cnt_slot() { return blockIdx.x & (CNT_SLOTS - 1); }

// Particles created during a step (ionisation products, re-injected pairs)
// wait in a staging buffer; a commit pass sorts them by their creator's key
// (so the layout and the new ids are deterministic), appends them to a
// particle buffer and optionally deposits their charge.
// Candidates of the reference's steady walls for one species and one axis:
// bucket counts, bucket starts (2 nb + 1) and scatter cursors over nb id
// ranges per side; the candidates (id, index into the store) in bucket
// order, then sorted by id within each bucket.  cap = the species' store.
struct SteadyBufs {
    uint32_t *hist, *start, *cursor;
    uint32_t *id_in, *idx_in, *id_out, *idx_out;
    uint32_t nb;
};
struct SteadyArgs {
    SpView sp[2];
    const uint32_t *n[2];
    const SpCounters *cnt[2];
    SteadyBufs b[2];
    int64_t cap[2];
};

struct StageView {
    uint32_t *cell;
    float *ox, *oy, *vx, *vy, *vz;
    float *w;                 // weight factor (null when the species is unweighted)
    unsigned long long *key;  // creator key (source << 32 | creator id)
    uint32_t *idx;            // sort permutation
    uint32_t cap;             // entries (power of two)
};

// Multigrid level storage on the device.  Level l has (2^(l+1)+1)^2 nodes;
// level ng-1 is the finest.  Levels <= lc are solved inside one workgroup.
template <typename T>
struct MgArrays {
    T *rhsF[16] = {};  // full-multigrid right-hand side per level
    T *u[16] = {};     // full-multigrid solution per level (u[ng-1]: result)
    T *cres[16] = {};  // V-cycle coarse right-hand side
    T *cv[16] = {};    // V-cycle coarse correction
    T *us[16] = {};    // pre-smoothed iterate of the fused V-cycle (fp32 levels > lc)
};

// host_flags[2] of the tolerance solve: residual passes published so far,
// and the convergence bit
#define MG_PUB_COUNT 0x3FFFFFFF
#define MG_PUB_CONV 0x40000000

struct MgLevels {
    int ng = 0, lc = 0;
    long lds = 0;        // LDS bytes a workgroup may allocate on this device
    MgArrays<double> d;  // FP64 full multigrid (fixed mode) and the FP64 iterate
    MgArrays<float> f;   // FP32 correction solves (tolerance mode)
    // tolerance mode state (device): flags[0] done, flags[1] V-cycles applied,
    // flags[3] residual passes published ([2] unused)
    int *flags = nullptr;
    double *partials = nullptr;  // per-block sums of r^2 and f^2
    double *ualt = nullptr;      // second FP64 iterate: cycles alternate u <-> ualt
    float *r32alt = nullptr;     // second FP32 residual: cycles alternate with f.rhsF[top]
    double *uprev = nullptr;     // the previous step's solution (extrapolated start)
    bool have_prev = false;      // uprev holds a solution
    int nparts = 0;
    double *relres = nullptr;    // pinned host: last relative residual
    int *host_flags = nullptr;   // pinned host copy of flags
    int pred = 4;                // predicted V-cycles per solve
    int pending = 0;             // passes a queued one-workgroup solve publishes (0: none)
    // exact coarsest-level solve of the tolerance mode: discrete sine
    // transform tables of level lc, M x M with M = 2^(lc+1) (padded by one)
    float *dst_S = nullptr;      // S[p][q] = sin(pi (p+1)(q+1) / M); M = 64: the folded halves
    float *dst_D = nullptr;      // (2/M)^2 / (lambda_p + lambda_q), h^2 applied in the kernel
                                 // (M = 64: rows and columns in the folded, split order)
    int nu = 2;                  // red-black sweeps per leg of the fused V-cycle
    hipEvent_t done = nullptr;   // end of a queued batch of cycles (host wait)
    // tolerance mode: a batch of k cycles from a given parity and iterate
    // buffer is replayed from a captured hipGraph (captured on first use)
    bool use_graphs = false;
    struct CycleGraph {
        int k, parity;
        bool first;  // the batch after the solve's initial residual pass
        const double *u0;
        hipGraphExec_t exec;
    };
    std::vector<CycleGraph> graphs;
};

// where created particles go during a pass (see k_particles.hip)
struct Spawn {
    StageView se, si;
    SpCounters *ce, *ci;
    int plop;
    float se_sigma, si_sigma;
    int plop_ref;
    // surface physics: particles reaching a wall are queued (state after the
    // push, key = id) for the wall pass (k_wall.hip) instead of dropped
    int wall;
    StageView wq;
    SpCounters *cw;
    // re-injection after the pass: a particle lost through a wall queues only
    // its id and weight in wq (k_plop_drain samples the pair)
    int lost_ids;
    // deferred collisions (in-place pass): a candidate is queued here by
    // store index, untouched, for k_collide_q to collide, push and deposit
    uint32_t *mq, *mq_cnt;
    uint32_t mq_seg;
};

// ---- k_particles.hip
// words between consecutive counters of the sort histogram (binCount): each
// key's counter has a 128-byte line of its own, so the per-block atomics of
// different keys do not queue on a shared line
#define BIN_STRIDE 32
void launch_key_count(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                      const PushConst &pc, uint32_t *binCount, int do_push);
void launch_push_scatter(hipStream_t st, const SpView &src, const SpView &dst, const uint32_t *n,
                         int64_t cap, const float2 *E, const PushConst &pc, const mcc::MccConst &mc,
                         uint32_t step, uint32_t *binCount, const uint32_t *binStart,
                         unsigned long long *q, int do_push, int deposit, const Spawn &spn,
                         unsigned long long *events, uint32_t *lost6, uint32_t *holes);
void launch_push_inplace(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                         const float2 *E, const PushConst &pc, const mcc::MccConst &mc,
                         uint32_t step, unsigned long long *q, int deposit, const Spawn &spn,
                         unsigned long long *events, uint32_t *lost6, uint32_t *holes,
                         uint32_t *hist = nullptr);
// one species' arguments of an in-place pass (launch_push_inplace's)
struct InplaceArgs {
    SpView sp;
    const uint32_t *n;
    int64_t cap;
    PushConst pc;
    mcc::MccConst mc;
    unsigned long long *q;
    Spawn spn;
    unsigned long long *events;
    uint32_t *lost6, *holes, *hist;
};
// both species' in-place passes in one launch, for two passes that
// inplace_pair_ok accepts (both collisionless: the PV_PLAIN variant)
bool inplace_pair_ok(const InplaceArgs &a, const InplaceArgs &b);
void launch_push_inplace_pair(hipStream_t st, const InplaceArgs &a, const InplaceArgs &b,
                              const float2 *E, uint32_t step, int deposit);
// candidate queue of the collisions run ahead: MQ_SEGS segments of
// collide_seg_cap(cap) indices, one counter per segment every MQ_STRIDE words
#define MQ_SEGS 1024
#define MQ_STRIDE 32
uint32_t collide_seg_cap(int64_t cap);
void launch_collide(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                    const float2 *E, const PushConst &pc, const mcc::MccConst &mc, uint32_t step,
                    const Spawn &spn, unsigned long long *events, uint32_t *lost6, uint32_t *mq,
                    uint32_t *mq_cnt);
void launch_collide_deferred(hipStream_t st, const SpView &sp, const float2 *E,
                             const PushConst &pc, const mcc::MccConst &mc, uint32_t step,
                             const Spawn &spn, unsigned long long *events, uint32_t *lost6,
                             unsigned long long *q, uint32_t *holes, uint32_t *hist = nullptr);
void launch_plop_drain(hipStream_t st, const StageView &wq, SpCounters *cw, const Spawn &sp,
                       const mcc::MccConst &mc, uint32_t step, int N);
void launch_mcc(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                const PushConst &pc, const mcc::MccConst &mc, uint32_t step, const Spawn &spn,
                unsigned long long *events);
// one species' commit of its staging buffer (k_commit)
struct CommitArgs {
    StageView sv;
    SpCounters *scnt;
    SpView dst;
    uint32_t *n_dst;
    uint32_t cap_dst;
    SpCounters *dcnt;
    int deposit, mode;
    unsigned long long *q;
    // histogram of a pending sort pass of dst (null: none) and its key shape
    uint32_t *hist;
    int tilesX, keyshift;
};
// both species' commits in one launch (one workgroup each)
void launch_commit_pair(hipStream_t st, const CommitArgs &a, const CommitArgs &b, int N);
void launch_commit(hipStream_t st, const StageView &sv, SpCounters *scnt, const SpView &dst,
                   uint32_t *n_dst, int64_t cap_dst, SpCounters *dcnt, int deposit, int N, int mode,
                   unsigned long long *q);
void launch_scan(hipStream_t st, const uint32_t *count, uint32_t *start, uint32_t nbins_pad,
                 uint32_t *partial, uint32_t *n_out);
uint32_t scan_pad(uint32_t nbins);
void launch_gather(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                   const float2 *E, const PushConst &pc, float *exp_, float *eyp);
void launch_move(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                 const float *exp_, const float *eyp, const PushConst &pc);
void launch_deposit(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap, int N,
                    int mode, unsigned long long *q);
void launch_absorb(hipStream_t st, const SpView &src, const SpView &dst, const uint32_t *n,
                   uint32_t *nout, int64_t cap, int N, int mode, uint32_t *lost6,
                   const mcc::MccConst &mc, uint32_t step, const Spawn &spn, uint32_t *bcount,
                   uint32_t *bstart, uint32_t *partial);
// block counts (and starts, + 1) the compaction of a store of `cap` needs
uint32_t absorb_blocks_pad(int64_t cap);
// reference steady-state walls: both axes' pairing passes (k_particles.hip)
void launch_steady(hipStream_t st, const SteadyArgs &a, const mcc::MccConst &mc, uint32_t step,
                   int N, float se, float si, unsigned long long *qe, unsigned long long *qi,
                   int mode);
void launch_load_random(hipStream_t st, const SpView &sp, uint32_t n, uint64_t seed,
                        uint32_t species, int N, double sigma);

// ---- k_reduce.hip: population reduction (particle.cpp:520-653)
#define RD_MAX_GROUPS 8
struct ReduceConst {
    int ng;                           // velocity groups
    uint32_t size_check;              // reduce only above this many particles
    uint32_t species;                 // RNG stream of the selection keys
    uint32_t key0, key1;
    double vlim[RD_MAX_GROUPS];       // group upper speed limits [m/s]
    uint32_t glimit[RD_MAX_GROUPS];   // group size that triggers a reduction
    uint32_t gnew[RD_MAX_GROUPS];     // group size after it
};
struct ReduceState {
    uint32_t count[RD_MAX_GROUPS];    // Ngroup
    uint32_t nrem[RD_MAX_GROUPS];     // particles to remove
    uint32_t rank[RD_MAX_GROUPS];     // radix select: rank within the prefix
    unsigned long long prefix[RD_MAX_GROUPS];
    double gw[RD_MAX_GROUPS];         // survivors' weight factor
    uint32_t active;
    uint32_t hist[RD_MAX_GROUPS * 65536];
};
void launch_reduce(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                   const ReduceConst &rc, ReduceState *state, SpCounters *cnt, uint32_t step,
                   unsigned long long *q, int N, int mode);
void launch_inject(hipStream_t st, const StageView &sv, SpCounters *cnt, const float2 *E,
                   const PushConst &pc, uint32_t key0, uint32_t key1, uint32_t step, uint32_t n_inj,
                   float w, double dx, double L, double vt, uint32_t species, int push,
                   uint32_t *lost6);

// ---- k_wall.hip: secondary emission, reflection, backscatter (boundary.cpp:281-529, 687-756)
struct WallConst {
    int mode;               // PIC_DEV_MODE_REF / PIC_DEV_MODE_PHYS
    int N;                  // nodes per side
    double dx;              // node spacing
    double Lw;              // re-emission wall position: N dx (reference), L (physical)
    double dt;              // boundary::dt (boundary.h:62: DT)
    double me, mi, q;       // ELECTRON_MASS, ION_MASS, CHARGE
    double max_yield, E0, Emax0;  // Vaughan parameters (pic.cpp:459-463)
    double thr, wf;         // THRESHOLD, WORK_FUNCTION [eV]
    int see_e, refl, back, see_i;  // E_IMPACT_SEE, REFLECTION, BACKSCATTER, ION_IMPACT_SEE
    uint32_t key0, key1;
};
void launch_wall(hipStream_t st, int species, const StageView &wq, SpCounters *cw,
                 const StageView &se, SpCounters *ce, const WallConst &wc, uint32_t step);

// ---- k_refcx.hip: the reference's per-column ion charge exchange
struct RefCxBuffers {
    uint32_t *id[2] = {};             // eligible ions' ids: bucket order, then sorted
    unsigned long long *val[2] = {};  // probability bits << 32 | store index
    uint32_t *hist = nullptr;         // N nb bucket counts (zero between steps)
    uint32_t *start = nullptr;        // N nb + 1 bucket starts
    uint32_t *cursor = nullptr;       // N nb scatter cursors
    uint32_t nb = 1;                  // id ranges per column (refcx_ranges)
    float *pcol = nullptr;            // probabilityCX[N], kept across steps
};
uint32_t refcx_ranges(int N);
void launch_refcx(hipStream_t st, const SpView &sp, const uint32_t *n, uint32_t cap, int N,
                  const mcc::MccConst &mc, uint32_t step, RefCxBuffers &B, const SpCounters *cnt,
                  unsigned long long *events);

// ---- k_diag.hip: batch-mode diagnostics
#define DIAG_TCOLS 8  // step, n_e, n_i, phi(P1), phi(P2), phi(P3), q_e(P3), q_i(P3)
struct Probe6 {
    int x0, y0, x1, y1, x2, y2;  // POINT1..3 (x, y)
};
struct DiagDev {
    double *rho = nullptr, *phi = nullptr, *ex = nullptr, *ey = nullptr;  // N^2 window sums
    double *q[2] = {};                      // N^2: density (charge / particle charge) sums
    long long *vsx[2] = {}, *vsy[2] = {};   // N: this step's fixed-point velocity sums
    uint32_t *cntx[2] = {}, *cnty[2] = {};  // N: this step's counts
    double *vpx[2] = {}, *vpy[2] = {};      // N: window sums of per-cell mean velocity
    unsigned long long *dist[2] = {};       // 4 x no_bins: X1, Y1, X2, Y2 counts
    double *tdata = nullptr;                // rows x DIAG_TCOLS
};
void launch_diag_step(hipStream_t st, DiagDev &D, bool graphs, bool time_row, int row,
                      long long step, int N, const double *rho, const double *phi,
                      const double *ex64, const double *ey64, const unsigned long long *qe,
                      const unsigned long long *qi, double scale_e, double scale_i, double corner,
                      double qc_e, double qc_i, const SpView &e, const uint32_t *ne, int64_t cap_e,
                      const SpView &ion, const uint32_t *ni, int64_t cap_i, double vt_e,
                      double vt_i, int no_bins, int divide_bins_by, int spread, int4 loc,
                      const SpCounters *cnt, int cur_e, int cur_i, Probe6 probe);

void launch_subset(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap_sp,
                   uint32_t stride, double dx, double weight, double *out, uint32_t cap,
                   uint32_t *count);

// ---- k_fields.hip
void launch_rho(hipStream_t st, const unsigned long long *qi, const unsigned long long *qe, int N,
                double scale_i, double scale_e, double corner, double *rho);
void launch_fold(hipStream_t st, const double *rho, int N, double coef, double phi0X, double phiR,
                 double phi0Y, double phiLY, double dxdy_inv, double *rhs);
void launch_phi_bc(hipStream_t st, const double *u, int N, double phi0X, double phiR,
                   double phi0Y, double phiLY, double *phi);
// u1, flags (may be null): the iterate is u1 when flags[1] is odd
void launch_phi_efield(hipStream_t st, const double *u, int N, double phi0X, double phiR,
                       double phi0Y, double phiLY, double dx, double *phi, double *ex64,
                       double *ey64, float2 *E32, const double *u1 = nullptr,
                       const int *flags = nullptr);
// the fused timestep's field set-up per node: rho (k_rho), the folded rhs
// (k_fold), the tolerance solve's start (extrap 1: uprev <- u; 2: also
// u <- 2 u - uprev), with flags the solve's flag reset, and zi, ze (the
// charge buffers of the next step, N*N each) cleared
void launch_field_start(hipStream_t st, const unsigned long long *qi, const unsigned long long *qe,
                        int N, double scale_i, double scale_e, double corner, double *rho,
                        double coef, double phi0X, double phiR, double phi0Y, double phiLY,
                        double dxdy_inv, double *rhs, double *u, double *uprev, int extrap,
                        int *flags, unsigned long long *zi, unsigned long long *ze);
void launch_efield(hipStream_t st, const double *phi, int N, double dx, double *ex64, double *ey64,
                   float2 *E32);
template <typename T>
void launch_fill(hipStream_t st, T *p, long n, T v);
int mg_coarse_level_for(long lds_bytes, int ng);
int mg_dst_tables(MgLevels &M);
void mg_set_coarse_attr(long lds_bytes);
void mg_fixed_solve(hipStream_t st, MgLevels &M, int ncycle, double length);
// FP64 reference solve of R.d.rhsF[top] into R.d.u[top]: full multigrid, then
// FP64 V(1,1) cycles to ||f - Au|| <= tol ||f|| (at most max_cycles);
// partials: device, >= 2 nb doubles; host: >= 6 nb doubles.  Returns the
// V-cycles after the full multigrid (-1 on a stream error); synchronises.
int mg_fp64_reference(hipStream_t st, MgLevels &R, double length, double tol, int max_cycles,
                      double *partials, double *host, int nb, double *relres);
// adds the non-finite values among a[0..na) and b, c[0..nbc) to *count
void launch_count_nonfinite(hipStream_t st, const double *a, long na, const double *b,
                            const double *c, long nbc, unsigned *count);
// sums of (a_k - b_k)^2 and b_k^2 over n entries for three pairs (out6);
// partials: device, >= 6 nb doubles; synchronises
int field_diff_sums(hipStream_t st, const double *a0, const double *b0, const double *a1,
                    const double *b1, const double *a2, const double *b2, long n,
                    double *partials, double *host, int nb, double *out6);
// residual-controlled solve of A u = M.d.rhsF[ng-1] into M.d.u[ng-1] (warm
// started from its current content); returns V-cycles used, -1 on HIP error
// after_batch(u0, u1) is queued after each batch of cycles, before the host
// waits for it; the final iterate is u1 when the applied cycle count
// (flags[1]) is odd, else u0.  flags_reset: the caller has cleared flags.
int mg_tol_solve(hipStream_t st, MgLevels &M, double length, double tol, int max_iter,
                 double *relres_out, bool flags_reset = false,
                 const std::function<void(const double *, const double *)> &after_batch = nullptr);
// A grid whose finest level is the DST level (lf == lc) is solved whole by
// one kernel that leaves the final iterate in M.d.u[ng-1]: mg_tol_solve then
// queues it, and after_batch, without waiting and returns MG_PENDING;
// mg_tol_verdict waits for that solve's verdict and returns its V-cycles (-1
// on a HIP error or with no solve pending).  mg_tol_solve reads a pending
// verdict itself (and drops it) before it starts another solve.
constexpr int MG_PENDING = -2;
int mg_tol_verdict(MgLevels &M, double *relres_out);


}  // namespace pic
