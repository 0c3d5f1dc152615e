// pic_host.h -- host-side state behind the C ABI handle.
#pragma once

#include <string>
#include <vector>

#include "../../include/pic2d.h"
#include "pic_internal.h"

namespace pic_host {

int set_global_error(int code, const std::string &msg);
const char *global_error();

// One particle species: double-buffered SoA store on the device.
struct Species {
    pic::SpView buf[2] = {};
    int cur = 0;
    int64_t cap = 0;
    unsigned long long *q = nullptr;  // fixed-point charge, N*N: qbuf[qcur]
    // a timestep deposits into the buffer its start switches to; the field
    // step zeroes the other one for the next step (it is zero whenever a
    // step starts)
    unsigned long long *qbuf[2] = {};
    int qcur = 0;
    float *exp_ = nullptr, *eyp = nullptr;  // per-phase gather output
    bool gathered = false;
    double q_over_m = 0, qc = 0, mass = 0, dt = 0, weight = 1.0, sigma_load = 0;
    double scale = 0;  // qc/(dx*dy) * weight * 2^-48
    PushConst pc = {};
    mcc::MccConst mc = {};
    pic::StageView stage = {};  // particles created during a step
    int sort_interval = 1;  // particle passes per counting sort
    int since_sort = 0;   // in-place steps since the last sort pass
    bool holes = false;   // the current buffer may hold empty slots
    // the last in-place pass (and the commits after it) counted the keys of
    // the next sort pass into this species' histogram
    bool hist_pending = false;
    const char *type = "";  // dataDump suffix (parameters.h:316-318)
    // per-particle weights (buf[b].w, stage.w): needed once weights can
    // differ between particles (source, reduction, a dump with mixed
    // weights); then `weight` is 1 and w holds each particle's weight
    bool weighted = false;
    bool reduce = false;             // contain_particle_growth every step
    pic::ReduceConst rc = {};
    pic::ReduceState *rstate = nullptr;
    int64_t injected = 0;            // particles created by the source
    pic::StageView wq = {};          // particles that reached a wall (surface physics)
    // candidate queue and its segment counters (collisions run ahead)
    uint32_t *mq = nullptr, *mq_cnt = nullptr;
    float2 *mc_bins = nullptr;  // energy-bin null-collision bounds (mc.bins)
    pic::SteadyBufs steady = {};  // the reference's steady walls (REF mode)
};

struct ProfPair {
    hipEvent_t a, b;
    int tag;
};

}  // namespace pic_host

struct pic_handle;
namespace pic_host {
// diag.cpp
bool diag_graphs_on(const pic_handle *h);
bool diag_time_on(const pic_handle *h);
int diag_alloc(pic_handle *h);
void diag_free(pic_handle *h);
int diag_prepare_time(pic_handle *h, int32_t last_step);
void diag_step(pic_handle *h, int64_t t);
int diag_fields(pic_handle *h, double *rho, double *phi, double *ex, double *ey);
int diag_density(pic_handle *h, int s, double *out);
int diag_velocities(pic_handle *h, int s, double *vx, double *vy);
int diag_distributions(pic_handle *h, int s, double *out);
int diag_time_rows(pic_handle *h, std::vector<double> &rows, int &n);
int diag_write_graphs(pic_handle *h, const std::string &dir);
int diag_write_time(pic_handle *h, const std::string &dir);
}  // namespace pic_host

struct pic_handle {
    pic_params P{};
    int N = 0;
    double dx = 0;
    int tilesX = 0;
    int keyshift = 0;  // sort key = cell key >> keyshift
    uint32_t nbins = 0, nbins_pad = 0;
    hipStream_t st = nullptr;
    pic_host::Species sp[2];
    pic::SpCounters *cnt = nullptr;       // device, [2]
    pic::SpCounters *cnt_host = nullptr;  // pinned mirror
    // particles dropped for lack of capacity: both species' overflow counters
    // copied (asynchronously) after every pic_advance, checked at the next
    // call or at pic_sync; an increase is reported as PIC_ERR_CAPACITY
    uint32_t *ovf_host = nullptr;  // pinned, [2]
    hipEvent_t ovf_ev = nullptr;
    bool ovf_pending = false;
    uint64_t ovf_reported = 0;
    uint32_t *binCount = nullptr, *binStart = nullptr, *scanPartial = nullptr;
    // the absorbing walls' compaction: block counts, starts, scan partials
    uint32_t *abs_count = nullptr, *abs_start = nullptr, *abs_partial = nullptr;
    double *rho = nullptr, *phi = nullptr, *ex64 = nullptr, *ey64 = nullptr;
    float2 *E32 = nullptr;
    float *ntab = nullptr;       // reference neutral velocity table (collisions.cpp:66-89)
    bool spawning = false;       // collisions or re-injection create particles
    pic::MgLevels M;
    bool refcx = false;         // the reference's per-column ion charge exchange
    bool wall_on = false;       // surface physics at the non-steady walls
    pic::WallConst wc = {};
    pic::RefCxBuffers cx;
    // batch-mode diagnostics
    pic::DiagDev diag;
    int diag_rows = 0, diag_rows_cap = 0;  // timeData rows written / allocated
    int64_t diag_graph_steps = 0;          // steps summed into the graph window
    std::string out_dir;                   // where pic_run writes its files ("" = none)
    double last_phiR = 0.0;
    double last_bc[4] = {0, 0, 0, 0};  // phi0X, phi_right(t), phi0Y, phiLY of the last solve
    // accuracy report: FP64 reference solve of the last right-hand side
    pic::MgLevels ref;                 // FP64 levels only (allocated on first use)
    double *ref_phi = nullptr, *ref_ex = nullptr, *ref_ey = nullptr;
    float2 *ref_E32 = nullptr;
    double *ref_partials = nullptr;    // device block sums
    std::vector<double> ref_host;      // their host copy
    int32_t acc_every = 0;             // report every acc_every steps (0: off)
    double acc_tol = 1e-12;
    std::vector<double> acc_rows;      // step, solver relres, err phi, Ex, Ey, ref relres
    uint32_t *nonfinite = nullptr;     // device counter of pic_field_finite
    int mg_cycles = 0;          // V-cycles (tolerance mode) of the last solve
    int64_t mg_cycles_total = 0;
    double mg_relres = 0.0;     // relative residual after the last solve
    bool have_rho = false, have_phi = false;
    int64_t step = 0;
    int64_t dev_bytes = 0;      // device memory allocated through dalloc
    int32_t run_time = 0;
    std::string err;
    // profiling
    bool prof = false;
    std::vector<pic_host::ProfPair> prof_live, prof_pool;
    double prof_acc[12] = {};
    std::vector<double> prof_steps;  // per-step durations (ms)
};
