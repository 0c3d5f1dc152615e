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
    unsigned long long *q = nullptr;  // fixed-point charge, N*N
    float *exp_ = nullptr, *eyp = nullptr;  // per-phase gather output
    bool gathered = false;
    double q_over_m = 0, qc = 0, mass = 0, dt = 0, weight = 1.0, sigma_load = 0;
    double scale = 0;  // qc/(dx*dy) * weight * 2^-40
    PushConst pc = {};
    mcc::MccConst mc = {};
    pic::StageView stage = {};  // particles created during a step
    int sort_interval = 1;  // particle passes per counting sort
    int since_sort = 0;   // in-place steps since the last sort pass
    bool holes = false;   // the current buffer may hold empty slots
    const char *type = "";  // dataDump suffix (parameters.h:316-318)
};

struct ProfPair {
    hipEvent_t a, b;
    int tag;
};

}  // namespace pic_host

struct pic_handle {
    pic_params P{};
    int N = 0;
    double dx = 0;
    int tilesX = 0;
    uint32_t nbins = 0, nbins_pad = 0;
    hipStream_t st = nullptr;
    pic_host::Species sp[2];
    pic::SpCounters *cnt = nullptr;       // device, [2]
    pic::SpCounters *cnt_host = nullptr;  // pinned mirror
    uint32_t *binCount = nullptr, *binStart = nullptr, *scanPartial = nullptr;
    double *rho = nullptr, *phi = nullptr, *ex64 = nullptr, *ey64 = nullptr;
    float2 *E32 = nullptr;
    float *ntab = nullptr;       // reference neutral velocity table (collisions.cpp:66-89)
    bool spawning = false;       // collisions or re-injection create particles
    pic::MgLevels M;
    bool refcx = false;         // the reference's per-column ion charge exchange
    pic::RefCxBuffers cx;
    double last_phiR = 0.0;
    int mg_cycles = 0;          // V-cycles (tolerance mode) of the last solve
    int64_t mg_cycles_total = 0;
    double mg_relres = 0.0;     // relative residual after the last solve
    bool have_rho = false, have_phi = false;
    int64_t step = 0;
    int32_t run_time = 0;
    std::string err;
    // profiling
    bool prof = false;
    std::vector<pic_host::ProfPair> prof_live, prof_pool;
    double prof_acc[12] = {};
    std::vector<double> prof_steps;  // per-step durations (ms)
};
