// engine.cpp -- the C ABI: handle lifecycle, the fused timestep, the
// reference-shaped per-phase entry points, getters and dataDump I/O.
//
// One handle = one simulation (the reference's pic object, pic.h:89-214).
// All state lives in device memory; every call is enqueued on the handle's
// stream.  Getters synchronise the stream and copy into caller buffers.
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "pic_host.h"
#include "tuned.h"

using namespace pic;
using namespace pic_host;

namespace pic_host {
// the last error of a call without a handle, per calling thread: the
// pointer pic_last_error(NULL) returns stays valid until that thread's next
// failing call, whatever other threads do meanwhile
static thread_local std::string g_err;
int set_global_error(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
const char *global_error() { return g_err.c_str(); }
}  // namespace pic_host

namespace {

int seterr(pic_handle *h, int code, const std::string &m) {
    if (h) h->err = m;
    return set_global_error(code, m);
}

#define HIPCHK(h, expr)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return seterr((h), PIC_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CHECK(expr)              \
    do {                         \
        int rc_ = (expr);        \
        if (rc_ != PIC_OK) return rc_; \
    } while (0)

template <class T>
int dalloc(pic_handle *h, T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess)
        return seterr(h, PIC_ERR_ALLOC, std::string("hipMalloc ") + std::to_string(n * sizeof(T)) +
                                            " bytes: " + hipGetErrorString(e));
    h->dev_bytes += (int64_t)(n * sizeof(T));
    return PIC_OK;
}

bool grid_ok(int n) {
    if (n < 5) return false;
    int m = n - 1;
    return (m & (m - 1)) == 0;
}

int levels_of(int n) {
    int ng = 0, nn = n;
    while (nn >>= 1) ng++;
    return ng;
}

long lds_limit() {
    int dev = 0, v = 0;
    hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        v <= 0)
        v = 65536;
    return v;
}

int alloc_mg(pic_handle *h, MgLevels &M, int N) {
    M.ng = levels_of(N);
    long lds = lds_limit();
    M.lds = lds;
    M.lc = mg_coarse_level_for(lds, M.ng);
    mg_set_coarse_attr(lds);
    if (mg_dst_tables(M) != 0) return seterr(h, PIC_ERR_ALLOC, "multigrid: coarse tables");
    for (int l = M.lc; l < M.ng; l++) {
        size_t s = (size_t)((1 << (l + 1)) + 1) * ((1 << (l + 1)) + 1);
        CHECK(dalloc(h, &M.d.rhsF[l], s));
        CHECK(dalloc(h, &M.d.u[l], s));
        CHECK(dalloc(h, &M.f.rhsF[l], s));
        CHECK(dalloc(h, &M.f.u[l], s));
        if (l > M.lc) CHECK(dalloc(h, &M.f.us[l], s));
        if (l < M.ng - 1) {
            CHECK(dalloc(h, &M.d.cres[l], s));
            CHECK(dalloc(h, &M.d.cv[l], s));
            CHECK(dalloc(h, &M.f.cres[l], s));
            CHECK(dalloc(h, &M.f.cv[l], s));
        }
    }
    size_t nf = (size_t)N * N;
    HIPCHK(h, hipMemset(M.d.u[M.ng - 1], 0, nf * sizeof(double)));  // cold start of the iterate
    M.nparts = (int)std::min<size_t>(1024, (nf + 255) / 256);
    // block partial sums: the residual pass's blocks, or one per tile of the
    // top level (the fused post-smoothing and residual kernel; tiles of 16 x
    // 16 or larger)
    const size_t tiles = (size_t)((N + 15) / 16) * (size_t)((N + 15) / 16);
    CHECK(dalloc(h, &M.partials, 2 * std::max((size_t)M.nparts, tiles)));
    CHECK(dalloc(h, &M.r32alt, nf));
    CHECK(dalloc(h, &M.uprev, nf));
    CHECK(dalloc(h, &M.flags, 4));
    HIPCHK(h, hipMemset(M.flags, 0, 4 * sizeof(int)));
    CHECK(dalloc(h, &M.ualt, nf));
    // coherent: the device's writes reach the host while kernels still run
    HIPCHK(h, hipHostMalloc((void **)&M.relres, sizeof(double), hipHostMallocCoherent));
    HIPCHK(h, hipHostMalloc((void **)&M.host_flags, 4 * sizeof(int), hipHostMallocCoherent));
    *M.relres = 0.0;
    M.host_flags[0] = M.host_flags[1] = M.host_flags[2] = M.host_flags[3] = 0;
    HIPCHK(h, hipEventCreateWithFlags(&M.done, hipEventDisableTiming));
    return PIC_OK;
}

void free_mg(MgLevels &M) {
    for (int l = 0; l < 16; l++) {
        void *ptrs[] = {M.d.rhsF[l], M.d.u[l], M.d.cres[l], M.d.cv[l],
                        M.f.rhsF[l], M.f.u[l], M.f.cres[l], M.f.cv[l], M.f.us[l]};
        for (void *p : ptrs)
            if (p) hipFree(p);
    }
    M.d = MgArrays<double>();
    M.f = MgArrays<float>();
    if (M.dst_S) hipFree(M.dst_S);
    if (M.dst_D) hipFree(M.dst_D);
    M.dst_S = M.dst_D = nullptr;
    if (M.partials) hipFree(M.partials);
    if (M.flags) hipFree(M.flags);
    if (M.ualt) hipFree(M.ualt);
    M.ualt = nullptr;
    if (M.r32alt) hipFree(M.r32alt);
    M.r32alt = nullptr;
    if (M.uprev) hipFree(M.uprev);
    M.uprev = nullptr;
    M.have_prev = false;
    if (M.relres) hipHostFree(M.relres);
    if (M.host_flags) hipHostFree(M.host_flags);
    if (M.done) hipEventDestroy(M.done);
    M.done = nullptr;
    for (MgLevels::CycleGraph &g : M.graphs) hipGraphExecDestroy(g.exec);
    M.graphs.clear();
    M.partials = nullptr;
    M.flags = nullptr;
    M.relres = nullptr;
    M.host_flags = nullptr;
}

// ----------------------------------------------------------- profiling
int prof_begin(pic_handle *h, int tag) {
    if (!h->prof) return -1;
    ProfPair p;
    if (!h->prof_pool.empty()) {
        p = h->prof_pool.back();
        h->prof_pool.pop_back();
    } else {
        hipEventCreate(&p.a);
        hipEventCreate(&p.b);
    }
    p.tag = tag;
    hipEventRecord(p.a, h->st);
    h->prof_live.push_back(p);
    return (int)h->prof_live.size() - 1;
}
void prof_end(pic_handle *h, int idx) {
    if (idx < 0) return;
    hipEventRecord(h->prof_live[idx].b, h->st);
}
void prof_flush(pic_handle *h) {
    hipStreamSynchronize(h->st);
    for (ProfPair &p : h->prof_live) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, p.a, p.b);
        if (p.tag == 9) {  // both species' in-place passes in one launch
            h->prof_acc[6] += ms;
            h->prof_acc[8] += 2.0;
        } else {
            h->prof_acc[p.tag] += ms;
        }
        if (p.tag == 2) h->prof_acc[7] += 1.0;
        if (p.tag == 6) h->prof_acc[8] += 1.0;
        if (p.tag == 0) h->prof_steps.push_back(ms);
        h->prof_pool.push_back(p);
    }
    h->prof_live.clear();
}

// ------------------------------------------------------------ species
int species_alloc(pic_handle *h, Species &s, int64_t cap) {
    s.cap = cap;
    for (int b = 0; b < 2; b++) {
        CHECK(dalloc(h, &s.buf[b].cell, cap));
        CHECK(dalloc(h, &s.buf[b].ox, cap));
        CHECK(dalloc(h, &s.buf[b].oy, cap));
        CHECK(dalloc(h, &s.buf[b].vx, cap));
        CHECK(dalloc(h, &s.buf[b].vy, cap));
        CHECK(dalloc(h, &s.buf[b].vz, cap));
        CHECK(dalloc(h, &s.buf[b].id, cap));
    }
    for (int b = 0; b < 2; b++) {
        CHECK(dalloc(h, &s.qbuf[b], (size_t)h->N * h->N));
        HIPCHK(h, hipMemsetAsync(s.qbuf[b], 0, sizeof(unsigned long long) * h->N * h->N, h->st));
    }
    s.qcur = 0;
    s.q = s.qbuf[0];
    return PIC_OK;
}

void species_free(Species &s) {
    for (int b = 0; b < 2; b++) {
        void *ptrs[] = {s.buf[b].cell, s.buf[b].ox, s.buf[b].oy, s.buf[b].vx,
                        s.buf[b].vy,   s.buf[b].vz, s.buf[b].id};
        for (void *p : ptrs)
            if (p) hipFree(p);
    }
    for (int b = 0; b < 2; b++)
        if (s.buf[b].w) hipFree(s.buf[b].w);
    if (s.stage.w) hipFree(s.stage.w);
    void *wq[] = {s.wq.cell, s.wq.ox, s.wq.oy, s.wq.vx, s.wq.vy, s.wq.vz, s.wq.w, s.wq.key};
    for (void *p : wq)
        if (p) hipFree(p);
    if (s.rstate) hipFree(s.rstate);
    if (s.mq) hipFree(s.mq);
    if (s.mq_cnt) hipFree(s.mq_cnt);
    if (s.mc_bins) hipFree(s.mc_bins);
    for (int b = 0; b < 2; b++)
        if (s.qbuf[b]) hipFree(s.qbuf[b]);
    if (s.exp_) hipFree(s.exp_);
    if (s.eyp) hipFree(s.eyp);
    void *st[] = {s.stage.cell, s.stage.ox, s.stage.oy, s.stage.vx,
                  s.stage.vy,   s.stage.vz, s.stage.key, s.stage.idx};
    for (void *p : st)
        if (p) hipFree(p);
    const pic::SteadyBufs &c = s.steady;
    void *sb[] = {c.hist, c.start, c.cursor, c.id_in, c.idx_in, c.id_out, c.idx_out};
    for (void *p : sb)
        if (p) hipFree(p);
    s = Species();
}

void setup_species(pic_handle *h, Species &s, bool electron) {
    const pic_params &P = h->P;
    s.q_over_m = electron ? P.electron_charge / P.electron_mass : P.ion_charge / P.ion_mass;
    s.qc = electron ? P.qce : P.qci;
    s.mass = electron ? P.electron_mass : P.ion_mass;
    s.dt = electron ? P.dt : P.dti;
    s.sigma_load = electron ? P.sigma_load_e : P.sigma_load_i;
    s.type = electron ? "e" : "i1";
    s.weight = 1.0;
    PushConst &pc = s.pc;
    pc.N = h->N;
    pc.mode = P.mode == PIC_MODE_PHYSICAL
                  ? PIC_DEV_MODE_PHYS
                  : (P.wall == PIC_WALL_STEADY ? PIC_DEV_MODE_REF_STEADY : PIC_DEV_MODE_REF);
    pc.bfield = P.magnetic_field > 0.0;
    pc.tilesX = h->tilesX;
    pc.keyshift = h->keyshift;
    double coef = 0.5 * s.dt * s.q_over_m;  // particle.cpp:1103
    double omega_dt = s.dt * s.q_over_m * P.magnetic_field;
    // sort interval: explicit, or chosen from the thermal displacement per
    // step d = sigma dt / dx (cells): about c / d passes, 1..16, c from the
    // tuned table (hand-set 7: 8 passes at d = 0.85, 2 at d = 3.4, 16 for
    // slow ions; the in-place pass loses little to disorder, its deposit
    // window falls back to a fixed square around the block's mean cell)
    if (P.sort_interval > 0) {
        s.sort_interval = P.sort_interval;
    } else {
        const double c = tuned_row(h->N).sort_c[electron ? 0 : 1];
        double d = s.sigma_load * s.dt / h->dx;
        int k = (int)std::lround(c / std::max(d, 0.05));
        s.sort_interval = std::max(1, std::min(16, k));
    }
    pc.cE = (float)coef;
    pc.dtdx = (float)(s.dt / h->dx);
    pc.dtdy = (float)(s.dt / h->dx);
    pc.cos_a = (float)cos(P.angle);
    pc.sin_a = (float)sin(P.angle);
    pc.cos_w = (float)cos(omega_dt);
    pc.sin_w = (float)sin(omega_dt);
}

const double kCharge = 1.602177e-19, kAmu = 1.66054e-27;

// null-collision bound: max over energy of n_g * sigma_total(E) * g(E)
double nu_max_scan(const mcc::MccConst &c) {
    // electrons: 1e-6 .. 3e4 eV; ions: 1e-6 .. 1e3 eV (an ion above the bound
    // would have its collision probability clipped; such events are counted)
    double best = 0.0, top = c.electron ? 4.5 : 3.0;
    for (int k = 0; k <= 200000; k++) {
        double E = pow(10.0, -6.0 + (top + 6.0) * k / 200000.0);
        double g = sqrt(2.0 * c.q * E / c.mass);
        double a, b, d = 0.0;
        if (c.electron)
            mcc::e_nu_parts<double>(c, E, g, a, b, d);
        else
            mcc::i_nu_parts<double>(c, E, g, a, b);
        double nu = a + b + d;
        if (nu > best) best = nu;
    }
    return best * 1.02;
}

// per-bin null-collision bounds of the physical-mode electrons (the bin
// rule is energy_bin's, mcc.h): the largest total collision frequency over
// the relative speeds a bin's particles reach against neutrals up to 10
// sigma fast, scanned at 257 speeds, times 1.02; nu[b] and p[b] = 1 -
// exp(-nu[b] dt)
void electron_bins(const mcc::MccConst &c, double *nu, double *p) {
    for (int b = 0; b < MCC_BINS; b++) {
        const double vlo = sqrt(2.0 * c.q * mcc::mcc_bin_lo(b) / c.mass);
        const double vhi = sqrt(2.0 * c.q * mcc::mcc_bin_hi(b) / c.mass);
        const double glo = std::max(0.0, vlo - 10.0 * c.vtn), ghi = vhi + 10.0 * c.vtn;
        double best = 0.0;
        for (int k = 0; k <= 256; k++) {
            const double g = glo + (ghi - glo) * k / 256.0;
            const double E = 0.5 * c.mass * g * g / c.q;
            double a, e, d;
            mcc::e_nu_parts<double>(c, E, g, a, e, d);
            best = std::max(best, a + e + d);
        }
        nu[b] = best * 1.02;
        p[b] = 1.0 - exp(-nu[b] * c.dt);
    }
}

int setup_mcc(pic_handle *h, Species &s, bool electron) {
    const pic_params &P = h->P;
    mcc::MccConst &c = s.mc;
    c = mcc::MccConst();
    c.mode = P.mode == PIC_MODE_PHYSICAL ? PIC_DEV_MODE_PHYS : PIC_DEV_MODE_REF;
    c.gas = P.gas == PIC_GAS_HYDROGEN ? mcc::GAS_H : mcc::GAS_AR;
    c.electron = electron;
    c.sig_el_energy = P.sigma_elastic == PIC_SIGMA_ENERGY;
    c.sig_iz_energy = P.sigma_ionization == PIC_SIGMA_ENERGY;
    c.sig_cx_energy = P.sigma_cx == PIC_SIGMA_ENERGY;
    c.sig_el_const = (float)P.elastic_sigma_value;
    c.sig_iz_const = (float)P.ionization_sigma_value;
    c.sig_cx_const = (float)P.sigma_value;
    c.N = h->N;
    c.ng = (float)P.neutral_density;
    c.dt = (float)s.dt;
    c.mass = (float)s.mass;
    c.M = (float)P.ion_mass;
    c.q = (float)kCharge;
    c.vtn = (float)P.vtn;
    c.e_exc = 11.55f;
    c.e_iz = c.gas == mcc::GAS_AR ? 15.76f : 13.6f;
    c.opal_B = c.gas == mcc::GAS_AR ? 10.0f : 8.3f;
    c.kev_amu_div = (float)(P.ion_mass / kAmu);
    c.ntab = h->ntab;
    c.ntab_n = 10000;
    c.key0 = (uint32_t)P.seed;
    c.key1 = (uint32_t)(P.seed >> 32);
    c.on = P.have_collisions != 0;
    // reference-mode ions collide through the per-column accumulator (k_refcx.hip)
    if (!electron && c.mode == PIC_DEV_MODE_REF) c.on = 0;
    c.bins = nullptr;
    if (c.on && c.mode == PIC_DEV_MODE_PHYS) {
        double nu = nu_max_scan(c);
        c.nu_max = (float)nu;
        c.p_null = (float)(1.0 - exp(-nu * s.dt));
        if (electron && P.null_bound != PIC_NULL_BOUND_GLOBAL) {
            double nub[MCC_BINS], pb[MCC_BINS];
            electron_bins(c, nub, pb);
            float2 t[MCC_BINS];
            for (int b = 0; b < MCC_BINS; b++) t[b] = make_float2((float)pb[b], (float)nub[b]);
            CHECK(dalloc(h, &s.mc_bins, MCC_BINS));
            HIPCHK(h, hipMemcpy(s.mc_bins, t, sizeof(t), hipMemcpyHostToDevice));
            c.bins = s.mc_bins;
        }
    }
    return PIC_OK;
}

// collisions.cpp:66-89: the reference's neutral velocity table, drawn from
// the seed's Philox stream instead of rand()
void make_neutral_table(const pic_params &P, std::vector<float> &t) {
    const int n = 10000;
    t.assign(3 * n, 0.0f);
    uint32_t k0 = (uint32_t)P.seed, k1 = (uint32_t)(P.seed >> 32);
    for (int i = 0; i < n; i++) {
        uint32_t r[4];
        philox4x32_10((uint32_t)i, 0x7AB1Eu, 0u, 0x4E455554u, k0, k1, r);
        double R1 = u01d(r[0], r[1]), R2 = u01d(r[2], r[3]);
        t[i] = (float)(P.vtn * cos(3.141592654 * R2) * sqrt(fabs(log(R1))));
        philox4x32_10((uint32_t)i, 0x7AB1Eu, 1u, 0x4E455554u, k0, k1, r);
        R1 = u01d(r[0], r[1]);
        R2 = u01d(r[2], r[3]);
        t[n + i] = (float)(P.vtn * cos(3.141592654 * R2) * sqrt(fabs(log(R1))));
        t[2 * n + i] = (float)(P.vtn * sin(3.141592654 * R2) * sqrt(fabs(log(R1))));
    }
}

int alloc_stage(pic_handle *h, Species &s, int64_t n0, int64_t extra) {
    int64_t want = std::max<int64_t>(16384, std::min<int64_t>(n0 / 16, (int64_t)1 << 22)) + extra;
    uint32_t cap = 1;
    while ((int64_t)cap < want) cap <<= 1;
    StageView &v = s.stage;
    v.cap = cap;
    CHECK(dalloc(h, &v.cell, cap));
    CHECK(dalloc(h, &v.ox, cap));
    CHECK(dalloc(h, &v.oy, cap));
    CHECK(dalloc(h, &v.vx, cap));
    CHECK(dalloc(h, &v.vy, cap));
    CHECK(dalloc(h, &v.vz, cap));
    CHECK(dalloc(h, &v.key, cap));
    CHECK(dalloc(h, &v.idx, cap));
    return PIC_OK;
}

// si >= 0: the pass moves species si; its particles that reach a wall are
// queued for the surface pass when surface physics is on
Spawn make_spawn(pic_handle *h, bool plop, int si) {
    Spawn sp;
    sp.se = h->sp[0].stage;
    sp.si = h->sp[1].stage;
    sp.ce = &h->cnt[0];
    sp.ci = &h->cnt[1];
    sp.plop = plop ? 1 : 0;
    sp.se_sigma = (float)h->P.vte;
    sp.si_sigma = (float)h->P.vti;
    sp.plop_ref = 0;
    sp.wall = (h->wall_on && si >= 0) ? 1 : 0;
    sp.lost_ids = 0;
    sp.mq = sp.mq_cnt = nullptr;
    sp.mq_seg = 0;
    if (sp.wall) {
        sp.wq = h->sp[si].wq;
        sp.cw = &h->cnt[si];
    } else {
        sp.wq = StageView{};
        sp.cw = nullptr;
    }
    return sp;
}

int alloc_wall_queue(pic_handle *h, Species &s, int64_t n0) {
    StageView &v = s.wq;
    v.cap = (uint32_t)std::min<int64_t>(std::max<int64_t>(65536, n0 / 8), (int64_t)1 << 26);
    CHECK(dalloc(h, &v.cell, v.cap));
    CHECK(dalloc(h, &v.ox, v.cap));
    CHECK(dalloc(h, &v.oy, v.cap));
    CHECK(dalloc(h, &v.vx, v.cap));
    CHECK(dalloc(h, &v.vy, v.cap));
    CHECK(dalloc(h, &v.vz, v.cap));
    CHECK(dalloc(h, &v.key, v.cap));
    return PIC_OK;
}

// boundary::e_boundary / ion_boundary on the queued particles of species si
// (their created and returning electrons go to the electron stage)
void wall_pass(pic_handle *h, int si) {
    if (!h->wall_on) return;
    launch_wall(h->st, si, h->sp[si].wq, &h->cnt[si], h->sp[0].stage, &h->cnt[0], h->wc,
                (uint32_t)h->step);
}

bool plop_on(pic_handle *h) {
    return h->P.wall == PIC_WALL_STEADY && h->P.mode == PIC_MODE_PHYSICAL;
}

// Spawn of a fused particle pass of species si.  With the physical-mode
// steady walls the ion pass queues its ions lost through a wall in the wall
// queue (unused by these walls otherwise) and plop_drain stages their
// re-injected pairs after the pass.
Spawn pass_spawn(pic_handle *h, int si, int do_push) {
    Spawn sp = make_spawn(h, false, do_push ? si : -1);
    if (do_push && si == 1 && plop_on(h)) {
        sp.lost_ids = 1;
        sp.wq = h->sp[1].wq;
        sp.cw = &h->cnt[1];
    }
    return sp;
}

void plop_drain(pic_handle *h, int si, int do_push) {
    if (!(do_push && si == 1 && plop_on(h))) return;
    launch_plop_drain(h->st, h->sp[1].wq, &h->cnt[1], make_spawn(h, true, si), h->sp[1].mc,
                      (uint32_t)h->step, h->N);
}

// the reference's steady-state walls (pairing of outside electrons and ions)
bool ref_steady(pic_handle *h) {
    return h->P.wall == PIC_WALL_STEADY && h->P.mode == PIC_MODE_REF;
}

// ion_collisions::collision in reference mode (ion_collisions.cpp:255-308)
int ref_ion_cx(pic_handle *h) {
    Species &s = h->sp[1];
    launch_refcx(h->st, s.buf[s.cur], &h->cnt[1].n[s.cur], (uint32_t)s.cap, h->N, s.mc,
                 (uint32_t)h->step, h->cx, &h->cnt[1], &h->cnt[1].events[0][0]);
    return PIC_OK;
}

// steady_e_boundary (boundary.cpp:1094-1097) on the current buffers; the
// re-injected pairs are deposited into this step's charge
void steady_boundary(pic_handle *h) {
    Species &e = h->sp[0], &i = h->sp[1];
    SteadyArgs a;
    for (int s = 0; s < 2; s++) {
        Species &sp = h->sp[s];
        a.sp[s] = sp.buf[sp.cur];
        a.n[s] = &h->cnt[s].n[sp.cur];
        a.cnt[s] = &h->cnt[s];
        a.b[s] = sp.steady;
        a.cap[s] = sp.cap;
    }
    launch_steady(h->st, a, i.mc, (uint32_t)h->step, h->N, (float)h->P.vte, (float)h->P.vti, e.q,
                  i.q, e.pc.mode);
}

// candidate buffers of the steady walls: every particle of the store may be
// outside on an axis; nb id ranges per side keep the buckets near 64 entries
int alloc_steady(pic_handle *h, Species &s, uint32_t nb) {
    SteadyBufs &b = s.steady;
    b.nb = nb;
    CHECK(dalloc(h, &b.hist, 2 * (size_t)nb));
    CHECK(dalloc(h, &b.start, 2 * (size_t)nb + 1));
    CHECK(dalloc(h, &b.cursor, 2 * (size_t)nb));
    if (hipMemsetAsync(b.hist, 0, sizeof(uint32_t) * 2 * nb, h->st) != hipSuccess)
        return seterr(h, PIC_ERR_DEVICE, "hipMemsetAsync failed");
    CHECK(dalloc(h, &b.id_in, (size_t)s.cap));
    CHECK(dalloc(h, &b.idx_in, (size_t)s.cap));
    CHECK(dalloc(h, &b.id_out, (size_t)s.cap));
    CHECK(dalloc(h, &b.idx_out, (size_t)s.cap));
    return PIC_OK;
}

// Histogram of a species' sort pass (binCount runs back to zero in each
// sort pass, so it is zero between them).
uint32_t *bin_count(pic_handle *h, int si) {
    return h->binCount + (size_t)si * h->nbins_pad * BIN_STRIDE;
}

// The in-place pass ahead of a sort pass can count the sort pass's keys (the
// cells its particles end the step in) when nothing but the fused step's
// commits, which count what they append, changes the store in between:
// physical mode without the particle source, the population reduction or
// surface physics.
bool hist_fusable(const pic_handle *h, int si) {
    const Species &s = h->sp[si];
    return h->P.mode == PIC_MODE_PHYSICAL && s.pc.mode == PIC_DEV_MODE_PHYS &&
           h->P.number_to_inject <= 0 && !s.reduce && !h->sp[0].reduce && !h->sp[1].reduce &&
           !h->wall_on && !h->refcx;
}

// A pending histogram that will not be used (any path but the fused step's
// next sort pass of the species) is cleared, and its sort pass counts anew.
void drop_hist(pic_handle *h, int si) {
    Species &s = h->sp[si];
    if (!s.hist_pending) return;
    launch_fill<uint32_t>(h->st, bin_count(h, si), (size_t)h->nbins_pad * BIN_STRIDE, 0u);
    s.hist_pending = false;
}

// The fused steps of one pic_run / pic_advance call may leave a sort pass's
// histogram pending: it is cleared when the call returns, so any other entry
// point (or the next call) sees a zero histogram.
struct HistScope {
    pic_handle *h;
    ~HistScope() {
        drop_hist(h, 0);
        drop_hist(h, 1);
    }
};

// append species `src`'s staged particles to species `dst_s`'s current buffer
void commit(pic_handle *h, int src, int dst_s, int deposit) {
    drop_hist(h, dst_s);  // (not on the fused step's path: count its keys again)
    Species &d = h->sp[dst_s];
    launch_commit(h->st, h->sp[src].stage, &h->cnt[src], d.buf[d.cur], &h->cnt[dst_s].n[d.cur],
                  d.cap, &h->cnt[dst_s], deposit, h->N, d.pc.mode, d.q);
}

// commit(h, 1, 1, dep_i) and commit(h, 0, 0, dep_e) in one launch
void commit_both(pic_handle *h, int dep_i, int dep_e) {
    auto args = [&](int s, int dep) {
        Species &d = h->sp[s];
        // appended particles join the histogram of a pending sort pass
        return CommitArgs{d.stage, &h->cnt[s], d.buf[d.cur], &h->cnt[s].n[d.cur], (uint32_t)d.cap,
                          &h->cnt[s], dep, d.pc.mode, d.q,
                          d.hist_pending ? bin_count(h, s) : nullptr, d.pc.tilesX,
                          d.pc.keyshift};
    };
    launch_commit_pair(h->st, args(1, dep_i), args(0, dep_e), h->N);
}

void update_scale(pic_handle *h, Species &s) {
    s.scale = ldexp(s.qc / (h->dx * h->dx) * s.weight, -2 * FX_BITS);
}

// switch a species to per-particle weights: every particle takes the
// species weight as its own, the species factor becomes 1
int enable_weights(pic_handle *h, Species &s) {
    if (s.weighted) return PIC_OK;
    for (int b = 0; b < 2; b++) {
        CHECK(dalloc(h, &s.buf[b].w, (size_t)s.cap));
        launch_fill<float>(h->st, s.buf[b].w, (long)s.cap, (float)s.weight);
    }
    CHECK(dalloc(h, &s.stage.w, (size_t)s.stage.cap));
    if (s.wq.cap) CHECK(dalloc(h, &s.wq.w, (size_t)s.wq.cap));
    s.weight = 1.0;
    s.weighted = true;
    update_scale(h, s);
    return PIC_OK;
}

// inject_more_particles (particle.cpp:491-510) into the species' staging
// buffer; push: move and wall-test them with this step's field first
void inject(pic_handle *h, int si, int push) {
    Species &s = h->sp[si];
    const pic_params &P = h->P;
    if (P.number_to_inject <= 0) return;
    // particleVec() value-initialises the weight to 0 (particleVec.h:26-35):
    // the reference's injected particles carry no charge; physical mode
    // gives them weight 1
    const float w = P.mode == PIC_MODE_REF ? 0.0f : 1.0f;
    const double vt = si == 0 ? P.vte : P.vti;
    launch_inject(h->st, s.stage, &h->cnt[si], h->E32, s.pc, s.mc.key0, s.mc.key1,
                  (uint32_t)h->step, (uint32_t)P.number_to_inject, w, h->dx, P.length, vt,
                  (uint32_t)si, push, &h->cnt[si].lost[0][0]);
    s.injected += P.number_to_inject;
}

// contain_particle_growth (particle.cpp:640-653) on the current buffer;
// fix_charge: correct the charge deposited this step for the removed and
// re-weighted particles (the fused step deposits before the reduction)
void reduce_species(pic_handle *h, int si, bool fix_charge) {
    Species &s = h->sp[si];
    launch_reduce(h->st, s.buf[s.cur], &h->cnt[si].n[s.cur], s.cap, s.rc, s.rstate, &h->cnt[si],
                  (uint32_t)h->step, fix_charge ? s.q : nullptr, h->N, s.pc.mode);
    s.holes = true;
}

// absolute position (m) -> cell + offset
void to_cell(double x, double dx, int &j, float &o) {
    double t = x / dx;
    if (!(t > -8192.0 && t < 8192.0)) {
        j = (t > 0.0) ? 16000 : -16000;
        o = 0.0f;
        return;
    }
    double k = floor(t);
    float f = (float)(t - k);
    if (f >= 1.0f) {
        f = 0.0f;
        k += 1.0;
    }
    j = (int)k;
    o = f;
}

int upload_particles(pic_handle *h, int si, const std::vector<double> &v) {
    Species &s = h->sp[si];
    size_t n = v.size() / 6;
    if ((int64_t)n > s.cap) return seterr(h, PIC_ERR_CAPACITY, "particle store too small");
    std::vector<uint32_t> cell(n), id(n);
    std::vector<float> ox(n), oy(n), vx(n), vy(n), vz(n), w(n);
    for (size_t i = 0; i < n; i++) {
        w[i] = (float)v[6 * i + 5];
        int jx, jy;
        to_cell(v[6 * i], h->dx, jx, ox[i]);
        to_cell(v[6 * i + 1], h->dx, jy, oy[i]);
        cell[i] = cell_pack(jx, jy);
        vx[i] = (float)v[6 * i + 2];
        vy[i] = (float)v[6 * i + 3];
        vz[i] = (float)v[6 * i + 4];
        id[i] = (uint32_t)i;
    }
    SpView &b = s.buf[s.cur];
    HIPCHK(h, hipMemcpyAsync(b.cell, cell.data(), 4 * n, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(b.ox, ox.data(), 4 * n, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(b.oy, oy.data(), 4 * n, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(b.vx, vx.data(), 4 * n, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(b.vy, vy.data(), 4 * n, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(b.vz, vz.data(), 4 * n, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(b.id, id.data(), 4 * n, hipMemcpyHostToDevice, h->st));
    if (b.w) HIPCHK(h, hipMemcpyAsync(b.w, w.data(), 4 * n, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return PIC_OK;
}

int set_count(pic_handle *h, int si, int buf, uint32_t n) {
    launch_fill<uint32_t>(h->st, &h->cnt[si].n[buf], 1, n);
    return PIC_OK;
}

void clear_holes(pic_handle *h, int si) {
    if (h->sp[si].holes) launch_fill<uint32_t>(h->st, &h->cnt[si].holes[0], CNT_SLOTS, 0u);
    h->sp[si].holes = false;
}

int read_counters(pic_handle *h) {
    HIPCHK(h, hipMemcpyAsync(h->cnt_host, h->cnt, 2 * sizeof(SpCounters), hipMemcpyDeviceToHost,
                             h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return PIC_OK;
}

// Particles dropped for lack of capacity (store or staging buffer full) make
// the next call that sees them fail with PIC_ERR_CAPACITY, once per increase
// of the count.  The counters travel to pinned memory behind the queued work
// (no wait); `wait` blocks until the last copy has landed.
void queue_overflow_copy(pic_handle *h) {
    for (int s = 0; s < 2; s++)
        hipMemcpyAsync(&h->ovf_host[s], &h->cnt[s].overflow, sizeof(uint32_t),
                       hipMemcpyDeviceToHost, h->st);
    if (hipEventRecord(h->ovf_ev, h->st) == hipSuccess) h->ovf_pending = true;
}

int check_overflow(pic_handle *h, bool wait) {
    if (!h->ovf_pending) return PIC_OK;
    if (wait) {
        HIPCHK(h, hipEventSynchronize(h->ovf_ev));
    } else if (hipEventQuery(h->ovf_ev) != hipSuccess) {
        return PIC_OK;  // not landed yet: a later call reports it
    }
    h->ovf_pending = false;
    const uint64_t dropped = (uint64_t)h->ovf_host[0] + h->ovf_host[1];
    if (dropped > h->ovf_reported) {
        h->ovf_reported = dropped;
        return seterr(h, PIC_ERR_CAPACITY,
                      std::to_string(dropped) + " particles dropped: particle store or staging "
                      "buffer full (electrons " + std::to_string(h->ovf_host[0]) + ", ions " +
                      std::to_string(h->ovf_host[1]) + ")");
    }
    return PIC_OK;
}

// Physical-mode collisions run in a pass of their own ahead of the particle
// pass when few particles are candidates (the argon ions: a fraction of a
// percent): that pass then has the registers of a collisionless one, for 8
// B per particle read ahead.  Where many are (the electrons: ~13 % under the
// global bound), the candidates' velocities would cost a read and write of
// nearly every line, and the particle pass collides them in its block queue
// instead.  With energy-binned bounds (about 2 % candidates in a 2 eV argon
// discharge) the deferred form was measured too: the electrons' in-place
// pass fell from 1.69 to 1.35 ms at 1025^2, but colliding, pushing and
// depositing the 2 M queued candidates took 0.43 ms, so they collide in the
// pass.
bool collide_ahead(const Species &s) {
    return s.mc.on && s.mc.mode == PIC_DEV_MODE_PHYS && s.mc.p_null < 0.02f && !s.mc.bins;
}

// counting sort + (optional) collisions and push + deposit of one species.
// The keys are the cells at the start of the pass; with a push, the
// particles lost in the pass leave holes in the sorted store (counted in the
// holes counter, dropped by the next sort pass); without one the output is
// compact.
void push_species(pic_handle *h, int si, int do_push, int deposit) {
    Species &s = h->sp[si];
    int c = s.cur;
    uint32_t step = (uint32_t)h->step;
    mcc::MccConst mc = s.mc;
    if (!do_push) mc.on = 0;
    Spawn spn = pass_spawn(h, si, do_push);
    if (do_push && collide_ahead(s) && s.mq) {
        launch_collide(h->st, s.buf[c], &h->cnt[si].n[c], s.cap, h->E32, s.pc, mc, step, spn,
                       &h->cnt[si].events[0][0], &h->cnt[si].lost[0][0], s.mq, s.mq_cnt);
        mc.on = 0;
    }
    clear_holes(h, si);  // the pass drops the old holes
    uint32_t *bins = bin_count(h, si);
    // the in-place pass before a timestep's sort pass has made its histogram
    if (!(do_push && s.hist_pending)) {
        drop_hist(h, si);
        int p1 = prof_begin(h, 4);
        launch_key_count(h->st, s.buf[c], &h->cnt[si].n[c], s.cap, s.pc, bins, do_push);
        prof_end(h, p1);
    }
    s.hist_pending = false;
    int p2 = prof_begin(h, 5);
    launch_scan(h->st, bins, h->binStart, h->nbins_pad, h->scanPartial, &h->cnt[si].n[1 - c]);
    prof_end(h, p2);
    // a store that was never sorted (freshly loaded or uploaded, or
    // scrambled by a compaction) deposits from its sorted slots after the
    // pass, where a block's particles are neighbours: in the unsorted order
    // nearly every particle would take four global atomics
    const bool cold = s.since_sort >= (1 << 20);
    int p3 = prof_begin(h, 2);
    launch_push_scatter(h->st, s.buf[c], s.buf[1 - c], &h->cnt[si].n[c], s.cap, h->E32, s.pc, mc,
                        step, bins, h->binStart, s.q, do_push, deposit && !cold, spn,
                        &h->cnt[si].events[0][0], &h->cnt[si].lost[0][0],
                        &h->cnt[si].holes[0]);
    if (deposit && cold)
        launch_deposit(h->st, s.buf[1 - c], &h->cnt[si].n[1 - c], s.cap, h->N, s.pc.mode, s.q);
    prof_end(h, p3);
    plop_drain(h, si, do_push);
    s.cur = 1 - c;
    s.holes = do_push != 0;
    s.since_sort = 0;
}

// collisions + push + deposit of one species in place (no sort): removed
// particles leave holes that the next sort pass compacts
void push_species_inplace(pic_handle *h, int si) {
    Species &s = h->sp[si];
    int c = s.cur;
    Spawn spn = pass_spawn(h, si, 1);
    // few candidates: the pass makes the candidate test only and leaves each
    // candidate to k_collide_q (collide, push, store, deposit) after it
    const bool defer = collide_ahead(s) && s.mq;
    // the next pass of this species sorts: this one (and the deferred
    // collisions after it) count its keys
    drop_hist(h, si);
    const bool hist = s.since_sort + 2 >= s.sort_interval && hist_fusable(h, si);
    if (defer) {  // the queue counters are zero (k_collide_q leaves them so)
        spn.mq = s.mq;
        spn.mq_cnt = s.mq_cnt;
        spn.mq_seg = collide_seg_cap(s.cap);
    }
    int p = prof_begin(h, 6);
    launch_push_inplace(h->st, s.buf[c], &h->cnt[si].n[c], s.cap, h->E32, s.pc, s.mc,
                        (uint32_t)h->step, s.q, 1, spn, &h->cnt[si].events[0][0],
                        &h->cnt[si].lost[0][0], &h->cnt[si].holes[0],
                        hist ? bin_count(h, si) : nullptr);
    s.hist_pending = hist;
    prof_end(h, p);
    if (defer)
        launch_collide_deferred(h->st, s.buf[c], h->E32, s.pc, s.mc, (uint32_t)h->step, spn,
                                &h->cnt[si].events[0][0], &h->cnt[si].lost[0][0], s.q,
                                &h->cnt[si].holes[0], hist ? bin_count(h, si) : nullptr);
    plop_drain(h, si, 1);
    s.holes = true;
    s.since_sort++;
}

// Both species' in-place passes of a timestep in one launch, where nothing
// runs between them (no surface physics, spawns, re-injection or
// reference-mode charge exchange), neither sorts this step and both are
// collisionless; false: nothing was launched (the passes run one by one).
bool push_pair_inplace(pic_handle *h) {
    if (h->spawning || h->wall_on || h->refcx || plop_on(h)) return false;
    for (int si = 0; si < 2; si++) {
        const Species &s = h->sp[si];
        if (s.since_sort + 1 >= s.sort_interval || s.mc.on) return false;
    }
    InplaceArgs a[2];
    bool hist[2];
    for (int si = 0; si < 2; si++) {
        Species &s = h->sp[si];
        drop_hist(h, si);
        hist[si] = s.since_sort + 2 >= s.sort_interval && hist_fusable(h, si);
        const int c = s.cur;
        a[si] = InplaceArgs{s.buf[c],
                            &h->cnt[si].n[c],
                            s.cap,
                            s.pc,
                            s.mc,
                            s.q,
                            pass_spawn(h, si, 1),
                            &h->cnt[si].events[0][0],
                            &h->cnt[si].lost[0][0],
                            &h->cnt[si].holes[0],
                            hist[si] ? bin_count(h, si) : nullptr};
    }
    if (!inplace_pair_ok(a[0], a[1])) return false;
    int p = prof_begin(h, 9);  // two species passes
    launch_push_inplace_pair(h->st, a[0], a[1], h->E32, (uint32_t)h->step, 1);
    prof_end(h, p);
    for (int si = 0; si < 2; si++) {
        Species &s = h->sp[si];
        s.hist_pending = hist[si];
        s.holes = true;
        s.since_sort++;
    }
    return true;
}

// one timestep's particle pass of a species: a sort pass every
// sort_interval-th pass, in-place passes in between
void advance_species(pic_handle *h, int si) {
    Species &s = h->sp[si];
    const int k = s.sort_interval;
    if (s.since_sort + 1 >= k)
        push_species(h, si, 1, 1);
    else
        push_species_inplace(h, si);
}

// drop the holes of the current buffer (before host reads and the per-phase
// entry points, which expect a compact store)
void compact_species(pic_handle *h, int si) {
    if (h->sp[si].holes) push_species(h, si, 0, 0);
}

int absorb_species(pic_handle *h, int si, bool plop) {
    Species &s = h->sp[si];
    int c = s.cur;
    // block counts and starts of the compaction, allocated on first use for
    // the larger store
    if (!h->abs_count) {
        const uint32_t nb = absorb_blocks_pad(std::max(h->sp[0].cap, h->sp[1].cap));
        CHECK(dalloc(h, &h->abs_count, nb));
        CHECK(dalloc(h, &h->abs_start, (size_t)nb + 1));
        CHECK(dalloc(h, &h->abs_partial, (size_t)nb / 4096 + 1));
    }
    launch_absorb(h->st, s.buf[c], s.buf[1 - c], &h->cnt[si].n[c], &h->cnt[si].n[1 - c], s.cap,
                  h->N, s.pc.mode, &h->cnt[si].lost[0][0], h->sp[1].mc, (uint32_t)h->step,
                  make_spawn(h, plop, si), h->abs_count, h->abs_start, h->abs_partial);
    s.cur = 1 - c;
    clear_holes(h, si);  // the compaction dropped them
    // the survivors keep their order: a sorted store stays sorted
    return PIC_OK;
}

// phi from the solve with the boundary values; with_e: E (and its fp32 copy
// for the gather) from the same values in the same launch.  timestep: the
// fused step's field_start has already formed the rhs, the extrapolated start
// and the flag reset, and phi/E are queued before the host waits for the
// last cycles (they take the final iterate by the cycle count's parity).
// the verdict of a one-workgroup tolerance solve queued without a wait:
// read before anything reports the solver's figures or starts another solve
int mg_harvest(pic_handle *h) {
    if (!h->M.pending) return PIC_OK;
    const int used = mg_tol_verdict(h->M, &h->mg_relres);
    if (used < 0) return seterr(h, PIC_ERR_DEVICE, "multigrid: stream error");
    h->mg_cycles = used;
    h->mg_cycles_total += used;
    return PIC_OK;
}

int solve_potential(pic_handle *h, double phi0X, double phiR, double phi0Y, double phiLY,
                    bool with_e = false, bool timestep = false) {
    const pic_params &P = h->P;
    int N = h->N;
    MgLevels &M = h->M;
    h->last_bc[0] = phi0X;
    h->last_bc[1] = phiR;
    h->last_bc[2] = phi0Y;
    h->last_bc[3] = phiLY;
    if (!timestep)
        launch_fold(h->st, h->rho, N, 1.0 / P.epsilon, phi0X, phiR, phi0Y, phiLY,
                    1.0 / (h->dx * h->dx), M.d.rhsF[M.ng - 1]);
    bool queued_e = false;
    if (P.mg_mode == PIC_MG_TOL) {
        const bool early = timestep && with_e;
        std::function<void(const double *, const double *)> phi_e;
        if (early)
            phi_e = [&](const double *u0, const double *u1) {
                launch_phi_efield(h->st, u0, N, phi0X, phiR, phi0Y, phiLY, h->dx, h->phi,
                                  h->ex64, h->ey64, h->E32, u1, M.flags);
            };
        CHECK(mg_harvest(h));
        int used = mg_tol_solve(h->st, M, P.length, P.mg_tol, P.mg_max_iter, &h->mg_relres,
                                timestep, phi_e);
        if (used < 0 && used != MG_PENDING)
            return seterr(h, PIC_ERR_DEVICE, "multigrid: stream error");
        if (used >= 0) {
            h->mg_cycles = used;
            h->mg_cycles_total += used;
        }
        queued_e = early;
    } else {
        mg_fixed_solve(h->st, M, P.mg_ncycle, P.length);
        h->mg_cycles = P.mg_ncycle;
    }
    if (with_e && !queued_e)
        launch_phi_efield(h->st, M.d.u[M.ng - 1], N, phi0X, phiR, phi0Y, phiLY, h->dx, h->phi,
                          h->ex64, h->ey64, h->E32);
    else if (!with_e)
        launch_phi_bc(h->st, M.d.u[M.ng - 1], N, phi0X, phiR, phi0Y, phiLY, h->phi);
    h->have_phi = true;
    return PIC_OK;
}

int field_step(pic_handle *h, int64_t step) {
    const pic_params &P = h->P;
    MgLevels &M = h->M;
    const int top = M.ng - 1;
    double phiR = P.phi_right * cos(P.phi_period * (double)step);
    int p = prof_begin(h, 3);
    // rho, rhs, the tolerance solve's extrapolated start and flag reset
    const bool tol = P.mg_mode == PIC_MG_TOL;
    int extrap = 0;
    if (tol) {
        extrap = M.have_prev ? 2 : 1;
        M.have_prev = true;
    }
    launch_field_start(h->st, h->sp[1].q, h->sp[0].q, h->N, h->sp[1].scale, h->sp[0].scale,
                       P.corner8 ? 8.0 : 4.0, h->rho, 1.0 / P.epsilon, P.phi_left, phiR,
                       P.phi_bottom, P.phi_top, 1.0 / (h->dx * h->dx), M.d.rhsF[top], M.d.u[top],
                       M.uprev, extrap, tol ? M.flags : nullptr, h->sp[1].qbuf[h->sp[1].qcur ^ 1],
                       h->sp[0].qbuf[h->sp[0].qcur ^ 1]);
    CHECK(solve_potential(h, P.phi_left, phiR, P.phi_bottom, P.phi_top, true, true));
    prof_end(h, p);
    return PIC_OK;
}

#define REF_BLOCKS 512  // blocks of the report's reduction kernels

// FP64 levels of the accuracy report (rhsF, u and the V-cycle arrays), the
// reference potential and field, and the block sums
int alloc_ref(pic_handle *h) {
    if (h->ref_phi) return PIC_OK;
    MgLevels &R = h->ref;
    R.ng = h->M.ng;
    // the refinement V-cycles of the reference start above the level solved
    // in one workgroup (a grid that fits it whole, 65^2 with 160 KB of LDS,
    // would have no coarse-grid arrays for the V-cycle to restrict into);
    // a lower coarsest level fits the LDS as well
    R.lc = std::min(h->M.lc, R.ng - 2);
    for (int l = R.lc; l < R.ng; l++) {
        size_t s = (size_t)((1 << (l + 1)) + 1) * ((1 << (l + 1)) + 1);
        CHECK(dalloc(h, &R.d.rhsF[l], s));
        CHECK(dalloc(h, &R.d.u[l], s));
        if (l < R.ng - 1) {
            CHECK(dalloc(h, &R.d.cres[l], s));
            CHECK(dalloc(h, &R.d.cv[l], s));
        }
    }
    const size_t N = (size_t)h->N, E = (N + 1) * (N + 1);
    CHECK(dalloc(h, &h->ref_ex, E));
    CHECK(dalloc(h, &h->ref_ey, E));
    CHECK(dalloc(h, &h->ref_E32, E));
    CHECK(dalloc(h, &h->ref_partials, 6 * (size_t)REF_BLOCKS));
    CHECK(dalloc(h, &h->ref_phi, N * N));
    h->ref_host.assign(6 * REF_BLOCKS, 0.0);
    return PIC_OK;
}

// relative L2 errors of phi, Ex, Ey against the FP64 reference solve of the
// last solve's right-hand side (e[0..2]), its relative residual (e[3]) and
// V-cycles after the full multigrid (e[4]); e[5] = 0
int field_error(pic_handle *h, double tol, double *e) {
    if (!h->have_phi) return seterr(h, PIC_ERR_ARG, "no potential solved yet");
    CHECK(alloc_ref(h));
    MgLevels &R = h->ref;
    const int N = h->N, lf = R.ng - 1;
    HIPCHK(h, hipMemcpyAsync(R.d.rhsF[lf], h->M.d.rhsF[lf], sizeof(double) * N * N,
                             hipMemcpyDeviceToDevice, h->st));
    double rr = 0.0;
    const int cyc = mg_fp64_reference(h->st, R, h->P.length, tol, 200, h->ref_partials,
                                      h->ref_host.data(), REF_BLOCKS, &rr);
    if (cyc < 0) return seterr(h, PIC_ERR_DEVICE, "reference solve: stream error");
    const double *b = h->last_bc;
    launch_phi_bc(h->st, R.d.u[lf], N, b[0], b[1], b[2], b[3], h->ref_phi);
    launch_efield(h->st, h->ref_phi, N, h->dx, h->ref_ex, h->ref_ey, h->ref_E32);
    double s[6];
    if (field_diff_sums(h->st, h->phi, h->ref_phi, h->ex64, h->ref_ex, h->ey64, h->ref_ey,
                        (long)N * N, h->ref_partials, h->ref_host.data(), REF_BLOCKS, s) != 0)
        return seterr(h, PIC_ERR_DEVICE, "reference comparison: stream error");
    for (int k = 0; k < 3; k++) e[k] = s[2 * k + 1] > 0 ? sqrt(s[2 * k] / s[2 * k + 1]) : sqrt(s[2 * k]);
    e[3] = rr;
    e[4] = cyc;
    e[5] = 0.0;
    return PIC_OK;
}

int fused_step(pic_handle *h) {
    int N = h->N;
    int pt = prof_begin(h, 0);
    int pp = prof_begin(h, 1);
    // this step's charge goes into the buffers the last field step zeroed
    for (int s = 0; s < 2; s++) {
        Species &sp = h->sp[s];
        sp.qcur ^= 1;
        sp.q = sp.qbuf[sp.qcur];
    }
    // particle source after the electron collisions, before the gather
    // (pic.cpp:1071-1082): electrons are staged already moved, ions unmoved
    inject(h, 0, 1);
    inject(h, 1, 0);
    const bool ions_move = h->step % h->P.dt_dti == 0;
    if (!(ions_move && push_pair_inplace(h))) {
        advance_species(h, 0);
        wall_pass(h, 0);  // boundary::e_boundary on the electrons that reached a wall
        if (h->spawning) {
            // ionisation products: new ions join this step's ion pass (they
            // were appended before the ion push in the reference,
            // pic.cpp:1071-1111); new electrons were pushed with their
            // parents and are deposited here
            commit_both(h, 0, 1);
        }
        if (ions_move) {
            if (h->refcx) CHECK(ref_ion_cx(h));
            advance_species(h, 1);
            wall_pass(h, 1);  // boundary::ion_boundary (secondaries join the electrons)
        } else {
            Species &s = h->sp[1];
            launch_deposit(h->st, s.buf[s.cur], &h->cnt[1].n[s.cur], s.cap, N, s.pc.mode, s.q);
        }
    }
    if (h->spawning && (plop_on(h) || h->wall_on)) {
        // pairs re-injected for ions lost this step; ion-impact secondaries
        commit_both(h, 1, 1);
    }
    if (ref_steady(h)) steady_boundary(h);
    // pic.cpp:1130-1141: electrons, then ions, before the charge is used
    for (int s = 0; s < 2; s++)
        if (h->sp[s].reduce) reduce_species(h, s, true);
    prof_end(h, pp);
    CHECK(field_step(h, h->step));
    diag_step(h, h->step);
    prof_end(h, pt);
    h->have_rho = true;
    if (h->acc_every > 0 && (h->step + 1) % h->acc_every == 0) {
        double e[6];
        CHECK(field_error(h, h->acc_tol, e));
        CHECK(mg_harvest(h));
        const double row[6] = {(double)h->step, h->mg_relres, e[0], e[1], e[2], e[3]};
        h->acc_rows.insert(h->acc_rows.end(), row, row + 6);
    }
    h->step++;
    return PIC_OK;
}

int check_species(pic_handle *h, int s) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    if (s < 0 || s > 1) return seterr(h, PIC_ERR_ARG, "species index must be 0 (e) or 1 (ion)");
    return PIC_OK;
}

// lattice positions as particle.cpp:410-443 (start at dxp, step dxp, row by row,
// wrapping to a second layer if the lattice is exhausted)
void lattice_positions(int64_t n, double dxp, double L, std::vector<double> &xy) {
    xy.resize(2 * n);
    int64_t i = 0;
    while (i < n) {
        double y = dxp;
        while (y < L && i < n) {
            double x = dxp;
            while (x < L && i < n) {
                xy[2 * i] = x;
                xy[2 * i + 1] = y;
                x += dxp;
                ++i;
            }
            y += dxp;
        }
    }
}

}  // namespace

// dump_io.cpp
int read_dump_file(const std::string &path, std::vector<double> &vals, std::string &err);
int write_dump_file(const std::string &path, const std::vector<pic_particle> &p, int precision);

extern "C" {

const char *pic_last_error(pic_handle *h) {
    if (h) return h->err.c_str();
    return global_error();
}

int pic_create(const pic_params *Pin, int restart, const char *dump_dir, pic_handle **out) {
    if (!Pin || !out) return set_global_error(PIC_ERR_ARG, "null argument");
    *out = nullptr;
    const pic_params &P = *Pin;
    if (!grid_ok(P.nodes))
        return set_global_error(PIC_ERR_GRID, "nodes-1 must be a power of 2 (and nodes >= 5)");
    if (P.number_to_inject < 0) return set_global_error(PIC_ERR_ARG, "NUMBER_TO_INJECT < 0");
    if (P.sort_key < -1 || P.sort_key > 10) return set_global_error(PIC_ERR_ARG, "sort_key must be -1..10");
    if (P.mg_sweeps != -1 && (P.mg_sweeps < 1 || P.mg_sweeps > 3))
        return set_global_error(PIC_ERR_ARG, "mg_sweeps must be -1 or 1..3");
    if (P.reduce) {
        if (P.no_groups < 1 || P.no_groups > RD_MAX_GROUPS)
            return set_global_error(PIC_ERR_ARG, "NO_GROUPS must be 1..8");
        for (int s = 0; s < 2; s++) {
            if (P.size_to_check[s] < 0)
                return set_global_error(PIC_ERR_ARG, "SIZE_TO_CHECK must be >= 0");
            for (int g = 0; g < P.no_groups; g++)
                if (P.group_limit[s][g] < 0 || P.new_group_size[s][g] < 0 ||
                    !(P.v_limit[s][g] >= 0.0))
                    return set_global_error(PIC_ERR_ARG, "bad velocity-group limits");
        }
    }
    if (P.no_ion_species != 1)
        return set_global_error(PIC_ERR_UNSUPPORTED, "NO_ION_SPECIES must be 1");
    if (P.null_bound != PIC_NULL_BOUND_BINNED && P.null_bound != PIC_NULL_BOUND_GLOBAL)
        return set_global_error(PIC_ERR_ARG, "null_bound must be PIC_NULL_BOUND_BINNED or _GLOBAL");
    if (P.mg_mode != PIC_MG_FIXED && P.mg_mode != PIC_MG_TOL)
        return set_global_error(PIC_ERR_ARG, "unknown multigrid mode");
    if (P.mg_mode == PIC_MG_TOL && (!(P.mg_tol > 0.0) || P.mg_max_iter < 1))
        return set_global_error(PIC_ERR_ARG, "PIC_MG_TOL needs mg_tol > 0 and mg_max_iter >= 1");
    if (P.dt_dti < 1) return set_global_error(PIC_ERR_ARG, "DT_DTI must be >= 1");
    if (P.init_particle_no < 0) return set_global_error(PIC_ERR_ARG, "negative particle count");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_global_error(PIC_ERR_DEVICE, "no HIP device available");

    pic_handle *h = new pic_handle();
    h->P = P;
    h->N = P.nodes;
    h->dx = P.length / (double)(P.nodes - 1);  // parameters.h:193
    h->run_time = P.run_time;
    h->tilesX = (h->N + 1 + 7) / 8;
    // 8x8-cell tile keys by default: a block's particles then share one to a
    // few keys, so the claims are few and the scattered writes contiguous
    // (measured 1.6x faster than per-cell keys at 1025^2, 100 ppc)
    h->keyshift = P.sort_key >= 0 ? P.sort_key : tuned_row(h->N).sort_key;
    h->nbins = ((uint32_t)h->tilesX * h->tilesX * 64u) >> h->keyshift;
    h->nbins_pad = scan_pad(h->nbins);
    int rc = PIC_OK;
    auto fail = [&](int code) {
        pic_destroy(h);
        return code;
    };
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess)
        return fail(set_global_error(PIC_ERR_DEVICE, "hipStreamCreate failed"));
    setup_species(h, h->sp[0], true);
    setup_species(h, h->sp[1], false);

    // initial particles
    std::vector<double> dump[2];
    bool mixed[2] = {false, false};  // a dump whose particles' weights differ
    int64_t n0[2] = {P.init_particle_no, P.init_particle_no};
    if (restart) {
        std::string dir = dump_dir ? dump_dir : ".";
        for (int s = 0; s < 2; s++) {
            std::string path = dir + "/dataDump_" + h->sp[s].type + ".txt";
            std::string err;
            rc = read_dump_file(path, dump[s], err);
            if (rc != PIC_OK) return fail(set_global_error(rc, err));
            n0[s] = (int64_t)(dump[s].size() / 6);
            double w0 = n0[s] ? dump[s][5] : 1.0;
            for (int64_t i = 0; i < n0[s]; i++)
                if (dump[s][6 * i + 5] != w0) mixed[s] = true;
            h->sp[s].weight = mixed[s] ? 1.0 : w0;
        }
    }
    // surface physics acts at the non-steady walls only (pic.cpp:516-528);
    // Ar+ impact emission needs argon (pic.cpp:498)
    h->wall_on = P.wall == PIC_WALL_ABSORB &&
                 (P.e_impact_see || P.reflection || P.backscatter ||
                  (P.ion_impact_see && P.gas == PIC_GAS_ARGON));
    h->spawning = (P.have_collisions && P.mode == PIC_MODE_PHYSICAL) ||
                  (P.wall == PIC_WALL_STEADY && P.mode == PIC_MODE_PHYSICAL) ||
                  P.number_to_inject > 0 || h->wall_on;
    const int64_t n_inj = P.number_to_inject;
    for (int s = 0; s < 2; s++) {
        // created particles need headroom; otherwise counts only shrink
        int64_t cap = n0[s] + (h->spawning ? std::max<int64_t>(n0[s] / 2, 65536) : 0) + 4096 * n_inj;
        if (cap > 0x7fffffffLL) return fail(set_global_error(PIC_ERR_CAPACITY, "too many particles"));
        rc = species_alloc(h, h->sp[s], cap > 0 ? cap : 1);
        if (rc != PIC_OK) return fail(rc);
        int64_t extra = 2 * n_inj + (h->wall_on ? n0[0] / 8 + n0[1] / 8 : 0);
        if ((rc = alloc_stage(h, h->sp[s], n0[s], extra)) != PIC_OK) return fail(rc);
        update_scale(h, h->sp[s]);
        Species &sp = h->sp[s];
        if (h->wall_on && (rc = alloc_wall_queue(h, sp, n0[s])) != PIC_OK) return fail(rc);
        if (ref_steady(h)) {
            const int64_t cmax = std::max(n0[0], n0[1]) + 4096 * n_inj;
            uint32_t nb = 16;
            while (nb < 65536 && (int64_t)nb * 64 < cmax) nb <<= 1;
            if ((rc = alloc_steady(h, sp, nb)) != PIC_OK) return fail(rc);
        }
        // the physical-mode steady walls queue the lost ions there for
        // re-injection (enable_weights adds the weights when needed)
        if (!h->wall_on && s == 1 && plop_on(h) &&
            (rc = alloc_wall_queue(h, sp, n0[s])) != PIC_OK)
            return fail(rc);
        sp.reduce = P.reduce != 0;
        if (sp.reduce || n_inj > 0 || mixed[s])
            if ((rc = enable_weights(h, sp)) != PIC_OK) return fail(rc);
        if (sp.reduce) {
            ReduceConst &c = sp.rc;
            c.ng = P.no_groups;
            c.size_check = (uint32_t)P.size_to_check[s];
            c.species = (uint32_t)s;
            c.key0 = (uint32_t)P.seed;  // the collision streams' key
            c.key1 = (uint32_t)(P.seed >> 32);
            for (int g = 0; g < RD_MAX_GROUPS; g++) {
                const bool on = g < P.no_groups;
                c.vlim[g] = on ? P.v_limit[s][g] : 0.0;
                c.glimit[g] = on ? (uint32_t)P.group_limit[s][g] : 0u;
                c.gnew[g] = on ? (uint32_t)P.new_group_size[s][g] : 0u;
            }
            if ((rc = dalloc(h, (char **)&sp.rstate, sizeof(ReduceState))) != PIC_OK) return fail(rc);
            if (hipMemset(sp.rstate, 0, sizeof(ReduceState)) != hipSuccess)
                return fail(set_global_error(PIC_ERR_DEVICE, "memset failed"));
        }
    }
    if (P.have_collisions) {
        std::vector<float> tab;
        make_neutral_table(P, tab);
        if ((rc = dalloc(h, &h->ntab, tab.size())) != PIC_OK) return fail(rc);
        hipMemcpy(h->ntab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice);
    }
    if ((rc = setup_mcc(h, h->sp[0], true)) != PIC_OK || (rc = setup_mcc(h, h->sp[1], false)) != PIC_OK)
        return fail(rc);
    for (int s = 0; s < 2; s++) {
        Species &sp = h->sp[s];
        if (!collide_ahead(sp)) continue;
        if ((rc = dalloc(h, &sp.mq, (size_t)MQ_SEGS * collide_seg_cap(sp.cap))) != PIC_OK ||
            (rc = dalloc(h, &sp.mq_cnt, (size_t)MQ_SEGS * MQ_STRIDE)) != PIC_OK)
            return fail(rc);
        launch_fill<uint32_t>(h->st, sp.mq_cnt, MQ_SEGS * MQ_STRIDE, 0u);
    }
    if (h->wall_on) {
        WallConst &c = h->wc;
        c.mode = P.mode == PIC_MODE_PHYSICAL ? PIC_DEV_MODE_PHYS : PIC_DEV_MODE_REF;
        c.N = h->N;
        c.dx = h->dx;
        // boundary.cpp:373-374: lengthX = DX * X_NO_OF_CELLS, one cell beyond L
        c.Lw = P.mode == PIC_MODE_PHYSICAL ? P.length : h->dx * h->N;
        c.dt = P.dt;  // boundary(double dt = DT), boundary.h:62
        c.me = P.electron_mass;
        c.mi = P.ion_mass;
        c.q = kCharge;
        c.max_yield = P.see_max_yield;
        c.E0 = P.see_E0;
        c.Emax0 = P.see_Emax0;
        c.thr = P.threshold;
        c.wf = P.work_function;
        c.see_e = P.e_impact_see != 0;
        c.refl = P.reflection != 0;
        c.back = P.backscatter != 0;
        c.see_i = P.ion_impact_see != 0 && P.gas == PIC_GAS_ARGON;
        c.key0 = (uint32_t)P.seed;
        c.key1 = (uint32_t)(P.seed >> 32);
    }
    // reference mode: ions collide through the per-column accumulator
    h->refcx = P.have_collisions && P.mode == PIC_MODE_REF &&
               !(P.sigma_cx == PIC_SIGMA_CONSTANT && P.sigma_value == 0.0);
    if (h->refcx) {
        RefCxBuffers &B = h->cx;
        const size_t cap = (size_t)h->sp[1].cap;
        for (int k = 0; k < 2; k++) {
            if ((rc = dalloc(h, &B.id[k], cap)) != PIC_OK) return fail(rc);
            if ((rc = dalloc(h, &B.val[k], cap)) != PIC_OK) return fail(rc);
        }
        B.nb = refcx_ranges(h->N);
        const size_t nt = (size_t)h->N * B.nb;
        if ((rc = dalloc(h, &B.hist, nt)) != PIC_OK || (rc = dalloc(h, &B.start, nt + 1)) != PIC_OK ||
            (rc = dalloc(h, &B.cursor, nt)) != PIC_OK)
            return fail(rc);
        if (hipMemset(B.hist, 0, sizeof(uint32_t) * nt) != hipSuccess)
            return fail(set_global_error(PIC_ERR_DEVICE, "memset failed"));
        if ((rc = dalloc(h, &B.pcol, (size_t)h->N)) != PIC_OK) return fail(rc);
        if (hipMemset(B.pcol, 0, sizeof(float) * h->N) != hipSuccess)
            return fail(set_global_error(PIC_ERR_DEVICE, "memset failed"));
    }
    if ((rc = dalloc(h, &h->cnt, 2)) != PIC_OK) return fail(rc);
    if ((rc = diag_alloc(h)) != PIC_OK) return fail(rc);
    if (hipHostMalloc((void **)&h->cnt_host, 2 * sizeof(SpCounters), 0) != hipSuccess ||
        hipHostMalloc((void **)&h->ovf_host, 2 * sizeof(uint32_t), 0) != hipSuccess)
        return fail(set_global_error(PIC_ERR_ALLOC, "hipHostMalloc failed"));
    if (hipEventCreateWithFlags(&h->ovf_ev, hipEventDisableTiming) != hipSuccess)
        return fail(set_global_error(PIC_ERR_DEVICE, "hipEventCreate failed"));
    memset(h->cnt_host, 0, 2 * sizeof(SpCounters));
    h->ovf_host[0] = h->ovf_host[1] = 0;
    for (int s = 0; s < 2; s++) {
        h->cnt_host[s].n[0] = (uint32_t)n0[s];
        h->cnt_host[s].next_id = (uint32_t)n0[s];
    }
    if (hipMemcpy(h->cnt, h->cnt_host, 2 * sizeof(SpCounters), hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_global_error(PIC_ERR_DEVICE, "counter upload failed"));
    int N = h->N;
    // one histogram per species: an in-place pass may count the keys of its
    // species' next sort pass while the other species' sort pass runs
    if ((rc = dalloc(h, &h->binCount, 2 * (size_t)h->nbins_pad * BIN_STRIDE)) != PIC_OK ||
        (rc = dalloc(h, &h->binStart, h->nbins_pad + 1)) != PIC_OK ||
        (rc = dalloc(h, &h->scanPartial, h->nbins_pad / 4096 + 1)) != PIC_OK ||
        (rc = dalloc(h, &h->rho, (size_t)N * N)) != PIC_OK ||
        (rc = dalloc(h, &h->phi, (size_t)N * N)) != PIC_OK ||
        (rc = dalloc(h, &h->ex64, (size_t)(N + 1) * (N + 1))) != PIC_OK ||
        (rc = dalloc(h, &h->ey64, (size_t)(N + 1) * (N + 1))) != PIC_OK ||
        (rc = dalloc(h, &h->E32, (size_t)(N + 1) * N + 1)) != PIC_OK)
        return fail(rc);
    hipMemsetAsync(h->binCount, 0, 2 * 4ull * h->nbins_pad * BIN_STRIDE, h->st);
    hipMemsetAsync(h->rho, 0, 8ull * N * N, h->st);
    hipMemsetAsync(h->phi, 0, 8ull * N * N, h->st);
    hipMemsetAsync(h->ex64, 0, 8ull * (N + 1) * (N + 1), h->st);
    hipMemsetAsync(h->ey64, 0, 8ull * (N + 1) * (N + 1), h->st);
    hipMemsetAsync(h->E32, 0, 8ull * ((N + 1) * N + 1), h->st);  // pic.cpp:1086: first gather sees E = 0
    if ((rc = alloc_mg(h, h->M, N)) != PIC_OK) return fail(rc);
    h->M.nu = P.mg_sweeps > 0 ? P.mg_sweeps : tuned_row(h->N).mg_sweeps;
    h->M.use_graphs = P.graphs != 0;

    for (int s = 0; s < 2; s++) {
        Species &sp = h->sp[s];
        if (restart) {
            if ((rc = upload_particles(h, s, dump[s])) != PIC_OK) return fail(rc);
            continue;
        }
        if (n0[s] == 0) continue;
        launch_load_random(h->st, sp.buf[sp.cur], (uint32_t)n0[s], P.seed, (uint32_t)s, N,
                           sp.sigma_load);
        if (P.loader == PIC_LOAD_LATTICE) {
            // particle.cpp:410-443: DXPE = XLENGTH/(sqrt(INIT_PARTICLE_NO)+1)
            double dxp = P.length / (sqrt((double)P.init_particle_no) + 1.0);
            std::vector<double> xy;
            lattice_positions(n0[s], dxp, P.length, xy);
            std::vector<uint32_t> cell(n0[s]);
            std::vector<float> ox(n0[s]), oy(n0[s]);
            for (int64_t i = 0; i < n0[s]; i++) {
                int jx, jy;
                to_cell(xy[2 * i], h->dx, jx, ox[i]);
                to_cell(xy[2 * i + 1], h->dx, jy, oy[i]);
                cell[i] = cell_pack(jx, jy);
            }
            hipMemcpyAsync(sp.buf[sp.cur].cell, cell.data(), 4 * n0[s], hipMemcpyHostToDevice, h->st);
            hipMemcpyAsync(sp.buf[sp.cur].ox, ox.data(), 4 * n0[s], hipMemcpyHostToDevice, h->st);
            hipMemcpyAsync(sp.buf[sp.cur].oy, oy.data(), 4 * n0[s], hipMemcpyHostToDevice, h->st);
            hipStreamSynchronize(h->st);
        }
        // sort into cell order (no push, no deposit); generated particles are all inside
        push_species(h, s, 0, 0);
    }
    // the loader's sort pass must not count as losses
    if (read_counters(h) != PIC_OK) return fail(PIC_ERR_DEVICE);
    for (int s = 0; s < 2; s++)
        for (int k = 0; k < CNT_SLOTS; k++)
            for (int w = 0; w < 6; w++) h->cnt_host[s].lost[k][w] = 0;
    if (hipMemcpy(h->cnt, h->cnt_host, 2 * sizeof(SpCounters), hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_global_error(PIC_ERR_DEVICE, "counter upload failed"));
    hipError_t e = hipStreamSynchronize(h->st);
    if (e != hipSuccess)
        return fail(set_global_error(PIC_ERR_DEVICE, std::string("setup: ") + hipGetErrorString(e)));
    *out = h;
    return PIC_OK;
}

int pic_destroy(pic_handle *h) {
    if (!h) return PIC_ERR_ARG;
    if (h->st) hipStreamSynchronize(h->st);
    for (int s = 0; s < 2; s++) species_free(h->sp[s]);
    void *ptrs[] = {h->cnt,  h->binCount, h->binStart,  h->scanPartial, h->rho,
                    h->phi,  h->ex64,     h->ey64,      h->E32,         h->ntab,
                    h->abs_count, h->abs_start, h->abs_partial};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (h->cnt_host) hipHostFree(h->cnt_host);
    if (h->ovf_host) hipHostFree(h->ovf_host);
    if (h->ovf_ev) hipEventDestroy(h->ovf_ev);
    diag_free(h);
    void *cx[] = {h->cx.id[0], h->cx.id[1], h->cx.val[0], h->cx.val[1], h->cx.hist,
                  h->cx.start, h->cx.cursor, h->cx.pcol};
    for (void *p : cx)
        if (p) hipFree(p);
    free_mg(h->M);
    free_mg(h->ref);
    if (h->nonfinite) hipFree(h->nonfinite);
    for (void *p : {(void *)h->ref_phi, (void *)h->ref_ex, (void *)h->ref_ey, (void *)h->ref_E32,
                    (void *)h->ref_partials})
        if (p) hipFree(p);
    for (ProfPair &p : h->prof_live) {
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    for (ProfPair &p : h->prof_pool) {
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    if (h->st) hipStreamDestroy(h->st);
    delete h;
    return PIC_OK;
}

int pic_set_runtime(pic_handle *h, int32_t run_time) {
    if (!h || run_time < 0) return set_global_error(PIC_ERR_ARG, "bad argument");
    h->run_time = run_time;
    return PIC_OK;
}

int pic_run(pic_handle *h) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    // pic.cpp:1049-1057: wall pass before the loop (ion wall, then electron wall)
    if (ref_steady(h)) {
        steady_boundary(h);
    } else {
        CHECK(absorb_species(h, 1, plop_on(h)));
        wall_pass(h, 1);
        CHECK(absorb_species(h, 0, false));
        wall_pass(h, 0);
    }
    if (h->spawning && (plop_on(h) || h->wall_on)) {
        commit_both(h, 0, 0);
    }
    CHECK(diag_prepare_time(h, (int32_t)std::min<int64_t>(h->step + h->run_time, INT32_MAX)));
    {
        HistScope hs{h};
        for (int32_t t = 0; t < h->run_time; t++) CHECK(fused_step(h));
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return seterr(h, PIC_ERR_DEVICE, std::string("launch: ") + hipGetErrorString(e));
    // pic::pic_run returns when the run is complete
    queue_overflow_copy(h);
    CHECK(check_overflow(h, true));
    // pic.cpp:1262-1274: timeData.txt at the end of the run
    if (diag_time_on(h) && !h->out_dir.empty()) CHECK(diag_write_time(h, h->out_dir));
    return PIC_OK;
}

int pic_advance(pic_handle *h, int32_t nsteps) {
    if (!h || nsteps < 0) return set_global_error(PIC_ERR_ARG, "bad argument");
    CHECK(check_overflow(h, false));  // drops of an earlier call that have landed
    CHECK(diag_prepare_time(h, (int32_t)std::min<int64_t>(h->step + nsteps, INT32_MAX)));
    {
        HistScope hs{h};
        for (int32_t t = 0; t < nsteps; t++) CHECK(fused_step(h));
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return seterr(h, PIC_ERR_DEVICE, std::string("launch: ") + hipGetErrorString(e));
    queue_overflow_copy(h);
    return PIC_OK;
}

int pic_sync(pic_handle *h) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    HIPCHK(h, hipStreamSynchronize(h->st));
    CHECK(mg_harvest(h));
    return check_overflow(h, true);
}

// ------------------------------------------------------ per-phase entry
int pic_species_gather(pic_handle *h, int32_t s) {
    CHECK(check_species(h, s));
    compact_species(h, s);
    Species &sp = h->sp[s];
    if (!sp.exp_) {
        CHECK(dalloc(h, &sp.exp_, sp.cap));
        CHECK(dalloc(h, &sp.eyp, sp.cap));
    }
    launch_gather(h->st, sp.buf[sp.cur], &h->cnt[s].n[sp.cur], sp.cap, h->E32, sp.pc, sp.exp_,
                  sp.eyp);
    sp.gathered = true;
    return PIC_OK;
}

int pic_species_move(pic_handle *h, int32_t s) {
    CHECK(check_species(h, s));
    compact_species(h, s);
    Species &sp = h->sp[s];
    if (!sp.gathered)
        return seterr(h, PIC_ERR_ARG, "pic_species_move needs pic_species_gather first");
    launch_move(h->st, sp.buf[sp.cur], &h->cnt[s].n[sp.cur], sp.cap, sp.exp_, sp.eyp, sp.pc);
    sp.gathered = false;  // particle.cpp:1336-1339 frees the per-particle field
    return PIC_OK;
}

int pic_species_collect_charge(pic_handle *h, int32_t s) {
    CHECK(check_species(h, s));
    compact_species(h, s);
    Species &sp = h->sp[s];
    launch_fill<unsigned long long>(h->st, sp.q, (long)h->N * h->N, 0ull);
    launch_deposit(h->st, sp.buf[sp.cur], &h->cnt[s].n[sp.cur], sp.cap, h->N, sp.pc.mode, sp.q);
    return PIC_OK;
}

int pic_fields_grid_rho(pic_handle *h) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    launch_rho(h->st, h->sp[1].q, h->sp[0].q, h->N, h->sp[1].scale, h->sp[0].scale,
               h->P.corner8 ? 8.0 : 4.0, h->rho);
    h->have_rho = true;
    return PIC_OK;
}

int pic_fields_grid_potential(pic_handle *h, double phi0X, double phiLX, double phi0Y,
                              double phiLY, double timestep, double phi_period) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    double phiR = phiLX * cos(phi_period * timestep);  // fields.cpp:177
    return solve_potential(h, phi0X, phiR, phi0Y, phiLY);
}

int pic_fields_grid_electric_field(pic_handle *h) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    launch_efield(h->st, h->phi, h->N, h->dx, h->ex64, h->ey64, h->E32);
    h->step++;  // the last phase of pic_run's loop body: the step is complete
    return PIC_OK;
}

int pic_e_collisions(pic_handle *h) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    compact_species(h, 0);
    compact_species(h, 1);
    Species &s = h->sp[0];
    if (!s.mc.on) return PIC_OK;
    launch_mcc(h->st, s.buf[s.cur], &h->cnt[0].n[s.cur], s.cap, s.pc, s.mc, (uint32_t)h->step,
               make_spawn(h, false, -1), &h->cnt[0].events[0][0]);
    if (s.mc.mode == PIC_DEV_MODE_PHYS) {  // e_collisions.cpp:405-441 appends the products
        commit_both(h, 0, 0);
    }
    return PIC_OK;
}

int pic_ion_collisions(pic_handle *h) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    compact_species(h, 1);
    if (h->refcx) return ref_ion_cx(h);
    Species &s = h->sp[1];
    if (!s.mc.on) return PIC_OK;
    launch_mcc(h->st, s.buf[s.cur], &h->cnt[1].n[s.cur], s.cap, s.pc, s.mc, (uint32_t)h->step,
               make_spawn(h, false, -1), &h->cnt[1].events[0][0]);
    return PIC_OK;
}

int pic_field_error(pic_handle *h, double tol, double *out5) {
    if (!h || !out5 || !(tol > 0.0)) return set_global_error(PIC_ERR_ARG, "bad argument");
    double e[6];
    CHECK(field_error(h, tol, e));
    for (int k = 0; k < 5; k++) out5[k] = e[k];
    return PIC_OK;
}

int pic_field_finite(pic_handle *h, int32_t *finite) {
    if (!h || !finite) return set_global_error(PIC_ERR_ARG, "bad argument");
    if (!h->nonfinite) {
        CHECK(dalloc(h, &h->nonfinite, 1));
    }
    launch_fill<uint32_t>(h->st, h->nonfinite, 1, 0u);
    const long n = (long)h->N * h->N, ne = (long)(h->N + 1) * (h->N + 1);
    launch_count_nonfinite(h->st, h->phi, n, h->ex64, h->ey64, ne, h->nonfinite);
    unsigned bad = 0;
    HIPCHK(h, hipMemcpyAsync(&bad, h->nonfinite, sizeof(unsigned), hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    *finite = bad == 0;
    return PIC_OK;
}

int pic_set_accuracy_report(pic_handle *h, int32_t every, double tol) {
    if (!h || every < 0 || !(tol > 0.0)) return set_global_error(PIC_ERR_ARG, "bad argument");
    h->acc_every = every;
    h->acc_tol = tol;
    return PIC_OK;
}

int pic_get_accuracy_report(pic_handle *h, double *out, int32_t cap_rows, int32_t *rows) {
    if (!h || !rows || cap_rows < 0 || (cap_rows > 0 && !out))
        return set_global_error(PIC_ERR_ARG, "bad argument");
    const int32_t n = (int32_t)(h->acc_rows.size() / 6);
    *rows = n;
    const int32_t m = std::min(n, cap_rows);
    for (int64_t k = 0; k < (int64_t)m * 6; k++) out[k] = h->acc_rows[k];
    return PIC_OK;
}

int pic_cross_section(int32_t gas, int32_t channel, int32_t mode, double E, double *sigma) {
    if (!sigma || (gas != PIC_GAS_ARGON && gas != PIC_GAS_HYDROGEN) ||
        (mode != PIC_MODE_REF && mode != PIC_MODE_PHYSICAL) || !(E >= 0.0))
        return set_global_error(PIC_ERR_ARG, "bad argument");
    const bool ref = mode == PIC_MODE_REF;
    double s = 0.0;
    if (gas == PIC_GAS_ARGON) {
        switch (channel) {
            case PIC_XS_ELASTIC: s = ref ? mcc::ar_el_ref<double>(E) : mcc::ar_el_phys<double>(E); break;
            case PIC_XS_EXCITATION:
                if (ref) return set_global_error(PIC_ERR_UNSUPPORTED, "no excitation channel in reference mode");
                s = mcc::ar_exc_phys<double>(E, 11.55);
                break;
            case PIC_XS_IONIZATION:
                s = ref ? mcc::ar_iz_ref<double>(E) : mcc::ar_iz_phys<double>(E, 15.76);
                break;
            case PIC_XS_CX: s = mcc::ar_cx<double>(E); break;
            case PIC_XS_ION_ISO:
                if (ref) return set_global_error(PIC_ERR_UNSUPPORTED, "no ion elastic channel in reference mode");
                s = mcc::ar_iso_phys<double>(E);
                break;
            default: return set_global_error(PIC_ERR_ARG, "unknown channel");
        }
    } else {
        switch (channel) {
            case PIC_XS_ELASTIC: {
                const double v = mcc::h_el<double>(E);
                s = ref ? v : (v > 0.0 ? v : 0.0);
                break;
            }
            case PIC_XS_IONIZATION:  // the Janev bracket; physical mode converts 1e-17 -> 1e-20 m^2
                s = (ref ? 1.0 : 1000.0) * mcc::h_iz_janev<double>(E);
                break;
            case PIC_XS_CX:  // keV/amu of H+ (the reference divides by AMU in kg: sigma 0)
                s = mcc::h_cx_janev<double>(ref ? E / 1000.0 / kAmu : E / 1000.0 / 1.00728);
                break;
            case PIC_XS_EXCITATION:
            case PIC_XS_ION_ISO: return set_global_error(PIC_ERR_UNSUPPORTED, "no such hydrogen channel");
            default: return set_global_error(PIC_ERR_ARG, "unknown channel");
        }
    }
    *sigma = s;
    return PIC_OK;
}

int pic_set_step(pic_handle *h, int64_t step) {
    if (!h || step < 0) return set_global_error(PIC_ERR_ARG, "bad argument");
    h->step = step;
    return PIC_OK;
}

int pic_set_tuning(pic_handle *h, int32_t sort_interval_e, int32_t sort_interval_i,
                   int32_t mg_sweeps) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    const int32_t si[2] = {sort_interval_e, sort_interval_i};
    for (int s = 0; s < 2; s++)
        if (si[s] < 0 || si[s] > 64) return seterr(h, PIC_ERR_ARG, "sort interval must be 0..64");
    if (mg_sweeps != 0 && (mg_sweeps < 1 || mg_sweeps > 3))
        return seterr(h, PIC_ERR_ARG, "mg_sweeps must be 0 or 1..3");
    for (int s = 0; s < 2; s++)
        if (si[s] > 0) h->sp[s].sort_interval = si[s];
    if (mg_sweeps > 0 && mg_sweeps != h->M.nu) {
        HIPCHK(h, hipStreamSynchronize(h->st));
        for (MgLevels::CycleGraph &g : h->M.graphs) hipGraphExecDestroy(g.exec);
        h->M.graphs.clear();  // captured with the old sweep count
        h->M.nu = mg_sweeps;
    }
    return PIC_OK;
}

int pic_get_tuning(pic_handle *h, int32_t *out4) {
    if (!h || !out4) return set_global_error(PIC_ERR_ARG, "null argument");
    out4[0] = h->sp[0].sort_interval;
    out4[1] = h->sp[1].sort_interval;
    out4[2] = h->keyshift;
    out4[3] = h->M.nu;
    return PIC_OK;
}

int pic_get_collision_counts(pic_handle *h, int32_t s, int64_t *out8) {
    CHECK(check_species(h, s));
    if (!out8) return seterr(h, PIC_ERR_ARG, "null output");
    CHECK(read_counters(h));
    for (int k = 0; k < 8; k++) {
        out8[k] = 0;
        for (int b = 0; b < CNT_SLOTS; b++) out8[k] += (int64_t)h->cnt_host[s].events[b][k];
    }
    return PIC_OK;
}

int pic_get_overflow(pic_handle *h, int64_t *dropped) {
    if (!h || !dropped) return set_global_error(PIC_ERR_ARG, "null argument");
    CHECK(read_counters(h));
    *dropped = (int64_t)h->cnt_host[0].overflow + h->cnt_host[1].overflow;
    return PIC_OK;
}

int pic_boundary(pic_handle *h, int32_t which) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    if (which != 0 && which != 1) return seterr(h, PIC_ERR_ARG, "which must be 0 (ion) or 1 (electron)");
    if (ref_steady(h)) {  // steady_i_boundary is empty; steady_e_boundary pairs
        if (which == 1) steady_boundary(h);
        return PIC_OK;
    }
    bool plop = which == 0 && plop_on(h);
    const int si = which == 0 ? 1 : 0;
    CHECK(absorb_species(h, si, plop));
    wall_pass(h, si);
    if (plop || h->wall_on) {
        commit_both(h, 0, 0);
    }
    return PIC_OK;
}

int pic_mglin(double *u, int32_t n, int32_t ncycle, double length) {
    if (!u || ncycle < 1) return set_global_error(PIC_ERR_ARG, "bad argument");
    int m = n - 1;
    if (n < 3 || (m & (m - 1)) != 0)
        return set_global_error(PIC_ERR_GRID, "n-1 must be a power of 2 in mglin.");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_global_error(PIC_ERR_DEVICE, "no HIP device available");
    pic_handle tmp;
    int rc = alloc_mg(&tmp, tmp.M, n);
    hipStream_t st = nullptr;
    if (rc == PIC_OK && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        rc = set_global_error(PIC_ERR_DEVICE, "stream");
    if (rc == PIC_OK) {
        size_t bytes = sizeof(double) * n * n;
        MgLevels &M = tmp.M;
        hipMemcpyAsync(M.d.rhsF[M.ng - 1], u, bytes, hipMemcpyHostToDevice, st);
        mg_fixed_solve(st, M, ncycle, length);
        hipMemcpyAsync(u, M.d.u[M.ng - 1], bytes, hipMemcpyDeviceToHost, st);
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = set_global_error(PIC_ERR_DEVICE, hipGetErrorString(e));
    }
    if (st) hipStreamDestroy(st);
    free_mg(tmp.M);
    return rc;
}

int pic_poisson_tol(const double *f, double *u, int32_t n, double length, double tol,
                    int32_t max_iter, int32_t *cycles, double *relres) {
    if (!f || !u || !(tol > 0.0) || max_iter < 1) return set_global_error(PIC_ERR_ARG, "bad argument");
    int m = n - 1;
    if (n < 5 || (m & (m - 1)) != 0)
        return set_global_error(PIC_ERR_GRID, "n-1 must be a power of 2 in mglin.");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_global_error(PIC_ERR_DEVICE, "no HIP device available");
    pic_handle tmp;
    int rc = alloc_mg(&tmp, tmp.M, n);
    hipStream_t st = nullptr;
    if (rc == PIC_OK && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        rc = set_global_error(PIC_ERR_DEVICE, "stream");
    if (rc == PIC_OK) {
        size_t bytes = sizeof(double) * n * n;
        MgLevels &M = tmp.M;
        hipMemcpyAsync(M.d.rhsF[M.ng - 1], f, bytes, hipMemcpyHostToDevice, st);
        hipMemcpyAsync(M.d.u[M.ng - 1], u, bytes, hipMemcpyHostToDevice, st);
        double rr = 0.0;
        int used = mg_tol_solve(st, M, length, tol, max_iter, &rr);
        if (used == MG_PENDING) used = mg_tol_verdict(M, &rr);
        hipMemcpyAsync(u, M.d.u[M.ng - 1], bytes, hipMemcpyDeviceToHost, st);
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess || used < 0)
            rc = set_global_error(PIC_ERR_DEVICE, e != hipSuccess ? hipGetErrorString(e) : "mg");
        if (cycles) *cycles = used;
        if (relres) *relres = rr;
    }
    if (st) hipStreamDestroy(st);
    free_mg(tmp.M);
    return rc;
}

int pic_get_mg_stats(pic_handle *h, int32_t *last_cycles, int64_t *total_cycles, double *relres) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    CHECK(mg_harvest(h));
    if (last_cycles) *last_cycles = h->mg_cycles;
    if (total_cycles) *total_cycles = h->mg_cycles_total;
    if (relres) *relres = h->mg_relres;
    return PIC_OK;
}

// ------------------------------------------------------------ getters
int pic_particle_size(pic_handle *h, int32_t s, int64_t *n) {
    CHECK(check_species(h, s));
    if (!n) return seterr(h, PIC_ERR_ARG, "null output");
    compact_species(h, s);
    CHECK(read_counters(h));
    *n = h->cnt_host[s].n[h->sp[s].cur];
    return PIC_OK;
}

int pic_get_state(pic_handle *h, int32_t s, uint32_t *cell, float *ox, float *oy, float *vx,
                  float *vy, float *vz, uint32_t *id, int64_t cap) {
    int64_t n;
    CHECK(pic_particle_size(h, s, &n));
    if (n > cap) return seterr(h, PIC_ERR_CAPACITY, "output buffer too small");
    const SpView &b = h->sp[s].buf[h->sp[s].cur];
    size_t bytes = 4 * (size_t)n;
    if (cell) HIPCHK(h, hipMemcpyAsync(cell, b.cell, bytes, hipMemcpyDeviceToHost, h->st));
    if (ox) HIPCHK(h, hipMemcpyAsync(ox, b.ox, bytes, hipMemcpyDeviceToHost, h->st));
    if (oy) HIPCHK(h, hipMemcpyAsync(oy, b.oy, bytes, hipMemcpyDeviceToHost, h->st));
    if (vx) HIPCHK(h, hipMemcpyAsync(vx, b.vx, bytes, hipMemcpyDeviceToHost, h->st));
    if (vy) HIPCHK(h, hipMemcpyAsync(vy, b.vy, bytes, hipMemcpyDeviceToHost, h->st));
    if (vz) HIPCHK(h, hipMemcpyAsync(vz, b.vz, bytes, hipMemcpyDeviceToHost, h->st));
    if (id) HIPCHK(h, hipMemcpyAsync(id, b.id, bytes, hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return PIC_OK;
}

int pic_set_state(pic_handle *h, int32_t s, const uint32_t *cell, const float *ox, const float *oy,
                  const float *vx, const float *vy, const float *vz, const uint32_t *id,
                  int64_t n) {
    CHECK(check_species(h, s));
    if (n < 0 || n > h->sp[s].cap) return seterr(h, PIC_ERR_CAPACITY, "more particles than capacity");
    if (!cell || !ox || !oy || !vx || !vy || !vz || !id) return seterr(h, PIC_ERR_ARG, "null array");
    Species &sp = h->sp[s];
    const SpView &b = sp.buf[sp.cur];
    size_t bytes = 4 * (size_t)n;
    HIPCHK(h, hipMemcpyAsync(b.cell, cell, bytes, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(b.ox, ox, bytes, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(b.oy, oy, bytes, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(b.vx, vx, bytes, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(b.vy, vy, bytes, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(b.vz, vz, bytes, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(b.id, id, bytes, hipMemcpyHostToDevice, h->st));
    if (b.w) launch_fill<float>(h->st, b.w, (long)n, 1.0f);  // pic_set_weights sets others
    set_count(h, s, sp.cur, (uint32_t)n);
    clear_holes(h, s);
    sp.since_sort = 1 << 20;  // unsorted upload: the next pass sorts
    // ids handed out later must not collide with the uploaded ones
    uint32_t top = 0;
    for (int64_t k = 0; k < n; k++) top = id[k] >= top ? id[k] + 1u : top;
    CHECK(read_counters(h));
    if (h->cnt_host[s].next_id < top) {
        h->cnt_host[s].next_id = top;
        HIPCHK(h, hipMemcpy(&h->cnt[s].next_id, &top, sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    HIPCHK(h, hipStreamSynchronize(h->st));
    return PIC_OK;
}

int pic_get_particles(pic_handle *h, int32_t s, pic_particle *out, int64_t cap) {
    int64_t n;
    CHECK(pic_particle_size(h, s, &n));
    if (n > cap || !out) return seterr(h, PIC_ERR_CAPACITY, "output buffer too small");
    std::vector<uint32_t> cell(n);
    std::vector<float> ox(n), oy(n), vx(n), vy(n), vz(n), wt;
    CHECK(pic_get_state(h, s, cell.data(), ox.data(), oy.data(), vx.data(), vy.data(), vz.data(),
                        nullptr, n));
    const Species &sp = h->sp[s];
    if (sp.weighted) {
        wt.resize(n);
        HIPCHK(h, hipMemcpy(wt.data(), sp.buf[sp.cur].w, 4 * (size_t)n, hipMemcpyDeviceToHost));
    }
    double dx = h->dx, w = sp.weight;
    for (int64_t i = 0; i < n; i++) {
        out[i].x = ((double)cell_jx(cell[i]) + (double)ox[i]) * dx;
        out[i].y = ((double)cell_jy(cell[i]) + (double)oy[i]) * dx;
        out[i].vx = vx[i];
        out[i].vy = vy[i];
        out[i].vz = vz[i];
        out[i].weight = sp.weighted ? w * (double)wt[i] : w;
    }
    return PIC_OK;
}

// per-particle weights, in the order of pic_get_state (particle.h: weight)
int pic_get_weights(pic_handle *h, int32_t s, float *w, int64_t cap) {
    int64_t n;
    CHECK(pic_particle_size(h, s, &n));
    if (n > cap || (!w && n > 0)) return seterr(h, PIC_ERR_CAPACITY, "output buffer too small");
    const Species &sp = h->sp[s];
    if (!sp.weighted) {
        for (int64_t i = 0; i < n; i++) w[i] = (float)sp.weight;
        return PIC_OK;
    }
    HIPCHK(h, hipMemcpyAsync(w, sp.buf[sp.cur].w, 4 * (size_t)n, hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return PIC_OK;
}

int pic_set_weights(pic_handle *h, int32_t s, const float *w, int64_t n) {
    int64_t cur;
    CHECK(pic_particle_size(h, s, &cur));
    if (n != cur) return seterr(h, PIC_ERR_ARG, "weight count must equal the particle count");
    if (!w && n > 0) return seterr(h, PIC_ERR_ARG, "null array");
    Species &sp = h->sp[s];
    CHECK(enable_weights(h, sp));
    HIPCHK(h, hipMemcpyAsync(sp.buf[sp.cur].w, w, 4 * (size_t)n, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return PIC_OK;
}

// particle::inject_more_particles (particle.cpp:491-510) as a phase of its
// own: NUMBER_TO_INJECT new particles appended unmoved
int pic_inject_particles(pic_handle *h, int32_t s) {
    CHECK(check_species(h, s));
    if (h->P.number_to_inject <= 0) return PIC_OK;
    compact_species(h, s);
    inject(h, s, 0);
    commit(h, s, s, 0);
    h->sp[s].since_sort = 1 << 20;  // appended unsorted: the next pass sorts
    return PIC_OK;
}

// particle::contain_particle_growth (particle.cpp:640-653) as a phase of its own
int pic_contain_particle_growth(pic_handle *h, int32_t s) {
    CHECK(check_species(h, s));
    Species &sp = h->sp[s];
    if (!sp.reduce) return seterr(h, PIC_ERR_ARG, "population reduction is off (pic_params.reduce)");
    compact_species(h, s);
    reduce_species(h, s, false);  // collect_charge follows (pic.cpp:1139-1150)
    compact_species(h, s);
    return PIC_OK;
}

// surface events of species s: reached a wall, absorbed, reflected,
// backscattered, emitting true secondaries, secondary electrons created
// (the reference's count_absorb / count_ref / count_scat / count_sec,
// boundary.cpp:591, kept here across steps)
// live particles of species s without compacting the store (the count minus
// the holes of in-place passes): a cheap read for drivers between steps
int pic_particle_count(pic_handle *h, int32_t s, int64_t *n) {
    CHECK(check_species(h, s));
    if (!n) return seterr(h, PIC_ERR_ARG, "null output");
    CHECK(read_counters(h));
    int64_t holes = 0;
    for (int b = 0; b < CNT_SLOTS; b++) holes += h->cnt_host[s].holes[b];
    *n = (int64_t)h->cnt_host[s].n[h->sp[s].cur] - (h->sp[s].holes ? holes : 0);
    return PIC_OK;
}

int pic_get_device_bytes(pic_handle *h, int64_t *bytes) {
    if (!h || !bytes) return set_global_error(PIC_ERR_ARG, "null argument");
    *bytes = h->dev_bytes;
    return PIC_OK;
}

// Upper estimate of the device memory pic_create allocates for p (the same
// capacities as pic_create, without a restart file): used to admit runs
// against a memory budget before creating them.
int pic_estimate_bytes(const pic_params *pp, int64_t *bytes) {
    if (!pp || !bytes) return set_global_error(PIC_ERR_ARG, "null argument");
    const pic_params &P = *pp;
    if (!grid_ok(P.nodes)) return set_global_error(PIC_ERR_GRID, "nodes-1 must be a power of 2");
    const int64_t N = P.nodes, n0 = std::max<int64_t>(P.init_particle_no, 0);
    const int64_t n_inj = std::max<int64_t>(P.number_to_inject, 0);
    const bool wall_on = P.wall == PIC_WALL_ABSORB &&
                         (P.e_impact_see || P.reflection || P.backscatter || P.ion_impact_see);
    const bool spawning = (P.have_collisions && P.mode == PIC_MODE_PHYSICAL) ||
                          (P.wall == PIC_WALL_STEADY && P.mode == PIC_MODE_PHYSICAL) ||
                          n_inj > 0 || wall_on;
    const bool weighted = P.reduce || n_inj > 0;
    int64_t b = 0;
    for (int s = 0; s < 2; s++) {
        int64_t cap = n0 + (spawning ? std::max<int64_t>(n0 / 2, 65536) : 0) + 4096 * n_inj;
        cap = std::max<int64_t>(cap, 1);
        b += 2 * cap * (28 + (weighted ? 4 : 0)) + N * N * 16;  // two charge buffers
        int64_t want = std::max<int64_t>(16384, std::min<int64_t>(n0 / 16, (int64_t)1 << 22)) +
                       2 * n_inj + (wall_on ? n0 / 4 : 0);
        int64_t sc = 1;
        while (sc < want) sc <<= 1;
        b += sc * (36 + (weighted ? 4 : 0));
        if (wall_on) b += std::min<int64_t>(std::max<int64_t>(65536, n0 / 8), (int64_t)1 << 26) * 36;
        if (P.reduce) b += (int64_t)sizeof(ReduceState);
        // the reference steady walls' candidate buffers
        if (P.wall == PIC_WALL_STEADY && P.mode == PIC_MODE_REF) b += cap * 16 + (1 << 20);
        if (s == 1 && P.have_collisions && P.mode == PIC_MODE_REF)
            b += cap * 24 + 12 * N * (int64_t)refcx_ranges((int)N) + N * 4;
        // the physical-mode steady walls' queue of lost ions
        if (s == 1 && !wall_on && P.wall == PIC_WALL_STEADY && P.mode == PIC_MODE_PHYSICAL)
            b += std::min<int64_t>(std::max<int64_t>(65536, n0 / 8), (int64_t)1 << 26) *
                 (36 + (weighted ? 4 : 0));
        // candidate queue of collisions run ahead (counted for both species)
        if (P.have_collisions && P.mode == PIC_MODE_PHYSICAL)
            b += (int64_t)MQ_SEGS * collide_seg_cap(cap) * 4 + MQ_SEGS * MQ_STRIDE * 4;
    }
    const int64_t tiles = (N + 1 + 7) / 8;
    {  // sort histogram (one counter per 128-byte line) and starts
        const int ks = P.sort_key >= 0 ? P.sort_key : tuned_row(N).sort_key;
        const int64_t nb = scan_pad((uint32_t)((tiles * tiles * 64) >> ks));
        b += 2 * nb * BIN_STRIDE * 4 + (nb + 1) * 4 + (nb / 4096 + 1) * 4;  // per species
    }
    {  // the absorbing walls' compaction: block counts and starts
        const int64_t cap = n0 + (spawning ? std::max<int64_t>(n0 / 2, 65536) : 0) + 4096 * n_inj;
        const int64_t nb = absorb_blocks_pad(std::max<int64_t>(cap, 1));
        b += (2 * nb + 1 + nb / 4096 + 1) * 4;
    }
    b += N * N * 16 + (N + 1) * (N + 1) * 16 + ((N + 1) * N + 1) * 8;  // rho, phi, E
    b += (int64_t)(60.0 * 4.0 / 3.0 * (double)(N * N)) + (1 << 20);   // multigrid levels
    if (P.graph_steps > 0 || P.time_data_freq > 0)
        b += N * N * 8 * 6 + 4 * (int64_t)std::max(P.no_bins, 1) * 8 * 2 + (1 << 20);
    b += 3 * 10000 * 4 + 2 * (int64_t)sizeof(SpCounters);
    *bytes = b + b / 20 + (8 << 20);
    return PIC_OK;
}

int pic_get_wall_counts(pic_handle *h, int32_t s, int64_t *out6) {
    CHECK(check_species(h, s));
    if (!out6) return seterr(h, PIC_ERR_ARG, "null output");
    CHECK(read_counters(h));
    for (int k = 0; k < 6; k++) {
        out6[k] = 0;
        for (int b = 0; b < CNT_SLOTS; b++) out6[k] += h->cnt_host[s].wall[b][k];
    }
    return PIC_OK;
}

// out3 = reductions performed (particle.cpp:648 count_reduce), particles
// removed by them, particles created by the source
int pic_get_population_stats(pic_handle *h, int32_t s, int64_t *out3) {
    CHECK(check_species(h, s));
    if (!out3) return seterr(h, PIC_ERR_ARG, "null output");
    CHECK(read_counters(h));
    out3[0] = h->cnt_host[s].reductions;
    out3[1] = 0;
    for (int b = 0; b < CNT_SLOTS; b++) out3[1] += h->cnt_host[s].lost[b][5];
    out3[2] = h->sp[s].injected;
    return PIC_OK;
}

// pic::getElectrons / getIons (pic.cpp:901-910) thinned on the device: the
// particles whose id is a multiple of stride, as 48-byte records
int pic_get_particle_subset(pic_handle *h, int32_t s, int32_t stride, pic_particle *out,
                            int64_t cap, int64_t *n) {
    CHECK(check_species(h, s));
    if (stride < 1 || !n || cap < 0) return seterr(h, PIC_ERR_ARG, "bad argument");
    Species &sp = h->sp[s];
    const uint32_t room = (uint32_t)std::min<int64_t>(cap, sp.cap);
    double *buf = nullptr;
    uint32_t *cnt = nullptr;
    CHECK(dalloc(h, &buf, 6 * (size_t)std::max<uint32_t>(room, 1)));
    if (dalloc(h, &cnt, 1) != PIC_OK) {
        hipFree(buf);
        return PIC_ERR_ALLOC;
    }
    launch_fill<uint32_t>(h->st, cnt, 1, 0u);
    launch_subset(h->st, sp.buf[sp.cur], &h->cnt[s].n[sp.cur], sp.cap, (uint32_t)stride, h->dx,
                  sp.weight, buf, room, cnt);
    uint32_t c = 0;
    hipError_t e = hipMemcpyAsync(&c, cnt, sizeof(c), hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    const uint32_t m = std::min(c, room);
    if (e == hipSuccess && m > 0 && out)
        e = hipMemcpy(out, buf, (size_t)m * sizeof(pic_particle), hipMemcpyDeviceToHost);
    hipFree(buf);
    hipFree(cnt);
    if (e != hipSuccess) return seterr(h, PIC_ERR_DEVICE, hipGetErrorString(e));
    *n = c;
    if (c > room) return seterr(h, PIC_ERR_CAPACITY, "output buffer too small");
    return PIC_OK;
}

static int copy_out(pic_handle *h, void *dst, const void *src, size_t bytes) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    if (!dst) return seterr(h, PIC_ERR_ARG, "null output");
    HIPCHK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return PIC_OK;
}

int pic_get_phi(pic_handle *h, double *out) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    return copy_out(h, out, h->phi, 8ull * h->N * h->N);
}
int pic_get_rho(pic_handle *h, double *out) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    return copy_out(h, out, h->rho, 8ull * h->N * h->N);
}
int pic_get_E(pic_handle *h, double *ex, double *ey) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    size_t b = 8ull * (h->N + 1) * (h->N + 1);
    CHECK(copy_out(h, ex, h->ex64, b));
    return copy_out(h, ey, h->ey64, b);
}
int pic_get_E32(pic_handle *h, float *out) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    return copy_out(h, out, h->E32, 8ull * ((h->N + 1) * h->N + 1));
}
int pic_set_E32(pic_handle *h, const float *in) {
    if (!h || !in) return set_global_error(PIC_ERR_ARG, "null argument");
    HIPCHK(h, hipMemcpyAsync(h->E32, in, 8ull * ((h->N + 1) * h->N + 1), hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return PIC_OK;
}
int pic_get_qfixed(pic_handle *h, int32_t s, int64_t *out) {
    CHECK(check_species(h, s));
    return copy_out(h, out, h->sp[s].q, 8ull * h->N * h->N);
}
int pic_get_q(pic_handle *h, int32_t s, double *out) {
    CHECK(check_species(h, s));
    int N = h->N;
    std::vector<int64_t> q((size_t)N * N);
    CHECK(pic_get_qfixed(h, s, q.data()));
    double corner = h->P.corner8 ? 8.0 : 4.0, sc = h->sp[s].scale;
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            bool ex = (i == 0 || i == N - 1), ey = (j == 0 || j == N - 1);
            double fac = (ex && ey) ? corner : ((ex || ey) ? 2.0 : 1.0);
            out[(size_t)j * N + i] = ((double)q[(size_t)j * N + i] * sc) * fac;
        }
    return PIC_OK;
}
int pic_get_step(pic_handle *h, int64_t *step) {
    if (!h || !step) return set_global_error(PIC_ERR_ARG, "null argument");
    *step = h->step;
    return PIC_OK;
}
int pic_get_nodes(pic_handle *h, int32_t *nodes) {
    if (!h || !nodes) return set_global_error(PIC_ERR_ARG, "null argument");
    *nodes = h->N;
    return PIC_OK;
}
int pic_get_loss_counts(pic_handle *h, int32_t s, int64_t *out4) {
    CHECK(check_species(h, s));
    if (!out4) return seterr(h, PIC_ERR_ARG, "null output");
    CHECK(read_counters(h));
    for (int w = 0; w < 5; w++) {
        out4[w] = 0;
        for (int b = 0; b < CNT_SLOTS; b++) out4[w] += h->cnt_host[s].lost[b][w];
    }
    return PIC_OK;
}

// ------------------------------------------------------- diagnostics
int pic_set_output_dir(pic_handle *h, const char *dir) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    h->out_dir = dir ? dir : "";
    return PIC_OK;
}

int pic_get_field_averages(pic_handle *h, double *rho, double *phi, double *ex, double *ey) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    return diag_fields(h, rho, phi, ex, ey);
}

int pic_get_density_average(pic_handle *h, int32_t s, double *out) {
    CHECK(check_species(h, s));
    if (!out) return seterr(h, PIC_ERR_ARG, "null output");
    return diag_density(h, s, out);
}

int pic_get_velocity_slices(pic_handle *h, int32_t s, double *vx, double *vy) {
    CHECK(check_species(h, s));
    return diag_velocities(h, s, vx, vy);
}

int pic_get_distributions(pic_handle *h, int32_t s, double *out) {
    CHECK(check_species(h, s));
    if (!out) return seterr(h, PIC_ERR_ARG, "null output");
    return diag_distributions(h, s, out);
}

int pic_get_time_data(pic_handle *h, double *out, int32_t cap_rows, int32_t *rows) {
    if (!h || !rows || cap_rows < 0) return set_global_error(PIC_ERR_ARG, "bad argument");
    std::vector<double> r;
    int n = 0;
    CHECK(diag_time_rows(h, r, n));
    *rows = n;
    if (out) {
        int m = n < cap_rows ? n : cap_rows;
        for (size_t k = 0; k < (size_t)m * DIAG_TCOLS; k++) out[k] = r[k];
    }
    return PIC_OK;
}

int pic_write_graphs(pic_handle *h, const char *dir) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    return diag_write_graphs(h, dir ? dir : ".");
}

// graphs::getPhiMax / getExMax / getEyMax / getQMax (graphs.cpp:1000-1112):
// out5 = max of the averaged phi, Ex, Ey, electron and ion density; each
// maximum starts from the reference's first element ([0][0]; [1][1] for q)
int pic_get_graph_max(pic_handle *h, double *out5) {
    if (!h || !out5) return set_global_error(PIC_ERR_ARG, "null argument");
    const int N = h->N;
    const size_t nn = (size_t)N * N;
    std::vector<double> a[4];
    for (auto &v : a) v.resize(nn);
    CHECK(diag_fields(h, nullptr, a[0].data(), a[1].data(), a[2].data()));
    for (int k = 0; k < 3; k++) {
        double m = a[k][0];
        for (double v : a[k]) m = v > m ? v : m;
        out5[k] = m;
    }
    for (int s = 0; s < 2; s++) {
        CHECK(diag_density(h, s, a[3].data()));
        double m = a[3][(size_t)1 * N + 1];
        for (double v : a[3]) m = v > m ? v : m;
        out5[3 + s] = m;
    }
    return PIC_OK;
}

// snapshot for offline plotting (replaces the OpenGL views): dir/snapshot_<step>_phi.txt
// (nodes x nodes, "%e ") and dir/snapshot_<step>_<e|i1>.txt ("x y vx vy vz weight"
// per particle of the id-strided subset)
int pic_write_snapshot(pic_handle *h, const char *dir, int32_t stride) {
    if (!h || stride < 1) return set_global_error(PIC_ERR_ARG, "bad argument");
    std::string d = dir ? dir : ".";
    if (!d.empty() && d.back() != '/') d += "/";
    const int N = h->N;
    std::vector<double> phi((size_t)N * N);
    CHECK(pic_get_phi(h, phi.data()));
    std::string tag = "snapshot_" + std::to_string((long long)h->step);
    FILE *f = fopen((d + tag + "_phi.txt").c_str(), "w");
    if (!f) return seterr(h, PIC_ERR_FILE, "cannot write " + d + tag + "_phi.txt");
    for (int j = 0; j < N; j++) {
        for (int i = 0; i < N; i++) fprintf(f, "%e ", phi[(size_t)j * N + i]);
        fprintf(f, "\n");
    }
    fclose(f);
    const char *type[2] = {"e", "i1"};
    for (int s = 0; s < 2; s++) {
        int64_t n = 0, cap = h->sp[s].cap / stride + 64;
        std::vector<pic_particle> p((size_t)cap);
        CHECK(pic_get_particle_subset(h, s, stride, p.data(), cap, &n));
        std::string path = d + tag + "_" + type[s] + ".txt";
        FILE *g = fopen(path.c_str(), "w");
        if (!g) return seterr(h, PIC_ERR_FILE, "cannot write " + path);
        for (int64_t k = 0; k < n; k++)
            fprintf(g, "%e %e %e %e %e %e\n", p[k].x, p[k].y, p[k].vx, p[k].vy, p[k].vz, p[k].weight);
        fclose(g);
    }
    return PIC_OK;
}

int pic_write_time_data(pic_handle *h, const char *dir) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    return diag_write_time(h, dir ? dir : ".");
}

int pic_dump_data(pic_handle *h, const char *dir, int32_t precision) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    if (precision < 1 || precision > 17) return seterr(h, PIC_ERR_ARG, "precision must be 1..17");
    std::string d = dir ? dir : ".";
    for (int s = 0; s < 2; s++) {
        int64_t n;
        CHECK(pic_particle_size(h, s, &n));
        std::vector<pic_particle> p(n);
        CHECK(pic_get_particles(h, s, p.data(), n));
        int rc = write_dump_file(d + "/dataDump_" + h->sp[s].type + ".txt", p, precision);
        if (rc != PIC_OK) return seterr(h, rc, "cannot write dataDump in " + d);
    }
    return PIC_OK;
}

int pic_profile_enable(pic_handle *h, int32_t on) {
    if (!h) return set_global_error(PIC_ERR_ARG, "null handle");
    if (!on && h->prof) prof_flush(h);
    h->prof = on != 0;
    for (double &v : h->prof_acc) v = 0.0;
    h->prof_steps.clear();
    return PIC_OK;
}

int pic_profile_steps(pic_handle *h, double *out, int32_t cap, int32_t *n) {
    if (!h || !n || (cap > 0 && !out)) return set_global_error(PIC_ERR_ARG, "null argument");
    prof_flush(h);
    int32_t k = 0;
    for (double v : h->prof_steps)
        if (k < cap) out[k++] = v;
    *n = k;
    h->prof_steps.clear();
    return PIC_OK;
}

int pic_profile_read(pic_handle *h, double *out, int32_t n) {
    if (!h || !out || n < 0) return set_global_error(PIC_ERR_ARG, "bad argument");
    prof_flush(h);
    const int m = n < 12 ? n : 12;
    for (int k = 0; k < m; k++) out[k] = h->prof_acc[k];
    for (double &v : h->prof_acc) v = 0.0;
    return PIC_OK;
}

}  // extern "C"

// ------------------------------------------------------------ checkpoint
// Exact-resume file (no counterpart in the reference, whose restart reads the
// lossy dataDump text, particle.cpp:359-401): the run's parameters, its step
// and host-side schedule state, then every array the next step reads -- both
// species' stores as they lie (empty slots included, so the layout and the
// sort schedule continue unchanged), their counters and charge, rho, phi, E,
// the tolerance solve's iterate and previous solution, the diagnostics'
// window sums and timeData rows and the accuracy-report rows.  Little-endian;
// an FNV-1a 64 checksum of all preceding bytes ends the file.
namespace {

const char CK_MAGIC[8] = {'P', 'I', 'C', '2', 'D', 'C', 'K', '1'};
constexpr uint32_t CK_VERSION = 1;

struct CkSpecies {
    int64_t n;          // particles in the store (empty slots included)
    int32_t cur;        // which of the two buffers holds them
    int32_t since_sort, holes, weighted;
    double weight, scale;
    int64_t injected;
};

struct CkHeader {
    char magic[8];
    uint32_t version, params_bytes, species_bytes, nblobs;
    int64_t step, mg_cycles_total, diag_graph_steps;
    uint64_t ovf_reported;
    double mg_relres, last_phiR, last_bc[4], acc_tol;
    int32_t mg_cycles, have_rho, have_phi, have_prev, pred, diag_rows, diag_rows_cap, acc_every;
    int64_t acc_rows;   // doubles in the accuracy-report rows
    CkSpecies sp[2];
};

struct CkBlob {
    uint32_t tag;
    void *ptr;     // device pointer, or host when host is set
    size_t bytes;
    bool host;
};

// the arrays of a checkpoint, in file order: n[s] particles per species,
// rows_cap timeData rows, acc_rows doubles of accuracy-report rows at acc
std::vector<CkBlob> ck_blobs(pic_handle *h, const int64_t n[2], int rows_cap, double *acc,
                             int64_t acc_rows) {
    std::vector<CkBlob> b;
    auto add = [&](void *p, size_t bytes, bool host = false) {
        b.push_back({(uint32_t)b.size() + 1, p, bytes, host});
    };
    const size_t N = (size_t)h->N, nn = N * N, E = (N + 1) * (N + 1);
    for (int s = 0; s < 2; s++) {
        Species &sp = h->sp[s];
        const SpView &v = sp.buf[sp.cur];
        const size_t m = 4 * (size_t)n[s];
        for (void *p : {(void *)v.cell, (void *)v.ox, (void *)v.oy, (void *)v.vx, (void *)v.vy,
                        (void *)v.vz, (void *)v.id})
            add(p, m);
        if (sp.weighted) add(v.w, m);
        add(sp.q, nn * sizeof(unsigned long long));
    }
    add(h->cnt, 2 * sizeof(SpCounters));
    add(h->rho, nn * sizeof(double));
    add(h->phi, nn * sizeof(double));
    add(h->ex64, E * sizeof(double));
    add(h->ey64, E * sizeof(double));
    add(h->E32, ((N + 1) * N + 1) * sizeof(float2));
    add(h->M.d.u[h->M.ng - 1], nn * sizeof(double));
    if (h->M.uprev) add(h->M.uprev, nn * sizeof(double));
    const DiagDev &D = h->diag;
    if (D.rho) {
        for (double *p : {D.rho, D.phi, D.ex, D.ey}) add(p, nn * sizeof(double));
        for (int s = 0; s < 2; s++) {
            add(D.q[s], nn * sizeof(double));
            add(D.vsx[s], N * sizeof(long long));
            add(D.vsy[s], N * sizeof(long long));
            add(D.cntx[s], N * sizeof(uint32_t));
            add(D.cnty[s], N * sizeof(uint32_t));
            add(D.vpx[s], N * sizeof(double));
            add(D.vpy[s], N * sizeof(double));
            add(D.dist[s], 4 * (size_t)h->P.no_bins * sizeof(unsigned long long));
        }
    }
    if (D.tdata) add(D.tdata, (size_t)rows_cap * DIAG_TCOLS * sizeof(double));
    if (acc_rows > 0) add(acc, (size_t)acc_rows * sizeof(double), true);
    return b;
}

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void *p, size_t n) {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        for (size_t k = 0; k < n; k++) h = (h ^ c[k]) * 1099511628211ull;
    }
};

constexpr size_t CK_CHUNK = (size_t)64 << 20;

// parameters that may differ between the writing and the reading handle:
// how long pic_run runs, whether it writes graph files, graph replay
pic_params ck_comparable(const pic_params &p) {
    pic_params q = p;
    q.run_time = 0;
    q.write_graphs = 0;
    q.graphs = 0;
    return q;
}

}  // namespace

int pic_checkpoint_write(pic_handle *h, const char *path) {
    if (!h || !path) return set_global_error(PIC_ERR_ARG, "null argument");
    HIPCHK(h, hipStreamSynchronize(h->st));
    CHECK(check_overflow(h, true));
    CHECK(read_counters(h));
    CHECK(mg_harvest(h));
    CkHeader H;
    memset(&H, 0, sizeof(H));
    memcpy(H.magic, CK_MAGIC, 8);
    H.version = CK_VERSION;
    H.params_bytes = sizeof(pic_params);
    H.species_bytes = sizeof(CkSpecies);
    H.step = h->step;
    H.mg_cycles_total = h->mg_cycles_total;
    H.diag_graph_steps = h->diag_graph_steps;
    H.ovf_reported = h->ovf_reported;
    H.mg_relres = h->mg_relres;
    H.last_phiR = h->last_phiR;
    for (int k = 0; k < 4; k++) H.last_bc[k] = h->last_bc[k];
    H.acc_tol = h->acc_tol;
    H.mg_cycles = h->mg_cycles;
    H.have_rho = h->have_rho;
    H.have_phi = h->have_phi;
    H.have_prev = h->M.have_prev;
    H.pred = h->M.pred;
    H.diag_rows = h->diag_rows;
    H.diag_rows_cap = h->diag_rows_cap;
    H.acc_every = h->acc_every;
    H.acc_rows = (int64_t)h->acc_rows.size();
    int64_t n[2];
    for (int s = 0; s < 2; s++) {
        const Species &sp = h->sp[s];
        n[s] = h->cnt_host[s].n[sp.cur];
        H.sp[s] = CkSpecies{n[s], sp.cur, sp.since_sort, sp.holes ? 1 : 0, sp.weighted ? 1 : 0,
                            sp.weight, sp.scale, sp.injected};
    }
    const std::vector<CkBlob> blobs =
        ck_blobs(h, n, h->diag_rows_cap, h->acc_rows.data(), H.acc_rows);
    H.nblobs = (uint32_t)blobs.size();
    FILE *f = fopen(path, "wb");
    if (!f) return seterr(h, PIC_ERR_FILE, std::string("cannot write ") + path);
    Fnv fnv;
    bool ok = true;
    auto put = [&](const void *p, size_t bytes) {
        if (ok && fwrite(p, 1, bytes, f) != bytes) ok = false;
        fnv.add(p, bytes);
    };
    put(&H, sizeof(H));
    put(&h->P, sizeof(pic_params));
    std::vector<char> buf;
    for (const CkBlob &b : blobs) {
        const uint64_t hdr[2] = {b.tag, b.bytes};
        put(hdr, sizeof(hdr));
        for (size_t off = 0; off < b.bytes && ok; off += CK_CHUNK) {
            const size_t k = std::min(CK_CHUNK, b.bytes - off);
            if (b.host) {
                put(static_cast<const char *>(b.ptr) + off, k);
                continue;
            }
            buf.resize(k);
            if (hipMemcpy(buf.data(), static_cast<const char *>(b.ptr) + off, k,
                          hipMemcpyDeviceToHost) != hipSuccess) {
                fclose(f);
                return seterr(h, PIC_ERR_DEVICE, "checkpoint: device copy failed");
            }
            put(buf.data(), k);
        }
    }
    const uint64_t sum = fnv.h;
    if (ok && fwrite(&sum, 1, sizeof(sum), f) != sizeof(sum)) ok = false;
    if (fclose(f) != 0) ok = false;
    if (!ok) return seterr(h, PIC_ERR_FILE, std::string("write failed: ") + path);
    return PIC_OK;
}

int pic_checkpoint_params(const char *path, pic_params *out) {
    if (!path || !out) return set_global_error(PIC_ERR_ARG, "null argument");
    FILE *f = fopen(path, "rb");
    if (!f) return set_global_error(PIC_ERR_FILE, std::string("cannot open ") + path);
    CkHeader H;
    const bool ok = fread(&H, 1, sizeof(H), f) == sizeof(H) && memcmp(H.magic, CK_MAGIC, 8) == 0 &&
                    H.version == CK_VERSION && H.params_bytes == sizeof(pic_params) &&
                    fread(out, 1, sizeof(pic_params), f) == sizeof(pic_params);
    fclose(f);
    if (!ok) return set_global_error(PIC_ERR_FORMAT, std::string("not a checkpoint: ") + path);
    return PIC_OK;
}

int pic_checkpoint_read(pic_handle *h, const char *path) {
    if (!h || !path) return set_global_error(PIC_ERR_ARG, "null argument");
    FILE *f = fopen(path, "rb");
    if (!f) return seterr(h, PIC_ERR_FILE, std::string("cannot open ") + path);
    struct Closer {
        FILE *f;
        ~Closer() { fclose(f); }
    } closer{f};
    Fnv fnv;
    auto get = [&](void *p, size_t bytes) {
        if (fread(p, 1, bytes, f) != bytes) return false;
        fnv.add(p, bytes);
        return true;
    };
    HIPCHK(h, hipStreamSynchronize(h->st));
    CkHeader H;
    pic_params P;
    if (!get(&H, sizeof(H)) || memcmp(H.magic, CK_MAGIC, 8) != 0 || H.version != CK_VERSION ||
        H.params_bytes != sizeof(pic_params) || H.species_bytes != sizeof(CkSpecies) ||
        !get(&P, sizeof(P)))
        return seterr(h, PIC_ERR_FORMAT, std::string("not a checkpoint: ") + path);
    const pic_params a = ck_comparable(P), b = ck_comparable(h->P);
    if (memcmp(&a, &b, sizeof(pic_params)) != 0)
        return seterr(h, PIC_ERR_ARG, "checkpoint: the handle was created with other parameters");
    int64_t n[2];
    for (int s = 0; s < 2; s++) {
        const CkSpecies &c = H.sp[s];
        if (c.n < 0 || c.n > h->sp[s].cap || c.cur < 0 || c.cur > 1)
            return seterr(h, PIC_ERR_FORMAT, "checkpoint: particle count beyond the store");
        n[s] = c.n;
    }
    if (H.acc_rows < 0 || H.acc_rows % 6 != 0 || H.acc_rows > ((int64_t)1 << 30) ||
        H.diag_rows < 0 || H.diag_rows > H.diag_rows_cap)
        return seterr(h, PIC_ERR_FORMAT, "checkpoint: bad header");
    // the reading handle takes the writer's schedule, weights and row space
    for (int s = 0; s < 2; s++) {
        Species &sp = h->sp[s];
        if (H.sp[s].weighted) CHECK(enable_weights(h, sp));
        sp.cur = H.sp[s].cur;
    }
    if (diag_time_on(h) && H.diag_rows_cap > h->diag_rows_cap)
        CHECK(diag_prepare_time(h, (H.diag_rows_cap - 1) * h->P.time_data_freq));
    std::vector<double> acc((size_t)H.acc_rows);
    std::vector<CkBlob> blobs = ck_blobs(h, n, H.diag_rows_cap, acc.data(), H.acc_rows);
    if (H.nblobs != blobs.size())
        return seterr(h, PIC_ERR_FORMAT, "checkpoint: array list differs from the run's");
    std::vector<char> buf;
    for (const CkBlob &bl : blobs) {
        uint64_t hdr[2];
        if (!get(hdr, sizeof(hdr)) || hdr[0] != bl.tag || hdr[1] != bl.bytes)
            return seterr(h, PIC_ERR_FORMAT,
                          "checkpoint: array " + std::to_string(bl.tag) + " differs from the run's");
        for (size_t off = 0; off < bl.bytes; off += CK_CHUNK) {
            const size_t k = std::min(CK_CHUNK, bl.bytes - off);
            char *dst = static_cast<char *>(bl.ptr) + off;
            if (!get(bl.host ? dst : (buf.resize(k), buf.data()), k))
                return seterr(h, PIC_ERR_FORMAT, "checkpoint: truncated");
            if (!bl.host) HIPCHK(h, hipMemcpy(dst, buf.data(), k, hipMemcpyHostToDevice));
        }
    }
    uint64_t sum = 0;
    if (fread(&sum, 1, sizeof(sum), f) != sizeof(sum) || sum != fnv.h)
        return seterr(h, PIC_ERR_FORMAT, "checkpoint: checksum mismatch");
    h->step = H.step;
    h->mg_cycles_total = H.mg_cycles_total;
    h->diag_graph_steps = H.diag_graph_steps;
    h->ovf_reported = H.ovf_reported;
    h->mg_relres = H.mg_relres;
    h->last_phiR = H.last_phiR;
    for (int k = 0; k < 4; k++) h->last_bc[k] = H.last_bc[k];
    h->acc_tol = H.acc_tol;
    h->mg_cycles = H.mg_cycles;
    h->have_rho = H.have_rho != 0;
    h->have_phi = H.have_phi != 0;
    h->M.have_prev = H.have_prev != 0;
    h->M.pred = H.pred;
    h->diag_rows = H.diag_rows;
    h->acc_every = H.acc_every;
    h->acc_rows.swap(acc);
    for (int s = 0; s < 2; s++) {
        Species &sp = h->sp[s];
        sp.since_sort = H.sp[s].since_sort;
        sp.holes = H.sp[s].holes != 0;
        sp.weight = H.sp[s].weight;
        sp.scale = H.sp[s].scale;
        sp.injected = H.sp[s].injected;
    }
    CHECK(read_counters(h));
    return PIC_OK;
}
