// k_wall.hip -- surface physics of the non-steady walls: electron-impact
// secondary emission (Vaughan), reflection and backscatter of electrons,
// Ar+ impact secondary emission (boundary::e_boundary / ion_boundary,
// boundary.cpp:281-529, 687-756, set up at pic.cpp:455-505).
//
// The particle passes queue every particle that reached a wall (state after
// its push); this pass decides each one's fate with draws from its own
// Philox stream (counter (id, step, species*16 + RS_WALL + 256 k,
// 0x4D434331)) and stages what returns to the plasma:
//   reflected / backscattered electrons: key 4 << 32 | id, 5 << 32 | id (the
//     commit keeps their id);
//   true secondaries of an electron: key (8 + j) << 32 | id;
//   secondaries of an ion: key (16 + j) << 32 | ion id, in the electron stage.
// The commit sorts the stage by key, so the result does not depend on the
// order of the queue.
//
// Reference mode keeps the reference's arithmetic as written: the impact
// angle acos(v/(sqrt(2E/m) + 1e-6)) with E in eV (so about pi/2), the wall
// at N dx, re-emitted secondaries at y = 0 (x walls) or x = 0 (y walls), since
// the new particleVec's other coordinate is left at 0 (boundary.cpp:488-504,
// 696-701).  Physical mode uses the angle to the wall normal, the wall at L
// and the incident's position along the wall.
#include "pic_internal.h"

namespace pic {

#define RS_WALL 12
#define WALL_MAX_CHILDREN 8

__device__ __forceinline__ void wall_draw(const WallConst &c, int species, uint32_t id,
                                          uint32_t step, int k, uint32_t r[4]) {
    philox4x32_10(id, step, (uint32_t)species * 16u + RS_WALL + 256u * (uint32_t)k,
                  RNG_DOMAIN_MCC, c.key0, c.key1, r);
}

// uniform in (0,1) from one 32-bit word
__device__ __forceinline__ double u01w(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }

// which walls a position is outside of (the reference's (int)(x/DX) tests,
// boundary.cpp:595-599; physical mode: x < 0 or x >= L)
struct WallSides {
    bool left, right, bottom, top;
};

__device__ __forceinline__ WallSides wall_sides(int jx, float ox, int jy, float oy, const WallConst &c) {
    WallSides w;
    if (c.mode == PIC_DEV_MODE_PHYS) {
        w.left = jx < 0;
        w.right = jx > c.N - 2;
        w.bottom = jy < 0;
        w.top = jy > c.N - 2;
    } else {
        w.left = jx < -1 || (jx == -1 && ox == 0.0f);
        w.right = jx >= c.N;
        w.bottom = jy < -1 || (jy == -1 && oy == 0.0f);
        w.top = jy >= c.N;
    }
    return w;
}

// absolute position -> cell + fp32 offset
__device__ __forceinline__ void wall_cell(double x, double dx, int &j, float &o) {
    const double t = x / dx;
    const double k = floor(t);
    float f = (float)(t - k);
    int jj = (int)k;
    if (f >= 1.0f) {
        f = 0.0f;
        jj += 1;
    }
    j = jj;
    o = f;
}

struct WState {
    double x, y, vx, vy, vz;
};

// secondary_vel / i_secondary_vel (boundary.cpp:423-482, 696-760): emission
// angles chi = asin(sqrt(Rchi)), phi = 2 pi Rphi (Krimke eqs. 5-11), the
// particle placed R dt |v_n| inside the wall it is outside of.  `sw`: the
// walls the reference's test finds for the particle's position; `xb`, `yb`:
// the draws of the x and y blocks (R, Rphi, Rchi); speed_from_R:
// i_secondary_vel's speed sqrt(2 R q (thr - 2 wf) / me) from the block's R.
__device__ __forceinline__ void emit(WState &p, double speed, const WallSides &sw,
                                     const uint32_t xb[4], const uint32_t yb[4], bool speed_from_R,
                                     const WallConst &c) {
    const double two_pi = 6.283185307179586;
    if (sw.left || sw.right) {
        const double R = u01d(xb[0], xb[1]), Rphi = u01w(xb[2]), Rchi = u01w(xb[3]);
        const double v = speed_from_R ? sqrt(2.0 * R * c.q * (c.thr - 2.0 * c.wf) / c.me) : speed;
        const double chi = asin(sqrt(Rchi)), phi = two_pi * Rphi;
        p.vx = fabs(v * cos(chi));
        p.vy = v * sin(chi) * cos(phi);
        p.vz = v * sin(chi) * sin(phi);
        p.x = R * c.dt * fabs(p.vx);
        if (sw.right) {
            p.x = c.Lw - p.x;
            p.vx = -fabs(v * cos(chi));
        }
    }
    if (sw.bottom || sw.top) {
        const double R = u01d(yb[0], yb[1]), Rphi = u01w(yb[2]), Rchi = u01w(yb[3]);
        const double v = speed_from_R ? sqrt(2.0 * R * c.q * (c.thr - 2.0 * c.wf) / c.me) : speed;
        const double chi = asin(sqrt(Rchi)), phi = two_pi * Rphi;
        p.vy = fabs(v * cos(chi));
        p.vx = v * sin(chi) * cos(phi);
        p.vz = v * sin(chi) * sin(phi);
        p.y = R * c.dt * fabs(p.vy);
        if (sw.top) {
            p.y = c.Lw - p.y;
            p.vy = -fabs(v * cos(chi));
        }
    }
}

__device__ __forceinline__ PState to_pstate(const WState &p, double dx) {
    PState s;
    wall_cell(p.x, dx, s.jx, s.ox);
    wall_cell(p.y, dx, s.jy, s.oy);
    s.vx = (float)p.vx;
    s.vy = (float)p.vy;
    s.vz = (float)p.vz;
    return s;
}

__device__ __forceinline__ void wall_stage(const StageView &sv, SpCounters *cnt, const PState &s,
                                           unsigned long long key, float w) {
    const uint32_t slot = atomicAdd(&cnt->stage_count, 1u);
    if (slot >= sv.cap) {
        atomicAdd(&cnt->overflow, 1u);
        return;
    }
    sv.cell[slot] = cell_pack(s.jx, s.jy);
    sv.ox[slot] = s.ox;
    sv.oy[slot] = s.oy;
    sv.vx[slot] = s.vx;
    sv.vy[slot] = s.vy;
    sv.vz[slot] = s.vz;
    sv.key[slot] = key;
    if (sv.w) sv.w[slot] = w;
}

// Vaughan's electron-impact yield (boundary.cpp:261-282; Radmilovic eq. 14)
__device__ __forceinline__ double vaughan(double E, double theta, const WallConst &c) {
    const double two_pi = 6.283185307179586;
    const double v = (E - c.E0) / (c.Emax0 - c.E0);
    const double k = (v > 1.0) ? 0.25 : 0.62;
    const double fac = 1.0 + k * (theta * theta / two_pi);
    const double w = (E - c.E0) / (c.Emax0 * fac - c.E0);
    return c.max_yield * fac * pow(w * exp(1.0 - w), k);
}

// Ar+ impact yield (boundary.cpp:292-299; Radmilovic eq. 16)
__device__ __forceinline__ double ar_ion_yield(double E) {
    if (!(E >= 80.0)) return 0.0;
    return 0.006 * E / (1.0 + E / 10.0) + 1.05e-4 * (sqrt(E - 80.0) / (1.0 + pow(E / 8000.0, 1.5)));
}

// one queued electron (boundary::e_boundary, boundary.cpp:586-677);
// returns the event: 1 absorbed, 2 reflected, 3 backscattered, 4 emitting
__device__ int wall_electron(const PState &s, uint32_t id, float w, uint32_t step,
                             const WallConst &c, const StageView &se, SpCounters *ce,
                             uint32_t &created) {
    created = 0;
    const WallSides sw = wall_sides(s.jx, s.ox, s.jy, s.oy, c);
    const bool onx = sw.left || sw.right;
    WState p{((double)s.jx + (double)s.ox) * c.dx, ((double)s.jy + (double)s.oy) * c.dx,
             (double)s.vx, (double)s.vy, (double)s.vz};
    const double u2 = p.vx * p.vx + p.vy * p.vy + p.vz * p.vz;
    const double E = (0.5 / c.q) * c.me * u2;  // calc_energy: eV
    double yield = 0.0;
    if (c.see_e) {
        double theta;
        if (c.mode == PIC_DEV_MODE_PHYS) {
            const double vn = onx ? fabs(p.vx) : fabs(p.vy), vm = sqrt(u2);
            theta = vm > 0.0 ? acos(fmin(1.0, vn / vm)) : 0.0;
            yield = E > c.E0 ? vaughan(E, theta, c) : 0.0;
        } else {
            // calc_angleX/Y (boundary.cpp:197-218): 2 E / m with E in eV
            const double uu = 2.0 * E / c.me;
            theta = onx ? acos(p.vx / (sqrt(uu) + 1.0e-6)) : acos(p.vy / (sqrt(uu) + 1.0e-6));
            yield = vaughan(E, theta, c);  // NaN below E0: every test below fails
        }
    }
    uint32_t a[4];
    wall_draw(c, 0, id, step, 0, a);
    const double R = u01d(a[0], a[1]), R2 = u01d(a[2], a[3]);
    if (!(R < yield)) return 1;  // absorbed
    const double refl = (c.refl && E > 50.0) ? 0.1 * yield : 0.0;
    const double back = (c.back && E > 50.0) ? 0.1 * yield : 0.0;
    int no_inject = (int)floor(fabs(yield - refl - back)) + 1;
    if (no_inject > WALL_MAX_CHILDREN) no_inject = WALL_MAX_CHILDREN;
    uint32_t b[4];
    wall_draw(c, 0, id, step, 1, b);
    if (R2 < refl / (yield + 1.0e-6)) {
        // reflect_electron (boundary.cpp:367-411): one R for both axes
        const double Rp = u01d(b[0], b[1]);
        if (sw.right) {
            p.vx = -p.vx;
            p.x = c.Lw - Rp * c.dt * fabs(p.vx);
        } else if (sw.left) {
            p.vx = -p.vx;
            p.x = Rp * c.dt * fabs(p.vx);
        }
        if (sw.top) {
            p.vy = -p.vy;
            p.y = c.Lw - Rp * c.dt * fabs(p.vy);
        } else if (sw.bottom) {
            p.vy = -p.vy;
            p.y = Rp * c.dt * fabs(p.vy);
        }
        wall_stage(se, ce, to_pstate(p, c.dx), (4ull << 32) | id, w);
        return 2;
    }
    if (R2 < (refl + back) / (yield + 1.0e-6)) {
        // backscatter: speed (rand() % 10000) / 10000 of the incident's
        const double R4 = floor(u01d(b[2], b[3]) * 10000.0) / 10000.0;
        const double speed = R4 * sqrt((2.0 * c.q * E) / c.me);
        uint32_t xb[4], yb[4];
        wall_draw(c, 0, id, step, 2, xb);
        wall_draw(c, 0, id, step, 3, yb);
        emit(p, speed, sw, xb, yb, false, c);
        wall_stage(se, ce, to_pstate(p, c.dx), (5ull << 32) | id, w);
        return 3;
    }
    // true secondaries (secondary_electrons, boundary.cpp:487-529), the
    // incident electron removed
    for (int j = 0; j < no_inject; j++) {
        uint32_t v[4], xb[4], yb[4];
        wall_draw(c, 0, id, step, 4 + 3 * j, v);
        wall_draw(c, 0, id, step, 5 + 3 * j, xb);
        wall_draw(c, 0, id, step, 6 + 3 * j, yb);
        // the new particle's walls: the incident's (physical); in the
        // reference its position is (x, 0) or (0, y), outside one wall only
        WState q;
        WallSides qs = sw;
        if (c.mode == PIC_DEV_MODE_PHYS) {
            q = p;
        } else {
            q = WState{0.0, 0.0, 0.0, 0.0, 0.0};
            if (onx) {
                q.x = p.x;
                qs.bottom = qs.top = false;
            } else {
                q.y = p.y;
                qs.left = qs.right = false;
            }
        }
        const double Rv = u01d(v[0], v[1]);
        const double speed = sqrt((2.0 * Rv * c.q * (c.thr - 2.0 * c.wf) / c.me));
        emit(q, speed, qs, xb, yb, false, c);
        wall_stage(se, ce, to_pstate(q, c.dx), ((unsigned long long)(8 + j) << 32) | id, w);
    }
    created = (uint32_t)no_inject;
    return 4;
}

// one queued ion (boundary::ion_boundary, boundary.cpp:771-824): absorbed,
// after emitting floor(yield) + 1 electrons with probability yield;
// returns 1 absorbed, 4 emitting
__device__ int wall_ion(const PState &s, uint32_t id, float w, uint32_t step, const WallConst &c,
                        const StageView &se, SpCounters *ce, uint32_t &created) {
    created = 0;
    if (!c.see_i) return 1;
    const double x = ((double)s.jx + (double)s.ox) * c.dx, y = ((double)s.jy + (double)s.oy) * c.dx;
    const double vx = s.vx, vy = s.vy, vz = s.vz;
    const double E = (0.5 / c.q) * c.mi * (vx * vx + vy * vy + vz * vz);
    const double yield = ar_ion_yield(E);
    uint32_t a[4];
    wall_draw(c, 1, id, step, 0, a);
    if (!(u01d(a[0], a[1]) < yield)) return 1;
    int no_inject = (int)floor(yield) + 1;
    if (no_inject > WALL_MAX_CHILDREN) no_inject = WALL_MAX_CHILDREN;
    const WallSides sw = wall_sides(s.jx, s.ox, s.jy, s.oy, c);
    for (int j = 0; j < no_inject; j++) {
        uint32_t xb[4], yb[4];
        wall_draw(c, 1, id, step, 5 + 3 * j, xb);
        wall_draw(c, 1, id, step, 6 + 3 * j, yb);
        // reference: the new electron's coordinates start at 0 and only the
        // wall coordinate is set; physical: it starts at the ion's position
        WState q = c.mode == PIC_DEV_MODE_PHYS ? WState{x, y, 0.0, 0.0, 0.0}
                                                : WState{0.0, 0.0, 0.0, 0.0, 0.0};
        emit(q, 0.0, sw, xb, yb, true, c);
        wall_stage(se, ce, to_pstate(q, c.dx), ((unsigned long long)(16 + j) << 32) | id, w);
    }
    created = (uint32_t)no_inject;
    return 4;
}

__global__ void __launch_bounds__(256)
    k_wall(int species, StageView wq, SpCounters *__restrict__ cw, StageView se,
           SpCounters *__restrict__ ce, WallConst wc, uint32_t step) {
    __shared__ uint32_t sc[6];
    if (threadIdx.x < 6) sc[threadIdx.x] = 0;
    __syncthreads();
    uint32_t n = cw->wall_count;
    if (n > wq.cap) n = wq.cap;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        PState s;
        const uint32_t c = wq.cell[i];
        s.jx = cell_jx(c);
        s.jy = cell_jy(c);
        s.ox = wq.ox[i];
        s.oy = wq.oy[i];
        s.vx = wq.vx[i];
        s.vy = wq.vy[i];
        s.vz = wq.vz[i];
        const uint32_t id = (uint32_t)wq.key[i];
        const float w = wq.w ? wq.w[i] : 1.0f;
        uint32_t created = 0;
        const int ev = species == 0 ? wall_electron(s, id, w, step, wc, se, ce, created)
                                    : wall_ion(s, id, w, step, wc, se, ce, created);
        atomicAdd(&sc[0], 1u);
        atomicAdd(&sc[ev], 1u);
        if (created) atomicAdd(&sc[5], created);
    }
    __syncthreads();
    if (threadIdx.x < 6 && sc[threadIdx.x])
        atomicAdd(&cw->wall[cnt_slot()][threadIdx.x], sc[threadIdx.x]);
}

// decide the fate of every queued particle of `species`; the queue is then
// emptied (a fill kernel, so the sequence can be captured in a graph)
void launch_wall(hipStream_t st, int species, const StageView &wq, SpCounters *cw,
                 const StageView &se, SpCounters *ce, const WallConst &wc, uint32_t step) {
    const unsigned blocks = (unsigned)std::min<int64_t>(((int64_t)wq.cap + 255) / 256, 1024);
    hipLaunchKernelGGL(k_wall, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, st, species, wq, cw, se,
                       ce, wc, step);
    launch_fill<uint32_t>(st, &cw->wall_count, 1, 0u);
}

}  // namespace pic
