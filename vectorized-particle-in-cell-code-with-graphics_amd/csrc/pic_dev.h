// pic_dev.h -- device helpers shared by the particle kernels.
//
// Particle position = packed int16 cell (jx | jy << 16) + fp32 offset in
// [0,1) per axis, i.e. x = (jx + ox) * dx.  The reference keeps absolute FP64
// positions and recomputes jp = (int)(x/dx) in every kernel (pic.cl:31-37,
// 66-76, 108-112); storing the cell makes the per-cell counting sort free of
// divisions and keeps the offsets at full fp32 resolution.
//
// All arithmetic below is written out operation by operation; the library is
// compiled with -ffp-contract=off so every product and sum rounds once, as
// IEEE single/double precision.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PIC_DEV_MODE_REF 0
#define PIC_DEV_MODE_PHYS 1
// reference semantics with the steady-state walls: particles outside are
// never removed (they wait, frozen, for the pairing pass); otherwise REF
#define PIC_DEV_MODE_REF_STEADY 2

#define ST_KEPT 0
#define ST_LEFT 1
#define ST_RIGHT 2
#define ST_BOTTOM 3
#define ST_TOP 4
#define ST_DEAD 5  // removed by a collision (reference ionisation rule)
#define ST_HOLE 6  // empty slot of the store (see CELL_HOLE)

// cell index of a particle a collision removed (outside every wall)
#define CELL_DEAD (-16001)
// cell index of an empty slot: a particle removed by an in-place step stays
// as a hole until the next sort pass compacts the store; holes are never
// collided, moved, deposited or counted again
#define CELL_HOLE (-16002)

#define POS_GUARD 8192.0f
#define CELL_FAR 16000

// fixed-point deposit: weights quantised to 2^-24 of a cell, 2^48 per
// particle (an fp32 offset in [0.5, 1) has exactly 24 fraction bits, so the
// weights keep every bit of the offset there); an int64 node accumulator
// holds 2^15 unit-weight particles of one species without overflow
#define FX_BITS 24
#define FX_SCALE_F 16777216.0f
#define FX_ONE (1ll << FX_BITS)

// per-species constants of one push (particle.cpp:1103-1112 in fp32)
struct PushConst {
    int N;      // nodes per side
    int mode;   // PIC_DEV_MODE_*
    int bfield; // magnetic rotation on/off (pic.cl:121)
    int tilesX; // sort-key tiles per side
    int keyshift; // sort key = cell key >> keyshift (0: cells, 6: 8x8-cell tiles)
    float cE;   // 0.5 * dt * q/m
    float dtdx, dtdy;
    float cos_a, sin_a, cos_w, sin_w;
};

__host__ __device__ inline uint32_t cell_pack(int jx, int jy) {
    return (uint32_t)(uint16_t)(int16_t)jx | ((uint32_t)(uint16_t)(int16_t)jy << 16);
}
__host__ __device__ inline int cell_jx(uint32_t c) { return (int)(int16_t)(c & 0xFFFFu); }
__host__ __device__ inline int cell_jy(uint32_t c) { return (int)(int16_t)(c >> 16); }

// new cell and offset from cell j and unnormalised offset t (cells)
__device__ __forceinline__ void reposition(int j, float t, int &jn, float &on) {
    if (!(t > -POS_GUARD && t < POS_GUARD)) {  // NaN or a jump far off the grid
        jn = (t > 0.0f) ? CELL_FAR : -CELL_FAR;
        on = 0.0f;
        return;
    }
    float k = floorf(t);
    float f = t - k;
    if (f >= 1.0f) {  // t just below an integer: t - k rounded up to 1
        f = 0.0f;
        k = k + 1.0f;
    }
    int jj = j + (int)k;
    jj = jj > CELL_FAR ? CELL_FAR : jj;
    jj = jj < -CELL_FAR ? -CELL_FAR : jj;
    jn = jj;
    on = f;
}

// pic.cl:74 / 112: jp < N && x > 0  (truncation of x/dx, so x in (0, N*dx))
__device__ __forceinline__ bool ref_inside_push(int j, float o, int N) {
    return j >= 0 && j <= N - 1 && !(j == 0 && o == 0.0f);
}

struct PState {
    int jx, jy;
    float ox, oy, vx, vy, vz;
};

// pic.cl:74 / pic.cl:112: only particles strictly inside (0, N*dx) move (REF);
// in PHYS mode every particle inside [0, L) moves (others leave at the next
// wall pass and are never gathered, so no read leaves the field array).
__device__ __forceinline__ bool can_move(const PState &s, const PushConst &pc) {
    if (pc.mode == PIC_DEV_MODE_PHYS)
        return s.jx >= 0 && s.jx <= pc.N - 2 && s.jy >= 0 && s.jy <= pc.N - 2;
    return ref_inside_push(s.jx, s.ox, pc.N) && ref_inside_push(s.jy, s.oy, pc.N);
}

// CIC gather of (Ex, Ey) at the particle (pic.cl:75-88).  E is the reference's
// padded stride-N node layout, so a particle in the last column reads the
// next row's first entry exactly as the reference does.
__device__ __forceinline__ void gather_one(const PState &s, const float2 *__restrict__ E,
                                           const PushConst &pc, float &epx, float &epy) {
    const int N = pc.N;
    float wj = s.ox, wk = s.oy;
    float a4 = (1.0f - wj) * (1.0f - wk);
    float a3 = wj * (1.0f - wk);
    float a2 = wk * (1.0f - wj);
    float a1 = wj * wk;
    long b = (long)s.jy * N + s.jx;
    float2 e00 = E[b], e01 = E[b + 1], e10 = E[b + N], e11 = E[b + N + 1];
    epx = a4 * e00.x + a3 * e01.x + a2 * e10.x + a1 * e11.x;
    epy = a4 * e00.y + a3 * e01.y + a2 * e10.y + a1 * e11.y;
}

// half kick, optional rotation about B, half kick, drift (pic.cl:114-146)
__device__ __forceinline__ void move_one(PState &s, float epx, float epy, const PushConst &pc) {
    float ux = s.vx + pc.cE * epx;
    float uy = s.vy + pc.cE * epy;
    float uz = s.vz;
    if (pc.bfield) {
        float vpar = pc.cos_a * ux + pc.sin_a * uy;
        float vperp = -pc.sin_a * ux + pc.cos_a * uy;
        float vperp_after = vperp * pc.cos_w + pc.sin_w * uz;
        uz = -vperp * pc.sin_w + pc.cos_w * uz;
        ux = pc.cos_a * vpar - pc.sin_a * vperp_after;
        uy = pc.sin_a * vpar + pc.cos_a * vperp_after;
    }
    ux = ux + pc.cE * epx;
    uy = uy + pc.cE * epy;
    float tx = s.ox + ux * pc.dtdx;
    float ty = s.oy + uy * pc.dtdy;
    reposition(s.jx, tx, s.jx, s.ox);
    reposition(s.jy, ty, s.jy, s.oy);
    s.vx = ux;
    s.vy = uy;
    s.vz = uz;
}

__device__ __forceinline__ bool push_one(PState &s, const float2 *__restrict__ E,
                                         const PushConst &pc) {
    if (!can_move(s, pc)) return false;
    float epx, epy;
    gather_one(s, E, pc, epx, epy);
    move_one(s, epx, epy, pc);
    return true;
}

// Wall test after the push.  REF: boundary.cpp:595-599 / 774-778, lost iff
// (int)(x/dx) < 0 or >= N, i.e. x <= -dx or x >= N*dx.  PHYS: x < 0 or x >= L.
__device__ __forceinline__ int classify(const PState &s, int N, int mode) {
    if (s.jx == CELL_HOLE) return ST_HOLE;
    if (s.jx == CELL_DEAD) return ST_DEAD;
    if (mode == PIC_DEV_MODE_REF_STEADY) return ST_KEPT;
    bool lx, ly;
    if (mode == PIC_DEV_MODE_PHYS) {
        lx = (s.jx < 0) || (s.jx > N - 2);
        ly = (s.jy < 0) || (s.jy > N - 2);
    } else {
        lx = (s.jx < -1) || (s.jx == -1 && s.ox == 0.0f) || (s.jx >= N);
        ly = (s.jy < -1) || (s.jy == -1 && s.oy == 0.0f) || (s.jy >= N);
    }
    if (lx) return s.jx < 0 ? ST_LEFT : ST_RIGHT;
    if (ly) return s.jy < 0 ? ST_BOTTOM : ST_TOP;
    return ST_KEPT;
}

// pic.cl:35: deposit only for jp < N-1 && x > 0 (REF); every particle inside
// [0, L) (PHYS)
__device__ __forceinline__ bool deposit_ok(const PState &s, int N, int mode) {
    if (mode == PIC_DEV_MODE_PHYS) return s.jx >= 0 && s.jx <= N - 2 && s.jy >= 0 && s.jy <= N - 2;
    return s.jx >= 0 && s.jx <= N - 2 && !(s.jx == 0 && s.ox == 0.0f) && s.jy >= 0 &&
           s.jy <= N - 2 && !(s.jy == 0 && s.oy == 0.0f);
}

// fixed-point CIC corner weights (Birdsall & Langdon bilinear weighting, as in
// pic.cl:40-43), summing to exactly 2^48
struct FxW {
    unsigned long long a4, a3, a2, a1;
};
// With S = 2^24 and a1 = wx wy (one 32 x 32 -> 64-bit product), the other
// corners are exactly S wx - a1, S wy - a1 and S^2 - S wx - S wy + a1.
__device__ __forceinline__ FxW fx_weights(float ox, float oy) {
    const int wx = (int)rintf(ox * FX_SCALE_F);
    const int wy = (int)rintf(oy * FX_SCALE_F);
    const long long a1 = (long long)wx * (long long)wy;
    const long long sx = (long long)wx << FX_BITS, sy = (long long)wy << FX_BITS;
    FxW w;
    w.a4 = (unsigned long long)(FX_ONE * FX_ONE - sx - sy + a1);
    w.a3 = (unsigned long long)(sx - a1);
    w.a2 = (unsigned long long)(sy - a1);
    w.a1 = (unsigned long long)a1;
    return w;
}

// weighted fixed-point corner weights: the exact integer products scaled by
// the particle's weight factor and rounded once (w == 1 gives fx_weights)
__device__ __forceinline__ FxW fx_weights_w(float ox, float oy, float w) {
    FxW f = fx_weights(ox, oy);
    if (w != 1.0f) {
        const double d = (double)w;
        f.a4 = (unsigned long long)llrint((double)f.a4 * d);
        f.a3 = (unsigned long long)llrint((double)f.a3 * d);
        f.a2 = (unsigned long long)llrint((double)f.a2 * d);
        f.a1 = (unsigned long long)llrint((double)f.a1 * d);
    }
    return f;
}

// four corner atomics into the fixed-point charge
__device__ __forceinline__ void dep_global(unsigned long long *q, int N, int jx, int jy,
                                           const FxW &w) {
    long b = (long)jy * N + jx;
    atomicAdd(&q[b], w.a4);
    atomicAdd(&q[b + 1], w.a3);
    atomicAdd(&q[b + N], w.a2);
    atomicAdd(&q[b + N + 1], w.a1);
}

// counting-sort key on the (N+1)^2 halo grid (jx, jy in [-1, N-1]): one bin
// per cell, ordered tile-major with 8x8 tiles so a contiguous key range is a
// compact patch (the sort key drops the low bits: one bin per tile).  A cell
// off the tile grid (a particle the steady-state walls leave outside) takes
// the nearest edge cell's key, so every key stays inside the histogram.
__device__ __forceinline__ uint32_t bin_key(int jx, int jy, int tilesX) {
    const int top = 8 * tilesX - 1;
    const uint32_t hx = (uint32_t)min(max(jx + 1, 0), top);
    const uint32_t hy = (uint32_t)min(max(jy + 1, 0), top);
    return ((((hy >> 3) * (uint32_t)tilesX + (hx >> 3)) << 6) | ((hy & 7u) << 3) | (hx & 7u));
}

// ----------------------------------------------------------- Philox4x32-10
__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Philox2x32-10 (Salmon et al., SC'11): two words from a two-word counter and
// a one-word key, one 32 x 32 -> 64-bit product per round
__host__ __device__ inline void philox2x32_10(uint32_t c0, uint32_t c1, uint32_t k, uint32_t out[2]) {
    for (int r = 0; r < 10; r++) {
        uint64_t p = (uint64_t)0xD256D193u * c0;
        uint32_t hi = (uint32_t)(p >> 32), lo = (uint32_t)p;
        c0 = hi ^ k ^ c1;
        c1 = lo;
        k += 0x9E3779B9u;
    }
    out[0] = c0;
    out[1] = c1;
}

// uniform in (0,1), exact in fp32
__host__ __device__ inline float u01(uint32_t x) {
    return ((float)(x >> 9) + 0.5f) * (1.0f / 8388608.0f);
}
__host__ __device__ inline double u01d(uint32_t hi, uint32_t lo) {
    // 53-bit uniform in (0,1): 27 bits of hi above 26 bits of lo
    uint64_t m = ((uint64_t)(hi >> 5) << 26) | (uint64_t)(lo >> 6);
    return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}
