// k_refcx.hip -- the reference's ion charge exchange (reference mode).
//
// ion_collisions::collision (ion_collisions.cpp:255-308) is not a Monte-Carlo
// test per ion: every eligible ion adds 1 - exp(-u dt 1e-20 n_g sigma) to an
// accumulator of its grid column (probabilityCX[jp], N entries, kept across
// steps); when the accumulator exceeds 1 the ion takes a neutral's velocity
// from the table and the accumulator restarts at 0.  The result depends on
// the order in which ions are visited, so the order is defined here as
// (column, id): a radix sort of (jp << 32 | id) keys, then one wavefront per
// column runs the accumulator sequentially in fp32.
#include <hipcub/hipcub.hpp>

#include "pic_internal.h"

namespace pic {

#define REFCX_INVALID (~0ull)

// sigma [1e-20 m^2] of the reference's charge-exchange choice
// (pic.cpp:363-378; ion_collisions.cpp:157-199)
__device__ __forceinline__ float refcx_sigma(const mcc::MccConst &c, float E) {
    if (!c.sig_cx_energy) return c.sig_cx_const;
    if (c.gas == mcc::GAS_AR) return mcc::ar_cx(E);
    // sigma_H_cx divides the energy by AMU in kilograms (ion_collisions.cpp:188-189)
    return mcc::h_cx_janev(E / 1000.0f / 1.66054e-27f);
}

// keys (jp << 32 | id) of eligible ions (ion_collisions.cpp:273-277: (int)(x/dx)
// and (int)(y/dy) in [0, N)), value = probability bits << 32 | store index
__global__ void __launch_bounds__(256)
    k_refcx_keys(SpView sp, const uint32_t *__restrict__ nptr, uint32_t cap, int N,
                 mcc::MccConst mc, unsigned long long *__restrict__ key,
                 unsigned long long *__restrict__ val) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap) return;
    unsigned long long k = REFCX_INVALID, v = 0ull;
    if (i < *nptr) {
        const uint32_t c = sp.cell[i];
        const int jx = cell_jx(c), jy = cell_jy(c);
        const float ox = sp.ox[i], oy = sp.oy[i];
        // truncation of x/dx: x in (-dx, N dx) gives jp = 0 .. N-1
        const bool inx = jx >= -1 && jx <= N - 1 && !(jx == -1 && ox == 0.0f);
        const bool iny = jy >= -1 && jy <= N - 1 && !(jy == -1 && oy == 0.0f);
        if (jx != CELL_HOLE && jx != CELL_DEAD && inx && iny) {
            const float vx = sp.vx[i], vy = sp.vy[i], vz = sp.vz[i];
            const float E = (0.5f / mc.q) * mc.mass * (vx * vx + vy * vy + vz * vz);
            const float u = sqrtf(2.0f * mc.q * E / mc.mass);
            const float P = 1.0f - expf(-(u * mc.dt * 1.0e-20f * mc.ng * refcx_sigma(mc, E)));
            const uint32_t jp = jx < 0 ? 0u : (uint32_t)jx;  // (int) truncates toward zero
            k = ((unsigned long long)jp << 32) | sp.id[i];
            v = ((unsigned long long)__float_as_uint(P) << 32) | i;
        }
    }
    key[i] = k;
    val[i] = v;
}

// first position in sorted key[0, n) whose key >= x
__device__ __forceinline__ uint32_t refcx_lower(const unsigned long long *key, uint32_t n,
                                                unsigned long long x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (key[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// one wavefront per column: the sequential accumulator over the column's
// ions in id order (ion_collisions.cpp:297-303); a firing ion takes the
// velocity of table entry rand() % 9999 (ion_collisions.cpp:209-221), drawn
// here from the ion's Philox stream
__global__ void __launch_bounds__(64)
    k_refcx_columns(const unsigned long long *__restrict__ key,
                    const unsigned long long *__restrict__ val, uint32_t n, int N,
                    float *__restrict__ pcol, SpView sp, mcc::MccConst mc, uint32_t step,
                    unsigned long long *__restrict__ events) {
    const uint32_t col = blockIdx.x;
    if ((int)col >= N) return;
    const int lane = threadIdx.x;
    const uint32_t b = refcx_lower(key, n, (unsigned long long)col << 32);
    const uint32_t e = refcx_lower(key, n, (unsigned long long)(col + 1) << 32);
    float p = pcol[col];
    uint32_t fired = 0;
    for (uint32_t r0 = b; r0 < e; r0 += 64) {
        const uint32_t r = r0 + lane;
        const bool have = r < e;
        const unsigned long long vv = have ? val[r] : 0ull;
        const float P = __uint_as_float((uint32_t)(vv >> 32));
        const uint32_t cnt = (e - r0) < 64u ? (e - r0) : 64u;
        bool fire = false;
        for (uint32_t k = 0; k < cnt; k++) {  // wave-uniform sequential sum
            const float Pk = __shfl(P, (int)k, 64);
            p = p + Pk;
            if (p > 1.0f) {
                if ((uint32_t)lane == k) fire = true;
                p = 0.0f;
            }
        }
        if (fire) {
            const uint32_t i = (uint32_t)vv;
            const uint32_t id = (uint32_t)key[r];
            uint32_t q[4];
            philox4x32_10(id, step, 16u + RS_REFCX, RNG_DOMAIN_MCC, mc.key0, mc.key1, q);
            int idx = (int)(u01(q[0]) * (float)(mc.ntab_n - 1));
            idx = idx < 0 ? 0 : (idx > mc.ntab_n - 2 ? mc.ntab_n - 2 : idx);
            sp.vx[i] = mc.ntab[idx];
            sp.vy[i] = mc.ntab[mc.ntab_n + idx];
            sp.vz[i] = mc.ntab[2 * mc.ntab_n + idx];
        }
        fired += (uint32_t)__popcll(__ballot(fire));
    }
    if (lane == 0) {
        pcol[col] = p;
        if (fired) atomicAdd(&events[cnt_slot() * 8 + mcc::EV_CX], (unsigned long long)fired);
    }
}

size_t refcx_temp_bytes(uint32_t cap) {
    size_t bytes = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const unsigned long long *)nullptr,
                                       (unsigned long long *)nullptr,
                                       (const unsigned long long *)nullptr,
                                       (unsigned long long *)nullptr, (int)cap, 0, 48);
    return bytes;
}

int launch_refcx(hipStream_t st, const SpView &sp, const uint32_t *n, uint32_t cap, int N,
                 const mcc::MccConst &mc, uint32_t step, RefCxBuffers &B,
                 unsigned long long *events) {
    hipLaunchKernelGGL(k_refcx_keys, dim3((cap + 255) / 256), dim3(256), 0, st, sp, n, cap, N, mc,
                       B.key[0], B.val[0]);
    size_t bytes = B.temp_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(B.temp, bytes, B.key[0], B.key[1], B.val[0], B.val[1],
                                           (int)cap, 0, 48, st) != hipSuccess)
        return -1;
    hipLaunchKernelGGL(k_refcx_columns, dim3(N), dim3(64), 0, st, B.key[1], B.val[1], cap, N,
                       B.pcol, sp, mc, step, events);
    return 0;
}

}  // namespace pic
