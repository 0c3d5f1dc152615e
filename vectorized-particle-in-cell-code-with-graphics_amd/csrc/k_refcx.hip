// k_refcx.hip -- the reference's ion charge exchange (reference mode).
//
// ion_collisions::collision (ion_collisions.cpp:255-308) is not a Monte-Carlo
// test per ion: every eligible ion adds 1 - exp(-u dt 1e-20 n_g sigma) to an
// accumulator of its grid column (probabilityCX[jp], N entries, kept across
// steps); when the accumulator exceeds 1 the ion takes a neutral's velocity
// from the table and the accumulator restarts at 0.  The result depends on
// the order in which ions are visited, so the order is defined here as
// (column, id): the eligible ions are bucketed by (column, id range) -- a
// histogram, one workgroup's scan, a scatter -- each bucket is sorted by id
// in one workgroup, then one wavefront per column runs the accumulator
// sequentially in fp32.
#include "pic_internal.h"

namespace pic {

// sigma [1e-20 m^2] of the reference's charge-exchange choice
// (pic.cpp:363-378; ion_collisions.cpp:157-199)
__device__ __forceinline__ float refcx_sigma(const mcc::MccConst &c, float E) {
    if (!c.sig_cx_energy) return c.sig_cx_const;
    if (c.gas == mcc::GAS_AR) return mcc::ar_cx(E);
    // sigma_H_cx divides the energy by AMU in kilograms (ion_collisions.cpp:188-189)
    return mcc::h_cx_janev(E / 1000.0f / 1.66054e-27f);
}

// an ion's column if it takes part (ion_collisions.cpp:273-277: (int)(x/dx)
// and (int)(y/dy) in [0, N)), else -1
__device__ __forceinline__ int refcx_column(const SpView &sp, uint32_t i, int N) {
    const uint32_t c = sp.cell[i];
    const int jx = cell_jx(c), jy = cell_jy(c);
    if (jx == CELL_HOLE || jx == CELL_DEAD) return -1;
    const float ox = sp.ox[i], oy = sp.oy[i];
    // truncation of x/dx: x in (-dx, N dx) gives jp = 0 .. N-1
    const bool inx = jx >= -1 && jx <= N - 1 && !(jx == -1 && ox == 0.0f);
    const bool iny = jy >= -1 && jy <= N - 1 && !(jy == -1 && oy == 0.0f);
    if (!inx || !iny) return -1;
    return jx < 0 ? 0 : jx;  // (int) truncates toward zero
}

// bucket of (column, id): column-major, then nb equal ranges of [0, idmax)
__device__ __forceinline__ uint32_t refcx_bucket(int col, uint32_t id, uint32_t idmax,
                                                 uint32_t nb) {
    const uint32_t r = (uint32_t)(((unsigned long long)id * nb) / idmax);
    return (uint32_t)col * nb + (r < nb ? r : nb - 1);
}

__global__ void __launch_bounds__(256)
    k_refcx_hist(SpView sp, const uint32_t *__restrict__ nptr, int N,
                 const SpCounters *__restrict__ cnt, uint32_t *__restrict__ hist, uint32_t nb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nptr) return;
    const int col = refcx_column(sp, i, N);
    if (col < 0) return;
    const uint32_t idmax = cnt->next_id ? cnt->next_id : 1u;
    atomicAdd(&hist[refcx_bucket(col, sp.id[i], idmax, nb)], 1u);
}

// exclusive scan of the nt bucket counts into start[0..nt] and the scatter
// cursors (one workgroup); the counts are cleared for the next step
__global__ void __launch_bounds__(1024)
    k_refcx_scan(uint32_t *__restrict__ hist, uint32_t *__restrict__ start,
                 uint32_t *__restrict__ cursor, uint32_t nt) {
    __shared__ uint32_t wsum[16], carry_s;
    uint32_t carry = 0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t r0 = 0; r0 < nt; r0 += blockDim.x) {
        const uint32_t r = r0 + threadIdx.x;
        const uint32_t v = r < nt ? hist[r] : 0u;
        uint32_t x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
            if (k < w) pre += wsum[k];
            tot += wsum[k];
        }
        if (r < nt) {
            start[r] = carry + pre + x - v;
            cursor[r] = carry + pre + x - v;
            hist[r] = 0u;
        }
        if (threadIdx.x == 0) carry_s = carry + tot;
        __syncthreads();
        carry = carry_s;
    }
    if (threadIdx.x == 0) start[nt] = carry;
}

// the eligible ions into bucket order: id, and the collision probability
// 1 - exp(-u dt 1e-20 n_g sigma) (ion_collisions.cpp:279-297) bits << 32 |
// store index
__global__ void __launch_bounds__(256)
    k_refcx_scatter(SpView sp, const uint32_t *__restrict__ nptr, int N, mcc::MccConst mc,
                    const SpCounters *__restrict__ cnt, uint32_t *__restrict__ cursor, uint32_t nb,
                    uint32_t *__restrict__ id_out, unsigned long long *__restrict__ val_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nptr) return;
    const int col = refcx_column(sp, i, N);
    if (col < 0) return;
    const uint32_t id = sp.id[i];
    const uint32_t idmax = cnt->next_id ? cnt->next_id : 1u;
    const float vx = sp.vx[i], vy = sp.vy[i], vz = sp.vz[i];
    const float E = (0.5f / mc.q) * mc.mass * (vx * vx + vy * vy + vz * vz);
    const float u = sqrtf(2.0f * mc.q * E / mc.mass);
    const float P = 1.0f - expf(-(u * mc.dt * 1.0e-20f * mc.ng * refcx_sigma(mc, E)));
    const uint32_t pos = atomicAdd(&cursor[refcx_bucket(col, id, idmax, nb)], 1u);
    id_out[pos] = id;
    val_out[pos] = ((unsigned long long)__float_as_uint(P) << 32) | i;
}

// one workgroup per bucket: its entries in id order (ids are unique); up to
// REFCX_LDS entries by a bitonic network in LDS, larger buckets by counting
// each entry's rank
#define REFCX_LDS 2048
__global__ void __launch_bounds__(256)
    k_refcx_sort(const uint32_t *__restrict__ start, const uint32_t *__restrict__ id_in,
                 const unsigned long long *__restrict__ val_in, uint32_t *__restrict__ id_out,
                 unsigned long long *__restrict__ val_out) {
    const uint32_t lo = start[blockIdx.x], k = start[blockIdx.x + 1] - lo;
    if (k == 0) return;
    if (k > REFCX_LDS) {
        for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) {
            const uint32_t v = id_in[lo + j];
            uint32_t r = 0;
            for (uint32_t t = 0; t < k; t++) r += id_in[lo + t] < v ? 1u : 0u;
            id_out[lo + r] = v;
            val_out[lo + r] = val_in[lo + j];
        }
        return;
    }
    __shared__ uint32_t sk[REFCX_LDS];
    __shared__ unsigned long long sv[REFCX_LDS];
    uint32_t P = 1;
    while (P < k) P <<= 1;
    for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) {
        sk[t] = t < k ? id_in[lo + t] : ~0u;
        sv[t] = t < k ? val_in[lo + t] : 0ull;
    }
    __syncthreads();
    for (uint32_t m = 2; m <= P; m <<= 1) {
        for (uint32_t j = m >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) {
                const uint32_t o = t ^ j;
                if (o > t) {
                    const uint32_t x = sk[t], y = sk[o];
                    if ((x > y) == ((t & m) == 0)) {
                        sk[t] = y;
                        sk[o] = x;
                        const unsigned long long w = sv[t];
                        sv[t] = sv[o];
                        sv[o] = w;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t t = threadIdx.x; t < k; t += blockDim.x) {
        id_out[lo + t] = sk[t];
        val_out[lo + t] = sv[t];
    }
}

// one wavefront per column: the sequential accumulator over the column's
// ions in id order (ion_collisions.cpp:297-303); a firing ion takes the
// velocity of table entry rand() % 9999 (ion_collisions.cpp:209-221), drawn
// here from the ion's Philox stream
__global__ void __launch_bounds__(64)
    k_refcx_columns(const uint32_t *__restrict__ ids, const unsigned long long *__restrict__ val,
                    const uint32_t *__restrict__ start, uint32_t nb, int N,
                    float *__restrict__ pcol, SpView sp, mcc::MccConst mc, uint32_t step,
                    unsigned long long *__restrict__ events) {
    const uint32_t col = blockIdx.x;
    if ((int)col >= N) return;
    const int lane = threadIdx.x;
    // the column's buckets, each in id order, are consecutive
    const uint32_t b = start[col * nb], e = start[(col + 1) * nb];
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
            const uint32_t id = ids[r];
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

// id ranges per column: the most that keep N nb within one workgroup's
// scan of 2^18 counters (about n / (N nb) ions per bucket)
uint32_t refcx_ranges(int N) {
    uint32_t nb = 1;
    while ((uint64_t)N * nb * 2 <= (1u << 18)) nb <<= 1;
    return nb;
}

void launch_refcx(hipStream_t st, const SpView &sp, const uint32_t *n, uint32_t cap, int N,
                  const mcc::MccConst &mc, uint32_t step, RefCxBuffers &B, const SpCounters *cnt,
                  unsigned long long *events) {
    const uint32_t nt = (uint32_t)N * B.nb;
    const dim3 grid((cap + 255) / 256);
    hipLaunchKernelGGL(k_refcx_hist, grid, dim3(256), 0, st, sp, n, N, cnt, B.hist, B.nb);
    hipLaunchKernelGGL(k_refcx_scan, dim3(1), dim3(1024), 0, st, B.hist, B.start, B.cursor, nt);
    hipLaunchKernelGGL(k_refcx_scatter, grid, dim3(256), 0, st, sp, n, N, mc, cnt, B.cursor, B.nb,
                       B.id[0], B.val[0]);
    hipLaunchKernelGGL(k_refcx_sort, dim3(nt), dim3(256), 0, st, B.start, B.id[0], B.val[0],
                       B.id[1], B.val[1]);
    hipLaunchKernelGGL(k_refcx_columns, dim3(N), dim3(64), 0, st, B.id[1], B.val[1], B.start,
                       B.nb, N, B.pcol, sp, mc, step, events);
}

}  // namespace pic
