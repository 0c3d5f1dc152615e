// k_reduce.hip -- population reduction with re-weighting (Kunhardt, Lu & Schoenbach,
// J. Comput. Phys. 67 (1986) 279), particle::contain_particle_growth and
// reduce_population (particle.cpp:520-653).
//
// When a species holds more than size_to_check particles, its particles are
// split into velocity groups by |(vx, vy)| (group g: the first g with
// |v| <= v_limit[g]); every group above its limit loses Ngroup - new_size
// particles chosen uniformly at random and its survivors' weights are
// multiplied by Ngroup / (Ngroup - removed) (particle.cpp:606-629).  The
// reference draws rand() % (size - 1) until every quota is met; here the
// removed particles of group g are the nrem[g] smallest keys
// (Philox(id, step) << 32 | id) of the group -- a uniform subset of exactly
// that size, found by a four-pass radix select over 16-bit digits, all on the
// device, so the step stays free of host round trips.  Removal leaves holes
// that the next sort pass compacts; the charge already deposited this step is
// corrected in fixed point (minus the removed particles, plus the weight
// change of the survivors), which equals depositing the reduced population
// afresh.  Particles faster than the last limit belong to no group and are
// neither removed nor re-weighted (the reference indexes past its arrays).
#include "pic_internal.h"

namespace pic {

#define RD_BINS 65536

__device__ __forceinline__ int rd_group(float vx, float vy, const ReduceConst &rc) {
    const double v = sqrt((double)vx * (double)vx + (double)vy * (double)vy);
    int g = 0;
    while (g < rc.ng && v > rc.vlim[g]) g++;
    return g;
}

__device__ __forceinline__ unsigned long long rd_key(uint32_t id, uint32_t step,
                                                     const ReduceConst &rc) {
    uint32_t r[4];
    philox4x32_10(id, step, rc.species * 16u + RS_REDUCE, RNG_DOMAIN_MCC, rc.key0, rc.key1, r);
    return ((unsigned long long)r[0] << 32) | id;
}

__device__ __forceinline__ bool rd_live(uint32_t c) {
    const int jx = cell_jx(c);
    return jx != CELL_HOLE && jx != CELL_DEAD;
}

// Ngroup of every group (particle.cpp:543-556); nothing when the species is
// not above size_to_check
__global__ void __launch_bounds__(256)
    k_reduce_count(SpView sp, const uint32_t *__restrict__ nptr, ReduceConst rc,
                   ReduceState *__restrict__ st) {
    const uint32_t n = *nptr;
    if (n <= rc.size_check) return;
    __shared__ uint32_t c[RD_MAX_GROUPS];
    if (threadIdx.x < RD_MAX_GROUPS) c[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!rd_live(sp.cell[i])) continue;
        const int g = rd_group(sp.vx[i], sp.vy[i], rc);
        if (g < rc.ng) atomicAdd(&c[g], 1u);
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)rc.ng && c[threadIdx.x]) atomicAdd(&st->count[threadIdx.x], c[threadIdx.x]);
}

// quotas (particle.cpp:561-569) and group weights (particle.cpp:606-613);
// clears the counts for the next step.  The size test counts the holes an
// in-place pass may have left: the reference's vector holds no holes, so
// the test uses the live count.
__global__ void k_reduce_plan(const uint32_t *__restrict__ nptr, const uint32_t *__restrict__ holes,
                              ReduceConst rc, ReduceState *__restrict__ st,
                              SpCounters *__restrict__ cnt) {
    if (threadIdx.x != 0) return;
    uint32_t h = 0;
    for (int b = 0; b < CNT_SLOTS; b++) h += holes[b];
    const uint32_t n = *nptr;
    const bool big = n > rc.size_check && n - h > rc.size_check;
    uint32_t active = 0;
    for (int g = 0; g < RD_MAX_GROUPS; g++) {
        uint32_t N = g < rc.ng ? st->count[g] : 0u;
        uint32_t r = 0;
        if (big && N > rc.glimit[g]) r = N > rc.gnew[g] ? N - rc.gnew[g] : 0u;
        st->nrem[g] = r;
        st->rank[g] = r;
        st->prefix[g] = 0ull;
        st->gw[g] = (r > 0 && N > r) ? (double)N / (double)(N - r) : 1.0;
        st->count[g] = 0;
        active |= r;
    }
    st->active = active ? 1u : 0u;
    if (active) cnt->reductions += 1u;
}

// histogram of digit `pass` (bits 48-16 pass .. 63-16 pass) of the keys whose
// higher digits equal the prefix selected so far
__global__ void __launch_bounds__(256)
    k_reduce_hist(SpView sp, const uint32_t *__restrict__ nptr, ReduceConst rc,
                  ReduceState *__restrict__ st, uint32_t step, int pass) {
    if (!st->active) return;
    const uint32_t n = *nptr;
    const int shift = 48 - 16 * pass;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!rd_live(sp.cell[i])) continue;
        const int g = rd_group(sp.vx[i], sp.vy[i], rc);
        if (g >= rc.ng || st->nrem[g] == 0) continue;
        const unsigned long long key = rd_key(sp.id[i], step, rc);
        if (pass > 0 && (key >> (shift + 16)) != (st->prefix[g] >> (shift + 16))) continue;
        atomicAdd(&st->hist[(size_t)g * RD_BINS + ((key >> shift) & 0xFFFFull)], 1u);
    }
}

// one workgroup per group: the digit holding the rank-th smallest key; the
// histogram is cleared for the next pass
__global__ void __launch_bounds__(1024)
    k_reduce_select(ReduceState *__restrict__ st, int pass) {
    const int g = blockIdx.x;
    if (!st->active || st->nrem[g] == 0) return;
    uint32_t *H = st->hist + (size_t)g * RD_BINS;
    const int per = RD_BINS / 1024;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t s_digit, s_rank;
    if (t == 0) {  // overwritten by the one thread whose bins hold the rank
        s_digit = 0;
        s_rank = 1;
    }
    uint32_t sum = 0;
    for (int b = 0; b < per; b++) sum += H[t * per + b];
    // block exclusive scan of the per-thread sums
    uint32_t x = sum;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (int k = 0; k < w; k++) pre += wsum[k];
    const uint32_t ex = pre + x - sum;
    const uint32_t target = st->rank[g];
    if (ex < target && target <= ex + sum) {
        uint32_t cum = ex;
        for (int b = 0; b < per; b++) {
            const uint32_t hb = H[t * per + b];
            if (cum + hb >= target) {
                s_digit = (uint32_t)(t * per + b);
                s_rank = target - cum;
                break;
            }
            cum += hb;
        }
    }
    __syncthreads();
    for (int b = 0; b < per; b++) H[t * per + b] = 0u;
    if (t == 0) {
        st->prefix[g] |= (unsigned long long)s_digit << (48 - 16 * pass);
        st->rank[g] = s_rank;
    }
}

__device__ __forceinline__ FxW fx_neg(const FxW &a) {
    return FxW{0ull - a.a4, 0ull - a.a3, 0ull - a.a2, 0ull - a.a1};
}

// remove the selected particles, re-weight the survivors of reduced groups
// (particle.cpp:615-629) and correct the deposited charge (q null: leave it)
__global__ void __launch_bounds__(256)
    k_reduce_apply(SpView sp, const uint32_t *__restrict__ nptr, ReduceConst rc,
                   const ReduceState *__restrict__ st, uint32_t step, unsigned long long *__restrict__ q,
                   int N, int mode, uint32_t *__restrict__ lost6, uint32_t *__restrict__ holes) {
    if (!st->active) return;
    __shared__ uint32_t s_rm;
    if (threadIdx.x == 0) s_rm = 0;
    __syncthreads();
    const uint32_t n = *nptr;
    uint32_t rm = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t c = sp.cell[i];
        if (!rd_live(c)) continue;
        const int g = rd_group(sp.vx[i], sp.vy[i], rc);
        if (g >= rc.ng || st->nrem[g] == 0) continue;
        PState s;
        s.jx = cell_jx(c);
        s.jy = cell_jy(c);
        s.ox = sp.ox[i];
        s.oy = sp.oy[i];
        const float w0 = load_w(sp, i);
        const bool dep = q != nullptr && deposit_ok(s, N, mode);
        const FxW old = fx_weights_w(s.ox, s.oy, w0);
        if (rd_key(sp.id[i], step, rc) <= st->prefix[g]) {
            sp.cell[i] = cell_pack(CELL_HOLE, CELL_HOLE);
            if (dep) dep_global(q, N, s.jx, s.jy, fx_neg(old));
            rm++;
        } else {
            const float w1 = (float)((double)w0 * st->gw[g]);
            sp.w[i] = w1;
            if (dep) {
                const FxW nw = fx_weights_w(s.ox, s.oy, w1);
                dep_global(q, N, s.jx, s.jy,
                           FxW{nw.a4 - old.a4, nw.a3 - old.a3, nw.a2 - old.a2, nw.a1 - old.a1});
            }
        }
    }
    if (rm) atomicAdd(&s_rm, rm);
    __syncthreads();
    if (threadIdx.x == 0 && s_rm) {
        atomicAdd(&lost6[cnt_slot() * 6 + 5], s_rm);
        atomicAdd(&holes[cnt_slot()], s_rm);
    }
}

// the whole reduction of one species, queued on the stream (no host wait);
// a no-op unless the species is above size_to_check and a group above its limit
void launch_reduce(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                   const ReduceConst &rc, ReduceState *state, SpCounters *cnt, uint32_t step,
                   unsigned long long *q, int N, int mode) {
    int64_t b = (cap + 255) / 256;
    const unsigned grid = (unsigned)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
    hipLaunchKernelGGL(k_reduce_count, dim3(grid), dim3(256), 0, st, sp, n, rc, state);
    hipLaunchKernelGGL(k_reduce_plan, dim3(1), dim3(64), 0, st, n, &cnt->holes[0], rc, state, cnt);
    for (int pass = 0; pass < 4; pass++) {
        hipLaunchKernelGGL(k_reduce_hist, dim3(grid), dim3(256), 0, st, sp, n, rc, state, step, pass);
        hipLaunchKernelGGL(k_reduce_select, dim3(rc.ng), dim3(1024), 0, st, state, pass);
    }
    hipLaunchKernelGGL(k_reduce_apply, dim3(grid), dim3(256), 0, st, sp, n, rc, state, step, q, N,
                       mode, &cnt->lost[0][0], &cnt->holes[0]);
}

}  // namespace pic
