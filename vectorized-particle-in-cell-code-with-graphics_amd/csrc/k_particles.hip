// k_particles.hip -- particle kernels.
//
// Sort pass of one species (a timestep that also reorders the store):
//   pass 1  k_key_count:    histogram of the sort keys of the particles'
//                           cells at the start of the step (counting-sort
//                           key, aggregated per block in LDS); holes of the
//                           store are not counted, so they are dropped;
//   scan    k_scan_*:       exclusive scan of the histogram -> bin starts;
//   pass 2  k_push_scatter: collisions + gather + push + wall test, the
//                           particle written to its key's slot of the other
//                           buffer (one claim per key per block; a particle
//                           lost in this step leaves a hole there), per-wall
//                           loss counts, and the fixed-point CIC deposit
//                           accumulated in an LDS window around the block's
//                           cells, flushed with one 64-bit atomic per touched
//                           node.
// Between sort passes, k_push_inplace advances a species in its own slots.
// Per-phase kernels for the reference-shaped entry points (gather, move,
// collect_charge, wall pass) reuse the same device functions.
#include "mcc.h"
#include "pic_internal.h"

namespace pic {

__device__ __forceinline__ PState load_state(const SpView &v, uint32_t i, bool need_vz) {
    PState s;
    uint32_t c = v.cell[i];
    s.jx = cell_jx(c);
    s.jy = cell_jy(c);
    s.ox = v.ox[i];
    s.oy = v.oy[i];
    s.vx = v.vx[i];
    s.vy = v.vy[i];
    s.vz = need_vz ? v.vz[i] : 0.0f;
    return s;
}

__device__ __forceinline__ void store_state(const SpView &v, uint32_t i, const PState &s,
                                            uint32_t id) {
    v.cell[i] = cell_pack(s.jx, s.jy);
    v.ox[i] = s.ox;
    v.oy[i] = s.oy;
    v.vx[i] = s.vx;
    v.vy[i] = s.vy;
    v.vz[i] = s.vz;
    v.id[i] = id;
}


// ------------------------------------------------------------ pass 1

// e_collisions.cpp:507-511 / ion_collisions.cpp:273-277: collide only while
// (int)(x/dx) and (int)(y/dy) are in [0, N), i.e. x in (-dx, N*dx)
__device__ __forceinline__ bool ref_mcc_inside(const PState &s, int N) {
    return s.jx >= -1 && s.jx <= N - 1 && !(s.jx == -1 && s.ox == 0.0f) && s.jy >= -1 &&
           s.jy <= N - 1 && !(s.jy == -1 && s.oy == 0.0f);
}

// collision (if enabled and not already settled by block_collide_n) then push;
// returns the particle's status.  `ev` and `ch` report the collision.
__device__ __forceinline__ int advance_one(PState &s, uint32_t id, const float2 *__restrict__ E,
                                           const PushConst &pc, const mcc::MccConst &mc,
                                           uint32_t step, int do_push, int &ev, mcc::Child &ch,
                                           bool mcc_done = false) {
    if (!mcc_done) ev = mcc::EV_NONE;
    if (s.jx == CELL_HOLE) return ST_HOLE;
    if (mc.on && !mcc_done) {
        if (mc.electron) {
            ev = mcc::collide_electron(mc, s.vx, s.vy, s.vz, id, step, ref_mcc_inside(s, pc.N), ch);
            if (EV_KIND(ev) == mcc::EV_IONIZE && mc.mode == PIC_DEV_MODE_REF) {
                s.jx = s.jy = CELL_DEAD;  // e_collisions.cpp:377-381: the parent does not survive
                return ST_DEAD;
            }
        } else {
            ev = mcc::collide_ion(mc, s.vx, s.vy, s.vz, id, step);
        }
    }
    if (do_push) push_one(s, E, pc);
    return classify(s, pc.N, pc.mode);
}


// LDS shared by the block-level collision queue and the deposit window
#define BLOCK_LDS_U64 2048
#define CQ_CAP 256  // collision queue entries per block (11 words each)

// Block-level null-collision step (physical mode) for the P particles of every
// thread.  Every particle makes its candidate test (mcc::candidate); the
// candidates -- about 13% of the electrons in the argon discharge, some 130 of
// a block's 1024 particles -- go into one LDS queue so the collision
// kinematics run once per block on dense lanes.  qs[r] returns the particle's
// queue entry (>= 0), -1 (no candidate) or -2 (queue full: the particle
// collides on its own lane in advance_one).  The results stay in the queue
// until every round has read them (collide_take): the caller leaves the LDS
// alone until a barrier after the last round.  The draws are the per-particle
// Philox streams of collide_electron / collide_ion, so every result is the
// per-lane one bit for bit.  Every thread of the block must call this (uniform
// condition: mc.on in physical mode).
template <int P>
__device__ __forceinline__ void block_collide_n(const bool (&active)[P], const PState (&s)[P],
                                                const uint32_t (&id)[P], const mcc::MccConst &mc,
                                                uint32_t step, int (&qs)[P],
                                                unsigned long long *lds, const float2 *tab) {
    static_assert(11 * CQ_CAP * 4 <= BLOCK_LDS_U64 * 8, "collision queue size");
    __shared__ uint32_t wcnt[4];
    float *q = reinterpret_cast<float *>(lds);  // SoA: v (3), id, event, child (6)
    uint32_t *qid = reinterpret_cast<uint32_t *>(q + 3 * CQ_CAP);
    int *qev = reinterpret_cast<int *>(q + 4 * CQ_CAP);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t pos[P];
    uint32_t wtot = 0;
#pragma unroll
    for (int r = 0; r < P; r++) {
        const bool cand =
            active[r] && mcc::candidate(mc, id[r], step,
                                        s[r].vx * s[r].vx + s[r].vy * s[r].vy + s[r].vz * s[r].vz,
                                        tab);
        const uint64_t bal = __ballot(cand);
        pos[r] = cand ? wtot + (uint32_t)__popcll(bal & below) : 0xFFFFFFFFu;
        wtot += (uint32_t)__popcll(bal);
    }
    if (lane == 0) wcnt[w] = wtot;
    __syncthreads();
    uint32_t base = 0, total = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
        if (k < w) base += wcnt[k];
        total += wcnt[k];
    }
#pragma unroll
    for (int r = 0; r < P; r++) {
        qs[r] = -1;
        if (pos[r] != 0xFFFFFFFFu) {
            const uint32_t slot = base + pos[r];
            if (slot < CQ_CAP) {
                q[slot] = s[r].vx;
                q[CQ_CAP + slot] = s[r].vy;
                q[2 * CQ_CAP + slot] = s[r].vz;
                qid[slot] = id[r];
                qs[r] = (int)slot;
            } else {
                qs[r] = -2;
            }
        }
    }
    __syncthreads();
    const uint32_t nq = min(total, (uint32_t)CQ_CAP);
    for (uint32_t t = threadIdx.x; t < nq; t += blockDim.x) {
        float vx = q[t], vy = q[CQ_CAP + t], vz = q[2 * CQ_CAP + t];
        mcc::Child c{};
        const int e = mc.electron ? mcc::collide_electron(mc, vx, vy, vz, qid[t], step, true, c, tab)
                                  : mcc::collide_ion(mc, vx, vy, vz, qid[t], step);
        q[t] = vx;
        q[CQ_CAP + t] = vy;
        q[2 * CQ_CAP + t] = vz;
        qev[t] = e;
        if (EV_KIND(e) == mcc::EV_IONIZE) {  // the only event with children
            q[5 * CQ_CAP + t] = c.vx;
            q[6 * CQ_CAP + t] = c.vy;
            q[7 * CQ_CAP + t] = c.vz;
            q[8 * CQ_CAP + t] = c.ivx;
            q[9 * CQ_CAP + t] = c.ivy;
            q[10 * CQ_CAP + t] = c.ivz;
        }
    }
    __syncthreads();
}

// A round's collision after block_collide_n: a queued particle takes its
// result.  Returns whether the collision is settled (false: not in physical
// mode, or the queue was full and advance_one collides the particle itself).
__device__ __forceinline__ bool collide_take(bool blk, int qs, PState &s, int &ev, mcc::Child &ch,
                                             const unsigned long long *lds) {
    if (!blk) return false;
    if (qs >= 0) {
        const float *q = reinterpret_cast<const float *>(lds);
        s.vx = q[qs];
        s.vy = q[CQ_CAP + qs];
        s.vz = q[2 * CQ_CAP + qs];
        ev = reinterpret_cast<const int *>(q + 4 * CQ_CAP)[qs];
        if (EV_KIND(ev) == mcc::EV_IONIZE) {
            ch.vx = q[5 * CQ_CAP + qs];
            ch.vy = q[6 * CQ_CAP + qs];
            ch.vz = q[7 * CQ_CAP + qs];
            ch.ivx = q[8 * CQ_CAP + qs];
            ch.ivy = q[9 * CQ_CAP + qs];
            ch.ivz = q[10 * CQ_CAP + qs];
        }
    }
    return qs != -2;
}

// Block-level key aggregation for the sort pass: the keys of the block's
// 256 x IPP particles go into an LDS hash table (open addressing, at most half
// full); each distinct key then costs one global atomic per block, and the
// block's particles of one key take consecutive slots.  Every particle's rank
// within its key comes from the LDS counter of its entry.
#define IPP 4
#define HASH_BITS 11
#define HASH_SIZE (1 << HASH_BITS)  // >= 2 x 256 x IPP
#define HASH_EMPTY 0xFFFFFFFFu
// direct key table of the scatter pass: key ranges up to SCAT_WIN entries,
// SCAT_EPT entries per thread of the 256
#define SCAT_WIN 1024
#define SCAT_EPT (SCAT_WIN / 256)

struct BlockHash {
    uint32_t key[HASH_SIZE];   // pass 2, after the claims: slot of the entry's rank-0 particle
    uint32_t cnt[HASH_SIZE];   // pass 2, after the claims: lowest rank with a valid slot
};

__device__ __forceinline__ void hash_clear(BlockHash &H) {
    for (int e = threadIdx.x; e < HASH_SIZE; e += blockDim.x) {
        H.key[e] = HASH_EMPTY;
        H.cnt[e] = 0u;
    }
}

// Wavefront key grouping: lanes holding the same key form one group (one per
// distinct key in the wavefront; leader = lowest lane).  Sorted input keeps
// the number of groups per wavefront small.
struct WaveGroup {
    int leader;      // lane of this lane's group leader (-1 if inactive)
    uint32_t rank;   // position of this lane within its group
    uint32_t count;  // group size (meaningful on the leader)
};

__device__ __forceinline__ WaveGroup wave_group(uint32_t key, bool active) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    WaveGroup g{-1, 0u, 0u};
    uint64_t pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;  // wave-uniform
        const uint32_t lkey = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
        const uint64_t grp = __ballot(active && key == lkey) & pending;
        if ((grp >> lane) & 1ull) {
            g.leader = leader;
            g.rank = (uint32_t)__popcll(grp & below);
            g.count = (uint32_t)__popcll(grp);
        }
        pending &= ~grp;
    }
    return g;
}

// entry of `key` in the table (inserted if new)
__device__ __forceinline__ uint32_t hash_entry(BlockHash &H, uint32_t key) {
    uint32_t e = (key * 0x9E3779B1u) >> (32 - HASH_BITS);
    for (;;) {  // the table never fills: at most 256 x IPP keys in 2x as many entries
        const uint32_t old = atomicCAS(&H.key[e], HASH_EMPTY, key);
        if (old == HASH_EMPTY || old == key) break;
        e = (e + 1u) & (HASH_SIZE - 1);
    }
    return e;
}

// Called by every lane of the wavefront.  Active lanes get the entry of their
// key and their rank among the block's particles of that key: the lanes of
// one key in a wavefront take consecutive ranks in lane order (one LDS
// atomic per distinct key per wavefront), so their slots are contiguous.
__device__ __forceinline__ uint32_t hash_insert(BlockHash &H, uint32_t key, bool active,
                                                uint32_t &rank) {
    const WaveGroup g = wave_group(key, active);
    uint32_t e = 0u, b = 0u;
    if (active && g.rank == 0) {
        e = hash_entry(H, key);
        b = atomicAdd(&H.cnt[e], g.count);
    }
    const int src = g.leader < 0 ? 0 : g.leader;
    e = (uint32_t)__shfl((int)e, src, 64);
    b = (uint32_t)__shfl((int)b, src, 64);
    rank = b + g.rank;
    return e;
}

// Variants of the particle passes.  The generic kernel takes every option at
// run time; the others pin options that the host has checked (physical mode,
// no magnetic field, no per-particle weights, no pair sampling in the pass;
// then: no collisions in the pass and nothing staged (PLAIN); no collisions
// in the pass -- none, or run ahead in k_collide -- and lost particles
// queuing their ids for re-injection (SPAWN); or the block-queued collisions
// of one species (COLL_E, COLL_I); or, in place, the candidate test only,
// each candidate left untouched for k_collide_q (DEFER)), so the compiler
// drops the code and registers of everything else (the collisionless
// in-place pass: 42 instead of 94 VGPRs, 8 instead of 5 waves per SIMD).  A
// pinned value always equals the one passed in, so every variant computes
// what the generic kernel computes.
enum PassVariant {
    PV_GENERIC = 0,
    PV_PLAIN = 1,
    PV_COLL_E = 2,
    PV_COLL_I = 3,
    PV_SPAWN = 4,
    PV_DEFER = 5
};

template <int V>
__device__ __forceinline__ void pin_variant(PushConst &pc, mcc::MccConst &mc, SpView &sp) {
    if (V == PV_GENERIC) return;
    pc.mode = PIC_DEV_MODE_PHYS;
    pc.bfield = 0;
    sp.w = nullptr;
    if (V == PV_PLAIN || V == PV_SPAWN) {
        mc.on = 0;
    } else if (V == PV_DEFER) {
        mc.on = 1;
        mc.mode = PIC_DEV_MODE_PHYS;
    } else {
        mc.on = 1;
        mc.mode = PIC_DEV_MODE_PHYS;
        mc.electron = (V == PV_COLL_E) ? 1 : 0;
    }
}
// no variant but the generic one samples re-injected pairs in the pass (the
// fused step queues lost ions for k_plop_drain); PLAIN stages nothing at all
template <int V>
__device__ __forceinline__ void pin_spawn(Spawn &spn) {
    if (V != PV_GENERIC) spn.plop = 0;
    if (V == PV_PLAIN || V == PV_SPAWN || V == PV_DEFER) spn.wall = 0;
    if (V == PV_PLAIN) spn.lost_ids = 0;
}

// A deferred pass's candidates of one round: the wavefront appends the
// store indices of its candidates to its block's queue segment (one atomic
// per wavefront).  Called by every lane.
__device__ __forceinline__ void defer_queue(const Spawn &spn, bool cand, uint32_t i) {
    const uint64_t bal = __ballot(cand);
    if (!bal) return;  // uniform over the wavefront
    const int lane = threadIdx.x & 63, lead = __ffsll((unsigned long long)bal) - 1;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t seg = blockIdx.x % MQ_SEGS;
    uint32_t b = 0u;
    if (lane == lead) b = atomicAdd(&spn.mq_cnt[seg * MQ_STRIDE], (uint32_t)__popcll(bal));
    b = (uint32_t)__shfl((int)b, lead, 64);
    const uint32_t t = b + (uint32_t)__popcll(bal & below);
    if (cand && t < spn.mq_seg) spn.mq[(size_t)seg * spn.mq_seg + t] = i;
}

// a particle lost through a wall queues its id and weight for k_plop_drain
__device__ __forceinline__ void queue_lost(const Spawn &spn, uint32_t id, float w) {
    const uint32_t slot = atomicAdd(&spn.cw->wall_count, 1u);
    if (slot >= spn.wq.cap) {
        atomicAdd(&spn.cw->overflow, 1u);
        return;
    }
    spn.wq.key[slot] = id;
    if (spn.wq.w) spn.wq.w[slot] = w;
}

// wavefront reductions of the device library (DPP row shifts and lane reads:
// no LDS round trips); every lane gets the result.  Called with all lanes
// active.
extern "C" __device__ int __ockl_wfred_min_i32(int);
extern "C" __device__ int __ockl_wfred_max_i32(int);
extern "C" __device__ int __ockl_wfred_add_i32(int);
__device__ __forceinline__ int wave_min(int v) { return __ockl_wfred_min_i32(v); }
__device__ __forceinline__ int wave_max(int v) { return __ockl_wfred_max_i32(v); }
__device__ __forceinline__ int wave_sum(int v) { return __ockl_wfred_add_i32(v); }

#define KC_WIN (2 * HASH_SIZE)  // direct-histogram range: the hash table's LDS

// Adds the block's keys (those with in[r]) to binCount, one global atomic
// per distinct key.  The keys of a block of a sorted store lie in a narrow
// range (a few rows of tiles): then the block counts them in a direct LDS
// histogram over that range (one LDS atomic per particle); a range of
// KC_WIN keys or more (an unsorted store) takes the LDS hash table.  `hist` is KC_WIN words of LDS, free on entry; every
// thread of the block calls this.
template <int P>
__device__ __forceinline__ void block_key_hist(const bool (&in)[P], const uint32_t (&key)[P],
                                               uint32_t *__restrict__ binCount, uint32_t *hist) {
    static_assert(sizeof(BlockHash) == KC_WIN * sizeof(uint32_t), "histogram size");
    __shared__ int wmin[4], wmax[4];
    int kmin = 0x7FFFFFFF, kmax = -1;
#pragma unroll
    for (int r = 0; r < P; r++)
        if (in[r]) {
            kmin = min(kmin, (int)key[r]);
            kmax = max(kmax, (int)key[r]);
        }
    kmin = wave_min(kmin);
    kmax = wave_max(kmax);
    if ((threadIdx.x & 63) == 0) {
        wmin[threadIdx.x >> 6] = kmin;
        wmax[threadIdx.x >> 6] = kmax;
    }
    __syncthreads();
    kmin = min(min(wmin[0], wmin[1]), min(wmin[2], wmin[3]));
    kmax = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
    if (kmax < kmin) return;  // no particle counted (uniform)
    if (kmax - kmin < KC_WIN) {
        const int span = kmax - kmin + 1;
        for (int e = threadIdx.x; e < span; e += blockDim.x) hist[e] = 0u;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < P; r++)
            if (in[r]) atomicAdd(&hist[key[r] - (uint32_t)kmin], 1u);
        __syncthreads();
        for (int e = threadIdx.x; e < span; e += blockDim.x) {
            const uint32_t c = hist[e];
            if (c) atomicAdd(&binCount[(size_t)(kmin + e) * BIN_STRIDE], c);
        }
        return;
    }
    BlockHash &H = *reinterpret_cast<BlockHash *>(hist);
    hash_clear(H);
    __syncthreads();  // table cleared
#pragma unroll
    for (int r = 0; r < P; r++) {
        uint32_t rank;
        hash_insert(H, key[r], in[r], rank);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < HASH_SIZE; e += blockDim.x) {
        const uint32_t k = H.key[e];
        if (k != HASH_EMPTY) atomicAdd(&binCount[(size_t)k * BIN_STRIDE], H.cnt[e]);
    }
}

// Pass 1 of the sort pass: the histogram of the sort keys of the particles'
// current cells (block_key_hist).  In a timestep's sort pass (do_push) every
// particle that is not a hole takes a slot, keyed by its cell before the
// step: one that leaves during the step leaves a hole in its slot, which the
// next sort pass drops.  Without a push (compaction) only the particles
// inside the domain are counted, so the output is compact.  Reads 4 B per
// particle (12 B without a push).  An in-place pass ahead of a sort pass
// makes this histogram itself (k_push_inplace, `hist`).
__global__ void __launch_bounds__(256)
    k_key_count(SpView sp, const uint32_t *__restrict__ nptr, PushConst pc,
                uint32_t *__restrict__ binCount, int do_push) {
    const uint32_t n = *nptr;
    if (blockIdx.x * (blockDim.x * IPP) >= n) return;  // past the live count
    __shared__ uint32_t hist[KC_WIN];  // the direct histogram or the hash table
    const uint32_t base = blockIdx.x * (blockDim.x * IPP) + threadIdx.x;
    uint32_t key[IPP];
    bool in[IPP];
#pragma unroll
    for (int r = 0; r < IPP; r++) {
        const uint32_t i = base + r * blockDim.x;
        key[r] = 0u;
        in[r] = false;
        if (i < n) {
            const uint32_t c = sp.cell[i];
            PState s{};
            s.jx = cell_jx(c);
            s.jy = cell_jy(c);
            if (do_push) {
                in[r] = s.jx != CELL_HOLE;
            } else if (s.jx != CELL_HOLE) {
                s.ox = sp.ox[i];
                s.oy = sp.oy[i];
                in[r] = classify(s, pc.N, pc.mode) == ST_KEPT;
            }
            if (in[r]) key[r] = bin_key(s.jx, s.jy, pc.tilesX) >> pc.keyshift;
        }
    }
    block_key_hist<IPP>(in, key, binCount, hist);
}

// ------------------------------------------------------------ scan
// exclusive scan of SCAN_TILE-element tiles: partial sums, top-level scan in
// one block, then per-tile scan plus offset.
#define SCAN_THREADS 256
#define SCAN_PER_THREAD 16
#define SCAN_TILE (SCAN_THREADS * SCAN_PER_THREAD)

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int k = 0; k < SCAN_THREADS / 64; k++) {
        if (k < w) pre += wsum[k];
        tot += wsum[k];
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

// exclusive scan over a 1024-thread block
__device__ __forceinline__ uint32_t block_excl_scan_1024(uint32_t v, uint32_t *total) {
    __shared__ uint32_t wsum[16];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
        if (k < w) pre += wsum[k];
        tot += wsum[k];
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

// the counts are S words apart (BIN_STRIDE for the sort histogram, 1 for
// the compaction's block counts)
template <int S>
__global__ void __launch_bounds__(SCAN_THREADS)
    k_scan_partial(const uint32_t *__restrict__ in, uint32_t *__restrict__ partial) {
    const uint32_t *p =
        in + ((size_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER_THREAD) * S;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) s += p[k * S];
    uint32_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_THREADS)
    k_scan_top(uint32_t *__restrict__ partial, int nparts, uint32_t *__restrict__ start_total,
               uint32_t *__restrict__ n_out) {
    uint32_t carry = 0;
    for (int base = 0; base < nparts; base += SCAN_THREADS) {
        int k = base + threadIdx.x;
        uint32_t v = (k < nparts) ? partial[k] : 0u;
        uint32_t tot;
        uint32_t ex = block_excl_scan(v, &tot);
        if (k < nparts) partial[k] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        *start_total = carry;
        *n_out = carry;
    }
}

template <int S>
__global__ void __launch_bounds__(SCAN_THREADS)
    k_scan_final(const uint32_t *__restrict__ in, const uint32_t *__restrict__ partial,
                 uint32_t *__restrict__ out) {
    size_t base = (size_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER_THREAD;
    uint32_t v[SCAN_PER_THREAD];
    const uint32_t *p = in + base * S;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) v[k] = p[k * S];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) s += v[k];
    uint32_t tot;
    uint32_t run = partial[blockIdx.x] + block_excl_scan(s, &tot);
    uint4 *o = reinterpret_cast<uint4 *>(out + base);
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD / 4; k++) {
        uint4 q;
        q.x = run;
        run += v[4 * k];
        q.y = run;
        run += v[4 * k + 1];
        q.z = run;
        run += v[4 * k + 2];
        q.w = run;
        run += v[4 * k + 3];
        o[k] = q;
    }
}

template <int S>
static void launch_scan_s(hipStream_t st, const uint32_t *count, uint32_t *start,
                          uint32_t nbins_pad, uint32_t *partial, uint32_t *n_out) {
    int nparts = (int)(nbins_pad / SCAN_TILE);
    hipLaunchKernelGGL(k_scan_partial<S>, dim3(nparts), dim3(SCAN_THREADS), 0, st, count, partial);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_THREADS), 0, st, partial, nparts,
                       start + nbins_pad, n_out);
    hipLaunchKernelGGL(k_scan_final<S>, dim3(nparts), dim3(SCAN_THREADS), 0, st, count, partial,
                       start);
}

void launch_scan(hipStream_t st, const uint32_t *count, uint32_t *start, uint32_t nbins_pad,
                 uint32_t *partial, uint32_t *n_out) {
    launch_scan_s<BIN_STRIDE>(st, count, start, nbins_pad, partial, n_out);
}

uint32_t scan_pad(uint32_t nbins) { return (nbins + SCAN_TILE - 1) / SCAN_TILE * SCAN_TILE; }

// ------------------------------------------------------------ pass 2
#define DEP_WCAP BLOCK_LDS_U64  // LDS window capacity (nodes), 16 KB of uint64


// Deposit the block's particles (dep true) through an LDS window spanning the
// bounding box of their cells (block reduction), flushed with one 64-bit
// atomic per touched node.  Sorted particles put most lanes of a
// wavefront on the same four nodes, and LDS atomics to one address serialise,
// so the window is replicated (up to DEP_COPIES copies, as many as fit) and
// lane l adds into copy l mod copies; the flush sums the copies.  Integer
// sums: the result does not depend on the split.  Every thread of the block
// must call this.
#define DEP_COPIES 8
// block_deposit_n for P particles per thread (dep[r] true: deposit s[r]); one
// window and one flush for the whole block.  When the bounding box of the
// block's cells exceeds the window (a few fast particles far from the rest),
// the window is a fixed square around the block's mean cell: the particles
// inside it go through LDS, the others deposit with global atomics.
// The window holds WCAP nodes (DEP_WCAP by default); the fallback square has
// edge 2 half + 2 with half = (floor(sqrt(WCAP)) - 2) / 2 (21 for 2048 nodes:
// 44 x 44 = 1936; 31 for 4096: 64 x 64).
constexpr int dep_half(int wcap) {
    int r = 0;
    while ((r + 1) * (r + 1) <= wcap) r++;
    return (r - 2) / 2;
}
template <int P, int WCAP = DEP_WCAP>
__device__ __forceinline__ void block_deposit_n(const bool (&dep)[P], const PState (&s)[P],
                                                const float (&pw)[P],
                                                unsigned long long *__restrict__ q, int N,
                                                unsigned long long *sq) {
    constexpr int DEP_HALF = dep_half(WCAP);
    static_assert((2 * DEP_HALF + 2) * (2 * DEP_HALF + 2) <= WCAP, "fallback window size");
    __shared__ int red[4][6];
    const int big = 1 << 30;
    int mnx = big, mny = big, mxx = -big, mxy = -big, sx = 0, sy = 0, cnt = 0;
#pragma unroll
    for (int r = 0; r < P; r++)
        if (dep[r]) {
            mnx = min(mnx, s[r].jx);
            mny = min(mny, s[r].jy);
            mxx = max(mxx, s[r].jx);
            mxy = max(mxy, s[r].jy);
            sx += s[r].jx;
            sy += s[r].jy;
            cnt++;
        }
    mnx = wave_min(mnx);
    mny = wave_min(mny);
    mxx = wave_max(mxx);
    mxy = wave_max(mxy);
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    cnt = wave_sum(cnt);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[w][0] = mnx;
        red[w][1] = mny;
        red[w][2] = mxx;
        red[w][3] = mxy;
        red[w][4] = sx;
        red[w][5] = sy;
    }
    __shared__ int redc[4];
    if ((threadIdx.x & 63) == 0) redc[w] = cnt;
    __syncthreads();
    const int nw = blockDim.x >> 6;
    mnx = red[0][0];
    mny = red[0][1];
    mxx = red[0][2];
    mxy = red[0][3];
    sx = red[0][4];
    sy = red[0][5];
    cnt = redc[0];
    for (int k = 1; k < nw; k++) {
        mnx = min(mnx, red[k][0]);
        mny = min(mny, red[k][1]);
        mxx = max(mxx, red[k][2]);
        mxy = max(mxy, red[k][3]);
        sx += red[k][4];
        sy += red[k][5];
        cnt += redc[k];
    }
    __syncthreads();  // red[] is reused by the next call in the same kernel
    if (cnt == 0) return;  // nothing to deposit (uniform: all threads)
    // a depositing particle has jx in [0, N-2]: its nodes jx, jx+1 <= N-1
    int x0 = mnx, y0 = mny, ww = mxx - mnx + 2, wh = mxy - mny + 2;
    if (ww * wh > WCAP) {
        const int cx = sx / cnt, cy = sy / cnt;
        x0 = cx - DEP_HALF;
        y0 = cy - DEP_HALF;
        ww = wh = 2 * DEP_HALF + 2;
    }
    const int area = ww * wh;
    int copies = DEP_COPIES;
    while (copies > 1 && copies * area > WCAP) copies >>= 1;
    for (int k = threadIdx.x; k < copies * area; k += blockDim.x) sq[k] = 0ull;
    __syncthreads();
    const int cb = (int)(threadIdx.x & (copies - 1)) * area;
#pragma unroll
    for (int r = 0; r < P; r++)
        if (dep[r]) {
            const FxW fw = fx_weights_w(s[r].ox, s[r].oy, pw[r]);
            const int dx = s[r].jx - x0, dy = s[r].jy - y0;
            if (dx >= 0 && dx <= ww - 2 && dy >= 0 && dy <= wh - 2) {
                const int b = cb + dy * ww + dx;
                atomicAdd(&sq[b], fw.a4);
                atomicAdd(&sq[b + 1], fw.a3);
                atomicAdd(&sq[b + ww], fw.a2);
                atomicAdd(&sq[b + ww + 1], fw.a1);
            } else {
                dep_global(q, N, s[r].jx, s[r].jy, fw);
            }
        }
    __syncthreads();
    // window nodes outside the grid (fallback window at an edge) stay zero
    for (int k = threadIdx.x; k < area; k += blockDim.x) {
        unsigned long long v = 0ull;
        for (int c = 0; c < copies; c++) v += sq[c * area + k];
        if (v) atomicAdd(&q[(long)(y0 + k / ww) * N + x0 + k % ww], v);
    }
}

__device__ __forceinline__ void block_deposit(bool dep, const PState &s,
                                              unsigned long long *__restrict__ q, int N,
                                              unsigned long long *sq, float pw = 1.0f) {
    const bool d[1] = {dep};
    const PState ss[1] = {s};
    const float ww[1] = {pw};
    block_deposit_n<1>(d, ss, ww, q, N, sq);
}

// staging of a created particle (slot order is arbitrary; the commit pass
// sorts by key)
__device__ __forceinline__ void stage_put(const StageView &sv, uint32_t *__restrict__ count,
                                          uint32_t *__restrict__ overflow, const PState &s,
                                          unsigned long long key, float w = 1.0f) {
    uint32_t slot = atomicAdd(count, 1u);
    if (slot >= sv.cap) {
        atomicAdd(overflow, 1u);
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

__device__ __forceinline__ void count_events(int ev, unsigned long long *__restrict__ events) {
    __shared__ uint32_t se[8];
    if (threadIdx.x < 8) se[threadIdx.x] = 0;
    __syncthreads();
    if (EV_KIND(ev) > 0) atomicAdd(&se[EV_KIND(ev)], 1u);
    if (ev & EV_OVER_BOUND) atomicAdd(&se[6], 1u);
    __syncthreads();
    if (threadIdx.x < 8 && se[threadIdx.x])
        atomicAdd(&events[cnt_slot() * 8 + threadIdx.x], (unsigned long long)se[threadIdx.x]);
}

// Products of one particle's step: the ejected electron and the new ion of a
// physical-mode ionisation, and the electron-ion pair re-injected for an ion
// lost through a wall (sp.plop).  `pre` is the state before the push.
__device__ __forceinline__ void stage_products(int ev, const mcc::Child &ch, const PState &pre,
                                               int st, bool valid, uint32_t pid,
                                               const float2 *__restrict__ E, const PushConst &pc,
                                               const mcc::MccConst &mc, uint32_t step,
                                               const Spawn &sp, uint32_t *__restrict__ lost6,
                                               float w, const PState &post) {
    if (EV_KIND(ev) == mcc::EV_IONIZE && mc.mode == PIC_DEV_MODE_PHYS) {
        // ejected electron: born at the parent's position, pushed this step
        // as the reference pushes particles appended before move_particles
        PState ce = pre;
        ce.vx = ch.vx;
        ce.vy = ch.vy;
        ce.vz = ch.vz;
        push_one(ce, E, pc);
        int cst = classify(ce, pc.N, pc.mode);
        if (cst != ST_KEPT) {
            // left the domain during its first push: it still takes its id
            // (as it does when created before the push), but is not stored
            atomicAdd(&lost6[cnt_slot() * 6 + cst - 1], 1u);
            ce.jx = ce.jy = CELL_DEAD;
        }
        stage_put(sp.se, &sp.ce->stage_count, &sp.ce->overflow, ce, (unsigned long long)pid, w);
        // new ion at the same place; the ion pass of this step moves it
        PState ci = pre;
        ci.vx = ch.ivx;
        ci.vy = ch.ivy;
        ci.vz = ch.ivz;
        // products take the parent's weight (e_collisions.cpp:413,441)
        stage_put(sp.si, &sp.ci->stage_count, &sp.ci->overflow, ci, (unsigned long long)pid, w);
    }
    if (sp.plop && valid && st >= ST_LEFT && st <= ST_TOP) {
        PState pe, pi;
        mcc::plop_pair(mc, pid, step, pc.N, sp.se_sigma, sp.si_sigma, sp.plop_ref, pe, pi);
        unsigned long long k = (1ull << 32) | pid;
        // the re-injected pair replaces the lost ion: it carries its weight
        stage_put(sp.se, &sp.ce->stage_count, &sp.ce->overflow, pe, k, w);
        stage_put(sp.si, &sp.ci->stage_count, &sp.ci->overflow, pi, k, w);
    }
    // a particle that reached a wall waits for the surface pass
    if (sp.wall && valid && st >= ST_LEFT && st <= ST_TOP)
        stage_put(sp.wq, &sp.cw->wall_count, &sp.cw->overflow, post, (unsigned long long)pid, w);
}

// one particle of k_push_scatter (before the claims)
__device__ __forceinline__ void scatter_round(PState &s, uint32_t pid, float w, bool valid,
                                              bool claim, bool &kept, bool &dep,
                                              const float2 *__restrict__ E, const PushConst &pc,
                                              const mcc::MccConst &mc, uint32_t step, int do_push,
                                              bool blk, int qs, uint32_t *se, uint32_t *sl,
                                              uint32_t *shole, const unsigned long long *lds,
                                              const Spawn &sp, uint32_t *__restrict__ lost6) {
    const PState pre = s;
    int ev = mcc::EV_NONE, st = -1;
    mcc::Child ch{};
    const bool done = collide_take(blk, qs, s, ev, ch, lds);
    if (valid) st = advance_one(s, pid, E, pc, mc, step, do_push, ev, ch, done);
    kept = st == ST_KEPT;
    dep = kept && deposit_ok(s, pc.N, pc.mode);
    if (valid && st != ST_KEPT && st != ST_HOLE) atomicAdd(&sl[st - 1], 1u);
    if (claim && !kept) atomicAdd(shole, 1u);
    if (sp.lost_ids && valid && st >= ST_LEFT && st <= ST_TOP) queue_lost(sp, pid, w);
    if (mc.on) {
        if (EV_KIND(ev) > 0) atomicAdd(&se[EV_KIND(ev)], 1u);
        if (ev & EV_OVER_BOUND) atomicAdd(&se[6], 1u);
    }
    stage_products(ev, ch, pre, st, valid, pid, E, pc, mc, step, sp, lost6, w, s);
}

// Pass 2 of the sort pass: the collide + push of IPP particles per thread,
// one claim (atomicSub on the bin count, which runs back to zero) per
// distinct key of the block, the deposit through one LDS window, the products
// staged.  The keys are the cells before the step, so the ranks and the
// claims come first and the claims' latency overlaps the collisions and the
// push.  A block of a sorted store ranks its particles in a direct LDS table
// over its narrow key range (up to SCAT_WIN keys); a block with a wider range
// (an unsorted store, or the particles appended since the last sort) ranks
// them in an LDS hash table and makes its claims before the collisions.  The
// slots a direct-table block claimed (survivors, and holes for the particles
// lost in this step) are staged in LDS in slot order (each key's claimed
// range contiguous) and written out by consecutive threads, so every store
// instruction covers consecutive slots; a wide block writes its particles
// directly.
//
// LDS: one arena serves three phases -- the rank table (8 KB) beside the
// collision queue (11 KB), or the hash table alone; then the deposit window
// (as many nodes as the arena holds); then the write staging in two halves
// of 512 particles.  Its size follows the waves per SIMD each variant's
// registers allow: 25 KB for the collisionless variants (6 blocks per CU),
// 32 KB for the collisional ones (4-5 blocks per CU).  The pass is bound by
// memory latency: with 3 instead of 4 blocks per CU it took 20 % longer.
#ifndef SCAT_WAVES_PLAIN
#define SCAT_WAVES_PLAIN 6
#endif
#ifndef SCAT_ARENA_COLL
#define SCAT_ARENA_COLL 4096
#endif
template <int V>
constexpr int scat_arena_u64() {
    return (V == PV_PLAIN || V == PV_SPAWN) ? 3200 : SCAT_ARENA_COLL;
}
static_assert(SCAT_WIN + (11 * CQ_CAP * 4) / 8 <= 3200, "table + queue");
static_assert(8 * (256 * IPP / 2) * 4 <= 3200 * 8, "staging half");
static_assert(sizeof(BlockHash) <= 3200 * 8, "hash table");
#define PUSH_SCATTER_PARAMS                                                                   \
    SpView src, SpView dst, const uint32_t *__restrict__ nptr, uint32_t cap,                  \
        const float2 *__restrict__ E, PushConst pc, mcc::MccConst mc, uint32_t step,          \
        uint32_t *__restrict__ binCount, const uint32_t *__restrict__ binStart,               \
        unsigned long long *__restrict__ q, int do_push, int deposit, Spawn sp,               \
        unsigned long long *__restrict__ events, uint32_t *__restrict__ lost6,                \
        uint32_t *__restrict__ holes
#define PUSH_SCATTER_ARGS                                                                     \
    src, dst, nptr, cap, E, pc, mc, step, binCount, binStart, q, do_push, deposit, sp, events, \
        lost6, holes
template <int V>
__device__ __forceinline__ void push_scatter_body(PUSH_SCATTER_PARAMS) {
    pin_variant<V>(pc, mc, src);
    pin_spawn<V>(sp);
    if (V != PV_GENERIC) dst.w = nullptr;
    if (blockIdx.x * (blockDim.x * IPP) >= *nptr) return;  // past the live count
    constexpr int ARENA = scat_arena_u64<V>();
    __shared__ unsigned long long arena[ARENA];
    __shared__ uint32_t se[8], sl[6], s_total, shole;
    uint32_t *tcnt = reinterpret_cast<uint32_t *>(arena);  // per key: count, then staging base
    uint32_t *tkey = tcnt + SCAT_WIN;                      // per key: slot of its rank 0
    unsigned long long *cq = arena + SCAT_WIN;             // collision queue behind the table
    if (threadIdx.x < 8) se[threadIdx.x] = 0;
    if (threadIdx.x < 6) sl[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_total = shole = 0;
    __shared__ int wmin[4], wmax[4];
    const uint32_t n = *nptr;
    const uint32_t base = blockIdx.x * (blockDim.x * IPP) + threadIdx.x;
    const bool blk = mc.on && mc.mode == PIC_DEV_MODE_PHYS;
    PState s[IPP];
    uint32_t pid[IPP], ent[IPP], rank[IPP];
    float w[IPP];
    bool claim[IPP], kept[IPP], dep[IPP];
    int kmin = 0x7FFFFFFF, kmax = -1;
#pragma unroll
    for (int r = 0; r < IPP; r++) {  // every load first
        const uint32_t i = base + r * blockDim.x;
        s[r] = PState{};
        pid[r] = 0u;
        w[r] = 1.0f;
        if (i < n) {
            s[r] = load_state(src, i, true);
            pid[r] = src.id[i];
            w[r] = load_w(src, i);
        }
        // the slot k_key_count counted, keyed by the cell before the step
        // (without a push nothing moves: the particles inside keep slots)
        claim[r] = i < n && s[r].jx != CELL_HOLE &&
                   (do_push || classify(s[r], pc.N, pc.mode) == ST_KEPT);
        ent[r] = claim[r] ? bin_key(s[r].jx, s[r].jy, pc.tilesX) >> pc.keyshift : 0u;
        if (claim[r]) {
            kmin = min(kmin, (int)ent[r]);
            kmax = max(kmax, (int)ent[r]);
        }
    }
    kmin = wave_min(kmin);
    kmax = wave_max(kmax);
    if ((threadIdx.x & 63) == 0) {
        wmin[threadIdx.x >> 6] = kmin;
        wmax[threadIdx.x >> 6] = kmax;
    }
    // the electrons' null-collision bounds in LDS (loaded beside the state)
    __shared__ float2 sbins[MCC_BINS];
    if (mc.bins && threadIdx.x < MCC_BINS) sbins[threadIdx.x] = mc.bins[threadIdx.x];
    __syncthreads();  // counters cleared, key range known, bounds staged
    kmin = min(min(wmin[0], wmin[1]), min(wmin[2], wmin[3]));
    kmax = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
    const int span = kmax - kmin + 1;  // <= 0: no claims
    const bool direct = span <= SCAT_WIN;  // uniform over the block
    if (direct)
        for (int e = threadIdx.x; e < span; e += blockDim.x) tcnt[e] = 0u;
    __syncthreads();
    // table block: rank of each particle within its key; wide block: one
    // claim per particle, its slot start(k) + old - 1 (old = 0 on a corrupt
    // histogram gives a slot >= cap, which is skipped)
    uint32_t slot[IPP];
#pragma unroll
    for (int r = 0; r < IPP; r++) {
        rank[r] = 0u;
        slot[r] = 0xFFFFFFFFu;
        if (direct) {
            ent[r] -= (uint32_t)kmin;
            if (claim[r]) rank[r] = atomicAdd(&tcnt[ent[r]], 1u);
        }
    }
    if (!direct) {
        // wide block: ranks in the LDS hash table, one claim per distinct
        // key, and every slot known before the collision queue takes the
        // LDS (the claims' latency is not hidden: such blocks are few once
        // the store has been sorted)
        BlockHash &H = *reinterpret_cast<BlockHash *>(arena);
        hash_clear(H);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < IPP; r++) ent[r] = hash_insert(H, ent[r], claim[r], rank[r]);
        __syncthreads();
        constexpr int EPT = HASH_SIZE / 256;
        uint32_t kk[EPT], cc[EPT], ho[EPT], hs[EPT];
#pragma unroll
        for (int j = 0; j < EPT; j++) {
            kk[j] = H.key[threadIdx.x + j * 256];
            cc[j] = H.cnt[threadIdx.x + j * 256];
        }
#pragma unroll
        for (int j = 0; j < EPT; j++) {
            ho[j] = hs[j] = 0u;
            if (kk[j] != HASH_EMPTY) {
                ho[j] = atomicSub(&binCount[(size_t)kk[j] * BIN_STRIDE], cc[j]);
                hs[j] = binStart[kk[j]];
            }
        }
#pragma unroll
        for (int j = 0; j < EPT; j++)
            if (kk[j] != HASH_EMPTY) H.key[threadIdx.x + j * 256] = hs[j] + ho[j] - cc[j];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < IPP; r++)
            if (claim[r]) slot[r] = H.key[ent[r]] + rank[r];
    }
    __syncthreads();
    // one claim per distinct key of a table block: the bin count runs down
    // from its total, so the block's c particles of key k take the slots
    // start(k) + old - c + rank, rank < c.  The claims are issued here and
    // their results used after the push.
    uint32_t old[SCAT_EPT], st0[SCAT_EPT];
#pragma unroll
    for (int j = 0; j < SCAT_EPT; j++) {
        const int e = threadIdx.x + j * 256;
        old[j] = st0[j] = 0u;
        if (direct && e < span) {
            const uint32_t c = tcnt[e];
            if (c) {
                old[j] = atomicSub(&binCount[(size_t)(kmin + e) * BIN_STRIDE], c);
                st0[j] = binStart[kmin + e];
            }
        }
    }
    int qs[IPP];
    if (blk) {
        bool act[IPP];
#pragma unroll
        for (int r = 0; r < IPP; r++) act[r] = base + r * blockDim.x < n && s[r].jx != CELL_HOLE;
        block_collide_n<IPP>(act, s, pid, mc, step, qs, cq, mc.bins ? sbins : nullptr);
    }
    // the rounds are spelled out so every array index is a compile-time constant
    scatter_round(s[0], pid[0], w[0], base + 0 * blockDim.x < n, claim[0], kept[0],
                  dep[0], E, pc, mc, step, do_push, blk, qs[0], se, sl, &shole, cq, sp, lost6);
    scatter_round(s[1], pid[1], w[1], base + 1 * blockDim.x < n, claim[1], kept[1],
                  dep[1], E, pc, mc, step, do_push, blk, qs[1], se, sl, &shole, cq, sp, lost6);
    scatter_round(s[2], pid[2], w[2], base + 2 * blockDim.x < n, claim[2], kept[2],
                  dep[2], E, pc, mc, step, do_push, blk, qs[2], se, sl, &shole, cq, sp, lost6);
    scatter_round(s[3], pid[3], w[3], base + 3 * blockDim.x < n, claim[3], kept[3],
                  dep[3], E, pc, mc, step, do_push, blk, qs[3], se, sl, &shole, cq, sp, lost6);
    __syncthreads();  // every round has read its collision result
    if (direct) {
#pragma unroll
        for (int j = 0; j < SCAT_EPT; j++) {
            const int e = threadIdx.x + j * 256;
            if (e < span) {
                const uint32_t c = tcnt[e];
                if (c) {
                    tkey[e] = st0[j] + old[j] - c;
                    tcnt[e] = atomicAdd(&s_total, c);
                }
            }
        }
    }
    __syncthreads();
    uint32_t lidx[IPP];
#pragma unroll
    for (int r = 0; r < IPP; r++) {
        lidx[r] = 0u;
        if (direct && claim[r]) {
            slot[r] = tkey[ent[r]] + rank[r];
            lidx[r] = tcnt[ent[r]] + rank[r];
        }
        // weighted stores stay direct
        if (dst.w && claim[r] && slot[r] < cap) dst.w[slot[r]] = w[r];
    }
    __syncthreads();  // the table is read; the deposit window reuses the LDS
    if (deposit) block_deposit_n<IPP, ARENA>(dep, s, w, q, pc.N, arena);
    __syncthreads();
    if (direct) {
        // the claimed slots staged in slot order, half of the block's at a
        // time (a claimed slot of a particle that left holds a hole)
        uint32_t *stg = reinterpret_cast<uint32_t *>(arena);  // 8 words x SH, SoA
        constexpr uint32_t SH = 256 * IPP / 2;
        const uint32_t tot = s_total;
        for (uint32_t lo = 0; lo < tot; lo += SH) {  // uniform: tot is the block's
#pragma unroll
            for (int r = 0; r < IPP; r++) {
                if (claim[r] && lidx[r] >= lo && lidx[r] < lo + SH) {
                    const uint32_t l = lidx[r] - lo;
                    stg[l] = slot[r];
                    stg[SH + l] =
                        kept[r] ? cell_pack(s[r].jx, s[r].jy) : cell_pack(CELL_HOLE, CELL_HOLE);
                    stg[2 * SH + l] = __float_as_uint(s[r].ox);
                    stg[3 * SH + l] = __float_as_uint(s[r].oy);
                    stg[4 * SH + l] = __float_as_uint(s[r].vx);
                    stg[5 * SH + l] = __float_as_uint(s[r].vy);
                    stg[6 * SH + l] = __float_as_uint(s[r].vz);
                    stg[7 * SH + l] = pid[r];
                }
            }
            __syncthreads();
            const uint32_t m = min(SH, tot - lo);
            for (uint32_t l = threadIdx.x; l < m; l += blockDim.x) {
                const uint32_t sl = stg[l];
                if (sl < cap) {
                    dst.cell[sl] = stg[SH + l];
                    dst.ox[sl] = __uint_as_float(stg[2 * SH + l]);
                    dst.oy[sl] = __uint_as_float(stg[3 * SH + l]);
                    dst.vx[sl] = __uint_as_float(stg[4 * SH + l]);
                    dst.vy[sl] = __uint_as_float(stg[5 * SH + l]);
                    dst.vz[sl] = __uint_as_float(stg[6 * SH + l]);
                    dst.id[sl] = stg[7 * SH + l];
                }
            }
            __syncthreads();  // the half is written out before the next fills it
        }
    } else {
#pragma unroll
        for (int r = 0; r < IPP; r++) {
            if (claim[r] && slot[r] < cap) {
                PState o = s[r];
                if (!kept[r]) o.jx = o.jy = CELL_HOLE;
                store_state(dst, slot[r], o, pid[r]);
            }
        }
    }
    if (mc.on) {
        if (threadIdx.x < 8 && se[threadIdx.x])
            atomicAdd(&events[cnt_slot() * 8 + threadIdx.x], (unsigned long long)se[threadIdx.x]);
    }
    if (threadIdx.x < 6 && sl[threadIdx.x])
        atomicAdd(&lost6[cnt_slot() * 6 + threadIdx.x], sl[threadIdx.x]);
    if (threadIdx.x == 0 && shole) atomicAdd(&holes[cnt_slot()], shole);
}

template <int V>
__global__ void __launch_bounds__(256) k_push_scatter(PUSH_SCATTER_PARAMS) {
    push_scatter_body<V>(PUSH_SCATTER_ARGS);
}
// The collisionless variants fit 6 waves per SIMD in 76 VGPRs without spills
// (7 spill, and measured slower); left to itself the compiler gives them 110
// (4 waves).  The collisional ones keep their 117 (4 waves).
template <>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(SCAT_WAVES_PLAIN, SCAT_WAVES_PLAIN)))
    k_push_scatter<PV_PLAIN>(PUSH_SCATTER_PARAMS) {
    push_scatter_body<PV_PLAIN>(PUSH_SCATTER_ARGS);
}
template <>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(SCAT_WAVES_PLAIN, SCAT_WAVES_PLAIN)))
    k_push_scatter<PV_SPAWN>(PUSH_SCATTER_PARAMS) {
    push_scatter_body<PV_SPAWN>(PUSH_SCATTER_ARGS);
}
#ifdef SCAT_WAVES_COLL
template <>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(SCAT_WAVES_COLL, SCAT_WAVES_COLL)))
    k_push_scatter<PV_COLL_E>(PUSH_SCATTER_PARAMS) {
    push_scatter_body<PV_COLL_E>(PUSH_SCATTER_ARGS);
}
#endif


// one particle of k_push_inplace: collide (block-wide when blk), push, store
// in place or leave a hole, count, stage products; dep = deposit it
__device__ __forceinline__ void inplace_round(const SpView &sp, uint32_t i, uint32_t n, PState &s,
                                              uint32_t pid, bool live, bool &dep, bool &kept_out,
                                              const float2 *__restrict__ E, const PushConst &pc,
                                              const mcc::MccConst &mc, uint32_t step,
                                              const Spawn &spn, uint32_t *__restrict__ lost6,
                                              bool need_vz, bool blk, int qs, uint32_t *sl,
                                              uint32_t *shole, uint32_t *se,
                                              const unsigned long long *lds, float w,
                                              bool defer) {
    const PState pre = s;
    int ev = mcc::EV_NONE;
    mcc::Child ch{};
    // deferred: a candidate stays as it is for k_collide_q
    bool cand = false;
    if (defer) {
        // (a deferring species has no energy bins: its bound is global)
        cand = live && mcc::candidate(mc, pid, step, 0.0f);
        defer_queue(spn, cand, i);
    }
    const bool done = defer || collide_take(blk, qs, s, ev, ch, lds);
    int st = ST_HOLE;
    if (live && !cand) st = advance_one(s, pid, E, pc, mc, step, 1, ev, ch, done);
    const bool kept = (st == ST_KEPT);
    kept_out = kept;
    dep = kept && deposit_ok(s, pc.N, pc.mode);
    if (kept) {
        // the slot holds this particle's previous state: a word whose bits
        // did not change (the cell of a particle that stayed in its cell, vz
        // without a collision or rotation) is not written again, and a
        // wavefront with no changed lane issues no store
        const uint32_t c1 = cell_pack(s.jx, s.jy);
        if (c1 != cell_pack(pre.jx, pre.jy)) sp.cell[i] = c1;
        sp.ox[i] = s.ox;
        sp.oy[i] = s.oy;
        sp.vx[i] = s.vx;
        sp.vy[i] = s.vy;
        if (need_vz && __float_as_uint(s.vz) != __float_as_uint(pre.vz)) sp.vz[i] = s.vz;
    } else if (live && !cand) {
        sp.cell[i] = cell_pack(CELL_HOLE, CELL_HOLE);
        atomicAdd(&sl[st - 1], 1u);
        atomicAdd(shole, 1u);
        // the id is read only here: a pass that needs it for nothing else
        // does not load it for every particle
        if (spn.lost_ids && st >= ST_LEFT && st <= ST_TOP) queue_lost(spn, sp.id[i], w);
    }
    if (mc.on) {
        if (EV_KIND(ev) > 0) atomicAdd(&se[EV_KIND(ev)], 1u);
        if (ev & EV_OVER_BOUND) atomicAdd(&se[6], 1u);
    }
    stage_products(ev, ch, pre, st, i < n, pid, E, pc, mc, step, spn, lost6, w, s);
}

// In-place timestep of one species between sort passes: the same collide +
// push as the sort pass, the state written back to its own slot, the
// deposit through one LDS window per block.  A particle that leaves (or that
// a collision removes) becomes a hole; the next sort pass compacts the store.
// Each thread handles IPP particles (strided by the block size, so every
// load and store stays coalesced): one deposit window and one flush of the
// block's counters then serve 256 * IPP particles.  Reads 20-28 B and writes
// 20-24 B per particle, with no histogram or slot claims.
// LDS of an in-place pass (u64 words): the collision queue, then the
// deposit window, then the sort histogram.  The electrons' collisional
// variant runs at 5 waves per SIMD (its registers): 28 KB fit five blocks
// per CU and give its deposit a 58 x 58 fallback window for particles that
// moved up to three steps since their store was sorted; the others, at 8
// waves, keep 16 KB.
template <int V>
constexpr int inplace_lds_u64() {
    return V == PV_COLL_E ? 3584 : BLOCK_LDS_U64;
}

template <int V>
__device__ __forceinline__ void push_inplace_body(
    SpView sp, const uint32_t *__restrict__ nptr, const float2 *__restrict__ E, PushConst pc,
    mcc::MccConst mc, uint32_t step, unsigned long long *__restrict__ q, int deposit, Spawn spn,
    unsigned long long *__restrict__ events, uint32_t *__restrict__ lost6,
    uint32_t *__restrict__ holes, uint32_t *__restrict__ hist, uint32_t bid) {
    pin_variant<V>(pc, mc, sp);
    pin_spawn<V>(spn);
    const uint32_t n = *nptr;
    // the grid covers the capacity: blocks past the live count leave at once
    // (uniformly, before any barrier); bid: this block's index in the species
    if (bid * (blockDim.x * IPP) >= n) return;
    __shared__ uint32_t sl[6];
    __shared__ uint32_t shole;
    __shared__ uint32_t se[8];
    constexpr int LDS_U64 = inplace_lds_u64<V>();
    __shared__ unsigned long long lds[LDS_U64];
    if (threadIdx.x < 6) sl[threadIdx.x] = 0;
    if (threadIdx.x < 8) se[threadIdx.x] = 0;
    if (threadIdx.x == 0) shole = 0;
    const uint32_t base = bid * (blockDim.x * IPP) + threadIdx.x;
    // vz and the id are read only where something uses them: the magnetic
    // rotation, collisions, re-injection keyed by the id, and the wall
    // queue (surface physics needs the full velocity and the id's draws)
    // (a deferred pass moves no candidate: the others keep their vz)
    const bool need_vz = pc.bfield != 0 || (mc.on && V != PV_DEFER) || spn.wall;
    const bool need_id = mc.on || spn.plop || spn.wall;
    const bool blk = mc.on && mc.mode == PIC_DEV_MODE_PHYS && V != PV_DEFER;
    PState s[IPP];
    uint32_t pid[IPP];
    bool live[IPP];
    float w[IPP];
#pragma unroll
    for (int r = 0; r < IPP; r++) {  // every load first
        const uint32_t i = base + r * blockDim.x;
        s[r] = PState{};
        pid[r] = 0u;
        live[r] = false;
        w[r] = 1.0f;
        if (i < n) {
            s[r] = load_state(sp, i, need_vz);
            live[r] = s[r].jx != CELL_HOLE;
            if (live[r] && need_id) pid[r] = sp.id[i];
            if (live[r]) w[r] = load_w(sp, i);
        }
    }
    // the electrons' null-collision bounds in LDS (loaded beside the state)
    __shared__ float2 sbins[MCC_BINS];
    if (blk && mc.bins && threadIdx.x < MCC_BINS) sbins[threadIdx.x] = mc.bins[threadIdx.x];
    __syncthreads();  // counters initialised, bounds staged
    int qs[IPP];
    if (blk) block_collide_n<IPP>(live, s, pid, mc, step, qs, lds, mc.bins ? sbins : nullptr);
    bool dep[IPP], kept[IPP];
    // the rounds are spelled out so every array index is a compile-time constant
    inplace_round(sp, base + 0 * blockDim.x, n, s[0], pid[0], live[0], dep[0], kept[0], E, pc,
                  mc, step,
                  spn, lost6, need_vz, blk, qs[0], sl, &shole, se, lds, w[0],
                  V == PV_DEFER);
    inplace_round(sp, base + 1 * blockDim.x, n, s[1], pid[1], live[1], dep[1], kept[1], E, pc,
                  mc, step,
                  spn, lost6, need_vz, blk, qs[1], sl, &shole, se, lds, w[1],
                  V == PV_DEFER);
    inplace_round(sp, base + 2 * blockDim.x, n, s[2], pid[2], live[2], dep[2], kept[2], E, pc,
                  mc, step,
                  spn, lost6, need_vz, blk, qs[2], sl, &shole, se, lds, w[2],
                  V == PV_DEFER);
    inplace_round(sp, base + 3 * blockDim.x, n, s[3], pid[3], live[3], dep[3], kept[3], E, pc,
                  mc, step,
                  spn, lost6, need_vz, blk, qs[3], sl, &shole, se, lds, w[3],
                  V == PV_DEFER);
    // block_deposit_n passes a barrier before it writes the window, so every
    // round has read its collision result from the shared LDS by then
    if (deposit) block_deposit_n<IPP, LDS_U64>(dep, s, w, q, pc.N, lds);
    __syncthreads();
    // ahead of a sort pass: the histogram of the cells the particles end the
    // step in, the keys that pass claims by (no k_key_count then)
    if (hist) {
        uint32_t key[IPP];
#pragma unroll
        for (int r = 0; r < IPP; r++)
            key[r] = kept[r] ? bin_key(s[r].jx, s[r].jy, pc.tilesX) >> pc.keyshift : 0u;
        block_key_hist<IPP>(kept, key, hist, reinterpret_cast<uint32_t *>(lds));
    }
    if (threadIdx.x < 6 && sl[threadIdx.x])
        atomicAdd(&lost6[cnt_slot() * 6 + threadIdx.x], sl[threadIdx.x]);
    if (threadIdx.x < 8 && se[threadIdx.x])
        atomicAdd(&events[cnt_slot() * 8 + threadIdx.x], (unsigned long long)se[threadIdx.x]);
    if (threadIdx.x == 0 && shole) atomicAdd(&holes[cnt_slot()], shole);
}

#define PUSH_INPLACE_PARAMS                                                                    \
    SpView sp, const uint32_t *__restrict__ nptr, const float2 *__restrict__ E, PushConst pc, \
        mcc::MccConst mc, uint32_t step, unsigned long long *__restrict__ q, int deposit,     \
        Spawn spn, unsigned long long *__restrict__ events, uint32_t *__restrict__ lost6,     \
        uint32_t *__restrict__ holes, uint32_t *__restrict__ hist
template <int V>
__global__ void __launch_bounds__(256) k_push_inplace(PUSH_INPLACE_PARAMS) {
    push_inplace_body<V>(sp, nptr, E, pc, mc, step, q, deposit, spn, events, lost6, holes, hist,
                         blockIdx.x);
}
// The collisionless variants fit 8 waves per SIMD in 42 VGPRs; left to
// itself the scheduler gives SPAWN (PLAIN plus a rare queue write) 84 and 5
// waves, so both ask for 8 (no spills).
template <>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
    k_push_inplace<PV_PLAIN>(PUSH_INPLACE_PARAMS) {
    push_inplace_body<PV_PLAIN>(sp, nptr, E, pc, mc, step, q, deposit, spn, events, lost6, holes,
                                hist, blockIdx.x);
}
template <>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
    k_push_inplace<PV_SPAWN>(PUSH_INPLACE_PARAMS) {
    push_inplace_body<PV_SPAWN>(sp, nptr, E, pc, mc, step, q, deposit, spn, events, lost6, holes,
                                hist, blockIdx.x);
}
template <>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
    k_push_inplace<PV_DEFER>(PUSH_INPLACE_PARAMS) {
    push_inplace_body<PV_DEFER>(sp, nptr, E, pc, mc, step, q, deposit, spn, events, lost6, holes,
                                hist, blockIdx.x);
}

// Both species' collisionless in-place passes in one launch (the first nba
// blocks advance a, the rest b): the kernel of each species pass as it is,
// one launch's fixed cost instead of two on the small grids
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
    k_push_inplace_pair(InplaceArgs a, InplaceArgs b, uint32_t nba,
                        const float2 *__restrict__ E, uint32_t step, int deposit) {
    if (blockIdx.x < nba)
        push_inplace_body<PV_PLAIN>(a.sp, a.n, E, a.pc, a.mc, step, a.q, deposit, a.spn,
                                    a.events, a.lost6, a.holes, a.hist, blockIdx.x);
    else
        push_inplace_body<PV_PLAIN>(b.sp, b.n, E, b.pc, b.mc, step, b.q, deposit, b.spn,
                                    b.events, b.lost6, b.holes, b.hist, blockIdx.x - nba);
}

// Physical-mode null-collision MCC of one species as passes of their own,
// ahead of the particle pass of the same step (collisions precede the push,
// e_collisions.cpp:480-597 and ion_collisions.cpp:255-308 before
// move_particles), for a species with few candidates (the argon ions: about
// 0.5 %).  Keeping the collision code out of the particle pass gives that
// pass the registers (and occupancy) of the collisionless one.  The draws
// are the per-particle streams of collide_electron / collide_ion, so every
// result equals the per-lane one.
//
// Pass 1, k_cand: every particle's candidate test from its id (one
// Philox2x32 word; reads the cell and the id, 8 B).  A block appends its
// candidates' store indices to one of MQ_SEGS queue segments (by block
// index) with one atomic on that segment's counter, each counter on a
// 128-byte line of its own: no counter takes more than 1/MQ_SEGS of the
// blocks (atomics on one address queue behind each other: with 64 counters
// this pass took 50 us for the C2 ions).  A segment holds every particle of
// its blocks, so it cannot overflow.
#define CIPP 8
__global__ void __launch_bounds__(256)
    k_cand(SpView sp, const uint32_t *__restrict__ nptr, mcc::MccConst mc, uint32_t step,
           uint32_t *__restrict__ mq, uint32_t seg_cap, uint32_t *__restrict__ mq_cnt) {
    const uint32_t n = *nptr;
    if (blockIdx.x * (256 * CIPP) >= n) return;  // past the live count
    __shared__ uint32_t wcnt[4], sbase;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t base = blockIdx.x * (256 * CIPP) + threadIdx.x;
    uint32_t pos[CIPP], cl[CIPP], idv[CIPP];
#pragma unroll
    for (int r = 0; r < CIPP; r++) {  // every load first, none waiting on another
        const uint32_t i = base + r * 256;
        cl[r] = i < n ? sp.cell[i] : cell_pack(CELL_HOLE, CELL_HOLE);
        idv[r] = i < n ? sp.id[i] : 0u;
    }
    uint32_t wtot = 0;
#pragma unroll
    for (int r = 0; r < CIPP; r++) {
        const bool cand = cell_jx(cl[r]) != CELL_HOLE && mcc::candidate(mc, idv[r], step, 0.0f);
        const uint64_t bal = __ballot(cand);
        pos[r] = cand ? wtot + (uint32_t)__popcll(bal & below) : 0xFFFFFFFFu;
        wtot += (uint32_t)__popcll(bal);
    }
    if (lane == 0) wcnt[w] = wtot;
    __syncthreads();
    uint32_t off = 0, total = 0;
    for (int k = 0; k < 4; k++) {
        if (k < w) off += wcnt[k];
        total += wcnt[k];
    }
    const uint32_t seg = blockIdx.x % MQ_SEGS;
    if (threadIdx.x == 0) sbase = total ? atomicAdd(&mq_cnt[seg * MQ_STRIDE], total) : 0u;
    __syncthreads();
    uint32_t *q = mq + (size_t)seg * seg_cap;
#pragma unroll
    for (int r = 0; r < CIPP; r++) {
        const uint32_t t = sbase + off + pos[r];
        if (pos[r] != 0xFFFFFFFFu && t < seg_cap) q[t] = base + r * 256;
    }
}

// Pass 2, k_collide_q: the queued candidates collide on dense lanes (a
// stride loop over each segment), their velocities read and written back by
// store index (12 B each way).  An ionisation stages the ejected electron
// (born at the parent's position and pushed this step) and the new ion, as
// the particle passes do.  `defer`: the queue comes from a deferred in-place
// pass, which left its candidates as they were: each is then also pushed,
// stored in place or left as a hole (loss counted, lost id queued for
// re-injection), and deposited with global atomics -- what the pass does
// for the others.
__global__ void __launch_bounds__(256)
    k_collide_q(SpView sp, const float2 *__restrict__ E, PushConst pc, mcc::MccConst mc,
                uint32_t step, Spawn spn, const uint32_t *__restrict__ mq, uint32_t seg_cap,
                uint32_t *__restrict__ mq_cnt, unsigned long long *__restrict__ events,
                uint32_t *__restrict__ lost6, int defer, unsigned long long *__restrict__ q,
                uint32_t *__restrict__ holes, uint32_t *__restrict__ hist) {
    __shared__ uint32_t se[8];
    if (threadIdx.x < 8) se[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t seg = blockIdx.x % MQ_SEGS, chunk = blockIdx.x / MQ_SEGS;
    const uint32_t nch = gridDim.x / MQ_SEGS;
    const uint32_t cnt = min(mq_cnt[seg * MQ_STRIDE], seg_cap);
    const uint32_t *queue = mq + (size_t)seg * seg_cap;
    for (uint32_t t = chunk * 256 + threadIdx.x; t < cnt; t += nch * 256) {
        const uint32_t i = queue[t];
        const uint32_t pid = sp.id[i];
        float vx = sp.vx[i], vy = sp.vy[i], vz = sp.vz[i];
        mcc::Child ch{};
        const int ev = mc.electron ? mcc::collide_electron(mc, vx, vy, vz, pid, step, true, ch)
                                   : mcc::collide_ion(mc, vx, vy, vz, pid, step);
        if (EV_KIND(ev) > 0) atomicAdd(&se[EV_KIND(ev)], 1u);
        if (ev & EV_OVER_BOUND) atomicAdd(&se[6], 1u);
        const PState pre = load_state(sp, i, false);
        const float pw = load_w(sp, i);
        if (defer) {
            PState s = pre;
            s.vx = vx;
            s.vy = vy;
            s.vz = vz;
            push_one(s, E, pc);
            const int st = classify(s, pc.N, pc.mode);
            if (st == ST_KEPT) {
                sp.cell[i] = cell_pack(s.jx, s.jy);
                sp.ox[i] = s.ox;
                sp.oy[i] = s.oy;
                if (deposit_ok(s, pc.N, pc.mode))
                    dep_global(q, pc.N, s.jx, s.jy, fx_weights_w(s.ox, s.oy, pw));
                // the histogram of the next sort pass, which the deferring
                // in-place pass counted for every other particle
                if (hist)
                    atomicAdd(&hist[(size_t)(bin_key(s.jx, s.jy, pc.tilesX) >> pc.keyshift) *
                                    BIN_STRIDE],
                              1u);
            } else {
                sp.cell[i] = cell_pack(CELL_HOLE, CELL_HOLE);
                atomicAdd(&lost6[cnt_slot() * 6 + st - 1], 1u);
                atomicAdd(&holes[cnt_slot()], 1u);
                if (spn.lost_ids && st >= ST_LEFT && st <= ST_TOP) queue_lost(spn, pid, pw);
            }
            vx = s.vx;
            vy = s.vy;
            vz = s.vz;
        }
        sp.vx[i] = vx;
        sp.vy[i] = vy;
        sp.vz[i] = vz;
        if (EV_KIND(ev) == mcc::EV_IONIZE) {
            // ejected electron: born at the parent's position, pushed this
            // step as the reference pushes particles appended before
            // move_particles
            PState ce = pre;
            ce.vx = ch.vx;
            ce.vy = ch.vy;
            ce.vz = ch.vz;
            push_one(ce, E, pc);
            const int cst = classify(ce, pc.N, pc.mode);
            if (cst != ST_KEPT) {
                // left the domain during its first push: it still takes its
                // id, but is not stored
                atomicAdd(&lost6[cnt_slot() * 6 + cst - 1], 1u);
                ce.jx = ce.jy = CELL_DEAD;
            }
            stage_put(spn.se, &spn.ce->stage_count, &spn.ce->overflow, ce, (unsigned long long)pid,
                      pw);
            // new ion at the same place; the ion pass of this step moves it;
            // products take the parent's weight (e_collisions.cpp:413,441)
            PState ci = pre;
            ci.vx = ch.ivx;
            ci.vy = ch.ivy;
            ci.vz = ch.ivz;
            stage_put(spn.si, &spn.ci->stage_count, &spn.ci->overflow, ci, (unsigned long long)pid,
                      pw);
        }
    }
    __syncthreads();
    if (threadIdx.x < 8 && se[threadIdx.x])
        atomicAdd(&events[cnt_slot() * 8 + threadIdx.x], (unsigned long long)se[threadIdx.x]);
    // one block per segment (the launches below): the segment's counter is
    // read; clear it for the next queue
    if (threadIdx.x == 0 && nch == 1) mq_cnt[seg * MQ_STRIDE] = 0u;
}

// Re-injection for the ions lost through a wall in the particle pass just
// run (physical mode, steady walls): the pass queued each such ion (key = id,
// with its weight) in the wall queue, which these walls do not otherwise
// use, and here every queued ion stages the electron-ion pair drawn from its
// own stream (mcc::plop_pair), keyed (1 << 32 | id), with its weight -- what
// the pass itself staged before.  The pair sampler (three Philox blocks, six
// logs and cosines) then stays out of the pass and its registers.  One
// workgroup: a step loses few ions.  The queue is left empty.
__global__ void __launch_bounds__(1024)
    k_plop_drain(StageView wq, SpCounters *__restrict__ cw, Spawn sp, mcc::MccConst mc,
                 uint32_t step, int N) {
    uint32_t n = cw->wall_count;
    if (n > wq.cap) n = wq.cap;
    for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
        const uint32_t pid = (uint32_t)wq.key[e];
        const float w = wq.w ? wq.w[e] : 1.0f;
        PState pe, pi;
        mcc::plop_pair(mc, pid, step, N, sp.se_sigma, sp.si_sigma, sp.plop_ref, pe, pi);
        const unsigned long long k = (1ull << 32) | pid;
        // the re-injected pair replaces the lost ion: it carries its weight
        stage_put(sp.se, &sp.ce->stage_count, &sp.ce->overflow, pe, k, w);
        stage_put(sp.si, &sp.ci->stage_count, &sp.ci->overflow, pi, k, w);
    }
    __syncthreads();  // every thread has read the count
    if (threadIdx.x == 0) cw->wall_count = 0;
}

// ------------------------------------------------------- per-phase path
__global__ void __launch_bounds__(256)
    k_gather(SpView sp, const uint32_t *__restrict__ nptr, const float2 *__restrict__ E,
             PushConst pc, float *__restrict__ exp_, float *__restrict__ eyp) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nptr) return;
    PState s = load_state(sp, i, false);
    float ex = 0.0f, ey = 0.0f;  // pic.cl:70-71
    if (can_move(s, pc)) gather_one(s, E, pc, ex, ey);
    exp_[i] = ex;
    eyp[i] = ey;
}

__global__ void __launch_bounds__(256)
    k_move(SpView sp, const uint32_t *__restrict__ nptr, const float *__restrict__ exp_,
           const float *__restrict__ eyp, PushConst pc) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nptr) return;
    PState s = load_state(sp, i, true);
    if (!can_move(s, pc)) return;
    move_one(s, exp_[i], eyp[i], pc);
    sp.cell[i] = cell_pack(s.jx, s.jy);
    sp.ox[i] = s.ox;
    sp.oy[i] = s.oy;
    sp.vx[i] = s.vx;
    sp.vy[i] = s.vy;
    sp.vz[i] = s.vz;
}

__global__ void __launch_bounds__(256)
    k_deposit(SpView sp, const uint32_t *__restrict__ nptr, int N, int mode,
              unsigned long long *__restrict__ q) {
    uint32_t n = *nptr;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = i < n;
    PState s{};
    float w = 1.0f;
    if (valid) {
        s = load_state(sp, i, false);
        w = load_w(sp, i);
    }
    bool dep = valid && deposit_ok(s, N, mode);
    __shared__ unsigned long long lds[BLOCK_LDS_U64];
    block_deposit(dep, s, q, N, lds, w);
}

// e_collisions::collision / ion_collisions::collision as a phase of its own
// (e_collisions.cpp:480-597, ion_collisions.cpp:255-308): velocities are
// updated in place; created particles are staged at the parent's position.
__global__ void __launch_bounds__(256)
    k_mcc(SpView sp, const uint32_t *__restrict__ nptr, PushConst pc, mcc::MccConst mc,
          uint32_t step, Spawn spn, unsigned long long *__restrict__ events) {
    uint32_t n = *nptr;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int ev = mcc::EV_NONE;
    mcc::Child ch{};
    PState s{};
    uint32_t pid = 0;
    if (i < n) s = load_state(sp, i, true);
    if (i < n && s.jx != CELL_HOLE) {
        pid = sp.id[i];
        if (mc.electron)
            ev = mcc::collide_electron(mc, s.vx, s.vy, s.vz, pid, step, ref_mcc_inside(s, pc.N), ch);
        else
            ev = mcc::collide_ion(mc, s.vx, s.vy, s.vz, pid, step);
        sp.vx[i] = s.vx;
        sp.vy[i] = s.vy;
        sp.vz[i] = s.vz;
        if (EV_KIND(ev) == mcc::EV_IONIZE && mc.mode == PIC_DEV_MODE_REF)
            sp.cell[i] = cell_pack(CELL_DEAD, CELL_DEAD);
    }
    count_events(ev, events);
    if (EV_KIND(ev) == mcc::EV_IONIZE && mc.mode == PIC_DEV_MODE_PHYS) {
        PState ce = s, ci = s;
        ce.vx = ch.vx;
        ce.vy = ch.vy;
        ce.vz = ch.vz;
        ci.vx = ch.ivx;
        ci.vy = ch.ivy;
        ci.vz = ch.ivz;
        const float w = load_w(sp, i);
        stage_put(spn.se, &spn.ce->stage_count, &spn.ce->overflow, ce, (unsigned long long)pid, w);
        stage_put(spn.si, &spn.ci->stage_count, &spn.ci->overflow, ci, (unsigned long long)pid, w);
    }
}

// Wall pass of the absorbing walls (pic_run's pre-loop pass, pic_boundary;
// boundary.cpp:581-677, 765-824 with zero yields): the particles inside are
// kept, in store order, the others counted as losses (and, with spn.plop,
// re-injected as pairs, or queued for the surface pass).  An
// order-preserving compaction in three launches: k_absorb_count counts each
// block's survivors (12 B per particle read), a scan gives every block its
// output offset, k_absorb writes the survivors there in order (the store
// stays sorted if it was).
#define ABS_IPP 4
__device__ __forceinline__ int absorb_class(const SpView &src, uint32_t i, int N, int mode) {
    PState s{};
    const uint32_t c = src.cell[i];
    s.jx = cell_jx(c);
    s.jy = cell_jy(c);
    if (s.jx == CELL_HOLE) return ST_HOLE;
    s.ox = src.ox[i];
    s.oy = src.oy[i];
    return classify(s, N, mode);
}

// the grid covers every count the scan reads (blocks past the live count
// write 0)
__global__ void __launch_bounds__(256)
    k_absorb_count(SpView src, const uint32_t *__restrict__ nptr, int N, int mode,
                   uint32_t *__restrict__ bcount) {
    const uint32_t n = *nptr;
    const uint32_t base = blockIdx.x * (256 * ABS_IPP) + threadIdx.x;
    int c = 0;
#pragma unroll
    for (int r = 0; r < ABS_IPP; r++) {
        const uint32_t i = base + r * 256;
        if (i < n && absorb_class(src, i, N, mode) == ST_KEPT) c++;
    }
    c = wave_sum(c);
    __shared__ int wc[4];
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = (uint32_t)(wc[0] + wc[1] + wc[2] + wc[3]);
}

__global__ void __launch_bounds__(256)
    k_absorb(SpView src, SpView dst, const uint32_t *__restrict__ nptr,
             const uint32_t *__restrict__ bstart, int N, int mode, uint32_t *__restrict__ lost6,
             mcc::MccConst mc, uint32_t step, Spawn spn) {
    const uint32_t n = *nptr;
    if (blockIdx.x * (256 * ABS_IPP) >= n) return;  // past the live count
    __shared__ uint32_t sl[6];
    __shared__ uint32_t cnt[ABS_IPP][4];
    if (threadIdx.x < 6) sl[threadIdx.x] = 0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t base = blockIdx.x * (256 * ABS_IPP) + threadIdx.x;
    PState s[ABS_IPP];
    uint32_t pid[ABS_IPP], rank[ABS_IPP];
    float wt[ABS_IPP];
    int st[ABS_IPP];
#pragma unroll
    for (int r = 0; r < ABS_IPP; r++) {
        const uint32_t i = base + r * 256;
        s[r] = PState{};
        pid[r] = 0u;
        wt[r] = 1.0f;
        st[r] = ST_HOLE;
        if (i < n) {
            s[r] = load_state(src, i, true);
            pid[r] = src.id[i];
            wt[r] = load_w(src, i);
            st[r] = classify(s[r], N, mode);
        }
        const uint64_t bal = __ballot(st[r] == ST_KEPT);
        rank[r] = (uint32_t)__popcll(bal & below);
        if (lane == 0) cnt[r][w] = (uint32_t)__popcll(bal);
    }
    __syncthreads();
    // store order: round r's particles follow round r-1's, waves in order
#pragma unroll
    for (int r = 0; r < ABS_IPP; r++) {
        uint32_t pre = bstart[blockIdx.x];
        for (int k = 0; k < ABS_IPP * 4; k++)
            if (k < r * 4 + w) pre += cnt[k >> 2][k & 3];
        if (st[r] == ST_KEPT) {
            const uint32_t o = pre + rank[r];
            store_state(dst, o, s[r], pid[r]);
            if (dst.w) dst.w[o] = wt[r];
        } else if (st[r] != ST_HOLE) {
            atomicAdd(&sl[st[r] - 1], 1u);
            if (spn.plop && st[r] >= ST_LEFT && st[r] <= ST_TOP) {
                PState pe, pi;
                mcc::plop_pair(mc, pid[r], step, N, spn.se_sigma, spn.si_sigma, spn.plop_ref, pe,
                               pi);
                const unsigned long long k = (1ull << 32) | pid[r];
                stage_put(spn.se, &spn.ce->stage_count, &spn.ce->overflow, pe, k, wt[r]);
                stage_put(spn.si, &spn.ci->stage_count, &spn.ci->overflow, pi, k, wt[r]);
            }
            if (spn.wall && st[r] >= ST_LEFT && st[r] <= ST_TOP)
                stage_put(spn.wq, &spn.cw->wall_count, &spn.cw->overflow, s[r],
                          (unsigned long long)pid[r], wt[r]);
        }
    }
    __syncthreads();
    if (threadIdx.x < 6 && sl[threadIdx.x])
        atomicAdd(&lost6[cnt_slot() * 6 + threadIdx.x], sl[threadIdx.x]);
}

// Sort (key, val) pairs [0, c) ascending by key in one workgroup: bitonic
// network over the next power of two, padded with ~0 keys (the arrays hold
// at least that many entries: the staging capacity is a power of two).
#define SORT_LDS 4096  // pairs sorted in LDS (48 KB); larger batches sort in place
__device__ void block_sort_pairs(unsigned long long *key, uint32_t *val, uint32_t c) {
    uint32_t P = 1;
    while (P < c) P <<= 1;
    if (P <= SORT_LDS) {
        // the same network on an LDS copy: no global round trip per stage
        __shared__ unsigned long long sk[SORT_LDS];
        __shared__ uint32_t sv[SORT_LDS];
        for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) {
            sk[t] = t < c ? key[t] : ~0ull;
            sv[t] = t < c ? val[t] : 0u;
        }
        __syncthreads();
        for (uint32_t k = 2; k <= P; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) {
                    uint32_t o = t ^ j;
                    if (o > t) {
                        unsigned long long a = sk[t], b = sk[o];
                        bool up = (t & k) == 0;
                        if ((a > b) == up) {
                            sk[t] = b;
                            sk[o] = a;
                            uint32_t x = sv[t];
                            sv[t] = sv[o];
                            sv[o] = x;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (uint32_t t = threadIdx.x; t < c; t += blockDim.x) {
            key[t] = sk[t];
            val[t] = sv[t];
        }
        __syncthreads();
        return;
    }
    for (uint32_t t = c + threadIdx.x; t < P; t += blockDim.x) key[t] = ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) {
                uint32_t o = t ^ j;
                if (o > t) {
                    unsigned long long a = key[t], b = key[o];
                    bool up = (t & k) == 0;
                    if ((a > b) == up) {
                        key[t] = b;
                        key[o] = a;
                        uint32_t x = val[t];
                        val[t] = val[o];
                        val[o] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Commit the staged particles of one species: bitonic sort of (key, index)
// in one workgroup, then append in key order to dst with consecutive new
// ids, optionally depositing their charge.  Sorting makes the result
// independent of the order in which the staging atomics fired.
__device__ __forceinline__ void commit_body(StageView sv, SpCounters *__restrict__ scnt, SpView dst,
                                            uint32_t *__restrict__ n_dst, uint32_t cap_dst,
                                            SpCounters *__restrict__ dcnt, int deposit, int N,
                                            int mode, unsigned long long *__restrict__ q,
                                            uint32_t *__restrict__ hist = nullptr,
                                            int tilesX = 0, int keyshift = 0) {
    __shared__ uint32_t s_c, s_base, s_nid;
    if (threadIdx.x == 0) {
        uint32_t c = scnt->stage_count;
        s_c = c < sv.cap ? c : sv.cap;
        s_base = *n_dst;
        s_nid = dcnt->next_id;
    }
    __syncthreads();
    const uint32_t c = s_c;
    if (c == 0) return;
    for (uint32_t t = threadIdx.x; t < c; t += blockDim.x) sv.idx[t] = t;
    __syncthreads();
    block_sort_pairs(sv.key, sv.idx, c);
    // entries staged as CELL_DEAD take an id but are not stored: exclusive
    // scan of the stored flags gives each stored entry its slot
    __shared__ uint32_t chunk_total;
    const uint32_t base = s_base, nid = s_nid;
    const uint32_t dead = cell_pack(CELL_DEAD, CELL_DEAD);
    uint32_t stored_before = 0;
    for (uint32_t r0 = 0; r0 < c; r0 += blockDim.x) {
        uint32_t r = r0 + threadIdx.x;
        bool live = r < c && sv.cell[sv.idx[r]] != dead;
        uint32_t tot;
        uint32_t ex = block_excl_scan_1024(live ? 1u : 0u, &tot);
        if (threadIdx.x == 0) chunk_total = tot;
        __syncthreads();
        if (live) {
            uint32_t k = sv.idx[r];
            uint32_t slot = base + stored_before + ex;
            if (slot < cap_dst) {
                PState p;
                uint32_t cc = sv.cell[k];
                p.jx = cell_jx(cc);
                p.jy = cell_jy(cc);
                p.ox = sv.ox[k];
                p.oy = sv.oy[k];
                p.vx = sv.vx[k];
                p.vy = sv.vy[k];
                p.vz = sv.vz[k];
                // reflected and backscattered particles (sources 4, 5) keep their id
                const unsigned long long kk = sv.key[r];
                const uint32_t src = (uint32_t)(kk >> 32);
                store_state(dst, slot, p, (src == 4u || src == 5u) ? (uint32_t)kk : nid + r);
                const float w = sv.w ? sv.w[k] : 1.0f;
                if (dst.w) dst.w[slot] = w;
                if (deposit && deposit_ok(p, N, mode))
                    dep_global(q, N, p.jx, p.jy, fx_weights_w(p.ox, p.oy, w));
                // the store's histogram for a pending sort pass counts it too
                if (hist)
                    atomicAdd(&hist[(size_t)(bin_key(p.jx, p.jy, tilesX) >> keyshift) * BIN_STRIDE],
                              1u);
            }
        }
        stored_before += chunk_total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        uint32_t room = base < cap_dst ? cap_dst - base : 0u;
        uint32_t written = stored_before < room ? stored_before : room;
        *n_dst = base + written;
        dcnt->next_id = nid + c;
        uint32_t dropped = (stored_before - written) + (scnt->stage_count - c);
        if (dropped) dcnt->overflow += dropped;
        scnt->stage_count = 0;
    }
}

__global__ void __launch_bounds__(1024)
    k_commit(StageView sv, SpCounters *__restrict__ scnt, SpView dst, uint32_t *__restrict__ n_dst,
             uint32_t cap_dst, SpCounters *__restrict__ dcnt, int deposit, int N, int mode,
             unsigned long long *__restrict__ q) {
    commit_body(sv, scnt, dst, n_dst, cap_dst, dcnt, deposit, N, mode, q);
}

// two independent commits (different species: disjoint staging, stores,
// counters and charge), workgroup b takes the b-th
__global__ void __launch_bounds__(1024) k_commit_pair(CommitArgs a, CommitArgs b, int N) {
    const CommitArgs &c = blockIdx.x == 0 ? a : b;
    commit_body(c.sv, c.scnt, c.dst, c.n_dst, c.cap_dst, c.dcnt, c.deposit, N, c.mode, c.q, c.hist,
                c.tilesX, c.keyshift);
}

// ---------------------------------------------- reference steady-state walls
// steady_e_boundary (boundary.cpp:890-1097) pairs, per side, the m-th
// electron outside with the m-th ion outside and re-injects each pair; the
// x sides first (left x < 0, right (int)(x/dx) >= N, boundary.cpp:906-912),
// then the y sides among the particles the x pass left outside.  Order
// within a side = particle id (the reference's vector order depends on its
// removal history, which this store does not keep).
//
// Each axis is one pass over both species with any number of candidates:
// a histogram of (side, id range) buckets, an exclusive scan of the buckets,
// a scatter into bucket order, a sort of each bucket by id (one workgroup per
// bucket) and the pairing, one thread per pair.  A particle the x pass
// re-injects lands inside, so collecting the y candidates from the store
// after the x pass leaves out exactly the particles the x pass paired.

// side of a particle on an axis: 0 low (x < 0 / y < 0), 1 high (cell >= N),
// -1 inside or an empty slot
__device__ __forceinline__ int steady_side(uint32_t c, int N, int axis) {
    const int jx = cell_jx(c);
    if (jx == CELL_HOLE || jx == CELL_DEAD) return -1;
    const int j = axis == 0 ? jx : cell_jy(c);
    return j < 0 ? 0 : (j >= N ? 1 : -1);
}

// bucket of (side, id): side-major, then nb equal ranges of [0, idmax)
__device__ __forceinline__ uint32_t steady_bucket(int side, uint32_t id, uint32_t idmax,
                                                  uint32_t nb) {
    uint32_t r = (uint32_t)(((unsigned long long)id * nb) / idmax);
    return (uint32_t)side * nb + (r < nb ? r : nb - 1);
}

__device__ __forceinline__ uint32_t steady_idmax(const SpCounters *c) {
    const uint32_t m = c->next_id;
    return m ? m : 1u;
}

__global__ void __launch_bounds__(256) k_steady_hist(SteadyArgs a, int N, int axis) {
    const int s = blockIdx.y;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *a.n[s]) return;
    const int side = steady_side(a.sp[s].cell[i], N, axis);
    if (side < 0) return;
    const SteadyBufs &b = a.b[s];
    atomicAdd(&b.hist[steady_bucket(side, a.sp[s].id[i], steady_idmax(a.cnt[s]), b.nb)], 1u);
}

// exclusive scan of a species' 2 nb bucket counts into start[0..2 nb] (and
// the scatter cursors); the counts are cleared for the next pass
__global__ void __launch_bounds__(1024) k_steady_scan(SteadyArgs a) {
    const SteadyBufs &b = a.b[blockIdx.x];
    const uint32_t nt = 2 * b.nb;
    __shared__ uint32_t carry_s;
    uint32_t carry = 0;
    for (uint32_t r0 = 0; r0 < nt; r0 += blockDim.x) {
        const uint32_t r = r0 + threadIdx.x;
        const uint32_t v = r < nt ? b.hist[r] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan_1024(v, &tot);
        if (r < nt) {
            b.start[r] = carry + ex;
            b.cursor[r] = carry + ex;
            b.hist[r] = 0u;
        }
        if (threadIdx.x == 0) carry_s = carry + tot;
        __syncthreads();
        carry = carry_s;
        __syncthreads();
    }
    if (threadIdx.x == 0) b.start[nt] = carry;
}

__global__ void __launch_bounds__(256) k_steady_scatter(SteadyArgs a, int N, int axis) {
    const int s = blockIdx.y;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *a.n[s]) return;
    const int side = steady_side(a.sp[s].cell[i], N, axis);
    if (side < 0) return;
    const SteadyBufs &b = a.b[s];
    const uint32_t id = a.sp[s].id[i];
    const uint32_t pos = atomicAdd(&b.cursor[steady_bucket(side, id, steady_idmax(a.cnt[s]), b.nb)], 1u);
    b.id_in[pos] = id;
    b.idx_in[pos] = i;
}

// one workgroup per bucket: its entries in id order (ids are unique within a
// species); up to STEADY_LDS entries by a bitonic network in LDS, larger
// buckets by counting each entry's rank
#define STEADY_LDS 2048
__global__ void __launch_bounds__(256) k_steady_sort(SteadyArgs a) {
    const SteadyBufs &b = a.b[blockIdx.y];
    const uint32_t lo = b.start[blockIdx.x], k = b.start[blockIdx.x + 1] - lo;
    if (k == 0) return;
    if (k > STEADY_LDS) {
        for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) {
            const uint32_t v = b.id_in[lo + j];
            uint32_t r = 0;
            for (uint32_t t = 0; t < k; t++) r += b.id_in[lo + t] < v ? 1u : 0u;
            b.id_out[lo + r] = v;
            b.idx_out[lo + r] = b.idx_in[lo + j];
        }
        return;
    }
    __shared__ uint32_t sk[STEADY_LDS], sv[STEADY_LDS];
    uint32_t P = 1;
    while (P < k) P <<= 1;
    for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) {
        sk[t] = t < k ? b.id_in[lo + t] : ~0u;
        sv[t] = t < k ? b.idx_in[lo + t] : 0u;
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
                        const uint32_t w = sv[t];
                        sv[t] = sv[o];
                        sv[o] = w;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t t = threadIdx.x; t < k; t += blockDim.x) {
        b.id_out[lo + t] = sk[t];
        b.idx_out[lo + t] = sv[t];
    }
}

// move the pair (e[ie], ion[ii]) to one point with fresh velocities and
// deposit both (boundary.cpp:854-878); the ion's id keys the draws
__device__ __forceinline__ void plop_store(SpView e, SpView ion, uint32_t ie, uint32_t ii,
                                           const mcc::MccConst &mc, uint32_t step, int N,
                                           float se, float si, unsigned long long *qe,
                                           unsigned long long *qi, int mode) {
    PState pe, pi;
    mcc::plop_pair(mc, ion.id[ii], step, N, se, si, 1, pe, pi);
    e.cell[ie] = cell_pack(pe.jx, pe.jy);
    e.ox[ie] = pe.ox;
    e.oy[ie] = pe.oy;
    e.vx[ie] = pe.vx;
    e.vy[ie] = pe.vy;
    e.vz[ie] = pe.vz;
    ion.cell[ii] = cell_pack(pi.jx, pi.jy);
    ion.ox[ii] = pi.ox;
    ion.oy[ii] = pi.oy;
    ion.vx[ii] = pi.vx;
    ion.vy[ii] = pi.vy;
    ion.vz[ii] = pi.vz;
    // each particle keeps its own weight factor
    if (deposit_ok(pe, N, mode))
        dep_global(qe, N, pe.jx, pe.jy, fx_weights_w(pe.ox, pe.oy, load_w(e, ie)));
    if (deposit_ok(pi, N, mode))
        dep_global(qi, N, pi.jx, pi.jy, fx_weights_w(pi.ox, pi.oy, load_w(ion, ii)));
}

// the pairs of both sides of an axis: thread k < m0 pairs rank k of the low
// side, m0 <= k < m0 + m1 rank k - m0 of the high side
__global__ void __launch_bounds__(256)
    k_steady_pair(SteadyArgs a, mcc::MccConst mc, uint32_t step, int N, float se, float si,
                  unsigned long long *__restrict__ qe, unsigned long long *__restrict__ qi,
                  int mode) {
    const SteadyBufs &be = a.b[0], &bi = a.b[1];
    const uint32_t nb = be.nb;
    const uint32_t e0 = be.start[nb] - be.start[0], e1 = be.start[2 * nb] - be.start[nb];
    const uint32_t i0 = bi.start[nb] - bi.start[0], i1 = bi.start[2 * nb] - bi.start[nb];
    const uint32_t m0 = e0 < i0 ? e0 : i0, m1 = e1 < i1 ? e1 : i1;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m0 + m1) return;
    const int side = k < m0 ? 0 : 1;
    const uint32_t r = side == 0 ? k : k - m0;
    const uint32_t ie = be.idx_out[be.start[side * nb] + r];
    const uint32_t ii = bi.idx_out[bi.start[side * nb] + r];
    plop_store(a.sp[0], a.sp[1], ie, ii, mc, step, N, se, si, qe, qi, mode);
}

// ------------------------------------------------------------ loader
// Uniform positions in (0,L)^2 and a 3-V Maxwellian, from Philox4x32-10 keyed
// by the seed with counter (id, species, call, 0).
__global__ void __launch_bounds__(256)
    k_load_random(SpView sp, uint32_t n, uint32_t key0, uint32_t key1, uint32_t species, int N,
                  double sigma) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t r[4], g[4], z[4];
    philox4x32_10(i, species, 0u, 0u, key0, key1, r);
    philox4x32_10(i, species, 1u, 0u, key0, key1, g);
    philox4x32_10(i, species, 2u, 0u, key0, key1, z);
    double cells = (double)(N - 1);
    double tx = u01d(r[0], r[1]) * cells, ty = u01d(r[2], r[3]) * cells;
    int jx = (int)floor(tx), jy = (int)floor(ty);
    float ox = (float)(tx - (double)jx), oy = (float)(ty - (double)jy);
    if (ox >= 1.0f) { ox = 0.0f; jx += 1; }
    if (oy >= 1.0f) { oy = 0.0f; jy += 1; }
    double u1 = u01d(g[0], g[1]), u2 = u01d(g[2], g[3]);
    double rad = sqrt(-2.0 * log(u1));
    double u3 = u01d(z[0], z[1]), u4 = u01d(z[2], z[3]);
    double two_pi = 6.283185307179586;
    sp.cell[i] = cell_pack(jx, jy);
    sp.ox[i] = ox;
    sp.oy[i] = oy;
    sp.vx[i] = (float)(sigma * rad * cos(two_pi * u2));
    sp.vy[i] = (float)(sigma * rad * sin(two_pi * u2));
    sp.vz[i] = (float)(sigma * sqrt(-2.0 * log(u3)) * cos(two_pi * u4));
    sp.id[i] = i;
}

// ------------------------------------------------------------ particle source
// inject_more_particles (particle.cpp:491-510): n particles a step, uniform in
// [dx, L-dx]^2, each velocity component vt cos(2 pi R2) sqrt|ln R1|
// (particle.cpp:452-464).  Draws: Philox counter (k, step, species*16 +
// RS_INJECT + 256 d, 0x4D434331), d = 0 position, 1..3 vx, vy, vz.  In the
// reference the source runs after the electron collisions and before the
// gather (pic.cpp:1071-1082): injected electrons are pushed and wall-tested
// here (no collision this step) and staged at their new position; injected
// ions are staged unmoved and join the ion pass.  Staging key 2 << 32 | k.
__device__ __forceinline__ void inject_pos(double x, double dx, int &j, float &o) {
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

__device__ __forceinline__ float inject_vel(const uint32_t r[4], double vt) {
    const double R1 = u01d(r[0], r[1]), R2 = u01d(r[2], r[3]);
    return (float)(vt * cos(6.283185307179586 * R2) * sqrt(fabs(log(R1))));
}

__global__ void __launch_bounds__(256)
    k_inject(StageView sv, SpCounters *__restrict__ cnt, const float2 *__restrict__ E, PushConst pc,
             uint32_t key0, uint32_t key1, uint32_t step, uint32_t n_inj, float w, double dx,
             double L, double vt, uint32_t species, int push, uint32_t *__restrict__ lost6) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_inj) return;
    const uint32_t s0 = species * 16u + RS_INJECT;
    uint32_t a[4], b[4], c[4], d[4];
    philox4x32_10(k, step, s0, RNG_DOMAIN_MCC, key0, key1, a);
    philox4x32_10(k, step, s0 + 256u, RNG_DOMAIN_MCC, key0, key1, b);
    philox4x32_10(k, step, s0 + 512u, RNG_DOMAIN_MCC, key0, key1, c);
    philox4x32_10(k, step, s0 + 768u, RNG_DOMAIN_MCC, key0, key1, d);
    const double span = L - 2.0 * dx;
    PState s;
    inject_pos(fma(u01d(a[0], a[1]), span, dx), dx, s.jx, s.ox);
    inject_pos(fma(u01d(a[2], a[3]), span, dx), dx, s.jy, s.oy);
    s.vx = inject_vel(b, vt);
    s.vy = inject_vel(c, vt);
    s.vz = inject_vel(d, vt);
    int st = ST_KEPT;
    if (push) {
        push_one(s, E, pc);
        st = classify(s, pc.N, pc.mode);
    }
    if (st == ST_KEPT)
        stage_put(sv, &cnt->stage_count, &cnt->overflow, s, (2ull << 32) | k, w);
    else
        atomicAdd(&lost6[cnt_slot() * 6 + st - 1], 1u);
}

// ------------------------------------------------------------ launchers
static inline unsigned blocks_for(int64_t n, int t) {
    int64_t b = (n + t - 1) / t;
    return (unsigned)(b > 0 ? b : 1);
}

void launch_inject(hipStream_t st, const StageView &sv, SpCounters *cnt, const float2 *E,
                   const PushConst &pc, uint32_t key0, uint32_t key1, uint32_t step, uint32_t n_inj,
                   float w, double dx, double L, double vt, uint32_t species, int push,
                   uint32_t *lost6) {
    if (n_inj == 0) return;
    hipLaunchKernelGGL(k_inject, dim3(blocks_for(n_inj, 256)), dim3(256), 0, st, sv, cnt, E, pc, key0,
                       key1, step, n_inj, w, dx, L, vt, species, push, lost6);
}

// the variant whose pinned options match the pass's (PV_GENERIC otherwise)
static int pass_variant(const SpView &sp, const PushConst &pc, const mcc::MccConst &mc,
                        const Spawn *spn) {
    if (pc.mode != PIC_DEV_MODE_PHYS || pc.bfield != 0 || sp.w != nullptr || (spn && spn->plop))
        return PV_GENERIC;
    if (!mc.on) {
        if (spn && spn->wall) return PV_GENERIC;  // the wall queue takes the whole state
        return (spn && spn->lost_ids) ? PV_SPAWN : PV_PLAIN;
    }
    if (mc.mode != PIC_DEV_MODE_PHYS) return PV_GENERIC;
    if (spn && spn->mq) return spn->wall ? PV_GENERIC : PV_DEFER;
    return mc.electron ? PV_COLL_E : PV_COLL_I;
}

#define LAUNCH_VARIANT(v, kern, grid, ...)                                                  \
    switch (v) {                                                                           \
    case PV_PLAIN: hipLaunchKernelGGL(kern<PV_PLAIN>, grid, dim3(256), 0, __VA_ARGS__); break; \
    case PV_COLL_E: hipLaunchKernelGGL(kern<PV_COLL_E>, grid, dim3(256), 0, __VA_ARGS__); break; \
    case PV_COLL_I: hipLaunchKernelGGL(kern<PV_COLL_I>, grid, dim3(256), 0, __VA_ARGS__); break; \
    case PV_SPAWN: hipLaunchKernelGGL(kern<PV_SPAWN>, grid, dim3(256), 0, __VA_ARGS__); break; \
    case PV_DEFER: hipLaunchKernelGGL(kern<PV_DEFER>, grid, dim3(256), 0, __VA_ARGS__); break; \
    default: hipLaunchKernelGGL(kern<PV_GENERIC>, grid, dim3(256), 0, __VA_ARGS__); break;   \
    }

void launch_key_count(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                      const PushConst &pc, uint32_t *binCount, int do_push) {
    hipLaunchKernelGGL(k_key_count, dim3(blocks_for(cap, 256 * IPP)), dim3(256), 0, st, sp, n, pc,
                       binCount, do_push);
}

void launch_push_scatter(hipStream_t st, const SpView &src, const SpView &dst, const uint32_t *n,
                         int64_t cap, const float2 *E, const PushConst &pc, const mcc::MccConst &mc,
                         uint32_t step, uint32_t *binCount, const uint32_t *binStart,
                         unsigned long long *q, int do_push, int deposit, const Spawn &spn,
                         unsigned long long *events, uint32_t *lost6, uint32_t *holes) {
    const int v = (dst.w == nullptr) ? pass_variant(src, pc, mc, &spn) : PV_GENERIC;
    LAUNCH_VARIANT(v, k_push_scatter, dim3(blocks_for(cap, 256 * IPP)), st, src, dst, n,
                   (uint32_t)cap, E, pc, mc, step, binCount, binStart, q, do_push, deposit, spn,
                   events, lost6, holes);
}

void launch_push_inplace(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                         const float2 *E, const PushConst &pc, const mcc::MccConst &mc,
                         uint32_t step, unsigned long long *q, int deposit, const Spawn &spn,
                         unsigned long long *events, uint32_t *lost6, uint32_t *holes,
                         uint32_t *hist) {
    const int v = pass_variant(sp, pc, mc, &spn);
    LAUNCH_VARIANT(v, k_push_inplace, dim3(blocks_for(cap, 256 * IPP)), st, sp, n, E, pc, mc, step,
                   q, deposit, spn, events, lost6, holes, hist);
}

bool inplace_pair_ok(const InplaceArgs &a, const InplaceArgs &b) {
    return pass_variant(a.sp, a.pc, a.mc, &a.spn) == PV_PLAIN &&
           pass_variant(b.sp, b.pc, b.mc, &b.spn) == PV_PLAIN;
}

void launch_push_inplace_pair(hipStream_t st, const InplaceArgs &a, const InplaceArgs &b,
                              const float2 *E, uint32_t step, int deposit) {
    const uint32_t nba = (uint32_t)blocks_for(a.cap, 256 * IPP);
    const uint32_t nb = nba + (uint32_t)blocks_for(b.cap, 256 * IPP);
    hipLaunchKernelGGL(k_push_inplace_pair, dim3(nb), dim3(256), 0, st, a, b, nba, E, step,
                       deposit);
}

uint32_t collide_seg_cap(int64_t cap) {
    const int64_t blocks = blocks_for(cap, 256 * CIPP);
    return (uint32_t)((blocks + MQ_SEGS - 1) / MQ_SEGS * (256 * CIPP));
}

void launch_collide(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                    const float2 *E, const PushConst &pc, const mcc::MccConst &mc, uint32_t step,
                    const Spawn &spn, unsigned long long *events, uint32_t *lost6, uint32_t *mq,
                    uint32_t *mq_cnt) {
    const uint32_t seg_cap = collide_seg_cap(cap);
    // the counters are zero: set so at allocation, cleared by k_collide_q
    hipLaunchKernelGGL(k_cand, dim3(blocks_for(cap, 256 * CIPP)), dim3(256), 0, st, sp, n, mc, step,
                       mq, seg_cap, mq_cnt);
    hipLaunchKernelGGL(k_collide_q, dim3(MQ_SEGS), dim3(256), 0, st, sp, E, pc, mc, step, spn,
                       mq, seg_cap, mq_cnt, events, lost6, 0, (unsigned long long *)nullptr,
                       (uint32_t *)nullptr, (uint32_t *)nullptr);
}

void launch_collide_deferred(hipStream_t st, const SpView &sp, const float2 *E,
                             const PushConst &pc, const mcc::MccConst &mc, uint32_t step,
                             const Spawn &spn, unsigned long long *events, uint32_t *lost6,
                             unsigned long long *q, uint32_t *holes, uint32_t *hist) {
    hipLaunchKernelGGL(k_collide_q, dim3(MQ_SEGS), dim3(256), 0, st, sp, E, pc, mc, step, spn,
                       spn.mq, spn.mq_seg, spn.mq_cnt, events, lost6, 1, q, holes, hist);
}

void launch_plop_drain(hipStream_t st, const StageView &wq, SpCounters *cw, const Spawn &sp,
                       const mcc::MccConst &mc, uint32_t step, int N) {
    hipLaunchKernelGGL(k_plop_drain, dim3(1), dim3(1024), 0, st, wq, cw, sp, mc, step, N);
}

void launch_mcc(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                const PushConst &pc, const mcc::MccConst &mc, uint32_t step, const Spawn &spn,
                unsigned long long *events) {
    hipLaunchKernelGGL(k_mcc, dim3(blocks_for(cap, 256)), dim3(256), 0, st, sp, n, pc, mc, step,
                       spn, events);
}

void launch_commit_pair(hipStream_t st, const CommitArgs &a, const CommitArgs &b, int N) {
    hipLaunchKernelGGL(k_commit_pair, dim3(2), dim3(1024), 0, st, a, b, N);
}

void launch_commit(hipStream_t st, const StageView &sv, SpCounters *scnt, const SpView &dst,
                   uint32_t *n_dst, int64_t cap_dst, SpCounters *dcnt, int deposit, int N, int mode,
                   unsigned long long *q) {
    hipLaunchKernelGGL(k_commit, dim3(1), dim3(1024), 0, st, sv, scnt, dst, n_dst,
                       (uint32_t)cap_dst, dcnt, deposit, N, mode, q);
}

void launch_gather(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                   const float2 *E, const PushConst &pc, float *exp_, float *eyp) {
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(cap, 256)), dim3(256), 0, st, sp, n, E, pc, exp_,
                       eyp);
}

void launch_move(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap,
                 const float *exp_, const float *eyp, const PushConst &pc) {
    hipLaunchKernelGGL(k_move, dim3(blocks_for(cap, 256)), dim3(256), 0, st, sp, n, exp_, eyp, pc);
}

void launch_deposit(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap, int N,
                    int mode, unsigned long long *q) {
    hipLaunchKernelGGL(k_deposit, dim3(blocks_for(cap, 256)), dim3(256), 0, st, sp, n, N, mode, q);
}

uint32_t absorb_blocks_pad(int64_t cap) {
    return scan_pad((uint32_t)blocks_for(cap, 256 * ABS_IPP));
}

void launch_absorb(hipStream_t st, const SpView &src, const SpView &dst, const uint32_t *n,
                   uint32_t *nout, int64_t cap, int N, int mode, uint32_t *lost6,
                   const mcc::MccConst &mc, uint32_t step, const Spawn &spn, uint32_t *bcount,
                   uint32_t *bstart, uint32_t *partial) {
    const uint32_t nb = absorb_blocks_pad(cap);
    hipLaunchKernelGGL(k_absorb_count, dim3(nb), dim3(256), 0, st, src, n, N, mode, bcount);
    launch_scan_s<1>(st, bcount, bstart, nb, partial, nout);
    hipLaunchKernelGGL(k_absorb, dim3(blocks_for(cap, 256 * ABS_IPP)), dim3(256), 0, st, src, dst,
                       n, (const uint32_t *)bstart, N, mode, lost6, mc, step, spn);
}

void launch_steady(hipStream_t st, const SteadyArgs &a, const mcc::MccConst &mc, uint32_t step,
                   int N, float se, float si, unsigned long long *qe, unsigned long long *qi,
                   int mode) {
    const int64_t cap = std::max(a.cap[0], a.cap[1]);
    const uint32_t pairs = (uint32_t)std::min(a.cap[0], a.cap[1]);
    for (int axis = 0; axis < 2; axis++) {
        hipLaunchKernelGGL(k_steady_hist, dim3(blocks_for(cap, 256), 2), dim3(256), 0, st, a, N,
                           axis);
        hipLaunchKernelGGL(k_steady_scan, dim3(2), dim3(1024), 0, st, a);
        hipLaunchKernelGGL(k_steady_scatter, dim3(blocks_for(cap, 256), 2), dim3(256), 0, st, a, N,
                           axis);
        hipLaunchKernelGGL(k_steady_sort, dim3(2 * a.b[0].nb, 2), dim3(256), 0, st, a);
        hipLaunchKernelGGL(k_steady_pair, dim3(blocks_for(pairs, 256)), dim3(256), 0, st, a, mc,
                           step, N, se, si, qe, qi, mode);
    }
}

void launch_load_random(hipStream_t st, const SpView &sp, uint32_t n, uint64_t seed,
                        uint32_t species, int N, double sigma) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_load_random, dim3(blocks_for(n, 256)), dim3(256), 0, st, sp, n,
                       (uint32_t)seed, (uint32_t)(seed >> 32), species, N, sigma);
}

}  // namespace pic
