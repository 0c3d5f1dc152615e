// k_diag.hip -- batch-mode diagnostics (graphs.cpp:412-1369, pic.cpp:1167-1274).
//
// Over the averaging window the reference sums, every step: rho, phi, Ex, Ey
// (graphs::sum_fields), each species' charge density / particle charge
// (sum_density), per-cell mean velocities along the centre row and column
// (sum_av_velocities, kernel pic.cl:151-185) and velocity histograms at two
// locations (sum_av_dist, pic.cl:188-258); every TIME_DATA_FREQUENCY steps it
// records particle counts, phi at three probe points and the densities at the
// third (pic.cpp:1167-1180).  Here the sums stay on the device: grid sums are
// per-node FP64 additions in step order; the velocity sums of a step are
// 64-bit fixed point (2^-20 m/s) and the histograms integer counts, so every
// diagnostic is deterministic (the reference's OpenCL += on shared bins is a
// race, as in its deposit).
#include "pic_internal.h"

namespace pic {

#define DIAG_VSCALE 1048576.0  // fixed-point velocity units per m/s

// graphs::sum_fields and sum_density for both species, one node per thread
__global__ void k_diag_fields(DiagDev D, int N, const double *__restrict__ rho,
                              const double *__restrict__ phi, const double *__restrict__ ex64,
                              const double *__restrict__ ey64,
                              const unsigned long long *__restrict__ qe,
                              const unsigned long long *__restrict__ qi, double scale_e,
                              double scale_i, double corner, double qc_e, double qc_i) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= N) return;
    const long k = (long)j * N + i;
    D.rho[k] += rho[k];
    D.phi[k] += phi[k];
    D.ex[k] += ex64[k];  // EgridX[j*N + i]: the stride-N layout's first N^2 entries
    D.ey[k] += ey64[k];
    const bool bx = (i == 0 || i == N - 1), by = (j == 0 || j == N - 1);
    const double fac = (bx && by) ? corner : ((bx || by) ? 2.0 : 1.0);
    const double q_e = ((double)(long long)qe[k] * scale_e) * fac;  // particle::get_q
    const double q_i = ((double)(long long)qi[k] * scale_i) * fac;
    D.q[0][k] += q_e / qc_e;
    D.q[1][k] += q_i / qc_i;
}

// sum_av_velocities + sum_av_dist of one species for one step
__global__ void __launch_bounds__(256)
    k_diag_particles(DiagDev D, int s, SpView sp, const uint32_t *__restrict__ nptr, int N,
                     double vt, int no_bins, int divide_bins_by, int spread, int4 loc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nptr) return;
    const uint32_t c = sp.cell[i];
    const int jx = cell_jx(c), jy = cell_jy(c);
    if (jx == CELL_HOLE || jx == CELL_DEAD) return;
    const float ox = sp.ox[i], oy = sp.oy[i];
    // jp = (int)(x/dx), "x > 0" (pic.cl:164,170); truncation toward zero makes
    // jp = 0 for x in (-dx, 0), which x > 0 excludes anyway
    const bool xpos = jx > 0 || (jx == 0 && ox > 0.0f);
    const bool ypos = jy > 0 || (jy == 0 && oy > 0.0f);
    const int mid = N / 2;  // yPosition = no_y_cells/2; xPosition likewise (pic.cl:159-160,196)
    const float vx = sp.vx[i], vy = sp.vy[i];
    if (xpos && jx < N && jy == mid) {
        atomicAdd((unsigned long long *)&D.vsx[s][jx], (unsigned long long)llrint((double)vx * DIAG_VSCALE));
        atomicAdd(&D.cntx[s][jx], 1u);
    }
    if (ypos && jy < N && jx == mid) {
        atomicAdd((unsigned long long *)&D.vsy[s][jy], (unsigned long long)llrint((double)vy * DIAG_VSCALE));
        atomicAdd(&D.cnty[s][jy], 1u);
    }
    const int half = (no_bins - 1) / 2;
    unsigned long long *dist = D.dist[s];  // 4 x no_bins: X1, Y1, X2, Y2
    if (xpos && jx < N && jy == mid) {
        const double qv = ((double)vx / vt) * (double)divide_bins_by;
        const int vb = fabs(qv) < (double)half + 1.0 ? (int)qv : half + 1;  // (int) truncates
        if (abs(vb) <= half) {
            if (jx >= loc.x - spread && jx <= loc.x + spread) atomicAdd(&dist[vb + half], 1ull);
            if (jx >= loc.z - spread && jx <= loc.z + spread)
                atomicAdd(&dist[2 * no_bins + vb + half], 1ull);
        }
    }
    if (ypos && jy < N && jx == mid) {
        const double qv = ((double)vy / vt) * (double)divide_bins_by;
        const int vb = fabs(qv) < (double)half + 1.0 ? (int)qv : half + 1;
        if (abs(vb) <= half) {
            if (jy >= loc.y - spread && jy <= loc.y + spread)
                atomicAdd(&dist[no_bins + vb + half], 1ull);
            if (jy >= loc.w - spread && jy <= loc.w + spread)
                atomicAdd(&dist[3 * no_bins + vb + half], 1ull);
        }
    }
}

// end of a step: per-cell mean velocities added to the window sums
// (graphs.cpp:683-691: only cells holding a particle), step sums cleared
__global__ void k_diag_vfinal(DiagDev D, int N) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (j >= N) return;
    const uint32_t cx = D.cntx[s][j], cy = D.cnty[s][j];
    if (cx > 0) D.vpx[s][j] += ((double)D.vsx[s][j] / DIAG_VSCALE) / (double)cx;
    if (cy > 0) D.vpy[s][j] += ((double)D.vsy[s][j] / DIAG_VSCALE) / (double)cy;
    D.vsx[s][j] = 0;
    D.vsy[s][j] = 0;
    D.cntx[s][j] = 0;
    D.cnty[s][j] = 0;
}

// one timeData row (pic.cpp:1167-1180): step, n_e, n_i, phi at the three
// probes, electron and ion density (get_q) at the third probe
__global__ void k_diag_time(DiagDev D, int row, long long step, const SpCounters *__restrict__ cnt,
                            int cur_e, int cur_i, int N, const double *__restrict__ phi,
                            const unsigned long long *__restrict__ qe,
                            const unsigned long long *__restrict__ qi, double scale_e,
                            double scale_i, double corner, Probe6 probe) {
    if (threadIdx.x != 0) return;
    uint32_t he = 0, hi = 0;
    for (int b = 0; b < CNT_SLOTS; b++) {
        he += cnt[0].holes[b];
        hi += cnt[1].holes[b];
    }
    double *r = D.tdata + (long)row * DIAG_TCOLS;
    r[0] = (double)step;
    r[1] = (double)(cnt[0].n[cur_e] - he);
    r[2] = (double)(cnt[1].n[cur_i] - hi);
    r[3] = phi[(long)probe.y0 * N + probe.x0];
    r[4] = phi[(long)probe.y1 * N + probe.x1];
    r[5] = phi[(long)probe.y2 * N + probe.x2];
    const int i3 = probe.x2, j3 = probe.y2;
    const bool bx = (i3 == 0 || i3 == N - 1), by = (j3 == 0 || j3 == N - 1);
    const double fac = (bx && by) ? corner : ((bx || by) ? 2.0 : 1.0);
    const long k = (long)j3 * N + i3;
    r[6] = ((double)(long long)qe[k] * scale_e) * fac;
    r[7] = ((double)(long long)qi[k] * scale_i) * fac;
}

void launch_diag_step(hipStream_t st, DiagDev &D, bool graphs, bool time_row, int row,
                      long long step, int N, const double *rho, const double *phi,
                      const double *ex64, const double *ey64, const unsigned long long *qe,
                      const unsigned long long *qi, double scale_e, double scale_i, double corner,
                      double qc_e, double qc_i, const SpView &e, const uint32_t *ne, int64_t cap_e,
                      const SpView &ion, const uint32_t *ni, int64_t cap_i, double vt_e,
                      double vt_i, int no_bins, int divide_bins_by, int spread, int4 loc,
                      const SpCounters *cnt, int cur_e, int cur_i, Probe6 probe) {
    if (graphs) {
        dim3 g((N + 127) / 128, N);
        hipLaunchKernelGGL(k_diag_fields, g, dim3(128), 0, st, D, N, rho, phi, ex64, ey64, qe, qi,
                           scale_e, scale_i, corner, qc_e, qc_i);
        hipLaunchKernelGGL(k_diag_particles, dim3((unsigned)((cap_e + 255) / 256)), dim3(256), 0, st,
                           D, 0, e, ne, N, vt_e, no_bins, divide_bins_by, spread, loc);
        hipLaunchKernelGGL(k_diag_particles, dim3((unsigned)((cap_i + 255) / 256)), dim3(256), 0, st,
                           D, 1, ion, ni, N, vt_i, no_bins, divide_bins_by, spread, loc);
        hipLaunchKernelGGL(k_diag_vfinal, dim3((N + 127) / 128, 2), dim3(128), 0, st, D, N);
    }
    if (time_row)
        hipLaunchKernelGGL(k_diag_time, dim3(1), dim3(64), 0, st, D, row, step, cnt, cur_e, cur_i, N,
                           phi, qe, qi, scale_e, scale_i, corner, probe);
}

}  // namespace pic

namespace pic {

// particles whose id is a multiple of `stride` (a subset independent of the
// store's order), compacted into 48-byte records (x, y, vx, vy, vz, weight)
__global__ void __launch_bounds__(256)
    k_subset(SpView sp, const uint32_t *__restrict__ nptr, uint32_t stride, double dx,
             double weight, double *__restrict__ out, uint32_t cap, uint32_t *__restrict__ count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool take = false;
    int jx = 0, jy = 0;
    if (i < *nptr) {
        const uint32_t c = sp.cell[i];
        jx = cell_jx(c);
        jy = cell_jy(c);
        take = jx != CELL_HOLE && jx != CELL_DEAD && (sp.id[i] % stride) == 0u;
    }
    const int lane = threadIdx.x & 63;
    const uint64_t bal = __ballot(take);
    if (!bal) return;
    const int leader = __ffsll((unsigned long long)bal) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(bal));
    base = __shfl(base, leader, 64);
    if (!take) return;
    const uint32_t slot = base + (uint32_t)__popcll(bal & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    if (slot >= cap) return;
    double *r = out + 6 * (size_t)slot;
    r[0] = ((double)jx + (double)sp.ox[i]) * dx;
    r[1] = ((double)jy + (double)sp.oy[i]) * dx;
    r[2] = sp.vx[i];
    r[3] = sp.vy[i];
    r[4] = sp.vz[i];
    r[5] = weight * (double)load_w(sp, i);
}

void launch_subset(hipStream_t st, const SpView &sp, const uint32_t *n, int64_t cap_sp,
                   uint32_t stride, double dx, double weight, double *out, uint32_t cap,
                   uint32_t *count) {
    hipLaunchKernelGGL(k_subset, dim3((unsigned)((cap_sp + 255) / 256)), dim3(256), 0, st, sp, n,
                       stride, dx, weight, out, cap, count);
}

}  // namespace pic
