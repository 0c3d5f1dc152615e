// k_fields.hip -- charge -> rhs, multigrid Poisson solve, E = -grad(phi).
//
// Multigrid (mglin.cpp:21-196): full multigrid with ncycle V(1,1) cycles per
// level, red-black Gauss-Seidel, half-weighting restriction, bilinear
// prolongation.  Every per-node expression keeps the reference's FP64
// association, so FIXED mode reproduces Mglin bit for bit.  Levels up to
// 65x65 (or what the LDS holds) run inside one workgroup with all arrays in
// LDS; finer levels use one launch per operation.
#include <chrono>
#include <cmath>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "pic_internal.h"

namespace pic {

// ---------------------------------------------------------------- rho
// particle.cpp:830-845 (edges x2, corners x8) applied to each species'
// fixed-point charge, then fields.cpp:100-143 (rho = ion + electron).
__global__ void k_rho(const unsigned long long *__restrict__ qi,
                      const unsigned long long *__restrict__ qe, int N, double scale_i,
                      double scale_e, double corner, double *__restrict__ rho) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (i >= N) return;
    bool ex = (i == 0 || i == N - 1), ey = (j == 0 || j == N - 1);
    double fac = (ex && ey) ? corner : ((ex || ey) ? 2.0 : 1.0);
    long k = (long)j * N + i;
    double ri = ((double)(long long)qi[k] * scale_i) * fac;
    double re = ((double)(long long)qe[k] * scale_e) * fac;
    rho[k] = ri + re;
}

// fields.cpp:184-204: rhs = -rho/EPSILON, then the Dirichlet values folded
// into the rows/columns next to the walls (x folds first, then y folds).
__global__ void k_fold(const double *__restrict__ rho, int N, double coef, double phi0X,
                       double phiR, double phi0Y, double phiLY, double dxdy_inv,
                       double *__restrict__ rhs) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (i >= N) return;
    long k = (long)j * N + i;
    double f = -coef * rho[k];
    if (i == 1) f = f - phi0X * dxdy_inv;
    if (i == N - 2) f = f - phiR * dxdy_inv;
    if (j == 1) f = f - phi0Y * dxdy_inv;
    if (j == N - 2) f = f - phiLY * dxdy_inv;
    rhs[k] = f;
}

void launch_rho(hipStream_t st, const unsigned long long *qi, const unsigned long long *qe, int N,
                double scale_i, double scale_e, double corner, double *rho) {
    dim3 g((N + 127) / 128, N);
    hipLaunchKernelGGL(k_rho, g, dim3(128), 0, st, qi, qe, N, scale_i, scale_e, corner, rho);
}

__global__ void k_field_start(const unsigned long long *__restrict__ qi,
                              const unsigned long long *__restrict__ qe, int N, double scale_i,
                              double scale_e, double corner, double *__restrict__ rho, double coef,
                              double phi0X, double phiR, double phi0Y, double phiLY,
                              double dxdy_inv, double *__restrict__ rhs, double *__restrict__ u,
                              double *__restrict__ uprev, int extrap, int *__restrict__ flags,
                              unsigned long long *__restrict__ zi,
                              unsigned long long *__restrict__ ze) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (flags && i == 0 && j == 0) flags[0] = flags[1] = flags[2] = flags[3] = 0;
    if (i >= N) return;
    // the charge buffers of the next step start from zero
    zi[(long)j * N + i] = 0ull;
    ze[(long)j * N + i] = 0ull;
    // k_rho
    bool ex = (i == 0 || i == N - 1), ey = (j == 0 || j == N - 1);
    double fac = (ex && ey) ? corner : ((ex || ey) ? 2.0 : 1.0);
    long k = (long)j * N + i;
    double ri = ((double)(long long)qi[k] * scale_i) * fac;
    double re = ((double)(long long)qe[k] * scale_e) * fac;
    const double r = ri + re;
    rho[k] = r;
    // k_fold
    double f = -coef * r;
    if (i == 1) f = f - phi0X * dxdy_inv;
    if (i == N - 2) f = f - phiR * dxdy_inv;
    if (j == 1) f = f - phi0Y * dxdy_inv;
    if (j == N - 2) f = f - phiLY * dxdy_inv;
    rhs[k] = f;
    // the solve's start: the linear extrapolation of the last two solutions
    if (extrap) {
        const double t = u[k];
        if (extrap == 2) u[k] = 2.0 * t - uprev[k];
        uprev[k] = t;
    }
}

void launch_field_start(hipStream_t st, const unsigned long long *qi, const unsigned long long *qe,
                        int N, double scale_i, double scale_e, double corner, double *rho,
                        double coef, double phi0X, double phiR, double phi0Y, double phiLY,
                        double dxdy_inv, double *rhs, double *u, double *uprev, int extrap,
                        int *flags, unsigned long long *zi, unsigned long long *ze) {
    dim3 g((N + 127) / 128, N);
    hipLaunchKernelGGL(k_field_start, g, dim3(128), 0, st, qi, qe, N, scale_i, scale_e, corner,
                       rho, coef, phi0X, phiR, phi0Y, phiLY, dxdy_inv, rhs, u, uprev, extrap,
                       flags, zi, ze);
}

void launch_fold(hipStream_t st, const double *rho, int N, double coef, double phi0X, double phiR,
                 double phi0Y, double phiLY, double dxdy_inv, double *rhs) {
    dim3 g((N + 127) / 128, N);
    hipLaunchKernelGGL(k_fold, g, dim3(128), 0, st, rho, N, coef, phi0X, phiR, phi0Y, phiLY,
                       dxdy_inv, rhs);
}

// ------------------------------------------------------------ phi, E
// fields.cpp:209-219: boundary columns, then boundary rows (rows win corners).
__global__ void k_phi_bc(const double *__restrict__ u, int N, double phi0X, double phiR,
                         double phi0Y, double phiLY, double *__restrict__ phi) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (i >= N) return;
    double v = u[(long)j * N + i];
    if (i == 0) v = phi0X;
    if (i == N - 1) v = phiR;
    if (j == 0) v = phi0Y;
    if (j == N - 1) v = phiLY;
    phi[(long)j * N + i] = v;
}

void launch_phi_bc(hipStream_t st, const double *u, int N, double phi0X, double phiR,
                   double phi0Y, double phiLY, double *phi) {
    dim3 g((N + 127) / 128, N);
    hipLaunchKernelGGL(k_phi_bc, g, dim3(128), 0, st, u, N, phi0X, phiR, phi0Y, phiLY, phi);
}

// fields.cpp:230-268 evaluated per node: central differences inside, linear
// extrapolation at the walls, never-written entries (Ex in rows 0 and N-1, Ey
// in columns 0 and N-1, the padding row) are 0.
__device__ __forceinline__ double exi(const double *phi, int N, int j, int i, double twoDX) {
    if (j < 1 || j > N - 2 || i < 1 || i > N - 2) return 0.0;
    return (phi[(long)j * N + i - 1] - phi[(long)j * N + i + 1]) / twoDX;
}
__device__ __forceinline__ double eyi(const double *phi, int N, int j, int i, double twoDX) {
    if (j < 1 || j > N - 2 || i < 1 || i > N - 2) return 0.0;
    return (phi[(long)(j - 1) * N + i] - phi[(long)(j + 1) * N + i]) / twoDX;
}

__global__ void k_efield(const double *__restrict__ phi, int N, double twoDX,
                         double *__restrict__ ex64, double *__restrict__ ey64,
                         float2 *__restrict__ E32) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long tot32 = (long)(N + 1) * N + 1, tot64 = (long)(N + 1) * (N + 1);
    if (k >= tot64) return;
    if (k >= (long)N * N) {
        if (k < tot32) E32[k] = make_float2(0.0f, 0.0f);
        ex64[k] = 0.0;
        ey64[k] = 0.0;
        return;
    }
    int j = (int)(k / N), i = (int)(k % N);
    double ex, ey;
    if (i == 0)
        ex = 2.0 * exi(phi, N, j, 1, twoDX) - exi(phi, N, j, 2, twoDX);
    else if (i == N - 1)
        ex = 2.0 * exi(phi, N, j, N - 2, twoDX) - exi(phi, N, j, N - 3, twoDX);
    else
        ex = exi(phi, N, j, i, twoDX);
    if (j == 0)
        ey = 2.0 * eyi(phi, N, 1, i, twoDX) - eyi(phi, N, 2, i, twoDX);
    else if (j == N - 1)
        ey = 2.0 * eyi(phi, N, N - 2, i, twoDX) - eyi(phi, N, N - 3, i, twoDX);
    else
        ey = eyi(phi, N, j, i, twoDX);
    ex64[k] = ex;
    ey64[k] = ey;
    E32[k] = make_float2((float)ex, (float)ey);
}

// phi and E in one pass: phi at any node from the solver's iterate and the
// boundary values (k_phi_bc's rule), E as k_efield computes it from phi.  The
// values are those of the two-kernel path, bit for bit.
struct PhiBc {
    double x0, xR, y0, yL;
};
__device__ __forceinline__ double phi_at(const double *u, int N, int j, int i, const PhiBc &b) {
    double v = u[(long)j * N + i];
    if (i == 0) v = b.x0;
    if (i == N - 1) v = b.xR;
    if (j == 0) v = b.y0;
    if (j == N - 1) v = b.yL;
    return v;
}
__device__ __forceinline__ double exu(const double *u, int N, int j, int i, double twoDX,
                                      const PhiBc &b) {
    if (j < 1 || j > N - 2 || i < 1 || i > N - 2) return 0.0;
    return (phi_at(u, N, j, i - 1, b) - phi_at(u, N, j, i + 1, b)) / twoDX;
}
__device__ __forceinline__ double eyu(const double *u, int N, int j, int i, double twoDX,
                                      const PhiBc &b) {
    if (j < 1 || j > N - 2 || i < 1 || i > N - 2) return 0.0;
    return (phi_at(u, N, j - 1, i, b) - phi_at(u, N, j + 1, i, b)) / twoDX;
}

__global__ void k_phi_efield(const double *__restrict__ u0, const double *__restrict__ u1,
                             const int *__restrict__ flags, int N, PhiBc b, double twoDX,
                             double *__restrict__ phi, double *__restrict__ ex64,
                             double *__restrict__ ey64, float2 *__restrict__ E32) {
    const double *__restrict__ u = (flags && (flags[1] & 1)) ? u1 : u0;
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long tot32 = (long)(N + 1) * N + 1, tot64 = (long)(N + 1) * (N + 1);
    if (k >= tot64) return;
    if (k >= (long)N * N) {
        if (k < tot32) E32[k] = make_float2(0.0f, 0.0f);
        ex64[k] = 0.0;
        ey64[k] = 0.0;
        return;
    }
    int j = (int)(k / N), i = (int)(k % N);
    phi[k] = phi_at(u, N, j, i, b);
    double ex, ey;
    if (i == 0)
        ex = 2.0 * exu(u, N, j, 1, twoDX, b) - exu(u, N, j, 2, twoDX, b);
    else if (i == N - 1)
        ex = 2.0 * exu(u, N, j, N - 2, twoDX, b) - exu(u, N, j, N - 3, twoDX, b);
    else
        ex = exu(u, N, j, i, twoDX, b);
    if (j == 0)
        ey = 2.0 * eyu(u, N, 1, i, twoDX, b) - eyu(u, N, 2, i, twoDX, b);
    else if (j == N - 1)
        ey = 2.0 * eyu(u, N, N - 2, i, twoDX, b) - eyu(u, N, N - 3, i, twoDX, b);
    else
        ey = eyu(u, N, j, i, twoDX, b);
    ex64[k] = ex;
    ey64[k] = ey;
    E32[k] = make_float2((float)ex, (float)ey);
}

void launch_phi_efield(hipStream_t st, const double *u, int N, double phi0X, double phiR,
                       double phi0Y, double phiLY, double dx, double *phi, double *ex64,
                       double *ey64, float2 *E32, const double *u1, const int *flags) {
    long tot = (long)(N + 1) * (N + 1);
    hipLaunchKernelGGL(k_phi_efield, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, u, u1,
                       flags, N, PhiBc{phi0X, phiR, phi0Y, phiLY}, 2.0 * dx, phi, ex64, ey64, E32);
}

void launch_efield(hipStream_t st, const double *phi, int N, double dx, double *ex64, double *ey64,
                   float2 *E32) {
    long tot = (long)(N + 1) * (N + 1);
    hipLaunchKernelGGL(k_efield, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, phi, N,
                       2.0 * dx, ex64, ey64, E32);
}

// ------------------------------------------------------ multigrid: nodes
// Arrays are row-major, u[i*n + j] == NR's u[i][j].

template <typename T>
__device__ __forceinline__ T relax_node(const T *u, const T *rhs, int n, int i, int j, T h2) {
    return (T)0.25 * (u[(i + 1) * n + j] + u[(i - 1) * n + j] + u[i * n + j + 1] +
                      u[i * n + j - 1] - h2 * rhs[i * n + j]);
}

template <typename T>
__device__ __forceinline__ T resid_node(const T *u, const T *rhs, int n, int i, int j, T h2i) {
    return -h2i * (u[(i + 1) * n + j] + u[(i - 1) * n + j] + u[i * n + j + 1] +
                   u[i * n + j - 1] - (T)4.0 * u[i * n + j]) +
           rhs[i * n + j];
}

// half weighting of the residual at coarse (ic,jc), residual evaluated on
// the fly (mglin.cpp:151-169 applied to mglin.cpp:133-144's output)
template <typename T>
__device__ __forceinline__ T resid_rstrct_node(const T *u, const T *rhs, int nf, int ic, int jc,
                                               T h2i) {
    int nc = (nf + 1) / 2;
    if (ic == 0 || jc == 0 || ic == nc - 1 || jc == nc - 1) return (T)0.0;
    int i = 2 * ic, j = 2 * jc;
    return (T)0.5 * resid_node(u, rhs, nf, i, j, h2i) +
           (T)0.125 * (resid_node(u, rhs, nf, i + 1, j, h2i) + resid_node(u, rhs, nf, i - 1, j, h2i) +
                       resid_node(u, rhs, nf, i, j + 1, h2i) + resid_node(u, rhs, nf, i, j - 1, h2i));
}

template <typename T>
__device__ __forceinline__ T rstrct_node(const T *uf, int nc, int ic, int jc) {
    int nf = 2 * nc - 1, ncc = 2 * nc - 2;
    if (jc == 0) return uf[(2 * ic) * nf + 0];
    if (jc == nc - 1 && ic != 0 && ic != nc - 1) return uf[(2 * ic) * nf + ncc];
    if (ic == 0) return uf[0 * nf + 2 * jc];
    if (ic == nc - 1) return uf[ncc * nf + 2 * jc];
    int i = 2 * ic, j = 2 * jc;
    return (T)0.5 * uf[i * nf + j] + (T)0.125 * (uf[(i + 1) * nf + j] + uf[(i - 1) * nf + j] +
                                                 uf[i * nf + j + 1] + uf[i * nf + j - 1]);
}

// bilinear prolongation value at fine (i,j) (mglin.cpp:70-82)
template <typename T>
__device__ __forceinline__ T interp_even_col(const T *uc, int nc, int i, int j) {
    if ((i & 1) == 0) return uc[(i >> 1) * nc + (j >> 1)];
    return (T)0.5 * (uc[((i + 1) >> 1) * nc + (j >> 1)] + uc[((i - 1) >> 1) * nc + (j >> 1)]);
}
template <typename T>
__device__ __forceinline__ T interp_node(const T *uc, int nc, int i, int j) {
    if ((j & 1) == 0) return interp_even_col(uc, nc, i, j);
    return (T)0.5 * (interp_even_col(uc, nc, i, j + 1) + interp_even_col(uc, nc, i, j - 1));
}

// ------------------------------------------- multigrid: grid-wide kernels
// `skip` (may be null) is the tolerance solver's converged flag: once set,
// the remaining queued iterations return immediately.
#define SKIP_IF_DONE(skip) \
    if ((skip) && *(skip)) return

template <typename T>
__global__ void k_mg_relax(T *__restrict__ u, const T *__restrict__ rhs, int n, T h2, int color,
                           const int *__restrict__ skip) {
    SKIP_IF_DONE(skip);
    int i = blockIdx.y + 1;  // interior row
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int j = 1 + ((i + 1 + color) & 1) + 2 * t;  // (i+j)%2 == color
    if (i > n - 2 || j > n - 2) return;
    u[i * n + j] = relax_node(u, rhs, n, i, j, h2);
}

template <typename T>
__global__ void k_mg_resid_rstrct(const T *__restrict__ u, const T *__restrict__ rhs, int nf, T h2i,
                                  T *__restrict__ uc, const int *__restrict__ skip) {
    SKIP_IF_DONE(skip);
    int nc = (nf + 1) / 2;
    int jc = blockIdx.x * blockDim.x + threadIdx.x, ic = blockIdx.y;
    if (jc >= nc) return;
    uc[ic * nc + jc] = resid_rstrct_node(u, rhs, nf, ic, jc, h2i);
}

template <typename T>
__global__ void k_mg_rstrct(T *__restrict__ uc, int nc, const T *__restrict__ uf) {
    int jc = blockIdx.x * blockDim.x + threadIdx.x, ic = blockIdx.y;
    if (jc >= nc) return;
    uc[ic * nc + jc] = rstrct_node(uf, nc, ic, jc);
}

template <typename T>
__global__ void k_mg_interp(T *__restrict__ uf, int nf, const T *__restrict__ uc) {
    int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= nf) return;
    uf[i * nf + j] = interp_node(uc, (nf + 1) / 2, i, j);
}

template <typename T>
__global__ void k_mg_addint(T *__restrict__ uf, int nf, const T *__restrict__ uc,
                            const int *__restrict__ skip) {
    SKIP_IF_DONE(skip);
    int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= nf) return;
    uf[i * nf + j] = uf[i * nf + j] + interp_node(uc, (nf + 1) / 2, i, j);
}

template <typename T>
__global__ void k_fill(T *__restrict__ p, long n, T v, const int *__restrict__ skip) {
    SKIP_IF_DONE(skip);
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

template <typename T>
void launch_fill(hipStream_t st, T *p, long n, T v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v,
                       (const int *)nullptr);
}
template void launch_fill<double>(hipStream_t, double *, long, double);
template void launch_fill<float>(hipStream_t, float *, long, float);
template void launch_fill<unsigned long long>(hipStream_t, unsigned long long *, long,
                                              unsigned long long);
template void launch_fill<uint32_t>(hipStream_t, uint32_t *, long, uint32_t);

// --------------------------------------- multigrid: one-workgroup coarse part

static __host__ __device__ inline int lvl_size(int l) { return (1 << (l + 1)) + 1; }

// LDS layout of the coarse part topped at level lc (levels 0..lc):
//   rhs[0..lc] | cres[0], cv[0], cres[1], cv[1], ... cres[lc-1], cv[lc-1] | u
// Offsets are computed arithmetically (wave-uniform) so no pointer table is
// kept per lane.
static __host__ __device__ inline long sq_below(int l) {
    long s = 0;
    for (int k = 0; k < l; k++) s += (long)lvl_size(k) * lvl_size(k);
    return s;
}

// doubles (or floats) of LDS needed for a coarse part topped at level lc
static __host__ __device__ inline long coarse_lds_elems(int lc) {
    return sq_below(lc + 1) + 2 * sq_below(lc) + (long)lvl_size(lc) * lvl_size(lc);
}

template <typename T>
struct CoarseLds {
    T *base;
    int lc;
    __device__ __forceinline__ T *rhs(int l) const { return base + sq_below(l); }
    __device__ __forceinline__ T *cres(int l) const {
        return base + sq_below(lc + 1) + 2 * sq_below(l);
    }
    __device__ __forceinline__ T *cv(int l) const { return cres(l) + (long)lvl_size(l) * lvl_size(l); }
    __device__ __forceinline__ T *u() const { return base + sq_below(lc + 1) + 2 * sq_below(lc); }
};

__device__ __forceinline__ double lvl_h2(int l, double length) {
    double h = length / (lvl_size(l) - 1);
    return h * h;
}

template <typename T>
__device__ void co_relax(T *u, const T *rhs, int n, T h2) {
    int m = n - 2;  // interior per row
    int half = (m + 1) / 2;
    for (int color = 0; color < 2; color++) {
        for (int k = threadIdx.x; k < m * half; k += blockDim.x) {
            int i = 1 + k / half, t = k % half;
            int j = 1 + ((i + 1 + color) & 1) + 2 * t;
            if (j <= n - 2) u[i * n + j] = relax_node(u, rhs, n, i, j, h2);
        }
        __syncthreads();
    }
}

// residual restricted to the coarse grid (mglin.cpp:189-190) and the coarse
// correction zeroed (mglin.cpp:185), one pass
template <typename T>
__device__ void co_resid_rstrct_zero(const T *u, const T *rhs, int nf, T h2i, T *cres, T *cv) {
    int nc = (nf + 1) / 2;
    for (int k = threadIdx.x; k < nc * nc; k += blockDim.x) {
        cres[k] = resid_rstrct_node(u, rhs, nf, k / nc, k % nc, h2i);
        cv[k] = (T)0.0;
    }
    __syncthreads();
}

template <typename T>
__device__ void co_rstrct(T *uc, int nc, const T *uf) {
    for (int k = threadIdx.x; k < nc * nc; k += blockDim.x) uc[k] = rstrct_node(uf, nc, k / nc, k % nc);
    __syncthreads();
}

template <typename T>
__device__ void co_interp(T *uf, int nf, const T *uc, bool add) {
    int nc = (nf + 1) / 2;
    for (int k = threadIdx.x; k < nf * nf; k += blockDim.x) {
        T v = interp_node(uc, nc, k / nf, k % nf);
        uf[k] = add ? uf[k] + v : v;
    }
    __syncthreads();
}

template <typename T>
__device__ void co_zero(T *p, int n) {
    for (int k = threadIdx.x; k < n; k += blockDim.x) p[k] = (T)0.0;
    __syncthreads();
}

template <typename T>
__device__ void co_slvsml(T *u, const T *rhs, double length) {
    // mglin.cpp:103-110
    if (threadIdx.x < 9) {
        T h = (T)(0.5 * length);
        u[threadIdx.x] = (threadIdx.x == 4) ? -h * h * rhs[4] / (T)4.0 : (T)0.0;
    }
    __syncthreads();
}

// V(1,1) cycle at level j on (U, R), recursion unrolled (mglin.cpp:176-196)
template <typename T>
__device__ void co_vcycle(const CoarseLds<T> &L, int j, T *U, const T *R, double length) {
    T *Ul = U;
    const T *Rl = R;
    for (int l = j; l >= 1; l--) {
        int n = lvl_size(l);
        double h2 = lvl_h2(l, length);
        co_relax(Ul, Rl, n, (T)h2);
        co_resid_rstrct_zero(Ul, Rl, n, (T)(1.0 / h2), L.cres(l - 1), L.cv(l - 1));
        Ul = L.cv(l - 1);
        Rl = L.cres(l - 1);
    }
    co_slvsml(L.cv(0), L.cres(0), length);
    for (int l = 1; l <= j; l++) {
        T *Uu = (l == j) ? U : L.cv(l);
        const T *Ru = (l == j) ? R : L.cres(l);
        co_interp(Uu, lvl_size(l), L.cv(l - 1), true);
        co_relax(Uu, Ru, lvl_size(l), (T)lvl_h2(l, length));
    }
}

// mode 0: full multigrid from rhs_in (level lc) -> u_out;  mode 1: ncycle
// V-cycles on a zero initial guess (the coarse-grid correction of a finer
// V-cycle) -> u_out.
template <typename T>
__global__ void __launch_bounds__(1024)
    k_mg_coarse(const T *__restrict__ rhs_in, T *__restrict__ u_out, int lc, int mode, int ncycle,
                double length, const int *__restrict__ skip) {
    SKIP_IF_DONE(skip);
    extern __shared__ double lds_raw[];
    CoarseLds<T> L{reinterpret_cast<T *>(lds_raw), lc};
    int nc = lvl_size(lc);
    T *top = L.rhs(lc);
    for (int k = threadIdx.x; k < nc * nc; k += blockDim.x) top[k] = rhs_in[k];
    __syncthreads();
    if (mode == 0) {
        for (int l = lc - 1; l >= 0; l--) co_rstrct(L.rhs(l), lvl_size(l), L.rhs(l + 1));
        // the full-multigrid iterate of level j < lc lives in cv(j) (free at
        // that point: a V-cycle at level j only touches cv(0..j-1))
        T *prev = (lc == 0) ? L.u() : L.cv(0);
        co_slvsml(prev, L.rhs(0), length);
        for (int j = 1; j <= lc; j++) {
            T *Uj = (j == lc) ? L.u() : L.cv(j);
            co_interp(Uj, lvl_size(j), prev, false);
            for (int c = 0; c < ncycle; c++) co_vcycle(L, j, Uj, L.rhs(j), length);
            prev = Uj;
        }
    } else {
        co_zero(L.u(), nc * nc);
        for (int c = 0; c < ncycle; c++) co_vcycle(L, lc, L.u(), L.rhs(lc), length);
    }
    T *u = L.u();
    for (int k = threadIdx.x; k < nc * nc; k += blockDim.x) u_out[k] = u[k];
}

// ------------------------------------- exact coarsest-level solve (TOL)
// The 5-point Laplacian with zero walls is diagonalised by the discrete sine
// transform: with S[p][q] = sin(pi (p+1)(q+1)/M) (S S = M/2 I) and
// lambda_k = -4 sin^2(pi (k+1) / 2M), the solution of A u = f on the m = M-1
// interior nodes is u = (2/M)^2 S ((S f S) ./ ((lambda_p + lambda_q)/h^2)) S:
// four M x M products, done here in LDS by one workgroup (2x2 register tiles,
// fp32).  It replaces the coarse V-cycle of the tolerance mode's correction
// solve; the FP64 refinement around it is unchanged.

// C = A B for M x M matrices in LDS, rows ld floats apart (M, ld multiples
// of 4, M <= 64): thread t < (M/4)^2 computes a 4x4 tile, per k four
// broadcast reads of A and one float4 of B feeding 16 FMAs.  Ends with a
// barrier.
__device__ void lds_gemm(const float *A, const float *B, float *C, int M, int ld) {
    const int T = M >> 2;
    const int t = threadIdx.x;
    if (t < T * T) {
        const int ti = t / T, tj = t - (t / T) * T;
        const float *a = A + (4 * ti) * ld;
        float c[4][4] = {};
        for (int k = 0; k < M; k += 4) {  // four k per step: float4 reads of A and B
            float4 ar[4], br[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                ar[r] = *reinterpret_cast<const float4 *>(a + r * ld + k);
                br[r] = *reinterpret_cast<const float4 *>(B + (k + r) * ld + 4 * tj);
            }
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const float4 b = br[kk];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float x = kk == 0 ? ar[r].x : kk == 1 ? ar[r].y : kk == 2 ? ar[r].z : ar[r].w;
                    c[r][0] = fmaf(x, b.x, c[r][0]);
                    c[r][1] = fmaf(x, b.y, c[r][1]);
                    c[r][2] = fmaf(x, b.z, c[r][2]);
                    c[r][3] = fmaf(x, b.w, c[r][3]);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 4; r++)
            *reinterpret_cast<float4 *>(C + (4 * ti + r) * ld + 4 * tj) =
                make_float4(c[r][0], c[r][1], c[r][2], c[r][3]);
    }
    __syncthreads();
}

// The same product on the matrix cores (M a multiple of 16): one 16 x 16
// block of C per wavefront, v_mfma_f32_16x16x4f32 over k.  Lane layout of
// the instruction: A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k,
// C[i][j] in lane j + 16 (i / 4), register i % 4.  A row stride ld = M + 4
// keeps the 16 rows of an A fragment on 8 banks (2-way at most).  Ends with
// a barrier.
typedef float mf4 __attribute__((ext_vector_type(4)));
__device__ void lds_gemm_mfma(const float *A, const float *B, float *C, int M, int ld) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int nb = M >> 4, li = lane & 15, lk = lane >> 4;
    for (int blk = w; blk < nb * nb; blk += nw) {
        const int bi = blk / nb, bj = blk - (blk / nb) * nb;
        const float *a = A + (16 * bi + li) * ld + lk;
        const float *b = B + lk * ld + 16 * bj + li;
        mf4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k = 0; k < M; k += 4)
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[k], b[k * ld], c, 0, 0, 0);
        float *o = C + (16 * bi + 4 * lk) * ld + 16 * bj + li;
        o[0] = c[0];
        o[ld] = c[1];
        o[2 * ld] = c[2];
        o[3 * ld] = c[3];
    }
    __syncthreads();
}

// the products of the unfolded DST (M < 64; M = 64 takes dstf_solve)
__device__ __forceinline__ void dst_gemm(const float *A, const float *B, float *C, int M, int ld) {
    if ((M & 15) == 0)
        lds_gemm_mfma(A, B, C, M, ld);
    else
        lds_gemm(A, B, C, M, ld);
}

// LDS floats of the DST solve: three M x M matrices with row stride M + 4
__host__ __device__ inline long dst_lds_floats(int M) { return 3L * M * (M + 4); }

// ---- the DST for M = 64 with the sine matrix folded in halves
// S[a][b] = sin(pi a b / 64) (a, b = 1..63, 0-based index a - 1) has
// S[a][64 - b] = (-1)^(a+1) S[a][b]: for odd a, (S x)_a sums
// S[a][b] (x_b + x_{64-b}) over b = 1..31 plus S[a][32] x_32, for even a
// S[a][b] (x_b - x_{64-b}) over b = 1..31.  So each 64 x 64 product becomes
// two 32 x 32 products of a folded operand (half the matrix-core work):
//   H1[i][p] = sin(pi (2i+1)(p+1) / 64), i, p = 0..31 (odd a = 2i+1, b = p+1)
//   H2[i][p] = sin(pi (2i+2)(p+1) / 64), i, p = 0..30, row and column 31 zero.
// The forward transform takes natural order to "split" order (odd
// frequencies 0..31, even 32..62, 63 zero); the inverse, the same tables
// transposed, takes split order back: y_c = U_c + V_c and
// y_{64-c} = U_c - V_c with U, V the odd and even frequencies' sums.
// Tables in LDS: H1 at Hs, H2 at Hs + 32 DSTF_LDH.  The eigenvalue scale is
// taken in split order (dst_tables writes it so).
#define DSTF_LDH 36
__host__ __device__ inline long dstf_table_floats() { return 2L * 32 * DSTF_LDH; }

// one 16 x 16 block of C = A B with k = 32 on the matrix cores; A(i, k) is
// A[k lda + i] when AT, else A[i lda + k]; B(k, j) is B[j ldb + k] when BT,
// else B[k ldb + j] (operands offset to the block); the lane layout of
// v_mfma_f32_16x16x4f32 as in lds_gemm_mfma
template <bool AT, bool BT>
__device__ __forceinline__ void dstf_block(const float *A, int lda, const float *B, int ldb,
                                           float *C, int ldc) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    mf4 c0 = {0.0f, 0.0f, 0.0f, 0.0f}, c1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const int k = 4 * t + lk;
        const float a = AT ? A[k * lda + li] : A[li * lda + k];
        const float b = BT ? B[li * ldb + k] : B[k * ldb + li];
        if (t & 1)
            c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c1, 0, 0, 0);
        else
            c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c0, 0, 0, 0);
    }
    const mf4 cs = c0 + c1;
    float *o = C + (4 * lk) * ldc + li;
    o[0] = cs[0];
    o[ldc] = cs[1];
    o[2 * ldc] = cs[2];
    o[3 * ldc] = cs[3];
}

// C rows 0..31 = Ha op B rows 0..31, rows 32..63 = Hb op B rows 32..63 (left
// products, 16 wavefronts: one 16 x 16 block each); HT: the tables transposed
template <bool HT>
__device__ __forceinline__ void dstf_left(const float *Hs, const float *B, float *C, int ld) {
    const int w = threadIdx.x >> 6, half = w >> 3, bi = (w >> 2) & 1, bj = w & 3;
    const float *H = Hs + half * 32 * DSTF_LDH;
    const float *A = HT ? H + 16 * bi : H + 16 * bi * DSTF_LDH;
    dstf_block<HT, false>(A, DSTF_LDH, B + 32 * half * ld + 16 * bj, ld,
                          C + (32 * half + 16 * bi) * ld + 16 * bj, ld);
    __syncthreads();
}

// C cols 0..31 = A cols 0..31 op Ha, cols 32..63 = A cols 32..63 op Hb
template <bool HT>
__device__ __forceinline__ void dstf_right(const float *Hs, const float *A, float *C, int ld) {
    const int w = threadIdx.x >> 6, half = w >> 3, bi = (w >> 1) & 3, bj = w & 1;
    const float *H = Hs + half * 32 * DSTF_LDH;
    const float *B = HT ? H + 16 * bj * DSTF_LDH : H + 16 * bj;
    dstf_block<false, HT>(A + 16 * bi * ld + 32 * half, ld, B, DSTF_LDH,
                          C + 16 * bi * ld + 32 * half + 16 * bj, ld);
    __syncthreads();
}

// fold (sums then differences) of the rows (ROWS) or columns of X into W
template <bool ROWS>
__device__ __forceinline__ void dstf_fold(const float *X, float *W, int ld) {
#pragma unroll 1  // elementwise: few registers beat the unrolled addresses
    for (int r = 0; r < 4; r++) {
        const int k = (int)threadIdx.x + r * 1024, p = k >> 6, q = k & 63;
        const int o = ROWS ? p : q;  // the folded index
        float v = 0.0f;
        auto at = [&](int idx) { return ROWS ? X[idx * ld + q] : X[p * ld + idx]; };
        if (o < 31)
            v = at(o) + at(62 - o);
        else if (o == 31)
            v = at(31);
        else if (o < 63)
            v = at(o - 32) - at(94 - o);
        W[p * ld + q] = v;
    }
    __syncthreads();
}

// unfold of the rows (ROWS) or columns of W (U in 0..31, V in 32..63) into X
template <bool ROWS>
__device__ __forceinline__ void dstf_unfold(const float *W, float *X, int ld) {
#pragma unroll 1  // elementwise: few registers beat the unrolled addresses
    for (int r = 0; r < 4; r++) {
        const int k = (int)threadIdx.x + r * 1024, p = k >> 6, q = k & 63;
        const int o = ROWS ? p : q;  // the natural index
        float v = 0.0f;
        auto at = [&](int idx) { return ROWS ? W[idx * ld + q] : W[p * ld + idx]; };
        if (o < 31)
            v = at(o) + at(32 + o);
        else if (o == 31)
            v = at(31);
        else if (o < 63)
            v = at(62 - o) - at(94 - o);
        X[p * ld + q] = v;
    }
    __syncthreads();
}

// X (64 x 64, natural order, row and column 63 zero) <- the DST solve of X;
// W: scratch of the same shape; dk: this thread's entries k = threadIdx.x +
// r 1024 of the split-order eigenvalue scale
__device__ __forceinline__ void dstf_solve(float *X, float *W, const float *Hs, const float dk[4],
                                           int ld) {
    dstf_fold<true>(X, W, ld);
    dstf_left<false>(Hs, W, X, ld);  // rows in split order
    dstf_fold<false>(X, W, ld);
    dstf_right<true>(Hs, W, X, ld);  // columns too
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int k = (int)threadIdx.x + r * 1024, p = k >> 6, q = k & 63;
        X[p * ld + q] = X[p * ld + q] * dk[r];
    }
    __syncthreads();
    dstf_left<true>(Hs, X, W, ld);
    dstf_unfold<true>(W, X, ld);  // rows in natural order
    dstf_right<false>(Hs, X, W, ld);
    dstf_unfold<false>(W, X, ld);  // columns too
}

// the DST solve by one workgroup of 1024 threads; lds_raw: dst_lds_floats(M)
__device__ void dst_solve(const float *__restrict__ rhs_in, float *__restrict__ u_out, int lc,
                          float h2, const float *__restrict__ Sg, const float *__restrict__ Dg,
                          double *lds_raw) {
    const int M = 1 << (lc + 1), n = M + 1, ld = M + 4;
    float *S = reinterpret_cast<float *>(lds_raw);
    float *X = S + M * ld, *Y = X + M * ld;
    // the eigenvalue scale, read now so its latency overlaps the products
    float dk[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int k = threadIdx.x + r * blockDim.x;
        dk[r] = k < M * M ? Dg[k] * h2 : 0.0f;
    }
    for (int k = threadIdx.x; k < M * M; k += blockDim.x) {
        const int p = k >> (lc + 1), q = k & (M - 1);
        X[p * ld + q] = (p < M - 1 && q < M - 1) ? rhs_in[(p + 1) * n + q + 1] : 0.0f;
    }
    if (M == 64) {  // the folded tables (dstf_solve), in S's space
        for (int k = threadIdx.x; k < 2048; k += blockDim.x)
            S[(k >> 10) * 32 * DSTF_LDH + ((k >> 5) & 31) * DSTF_LDH + (k & 31)] = Sg[k];
        __syncthreads();
        dstf_solve(X, Y, S, dk, ld);  // X = u on the interior
    } else {
        for (int k = threadIdx.x; k < M * M; k += blockDim.x)
            S[(k >> (lc + 1)) * ld + (k & (M - 1))] = Sg[k];
        __syncthreads();
        dst_gemm(S, X, Y, M, ld);  // Y = S f
        dst_gemm(Y, S, X, M, ld);  // X = S f S
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int k = threadIdx.x + r * blockDim.x;
            if (k < M * M) {
                const int p = k >> (lc + 1), q = k & (M - 1);
                X[p * ld + q] = X[p * ld + q] * dk[r];
            }
        }
        __syncthreads();
        dst_gemm(S, X, Y, M, ld);
        dst_gemm(Y, S, X, M, ld);  // X = u on the interior
    }
    for (int k = threadIdx.x; k < n * n; k += blockDim.x) {
        const int i = k / n, j = k - (k / n) * n;
        const bool in = i > 0 && j > 0 && i < n - 1 && j < n - 1;
        u_out[k] = in ? X[(i - 1) * ld + (j - 1)] : 0.0f;
    }
}

__global__ void __launch_bounds__(1024)
    k_mg_coarse_dst(const float *__restrict__ rhs_in, float *__restrict__ u_out, int lc, float h2,
                    const float *__restrict__ Sg, const float *__restrict__ Dg,
                    const int *__restrict__ skip) {
    SKIP_IF_DONE(skip);
    extern __shared__ double lds_raw[];
    dst_solve(rhs_in, u_out, lc, h2, Sg, Dg, lds_raw);
}

// The verdict of the tolerance solve's last residual pass, formed at the
// start of the next cycle's first kernel instead of by k_tol_finish (one
// launch fewer per cycle): every block sums the pass's partials as
// k_tol_finish does (256 threads, the same order, so the same residual) and
// leaves if the solve has converged; block (0, 0) also records and
// publishes the verdict as k_tol_finish does (applied: the pass followed a
// cycle, not the solve's start).  partials null: no verdict.
struct TolVerdict {
    const double *partials;
    int nb;
    double tol;
    int applied;
    int *flags;
    double *relres_host;
    int *host_flags;
};

// block sums of the residual passes (TOL_THREADS threads, fixed order)
#define TOL_THREADS 256

__device__ __forceinline__ double block_sum_d(double v) {
    __shared__ double ws[TOL_THREADS / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) ws[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < TOL_THREADS / 64; k++) t += ws[k];  // fixed order
    __syncthreads();
    return t;
}

// true: the solve goes on (this block does the cycle's work).  flags[0] may
// turn 1 during the kernel (block (0, 0) sets it on convergence, which every
// block then finds itself), so one thread reads it for the whole block.
__device__ __forceinline__ bool tol_verdict_go(const TolVerdict &vd) {
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = *(volatile int *)vd.flags;
    __syncthreads();
    if (s_done) return false;  // converged at an earlier pass: skipped, as k_tol_finish is
    double rr = 0.0, ff = 0.0;
    for (int k = threadIdx.x; k < vd.nb; k += blockDim.x) {
        rr += vd.partials[2 * k];
        ff += vd.partials[2 * k + 1];
    }
    rr = block_sum_d(rr);
    ff = block_sum_d(ff);
    const double rel = (ff > 0.0) ? sqrt(rr / ff) : sqrt(rr);
    const bool conv = rel <= vd.tol;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        int *flags = vd.flags;
        if (vd.applied) flags[1] += 1;
        if (conv) flags[0] = 1;
        *vd.relres_host = rel;
        vd.host_flags[0] = conv ? 1 : 0;
        vd.host_flags[1] = flags[1];
        __threadfence_system();
        flags[3] += 1;
        vd.host_flags[2] = flags[3] | (conv ? MG_PUB_CONV : 0);
    }
    return !conv;
}

// ------------------------------------------- fused fp32 V-cycle levels
// One V(nu,nu) visit of a level above lc takes two launches:
//   pre:  nu red-black Gauss-Seidel sweep pairs (mglin.cpp:116-127) from a
//         zero guess and the half-weighted residual on the coarse nodes
//         (mglin.cpp:133-169);
//   post: prolongation-add (mglin.cpp:70-97) then nu red-black pairs again.
// Each workgroup owns a 32 x 32 tile of fine nodes and recomputes the halo it
// needs in LDS (every sweep of one colour shrinks the valid region by a ring:
// 2 nu + 2 rings for pre, which ends with the residual, 2 nu for post), so no
// other tile's output is read; the per-node arithmetic and colour order are
// those of the grid-wide kernels, so the iterates are identical to theirs.
// nu = 1 is mglin's V(1,1); more sweeps per leg cost little here (a visit is
// bound by its loads and launch, not by the LDS sweeps) and cut the cycles a
// solve needs: the measured two-grid factor is 0.13 for V(1,1), 0.03 for
// V(2,2) (257^2, exact 65^2 coarse level).
// The tile edge T is 16 for levels up to 513^2, where the legs are bound by
// latency and more workgroups finish sooner despite the wider relative halo,
// and 32 above, where the halo's extra reads matter (mg_tile; measured at
// 257^2 and 1025^2).
template <int NU, int T> struct MgHalo {
    static constexpr int PRE = 2 * NU + 2, WPRE = T + 2 * PRE;
    static constexpr int POST = 2 * NU, WPOST = T + 2 * POST;
};

// Gauss-Seidel sweep of one colour over the tile plus `ring` halo rings
// (interior nodes only).  s, f: W x W LDS tiles whose (0,0) is global
// (y0 - H, x0 - H).
template <int T>
__device__ __forceinline__ void tile_sweep(float *s, const float *f, int W, int H, int x0, int y0,
                                           int ring, int n, float h2, int color) {
    const int lo = H - ring, w = T + 2 * ring;
    for (int k = threadIdx.x; k < w * w; k += blockDim.x) {
        const int ly = lo + k / w, lx = lo + k % w;
        const int i = y0 - H + ly, j = x0 - H + lx;
        if (i < 1 || j < 1 || i > n - 2 || j > n - 2 || ((i + j) & 1) != color) continue;
        const int c = ly * W + lx;
        s[c] = 0.25f * (s[c + W] + s[c - W] + s[c + 1] + s[c - 1] - h2 * f[c]);
    }
}

// nu red-black pairs; the valid region ends `last` rings outside the tile
template <int NU, int T>
__device__ __forceinline__ void tile_smooth(float *s, const float *f, int W, int H, int x0, int y0,
                                            int last, int n, float h2) {
#pragma unroll
    for (int k = 0; k < NU; k++) {
        const int ring = last + 2 * (NU - k) - 1;
        tile_sweep<T>(s, f, W, H, x0, y0, ring, n, h2, 0);
        __syncthreads();
        tile_sweep<T>(s, f, W, H, x0, y0, ring - 1, n, h2, 1);
        __syncthreads();
    }
}

__device__ __forceinline__ float tile_resid(const float *s, const float *f, int c, int W, float h2i) {
    return -h2i * (s[c + W] + s[c - W] + s[c + 1] + s[c - 1] - 4.0f * s[c]) + f[c];
}

// pre leg of one T x T tile (tx, ty); su, sf: WPRE^2 LDS floats each
template <int NU, int T>
__device__ __forceinline__ void mg_pre_tile(int tx, int ty, const float *__restrict__ u_in,
                                            const float *__restrict__ rhs, int n, float h2,
                                            float h2i, float *__restrict__ u_out,
                                            float *__restrict__ cres, float *su, float *sf) {
    constexpr int H = MgHalo<NU, T>::PRE, W = MgHalo<NU, T>::WPRE;
    const int x0 = tx * T, y0 = ty * T;
    for (int k = threadIdx.x; k < W * W; k += blockDim.x) {
        const int ly = k / W, lx = k % W;
        const int i = y0 - H + ly, j = x0 - H + lx;
        const bool in = i >= 0 && j >= 0 && i < n && j < n;
        su[k] = (in && u_in) ? u_in[(long)i * n + j] : 0.0f;  // null: zero initial guess
        sf[k] = in ? rhs[(long)i * n + j] : 0.0f;
    }
    __syncthreads();
    tile_smooth<NU, T>(su, sf, W, H, x0, y0, 2, n, h2);  // valid on the tile + 2 rings
    for (int k = threadIdx.x; k < T * T; k += blockDim.x) {
        const int ly = k / T, lx = k % T, i = y0 + ly, j = x0 + lx;
        if (i < n && j < n) u_out[(long)i * n + j] = su[(ly + H) * W + lx + H];
    }
    const int nc = (n + 1) / 2;
    for (int k = threadIdx.x; k < (T / 2) * (T / 2); k += blockDim.x) {
        const int ic = y0 / 2 + k / (T / 2), jc = x0 / 2 + k % (T / 2);
        if (ic >= nc || jc >= nc) continue;
        float v = 0.0f;
        if (ic > 0 && jc > 0 && ic < nc - 1 && jc < nc - 1) {
            const int c = (2 * ic - y0 + H) * W + (2 * jc - x0 + H);
            v = 0.5f * tile_resid(su, sf, c, W, h2i) +
                0.125f * (tile_resid(su, sf, c + W, W, h2i) + tile_resid(su, sf, c - W, W, h2i) +
                          tile_resid(su, sf, c + 1, W, h2i) + tile_resid(su, sf, c - 1, W, h2i));
        }
        cres[(long)ic * nc + jc] = v;
    }
}

template <int NU, int T>
__global__ void __launch_bounds__(256)
    k_mg_pre(const float *__restrict__ u_in, const float *__restrict__ rhs, int n, float h2,
             float h2i, float *__restrict__ u_out, float *__restrict__ cres,
             const int *__restrict__ skip, TolVerdict vd) {
    if (vd.partials) {
        if (!tol_verdict_go(vd)) return;
    } else {
        SKIP_IF_DONE(skip);
    }
    constexpr int W = MgHalo<NU, T>::WPRE;
    __shared__ float su[W * W], sf[W * W];
    mg_pre_tile<NU, T>(blockIdx.x, blockIdx.y, u_in, rhs, n, h2, h2i, u_out, cres, su, sf);
}

// post leg of one tile; su, sf: WPOST^2 LDS floats each
template <int NU, int T>
__device__ __forceinline__ void mg_post_tile(int tx, int ty, const float *__restrict__ us,
                                             const float *__restrict__ rhs,
                                             const float *__restrict__ cv, int n, float h2,
                                             float *__restrict__ u_out, float *su, float *sf) {
    constexpr int H = MgHalo<NU, T>::POST, W = MgHalo<NU, T>::WPOST;
    const int x0 = tx * T, y0 = ty * T, nc = (n + 1) / 2;
    for (int k = threadIdx.x; k < W * W; k += blockDim.x) {
        const int ly = k / W, lx = k % W;
        const int i = y0 - H + ly, j = x0 - H + lx;
        const bool in = i >= 0 && j >= 0 && i < n && j < n;
        su[k] = in ? us[(long)i * n + j] + interp_node(cv, nc, i, j) : 0.0f;
        sf[k] = in ? rhs[(long)i * n + j] : 0.0f;
    }
    __syncthreads();
    tile_smooth<NU, T>(su, sf, W, H, x0, y0, 0, n, h2);  // valid on the tile
    for (int k = threadIdx.x; k < T * T; k += blockDim.x) {
        const int ly = k / T, lx = k % T, i = y0 + ly, j = x0 + lx;
        if (i < n && j < n) u_out[(long)i * n + j] = su[(ly + H) * W + lx + H];
    }
}

template <int NU, int T>
__global__ void __launch_bounds__(256)
    k_mg_post(const float *__restrict__ us, const float *__restrict__ rhs,
              const float *__restrict__ cv, int n, float h2, float *__restrict__ u_out,
              const int *__restrict__ skip) {
    SKIP_IF_DONE(skip);
    constexpr int W = MgHalo<NU, T>::WPOST;
    __shared__ float su[W * W], sf[W * W];
    mg_post_tile<NU, T>(blockIdx.x, blockIdx.y, us, rhs, cv, n, h2, u_out, su, sf);
}

static void coarse_dst_launch(hipStream_t st, const MgLevels &M, const float *rhs, float *u,
                              double length, const int *skip);

static inline int mg_tile(int nf) { return nf <= 513 ? 16 : 32; }
// LAUNCH(nu, T) for the run-time sweep count nu (1..3) and tile edge T (16, 32)
#define MG_DISPATCH(NUV, TV, LAUNCH)                  \
    do {                                              \
        if ((TV) == 16) {                             \
            if ((NUV) == 3) LAUNCH(3, 16);            \
            else if ((NUV) == 2) LAUNCH(2, 16);       \
            else LAUNCH(1, 16);                       \
        } else {                                      \
            if ((NUV) == 3) LAUNCH(3, 32);            \
            else if ((NUV) == 2) LAUNCH(2, 32);       \
            else LAUNCH(1, 32);                       \
        }                                             \
    } while (0)

// fused V(nu,nu) at level j > lc on (Uj, Rj); Uj's initial guess is zero;
// vd: the verdict the first kernel forms first (TolVerdict)
static void vcycle_fused(hipStream_t st, MgLevels &M, int j, float *Uj, const float *Rj,
                         double length, const int *skip, const TolVerdict &vd = TolVerdict{}) {
    MgArrays<float> &A = M.f;
    const int nf = lvl_size(j);
    const double h = length / (nf - 1);
    const int T = mg_tile(nf), tiles = (nf + T - 1) / T;
    const float h2 = (float)(h * h), h2i = (float)(1.0 / (h * h));
    const dim3 g(tiles, tiles), b(256);
    const float *z = nullptr;
#define PRE(NU, TT)                                                                          \
    hipLaunchKernelGGL((k_mg_pre<NU, TT>), g, b, 0, st, z, Rj, nf, h2, h2i, A.us[j], A.cres[j - 1], \
                       skip, vd)
    MG_DISPATCH(M.nu, T, PRE);
#undef PRE
    if (j - 1 == M.lc)
        coarse_dst_launch(st, M, A.cres[j - 1], A.cv[j - 1], length, skip);
    else
        vcycle_fused(st, M, j - 1, A.cv[j - 1], A.cres[j - 1], length, skip);
#define POST(NU, TT) \
    hipLaunchKernelGGL((k_mg_post<NU, TT>), g, b, 0, st, A.us[j], Rj, A.cv[j - 1], nf, h2, Uj, skip)
    MG_DISPATCH(M.nu, T, POST);
#undef POST
}

static void coarse_dst_launch(hipStream_t st, const MgLevels &M, const float *rhs, float *u,
                              double length, const int *skip) {
    const int lc = M.lc, m = 1 << (lc + 1);
    double h = length / (double)m;
    hipLaunchKernelGGL(k_mg_coarse_dst, dim3(1), dim3(1024), (size_t)dst_lds_floats(m) * sizeof(float),
                       st, rhs, u, lc, (float)(h * h), M.dst_S, M.dst_D, skip);
}

int mg_dst_tables(MgLevels &M) {
    const int m = 1 << (M.lc + 1);
    const double pi = 3.14159265358979323846;
    std::vector<float> S((size_t)m * m, 0.0f), D((size_t)m * m, 0.0f);
    for (int p = 0; p < m - 1; p++) {
        double lp = -4.0 * std::pow(std::sin(pi * (p + 1) / (2.0 * m)), 2);
        for (int q = 0; q < m - 1; q++) {
            // (p+1)(q+1) mod 2m keeps the sine argument small
            long r = ((long)(p + 1) * (q + 1)) % (2L * m);
            S[(size_t)p * m + q] = (float)std::sin(pi * (double)r / (double)m);
            double lq = -4.0 * std::pow(std::sin(pi * (q + 1) / (2.0 * m)), 2);
            D[(size_t)p * m + q] = (float)((2.0 / m) * (2.0 / m) / (lp + lq));
        }
    }
    if (m == 64) {
        // the folded form (dstf_solve): H1 (odd frequencies a = 2i+1) and H2
        // (even a = 2i+2) over b = p+1, then the scale in split order
        auto sinab = [&](long a, long b) {
            return (float)std::sin(pi * (double)((a * b) % (2L * m)) / (double)m);
        };
        std::vector<float> H(2048, 0.0f), Ds((size_t)m * m, 0.0f);
        for (int i = 0; i < 32; i++)
            for (int p = 0; p < 32; p++) {
                H[i * 32 + p] = sinab(2 * i + 1, p + 1);
                if (i < 31 && p < 31) H[1024 + i * 32 + p] = sinab(2 * i + 2, p + 1);
            }
        // split index -> natural frequency index (a - 1), -1: none
        auto nat = [](int r) { return r < 32 ? 2 * r : (r < 63 ? 2 * (r - 32) + 1 : -1); };
        for (int r = 0; r < m; r++)
            for (int c = 0; c < m; c++)
                if (nat(r) >= 0 && nat(c) >= 0)
                    Ds[(size_t)r * m + c] = D[(size_t)nat(r) * m + nat(c)];
        S.swap(H);
        D.swap(Ds);
    }
    if (hipMalloc((void **)&M.dst_S, S.size() * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&M.dst_D, D.size() * sizeof(float)) != hipSuccess)
        return -1;
    if (hipMemcpy(M.dst_S, S.data(), S.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(M.dst_D, D.data(), D.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return -1;
    return 0;
}

// ------------------------------------------------------------ MG driver

template <typename T>
static void relax_launch(hipStream_t st, T *u, const T *rhs, int n, double length,
                         const int *skip) {
    double h = length / (n - 1);
    T h2 = (T)(h * h);
    int m = n - 2, half = (m + 1) / 2;
    dim3 g((half + 127) / 128, m);
    for (int c = 0; c < 2; c++)
        hipLaunchKernelGGL(k_mg_relax<T>, g, dim3(128), 0, st, u, rhs, n, h2, c, skip);
}

template <typename T>
static void coarse_launch(hipStream_t st, int lc, const T *rhs, T *u, int mode, int ncycle,
                          double length, const int *skip) {
    hipLaunchKernelGGL(k_mg_coarse<T>, dim3(1), dim3(1024), coarse_lds_elems(lc) * sizeof(T), st,
                       rhs, u, lc, mode, ncycle, length, skip);
}

// V(1,1) cycle at level j > lc on (Uj, Rj) (mglin.cpp:176-196)
// dst (fp32 only): solve the coarsest level exactly instead of by a V-cycle
template <typename T>
static void vcycle_launch(hipStream_t st, MgArrays<T> &A, int lc, int j, T *Uj, const T *Rj,
                          double length, const int *skip, const MgLevels *dst = nullptr) {
    int nf = lvl_size(j), nc = lvl_size(j - 1);
    relax_launch<T>(st, Uj, Rj, nf, length, skip);
    double h = length / (nf - 1);
    hipLaunchKernelGGL(k_mg_resid_rstrct<T>, dim3((nc + 127) / 128, nc), dim3(128), 0, st, Uj, Rj,
                       nf, (T)(1.0 / (h * h)), A.cres[j - 1], skip);
    if (j - 1 == lc) {
        if constexpr (std::is_same<T, float>::value) {
            if (dst) coarse_dst_launch(st, *dst, A.cres[j - 1], A.cv[j - 1], length, skip);
            else coarse_launch<T>(st, lc, A.cres[j - 1], A.cv[j - 1], 1, 1, length, skip);
        } else {
            coarse_launch<T>(st, lc, A.cres[j - 1], A.cv[j - 1], 1, 1, length, skip);
        }
    } else {
        long s = (long)nc * nc;
        hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((s + 255) / 256)), dim3(256), 0, st,
                           A.cv[j - 1], s, (T)0, skip);
        vcycle_launch<T>(st, A, lc, j - 1, A.cv[j - 1], A.cres[j - 1], length, skip, dst);
    }
    hipLaunchKernelGGL(k_mg_addint<T>, dim3((nf + 127) / 128, nf), dim3(128), 0, st, Uj, nf,
                       A.cv[j - 1], skip);
    relax_launch<T>(st, Uj, Rj, nf, length, skip);
}

int mg_coarse_level_for(long lds_bytes, int ng) {
    int lc = 0;
    for (int l = 1; l <= 5 && l < ng; l++)
        if (coarse_lds_elems(l) * (long)sizeof(double) <= lds_bytes) lc = l;
    return lc;
}

template <bool FOLD>
__global__ void __launch_bounds__(1024)
    k_tol_solve_dst(double *__restrict__ ug, const double *__restrict__ fg, int n, double h2i,
                    int lc, float h2, const float *__restrict__ Sg, const float *__restrict__ Dg,
                    double tol, int max_cycles, int *__restrict__ flags,
                    double *__restrict__ relres_host, int *__restrict__ host_flags);

void mg_set_coarse_attr(long lds_bytes) {
    hipFuncSetAttribute((const void *)k_mg_coarse<double>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    hipFuncSetAttribute((const void *)k_mg_coarse_dst, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)lds_bytes);
    hipFuncSetAttribute((const void *)k_tol_solve_dst<true>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    hipFuncSetAttribute((const void *)k_tol_solve_dst<false>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    hipFuncSetAttribute((const void *)k_mg_coarse<float>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
}

// FP64 full multigrid (mglin.cpp:21-55) on the device.  rhs = M.d.rhsF[ng-1]
// (finest) on entry; the solution is written to M.d.u[ng-1].
void mg_fixed_solve(hipStream_t st, MgLevels &M, int ncycle, double length) {
    MgArrays<double> &A = M.d;
    int lf = M.ng - 1, lc = M.lc;
    for (int l = lf - 1; l >= lc; l--) {
        int nc = lvl_size(l);
        hipLaunchKernelGGL(k_mg_rstrct<double>, dim3((nc + 127) / 128, nc), dim3(128), 0, st,
                           A.rhsF[l], nc, A.rhsF[l + 1]);
    }
    coarse_launch<double>(st, lc, A.rhsF[lc], A.u[lc], 0, ncycle, length, nullptr);
    for (int j = lc + 1; j <= lf; j++) {
        int nf = lvl_size(j);
        hipLaunchKernelGGL(k_mg_interp<double>, dim3((nf + 127) / 128, nf), dim3(128), 0, st,
                           A.u[j], nf, A.u[j - 1]);
        for (int c = 0; c < ncycle; c++)
            vcycle_launch<double>(st, A, lc, j, A.u[j], A.rhsF[j], length, nullptr);
    }
}

// ---------------------------------------------- tolerance (mixed precision)
// Iterative refinement: r = f - A u in FP64 (the reference's stencil and
// association), the correction equation A e = r solved approximately by one
// FP32 V(1,1) cycle from e = 0, u += e in FP64; repeated until
// ||r||_2 <= tol ||f||_2 over interior nodes.


// End of a residual pass: this block's sums of r^2 and f^2 go to
// partials[b].  The pass's total and verdict come from a later launch
// (k_tol_finish, or the next cycle's first kernel: TolVerdict): a
// last-block ticket inside the pass would need a device-scope fence in
// every block (an L2 write-back per block on a multi-XCD device), which
// cost more than the launch.
__device__ __forceinline__ void tol_partial(double rr, double ff, int b,
                                            double *__restrict__ partials) {
    rr = block_sum_d(rr);
    ff = block_sum_d(ff);
    if (threadIdx.x == 0) {
        partials[2 * b] = rr;
        partials[2 * b + 1] = ff;
    }
}

// Verdict of a residual pass of nb blocks (one workgroup): sums the partials
// in block order, sets flags[0] when ||r|| <= tol ||f||, counts an applied
// cycle in flags[1] when `applied`, and publishes both and the relative
// residual to the host-mapped copies, then (after a system-scope fence) the
// count of residual passes of this solve (flags[3]) with the convergence bit
// MG_PUB_CONV in host_flags[2], on which the host waits.  Skipped, as its
// pass, once the solve has converged.
__global__ void __launch_bounds__(TOL_THREADS)
    k_tol_finish(const double *__restrict__ partials, int nb, double tol,
                 int *__restrict__ flags, int applied, double *__restrict__ relres_host,
                 int *__restrict__ host_flags) {
    SKIP_IF_DONE(flags);
    double rr = 0.0, ff = 0.0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        rr += partials[2 * k];
        ff += partials[2 * k + 1];
    }
    rr = block_sum_d(rr);
    ff = block_sum_d(ff);
    if (threadIdx.x == 0) {
        const double rel = (ff > 0.0) ? sqrt(rr / ff) : sqrt(rr);
        if (applied) flags[1] += 1;
        if (rel <= tol) flags[0] = 1;
        *relres_host = rel;
        host_flags[0] = flags[0];
        host_flags[1] = flags[1];
        __threadfence_system();
        flags[3] += 1;
        // one word carries the pass count and the verdict, so the host that
        // acquires it also sees the applied-cycle count and the residual
        // written before the fence
        host_flags[2] = flags[3] | (flags[0] ? MG_PUB_CONV : 0);
    }
}

static void tol_finish_launch(hipStream_t st, MgLevels &M, int nb, double tol, bool applied) {
    hipLaunchKernelGGL(k_tol_finish, dim3(1), dim3(TOL_THREADS), 0, st,
                       (const double *)M.partials, nb, tol, M.flags, applied ? 1 : 0, M.relres,
                       M.host_flags);
}

// One residual pass of the refinement loop, fused: v = u_in + e (FP64
// accumulation of the FP32 correction; e null: v = u_in) written to u_out
// (null: not written), r = f - A v in FP32 for the next correction solve,
// block partial sums of r^2 and f^2 (k_tol_finish or the next cycle's
// first kernel forms the verdict).  u_in and u_out are different arrays
// (neighbours read u_in + e).
__global__ void __launch_bounds__(TOL_THREADS)
    k_tol_update(const double *__restrict__ u_in, const float *__restrict__ e,
                 double *__restrict__ u_out, const double *__restrict__ f, int n, double h2i,
                 float *__restrict__ r32, double *__restrict__ partials,
                 const int *__restrict__ flags) {
    SKIP_IF_DONE(flags);
    double rr = 0.0, ff = 0.0;
    // whole rows per block (no index division); threads stride along the
    // row, U nodes per thread at a time: the loads of all U are issued
    // before the first is used, so their latencies overlap (a thread still
    // visits its nodes in increasing j: the same sums)
    constexpr int U = 4;
    for (int i = (int)blockIdx.x; i < n; i += (int)gridDim.x) {
        const bool irow = i > 0 && i < n - 1;
        for (int j0 = (int)threadIdx.x; j0 < n; j0 += U * (int)blockDim.x) {
            double c[U], un[U], us[U], ue[U], uw[U], fv[U];
#pragma unroll
            for (int t = 0; t < U; t++) {
                const int j = j0 + t * (int)blockDim.x;
                const long k = (long)i * n + j;
                c[t] = un[t] = us[t] = ue[t] = uw[t] = fv[t] = 0.0;
                if (j < n) {
                    c[t] = u_in[k];
                    if (irow && j > 0 && j < n - 1) {
                        un[t] = u_in[k + n];
                        us[t] = u_in[k - n];
                        ue[t] = u_in[k + 1];
                        uw[t] = u_in[k - 1];
                        if (e) {
                            c[t] += (double)e[k];
                            un[t] += (double)e[k + n];
                            us[t] += (double)e[k - n];
                            ue[t] += (double)e[k + 1];
                            uw[t] += (double)e[k - 1];
                        }
                        fv[t] = f[k];
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < U; t++) {
                const int j = j0 + t * (int)blockDim.x;
                if (j >= n) break;
                const long k = (long)i * n + j;
                float out = 0.0f;
                if (irow && j > 0 && j < n - 1) {
                    // the stencil of resid_node: north, south, east, west
                    const double r = -h2i * (un[t] + us[t] + ue[t] + uw[t] - 4.0 * c[t]) + fv[t];
                    rr += r * r;
                    out = (float)r;
                    ff += fv[t] * fv[t];
                }
                if (u_out) u_out[k] = c[t];
                r32[k] = out;
            }
        }
    }
    tol_partial(rr, ff, (int)blockIdx.x, partials);
}

// The top level's post leg fused with the refinement's residual pass: the
// halo takes one more ring (2 nu + 1) so the correction e is final on the
// tile and its first ring; then, as k_tol_update, v = u_in + e in FP64 is
// written to u_out, r = f - A v (the same expression and association) to
// r_out as FP32, and the tile's sums of r^2 and f^2 go to the partials.  The
// iterate and the residual are those of k_mg_post + k_tol_update bit for
// bit; only the block split of the norm's sum differs.  r_out must not be
// the rhs this leg reads (the halo reads neighbouring tiles' nodes).
template <int NU, int T>
__global__ void __launch_bounds__(256)
    k_mg_post_update(const float *__restrict__ us, const float *__restrict__ rhs,
                     const float *__restrict__ cv, int n, float h2,
                     const double *__restrict__ u_in, double *__restrict__ u_out,
                     const double *__restrict__ f, double h2i, float *__restrict__ r_out,
                     double *__restrict__ partials, const int *__restrict__ flags) {
    SKIP_IF_DONE(flags);
    constexpr int H = 2 * NU + 1, W = T + 2 * H;
    __shared__ float su[W * W], sf[W * W];
    const int x0 = blockIdx.x * T, y0 = blockIdx.y * T, nc = (n + 1) / 2;
    // one node per thread (16 x 16 tiles): its FP64 operands are read now, so
    // their latency overlaps the sweeps
    constexpr bool ONE = T * T == 256;
    double pu = 0.0, pn = 0.0, ps = 0.0, pe = 0.0, pw = 0.0, pf = 0.0;
    if (ONE) {
        const int i = y0 + (int)threadIdx.x / T, j = x0 + (int)threadIdx.x % T;
        if (i < n && j < n) {
            const long g = (long)i * n + j;
            pu = u_in[g];
            if (i > 0 && j > 0 && i < n - 1 && j < n - 1) {
                pn = u_in[g + n];
                ps = u_in[g - n];
                pe = u_in[g + 1];
                pw = u_in[g - 1];
                pf = f[g];
            }
        }
    }
    for (int k = threadIdx.x; k < W * W; k += blockDim.x) {
        const int ly = k / W, lx = k % W;
        const int i = y0 - H + ly, j = x0 - H + lx;
        const bool in = i >= 0 && j >= 0 && i < n && j < n;
        su[k] = in ? us[(long)i * n + j] + interp_node(cv, nc, i, j) : 0.0f;
        sf[k] = in ? rhs[(long)i * n + j] : 0.0f;
    }
    __syncthreads();
    tile_smooth<NU, T>(su, sf, W, H, x0, y0, 1, n, h2);  // e valid on the tile + 1 ring
    double rr = 0.0, ff = 0.0;
    for (int k = threadIdx.x; k < T * T; k += blockDim.x) {
        const int ly = k / T, lx = k % T, i = y0 + ly, j = x0 + lx;
        if (i >= n || j >= n) continue;
        const long g = (long)i * n + j;
        const int c = (ly + H) * W + lx + H;
        double v = ONE ? pu : u_in[g];
        float out = 0.0f;
        if (i > 0 && j > 0 && i < n - 1 && j < n - 1) {
            const double fg = ONE ? pf : f[g];
            v += (double)su[c];
            const double un = (ONE ? pn : u_in[g + n]) + (double)su[c + W];
            const double us_ = (ONE ? ps : u_in[g - n]) + (double)su[c - W];
            const double ue = (ONE ? pe : u_in[g + 1]) + (double)su[c + 1];
            const double uw = (ONE ? pw : u_in[g - 1]) + (double)su[c - 1];
            const double r = -h2i * (un + us_ + ue + uw - 4.0 * v) + fg;
            rr += r * r;
            out = (float)r;
            ff += fg * fg;
        }
        u_out[g] = v;
        r_out[g] = out;
    }
    tol_partial(rr, ff, (int)(blockIdx.y * gridDim.x + blockIdx.x), partials);
}

__global__ void k_tol_reset(int *flags) {
    flags[0] = 0;
    flags[1] = 0;
    flags[2] = 0;
    flags[3] = 0;
}

// one refinement cycle above the coarsest level: the V-cycle on the
// correction equation A e = r_in from e = 0, its top post leg fused with
// the FP64 update u_out = u_in + e and the new residual r_out.  vd: the
// verdict of the previous residual pass, formed by the first kernel; fin:
// this pass's verdict by k_tol_finish (else by the next cycle's first kernel)
static void vcycle_top(hipStream_t st, MgLevels &M, int j, const float *r_in, const double *u_in,
                       double *u_out, const double *f, double h2i, float *r_out, double tol,
                       double length, const TolVerdict &vd, bool fin) {
    MgArrays<float> &A = M.f;
    const int *skip = M.flags;
    const int nf = lvl_size(j);
    const double h = length / (nf - 1);
    const int T = mg_tile(nf), tiles = (nf + T - 1) / T;
    const float h2 = (float)(h * h), h2f = (float)(1.0 / (h * h));
    const dim3 g(tiles, tiles), b(256);
    const float *z = nullptr;
#define PRE(NU, TT)                                                                             \
    hipLaunchKernelGGL((k_mg_pre<NU, TT>), g, b, 0, st, z, r_in, nf, h2, h2f, A.us[j],          \
                       A.cres[j - 1], skip, vd)
    MG_DISPATCH(M.nu, T, PRE);
#undef PRE
    if (j - 1 == M.lc)
        coarse_dst_launch(st, M, A.cres[j - 1], A.cv[j - 1], length, skip);
    else
        vcycle_fused(st, M, j - 1, A.cv[j - 1], A.cres[j - 1], length, skip);
#define POST_UPDATE(NU, TT)                                                                       \
    hipLaunchKernelGGL((k_mg_post_update<NU, TT>), g, b, 0, st, A.us[j], r_in, A.cv[j - 1], nf, h2, \
                       u_in, u_out, f, h2i, r_out, M.partials, M.flags)
    MG_DISPATCH(M.nu, T, POST_UPDATE);
#undef POST_UPDATE
    if (fin) tol_finish_launch(st, M, tiles * tiles, tol, true);
}

// Waits for the queued cycles' verdict: the device has published (host_flags
// in coherent host memory, the pass count last) either convergence or all
// `passes` residual passes of the solve.  The host reads the count as soon
// as the last pass's final block writes it, so the next launches queue up
// behind the work still running (phi and E) instead of after an idle GPU.
// After 2 ms of spinning (or if the device never publishes, e.g. on a
// fault) it blocks on the stream's event, which also reports errors.  The
// spin yields the core between reads: with many runs stepped from their own
// threads (ensembles, serving) the waiting threads leave the cores to the
// threads that are launching.
static hipError_t wait_flags(MgLevels &M, int passes) {
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        const int w = __atomic_load_n(&M.host_flags[2], __ATOMIC_ACQUIRE);
        const int p = w & MG_PUB_COUNT;
        if (p >= passes || (p > 0 && (w & MG_PUB_CONV))) return hipSuccess;
        if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(2000)) break;
        std::this_thread::yield();
    }
    return hipEventSynchronize(M.done);
}

static hipError_t wait_published(hipStream_t st, MgLevels &M, int passes) {
    hipError_t e = hipEventRecord(M.done, st);
    if (e != hipSuccess) return e;
    return wait_flags(M, passes);
}

int mg_tol_verdict(MgLevels &M, double *relres_out) {
    const int passes = M.pending;
    if (passes <= 0) return -1;
    M.pending = 0;
    if (wait_flags(M, passes) != hipSuccess) return -1;
    const bool conv = (__atomic_load_n(&M.host_flags[2], __ATOMIC_ACQUIRE) & MG_PUB_CONV) != 0;
    const int used = M.host_flags[1];
    M.pred = used + (conv ? 0 : 1);
    if (relres_out) *relres_out = *M.relres;
    return used;
}

// The whole tolerance solve of a grid whose finest level is the DST level
// (65^2 and below: C1, the small runs of ensembles and serving), in one
// workgroup of 1024 threads, everything resident in LDS: the FP64 iterate
// and right-hand side, the sine table, and the FP32 residual / correction
// matrix R with its product scratch W (the layout of dst_solve's X and Y:
// node (i, j) of the interior at R[(i-1) ld + (j-1)]).  The initial residual,
// then per cycle the exact DST solve of R in place, the update u += e and
// the residual of the updated iterate into R -- the arithmetic of
// k_mg_coarse_dst and k_tol_update node for node, each thread summing its
// nodes in ascending order as there, so iterates, residuals and cycle
// counts are the launched cycles' -- until ||f - Au|| <= tol ||f|| or
// max_cycles cycles.  The final iterate is written to u; the verdict is
// published as k_tol_finish publishes it, once, at the end.  One launch per
// solve and no host wait: these grids are bound by latency.
#define SMALL_NPT 5  // nodes per thread: n^2 <= 5 x 1024 for n <= 65

__host__ __device__ inline long tol_small_lds_bytes(int m) {
    const long n = m + 1;
    return (2 * n * n + 16) * (long)sizeof(double) + dst_lds_floats(m) * (long)sizeof(float);
}

__device__ __forceinline__ double block_sum_1024(double v, double *ws) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) ws[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += ws[k];  // fixed order
    __syncthreads();
    return t;
}

template <bool FOLD>
__global__ void __launch_bounds__(1024)
    k_tol_solve_dst(double *__restrict__ ug, const double *__restrict__ fg, int n, double h2i,
                    int lc, float h2, const float *__restrict__ Sg, const float *__restrict__ Dg,
                    double tol, int max_cycles, int *__restrict__ flags,
                    double *__restrict__ relres_host, int *__restrict__ host_flags) {
    const int m = 1 << (lc + 1), ld = m + 4, nn = n * n;
    // no static LDS: the attribute grants the dynamic region the whole
    // per-workgroup limit
    extern __shared__ double lds_raw[];
    double *U = lds_raw, *F = U + nn, *ws = F + nn;
    float *S = reinterpret_cast<float *>(ws + 16), *R = S + m * ld, *W = R + m * ld;
    // this thread's nodes: k = threadIdx.x + r * 1024; o[r] = its R slot, or
    // -1 off the interior (and past the grid)
    int o[SMALL_NPT];
    double ff = 0.0;
#pragma unroll
    for (int r = 0; r < SMALL_NPT; r++) {
        const int k = (int)threadIdx.x + r * 1024;
        o[r] = -1;
        if (k < nn) {
            const int i = k / n, j = k - (k / n) * n;
            if (i > 0 && j > 0 && i < n - 1 && j < n - 1) o[r] = (i - 1) * ld + (j - 1);
            const double fk = fg[k];
            U[k] = ug[k];
            F[k] = fk;
            if (o[r] >= 0) ff += fk * fk;
        }
    }
    float dk[4];  // the eigenvalue scale of this thread's R entries
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int k = (int)threadIdx.x + r * 1024;
        dk[r] = k < m * m ? Dg[k] * h2 : 0.0f;
    }
    for (int k = (int)threadIdx.x; k < m * ld; k += 1024) R[k] = 0.0f;
    if (FOLD)  // m = 64: the folded tables (dstf_solve)
        for (int k = (int)threadIdx.x; k < 2048; k += 1024)
            S[(k >> 10) * 32 * DSTF_LDH + ((k >> 5) & 31) * DSTF_LDH + (k & 31)] = Sg[k];
    else
        for (int k = (int)threadIdx.x; k < m * m; k += 1024)
            S[(k >> (lc + 1)) * ld + (k & (m - 1))] = Sg[k];
    ff = block_sum_1024(ff, ws);  // its barriers also publish U, F, S and R
    // the residual of U into R, its sum of squares
    auto residual = [&]() -> double {
        double rr = 0.0;
#pragma unroll
        for (int r = 0; r < SMALL_NPT; r++) {
            const int k = (int)threadIdx.x + r * 1024;
            if (o[r] >= 0) {
                // the stencil of resid_node: north, south, east, west
                const double res =
                    -h2i * (U[k + n] + U[k - n] + U[k + 1] + U[k - 1] - 4.0 * U[k]) + F[k];
                rr += res * res;
                R[o[r]] = (float)res;
            }
        }
        return block_sum_1024(rr, ws);
    };
    double rr = residual();
    double rel = (ff > 0.0) ? sqrt(rr / ff) : sqrt(rr);
    int q = 0;
    while (!(rel <= tol) && q < max_cycles) {  // uniform: every thread holds the same sums
        if (FOLD) {
            dstf_solve(R, W, S, dk, ld);  // R = e on the interior
        } else {
            dst_gemm(S, R, W, m, ld);  // W = S r
            dst_gemm(W, S, R, m, ld);  // R = S r S
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int k = (int)threadIdx.x + r * 1024;
                if (k < m * m) {
                    const int p = k >> (lc + 1), c = k & (m - 1);
                    R[p * ld + c] = R[p * ld + c] * dk[r];
                }
            }
            __syncthreads();
            dst_gemm(S, R, W, m, ld);
            dst_gemm(W, S, R, m, ld);  // R = e on the interior
        }
#pragma unroll
        for (int r = 0; r < SMALL_NPT; r++) {
            const int k = (int)threadIdx.x + r * 1024;
            if (o[r] >= 0) U[k] = U[k] + (double)R[o[r]];
        }
        __syncthreads();
        rr = residual();
        rel = (ff > 0.0) ? sqrt(rr / ff) : sqrt(rr);
        q++;
    }
#pragma unroll
    for (int r = 0; r < SMALL_NPT; r++) {
        const int k = (int)threadIdx.x + r * 1024;
        if (o[r] >= 0) ug[k] = U[k];
    }
    if (threadIdx.x == 0) {
        const int conv = rel <= tol;
        flags[0] = conv;
        flags[1] = q;
        flags[3] = 1 + q;
        *relres_host = rel;
        host_flags[0] = conv;
        host_flags[1] = q;
        __threadfence_system();
        host_flags[2] = (1 + q) | (conv ? MG_PUB_CONV : 0);
    }
}

int mg_tol_solve(hipStream_t st, MgLevels &M, double length, double tol, int max_iter,
                 double *relres_out, bool flags_reset,
                 const std::function<void(const double *, const double *)> &after_batch) {
    int lf = M.ng - 1, n = lvl_size(lf);
    double h = length / (n - 1), h2i = 1.0 / (h * h);
    double *u = M.d.u[lf];
    const double *f = M.d.rhsF[lf];
    float *r32 = M.f.rhsF[lf], *e32 = M.f.u[lf];
    const int *skip = M.flags;
    // cycle c reads the iterate from buf[c % 2] and writes buf[(c + 1) % 2];
    // above the coarsest level the residual alternates between rb[0], rb[1]
    double *buf[2] = {u, M.ualt};
    float *rb[2] = {r32, M.r32alt};
    int queued = 0;
    // the top post leg fused with the residual pass saves a launch per cycle
    // where the levels are latency-bound; at 1025^2 the wider halo's extra
    // reads cost more than the launch (measured), so the passes stay apart
    const bool fuse_top = lf > M.lc && n <= 513;
    // residual-pass blocks of a cycle (the partials its verdict sums)
    const int top_tiles = (n + mg_tile(n) - 1) / mg_tile(n);
    const int cyc_nb = (lf > M.lc && fuse_top) ? top_tiles * top_tiles : M.nparts;
    // vfirst: the verdict of the residual pass before this cycle (the
    // solve's initial pass, or the previous cycle's) is formed by this
    // cycle's first kernel, else k_tol_finish formed it after the pass; fin:
    // this cycle's own pass gets k_tol_finish (the last cycle of a batch, on
    // which the host waits)
    auto cycle = [&](bool vfirst, bool fin) {
        const TolVerdict vd = vfirst ? TolVerdict{M.partials, queued > 0 ? cyc_nb : M.nparts, tol,
                                                  queued > 0 ? 1 : 0, M.flags, M.relres,
                                                  M.host_flags}
                                     : TolVerdict{};
        if (lf > M.lc && !fuse_top) {
            vcycle_fused(st, M, lf, e32, r32, length, skip, vd);
            hipLaunchKernelGGL(k_tol_update, dim3(M.nparts), dim3(TOL_THREADS), 0, st,
                               (const double *)buf[queued & 1], (const float *)e32,
                               buf[(queued + 1) & 1], f, n, h2i, r32, M.partials, M.flags);
            if (fin) tol_finish_launch(st, M, M.nparts, tol, true);
        } else if (lf == M.lc) {
            coarse_dst_launch(st, M, r32, e32, length, skip);
            hipLaunchKernelGGL(k_tol_update, dim3(M.nparts), dim3(TOL_THREADS), 0, st,
                               (const double *)buf[queued & 1], (const float *)e32,
                               buf[(queued + 1) & 1], f, n, h2i, r32, M.partials, M.flags);
            tol_finish_launch(st, M, M.nparts, tol, true);
        } else {
            vcycle_top(st, M, lf, rb[queued & 1], buf[queued & 1], buf[(queued + 1) & 1], f, h2i,
                       rb[(queued + 1) & 1], tol, length, vd, fin);
        }
    };
    // the coarsest-level branch keeps k_tol_finish after every pass, and so
    // does a solve allowed no cycles
    const bool vfuse = lf > M.lc && max_iter > 0;
    // nothing on the device writes the host copies now (the last solve's
    // remaining cycles were skipped): clear them for this solve's verdict
    if (M.pending && mg_tol_verdict(M, nullptr) < 0) return -1;
    __atomic_store_n(&M.host_flags[2], 0, __ATOMIC_RELEASE);
    M.host_flags[0] = M.host_flags[1] = 0;
    if (lf == M.lc && n * n <= SMALL_NPT * 1024 && tol_small_lds_bytes(n - 1) <= M.lds) {
        // the whole solve in one workgroup (k_tol_solve_dst)
        const int m = n - 1;
        const double hc = length / (double)m;
        hipLaunchKernelGGL((m == 64 ? k_tol_solve_dst<true> : k_tol_solve_dst<false>), dim3(1),
                           dim3(1024), (size_t)tol_small_lds_bytes(m),
                           st, u, f, n, h2i, M.lc, (float)(hc * hc), M.dst_S, M.dst_D, tol,
                           max_iter, M.flags, M.relres, M.host_flags);
        if (after_batch) after_batch(u, u);
        if (hipEventRecord(M.done, st) != hipSuccess) return -1;
        M.pending = 1 + max_iter;
        return MG_PENDING;
    }
    if (!flags_reset) hipLaunchKernelGGL(k_tol_reset, dim3(1), dim3(1), 0, st, M.flags);
    hipLaunchKernelGGL(k_tol_update, dim3(M.nparts), dim3(TOL_THREADS), 0, st, (const double *)u,
                       (const float *)nullptr, (double *)nullptr, f, n, h2i, r32, M.partials,
                       M.flags);
    // its verdict: by the first cycle's first kernel where cycles form them
    if (!vfuse) tol_finish_launch(st, M, M.nparts, tol, false);
    // k cycles from `queued` on: launched one kernel at a time, or replayed
    // from the hipGraph of that batch (k, parity of queued, iterate buffer),
    // captured on the batch's first use; a capture records the launches of
    // this thread on st without running them
    auto queue_cycles = [&](int k) -> bool {
        // the batch right after the initial pass forms that pass's verdict too
        const bool first = queued == 0;
        if (!M.use_graphs) {
            for (int c = 0; c < k; c++, queued++)
                cycle(vfuse && (c > 0 || first), !vfuse || c == k - 1);
            return true;
        }
        const int par = queued & 1;
        hipGraphExec_t ex = nullptr;
        for (const MgLevels::CycleGraph &g : M.graphs)
            if (g.k == k && g.parity == par && g.first == first && g.u0 == buf[0]) ex = g.exec;
        if (!ex) {
            hipGraph_t gr = nullptr;
            if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return false;
            const int q0 = queued;
            for (int c = 0; c < k; c++, queued++)
                cycle(vfuse && (c > 0 || first), !vfuse || c == k - 1);
            queued = q0;
            if (hipStreamEndCapture(st, &gr) != hipSuccess) return false;
            const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
            hipGraphDestroy(gr);
            if (e != hipSuccess) return false;
            M.graphs.push_back({k, par, first, buf[0], ex});
        }
        if (hipGraphLaunch(ex, st) != hipSuccess) return false;
        queued += k;
        return true;
    };
    // queue the predicted number of cycles without waiting (converged cycles
    // skip their work on the device), then confirm on the host
    int batch = M.pred < 1 ? 1 : M.pred;
    bool conv = false;
    for (;;) {
        if (!queue_cycles(std::min(batch, max_iter - queued))) return -1;
        if (after_batch) after_batch(buf[0], buf[1]);
        // the initial residual pass plus one per queued cycle
        if (wait_published(st, M, 1 + queued) != hipSuccess) return -1;
        // the published word (acquired) orders the reads of host_flags[1]
        // and the residual after it
        conv = (__atomic_load_n(&M.host_flags[2], __ATOMIC_ACQUIRE) & MG_PUB_CONV) != 0;
        if (conv || queued >= max_iter) break;
        batch = 1;
    }
    int used = M.host_flags[1];
    // the iterate of an odd number of applied cycles is in ualt: it becomes u
    if (used & 1) std::swap(M.d.u[lf], M.ualt);
    M.pred = used + (conv ? 0 : 1);
    if (relres_out) *relres_out = *M.relres;
    return used;
}

// ------------------------------------------- FP64 reference solve (report)
// The accuracy report of the split-precision solver: the same folded
// right-hand side solved by the FP64 multigrid alone (the full multigrid of
// the fixed mode, then FP64 V(1,1) cycles, mglin.cpp:176-196) until
// ||f - Au||_2 <= tol ||f||_2 on the interior.

// per-block sums of the FP64 residual r = f - A u and of f, squared
__global__ void __launch_bounds__(TOL_THREADS)
    k_resid_sums(const double *__restrict__ u, const double *__restrict__ f, int n, double h2i,
                 double *__restrict__ partials) {
    double rr = 0.0, ff = 0.0;
    const long tot = (long)n * n;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < tot;
         k += (long)gridDim.x * blockDim.x) {
        const int i = (int)(k / n), j = (int)(k % n);
        if (i > 0 && j > 0 && i < n - 1 && j < n - 1) {
            const double r = -h2i * (u[k + n] + u[k - n] + u[k + 1] + u[k - 1] - 4.0 * u[k]) + f[k];
            rr += r * r;
            ff += f[k] * f[k];
        }
    }
    rr = block_sum_d(rr);
    ff = block_sum_d(ff);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = rr;
        partials[2 * blockIdx.x + 1] = ff;
    }
}

// per-block sums of (a - b)^2 and b^2 over n entries, for three pairs
__global__ void __launch_bounds__(TOL_THREADS)
    k_diff_sums3(const double *__restrict__ a0, const double *__restrict__ b0,
                 const double *__restrict__ a1, const double *__restrict__ b1,
                 const double *__restrict__ a2, const double *__restrict__ b2, long n,
                 double *__restrict__ partials) {
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long)gridDim.x * blockDim.x) {
        const double d0 = a0[k] - b0[k], d1 = a1[k] - b1[k], d2 = a2[k] - b2[k];
        s[0] += d0 * d0;
        s[1] += b0[k] * b0[k];
        s[2] += d1 * d1;
        s[3] += b1[k] * b1[k];
        s[4] += d2 * d2;
        s[5] += b2[k] * b2[k];
    }
    for (int m = 0; m < 6; m++) {
        const double t = block_sum_d(s[m]);
        if (threadIdx.x == 0) partials[6 * blockIdx.x + m] = t;
    }
}

// sums of the blocks' partials in block order (deterministic)
static int host_block_sums(hipStream_t st, const double *dev, double *host, int nb, int w,
                           double *out) {
    if (hipMemcpyAsync(host, dev, sizeof(double) * nb * w, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return -1;
    for (int m = 0; m < w; m++) out[m] = 0.0;
    for (int b = 0; b < nb; b++)
        for (int m = 0; m < w; m++) out[m] += host[w * b + m];
    return 0;
}

int mg_fp64_reference(hipStream_t st, MgLevels &R, double length, double tol, int max_cycles,
                      double *partials, double *host, int nb, double *relres) {
    const int lf = R.ng - 1, n = lvl_size(lf);
    // the V-cycle at lf restricts into cres[lf-1]: levels lc..lf-1 must exist
    if (R.lc >= lf || !R.d.cres[lf - 1] || !R.d.cv[lf - 1]) return -1;
    const double h = length / (n - 1), h2i = 1.0 / (h * h);
    double *u = R.d.u[lf];
    const double *f = R.d.rhsF[lf];
    mg_fixed_solve(st, R, 2, length);
    int c = 0;
    for (;;) {
        hipLaunchKernelGGL(k_resid_sums, dim3(nb), dim3(TOL_THREADS), 0, st, (const double *)u, f,
                           n, h2i, partials);
        double sums[2];
        if (host_block_sums(st, partials, host, nb, 2, sums) != 0) return -1;
        *relres = sums[1] > 0.0 ? sqrt(sums[0] / sums[1]) : sqrt(sums[0]);
        if (*relres <= tol || c >= max_cycles) break;
        vcycle_launch<double>(st, R.d, R.lc, lf, u, f, length, nullptr);
        c++;
    }
    return c;
}

// non-finite values among n entries of a, b, c (counted into *count)
__global__ void k_count_nonfinite(const double *__restrict__ a, long na, const double *__restrict__ b,
                                  const double *__restrict__ c, long nbc, unsigned *__restrict__ count) {
    unsigned bad = 0;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < na; k += (long)gridDim.x * blockDim.x)
        bad += !isfinite(a[k]);
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < nbc; k += (long)gridDim.x * blockDim.x)
        bad += !isfinite(b[k]) + !isfinite(c[k]);
    if (bad) atomicAdd(count, bad);
}

void launch_count_nonfinite(hipStream_t st, const double *a, long na, const double *b,
                            const double *c, long nbc, unsigned *count) {
    hipLaunchKernelGGL(k_count_nonfinite, dim3(256), dim3(256), 0, st, a, na, b, c, nbc, count);
}

int field_diff_sums(hipStream_t st, const double *a0, const double *b0, const double *a1,
                    const double *b1, const double *a2, const double *b2, long n,
                    double *partials, double *host, int nb, double *out6) {
    hipLaunchKernelGGL(k_diff_sums3, dim3(nb), dim3(TOL_THREADS), 0, st, a0, b0, a1, b1, a2, b2, n,
                       partials);
    return host_block_sums(st, partials, host, nb, 6, out6);
}

}  // namespace pic
