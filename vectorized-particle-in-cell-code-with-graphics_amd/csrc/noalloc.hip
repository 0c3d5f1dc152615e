// noalloc.hip -- pic_mglin_noalloc: the FP64 full multigrid of pic_mglin for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops).  Every multigrid level lives in one file-scope device
// array, `scratch`, sized for grids up to NOALLOC_MAX_N nodes per side; the
// caller's device buffer is the finest level's solution.  No call allocates or
// frees anything: the array is part of the code object's device image, held
// by every GPU the library runs on until the process ends.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pic_host.h"
#include "pic_internal.h"

namespace {

constexpr int NOALLOC_MAX_N = 1025;

constexpr int lvl_n(int l) { return (1 << (l + 1)) + 1; }
// arrays start on 256-byte boundaries (32 doubles)
constexpr size_t pad32(size_t n) { return (n + 31) & ~(size_t)31; }
// the finest level's right-hand side plus, for every coarser level down to
// 3 x 3, its right-hand side, full-multigrid iterate and V-cycle arrays
constexpr size_t scratch_doubles(int maxn) {
    int lf = 0;
    while (lvl_n(lf) < maxn) lf++;
    size_t t = pad32((size_t)lvl_n(lf) * lvl_n(lf));
    for (int l = 0; l < lf; l++) t += 4 * pad32((size_t)lvl_n(l) * lvl_n(l));
    return t;
}
constexpr size_t SCRATCH_DOUBLES = scratch_doubles(NOALLOC_MAX_N);  // 2 457 888 (19.7 MB)
static_assert(SCRATCH_DOUBLES * sizeof(double) < (20u << 20), "scratch stays modest");

}  // namespace

// the multigrid levels of pic_mglin_noalloc (one copy per GPU)
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_scratch[MAX_DEVICES] = {};  // device address of `scratch`, per device
long g_lds[MAX_DEVICES] = {};

long device_lds_limit(int dev) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        v <= 0)
        v = 65536;
    return v;
}

// set-up once per device, with that device current: the address of this
// device's copy of `scratch` (the host-side name is not a device address)
// and the coarsest level's LDS attribute
int scratch_for_current_device(double **out, long *lds) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
        return pic_host::set_global_error(PIC_ERR_DEVICE, "no current HIP device");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return pic_host::set_global_error(
                PIC_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_lds[dev] = device_lds_limit(dev);
        pic::mg_set_coarse_attr(g_lds[dev]);
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    *lds = g_lds[dev];
    return PIC_OK;
}

}  // namespace

extern "C" int pic_mglin_noalloc(double *u, int32_t n, int32_t ncycle, double length,
                                 void *stream) {
    if (!u || ncycle < 1 || !(length > 0.0))
        return pic_host::set_global_error(PIC_ERR_ARG, "bad argument");
    const int m = n - 1;
    if (n < 3 || (m & (m - 1)) != 0)
        return pic_host::set_global_error(PIC_ERR_GRID, "n-1 must be a power of 2 in mglin.");
    if (n > NOALLOC_MAX_N)
        return pic_host::set_global_error(
            PIC_ERR_CAPACITY, "pic_mglin_noalloc: n = " + std::to_string(n) +
                                  " exceeds the scratch array's largest grid, n <= 1025");
    double *base = nullptr;
    long lds = 0;
    int rc = scratch_for_current_device(&base, &lds);
    if (rc != PIC_OK) return rc;
    pic::MgLevels R;
    int ng = 0;
    for (int nn = n; nn >>= 1;) ng++;
    R.ng = ng;
    R.lc = pic::mg_coarse_level_for(lds, ng);
    const int lf = ng - 1;
    size_t off = 0;
    auto take = [&](int l) {
        double *p = base + off;
        off += pad32((size_t)lvl_n(l) * lvl_n(l));
        return p;
    };
    R.d.rhsF[lf] = take(lf);
    R.d.u[lf] = u;  // the solution goes straight to the caller's buffer
    for (int l = R.lc; l < lf; l++) {
        R.d.rhsF[l] = take(l);
        R.d.u[l] = take(l);
        R.d.cres[l] = take(l);
        R.d.cv[l] = take(l);
    }
    if (off > SCRATCH_DOUBLES)  // cannot happen for n <= NOALLOC_MAX_N
        return pic_host::set_global_error(PIC_ERR_CAPACITY, "pic_mglin_noalloc: scratch layout");
    hipStream_t st = static_cast<hipStream_t>(stream);
    // u holds the right-hand side on entry (mglin.cpp:21-55 solves in place)
    hipError_t e = hipMemcpyAsync(R.d.rhsF[lf], u, sizeof(double) * n * n,
                                  hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess)
        return pic_host::set_global_error(PIC_ERR_DEVICE, hipGetErrorString(e));
    pic::mg_fixed_solve(st, R, ncycle, length);
    e = hipGetLastError();
    if (e != hipSuccess)
        return pic_host::set_global_error(PIC_ERR_DEVICE, std::string("launch: ") + hipGetErrorString(e));
    return PIC_OK;
}
