// mcc.h -- Monte-Carlo collisions: cross sections and per-particle collision
// kernels (device), shared with the host for the null-collision bound.
//
// Cross sections are in units of 1e-20 m^2 as the reference uses them
// (e_collisions.cpp:533,551 and ion_collisions.cpp:297 multiply by 1e-20).
// Energies in eV.
//
// Two rule sets:
//  REF   -- the reference's decision rules (e_collisions.cpp:480-597): one
//           uniform R, elastic iff P_el > 2R, ionisation iff P_iz > 0 (which
//           in the reference deletes the parent: its excess-energy formula
//           mixes joules and eV and yields NaN, e_collisions.cpp:377-381);
//           scattering angle and azimuth from one shared uniform
//           (collisions.cpp:118-133); neutral velocity from a 10000-entry
//           table (collisions.cpp:66-89, 135-139).
//  PHYS  -- null-collision Monte Carlo (Vahedi & Surendra, Comput. Phys.
//           Commun. 87 (1995) 179): a candidate fraction 1-exp(-nu' dt) of
//           particles, channel chosen by nu_k(E)/nu'; the neutral partner is
//           sampled from a Maxwellian; electrons: elastic, excitation,
//           ionisation; ions: isotropic elastic and charge exchange.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "pic_dev.h"

#define HD __host__ __device__ inline

namespace mcc {

enum Gas { GAS_AR = 0, GAS_H = 1 };

// Math of the cross-section fits.  LibMath: the C library (host, and the
// reference rule set on the device).  HwMath: the device's hardware
// transcendentals (v_log / v_exp in base 2, v_sqrt, v_rcp; about 1 ulp each)
// for the physical rule set, whose results are checked against an FP64
// restatement to 2e-4 relative, not bit for bit.
struct LibMath {
    template <typename T> static HD T powr(T x, T y) { return pow(x, y); }
    template <typename T> static HD T logn(T x) { return log(x); }
    template <typename T> static HD T sqroot(T x) { return sqrt(x); }
    template <typename T> static HD T quot(T a, T b) { return a / b; }
};
struct HwMath {
    static __device__ __forceinline__ float powr(float x, float y) {
        return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
    }
    static __device__ __forceinline__ float logn(float x) {
        return __builtin_amdgcn_logf(x) * 0.693147181f;
    }
    static __device__ __forceinline__ float sqroot(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float quot(float a, float b) {
        return a * __builtin_amdgcn_rcpf(b);
    }
};
// cos(2 pi u), sin(2 pi u): v_cos / v_sin take the angle in revolutions
__device__ __forceinline__ float cos_rev(float u) { return __builtin_amdgcn_cosf(u); }
__device__ __forceinline__ float sin_rev(float u) { return __builtin_amdgcn_sinf(u); }

// -------------------------------------------------------- cross sections
// e-Ar elastic (e_collisions.cpp:205-224).  REF keeps the reference's exact
// branch tests (E exactly 0.35 or 15.0 falls in no branch: sigma 0).
template <typename T>
HD T ar_el_ref(T E) {
    if (E < (T)0.35)
        return (T)28019.0 * E * E * E * E * E * E - (T)32798.0 * E * E * E * E * E +
               (T)15048.0 * E * E * E * E - (T)34312.0 * E * E * E + (T)4084.0 * E * E -
               (T)249.0 * E + (T)7.576;
    if (E < (T)15.0 && E > (T)0.35)
        return -(T)0.013 * E * E * E + (T)0.232 * E * E + (T)1.066 * E + (T)0.085;
    if (E > (T)15.0) return (T)126.2 * pow(E, (T)-0.69);
    return (T)0.0;
}
// PHYS: same fit, breakpoints closed, negative values (0.061-0.349 eV) clamped
template <typename T, typename M = LibMath>
HD T ar_el_phys(T E) {
    T s;
    if (E < (T)0.35)
        s = (T)28019.0 * E * E * E * E * E * E - (T)32798.0 * E * E * E * E * E +
            (T)15048.0 * E * E * E * E - (T)34312.0 * E * E * E + (T)4084.0 * E * E -
            (T)249.0 * E + (T)7.576;
    else if (E < (T)15.0)
        s = -(T)0.013 * E * E * E + (T)0.232 * E * E + (T)1.066 * E + (T)0.085;
    else
        s = (T)126.2 * M::powr(E, (T)-0.69);
    return s > (T)0 ? s : (T)0;
}

// e-Ar ionisation (e_collisions.cpp:235-256); REF threshold 15.755, PHYS threshold thr
template <typename T>
HD T ar_iz_ref(T E) {
    if (E < (T)15.755) return (T)0.0;
    if (E < (T)100.0 && E > (T)15.755)
        return (T)-3.0e-7 * E * E * E * E + (T)8.0e-5 * E * E * E - (T)0.008 * E * E +
               (T)0.416 * E - (T)4.758;
    if (E > (T)100.0) return (T)130.0 * pow(E, (T)-0.74);
    return (T)0.0;
}
template <typename T, typename M = LibMath>
HD T ar_iz_phys(T E, T thr) {
    if (E < thr) return (T)0.0;
    T s;
    if (E < (T)100.0)
        s = (T)-3.0e-7 * E * E * E * E + (T)8.0e-5 * E * E * E - (T)0.008 * E * E + (T)0.416 * E -
            (T)4.758;
    else
        s = (T)130.0 * M::powr(E, (T)-0.74);
    return s > (T)0 ? s : (T)0;
}

// e-Ar electronic excitation, one lumped channel (PHYS only; the reference
// has no excitation channel).  Cross section: the sum of the six tabulated
// excitation levels (thresholds 11.549-14.09 eV; data of Chutjian &
// Cartwright, Phys. Rev. A 23 (1981) 2178) of the recommended Ar set that
// ships with the reference, References/PhelpsAr_e_cross_section.pdf,
// section 1, tables (2)-(7).  Each table is linear between its rows, zero
// below its threshold and held at its last row above its last energy (the
// usual convention of Boltzmann solvers for tabulated data); a sum of such
// functions is linear between the union of their energies, tabulated below
// (35 rows, sums exact to 1e-10).  The channel fires above thr (11.55 eV,
// the lumped level and energy loss of Vahedi & Surendra, Comput. Phys.
// Commun. 87 (1995) 179).  Units 1e-20 m^2 (= angstrom^2).
#define AR_EXC_ROWS 35
#define AR_EXC_TABLE_E                                                                          \
    {11.549, 11.624, 11.723, 11.828, 11.9, 12.0, 12.5, 12.8, 12.9, 13.0, 13.1, 13.4, 14.0, 14.09, \
     15.0,   16.0,   17.0,   18.0,   20.0, 22.0, 24.0, 25.0, 26.0, 28.0, 30.0, 35.0, 40.0, 45.0,  \
     50.0,   60.0,   70.0,   80.0,   90.0, 100.0, 120.0}
#define AR_EXC_TABLE_S                                                                      \
    {0.0,          0.0006196581, 0.0024907983, 0.0048312732, 0.0082177402, 0.0116559727,   \
     0.0325279863, 0.0452511945, 0.0485255973, 0.0548,       0.0666,       0.095,          \
     0.154,        0.164395,     0.2975,       0.4446,       0.5685,       0.6668666667,   \
     0.7946,       0.85176,      0.87392,      0.885,        0.8893,       0.8959,         \
     0.9005,       0.9144,       0.9172,       0.9045,       0.8898,       0.8691,         \
     0.8403,       0.7956,       0.76185,      0.7321,       0.6919}
static __constant__ const float ar_exc_E_dev[AR_EXC_ROWS] = AR_EXC_TABLE_E;
static __constant__ const float ar_exc_S_dev[AR_EXC_ROWS] = AR_EXC_TABLE_S;
static const double ar_exc_E_host[AR_EXC_ROWS] = AR_EXC_TABLE_E;
static const double ar_exc_S_host[AR_EXC_ROWS] = AR_EXC_TABLE_S;

template <typename T, typename M = LibMath>
HD T ar_exc_phys(T E, T thr) {
    if (!(E >= thr) || !(E >= (T)11.549)) return (T)0.0;
#if defined(__HIP_DEVICE_COMPILE__)
    const float *te = ar_exc_E_dev, *ts = ar_exc_S_dev;
#else
    const double *te = ar_exc_E_host, *ts = ar_exc_S_host;
#endif
    if (E >= (T)te[AR_EXC_ROWS - 1]) return (T)ts[AR_EXC_ROWS - 1];
    int lo = 0, hi = AR_EXC_ROWS - 1;  // te[lo] <= E < te[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((T)te[mid] <= E)
            lo = mid;
        else
            hi = mid;
    }
    const T e0 = (T)te[lo], e1 = (T)te[hi], s0 = (T)ts[lo], s1 = (T)ts[hi];
    return s0 + M::quot((s1 - s0) * (E - e0), e1 - e0);
}

// e-H elastic (e_collisions.cpp:267-283)
template <typename T, typename M = LibMath>
HD T h_el(T E) {
    if (E < (T)1.0e-2) return (T)3.834 * E * E - (T)20.82 * E + (T)40.98;
    if (E < (T)20.0) return (T)-6.15 * M::logn(E) + (T)20.91;
    return (T)102.2 * M::powr(E, (T)-1.19);
}
// e-H ionisation, Janev (e_collisions.cpp:293-308).  The bracket/(I E) is in
// units of 1e-17 m^2; REF uses the number as if it were 1e-20 m^2 (as the
// reference does), PHYS converts (x 1000).
template <typename T, typename M = LibMath>
HD T h_iz_janev(T E) {
    const T I = (T)13.6;
    if (E < I) return (T)0.0;
    T c = (T)1.0 - M::quot(I, E);
    return M::quot((T)1.0, I * E) * ((T)0.18450 * M::logn(M::quot(E, I)) + (T)-0.032226 * c +
                                 (T)-0.034539 * c * c + (T)1.4003 * c * c * c +
                                 (T)-2.8115 * c * c * c * c + (T)2.2986 * c * c * c * c * c);
}

// Ar+ - Ar charge exchange (ion_collisions.cpp:170-176)
template <typename T, typename M = LibMath>
HD T ar_cx(T E) {
    if (E > (T)4.0) return ((T)2.0 + M::quot((T)5.5, M::sqroot(E))) * (T)10.0;
    return ((T)-2.95 * M::sqroot(E) + (T)10.65) * (T)10.0;
}
// Ar+ - Ar isotropic elastic (absent from the reference; PHYS only):
// Phelps, J. Appl. Phys. 76 (1994) 747: 2e-19/(sqrt(e)(1+e)) + 3e-19 e/(1+e/3)^2 m^2
template <typename T, typename M = LibMath>
HD T ar_iso_phys(T E) {
    T e = E > (T)1e-8 ? E : (T)1e-8;
    T d = (T)1.0 + M::quot(e, (T)3.0);
    return M::quot((T)20.0, M::sqroot(e) * ((T)1.0 + e)) + M::quot((T)30.0 * e, d * d);
}
// H+ - H charge exchange, Janev (ion_collisions.cpp:185-199) in 1e-20 m^2
// (the fit's 1e-16 cm^2).  E_kev_amu = lab energy / 1000 / mass in amu; the
// reference divides by AMU in kilograms instead, which makes sigma vanish
// (REF reproduces that through kev_amu_div).
template <typename T, typename M = LibMath>
HD T h_cx_janev(T E_kev_amu) {
    if (E_kev_amu < (T)1.0e-4) return (T)0.0;
    T e = E_kev_amu;
    return M::quot((T)3.2345 * M::logn(M::quot((T)235.88, e) + (T)2.3713),
                   (T)1.0 + (T)0.038371 * e + (T)3.8068e-6 * M::powr(e, (T)3.5) +
                       (T)1.1832e-10 * M::powr(e, (T)5.4));
}

// --------------------------------------------------------- parameters
struct MccConst {
    int on;         // collisions enabled for this species
    int mode;       // PIC_DEV_MODE_REF / PIC_DEV_MODE_PHYS
    int gas;        // GAS_AR / GAS_H
    int electron;   // 1 electrons, 0 ions
    int sig_el_energy, sig_iz_energy, sig_cx_energy;  // REF: energy-dependent vs constant
    int N;          // nodes (REF in-domain test)
    float sig_el_const, sig_iz_const, sig_cx_const;    // REF constants (1e-20 m^2)
    float ng;       // neutral density [m^-3]
    float dt;       // species time step
    float mass;     // species mass [kg]
    float M;        // neutral (ion) mass [kg]
    float q;        // elementary charge [C]
    float vtn;      // neutral Maxwellian sigma per component [m/s] (PHYS)
    float nu_max;   // null-collision bound nu' [1/s] (PHYS)
    float p_null;   // 1 - exp(-nu' dt) (PHYS)
    float e_exc, e_iz;  // thresholds [eV] (PHYS)
    float opal_B;   // secondary-energy parameter [eV] (PHYS)
    float kev_amu_div;  // H+ CX energy divisor (eV -> keV/amu)
    const float *ntab;  // REF neutral velocity table (3 x 10000, SoA)
    int ntab_n;
    uint32_t key0, key1;
    // PHYS electrons: null-collision bound per energy bin of the particle,
    // {1 - exp(-nu_b dt), nu_b} (MCC_BINS entries, device), or null: the
    // global bound (nu_max, p_null) for every particle
    const float2 *bins;
};

// Energy bins of the null-collision bound: two per octave of the particle's
// kinetic energy E = m v^2 / 2q, bin b covering [2^(b/2 - 12), 2^((b+1)/2 - 12))
// eV; bin 0 takes everything below, the last everything above.  The bound of
// a bin is the largest n_g sigma_tot(E_rel) g over every relative speed g its
// particles can have against the neutral gas (the particle's speed range
// widened by 10 neutral sigmas), times 1.02: the method stays exact (a
// particle's collision probability is p_b nu(E_rel) / nu_b = the exact one
// to O(nu dt)^2, as with one global bound) while a slow electron, whose true
// rate is tens of times below the global bound set near 15 eV, is rarely a
// candidate.  Particles over their bin's bound are counted (EV_OVER_BOUND).
#define MCC_BINS 64
#define MCC_BIN_PER_OCT 2
#define MCC_BIN_E0_LOG2 (-12)
// lower and upper energy [eV] of bin b (host; the first bin starts at 0, the
// last ends at 1e7 eV)
inline double mcc_bin_lo(int b) { return b == 0 ? 0.0 : exp2((double)b / MCC_BIN_PER_OCT + MCC_BIN_E0_LOG2); }
inline double mcc_bin_hi(int b) {
    return b == MCC_BINS - 1 ? 1e7 : exp2((double)(b + 1) / MCC_BIN_PER_OCT + MCC_BIN_E0_LOG2);
}
// bin of a particle of speed^2 v2 (device: hardware log2, NaN and 0 -> bin 0)
__device__ __forceinline__ int energy_bin(const MccConst &c, float v2) {
    const float E = (0.5f / c.q) * c.mass * v2;
    float t = (__builtin_amdgcn_logf(E) - (float)MCC_BIN_E0_LOG2) * (float)MCC_BIN_PER_OCT;
    t = fminf(fmaxf(t, 0.0f), (float)(MCC_BINS - 1));
    return (int)t;
}
// the bound that applies to a particle: {p, nu}; tab: the bins (c.bins or a
// block's LDS copy of them)
__device__ __forceinline__ float2 null_bound(const MccConst &c, float v2, const float2 *tab) {
    return tab ? tab[energy_bin(c, v2)] : make_float2(c.p_null, c.nu_max);
}

// event codes (counters)
enum Event { EV_NONE = 0, EV_ELASTIC = 1, EV_EXCITE = 2, EV_IONIZE = 3, EV_CX = 4, EV_ISO = 5 };
// flag or-ed into the returned event when the particle's total collision
// frequency exceeded the null-collision bound (its probability was clipped)
#define EV_OVER_BOUND 0x100
#define EV_KIND(e) ((e) & 0xFF)

// total electron collision frequency at relative energy E and speed g (PHYS)
template <typename T, typename M = LibMath>
HD void e_nu_parts(const MccConst &c, T E, T g, T &nu_el, T &nu_ex, T &nu_iz) {
    T f = (T)c.ng * (T)1e-20 * g;
    if (c.gas == GAS_AR) {
        nu_el = f * ar_el_phys<T, M>(E);
        nu_ex = f * ar_exc_phys<T, M>(E, (T)c.e_exc);
        nu_iz = f * ar_iz_phys<T, M>(E, (T)c.e_iz);
    } else {
        T s = h_el<T, M>(E);
        nu_el = f * (s > (T)0 ? s : (T)0);
        nu_ex = (T)0;
        nu_iz = f * (T)1000.0 * h_iz_janev<T, M>(E);
    }
}

template <typename T, typename M = LibMath>
HD void i_nu_parts(const MccConst &c, T E, T g, T &nu_iso, T &nu_cx) {
    T f = (T)c.ng * (T)1e-20 * g;
    if (c.gas == GAS_AR) {
        nu_iso = f * ar_iso_phys<T, M>(E);
        nu_cx = f * ar_cx<T, M>(E);
    } else {
        nu_iso = (T)0;
        nu_cx = f * h_cx_janev<T, M>(M::quot(M::quot(E, (T)1000.0), (T)c.kev_amu_div));
    }
}

// ------------------------------------------------------- device helpers

// polar/azimuthal rotation of vector g by (chi, phi) (Nanbu, IEEE TPS 28
// (2000) 971, eq. 56; the same construction as collisions.cpp:141-169)
__device__ __forceinline__ void rotate(float gx, float gy, float gz, float cosc, float sinc,
                                       float cosp, float sinp, float &ox, float &oy, float &oz) {
    float g = HwMath::sqroot(gx * gx + gy * gy + gz * gz);
    float gp = HwMath::sqroot(gy * gy + gz * gz);
    if (gp < 1e-6f * g || gp == 0.0f) {  // g along x: any perpendicular pair
        ox = gx * cosc;
        oy = g * sinc * cosp;
        oz = g * sinc * sinp;
        return;
    }
    const float s = sinc * __builtin_amdgcn_rcpf(gp);
    ox = gx * cosc + gp * sinc * cosp;
    oy = gy * cosc - (gx * gy * cosp - g * gz * sinp) * s;
    oz = gz * cosc - (gx * gz * cosp + g * gy * sinp) * s;
}

// anisotropic scattering angle (Vahedi & Surendra eq. 13; collisions.cpp:121-124)
__device__ __forceinline__ float cos_chi(float E, float R) {
    if (E < 1e-4f) return 1.0f - 2.0f * R;
    float c = HwMath::quot(2.0f + E - 2.0f * HwMath::powr(1.0f + E, R), E);
    return fminf(1.0f, fmaxf(-1.0f, c));
}

// fourth Philox counter word of collision draws (the loader uses 0)
#define RNG_DOMAIN_MCC 0x4D434331u

__device__ __forceinline__ void normals3(uint32_t id, uint32_t step, uint32_t stream,
                                         const MccConst &c, float &a, float &b, float &d) {
    uint32_t r[4];
    philox4x32_10(id, step, stream, RNG_DOMAIN_MCC, c.key0, c.key1, r);
    // Box-Muller
    float r1 = HwMath::sqroot(-2.0f * HwMath::logn(u01(r[0])));
    float r2 = HwMath::sqroot(-2.0f * HwMath::logn(u01(r[2])));
    a = r1 * cos_rev(u01(r[1]));
    b = r1 * sin_rev(u01(r[1]));
    d = r2 * cos_rev(u01(r[3]));
}

// Children produced by one collision (PHYS ionisation).
struct Child {
    float vx, vy, vz;     // new electron velocity
    float ivx, ivy, ivz;  // new ion velocity
};

// RNG streams (third counter word = species * 16 + stream)
#define RS_MAIN 0
#define RS_NEUTRAL 1
#define RS_IONIZE 2
#define RS_PLOP 3
#define RS_PLOP2 4
// 5, 6: second and third draws of a re-injected pair
#define RS_REFCX 9   // reference-mode charge exchange: neutral table index
#define RS_INJECT 10  // particle source: + 256 k for the k-th draw of a particle
#define RS_REDUCE 11  // population reduction: selection key

// Physical-mode null-collision candidate test: u < 1 - exp(-nu' dt) with u
// from one Philox2x32-10 word (counter (id, step), the species' key); only a
// candidate draws its collision from Philox4x32-10 stream RS_MAIN.
#define RNG_KEY_CAND_E 0x43414E45u
#define RNG_KEY_CAND_I 0x43414E49u
// v2: the particle's speed^2 (selects its energy bin when c.bins is set);
// tab: c.bins or a block's LDS copy of them
__device__ __forceinline__ bool candidate(const MccConst &c, uint32_t id, uint32_t step, float v2,
                                          const float2 *tab) {
    uint32_t r[2];
    philox2x32_10(id, step, c.key0 ^ (c.electron ? RNG_KEY_CAND_E : RNG_KEY_CAND_I), r);
    return u01(r[0]) < (tab ? tab[energy_bin(c, v2)].x : c.p_null);
}
__device__ __forceinline__ bool candidate(const MccConst &c, uint32_t id, uint32_t step, float v2) {
    return candidate(c, id, step, v2, c.bins);
}

// Electron collision.  Returns the event; on EV_IONIZE in PHYS mode `ch`
// holds the ejected electron and the new ion; in REF mode EV_IONIZE means the
// parent is removed.  inside: the particle passes the reference's
// "inside the domain" test (e_collisions.cpp:507-511), used in REF mode.
__device__ __forceinline__ int collide_electron(const MccConst &c, float &vx, float &vy, float &vz,
                                                uint32_t id, uint32_t step, bool inside, Child &ch,
                                                const float2 *tab) {
    const uint32_t s0 = (c.electron ? 0u : 16u);
    uint32_t r[4];
    if (c.mode == PIC_DEV_MODE_REF) {
        philox4x32_10(id, step, s0 + RS_MAIN, RNG_DOMAIN_MCC, c.key0, c.key1, r);
        if (!inside) return EV_NONE;
        float v2 = vx * vx + vy * vy + vz * vz;
        float E = (0.5f / c.q) * c.mass * v2;
        float u = sqrtf(2.0f * c.q * E / c.mass);
        float sel = c.sig_el_energy ? (c.gas == GAS_AR ? ar_el_ref(E) : h_el(E)) : c.sig_el_const;
        float siz = c.sig_iz_energy ? (c.gas == GAS_AR ? ar_iz_ref(E) : h_iz_janev(E)) : c.sig_iz_const;
        float R = u01(r[0]);
        float pel = 1.0f - expf(-(u * c.dt * 1.0e-20f * c.ng * sel));
        int ev = EV_NONE;
        if (pel > 2.0f * R) {
            // collisions.cpp:107-171 with coll_factor = M/(m+M)
            float Rs = u01(r[1]);
            float cc = (E < 1.0e-2f) ? sqrtf(1.0f - Rs)
                                     : 2.0f * (1.0f - powf(1.0f + E, Rs)) / E + 1.0f;
            cc = fminf(1.0f, cc);
            float sc = sqrtf(fmaxf(0.0f, 1.0f - cc * cc));
            float cp = cosf(6.2831853f * Rs), sp = sinf(6.2831853f * Rs);
            int idx = (int)(u01(r[2]) * (float)(c.ntab_n - 1));
            idx = idx < 0 ? 0 : (idx > c.ntab_n - 2 ? c.ntab_n - 2 : idx);
            float nvx = c.ntab[idx], nvy = c.ntab[c.ntab_n + idx], nvz = c.ntab[2 * c.ntab_n + idx];
            float ux = vx - nvx, uy = vy - nvy, uz = vz - nvz;
            float uperp = sqrtf(uy * uy + uz * uz);
            float um = sqrtf(ux * ux + uperp * uperp);
            float uyp = 1.0f, uzp = 1.0f;
            if (!(uperp < 1.0f)) {
                uyp = uy / uperp;
                uzp = uz / uperp;
            }
            float cf = c.M / (c.mass + c.M);
            vz -= cf * (uz * (1.0f - cc) - (ux * uzp * cp - um * uyp * sp) * sc);
            vy -= cf * (uy * (1.0f - cc) - (ux * uyp * cp + um * uzp * sp) * sc);
            vx -= cf * (ux * (1.0f - cc) + uperp * cp * sc);
            ev = EV_ELASTIC;
        }
        float piz = 1.0f - expf(-(u * c.dt * 1.0e-20f * c.ng * siz));
        if (piz > 0.0f) ev = EV_IONIZE;
        return ev;
    }
    // PHYS null collision
    const float2 bound = null_bound(c, vx * vx + vy * vy + vz * vz, tab);
    {
        uint32_t rc[2];
        philox2x32_10(id, step, c.key0 ^ (c.electron ? RNG_KEY_CAND_E : RNG_KEY_CAND_I), rc);
        if (!(u01(rc[0]) < bound.x)) return EV_NONE;
    }
    philox4x32_10(id, step, s0 + RS_MAIN, RNG_DOMAIN_MCC, c.key0, c.key1, r);
    float nx, ny, nz;
    normals3(id, step, s0 + RS_NEUTRAL, c, nx, ny, nz);
    nx *= c.vtn;
    ny *= c.vtn;
    nz *= c.vtn;
    float gx = vx - nx, gy = vy - ny, gz = vz - nz;
    float g2 = gx * gx + gy * gy + gz * gz;
    float g = HwMath::sqroot(g2);
    float E = (0.5f / c.q) * c.mass * g2;
    float nel, nex, niz;
    e_nu_parts<float, HwMath>(c, E, g, nel, nex, niz);
    const int over = (nel + nex + niz > bound.y) ? EV_OVER_BOUND : 0;
    float x = u01(r[1]) * bound.y;
    float cp = cos_rev(u01(r[3])), sp = sin_rev(u01(r[3]));
    const float rg = __builtin_amdgcn_rcpf(g);
    if (x < nel) {
        float cc = cos_chi(E, u01(r[2]));
        float sc = HwMath::sqroot(fmaxf(0.0f, 1.0f - cc * cc));
        float hx, hy, hz;
        rotate(gx, gy, gz, cc, sc, cp, sp, hx, hy, hz);
        float cf = c.M / (c.mass + c.M);
        vx -= cf * (gx - hx);
        vy -= cf * (gy - hy);
        vz -= cf * (gz - hz);
        return EV_ELASTIC | over;
    }
    if (x < nel + nex) {
        float Ep = E - c.e_exc;
        if (!(Ep > 0.0f)) return EV_NONE;
        float cc = cos_chi(Ep, u01(r[2]));
        float sc = HwMath::sqroot(fmaxf(0.0f, 1.0f - cc * cc));
        float hx, hy, hz;
        rotate(gx, gy, gz, cc, sc, cp, sp, hx, hy, hz);
        float k = HwMath::sqroot(HwMath::quot(2.0f * c.q * Ep, c.mass)) * rg;
        vx = nx + hx * k;
        vy = ny + hy * k;
        vz = nz + hz * k;
        return EV_EXCITE | over;
    }
    if (x < nel + nex + niz) {
        float Ex = E - c.e_iz;
        if (!(Ex > 0.0f)) return EV_NONE;
        uint32_t q4[4];
        philox4x32_10(id, step, s0 + RS_IONIZE, RNG_DOMAIN_MCC, c.key0, c.key1, q4);
        // secondary energy (Opal, Peterson & Beaty 1971, as in Vahedi & Surendra)
        // tan(a) from the revolutions a / 2 pi
        const float ta = u01(q4[0]) * atanf(HwMath::quot(Ex, 2.0f * c.opal_B)) * 0.159154943f;
        float Eej = c.opal_B * HwMath::quot(sin_rev(ta), cos_rev(ta));
        float Esc = Ex - Eej;
        float cc = cos_chi(Esc, u01(r[2]));
        float sc = HwMath::sqroot(fmaxf(0.0f, 1.0f - cc * cc));
        float hx, hy, hz;
        rotate(gx, gy, gz, cc, sc, cp, sp, hx, hy, hz);
        float k = HwMath::sqroot(HwMath::quot(2.0f * c.q * Esc, c.mass)) * rg;
        vx = nx + hx * k;
        vy = ny + hy * k;
        vz = nz + hz * k;
        float cc2 = cos_chi(Eej, u01(q4[1]));
        float sc2 = HwMath::sqroot(fmaxf(0.0f, 1.0f - cc2 * cc2));
        rotate(gx, gy, gz, cc2, sc2, cos_rev(u01(q4[2])), sin_rev(u01(q4[2])), hx, hy, hz);
        float k2 = HwMath::sqroot(HwMath::quot(2.0f * c.q * Eej, c.mass)) * rg;
        ch.vx = nx + hx * k2;
        ch.vy = ny + hy * k2;
        ch.vz = nz + hz * k2;
        ch.ivx = nx;
        ch.ivy = ny;
        ch.ivz = nz;
        return EV_IONIZE | over;
    }
    return EV_NONE | over;
}
__device__ __forceinline__ int collide_electron(const MccConst &c, float &vx, float &vy, float &vz,
                                                uint32_t id, uint32_t step, bool inside, Child &ch) {
    return collide_electron(c, vx, vy, vz, id, step, inside, ch, c.bins);
}

// Ion collision (PHYS: isotropic elastic + charge exchange against a
// Maxwellian neutral).  REF ion charge exchange is a per-column sequential
// accumulator and is not handled here.
__device__ __forceinline__ int collide_ion(const MccConst &c, float &vx, float &vy, float &vz,
                                           uint32_t id, uint32_t step) {
    const uint32_t s0 = 16u;
    if (!candidate(c, id, step, vx * vx + vy * vy + vz * vz)) return EV_NONE;
    uint32_t r[4];
    philox4x32_10(id, step, s0 + RS_MAIN, RNG_DOMAIN_MCC, c.key0, c.key1, r);
    float nx, ny, nz;
    normals3(id, step, s0 + RS_NEUTRAL, c, nx, ny, nz);
    nx *= c.vtn;
    ny *= c.vtn;
    nz *= c.vtn;
    float gx = vx - nx, gy = vy - ny, gz = vz - nz;
    float g2 = gx * gx + gy * gy + gz * gz;
    float g = HwMath::sqroot(g2);
    float E = (0.5f / c.q) * c.mass * g2;
    float niso, ncx;
    i_nu_parts<float, HwMath>(c, E, g, niso, ncx);
    const int over = (niso + ncx > c.nu_max) ? EV_OVER_BOUND : 0;
    float x = u01(r[1]) * c.nu_max;
    if (x < niso) {
        // equal masses: centre-of-mass velocity plus half the relative
        // velocity turned to an isotropic direction
        float cc = 1.0f - 2.0f * u01(r[2]);
        float sc = HwMath::sqroot(fmaxf(0.0f, 1.0f - cc * cc));
        float cmx = 0.5f * (vx + nx), cmy = 0.5f * (vy + ny), cmz = 0.5f * (vz + nz);
        vx = cmx + 0.5f * g * sc * cos_rev(u01(r[3]));
        vy = cmy + 0.5f * g * sc * sin_rev(u01(r[3]));
        vz = cmz + 0.5f * g * cc;
        return EV_ISO | over;
    }
    if (x < niso + ncx) {
        vx = nx;
        vy = ny;
        vz = nz;
        return EV_CX | over;
    }
    return EV_NONE | over;
}

// Position uniform in [dx, L-dx]^2 (boundary.cpp:860-866) and Maxwellian
// velocities with per-component sigma se / si for a re-injected electron-ion
// pair, drawn from the lost ion's stream.  REF uses the reference's sampler
// v = vt cos(2 pi R2) sqrt|ln R1| (boundary.cpp:833-845), sigma = vt/sqrt 2.
__device__ __forceinline__ void plop_pair(const MccConst &c, uint32_t id, uint32_t step, int N,
                                          float se, float si, int ref, PState &e, PState &i) {
    // 14 words from four counters: 2 for the position, 12 for the six
    // Box-Muller pairs, none used twice
    uint32_t r[4], a[4], b[4], d[4];
    philox4x32_10(id, step, 16u + RS_PLOP, RNG_DOMAIN_MCC, c.key0, c.key1, r);
    philox4x32_10(id, step, 16u + RS_PLOP2, RNG_DOMAIN_MCC, c.key0, c.key1, a);
    philox4x32_10(id, step, 16u + RS_PLOP2 + 1u, RNG_DOMAIN_MCC, c.key0, c.key1, b);
    philox4x32_10(id, step, 16u + RS_PLOP2 + 2u, RNG_DOMAIN_MCC, c.key0, c.key1, d);
    float span = (float)(N - 3);
    float tx = 1.0f + u01(r[0]) * span, ty = 1.0f + u01(r[1]) * span;
    int jx = (int)floorf(tx), jy = (int)floorf(ty);
    float ox = tx - (float)jx, oy = ty - (float)jy;
    if (ox >= 1.0f) { ox = 0.0f; jx += 1; }
    if (oy >= 1.0f) { oy = 0.0f; jy += 1; }
    e.jx = i.jx = jx;
    e.jy = i.jy = jy;
    e.ox = i.ox = ox;
    e.oy = i.oy = oy;
    float n[6];
    uint32_t w[6] = {a[0], a[1], a[2], a[3], b[0], b[1]};
    uint32_t w2[6] = {b[2], b[3], r[2], r[3], d[0], d[1]};
    for (int k = 0; k < 6; k++) {
        float R1 = u01(w[k]), R2 = u01(w2[k]);
        n[k] = ref ? cosf(6.2831853f * R2) * sqrtf(fabsf(logf(R1)))
                   : sqrtf(-2.0f * logf(R1)) * cosf(6.2831853f * R2);
    }
    e.vx = se * n[0];
    e.vy = se * n[1];
    e.vz = se * n[2];
    i.vx = si * n[3];
    i.vy = si * n[4];
    i.vz = si * n[5];
}

}  // namespace mcc
