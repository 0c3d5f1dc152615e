/*
 * orc_control.c -- particle-count control in the oracle (TEST INFRASTRUCTURE
 * ONLY): the particle source (particle.cpp:491-510), the population reduction
 * with re-weighting (particle.cpp:520-653) and the weighted fixed-point
 * deposit (pic.cl:46-49 multiplies each corner by the particle's weight).
 *
 * The reference draws with rand(); the GPU library and this oracle draw from
 * Philox4x32-10 with counter (k or id, step, species*16 + stream,
 * 0x4D434331), key (key0, key1): stream 10 (+256 d) for the source, 11 for
 * the reduction keys.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_DOMAIN 0x4D434331u
#define ORC_RS_INJECT 10u
#define ORC_RS_REDUCE 11u

/* absolute coordinate -> cell + fp32 offset, as the source kernel */
static void inj_cell(double x, double dx, int *j, float *o) {
    double t = x / dx;
    double k = floor(t);
    float f = (float)(t - k);
    int jj = (int)k;
    if (f >= 1.0f) {
        f = 0.0f;
        jj += 1;
    }
    *j = jj;
    *o = f;
}

/* particle.cpp:452-464: v = vt cos(2 pi R2) sqrt|ln R1| (one (R1, R2) pair
 * per component; the rejection |v| > 10^4 vt never fires for R1 in (0,1)) */
static float inj_vel(const uint32_t r[4], double vt) {
    double R1 = orc_u01d(r[0], r[1]), R2 = orc_u01d(r[2], r[3]);
    return (float)(vt * cos(6.283185307179586 * R2) * sqrt(fabs(log(R1))));
}

/* particle.cpp:496-509: x = DX + R (XLENGTH - 2 DX), likewise y */
void orc_inject(long n, uint32_t species, uint32_t step, uint32_t key0, uint32_t key1, double dx,
                double L, double vt, uint32_t *cell, float *ox, float *oy, float *vx, float *vy,
                float *vz) {
    const uint32_t key[2] = {key0, key1};
    const uint32_t s0 = species * 16u + ORC_RS_INJECT;
    const double span = L - 2.0 * dx;
    for (long k = 0; k < n; k++) {
        uint32_t a[4], b[4], c[4], d[4];
        uint32_t ca[4] = {(uint32_t)k, step, s0, ORC_DOMAIN};
        uint32_t cb[4] = {(uint32_t)k, step, s0 + 256u, ORC_DOMAIN};
        uint32_t cc[4] = {(uint32_t)k, step, s0 + 512u, ORC_DOMAIN};
        uint32_t cd[4] = {(uint32_t)k, step, s0 + 768u, ORC_DOMAIN};
        orc_philox4x32_10(ca, key, a);
        orc_philox4x32_10(cb, key, b);
        orc_philox4x32_10(cc, key, c);
        orc_philox4x32_10(cd, key, d);
        int jx, jy;
        inj_cell(fma(orc_u01d(a[0], a[1]), span, dx), dx, &jx, &ox[k]);
        inj_cell(fma(orc_u01d(a[2], a[3]), span, dx), dx, &jy, &oy[k]);
        cell[k] = orc_cell_pack(jx, jy);
        vx[k] = inj_vel(b, vt);
        vy[k] = inj_vel(c, vt);
        vz[k] = inj_vel(d, vt);
    }
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

static int vgroup(float vx, float vy, const orc_reduce_params *c) {
    double v = sqrt((double)vx * (double)vx + (double)vy * (double)vy);
    int g = 0;
    while (g < c->ng && v > c->vlim[g]) g++;
    return g;
}

static uint64_t rkey(uint32_t id, uint32_t step, const orc_reduce_params *c) {
    const uint32_t key[2] = {c->key0, c->key1};
    uint32_t ctr[4] = {id, step, c->species * 16u + ORC_RS_REDUCE, ORC_DOMAIN};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    return ((uint64_t)r[0] << 32) | id;
}

/* reduce_population (particle.cpp:538-632) with the selection of the GPU
 * library: in every group above its limit, the Ngroup - new_size particles of
 * smallest key are removed; survivors of those groups take weight
 * (float)(w * Ngroup / (Ngroup - removed)).  Particles faster than the last
 * limit are in no group.  n > size_check gates it (particle.cpp:641). */
long orc_reduce(const orc_reduce_params *c, uint32_t step, long n, const float *vx,
                const float *vy, const uint32_t *id, float *w, uint8_t *keep, uint32_t *ngroup,
                uint32_t *nrem) {
    for (long i = 0; i < n; i++) keep[i] = 1;
    for (int g = 0; g < 8; g++) ngroup[g] = nrem[g] = 0;
    if (n <= (long)c->size_check) return 0;
    int *grp = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (long i = 0; i < n; i++) {
        grp[i] = vgroup(vx[i], vy[i], c);
        if (grp[i] < c->ng) ngroup[grp[i]]++;
    }
    long removed = 0;
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
    for (int g = 0; g < c->ng; g++) {
        uint32_t N = ngroup[g];
        if (N <= c->glimit[g]) continue;
        uint32_t r = N > c->gnew[g] ? N - c->gnew[g] : 0u;
        nrem[g] = r;
        if (r == 0) continue;
        long m = 0;
        for (long i = 0; i < n; i++)
            if (grp[i] == g) keys[m++] = rkey(id[i], step, c);
        if (m > 1) qsort(keys, (size_t)m, sizeof(uint64_t), cmp_u64);
        uint64_t thr = keys[r - 1];
        double gw = N > r ? (double)N / (double)(N - r) : 1.0;
        for (long i = 0; i < n; i++) {
            if (grp[i] != g) continue;
            if (rkey(id[i], step, c) <= thr) {
                keep[i] = 0;
                removed++;
            } else {
                w[i] = (float)((double)w[i] * gw);
            }
        }
    }
    free(keys);
    free(grp);
    return removed;
}

/* pic.cl:46-49 on the fixed-point layout: each corner integer of
 * orc32_deposit times the particle weight, rounded to nearest (even) once */
void orc32_deposit_w(long n, const uint32_t *cell, const float *ox, const float *oy,
                     const float *w, int N, int mode, int64_t *qfix) {
    const int64_t ONE = (int64_t)1 << 24;
    for (long g = 0; g < n; g++) {
        int jx = orc_cell_jx(cell[g]), jy = orc_cell_jy(cell[g]);
        float fx = ox[g], fy = oy[g];
        int ok;
        if (mode == ORC_MODE_PHYS)
            ok = jx >= 0 && jx <= N - 2 && jy >= 0 && jy <= N - 2;
        else
            ok = jx >= 0 && jx <= N - 2 && !(jx == 0 && fx == 0.0f) && jy >= 0 && jy <= N - 2 &&
                 !(jy == 0 && fy == 0.0f);
        if (!ok) continue;
        int64_t wx = (int64_t)rintf(fx * 16777216.0f);
        int64_t wy = (int64_t)rintf(fy * 16777216.0f);
        int64_t a[4] = {(ONE - wx) * (ONE - wy), wx * (ONE - wy), (ONE - wx) * wy, wx * wy};
        if (w[g] != 1.0f)
            for (int k = 0; k < 4; k++) a[k] = (int64_t)llrint((double)a[k] * (double)w[g]);
        long b = (long)jy * N + jx;
        qfix[b] += a[0];
        qfix[b + 1] += a[1];
        qfix[b + N] += a[2];
        qfix[b + N + 1] += a[3];
    }
}
