/*
 * orc_particles.c -- particle side of the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * orc_ref_*: FP64 AoS restatement of pic.cl (gather, push, deposit),
 *            particle.cpp (charge post-scale, dataDump I/O) and boundary.cpp
 *            (absorbing walls).
 * orc32_*:   the same steps on the GPU library's layout (fp32 SoA, packed
 *            int16 cell + fp32 offset, fixed-point charge), spelled in the
 *            operation order of the HIP kernels.
 */
#include "oracle.h"

#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

/* ---------------------------------------------------------------- Philox */

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

/* 53-bit uniform in (0,1): 27 bits of hi above 26 bits of lo */
double orc_u01d(uint32_t hi, uint32_t lo) {
    uint64_t m = ((uint64_t)(hi >> 5) << 26) | (uint64_t)(lo >> 6);
    return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

void orc_philox2x32_10(const uint32_t ctr[2], uint32_t key, uint32_t out[2]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], k = key;
    for (int r = 0; r < 10; r++) {
        uint32_t hi, lo;
        mulhilo32(0xD256D193u, c0, &hi, &lo);
        c0 = hi ^ k ^ c1;
        c1 = lo;
        k += PHILOX_W0;
    }
    out[0] = c0;
    out[1] = c1;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(PHILOX_M0, c0, &hi0, &lo0);
        mulhilo32(PHILOX_M1, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

float orc_u01(uint32_t x) { return ((float)(x >> 9) + 0.5f) * (1.0f / 8388608.0f); }

/* -------------------------------------------------------- FP64 reference */

/* C's (int) conversion as the reference's x86 build performs it: truncation,
 * and INT_MIN for NaN or out-of-range values (the x86 "integer indefinite"). */
static int ref_trunc(double v) {
    if (!(v > -2147483648.0 && v < 2147483647.0)) return INT_MIN;
    return (int)v;
}

void orc_ref_gather(const orc_particle *p, long n, const double *Ex, const double *Ey,
                    double dx, double dy, int nx, int ny, double *Exp, double *Eyp) {
    for (long g = 0; g < n; g++) {
        int jp = ref_trunc(p[g].x / dx);
        int kp = ref_trunc(p[g].y / dy);
        Exp[g] = 0.0;
        Eyp[g] = 0.0;
        if (jp < nx && p[g].x > 0.0 && kp < ny && p[g].y > 0.0) {
            double wj = p[g].x / dx - (double)jp;
            double wk = p[g].y / dy - (double)kp;
            double a4 = (1.0 - wj) * (1.0 - wk);
            double a3 = wj * (1.0 - wk);
            double a2 = wk * (1.0 - wj);
            double a1 = wj * wk;
            long b = (long)kp * nx + jp;
            Exp[g] = a4 * Ex[b] + a3 * Ex[b + 1] + a2 * Ex[b + nx] + a1 * Ex[b + nx + 1];
            Eyp[g] = a4 * Ey[b] + a3 * Ey[b + 1] + a2 * Ey[b + nx] + a1 * Ey[b + nx + 1];
        }
    }
}

void orc_ref_push(orc_particle *p, long n, const double *Exp, const double *Eyp,
                  double dx, double dy, int nx, int ny, double coef, double cos_a,
                  double sin_a, double dt, double cos_w, double sin_w, double bfield) {
    for (long g = 0; g < n; g++) {
        int jp = ref_trunc(p[g].x / dx);
        int kp = ref_trunc(p[g].y / dy);
        if (jp < nx && p[g].x > 0.0 && kp < ny && p[g].y > 0.0) {
            p[g].vx += coef * Exp[g];
            p[g].vy += coef * Eyp[g];
            if (bfield > 0.0) {
                double vpar = cos_a * p[g].vx + sin_a * p[g].vy;
                double vperp = -sin_a * p[g].vx + cos_a * p[g].vy;
                double vperp_after = vperp * cos_w + sin_w * p[g].vz;
                p[g].vz = -vperp * sin_w + cos_w * p[g].vz;
                p[g].vx = cos_a * vpar - sin_a * vperp_after;
                p[g].vy = sin_a * vpar + cos_a * vperp_after;
            }
            p[g].vx += coef * Exp[g];
            p[g].vy += coef * Eyp[g];
            p[g].x += dt * p[g].vx;
            p[g].y += dt * p[g].vy;
        }
    }
}

void orc_ref_deposit(const orc_particle *p, long n, double qc_dxdy, double dx, double dy,
                     int nx, int ny, double *q) {
    for (long g = 0; g < n; g++) {
        int jp = ref_trunc(p[g].x / dx);
        int kp = ref_trunc(p[g].y / dy);
        if (jp < nx - 1 && p[g].x > 0.0 && kp < ny - 1 && p[g].y > 0.0) {
            double wj = p[g].x / dx - (double)jp;
            double wk = p[g].y / dy - (double)kp;
            double a4 = (1.0 - wj) * (1.0 - wk);
            double a3 = wj * (1.0 - wk);
            double a2 = wk * (1.0 - wj);
            double a1 = wj * wk;
            double w = p[g].weight;
            long b = (long)kp * nx + jp;
            q[b] += qc_dxdy * a4 * w;
            q[b + 1] += qc_dxdy * a3 * w;
            q[b + nx] += qc_dxdy * a2 * w;
            q[b + nx + 1] += qc_dxdy * a1 * w;
        }
    }
}

void orc_ref_charge_postscale(double *q, int nx, int ny) {
    for (int j = 0; j < ny; j++) {
        q[j * nx] *= 2.0;
        q[j * nx + nx - 1] *= 2.0;
    }
    for (int i = 0; i < nx; i++) {
        q[i] *= 2.0;
        q[(ny - 1) * nx + i] *= 2.0;
    }
    q[0] *= 2.0;
    q[nx - 1] *= 2.0;
    q[(ny - 1) * nx] *= 2.0;
    q[(ny - 1) * nx + nx - 1] *= 2.0;
}

long orc_ref_absorb(orc_particle *p, uint32_t *id, long n, double dx, double dy, int nx, int ny) {
    long i = 0;
    while (i < n) {
        int jp = ref_trunc(p[i].x / dx);
        int kp = ref_trunc(p[i].y / dy);
        if (jp < 0 || jp >= nx || kp < 0 || kp >= ny) {
            orc_particle t = p[i];
            p[i] = p[n - 1];
            p[n - 1] = t;
            if (id) {
                uint32_t ti = id[i];
                id[i] = id[n - 1];
                id[n - 1] = ti;
            }
            n--;
        } else {
            i++;
        }
    }
    return n;
}

/* ------------------------------------------------- build-layout (fp32)  */

#define POS_GUARD 8192.0f
#define CELL_FAR 16000

/* new cell index and offset from cell j and unnormalised offset t */
static inline void reposition(int j, float t, int *jn, float *on) {
    if (!(t > -POS_GUARD && t < POS_GUARD)) {
        *jn = (t > 0.0f) ? CELL_FAR : -CELL_FAR;
        *on = 0.0f;
        return;
    }
    float k = floorf(t);
    float f = t - k;
    if (f >= 1.0f) {
        f = 0.0f;
        k = k + 1.0f;
    }
    int jj = j + (int)k;
    if (jj > CELL_FAR) jj = CELL_FAR;
    if (jj < -CELL_FAR) jj = -CELL_FAR;
    *jn = jj;
    *on = f;
}

static inline int ref_inside_push(int j, float o, int N) {
    return j >= 0 && j <= N - 1 && !(j == 0 && o == 0.0f);
}

void orc32_push(long n, uint32_t *cell, float *ox, float *oy, float *vx, float *vy, float *vz,
                const float *E32, const orc32_push_params *pp, uint8_t *status) {
    const int N = pp->N;
    for (long g = 0; g < n; g++) {
        int jx = orc_cell_jx(cell[g]), jy = orc_cell_jy(cell[g]);
        float fx = ox[g], fy = oy[g];
        int move = (pp->mode == ORC_MODE_PHYS)
                       ? (jx >= 0 && jx <= N - 2 && jy >= 0 && jy <= N - 2)
                       : (ref_inside_push(jx, fx, N) && ref_inside_push(jy, fy, N));
        if (move) {
            float wj = fx, wk = fy;
            float a4 = (1.0f - wj) * (1.0f - wk);
            float a3 = wj * (1.0f - wk);
            float a2 = wk * (1.0f - wj);
            float a1 = wj * wk;
            long b = (long)jy * N + jx;
            const float *e00 = E32 + 2 * b, *e01 = E32 + 2 * (b + 1);
            const float *e10 = E32 + 2 * (b + N), *e11 = E32 + 2 * (b + N + 1);
            float epx = a4 * e00[0] + a3 * e01[0] + a2 * e10[0] + a1 * e11[0];
            float epy = a4 * e00[1] + a3 * e01[1] + a2 * e10[1] + a1 * e11[1];
            float ux = vx[g], uy = vy[g], uz = vz[g];
            ux = ux + pp->cE * epx;
            uy = uy + pp->cE * epy;
            if (pp->bfield) {
                float vpar = pp->cos_a * ux + pp->sin_a * uy;
                float vperp = -pp->sin_a * ux + pp->cos_a * uy;
                float vperp_after = vperp * pp->cos_w + pp->sin_w * uz;
                uz = -vperp * pp->sin_w + pp->cos_w * uz;
                ux = pp->cos_a * vpar - pp->sin_a * vperp_after;
                uy = pp->sin_a * vpar + pp->cos_a * vperp_after;
            }
            ux = ux + pp->cE * epx;
            uy = uy + pp->cE * epy;
            float tx = fx + ux * pp->dtdx;
            float ty = fy + uy * pp->dtdy;
            int njx, njy;
            float nfx, nfy;
            reposition(jx, tx, &njx, &nfx);
            reposition(jy, ty, &njy, &nfy);
            jx = njx;
            jy = njy;
            fx = nfx;
            fy = nfy;
            vx[g] = ux;
            vy[g] = uy;
            vz[g] = uz;
            ox[g] = fx;
            oy[g] = fy;
            cell[g] = orc_cell_pack(jx, jy);
        }
        int lx, ly;
        if (pp->mode == ORC_MODE_REF_STEADY) {
            lx = ly = 0;
        } else if (pp->mode == ORC_MODE_PHYS) {
            lx = (jx < 0) || (jx > N - 2);
            ly = (jy < 0) || (jy > N - 2);
        } else {
            lx = (jx < -1) || (jx == -1 && fx == 0.0f) || (jx >= N);
            ly = (jy < -1) || (jy == -1 && fy == 0.0f) || (jy >= N);
        }
        uint8_t st = ORC_ST_KEPT;
        if (lx)
            st = (jx < 0) ? ORC_ST_LOST_LEFT : ORC_ST_LOST_RIGHT;
        else if (ly)
            st = (jy < 0) ? ORC_ST_LOST_BOTTOM : ORC_ST_LOST_TOP;
        status[g] = st;
    }
}

void orc32_deposit(long n, const uint32_t *cell, const float *ox, const float *oy, int N,
                   int mode, int64_t *qfix) {
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
        long b = (long)jy * N + jx;
        qfix[b] += (ONE - wx) * (ONE - wy);
        qfix[b + 1] += wx * (ONE - wy);
        qfix[b + N] += (ONE - wx) * wy;
        qfix[b + N + 1] += wx * wy;
    }
}

/* --------------------------------------------------------- dataDump I/O */

int orc_dump_write(const char *path, const orc_particle *p, long n, int precision) {
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    for (long i = 0; i < n; i++)
        fprintf(f, "%.*g %.*g %.*g %.*g %.*g %.*g\n", precision, p[i].x, precision, p[i].y,
                precision, p[i].vx, precision, p[i].vy, precision, p[i].vz, precision,
                p[i].weight);
    fclose(f);
    return 0;
}

long orc_dump_read(const char *path, orc_particle *p, long cap) {
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    long n = 0;
    orc_particle t;
    while (n < cap &&
           fscanf(f, "%lf %lf %lf %lf %lf %lf", &t.x, &t.y, &t.vx, &t.vy, &t.vz, &t.weight) == 6)
        p[n++] = t;
    fclose(f);
    return n;
}

/* ------------------------------------------------- batch diagnostics */

void orc32_diag_particles(long n, const uint32_t *cell, const float *ox, const float *oy,
                          const float *vx, const float *vy, int N, double vt, int no_bins,
                          int divide_bins_by, int spread, const int loc[4], int64_t *vsx,
                          int64_t *vsy, uint32_t *cntx, uint32_t *cnty, uint64_t *dist) {
    const int mid = N / 2, half = (no_bins - 1) / 2;
    for (long k = 0; k < n; k++) {
        int jx = orc_cell_jx(cell[k]), jy = orc_cell_jy(cell[k]);
        int xpos = jx > 0 || (jx == 0 && ox[k] > 0.0f);
        int ypos = jy > 0 || (jy == 0 && oy[k] > 0.0f);
        if (xpos && jx < N && jy == mid) {
            vsx[jx] += llrint((double)vx[k] * 1048576.0);
            cntx[jx]++;
            double q = ((double)vx[k] / vt) * (double)divide_bins_by;
            int vb = fabs(q) < (double)half + 1.0 ? (int)q : half + 1;
            if (abs(vb) <= half) {
                if (jx >= loc[0] - spread && jx <= loc[0] + spread) dist[vb + half]++;
                if (jx >= loc[2] - spread && jx <= loc[2] + spread) dist[2 * no_bins + vb + half]++;
            }
        }
        if (ypos && jy < N && jx == mid) {
            vsy[jy] += llrint((double)vy[k] * 1048576.0);
            cnty[jy]++;
            double q = ((double)vy[k] / vt) * (double)divide_bins_by;
            int vb = fabs(q) < (double)half + 1.0 ? (int)q : half + 1;
            if (abs(vb) <= half) {
                if (jy >= loc[1] - spread && jy <= loc[1] + spread) dist[no_bins + vb + half]++;
                if (jy >= loc[3] - spread && jy <= loc[3] + spread) dist[3 * no_bins + vb + half]++;
            }
        }
    }
}

void orc_diag_vfinal(int N, int64_t *vsx, int64_t *vsy, uint32_t *cntx, uint32_t *cnty,
                     double *vpx, double *vpy) {
    for (int j = 0; j < N; j++) {
        if (cntx[j] > 0) vpx[j] += ((double)vsx[j] / 1048576.0) / (double)cntx[j];
        if (cnty[j] > 0) vpy[j] += ((double)vsy[j] / 1048576.0) / (double)cnty[j];
        vsx[j] = vsy[j] = 0;
        cntx[j] = cnty[j] = 0;
    }
}
