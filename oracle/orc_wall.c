/*
 * orc_wall.c -- surface physics of the non-steady walls in the oracle (TEST
 * INFRASTRUCTURE ONLY): boundary::e_boundary and ion_boundary
 * (boundary.cpp:261-529, 586-824) with the yields set up at pic.cpp:455-505.
 *
 * Input: the particles that reached a wall (state after their push, build
 * layout).  Output: the particles the walls return or emit, each with the
 * staging key of the GPU library (4 << 32 | id reflected, 5 << 32 | id
 * backscattered, (8 + j) << 32 | id secondaries of an electron,
 * (16 + j) << 32 | id secondaries of an ion).  Draws: Philox4x32-10 with
 * counter (id, step, species*16 + 12 + 256 k, 0x4D434331).
 *
 * Reference mode follows the reference's arithmetic as written (impact angle
 * from 2E/m with E in eV, wall at N dx, secondaries at y = 0 or x = 0);
 * physical mode uses the angle to the wall normal, the wall at L and the
 * incident's position along the wall.
 */
#include "oracle.h"

#include <math.h>

#define W_DOMAIN 0x4D434331u
#define W_RS 12u
#define W_MAX_CHILDREN 8

static void wdraw(const orc_wall_params *c, int species, uint32_t id, uint32_t step, int k,
                  uint32_t r[4]) {
    const uint32_t key[2] = {c->key0, c->key1};
    uint32_t ctr[4] = {id, step, (uint32_t)species * 16u + W_RS + 256u * (uint32_t)k, W_DOMAIN};
    orc_philox4x32_10(ctr, key, r);
}

static double u01w(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }

typedef struct {
    int left, right, bottom, top;
} sides;

static sides wall_sides(int jx, float ox, int jy, float oy, const orc_wall_params *c) {
    sides w;
    if (c->mode == ORC_MODE_PHYS) {
        w.left = jx < 0;
        w.right = jx > c->N - 2;
        w.bottom = jy < 0;
        w.top = jy > c->N - 2;
    } else {  /* (int)(x/DX) < 0 or >= X_NO_OF_CELLS (boundary.cpp:595-599) */
        w.left = jx < -1 || (jx == -1 && ox == 0.0f);
        w.right = jx >= c->N;
        w.bottom = jy < -1 || (jy == -1 && oy == 0.0f);
        w.top = jy >= c->N;
    }
    return w;
}

static void wcell(double x, double dx, int *j, float *o) {
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

typedef struct {
    double x, y, vx, vy, vz;
} wstate;

/* secondary_vel / i_secondary_vel (boundary.cpp:423-482, 696-760) */
static void emit(wstate *p, double speed, sides sw, const uint32_t xb[4], const uint32_t yb[4],
                 int speed_from_R, const orc_wall_params *c) {
    const double two_pi = 6.283185307179586;
    if (sw.left || sw.right) {
        double R = orc_u01d(xb[0], xb[1]), Rphi = u01w(xb[2]), Rchi = u01w(xb[3]);
        double v = speed_from_R ? sqrt(2.0 * R * c->q * (c->thr - 2.0 * c->wf) / c->me) : speed;
        double chi = asin(sqrt(Rchi)), phi = two_pi * Rphi;
        p->vx = fabs(v * cos(chi));
        p->vy = v * sin(chi) * cos(phi);
        p->vz = v * sin(chi) * sin(phi);
        p->x = R * c->dt * fabs(p->vx);
        if (sw.right) {
            p->x = c->Lw - p->x;
            p->vx = -fabs(v * cos(chi));
        }
    }
    if (sw.bottom || sw.top) {
        double R = orc_u01d(yb[0], yb[1]), Rphi = u01w(yb[2]), Rchi = u01w(yb[3]);
        double v = speed_from_R ? sqrt(2.0 * R * c->q * (c->thr - 2.0 * c->wf) / c->me) : speed;
        double chi = asin(sqrt(Rchi)), phi = two_pi * Rphi;
        p->vy = fabs(v * cos(chi));
        p->vx = v * sin(chi) * cos(phi);
        p->vz = v * sin(chi) * sin(phi);
        p->y = R * c->dt * fabs(p->vy);
        if (sw.top) {
            p->y = c->Lw - p->y;
            p->vy = -fabs(v * cos(chi));
        }
    }
}

static long put(orc_wall_out *o, long m, long cap, const wstate *p, double dx, uint64_t key,
                float w) {
    if (m >= cap) return m + 1;
    int jx, jy;
    wcell(p->x, dx, &jx, &o->ox[m]);
    wcell(p->y, dx, &jy, &o->oy[m]);
    o->cell[m] = orc_cell_pack(jx, jy);
    o->vx[m] = (float)p->vx;
    o->vy[m] = (float)p->vy;
    o->vz[m] = (float)p->vz;
    o->w[m] = w;
    o->key[m] = key;
    return m + 1;
}

/* Vaughan (boundary.cpp:261-282) */
double orc_vaughan(double E, double theta, double max_yield, double E0, double Emax0) {
    const double two_pi = 6.283185307179586;
    double v = (E - E0) / (Emax0 - E0);
    double k = (v > 1.0) ? 0.25 : 0.62;
    double fac = 1.0 + k * (theta * theta / two_pi);
    double w = (E - E0) / (Emax0 * fac - E0);
    return max_yield * fac * pow(w * exp(1.0 - w), k);
}

/* boundary.cpp:292-299 */
double orc_ar_ion_yield(double E) {
    if (!(E >= 80.0)) return 0.0;
    return 0.006 * E / (1.0 + E / 10.0) + 1.05e-4 * (sqrt(E - 80.0) / (1.0 + pow(E / 8000.0, 1.5)));
}

long orc_wall(const orc_wall_params *c, int species, uint32_t step, long n, const uint32_t *cell,
              const float *ox, const float *oy, const float *vx, const float *vy, const float *vz,
              const float *w, const uint32_t *id, orc_wall_out *out, long cap, int64_t counts[6]) {
    long m = 0;
    for (int k = 0; k < 6; k++) counts[k] = 0;
    for (long i = 0; i < n; i++) {
        int jx = orc_cell_jx(cell[i]), jy = orc_cell_jy(cell[i]);
        sides sw = wall_sides(jx, ox[i], jy, oy[i], c);
        int onx = sw.left || sw.right;
        wstate p = {((double)jx + (double)ox[i]) * c->dx, ((double)jy + (double)oy[i]) * c->dx,
                    (double)vx[i], (double)vy[i], (double)vz[i]};
        float wt = w ? w[i] : 1.0f;
        counts[0]++;
        if (species == 1) {  /* ion_boundary (boundary.cpp:771-824) */
            double E = (0.5 / c->q) * c->mi * (p.vx * p.vx + p.vy * p.vy + p.vz * p.vz);
            double yield = c->see_i ? orc_ar_ion_yield(E) : 0.0;
            uint32_t a[4];
            wdraw(c, 1, id[i], step, 0, a);
            if (!c->see_i || !(orc_u01d(a[0], a[1]) < yield)) {
                counts[1]++;
                continue;
            }
            int no_inject = (int)floor(yield) + 1;
            if (no_inject > W_MAX_CHILDREN) no_inject = W_MAX_CHILDREN;
            for (int j = 0; j < no_inject; j++) {
                uint32_t xb[4], yb[4];
                wdraw(c, 1, id[i], step, 5 + 3 * j, xb);
                wdraw(c, 1, id[i], step, 6 + 3 * j, yb);
                wstate q = {0, 0, 0, 0, 0};
                if (c->mode == ORC_MODE_PHYS) {
                    q.x = p.x;
                    q.y = p.y;
                }
                emit(&q, 0.0, sw, xb, yb, 1, c);
                m = put(out, m, cap, &q, c->dx, ((uint64_t)(16 + j) << 32) | id[i], wt);
            }
            counts[4]++;
            counts[5] += no_inject;
            continue;
        }
        /* e_boundary (boundary.cpp:586-677) */
        double u2 = p.vx * p.vx + p.vy * p.vy + p.vz * p.vz;
        double E = (0.5 / c->q) * c->me * u2;
        double yield = 0.0;
        if (c->see_e) {
            if (c->mode == ORC_MODE_PHYS) {
                double vn = onx ? fabs(p.vx) : fabs(p.vy), vm = sqrt(u2);
                double theta = vm > 0.0 ? acos(fmin(1.0, vn / vm)) : 0.0;
                yield = E > c->E0 ? orc_vaughan(E, theta, c->max_yield, c->E0, c->Emax0) : 0.0;
            } else {
                double uu = 2.0 * E / c->me;  /* calc_angleX/Y, boundary.cpp:197-218 */
                double theta = onx ? acos(p.vx / (sqrt(uu) + 1.0e-6)) : acos(p.vy / (sqrt(uu) + 1.0e-6));
                yield = orc_vaughan(E, theta, c->max_yield, c->E0, c->Emax0);
            }
        }
        uint32_t a[4], b[4];
        wdraw(c, 0, id[i], step, 0, a);
        double R = orc_u01d(a[0], a[1]), R2 = orc_u01d(a[2], a[3]);
        if (!(R < yield)) {
            counts[1]++;
            continue;
        }
        double refl = (c->refl && E > 50.0) ? 0.1 * yield : 0.0;
        double back = (c->back && E > 50.0) ? 0.1 * yield : 0.0;
        int no_inject = (int)floor(fabs(yield - refl - back)) + 1;
        if (no_inject > W_MAX_CHILDREN) no_inject = W_MAX_CHILDREN;
        wdraw(c, 0, id[i], step, 1, b);
        if (R2 < refl / (yield + 1.0e-6)) {  /* reflect_electron, boundary.cpp:367-411 */
            double Rp = orc_u01d(b[0], b[1]);
            if (sw.right) {
                p.vx = -p.vx;
                p.x = c->Lw - Rp * c->dt * fabs(p.vx);
            } else if (sw.left) {
                p.vx = -p.vx;
                p.x = Rp * c->dt * fabs(p.vx);
            }
            if (sw.top) {
                p.vy = -p.vy;
                p.y = c->Lw - Rp * c->dt * fabs(p.vy);
            } else if (sw.bottom) {
                p.vy = -p.vy;
                p.y = Rp * c->dt * fabs(p.vy);
            }
            m = put(out, m, cap, &p, c->dx, ((uint64_t)4 << 32) | id[i], wt);
            counts[2]++;
            continue;
        }
        if (R2 < (refl + back) / (yield + 1.0e-6)) {  /* backscatter, boundary.cpp:640-648 */
            double R4 = floor(orc_u01d(b[2], b[3]) * 10000.0) / 10000.0;
            double speed = R4 * sqrt((2.0 * c->q * E) / c->me);
            uint32_t xb[4], yb[4];
            wdraw(c, 0, id[i], step, 2, xb);
            wdraw(c, 0, id[i], step, 3, yb);
            emit(&p, speed, sw, xb, yb, 0, c);
            m = put(out, m, cap, &p, c->dx, ((uint64_t)5 << 32) | id[i], wt);
            counts[3]++;
            continue;
        }
        for (int j = 0; j < no_inject; j++) {  /* secondary_electrons, boundary.cpp:487-529 */
            uint32_t v[4], xb[4], yb[4];
            wdraw(c, 0, id[i], step, 4 + 3 * j, v);
            wdraw(c, 0, id[i], step, 5 + 3 * j, xb);
            wdraw(c, 0, id[i], step, 6 + 3 * j, yb);
            wstate q = {0, 0, 0, 0, 0};
            sides qs = sw;
            if (c->mode == ORC_MODE_PHYS) {
                q = p;
            } else if (onx) {
                q.x = p.x;
                qs.bottom = qs.top = 0;
            } else {
                q.y = p.y;
                qs.left = qs.right = 0;
            }
            double Rv = orc_u01d(v[0], v[1]);
            double speed = sqrt((2.0 * Rv * c->q * (c->thr - 2.0 * c->wf) / c->me));
            emit(&q, speed, qs, xb, yb, 0, c);
            m = put(out, m, cap, &q, c->dx, ((uint64_t)(8 + j) << 32) | id[i], wt);
        }
        counts[4]++;
        counts[5] += no_inject;
    }
    return m;
}
