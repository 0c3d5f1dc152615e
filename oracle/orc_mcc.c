/*
 * orc_mcc.c -- physical-mode Monte-Carlo collisions of the oracle and a
 * collisional step loop in the reference's layout
 * (TEST INFRASTRUCTURE ONLY; also the collisional CPU baseline).
 *
 * Null-collision MCC after Vahedi & Surendra (Comput. Phys. Commun. 87
 * (1995) 179) with the reference's cross-section fits (orc_collisions.c) in
 * FP64, drawing the same Philox4x32-10 uniforms as the GPU library
 * (counter = (particle id, step, species*16 + stream, domain word)), so the
 * two agree event by event except where FP32 rounding moves a draw across a
 * decision boundary.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define DOMAIN_MCC 0x4D434331u
#define RS_MAIN 0u
#define RS_NEUTRAL 1u
#define RS_IONIZE 2u
#define RS_PLOP 3u
#define RS_PLOP2 4u

static void draw(const orc_mcc_params *c, uint32_t id, uint32_t step, uint32_t stream,
                 double u[4], uint32_t raw[4]) {
    uint32_t ctr[4] = {id, step, stream, DOMAIN_MCC}, key[2] = {c->key0, c->key1};
    orc_philox4x32_10(ctr, key, raw);
    for (int k = 0; k < 4; k++) u[k] = (double)orc_u01(raw[k]);
}

/* physical-mode candidate test: one Philox2x32-10 word, counter (id, step),
 * key = key0 xor a per-species constant; only candidates draw stream RS_MAIN */
#define KEY_CAND_E 0x43414E45u
#define KEY_CAND_I 0x43414E49u
int orc_energy_bin(const orc_mcc_params *c, double v2) {
    const double E = 0.5 * c->mass * v2 / c->q;
    double t = (log2(E) + 12.0) * 2.0;
    if (!(t > 0.0)) t = 0.0;
    if (t > ORC_MCC_BINS - 1) t = ORC_MCC_BINS - 1;
    return (int)t;
}

/* the bound of a particle of speed^2 v2: the bin's (electrons) or global */
static void bound(const orc_mcc_params *c, double v2, double *p, double *nu) {
    if (c->binned) {
        const int b = orc_energy_bin(c, v2);
        *p = c->bin_p[b];
        *nu = c->bin_nu[b];
    } else {
        *p = c->p_null;
        *nu = c->nu_max;
    }
}

static int candidate(const orc_mcc_params *c, uint32_t id, uint32_t step, int electron, double p) {
    uint32_t ctr[2] = {id, step}, r[2];
    orc_philox2x32_10(ctr, c->key0 ^ (electron ? KEY_CAND_E : KEY_CAND_I), r);
    return (double)orc_u01(r[0]) < p;
}

static void neutral(const orc_mcc_params *c, uint32_t id, uint32_t step, uint32_t s0, double *nx,
                    double *ny, double *nz) {
    double u[4];
    uint32_t r[4];
    draw(c, id, step, s0 + RS_NEUTRAL, u, r);
    double r1 = sqrt(-2.0 * log(u[0])), t1 = 6.2831853 * u[1];
    double r2 = sqrt(-2.0 * log(u[2])), t2 = 6.2831853 * u[3];
    *nx = c->vtn * r1 * cos(t1);
    *ny = c->vtn * r1 * sin(t1);
    *nz = c->vtn * r2 * cos(t2);
}

static void e_nu(const orc_mcc_params *c, double E, double g, double *nel, double *nex, double *niz) {
    double f = c->ng * 1e-20 * g;
    if (c->gas == 0) {
        *nel = f * orc_sigma(ORC_SIG_AR_EL_PHYS, E, 0.0);
        *nex = f * orc_sigma(ORC_SIG_AR_EXC, E, c->e_exc);
        *niz = f * orc_sigma(ORC_SIG_AR_IZ_PHYS, E, c->e_iz);
    } else {
        double s = orc_sigma(ORC_SIG_H_EL, E, 0.0);
        *nel = f * (s > 0.0 ? s : 0.0);
        *nex = 0.0;
        *niz = f * 1000.0 * orc_sigma(ORC_SIG_H_IZ, E, 0.0);
    }
}

static void i_nu(const orc_mcc_params *c, double E, double g, double *niso, double *ncx) {
    double f = c->ng * 1e-20 * g;
    if (c->gas == 0) {
        *niso = f * orc_sigma(ORC_SIG_AR_ISO, E, 0.0);
        *ncx = f * orc_sigma(ORC_SIG_AR_CX, E, 0.0);
    } else {
        *niso = 0.0;
        *ncx = f * orc_sigma(ORC_SIG_H_CX, E / 1000.0 / c->kev_amu_div, 0.0);
    }
}

void orc_mcc_setup(orc_mcc_params *c) {
    double best = 0.0, top = c->electron ? 4.5 : 3.0;
    for (int k = 0; k <= 200000; k++) {
        double E = pow(10.0, -6.0 + (top + 6.0) * k / 200000.0);
        double g = sqrt(2.0 * c->q * E / c->mass), a, b, d = 0.0;
        if (c->electron)
            e_nu(c, E, g, &a, &b, &d);
        else
            i_nu(c, E, g, &a, &b);
        if (a + b + d > best) best = a + b + d;
    }
    c->nu_max = best * 1.02;
    c->p_null = 1.0 - exp(-c->nu_max * c->dt);
    c->binned = c->electron;
    if (!c->binned) return;
    for (int b = 0; b < ORC_MCC_BINS; b++) {
        const double elo = b == 0 ? 0.0 : exp2(b / 2.0 - 12.0);
        const double ehi = b == ORC_MCC_BINS - 1 ? 1e7 : exp2((b + 1) / 2.0 - 12.0);
        const double vlo = sqrt(2.0 * c->q * elo / c->mass), vhi = sqrt(2.0 * c->q * ehi / c->mass);
        const double glo = fmax(0.0, vlo - 10.0 * c->vtn), ghi = vhi + 10.0 * c->vtn;
        double m = 0.0;
        for (int k = 0; k <= 256; k++) {
            const double g = glo + (ghi - glo) * k / 256.0;
            const double E = 0.5 * c->mass * g * g / c->q;
            double a, e, d;
            e_nu(c, E, g, &a, &e, &d);
            if (a + e + d > m) m = a + e + d;
        }
        c->bin_nu[b] = m * 1.02;
        c->bin_p[b] = 1.0 - exp(-c->bin_nu[b] * c->dt);
    }
}

static double cos_chi(double E, double R) {
    if (E < 1e-4) return 1.0 - 2.0 * R;
    double cc = (2.0 + E - 2.0 * pow(1.0 + E, R)) / E;
    return cc > 1.0 ? 1.0 : (cc < -1.0 ? -1.0 : cc);
}

/* Nanbu (IEEE TPS 28 (2000) 971) eq. 56 rotation of g by (chi, phi) */
static void rotate(double gx, double gy, double gz, double cc, double sc, double cp, double sp,
                   double *ox, double *oy, double *oz) {
    double g = sqrt(gx * gx + gy * gy + gz * gz), gp = sqrt(gy * gy + gz * gz);
    if (gp < 1e-6 * g || gp == 0.0) {
        *ox = gx * cc;
        *oy = g * sc * cp;
        *oz = g * sc * sp;
        return;
    }
    *ox = gx * cc + gp * sc * cp;
    *oy = gy * cc - (gx * gy * cp - g * gz * sp) / gp * sc;
    *oz = gz * cc - (gx * gz * cp + g * gy * sp) / gp * sc;
}

int orc_collide_electron(const orc_mcc_params *c, double *vx, double *vy, double *vz, uint32_t id,
                         uint32_t step, double child[6]) {
    double u[4];
    uint32_t raw[4];
    double pb, nub;
    bound(c, *vx * *vx + *vy * *vy + *vz * *vz, &pb, &nub);
    if (!candidate(c, id, step, 1, pb)) return 0;
    draw(c, id, step, RS_MAIN, u, raw);
    double nx, ny, nz;
    neutral(c, id, step, 0u, &nx, &ny, &nz);
    double gx = *vx - nx, gy = *vy - ny, gz = *vz - nz;
    double g2 = gx * gx + gy * gy + gz * gz, g = sqrt(g2);
    double E = (0.5 / c->q) * c->mass * g2;
    double nel, nex, niz;
    e_nu(c, E, g, &nel, &nex, &niz);
    double x = u[1] * nub;
    double cp = cos(6.2831853 * u[3]), sp = sin(6.2831853 * u[3]);
    double hx, hy, hz;
    if (x < nel) {
        double cc = cos_chi(E, u[2]), sc = sqrt(fmax(0.0, 1.0 - cc * cc));
        rotate(gx, gy, gz, cc, sc, cp, sp, &hx, &hy, &hz);
        double cf = c->M / (c->mass + c->M);
        *vx -= cf * (gx - hx);
        *vy -= cf * (gy - hy);
        *vz -= cf * (gz - hz);
        return 1;
    }
    if (x < nel + nex) {
        double Ep = E - c->e_exc;
        if (!(Ep > 0.0)) return 0;
        double cc = cos_chi(Ep, u[2]), sc = sqrt(fmax(0.0, 1.0 - cc * cc));
        rotate(gx, gy, gz, cc, sc, cp, sp, &hx, &hy, &hz);
        double k = sqrt(2.0 * c->q * Ep / c->mass) / g;
        *vx = nx + hx * k;
        *vy = ny + hy * k;
        *vz = nz + hz * k;
        return 2;
    }
    if (x < nel + nex + niz) {
        double Ex = E - c->e_iz;
        if (!(Ex > 0.0)) return 0;
        double q4[4];
        uint32_t r4[4];
        draw(c, id, step, RS_IONIZE, q4, r4);
        double Eej = c->opal_B * tan(q4[0] * atan(Ex / (2.0 * c->opal_B)));
        double Esc = Ex - Eej;
        double cc = cos_chi(Esc, u[2]), sc = sqrt(fmax(0.0, 1.0 - cc * cc));
        rotate(gx, gy, gz, cc, sc, cp, sp, &hx, &hy, &hz);
        double k = sqrt(2.0 * c->q * Esc / c->mass) / g;
        *vx = nx + hx * k;
        *vy = ny + hy * k;
        *vz = nz + hz * k;
        double cc2 = cos_chi(Eej, q4[1]), sc2 = sqrt(fmax(0.0, 1.0 - cc2 * cc2));
        double ph2 = 6.2831853 * q4[2];
        rotate(gx, gy, gz, cc2, sc2, cos(ph2), sin(ph2), &hx, &hy, &hz);
        double k2 = sqrt(2.0 * c->q * Eej / c->mass) / g;
        child[0] = nx + hx * k2;
        child[1] = ny + hy * k2;
        child[2] = nz + hz * k2;
        child[3] = nx;
        child[4] = ny;
        child[5] = nz;
        return 3;
    }
    return 0;
}

int orc_collide_ion(const orc_mcc_params *c, double *vx, double *vy, double *vz, uint32_t id,
                    uint32_t step) {
    double u[4];
    uint32_t raw[4];
    if (!candidate(c, id, step, 0, c->p_null)) return 0;
    draw(c, id, step, 16u + RS_MAIN, u, raw);
    double nx, ny, nz;
    neutral(c, id, step, 16u, &nx, &ny, &nz);
    double gx = *vx - nx, gy = *vy - ny, gz = *vz - nz;
    double g2 = gx * gx + gy * gy + gz * gz, g = sqrt(g2);
    double E = (0.5 / c->q) * c->mass * g2;
    double niso, ncx;
    i_nu(c, E, g, &niso, &ncx);
    double x = u[1] * c->nu_max;
    if (x < niso) {
        double cc = 1.0 - 2.0 * u[2], sc = sqrt(fmax(0.0, 1.0 - cc * cc)), ph = 6.2831853 * u[3];
        double cmx = 0.5 * (*vx + nx), cmy = 0.5 * (*vy + ny), cmz = 0.5 * (*vz + nz);
        *vx = cmx + 0.5 * g * sc * cos(ph);
        *vy = cmy + 0.5 * g * sc * sin(ph);
        *vz = cmz + 0.5 * g * cc;
        return 5;
    }
    if (x < niso + ncx) {
        *vx = nx;
        *vy = ny;
        *vz = nz;
        return 4;
    }
    return 0;
}

/* ---------------------------------------------------- collisional step loop */

typedef struct {
    uint64_t key;
    long src;
} orc_kv;

static int kv_cmp(const void *a, const void *b) {
    uint64_t x = ((const orc_kv *)a)->key, y = ((const orc_kv *)b)->key;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* append staged particles in creator-key order with consecutive ids */
static long append_sorted(orc_particle *dst, uint32_t *did, long n, long cap, const orc_particle *st,
                          orc_kv *kv, long ns, uint32_t *next_id) {
    if (ns > 1) qsort(kv, (size_t)ns, sizeof(orc_kv), kv_cmp);
    for (long r = 0; r < ns; r++) {
        if (n < cap) {
            dst[n] = st[kv[r].src];
            did[n] = *next_id + (uint32_t)r;
            n++;
        }
    }
    *next_id += (uint32_t)ns;
    return n;
}

/* parallel helpers of orc_step.c (any split gives the serial result) */
void orc_i_gather_push(orc_particle *p, long n, const double *Ex, const double *Ey,
                       const orc_ref_config *c, double coef, double dt, double cw, double sw,
                       double *scratch);
void orc_i_deposit_par(const orc_particle *p, long n, double qc_dxdy, const orc_ref_config *c,
                       double *q);

typedef struct {
    orc_particle *a, *b; /* ejected electron, new ion */
    orc_kv *ka, *kb;
    long n, cap;
} orc_stage;

static void stage_push(orc_stage *s, const orc_particle *pa, const orc_particle *pb,
                       uint64_t key) {
    if (s->n == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 1024;
        s->a = (orc_particle *)realloc(s->a, sizeof(orc_particle) * s->cap);
        s->b = (orc_particle *)realloc(s->b, sizeof(orc_particle) * s->cap);
        s->ka = (orc_kv *)realloc(s->ka, sizeof(orc_kv) * s->cap);
        s->kb = (orc_kv *)realloc(s->kb, sizeof(orc_kv) * s->cap);
    }
    s->a[s->n] = *pa;
    s->b[s->n] = *pb;
    s->ka[s->n].key = s->kb[s->n].key = key;
    s->ka[s->n].src = s->kb[s->n].src = s->n;
    s->n++;
}

/* merge a thread's staged pairs into the step's stage */
static void stage_merge(orc_stage *dst, const orc_stage *src) {
    for (long k = 0; k < src->n; k++) stage_push(dst, &src->a[k], &src->b[k], src->ka[k].key);
}

static void stage_free(orc_stage *s) {
    free(s->a);
    free(s->b);
    free(s->ka);
    free(s->kb);
}

int orc_ref_run_mcc(const orc_ref_config *c, const orc_mcc_params *me, const orc_mcc_params *mi,
                    int plop, double vte, double vti, int nsteps, long *ne, long cap_e,
                    orc_particle *e, uint32_t *eid, long *ni, long cap_i, orc_particle *ip,
                    uint32_t *iid, uint32_t *next_id, double *Ex, double *Ey, double *qe,
                    double *qi, double *rho, double *phi, long *events8) {
    int N = c->N;
    long capmax = cap_e > cap_i ? cap_e : cap_i;
    double *scratch = (double *)malloc(sizeof(double) * 2 * (capmax + 1));
    double L = c->L, dx = c->dx;
    for (int s = 0; s < nsteps; s++) {
        uint32_t step = (uint32_t)(c->start_step + s);
        /* e_collisions (pic.cpp:1071-1075): products appended before the push,
         * in parent-id order whatever the thread split */
        orc_stage iz = {0};
        long ev_e[8] = {0};
#pragma omp parallel
        {
            orc_stage loc = {0};
            long lev[8] = {0};
#pragma omp for schedule(static)
            for (long k = 0; k < *ne; k++) {
                double ch[6];
                int ev = orc_collide_electron(me, &e[k].vx, &e[k].vy, &e[k].vz, eid[k], step, ch);
                lev[ev]++;
                if (ev == 3) {
                    orc_particle a = e[k], b = e[k];
                    a.vx = ch[0], a.vy = ch[1], a.vz = ch[2];
                    b.vx = ch[3], b.vy = ch[4], b.vz = ch[5];
                    stage_push(&loc, &a, &b, eid[k]);
                }
            }
#pragma omp critical
            {
                stage_merge(&iz, &loc);
                for (int m = 0; m < 8; m++) ev_e[m] += lev[m];
            }
            stage_free(&loc);
        }
        for (int m = 0; m < 8; m++) events8[m] += ev_e[m];
        *ni = append_sorted(ip, iid, *ni, cap_i, iz.b, iz.kb, iz.n, &next_id[1]);
        *ne = append_sorted(e, eid, *ne, cap_e, iz.a, iz.ka, iz.n, &next_id[0]);
        stage_free(&iz);
        orc_i_gather_push(e, *ne, Ex, Ey, c, c->coef_e, c->dt_e, c->cos_we, c->sin_we, scratch);
        if (step % c->dt_dti == 0) {
            long ev_i[8] = {0};
#pragma omp parallel
            {
                long lev[8] = {0};
#pragma omp for schedule(static)
                for (long k = 0; k < *ni; k++)
                    lev[orc_collide_ion(mi, &ip[k].vx, &ip[k].vy, &ip[k].vz, iid[k], step)]++;
#pragma omp critical
                for (int m = 0; m < 8; m++) ev_i[m] += lev[m];
            }
            for (int m = 0; m < 8; m++) events8[m] += ev_i[m];
            orc_i_gather_push(ip, *ni, Ex, Ey, c, c->coef_i, c->dt_i, c->cos_wi, c->sin_wi,
                              scratch);
        }
        /* walls: x < 0 or x >= L leaves (physical mode); with plop every lost
         * ion re-injects an electron-ion pair (boundary.cpp:854-878 placement) */
        orc_stage pl = {0};
        long w = 0;
        for (long k = 0; k < *ni; k++) {
            int out = !(ip[k].x >= 0.0 && ip[k].x < L && ip[k].y >= 0.0 && ip[k].y < L);
            if (!out) {
                ip[w] = ip[k];
                iid[w] = iid[k];
                w++;
                continue;
            }
            if (plop) {
                double u0[4], u1[4], u2[4], u3[4];
                uint32_t r0[4], r1[4], r2[4], r3[4];
                draw(mi, iid[k], step, 16u + RS_PLOP, u0, r0);
                draw(mi, iid[k], step, 16u + RS_PLOP2, u1, r1);
                draw(mi, iid[k], step, 16u + RS_PLOP2 + 1u, u2, r2);
                draw(mi, iid[k], step, 16u + RS_PLOP2 + 2u, u3, r3);
                double cells = (double)(N - 3);
                double tx = 1.0 + u0[0] * cells, ty = 1.0 + u0[1] * cells;
                uint32_t wa[6] = {r1[0], r1[1], r1[2], r1[3], r2[0], r2[1]};
                uint32_t wb[6] = {r2[2], r2[3], r0[2], r0[3], r3[0], r3[1]};
                double nrm[6];
                for (int m = 0; m < 6; m++)
                    nrm[m] = sqrt(-2.0 * log((double)orc_u01(wa[m]))) *
                             cos(6.2831853 * (double)orc_u01(wb[m]));
                orc_particle pe = {tx * dx, ty * dx, vte * nrm[0], vte * nrm[1], vte * nrm[2], 1.0};
                orc_particle pi = {tx * dx, ty * dx, vti * nrm[3], vti * nrm[4], vti * nrm[5], 1.0};
                stage_push(&pl, &pe, &pi, (1ull << 32) | iid[k]);
            }
        }
        *ni = w;
        w = 0;
        for (long k = 0; k < *ne; k++) {
            int out = !(e[k].x >= 0.0 && e[k].x < L && e[k].y >= 0.0 && e[k].y < L);
            if (!out) {
                e[w] = e[k];
                eid[w] = eid[k];
                w++;
            }
        }
        *ne = w;
        *ni = append_sorted(ip, iid, *ni, cap_i, pl.b, pl.kb, pl.n, &next_id[1]);
        *ne = append_sorted(e, eid, *ne, cap_e, pl.a, pl.ka, pl.n, &next_id[0]);
        stage_free(&pl);
        orc_i_deposit_par(ip, *ni, c->qc_dxdy_i, c, qi);
        orc_ref_charge_postscale(qi, N, N);
        orc_i_deposit_par(e, *ne, c->qc_dxdy_e, c, qe);
        orc_ref_charge_postscale(qe, N, N);
        orc_ref_grid_rho(qi, qe, N, rho);
        if (orc_ref_potential_cfg(c, rho, (double)step, phi) != 0) {
            free(scratch);
            return -1;
        }
        orc_ref_grid_efield(phi, N, dx, Ex, Ey);
    }
    free(scratch);
    return 0;
}

void orc_mcc_apply(const orc_mcc_params *c, long n, double *vx, double *vy, double *vz,
                   const uint32_t *id, uint32_t step, int32_t *event, double *child6) {
    for (long k = 0; k < n; k++) {
        double ch[6] = {0, 0, 0, 0, 0, 0};
        int ev = c->electron ? orc_collide_electron(c, &vx[k], &vy[k], &vz[k], id[k], step, ch)
                             : orc_collide_ion(c, &vx[k], &vy[k], &vz[k], id[k], step);
        event[k] = ev;
        if (child6)
            for (int m = 0; m < 6; m++) child6[6 * k + m] = ch[m];
    }
}

/* ------------------------------------ reference-mode ion charge exchange */

typedef struct {
    uint64_t key; /* jp << 32 | id */
    long idx;
    float P;
} orc_cx_ent;

static int cx_cmp(const void *a, const void *b) {
    uint64_t x = ((const orc_cx_ent *)a)->key, y = ((const orc_cx_ent *)b)->key;
    return x < y ? -1 : (x > y ? 1 : 0);
}

float orc32_ref_cx_prob(const orc_refcx_params *c, float vx, float vy, float vz) {
    float E = (0.5f / c->q) * c->mass * (vx * vx + vy * vy + vz * vz);
    float u = sqrtf(2.0f * c->q * E / c->mass);
    float sig;
    if (!c->sig_energy) {
        sig = c->sig_const;
    } else if (c->gas == 0) {
        sig = (E > 4.0f) ? (2.0f + 5.5f / sqrtf(E)) * 10.0f : (-2.95f * sqrtf(E) + 10.65f) * 10.0f;
    } else {
        float e = E / 1000.0f / 1.66054e-27f;
        if (e < 1.0e-4f) {
            sig = 0.0f;
        } else {
            sig = (3.2345f * logf(235.88f / e + 2.3713f)) /
                  (1.0f + 0.038371f * e + 3.8068e-6f * powf(e, 3.5f) + 1.1832e-10f * powf(e, 5.4f));
        }
    }
    return 1.0f - expf(-(u * c->dt * 1.0e-20f * c->ng * sig));
}

long orc32_ref_cx(long n, const uint32_t *cell, const float *ox, const float *oy, float *vx,
                  float *vy, float *vz, const uint32_t *id, int N, const orc_refcx_params *c,
                  float *pcol, const float *ntab, int ntab_n, uint32_t step) {
    orc_cx_ent *v = (orc_cx_ent *)malloc(sizeof(orc_cx_ent) * (size_t)(n + 1));
    long m = 0;
    for (long k = 0; k < n; k++) {
        int jx = orc_cell_jx(cell[k]), jy = orc_cell_jy(cell[k]);
        int inx = jx >= -1 && jx <= N - 1 && !(jx == -1 && ox[k] == 0.0f);
        int iny = jy >= -1 && jy <= N - 1 && !(jy == -1 && oy[k] == 0.0f);
        if (!inx || !iny) continue;
        uint32_t jp = jx < 0 ? 0u : (uint32_t)jx;
        v[m].key = ((uint64_t)jp << 32) | id[k];
        v[m].idx = k;
        v[m].P = orc32_ref_cx_prob(c, vx[k], vy[k], vz[k]);
        m++;
    }
    if (m > 1) qsort(v, (size_t)m, sizeof(orc_cx_ent), cx_cmp);
    long fired = 0;
    const uint32_t key[2] = {c->key0, c->key1};
    for (long t = 0; t < m; t++) {
        uint32_t jp = (uint32_t)(v[t].key >> 32);
        pcol[jp] = pcol[jp] + v[t].P;
        if (pcol[jp] > 1.0f) {
            uint32_t ctr[4] = {(uint32_t)v[t].key, step, 16u + 9u, 0x4D434331u}, r[4];
            orc_philox4x32_10(ctr, key, r);
            int ix = (int)(orc_u01(r[0]) * (float)(ntab_n - 1));
            ix = ix < 0 ? 0 : (ix > ntab_n - 2 ? ntab_n - 2 : ix);
            long k = v[t].idx;
            vx[k] = ntab[ix];
            vy[k] = ntab[ntab_n + ix];
            vz[k] = ntab[2 * ntab_n + ix];
            pcol[jp] = 0.0f;
            fired++;
        }
    }
    free(v);
    return fired;
}

void orc_neutral_table(uint64_t seed, double vtn, float *t) {
    const int n = 10000;
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int i = 0; i < n; i++) {
        uint32_t c0[4] = {(uint32_t)i, 0x7AB1Eu, 0u, 0x4E455554u}, r[4];
        orc_philox4x32_10(c0, key, r);
        double R1 = orc_u01d(r[0], r[1]), R2 = orc_u01d(r[2], r[3]);
        t[i] = (float)(vtn * cos(3.141592654 * R2) * sqrt(fabs(log(R1))));
        uint32_t c1[4] = {(uint32_t)i, 0x7AB1Eu, 1u, 0x4E455554u};
        orc_philox4x32_10(c1, key, r);
        R1 = orc_u01d(r[0], r[1]);
        R2 = orc_u01d(r[2], r[3]);
        t[n + i] = (float)(vtn * cos(3.141592654 * R2) * sqrt(fabs(log(R1))));
        t[2 * n + i] = (float)(vtn * sin(3.141592654 * R2) * sqrt(fabs(log(R1))));
    }
}
