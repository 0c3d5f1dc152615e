/*
 * orc_step.c -- whole-timestep orchestration of the oracle
 * (TEST INFRASTRUCTURE ONLY; also the CPU baseline timed by bench.py).
 *
 * Phase order of one step follows pic::pic_run (pic.cpp:1069-1162) for the
 * collisionless case: electron gather + push, ion gather + push on ion steps,
 * ion then electron wall pass, ion then electron deposit, grid_rho,
 * grid_potential, grid_electric_field.  The field gathered at step s is the
 * one solved at the end of step s-1 (zero before the first solve).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------- FP64 reference */



int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* parallel gather + push over chunks (the reference's kernels are per-particle
 * independent, so any split gives the serial result bit for bit) */
void orc_i_gather_push(orc_particle *p, long n, const double *Ex, const double *Ey,
                            const orc_ref_config *c, double coef, double dt, double cw, double sw,
                            double *scratch) {
    double *Exp = scratch, *Eyp = scratch + n;
#pragma omp parallel
    {
        long nt = 1, t = 0;
#ifdef _OPENMP
        nt = omp_get_num_threads();
        t = omp_get_thread_num();
#endif
        long b = n * t / nt, e = n * (t + 1) / nt;
        orc_ref_gather(p + b, e - b, Ex, Ey, c->dx, c->dy, c->N, c->N, Exp + b, Eyp + b);
        orc_ref_push(p + b, e - b, Exp + b, Eyp + b, c->dx, c->dy, c->N, c->N, coef, c->cos_a,
                     c->sin_a, dt, cw, sw, c->bfield);
    }
}

/* deposit with per-thread private grids summed in thread order */
void orc_i_deposit_par(const orc_particle *p, long n, double qc_dxdy, const orc_ref_config *c,
                            double *q) {
    int N = c->N;
    memset(q, 0, sizeof(double) * N * N);
#ifdef _OPENMP
    int nt = omp_get_max_threads();
#else
    int nt = 1;
#endif
    if (nt == 1) {
        orc_ref_deposit(p, n, qc_dxdy, c->dx, c->dy, N, N, q);
        return;
    }
    double *priv = (double *)calloc((size_t)nt * N * N, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        long b = n * t / nt, e = n * (t + 1) / nt;
        orc_ref_deposit(p + b, e - b, qc_dxdy, c->dx, c->dy, N, N, priv + (size_t)t * N * N);
    }
    for (int t = 0; t < nt; t++)
        for (long k = 0; k < (long)N * N; k++) q[k] += priv[(size_t)t * N * N + k];
    free(priv);
}

int orc_ref_run(const orc_ref_config *c, int nsteps, long *ne, orc_particle *e, uint32_t *eid,
                long *ni, orc_particle *ip, uint32_t *iid, double *Ex, double *Ey, double *qe,
                double *qi, double *rho, double *phi) {
    int N = c->N;
    long cap = (*ne > *ni ? *ne : *ni);
    double *scratch = (double *)malloc(sizeof(double) * 2 * (cap > 0 ? cap : 1));
    for (int s = 0; s < nsteps; s++) {
        long step = c->start_step + s;
        orc_i_gather_push(e, *ne, Ex, Ey, c, c->coef_e, c->dt_e, c->cos_we, c->sin_we, scratch);
        if (step % c->dt_dti == 0) {
            orc_i_gather_push(ip, *ni, Ex, Ey, c, c->coef_i, c->dt_i, c->cos_wi, c->sin_wi, scratch);
            *ni = orc_ref_absorb(ip, iid, *ni, c->dx, c->dy, N, N);
        }
        *ne = orc_ref_absorb(e, eid, *ne, c->dx, c->dy, N, N);
        orc_i_deposit_par(ip, *ni, c->qc_dxdy_i, c, qi);
        orc_ref_charge_postscale(qi, N, N);
        orc_i_deposit_par(e, *ne, c->qc_dxdy_e, c, qe);
        orc_ref_charge_postscale(qe, N, N);
        orc_ref_grid_rho(qi, qe, N, rho);
        if (orc_ref_potential_cfg(c, rho, (double)step, phi) != 0) {
            free(scratch);
            return -1;
        }
        orc_ref_grid_efield(phi, N, c->dx, Ex, Ey);
    }
    free(scratch);
    return 0;
}

/* --------------------------------------------------- build layout (fp32) */



typedef struct {
    long *n;
    uint32_t *cell;
    float *ox, *oy, *vx, *vy, *vz;
    uint32_t *id;
} sp32;

/* stable removal of particles whose status is not KEPT; counts per wall */
static void compact32(sp32 *s, const uint8_t *st, long *lost4) {
    long w = 0, n = *s->n;
    for (long r = 0; r < n; r++) {
        if (st[r] != ORC_ST_KEPT) {
            lost4[st[r] - 1]++;
            continue;
        }
        s->cell[w] = s->cell[r];
        s->ox[w] = s->ox[r];
        s->oy[w] = s->oy[r];
        s->vx[w] = s->vx[r];
        s->vy[w] = s->vy[r];
        s->vz[w] = s->vz[r];
        s->id[w] = s->id[r];
        w++;
    }
    *s->n = w;
}

/* wall classification without moving (the pre-loop pass, pic.cpp:1049-1057) */
static void classify32(const sp32 *s, int N, int mode, uint8_t *st) {
    orc32_push_params pp;
    memset(&pp, 0, sizeof(pp));
    pp.N = N;
    pp.mode = mode;
    for (long g = 0; g < *s->n; g++) {
        int jx = orc_cell_jx(s->cell[g]), jy = orc_cell_jy(s->cell[g]);
        float fx = s->ox[g], fy = s->oy[g];
        int lx, ly;
        if (mode == ORC_MODE_REF_STEADY) {
            lx = ly = 0;
        } else if (mode == ORC_MODE_PHYS) {
            lx = (jx < 0) || (jx > N - 2);
            ly = (jy < 0) || (jy > N - 2);
        } else {
            lx = (jx < -1) || (jx == -1 && fx == 0.0f) || (jx >= N);
            ly = (jy < -1) || (jy == -1 && fy == 0.0f) || (jy >= N);
        }
        st[g] = lx ? ((jx < 0) ? ORC_ST_LOST_LEFT : ORC_ST_LOST_RIGHT)
                   : (ly ? ((jy < 0) ? ORC_ST_LOST_BOTTOM : ORC_ST_LOST_TOP) : ORC_ST_KEPT);
    }
}

int orc32_run(const orc32_config *c, int nsteps, long *ne, uint32_t *ecell, float *eox, float *eoy,
              float *evx, float *evy, float *evz, uint32_t *eid, long *ni, uint32_t *icell,
              float *iox, float *ioy, float *ivx, float *ivy, float *ivz, uint32_t *iid,
              float *E32, int64_t *qe, int64_t *qi, double *rho, double *phi, long *lost8,
              int *mg_cycles_out) {
    int N = c->N;
    sp32 se = {ne, ecell, eox, eoy, evx, evy, evz, eid};
    sp32 si = {ni, icell, iox, ioy, ivx, ivy, ivz, iid};
    long cap = (*ne > *ni ? *ne : *ni) + 1;
    uint8_t *st = (uint8_t *)malloc((size_t)cap);
    double *rhs = (double *)malloc(sizeof(double) * N * N);
    int rc = 0;
    const int steady = c->mode == ORC_MODE_REF_STEADY;
    if (c->preloop) {
        if (steady) {
            orc32_steady(*ne, ecell, eox, eoy, evx, evy, evz, eid, *ni, icell, iox, ioy, ivx, ivy,
                         ivz, iid, N, (uint32_t)c->start_step, c->key0, c->key1, c->steady_se,
                         c->steady_si);
        } else {
            classify32(&si, N, c->mode, st);
            compact32(&si, st, lost8 + 4);
            classify32(&se, N, c->mode, st);
            compact32(&se, st, lost8);
        }
    }
    for (int s = 0; s < nsteps; s++) {
        long step = c->start_step + s;
        orc32_push(*ne, ecell, eox, eoy, evx, evy, evz, E32, &c->pe, st);
        compact32(&se, st, lost8);
        if (step % c->dt_dti == 0) {
            orc32_push(*ni, icell, iox, ioy, ivx, ivy, ivz, E32, &c->pi, st);
            compact32(&si, st, lost8 + 4);
        }
        if (steady)
            orc32_steady(*ne, ecell, eox, eoy, evx, evy, evz, eid, *ni, icell, iox, ioy, ivx, ivy,
                         ivz, iid, N, (uint32_t)step, c->key0, c->key1, c->steady_se,
                         c->steady_si);
        memset(qe, 0, sizeof(int64_t) * N * N);
        memset(qi, 0, sizeof(int64_t) * N * N);
        orc32_deposit(*ni, icell, iox, ioy, N, c->mode, qi);
        orc32_deposit(*ne, ecell, eox, eoy, N, c->mode, qe);
        double phiR = c->phi_right * cos(c->phi_period * (double)step);
        double h = c->dx;
        orc_fixed_to_rhs(qi, qe, N, c->scale_i, c->scale_e, c->corner8, c->eps, c->phi_left, phiR,
                         c->phi_bottom, c->phi_top, 1.0 / (h * h), rho, rhs);
        int cyc = 0;
        if (c->mg_mode == 0) {
            if (orc_mglin(rhs, N, c->mg_ncycle, c->L) != 0) {
                rc = -1;
                break;
            }
        } else {
            cyc = orc_mg_tol(rhs, N, c->L, c->mg_tol, c->mg_max_cycles, NULL);
            if (cyc < 0) {
                rc = -1;
                break;
            }
        }
        if (mg_cycles_out) mg_cycles_out[s] = cyc;
        for (int j = 0; j < N; j++) {
            rhs[j * N + 0] = c->phi_left;
            rhs[j * N + N - 1] = phiR;
        }
        for (int i = 0; i < N; i++) {
            rhs[i] = c->phi_bottom;
            rhs[(N - 1) * N + i] = c->phi_top;
        }
        memcpy(phi, rhs, sizeof(double) * N * N);
        orc_efield_pack(phi, N, c->dx, NULL, NULL, E32);
    }
    free(st);
    free(rhs);
    return rc;
}

/* ------------------------------------------- reference steady-state walls */

typedef struct {
    int side;
    uint32_t id;
    long idx;
} orc_cand;

static int cand_cmp(const void *a, const void *b) {
    const orc_cand *x = (const orc_cand *)a, *y = (const orc_cand *)b;
    if (x->side != y->side) return x->side < y->side ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id ? 1 : 0);
}

/* candidates of one species sorted by (side, id); sides as boundary.cpp:906-912 */
static long collect_cands(long n, const uint32_t *cell, const uint32_t *id, int N, orc_cand *out) {
    long m = 0;
    for (long k = 0; k < n; k++) {
        int jx = orc_cell_jx(cell[k]), jy = orc_cell_jy(cell[k]);
        if (jx < 0) out[m++] = (orc_cand){0, id[k], k};
        else if (jx >= N) out[m++] = (orc_cand){1, id[k], k};
        if (jy < 0) out[m++] = (orc_cand){2, id[k], k};
        else if (jy >= N) out[m++] = (orc_cand){3, id[k], k};
    }
    if (m > 1) qsort(out, (size_t)m, sizeof(orc_cand), cand_cmp);
    return m;
}

static void plop32(uint32_t ion_id, uint32_t step, int N, uint32_t key0, uint32_t key1, int *jx,
                   int *jy, float *ox, float *oy, float nrm[6]) {
    const uint32_t key[2] = {key0, key1};
    uint32_t r[4], a[4], b[4], d[4];
    uint32_t c0[4] = {ion_id, step, 16u + 3u, 0x4D434331u};
    uint32_t c1[4] = {ion_id, step, 16u + 4u, 0x4D434331u};
    uint32_t c2[4] = {ion_id, step, 16u + 5u, 0x4D434331u};
    uint32_t c3[4] = {ion_id, step, 16u + 6u, 0x4D434331u};
    orc_philox4x32_10(c0, key, r);
    orc_philox4x32_10(c1, key, a);
    orc_philox4x32_10(c2, key, b);
    orc_philox4x32_10(c3, key, d);
    float span = (float)(N - 3);
    float tx = 1.0f + orc_u01(r[0]) * span, ty = 1.0f + orc_u01(r[1]) * span;
    int ix = (int)floorf(tx), iy = (int)floorf(ty);
    float fx = tx - (float)ix, fy = ty - (float)iy;
    if (fx >= 1.0f) { fx = 0.0f; ix += 1; }
    if (fy >= 1.0f) { fy = 0.0f; iy += 1; }
    *jx = ix;
    *jy = iy;
    *ox = fx;
    *oy = fy;
    uint32_t w[6] = {a[0], a[1], a[2], a[3], b[0], b[1]};
    uint32_t w2[6] = {b[2], b[3], r[2], r[3], d[0], d[1]};
    for (int k = 0; k < 6; k++) {
        float R1 = orc_u01(w[k]), R2 = orc_u01(w2[k]);
        nrm[k] = cosf(6.2831853f * R2) * sqrtf(fabsf(logf(R1)));
    }
}

long orc32_steady(long ne, uint32_t *ecell, float *eox, float *eoy, float *evx, float *evy,
                  float *evz, const uint32_t *eid, long ni, uint32_t *icell, float *iox,
                  float *ioy, float *ivx, float *ivy, float *ivz, const uint32_t *iid, int N,
                  uint32_t step, uint32_t key0, uint32_t key1, float se, float si) {
    orc_cand *ce = (orc_cand *)malloc(sizeof(orc_cand) * (2 * ne + 1));
    orc_cand *ci = (orc_cand *)malloc(sizeof(orc_cand) * (2 * ni + 1));
    uint8_t *pe = (uint8_t *)calloc((size_t)ne + 1, 1), *pi = (uint8_t *)calloc((size_t)ni + 1, 1);
    long me = collect_cands(ne, ecell, eid, N, ce), mi = collect_cands(ni, icell, iid, N, ci);
    long pairs = 0;
    for (int side = 0; side < 4; side++) {
        /* entries of this side still outside (y sides: not re-injected by x) */
        long ke = 0, ki = 0;
        long *le = (long *)malloc(sizeof(long) * (me + 1)), *li = (long *)malloc(sizeof(long) * (mi + 1));
        for (long k = 0; k < me; k++)
            if (ce[k].side == side && !pe[ce[k].idx]) le[ke++] = ce[k].idx;
        for (long k = 0; k < mi; k++)
            if (ci[k].side == side && !pi[ci[k].idx]) li[ki++] = ci[k].idx;
        long m = ke < ki ? ke : ki;
        for (long t = 0; t < m; t++) {
            long a = le[t], b = li[t];
            int jx, jy;
            float ox, oy, nrm[6];
            plop32(iid[b], step, N, key0, key1, &jx, &jy, &ox, &oy, nrm);
            ecell[a] = icell[b] = orc_cell_pack(jx, jy);
            eox[a] = iox[b] = ox;
            eoy[a] = ioy[b] = oy;
            evx[a] = se * nrm[0];
            evy[a] = se * nrm[1];
            evz[a] = se * nrm[2];
            ivx[b] = si * nrm[3];
            ivy[b] = si * nrm[4];
            ivz[b] = si * nrm[5];
            pe[a] = pi[b] = 1;
        }
        pairs += m;
        free(le);
        free(li);
    }
    free(ce);
    free(ci);
    free(pe);
    free(pi);
    return pairs;
}
