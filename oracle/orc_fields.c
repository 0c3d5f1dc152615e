/*
 * orc_fields.c -- FP64 field solve of the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Restates fields.cpp (grid_rho, grid_potential, grid_electric_field) and the
 * full-multigrid solver of mglin.cpp.  The multigrid keeps the reference's
 * floating-point association inside every stencil expression, so with
 * -ffp-contract=off its output is bit-identical to the reference's Mglin
 * (checked against the reference's own mglin.cpp compiled into oracle/_ref).
 * The data structures are this file's own: flat row-major arrays, one
 * allocation per level for the whole solve.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- helpers */

/* number of levels ng with n - 1 == 2^ng (mglin.cpp:24-27); -1 if invalid */
static int mg_levels(int n) {
    int ng = 0, nn = n;
    while (nn >>= 1) ng++;
    if (n < 3 || (n - 1) != (1 << ng)) return -1;
    return ng;
}

static int level_size(int j) { return (1 << (j + 1)) + 1; }

/* mglin.cpp:70-82: copies on even/even, vertical average on even columns,
 * then horizontal average on odd columns (which reads the even columns). */
static void mg_interp(double *uf, int nf, const double *uc) {
    int nc = nf / 2 + 1;
    for (int jc = 0; jc < nc; jc++)
        for (int ic = 0; ic < nc; ic++) uf[(2 * ic) * nf + 2 * jc] = uc[ic * nc + jc];
    for (int jf = 0; jf < nf; jf += 2)
        for (int i = 1; i < nf - 1; i += 2)
            uf[i * nf + jf] = 0.5 * (uf[(i + 1) * nf + jf] + uf[(i - 1) * nf + jf]);
    for (int jf = 1; jf < nf - 1; jf += 2)
        for (int i = 0; i < nf; i++)
            uf[i * nf + jf] = 0.5 * (uf[i * nf + jf + 1] + uf[i * nf + jf - 1]);
}

/* mglin.cpp:103-110: 3x3 grid, h = length/2, only the centre is unknown */
static void mg_slvsml(double *u, const double *rhs, double length) {
    double h = 0.5 * length;
    for (int k = 0; k < 9; k++) u[k] = 0.0;
    u[4] = -h * h * rhs[4] / 4.0;
}

/* mglin.cpp:116-127: red ((i+j) even) then black sweep over interior nodes */
static void mg_relax(double *u, const double *rhs, int n, double length) {
    double h = length / (n - 1);
    double h2 = h * h;
    for (int color = 0; color < 2; color++) {
        for (int j = 1; j < n - 1; j++) {
            int i0 = ((j + color) & 1) ? 1 : 2; /* color 0: i + j even */
            for (int i = i0; i < n - 1; i += 2)
                u[i * n + j] = 0.25 * (u[(i + 1) * n + j] + u[(i - 1) * n + j] + u[i * n + j + 1] +
                                       u[i * n + j - 1] - h2 * rhs[i * n + j]);
        }
    }
}

/* mglin.cpp:133-144: res = -(1/h^2)(sum nbr - 4u) + rhs, zero boundary */
static void mg_resid(double *res, const double *u, const double *rhs, int n, double length) {
    double h = length / (n - 1);
    double h2i = 1.0 / (h * h);
    for (int j = 1; j < n - 1; j++)
        for (int i = 1; i < n - 1; i++)
            res[i * n + j] = -h2i * (u[(i + 1) * n + j] + u[(i - 1) * n + j] + u[i * n + j + 1] +
                                     u[i * n + j - 1] - 4.0 * u[i * n + j]) +
                             rhs[i * n + j];
    for (int i = 0; i < n; i++) {
        res[i * n + 0] = 0.0;
        res[i * n + n - 1] = 0.0;
        res[0 * n + i] = 0.0;
        res[(n - 1) * n + i] = 0.0;
    }
}

/* mglin.cpp:151-169: half weighting inside, injection on the boundary */
static void mg_rstrct(double *uc, int nc, const double *uf) {
    int nf = 2 * nc - 1, ncc = 2 * nc - 2;
    for (int jc = 1, jf = 2; jc < nc - 1; jc++, jf += 2)
        for (int ic = 1, iff = 2; ic < nc - 1; ic++, iff += 2)
            uc[ic * nc + jc] = 0.5 * uf[iff * nf + jf] +
                               0.125 * (uf[(iff + 1) * nf + jf] + uf[(iff - 1) * nf + jf] +
                                        uf[iff * nf + jf + 1] + uf[iff * nf + jf - 1]);
    for (int ic = 0, jf = 0; ic < nc; ic++, jf += 2) {
        uc[ic * nc + 0] = uf[jf * nf + 0];
        uc[ic * nc + nc - 1] = uf[jf * nf + ncc];
    }
    for (int ic = 0, jf = 0; ic < nc; ic++, jf += 2) {
        uc[0 * nc + ic] = uf[0 * nf + jf];
        uc[(nc - 1) * nc + ic] = uf[ncc * nf + jf];
    }
}

typedef struct {
    int ng;
    double length;
    double *rhs[32];  /* rhs per level (level ng-1 = finest)             */
    double *tmp[32];  /* fine-level scratch (resid output / interp)      */
    double *cres[32]; /* coarse rhs built by restriction in the V-cycle  */
    double *cv[32];   /* coarse correction                               */
} mg_ws;

/* mglin.cpp:176-196: one V(1,1) cycle at level j (recursive) */
static void mg_vcycle(mg_ws *w, int j, double *u, const double *rhs) {
    int nf = level_size(j);
    if (j == 0) {
        mg_slvsml(u, rhs, w->length);
        return;
    }
    int nc = (nf + 1) / 2;
    double *temp = w->tmp[j], *res = w->cres[j - 1], *v = w->cv[j - 1];
    mg_relax(u, rhs, nf, w->length);
    mg_resid(temp, u, rhs, nf, w->length);
    mg_rstrct(res, nc, temp);
    memset(v, 0, sizeof(double) * nc * nc);
    mg_vcycle(w, j - 1, v, res);
    mg_interp(temp, nf, v); /* addint: interp into scratch, then add */
    for (int k = 0; k < nf * nf; k++) u[k] += temp[k];
    mg_relax(u, rhs, nf, w->length);
}

static int mg_ws_alloc(mg_ws *w, int n, double length) {
    memset(w, 0, sizeof(*w));
    w->ng = mg_levels(n);
    if (w->ng < 0) return -1;
    w->length = length;
    for (int j = 0; j < w->ng; j++) {
        size_t s = (size_t)level_size(j) * level_size(j);
        w->rhs[j] = (double *)calloc(s, sizeof(double));
        w->tmp[j] = (double *)calloc(s, sizeof(double));
        w->cres[j] = (double *)calloc(s, sizeof(double));
        w->cv[j] = (double *)calloc(s, sizeof(double));
    }
    return 0;
}

static void mg_ws_free(mg_ws *w) {
    for (int j = 0; j < w->ng; j++) {
        free(w->rhs[j]);
        free(w->tmp[j]);
        free(w->cres[j]);
        free(w->cv[j]);
    }
}

/* FMG driver (mglin.cpp:21-55).  On exit `u` holds the level-(ng-1) solution. */
static void mg_fmg(mg_ws *w, double *u, int ncycle) {
    int ng = w->ng;
    int n = level_size(ng - 1);
    memcpy(w->rhs[ng - 1], u, sizeof(double) * n * n);
    for (int j = ng - 1; j > 0; j--) mg_rstrct(w->rhs[j - 1], level_size(j - 1), w->rhs[j]);
    double *cur = (double *)calloc(9, sizeof(double));
    mg_slvsml(cur, w->rhs[0], w->length);
    for (int j = 1; j < ng; j++) {
        int nf = level_size(j);
        double *next = (double *)calloc((size_t)nf * nf, sizeof(double));
        mg_interp(next, nf, cur);
        free(cur);
        cur = next;
        for (int c = 0; c < ncycle; c++) mg_vcycle(w, j, cur, w->rhs[j]);
    }
    memcpy(u, cur, sizeof(double) * n * n);
    free(cur);
}

int orc_mglin(double *u, int n, int ncycle, double length) {
    mg_ws w;
    if (mg_ws_alloc(&w, n, length) != 0) return -1;
    mg_fmg(&w, u, ncycle);
    mg_ws_free(&w);
    return 0;
}

double orc_mg_relres(const double *u, const double *f, int n, double length) {
    double h = length / (n - 1);
    double h2i = 1.0 / (h * h);
    double rr = 0.0, ff = 0.0;
    for (int i = 1; i < n - 1; i++)
        for (int j = 1; j < n - 1; j++) {
            double au = h2i * (u[(i + 1) * n + j] + u[(i - 1) * n + j] + u[i * n + j + 1] +
                               u[i * n + j - 1] - 4.0 * u[i * n + j]);
            double r = f[i * n + j] - au;
            rr += r * r;
            ff += f[i * n + j] * f[i * n + j];
        }
    if (ff == 0.0) return sqrt(rr);
    return sqrt(rr / ff);
}

int orc_mg_tol(double *u_rhs, int n, double length, double tol, int max_cycles, double *rel_res_out) {
    mg_ws w;
    if (mg_ws_alloc(&w, n, length) != 0) return -1;
    double *f = (double *)malloc(sizeof(double) * n * n);
    memcpy(f, u_rhs, sizeof(double) * n * n);
    mg_fmg(&w, u_rhs, 2);
    int c = 0;
    double rel = orc_mg_relres(u_rhs, f, n, length);
    while (rel > tol && c < max_cycles) {
        mg_vcycle(&w, w.ng - 1, u_rhs, f);
        rel = orc_mg_relres(u_rhs, f, n, length);
        c++;
    }
    if (rel_res_out) *rel_res_out = rel;
    free(f);
    mg_ws_free(&w);
    return c;
}

/* ------------------------------------------------------------ fields.cpp */

void orc_ref_grid_rho(const double *q_ion, const double *q_e, int n, double *rho) {
    for (int k = 0; k < n * n; k++) rho[k] = q_ion[k];
    for (int k = 0; k < n * n; k++) rho[k] += q_e[k];
}

int orc_ref_grid_potential(const double *rho, int n, double dx, double dy, double eps,
                           double phi0X, double phiLX, double phi0Y, double phiLY,
                           double timestep, double phi_period, int ncycle, double length,
                           double *phi) {
    double coef = 1.0 / eps;
    double phi_right = phiLX * cos(phi_period * timestep);
    double divdxdy = 1.0 / (dx * dy);
    for (int k = 0; k < n * n; k++) phi[k] = -coef * rho[k];
    for (int j = 0; j < n; j++) {
        phi[j * n + 1] = phi[j * n + 1] - phi0X * divdxdy;
        phi[j * n + n - 2] = phi[j * n + n - 2] - phi_right * divdxdy;
    }
    for (int i = 0; i < n; i++) {
        phi[1 * n + i] = phi[1 * n + i] - phi0Y * divdxdy;
        phi[(n - 2) * n + i] = phi[(n - 2) * n + i] - phiLY * divdxdy;
    }
    if (orc_mglin(phi, n, ncycle, length) != 0) return -1;
    for (int j = 0; j < n; j++) {
        phi[j * n + 0] = phi0X;
        phi[j * n + n - 1] = phi_right;
    }
    for (int i = 0; i < n; i++) {
        phi[i] = phi0Y;
        phi[(n - 1) * n + i] = phiLY;
    }
    return 0;
}

int orc_ref_potential_cfg(const orc_ref_config *c, const double *rho, double timestep,
                          double *phi) {
    if (!(c->mg_tol > 0.0))
        return orc_ref_grid_potential(rho, c->N, c->dx, c->dy, c->eps, c->phi_left, c->phi_right,
                                      c->phi_bottom, c->phi_top, timestep, c->phi_period,
                                      c->ncycle, c->L, phi);
    /* the folded right-hand side and boundary values of fields.cpp:169-220,
     * with the converged solve in place of Mglin's fixed cycles */
    const int n = c->N;
    const double coef = 1.0 / c->eps, divdxdy = 1.0 / (c->dx * c->dy);
    const double phi_right = c->phi_right * cos(c->phi_period * timestep);
    for (int k = 0; k < n * n; k++) phi[k] = -coef * rho[k];
    for (int j = 0; j < n; j++) {
        phi[j * n + 1] = phi[j * n + 1] - c->phi_left * divdxdy;
        phi[j * n + n - 2] = phi[j * n + n - 2] - phi_right * divdxdy;
    }
    for (int i = 0; i < n; i++) {
        phi[1 * n + i] = phi[1 * n + i] - c->phi_bottom * divdxdy;
        phi[(n - 2) * n + i] = phi[(n - 2) * n + i] - c->phi_top * divdxdy;
    }
    if (orc_mg_tol(phi, n, c->L, c->mg_tol, 1000, NULL) < 0) return -1;
    for (int j = 0; j < n; j++) {
        phi[j * n + 0] = c->phi_left;
        phi[j * n + n - 1] = phi_right;
    }
    for (int i = 0; i < n; i++) {
        phi[i] = c->phi_bottom;
        phi[(n - 1) * n + i] = c->phi_top;
    }
    return 0;
}

void orc_ref_grid_efield(const double *phi, int n, double dx, double *Ex, double *Ey) {
    double twoDX = 2.0 * dx;
    for (int j = 1; j < n - 1; j++)
        for (int i = 1; i < n - 1; i++) {
            Ex[j * n + i] = (phi[j * n + i - 1] - phi[j * n + i + 1]) / twoDX;
            Ey[j * n + i] = (phi[(j - 1) * n + i] - phi[(j + 1) * n + i]) / twoDX;
        }
    for (int j = 0; j < n; j++) {
        Ex[j * n] = 2.0 * Ex[j * n + 1] - Ex[j * n + 2];
        Ex[j * n + n - 1] = 2.0 * Ex[j * n + n - 2] - Ex[j * n + n - 3];
        Ex[j * n + n] = 0.0;
    }
    for (int i = 0; i < n; i++) {
        Ey[i] = 2.0 * Ey[n + i] - Ey[2 * n + i];
        Ey[(n - 1) * n + i] = 2.0 * Ey[(n - 2) * n + i] - Ey[(n - 3) * n + i];
        Ey[n * n + i] = 0.0;
    }
}

/* ------------------------------------------------- build-layout field glue */

void orc_fixed_to_rhs(const int64_t *q_ion, const int64_t *q_e, int n, double scale_i,
                      double scale_e, int corner8, double eps, double phi0X, double phiR,
                      double phi0Y, double phiLY, double dxdy_inv, double *rho, double *rhs) {
    double coef = 1.0 / eps;
    double corner = corner8 ? 8.0 : 4.0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            int ex = (i == 0 || i == n - 1), ey = (j == 0 || j == n - 1);
            double fac = (ex && ey) ? corner : ((ex || ey) ? 2.0 : 1.0);
            int k = j * n + i;
            double ri = ((double)q_ion[k] * scale_i) * fac;
            double re = ((double)q_e[k] * scale_e) * fac;
            double r = ri + re;
            double f = -coef * r;
            if (i == 1) f = f - phi0X * dxdy_inv;
            if (i == n - 2) f = f - phiR * dxdy_inv;
            if (j == 1) f = f - phi0Y * dxdy_inv;
            if (j == n - 2) f = f - phiLY * dxdy_inv;
            rho[k] = r;
            rhs[k] = f;
        }
}

/* E at node (j,i) from phi with the reference's stencil and wall
 * extrapolation; entries the reference never writes are 0. */
static void efield_node(const double *phi, int n, double twoDX, int j, int i, double *ex, double *ey) {
    /* interior central differences as fields.cpp:240-247 */
#define EXI(jj, ii) (((jj) >= 1 && (jj) <= n - 2 && (ii) >= 1 && (ii) <= n - 2) \
                         ? (phi[(jj) * n + (ii) - 1] - phi[(jj) * n + (ii) + 1]) / twoDX : 0.0)
#define EYI(jj, ii) (((jj) >= 1 && (jj) <= n - 2 && (ii) >= 1 && (ii) <= n - 2) \
                         ? (phi[((jj) - 1) * n + (ii)] - phi[((jj) + 1) * n + (ii)]) / twoDX : 0.0)
    if (i == 0)
        *ex = 2.0 * EXI(j, 1) - EXI(j, 2);
    else if (i == n - 1)
        *ex = 2.0 * EXI(j, n - 2) - EXI(j, n - 3);
    else
        *ex = EXI(j, i);
    if (j == 0)
        *ey = 2.0 * EYI(1, i) - EYI(2, i);
    else if (j == n - 1)
        *ey = 2.0 * EYI(n - 2, i) - EYI(n - 3, i);
    else
        *ey = EYI(j, i);
#undef EXI
#undef EYI
}

void orc_efield_pack(const double *phi, int n, double dx, double *Ex64, double *Ey64, float *E32) {
    double twoDX = 2.0 * dx;
    long tot = (long)(n + 1) * n + 1;
    for (long k = 0; k < tot; k++) {
        E32[2 * k] = 0.0f;
        E32[2 * k + 1] = 0.0f;
    }
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            double ex, ey;
            efield_node(phi, n, twoDX, j, i, &ex, &ey);
            long k = (long)j * n + i;
            if (Ex64) Ex64[k] = ex;
            if (Ey64) Ey64[k] = ey;
            E32[2 * k] = (float)ex;
            E32[2 * k + 1] = (float)ey;
        }
}
