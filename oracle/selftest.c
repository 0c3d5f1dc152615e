/*
 * selftest.c -- drives the oracle's entry points on small inputs so that a
 * build with -fsanitize=address,undefined (make -C oracle selftest) checks
 * them for out-of-bounds accesses, leaks and undefined behaviour
 * (TEST INFRASTRUCTURE ONLY; tests/test_oracle_sanitized.py runs it).
 *
 * Covered: Philox streams, the cross sections, the FP64 multigrid (fixed
 * cycles and to a tolerance), the build-layout step loop in reference and
 * physical mode with the fixed and tolerance solves, the reference-layout
 * step loop, the FP64 Monte-Carlo step loop with pair re-injection, the
 * steady-wall pairing, the source/reduction/weighted-deposit control, the
 * surface physics of the walls and dataDump text I/O.  Exit status 0: every call
 * returned success and every result is finite.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int failures = 0;
#define CHECK(cond)                                                        \
    do {                                                                   \
        if (!(cond)) {                                                     \
            fprintf(stderr, "selftest: %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            failures++;                                                    \
        }                                                                  \
    } while (0)

static uint32_t lcg = 12345u;
static double urand(void) {
    lcg = lcg * 1664525u + 1013904223u;
    return ((lcg >> 8) + 0.5) / 16777216.0;
}

static void state32(long n, int N, double sigma, uint32_t *cell, float *ox, float *oy, float *vx,
                    float *vy, float *vz, uint32_t *id) {
    for (long k = 0; k < n; k++) {
        double tx = urand() * (N - 1), ty = urand() * (N - 1);
        int jx = (int)floor(tx), jy = (int)floor(ty);
        cell[k] = orc_cell_pack(jx, jy);
        ox[k] = (float)(tx - jx);
        oy[k] = (float)(ty - jy);
        if (ox[k] >= 1.0f) ox[k] = 0.0f;
        if (oy[k] >= 1.0f) oy[k] = 0.0f;
        vx[k] = (float)(sigma * (urand() - 0.5));
        vy[k] = (float)(sigma * (urand() - 0.5));
        vz[k] = (float)(sigma * (urand() - 0.5));
        id[k] = (uint32_t)k;
    }
}

static void run32(int mode, int mg_mode) {
    const int N = 17, steps = 3;
    const long n = 600, cap = n;
    orc32_config c;
    memset(&c, 0, sizeof(c));
    c.N = N;
    c.mode = mode;
    c.corner8 = mode == ORC_MODE_REF;
    c.mg_mode = mg_mode;
    c.mg_ncycle = 2;
    c.mg_max_cycles = 40;
    c.dt_dti = 2;
    c.preloop = 1;
    c.L = 0.01;
    c.dx = c.L / (N - 1);
    c.eps = 8.854e-12;
    c.mg_tol = 1e-9;
    c.phi_left = 1.0;
    c.phi_right = -1.0;
    c.phi_period = 0.2;
    c.scale_e = ldexp(-1.6e-19 * 1e5 / (c.dx * c.dx), -48);
    c.scale_i = -c.scale_e;
    c.pe.N = c.pi.N = N;
    c.pe.mode = c.pi.mode = mode;
    c.pe.cE = -0.02f;
    c.pi.cE = 1e-6f;
    c.pe.dtdx = c.pe.dtdy = 2e-6f;
    c.pi.dtdx = c.pi.dtdy = 4e-7f;
    c.pe.bfield = 1;
    c.pe.cos_a = 1.0f;
    c.pe.cos_w = 0.9f;
    c.pe.sin_w = 0.43589f;
    c.pi.cos_a = 1.0f;
    c.pi.cos_w = 1.0f;
    c.key0 = 7;
    c.steady_se = 1e5f;
    c.steady_si = 300.0f;
    uint32_t *ec = malloc(4 * cap), *ic = malloc(4 * cap), *eid = malloc(4 * cap), *iid = malloc(4 * cap);
    float *ef[5], *if_[5];
    for (int k = 0; k < 5; k++) {
        ef[k] = malloc(4 * cap);
        if_[k] = malloc(4 * cap);
    }
    state32(n, N, 5e5, ec, ef[0], ef[1], ef[2], ef[3], ef[4], eid);
    state32(n, N, 3e2, ic, if_[0], if_[1], if_[2], if_[3], if_[4], iid);
    float *E32 = calloc(2 * ((N + 1) * N + 1), sizeof(float));
    int64_t *qe = calloc(N * N, 8), *qi = calloc(N * N, 8);
    double *rho = calloc(N * N, 8), *phi = calloc(N * N, 8);
    long lost8[8] = {0}, ne = n, ni = n;
    int cyc[8] = {0};
    CHECK(orc32_run(&c, steps, &ne, ec, ef[0], ef[1], ef[2], ef[3], ef[4], eid, &ni, ic, if_[0],
                    if_[1], if_[2], if_[3], if_[4], iid, E32, qe, qi, rho, phi, lost8, cyc) == 0);
    for (int k = 0; k < N * N; k++) CHECK(isfinite(phi[k]));
    if (mode == ORC_MODE_REF) {
        long m = orc32_steady(ne, ec, ef[0], ef[1], ef[2], ef[3], ef[4], eid, ni, ic, if_[0],
                              if_[1], if_[2], if_[3], if_[4], iid, N, 5, 7, 0, 1e5f, 300.0f);
        CHECK(m >= 0);
    }
    free(ec), free(ic), free(eid), free(iid), free(E32), free(qe), free(qi), free(rho), free(phi);
    for (int k = 0; k < 5; k++) free(ef[k]), free(if_[k]);
}

static orc_ref_config ref_config(int N) {
    orc_ref_config c;
    memset(&c, 0, sizeof(c));
    c.N = N;
    c.ncycle = 2;
    c.dt_dti = 1;
    c.L = 0.01;
    c.dx = c.dy = c.L / (N - 1);
    c.eps = 8.854e-12;
    c.coef_e = -0.5 * 5.6e-11 * 1.76e11;
    c.coef_i = 0.5 * 5.6e-11 * 2.4e6;
    c.dt_e = c.dt_i = 5.6e-11;
    c.qc_dxdy_e = -1.6e-19 * 1e5 / (c.dx * c.dx);
    c.qc_dxdy_i = -c.qc_dxdy_e;
    c.cos_a = 1.0;
    c.cos_we = c.cos_wi = 1.0;
    return c;
}

static void aos(orc_particle *p, uint32_t *id, long n, double L, double v) {
    for (long k = 0; k < n; k++) {
        p[k].x = L * urand();
        p[k].y = L * urand();
        p[k].vx = v * (urand() - 0.5);
        p[k].vy = v * (urand() - 0.5);
        p[k].vz = v * (urand() - 0.5);
        p[k].weight = 1.0;
        id[k] = (uint32_t)k;
    }
}

static void run_ref(void) {
    const int N = 17;
    const long n = 500;
    orc_ref_config c = ref_config(N);
    c.mg_tol = 0.0;
    orc_particle *e = malloc(sizeof(orc_particle) * n), *ip = malloc(sizeof(orc_particle) * n);
    uint32_t *eid = malloc(4 * n), *iid = malloc(4 * n);
    aos(e, eid, n, c.L, 1e6);
    aos(ip, iid, n, c.L, 1e3);
    double *Ex = calloc((N + 1) * (N + 1), 8), *Ey = calloc((N + 1) * (N + 1), 8);
    double *qe = calloc(N * N, 8), *qi = calloc(N * N, 8), *rho = calloc(N * N, 8), *phi = calloc(N * N, 8);
    long ne = n, ni = n;
    CHECK(orc_ref_run(&c, 3, &ne, e, eid, &ni, ip, iid, Ex, Ey, qe, qi, rho, phi) == 0);
    c.mg_tol = 1e-10;
    CHECK(orc_ref_run(&c, 2, &ne, e, eid, &ni, ip, iid, Ex, Ey, qe, qi, rho, phi) == 0);
    for (int k = 0; k < N * N; k++) CHECK(isfinite(phi[k]));
    /* the FP64 Monte-Carlo loop with pair re-injection */
    orc_mcc_params me, mi;
    memset(&me, 0, sizeof(me));
    me.gas = 0;
    me.electron = 1;
    me.ng = 3.2e21;
    me.dt = 5.6e-11;
    me.mass = 9.10939e-31;
    me.M = 6.63e-26;
    me.q = 1.602177e-19;
    me.vtn = 250.0;
    me.e_exc = 11.55;
    me.e_iz = 15.76;
    me.opal_B = 10.0;
    me.kev_amu_div = 39.948;
    mi = me;
    mi.electron = 0;
    mi.mass = me.M;
    orc_mcc_setup(&me);
    orc_mcc_setup(&mi);
    CHECK(me.binned == 1 && mi.binned == 0 && me.p_null > 0.0);
    const long cap = 4 * n;
    orc_particle *e2 = malloc(sizeof(orc_particle) * cap), *i2 = malloc(sizeof(orc_particle) * cap);
    uint32_t *eid2 = malloc(4 * cap), *iid2 = malloc(4 * cap);
    aos(e2, eid2, n, c.L, 4e6);
    aos(i2, iid2, n, c.L, 1e3);
    uint32_t next[2] = {(uint32_t)n, (uint32_t)n};
    long ev[8] = {0};
    ne = n;
    ni = n;
    CHECK(orc_ref_run_mcc(&c, &me, &mi, 1, 1e6, 300.0, 4, &ne, cap, e2, eid2, &ni, cap, i2, iid2,
                          next, Ex, Ey, qe, qi, rho, phi, ev) == 0);
    CHECK(ev[1] > 0);
    free(e), free(ip), free(eid), free(iid), free(Ex), free(Ey), free(qe), free(qi), free(rho),
        free(phi), free(e2), free(i2), free(eid2), free(iid2);
}

static void multigrid(void) {
    for (int n = 3; n <= 65; n = 2 * n - 1) {
        double *u = malloc(sizeof(double) * n * n), *f = malloc(sizeof(double) * n * n);
        for (int k = 0; k < n * n; k++) u[k] = f[k] = urand() - 0.5;
        CHECK(orc_mglin(u, n, 2, 0.02) == 0);
        memcpy(u, f, sizeof(double) * n * n);
        double rr = 1.0;
        CHECK(orc_mg_tol(u, n, 0.02, 1e-10, 100, &rr) >= 0);
        CHECK(rr <= 1e-10);
        CHECK(orc_mg_relres(u, f, n, 0.02) <= 1e-10);
        free(u);
        free(f);
    }
}

static void misc(void) {
    uint32_t ctr[4] = {1, 2, 3, 4}, key[2] = {5, 6}, out[4];
    orc_philox4x32_10(ctr, key, out);
    uint32_t c2[2] = {1, 2}, o2[2];
    orc_philox2x32_10(c2, 9, o2);
    for (int kind = 0; kind <= 9; kind++)
        for (double E = 1e-4; E < 2e4; E *= 1.7) CHECK(isfinite(orc_sigma(kind, E, 11.55)));
    orc_particle p[3] = {{1e-3, 2e-3, 1, 2, 3, 1}, {0.5, 0.25, -1, -2, -3, 2}, {0, 0, 0, 0, 0, 1}};
    const char *path = "selftest_dump.txt";
    CHECK(orc_dump_write(path, p, 3, 17) == 0);
    orc_particle q[8];
    long m = orc_dump_read(path, q, 8);
    CHECK(m == 3 && q[1].x == 0.5 && q[0].vz == 3.0);
    remove(path);
}

static void control_and_walls(void) {
    const int N = 17;
    const long n = 400, cap = 4 * n;
    const double L = 0.01, dx = L / (N - 1);
    uint32_t *cell = malloc(4 * cap), *id = malloc(4 * cap);
    float *ox = malloc(4 * cap), *oy = malloc(4 * cap), *vx = malloc(4 * cap), *vy = malloc(4 * cap),
          *vz = malloc(4 * cap), *w = malloc(4 * cap);
    orc_inject(n, 0, 3, 11, 12, dx, L, 1e6, cell, ox, oy, vx, vy, vz);
    for (long k = 0; k < n; k++) {
        CHECK(orc_cell_jx(cell[k]) >= 0 && orc_cell_jx(cell[k]) < N - 1);
        id[k] = (uint32_t)k;
        w[k] = 1.0f;
    }
    orc_reduce_params rp;
    memset(&rp, 0, sizeof(rp));
    rp.ng = 4;
    rp.size_check = 100;
    rp.key0 = 3;
    for (int g = 0; g < 4; g++) {
        rp.vlim[g] = 3e5 * (g + 1);
        rp.glimit[g] = 40;
        rp.gnew[g] = 20;
    }
    uint8_t *keep = malloc(cap);
    uint32_t ngroup[8], nrem[8];
    CHECK(orc_reduce(&rp, 5, n, vx, vy, id, w, keep, ngroup, nrem) >= 0);
    int64_t *q = calloc(N * N, 8);
    orc32_deposit_w(n, cell, ox, oy, w, N, ORC_MODE_PHYS, q);
    free(q);
    /* every particle outside one of the four walls */
    for (long k = 0; k < n; k++) {
        int side = (int)(k & 3);
        int jx = side == 0 ? -1 : side == 1 ? N - 1 : (int)(k % (N - 1));
        int jy = side == 2 ? -1 : side == 3 ? N - 1 : (int)(k % (N - 1));
        cell[k] = orc_cell_pack(jx, jy);
        vx[k] *= 8.0f;
        vy[k] *= 8.0f;
    }
    orc_wall_params wp;
    memset(&wp, 0, sizeof(wp));
    wp.N = N;
    wp.dx = dx;
    wp.Lw = L;
    wp.dt = 5.6e-11;
    wp.me = 9.10939e-31;
    wp.mi = 6.63e-26;
    wp.q = 1.602177e-19;
    wp.max_yield = 2.5;
    wp.E0 = 5.0;
    wp.Emax0 = 300.0;
    wp.thr = 15.76;
    wp.see_e = wp.refl = wp.back = wp.see_i = 1;
    uint32_t *oc = malloc(4 * cap);
    float *of[7];
    for (int k = 0; k < 7; k++) of[k] = malloc(4 * cap);
    uint64_t *okey = malloc(8 * cap);
    orc_wall_out out = {oc, of[0], of[1], of[2], of[3], of[4], of[5], okey};
    int64_t counts[6];
    for (int mode = ORC_MODE_REF; mode <= ORC_MODE_PHYS; mode++)
        for (int sp = 0; sp < 2; sp++) {
            wp.mode = mode;
            long m = orc_wall(&wp, sp, 9, n, cell, ox, oy, vx, vy, vz, w, id, &out, cap, counts);
            CHECK(m >= 0 && m <= cap && counts[0] == n);
        }
    free(cell), free(id), free(ox), free(oy), free(vx), free(vy), free(vz), free(w), free(keep),
        free(oc), free(okey);
    for (int k = 0; k < 7; k++) free(of[k]);
}

int main(void) {
    misc();
    multigrid();
    run32(ORC_MODE_REF, 0);
    run32(ORC_MODE_PHYS, 0);
    run32(ORC_MODE_PHYS, 1);
    run32(ORC_MODE_REF_STEADY, 0);
    run_ref();
    control_and_walls();
    if (failures) {
        fprintf(stderr, "selftest: %d failures\n", failures);
        return 1;
    }
    printf("selftest: ok\n");
    return 0;
}
