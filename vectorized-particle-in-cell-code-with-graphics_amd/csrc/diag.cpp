// diag.cpp -- host side of the batch-mode diagnostics: allocation, the
// per-step hook of the timestep loop, averages (graphs.cpp:1000-1165) and the
// reference's output files (graphs.cpp:1167-1369, pic.cpp:1262-1274).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "pic_host.h"

using namespace pic;

namespace pic_host {

namespace {

template <class T>
int dzero(T **p, size_t n) {
    *p = nullptr;
    if (hipMalloc((void **)p, (n ? n : 1) * sizeof(T)) != hipSuccess) return PIC_ERR_ALLOC;
    if (hipMemset(*p, 0, (n ? n : 1) * sizeof(T)) != hipSuccess) return PIC_ERR_DEVICE;
    return PIC_OK;
}

bool probes_ok(const pic_params &P) {
    for (int k = 0; k < 6; k++)
        if (P.probe[k] < 0 || P.probe[k] >= P.nodes) return false;
    return true;
}

}  // namespace

bool diag_graphs_on(const pic_handle *h) { return h->P.graph_steps > 0; }
bool diag_time_on(const pic_handle *h) { return h->P.time_data_freq > 0; }

int diag_alloc(pic_handle *h) {
    const pic_params &P = h->P;
    if (!diag_graphs_on(h) && !diag_time_on(h)) return PIC_OK;
    if (diag_graphs_on(h) && (P.no_bins < 1 || (P.no_bins & 1) == 0 || P.divide_bins_by < 1))
        return set_global_error(PIC_ERR_ARG, "NO_BINS must be odd and DIVIDE_BINS_BY >= 1");
    if (diag_time_on(h) && !probes_ok(P))
        return set_global_error(PIC_ERR_ARG, "POINT1X..POINT3Y must lie on the grid");
    DiagDev &D = h->diag;
    const size_t nn = (size_t)h->N * h->N, N = (size_t)h->N;
    int rc = PIC_OK;
    if (diag_graphs_on(h)) {
        if ((rc = dzero(&D.rho, nn)) || (rc = dzero(&D.phi, nn)) || (rc = dzero(&D.ex, nn)) ||
            (rc = dzero(&D.ey, nn)))
            return set_global_error(rc, "diagnostics allocation failed");
        for (int s = 0; s < 2; s++)
            if ((rc = dzero(&D.q[s], nn)) || (rc = dzero(&D.vsx[s], N)) ||
                (rc = dzero(&D.vsy[s], N)) || (rc = dzero(&D.cntx[s], N)) ||
                (rc = dzero(&D.cnty[s], N)) || (rc = dzero(&D.vpx[s], N)) ||
                (rc = dzero(&D.vpy[s], N)) || (rc = dzero(&D.dist[s], 4 * (size_t)P.no_bins)))
                return set_global_error(rc, "diagnostics allocation failed");
    }
    return PIC_OK;
}

void diag_free(pic_handle *h) {
    DiagDev &D = h->diag;
    void *p[] = {D.rho, D.phi, D.ex, D.ey, D.q[0], D.q[1], D.vsx[0], D.vsx[1], D.vsy[0],
                 D.vsy[1], D.cntx[0], D.cntx[1], D.cnty[0], D.cnty[1], D.vpx[0], D.vpx[1],
                 D.vpy[0], D.vpy[1], D.dist[0], D.dist[1], D.tdata};
    for (void *x : p)
        if (x) hipFree(x);
    D = DiagDev();
}

// timeData rows for a run of run_time steps (SIZE_TIME_DATA, parameters.h:241)
int diag_prepare_time(pic_handle *h, int32_t run_time) {
    if (!diag_time_on(h)) return PIC_OK;
    const int rows = run_time / h->P.time_data_freq + 1;
    if (rows <= h->diag_rows_cap) return PIC_OK;
    double *t = nullptr;
    if (dzero(&t, (size_t)rows * DIAG_TCOLS) != PIC_OK)
        return set_global_error(PIC_ERR_ALLOC, "timeData allocation failed");
    if (h->diag.tdata) {
        hipMemcpyAsync(t, h->diag.tdata, sizeof(double) * DIAG_TCOLS * h->diag_rows_cap,
                       hipMemcpyDeviceToDevice, h->st);
        hipStreamSynchronize(h->st);
        hipFree(h->diag.tdata);
    }
    h->diag.tdata = t;
    h->diag_rows_cap = rows;
    return PIC_OK;
}

// after the fields of step t (pic.cpp:1167-1199)
void diag_step(pic_handle *h, int64_t t) {
    const pic_params &P = h->P;
    const bool graphs =
        diag_graphs_on(h) && t >= P.graph_start && t < (int64_t)P.graph_start + P.graph_steps;
    const bool trow = diag_time_on(h) && t % P.time_data_freq == 0 && h->diag_rows < h->diag_rows_cap;
    if (!graphs && !trow) return;
    Species &e = h->sp[0], &i = h->sp[1];
    int4 loc = make_int4(P.dist_loc[0], P.dist_loc[1], P.dist_loc[2], P.dist_loc[3]);
    Probe6 pr{P.probe[0], P.probe[1], P.probe[2], P.probe[3], P.probe[4], P.probe[5]};
    launch_diag_step(h->st, h->diag, graphs, trow, h->diag_rows, (long long)t, h->N, h->rho,
                     h->phi, h->ex64, h->ey64, e.q, i.q, e.scale, i.scale, P.corner8 ? 8.0 : 4.0,
                     P.electron_charge, P.ion_charge, e.buf[e.cur], &h->cnt[0].n[e.cur], e.cap,
                     i.buf[i.cur], &h->cnt[1].n[i.cur], i.cap, P.vte, P.vti, P.no_bins,
                     P.divide_bins_by, P.spread, loc, h->cnt, e.cur, i.cur, pr);
    if (trow) h->diag_rows++;
    if (graphs) h->diag_graph_steps++;
    if (graphs && t == (int64_t)P.graph_start + P.graph_steps - 1 && P.write_graphs &&
        !h->out_dir.empty())
        diag_write_graphs(h, h->out_dir);
}

// ------------------------------------------------------------- averages
int diag_fields(pic_handle *h, double *rho, double *phi, double *ex, double *ey) {
    if (!diag_graphs_on(h)) return set_global_error(PIC_ERR_ARG, "graph averages are off (NO_GRAPH_STEPS = 0)");
    const size_t nn = (size_t)h->N * h->N;
    const double inv = (double)h->P.graph_steps;
    double *src[4] = {h->diag.rho, h->diag.phi, h->diag.ex, h->diag.ey};
    double *dst[4] = {rho, phi, ex, ey};
    for (int k = 0; k < 4; k++) {
        if (!dst[k]) continue;
        if (hipMemcpyAsync(dst[k], src[k], nn * sizeof(double), hipMemcpyDeviceToHost, h->st) !=
            hipSuccess)
            return set_global_error(PIC_ERR_DEVICE, "copy failed");
    }
    hipStreamSynchronize(h->st);
    for (int k = 0; k < 4; k++)  // graphs::average_fields divides by NO_GRAPH_STEPS
        if (dst[k])
            for (size_t m = 0; m < nn; m++) dst[k][m] /= inv;
    return PIC_OK;
}

int diag_density(pic_handle *h, int s, double *out) {
    if (!diag_graphs_on(h)) return set_global_error(PIC_ERR_ARG, "graph averages are off (NO_GRAPH_STEPS = 0)");
    const size_t nn = (size_t)h->N * h->N;
    if (hipMemcpyAsync(out, h->diag.q[s], nn * sizeof(double), hipMemcpyDeviceToHost, h->st) !=
        hipSuccess)
        return set_global_error(PIC_ERR_DEVICE, "copy failed");
    hipStreamSynchronize(h->st);
    for (size_t m = 0; m < nn; m++) out[m] /= (double)h->P.graph_steps;
    return PIC_OK;
}

int diag_velocities(pic_handle *h, int s, double *vx, double *vy) {
    if (!diag_graphs_on(h)) return set_global_error(PIC_ERR_ARG, "graph averages are off (NO_GRAPH_STEPS = 0)");
    const size_t N = (size_t)h->N;
    std::vector<double> a(N), b(N);
    hipMemcpyAsync(a.data(), h->diag.vpx[s], N * sizeof(double), hipMemcpyDeviceToHost, h->st);
    hipMemcpyAsync(b.data(), h->diag.vpy[s], N * sizeof(double), hipMemcpyDeviceToHost, h->st);
    if (hipStreamSynchronize(h->st) != hipSuccess)
        return set_global_error(PIC_ERR_DEVICE, "copy failed");
    // graphs::average_velocities: both species normalised by VTE (graphs.cpp:1114-1121)
    const double d = h->P.vte * (double)h->P.graph_steps;
    for (size_t j = 0; j < N; j++) {
        if (vx) vx[j] = a[j] / d;
        if (vy) vy[j] = b[j] / d;
    }
    return PIC_OK;
}

int diag_distributions(pic_handle *h, int s, double *out) {
    if (!diag_graphs_on(h)) return set_global_error(PIC_ERR_ARG, "graph averages are off (NO_GRAPH_STEPS = 0)");
    const size_t nb = 4 * (size_t)h->P.no_bins;
    std::vector<unsigned long long> c(nb);
    hipMemcpyAsync(c.data(), h->diag.dist[s], nb * sizeof(unsigned long long),
                   hipMemcpyDeviceToHost, h->st);
    if (hipStreamSynchronize(h->st) != hipSuccess)
        return set_global_error(PIC_ERR_DEVICE, "copy failed");
    for (size_t k = 0; k < nb; k++) out[k] = (double)c[k] / (double)h->P.graph_steps;
    return PIC_OK;
}

int diag_time_rows(pic_handle *h, std::vector<double> &rows, int &n) {
    n = h->diag_rows;
    rows.assign((size_t)h->diag_rows_cap * DIAG_TCOLS, 0.0);
    if (!h->diag.tdata) return PIC_OK;
    hipMemcpyAsync(rows.data(), h->diag.tdata, rows.size() * sizeof(double), hipMemcpyDeviceToHost,
                   h->st);
    if (hipStreamSynchronize(h->st) != hipSuccess)
        return set_global_error(PIC_ERR_DEVICE, "copy failed");
    return PIC_OK;
}

// ---------------------------------------------------------------- files
namespace {
std::string path_of(const std::string &dir, const char *name) {
    if (dir.empty()) return name;
    return (dir.back() == '/') ? dir + name : dir + "/" + name;
}

int write_grid(const std::string &path, const double *a, int N) {  // graphs.cpp:1189-1275
    FILE *f = fopen(path.c_str(), "w");
    if (!f) return set_global_error(PIC_ERR_FILE, "cannot write " + path);
    for (int j = 0; j < N; j++) {
        for (int i = 0; i < N; i++) fprintf(f, "%e ", a[(size_t)j * N + i]);
        fprintf(f, "\n");
    }
    fclose(f);
    return PIC_OK;
}
}  // namespace

int diag_write_graphs(pic_handle *h, const std::string &dir) {
    if (!diag_graphs_on(h)) return set_global_error(PIC_ERR_ARG, "graph averages are off (NO_GRAPH_STEPS = 0)");
    const int N = h->N, mid = N / 2, nb = h->P.no_bins, half = (nb - 1) / 2;
    const size_t nn = (size_t)N * N;
    std::vector<double> rho(nn), phi(nn), ex(nn), ey(nn);
    int rc = diag_fields(h, rho.data(), phi.data(), ex.data(), ey.data());
    if (rc != PIC_OK) return rc;
    {  // graphs.cpp:1167-1187: centre row and centre column
        std::string p1 = path_of(dir, "rho_phi_EgridX.txt"), p2 = path_of(dir, "rho_phi_EgridY.txt");
        FILE *f1 = fopen(p1.c_str(), "w"), *f2 = fopen(p2.c_str(), "w");
        if (!f1 || !f2) {
            if (f1) fclose(f1);
            if (f2) fclose(f2);
            return set_global_error(PIC_ERR_FILE, "cannot write " + p1);
        }
        for (int i = 0; i < N; i++) {
            fprintf(f1, "%e %e %e %e\n", (double)i, rho[(size_t)mid * N + i],
                    phi[(size_t)mid * N + i], ex[(size_t)mid * N + i]);
            fprintf(f2, "%e %e %e %e\n", (double)i, rho[(size_t)i * N + mid],
                    phi[(size_t)i * N + mid], ey[(size_t)i * N + mid]);
        }
        fclose(f1);
        fclose(f2);
    }
    if ((rc = write_grid(path_of(dir, "rho2d.txt"), rho.data(), N)) ||
        (rc = write_grid(path_of(dir, "phi2d.txt"), phi.data(), N)) ||
        (rc = write_grid(path_of(dir, "egrid2dX.txt"), ex.data(), N)) ||
        (rc = write_grid(path_of(dir, "egrid2dY.txt"), ey.data(), N)))
        return rc;
    const char *type[2] = {"e", "i1"};  // ELECTRON_TYPE, ION_TYPE (parameters.h:316-318)
    std::vector<double> q(nn), vx(N), vy(N), d(4 * (size_t)nb);
    for (int s = 0; s < 2; s++) {
        if ((rc = diag_density(h, s, q.data())) ||
            (rc = write_grid(path_of(dir, ("density2d_" + std::string(type[s]) + ".txt").c_str()),
                             q.data(), N)))
            return rc;
        if ((rc = diag_velocities(h, s, vx.data(), vy.data()))) return rc;
        std::string pv = path_of(dir, ("av_velocity_" + std::string(type[s]) + ".txt").c_str());
        FILE *f = fopen(pv.c_str(), "w");
        if (!f) return set_global_error(PIC_ERR_FILE, "cannot write " + pv);
        for (int j = 0; j < N; j++) fprintf(f, "%e %e %e \n", (double)j, vx[j], vy[j]);
        fclose(f);
        if ((rc = diag_distributions(h, s, d.data()))) return rc;
        const char *names[4] = {"first_v_distX_", "first_v_distY_", "second_v_distX_",
                                "second_v_distY_"};
        for (int k = 0; k < 4; k++) {  // graphs.cpp:1320-1369
            std::string pd = path_of(dir, (std::string(names[k]) + type[s] + ".txt").c_str());
            FILE *g = fopen(pd.c_str(), "w");
            if (!g) return set_global_error(PIC_ERR_FILE, "cannot write " + pd);
            for (int j = 0; j < nb; j++)
                fprintf(g, "%e %e \n", (double)(j - half) / (double)h->P.divide_bins_by,
                        d[(size_t)k * nb + j]);
            fclose(g);
        }
    }
    return PIC_OK;
}

// pic.cpp:1262-1274: SIZE_TIME_DATA rows, ostream default formatting
int diag_write_time(pic_handle *h, const std::string &dir) {
    if (!diag_time_on(h)) return set_global_error(PIC_ERR_ARG, "timeData is off (TIME_DATA_FREQUENCY = 0)");
    std::vector<double> rows;
    int n = 0;
    int rc = diag_time_rows(h, rows, n);
    if (rc != PIC_OK) return rc;
    std::string p = path_of(dir, "timeData.txt");
    FILE *f = fopen(p.c_str(), "w");
    if (!f) return set_global_error(PIC_ERR_FILE, "cannot write " + p);
    for (int r = 0; r < h->diag_rows_cap; r++) {
        const double *x = &rows[(size_t)r * DIAG_TCOLS];
        fprintf(f, "%lld %lld %lld %g %g %g %g %g\n", (long long)x[0], (long long)x[1],
                (long long)x[2], x[3], x[4], x[5], x[6], x[7]);
    }
    fclose(f);
    return PIC_OK;
}

}  // namespace pic_host
