// serve.cpp -- batched ensembles and a continuous-batching serving loop over
// independent runs (no reference counterpart: the reference runs one
// compile-time configuration per process).
//
// A run is one pic_handle with its own HIP stream.  Runs share the GPU by
// being stepped from their own host threads: each thread issues its run's
// kernels and waits only on its own stream (the tolerance solve's convergence
// check), so the kernels of different runs overlap on the device while any one
// run waits.  Every run executes exactly the code path of a run alone, so its
// results are bit-identical whether it runs alone or in a batch.
//
// The server (pic_server_create / pic_serve_submit / poll / cancel / wait)
// keeps a FIFO of submitted requests; its scheduler thread admits the head of
// the queue whenever the estimated device memory of the active runs plus the
// request's fits the budget and fewer than max_active runs are active, so
// requests that arrive while others run join them as memory and slots free
// up.  Each admitted request runs on a worker thread; between steps it obeys
// a cancel request, and a run whose potential or field turns non-finite is
// stopped (PIC_ERR_NUMERIC) without touching the others.  A finished request
// reports its per-step latency, its potential's error against the FP64
// multigrid, and a fingerprint of its final charge and potential.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pic2d.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    double pos = q * (double)(v.size() - 1);
    size_t k = (size_t)pos;
    double f = pos - (double)k;
    return k + 1 < v.size() ? v[k] * (1.0 - f) + v[k + 1] * f : v[k];
}

int live_particles(pic_handle *h, int64_t *n) {
    int64_t a = 0, b = 0;
    int rc = pic_particle_count(h, 0, &a);
    if (rc == PIC_OK) rc = pic_particle_count(h, 1, &b);
    *n = a + b;
    return rc;
}

// one run: pre-loop wall pass, `steps` synchronised steps with their
// latencies, particle-steps summed from the count before each step
int run_request(const pic_serve_request &q, pic_serve_result &r, std::vector<double> &lat_ms) {
    pic_handle *h = nullptr;
    int rc = pic_create(&q.params, 0, nullptr, &h);
    if (rc != PIC_OK) return rc;
    rc = pic_set_runtime(h, 0);
    if (rc == PIC_OK) rc = pic_run(h);  // pre-loop wall pass only
    if (rc == PIC_OK) rc = pic_sync(h);
    int64_t ps = 0;
    for (int32_t t = 0; t < q.steps && rc == PIC_OK; t++) {
        int64_t n = 0;
        rc = live_particles(h, &n);
        if (rc != PIC_OK) break;
        Clock::time_point t0 = Clock::now();
        rc = pic_advance(h, 1);
        if (rc == PIC_OK) rc = pic_sync(h);
        lat_ms.push_back(1e3 * seconds_since(t0));
        ps += n;
    }
    r.particle_steps = ps;
    pic_destroy(h);
    return rc;
}

}  // namespace


namespace {

// FNV-1a over 64-bit words
uint64_t fnv64(uint64_t h, const void *data, size_t words) {
    const uint64_t *w = static_cast<const uint64_t *>(data);
    for (size_t k = 0; k < words; k++) {
        uint64_t x = w[k];
        for (int b = 0; b < 8; b++) {
            h ^= (x >> (8 * b)) & 0xFFu;
            h *= 1099511628211ull;
        }
    }
    return h;
}

// fingerprint of a run's last step: its fixed-point charge (both species)
// and its potential, bit for bit
int run_fingerprint(pic_handle *h, uint64_t *out) {
    int32_t N = 0;
    int rc = pic_get_nodes(h, &N);
    if (rc != PIC_OK) return rc;
    std::vector<int64_t> q((size_t)N * N);
    std::vector<double> phi((size_t)N * N);
    uint64_t f = 1469598103934665603ull;
    for (int s = 0; s < 2; s++) {
        if ((rc = pic_get_qfixed(h, s, q.data())) != PIC_OK) return rc;
        f = fnv64(f, q.data(), q.size());
    }
    if ((rc = pic_get_phi(h, phi.data())) != PIC_OK) return rc;
    *out = fnv64(f, phi.data(), phi.size());
    return PIC_OK;
}

}  // namespace

struct Request {
    pic_params params{};
    int32_t steps = 0;
    int64_t need = 0;
    int32_t state = PIC_REQ_QUEUED;
    int32_t error = PIC_OK;
    int32_t steps_done = 0;
    double submit_s = 0, admit_s = 0, finish_s = 0;
    std::vector<double> lat_ms;
    int64_t particle_steps = 0;
    double phi_error = -1, e_error = -1;
    uint64_t fingerprint = 0;
    int64_t device_bytes = 0;
    std::atomic<bool> cancel{false};
    std::thread worker;
};

struct pic_server {
    Clock::time_point t0 = Clock::now();
    std::mutex mu;
    std::condition_variable cv;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    std::deque<int64_t> queue;
    int64_t budget = 0, in_use = 0, peak = 0, next_id = 1;
    int32_t max_active = 1, active = 0, max_seen = 0;
    bool stop = false;
    std::thread scheduler;
};

namespace {

bool terminal(int32_t st) {
    return st == PIC_REQ_DONE || st == PIC_REQ_CANCELLED || st == PIC_REQ_FAILED;
}

void fill_status(const Request &q, pic_serve_status *o) {
    o->state = q.state;
    o->error = q.error;
    o->steps_done = q.steps_done;
    o->steps = q.steps;
    o->submit_s = q.submit_s;
    o->admit_s = q.admit_s;
    o->finish_s = q.finish_s;
    o->step_p50_ms = percentile(q.lat_ms, 0.50);
    o->step_p99_ms = percentile(q.lat_ms, 0.99);
    o->particle_steps = q.particle_steps;
    o->phi_error = q.phi_error;
    o->e_error = q.e_error;
    o->fingerprint = q.fingerprint;
    o->device_bytes = q.device_bytes;
}

// one admitted request, on its own thread
void serve_one(pic_server *srv, Request *q) {
    pic_handle *h = nullptr;
    int rc = pic_create(&q->params, 0, nullptr, &h);
    if (rc == PIC_OK) rc = pic_set_runtime(h, 0);
    if (rc == PIC_OK) rc = pic_run(h);  // the pre-loop wall pass
    if (rc == PIC_OK) pic_get_device_bytes(h, &q->device_bytes);
    bool cancelled = false;
    std::vector<double> lat;
    int64_t ps = 0;
    for (int32_t t = 0; t < q->steps && rc == PIC_OK; t++) {
        if (q->cancel.load()) {
            cancelled = true;
            break;
        }
        int64_t n = 0;
        if ((rc = live_particles(h, &n)) != PIC_OK) break;
        const Clock::time_point t0 = Clock::now();
        rc = pic_advance(h, 1);
        if (rc == PIC_OK) rc = pic_sync(h);
        lat.push_back(1e3 * seconds_since(t0));
        ps += n;
        int32_t finite = 1;
        if (rc == PIC_OK) rc = pic_field_finite(h, &finite);
        if (rc == PIC_OK && !finite) rc = PIC_ERR_NUMERIC;  // quarantined: this run stops
        std::lock_guard<std::mutex> g(srv->mu);
        q->steps_done = t + 1;
        q->particle_steps = ps;
    }
    double fe[5] = {-1, -1, -1, -1, -1};
    uint64_t fp = 0;
    if (rc == PIC_OK && !cancelled) {
        rc = pic_field_error(h, 1e-12, fe);
        if (rc == PIC_OK) rc = run_fingerprint(h, &fp);
    }
    if (h) pic_destroy(h);
    std::lock_guard<std::mutex> g(srv->mu);
    q->lat_ms.swap(lat);
    q->particle_steps = ps;
    q->phi_error = fe[0];
    q->e_error = std::max(fe[1], fe[2]);
    q->fingerprint = fp;
    q->error = rc;
    q->state = cancelled ? PIC_REQ_CANCELLED : (rc == PIC_OK ? PIC_REQ_DONE : PIC_REQ_FAILED);
    q->finish_s = seconds_since(srv->t0);
    srv->in_use -= q->need;
    srv->active--;
    srv->cv.notify_all();
}

// admits the head of the queue while it fits (FIFO); called with mu held
void admit(pic_server *srv) {
    while (!srv->queue.empty()) {
        Request &q = *srv->reqs.at(srv->queue.front());
        if (q.cancel.load()) {  // cancelled while queued
            srv->queue.pop_front();
            q.state = PIC_REQ_CANCELLED;
            q.finish_s = seconds_since(srv->t0);
            srv->cv.notify_all();
            continue;
        }
        if (srv->active >= srv->max_active || srv->in_use + q.need > srv->budget) return;
        srv->queue.pop_front();
        srv->in_use += q.need;
        srv->peak = std::max(srv->peak, srv->in_use);
        srv->active++;
        srv->max_seen = std::max(srv->max_seen, srv->active);
        q.state = PIC_REQ_RUNNING;
        q.admit_s = seconds_since(srv->t0);
        q.worker = std::thread(serve_one, srv, &q);
    }
}

void schedule(pic_server *srv) {
    std::unique_lock<std::mutex> lk(srv->mu);
    while (!srv->stop) {
        admit(srv);
        srv->cv.wait(lk);
    }
}

}  // namespace

extern "C" {

int pic_server_create(int64_t hbm_budget, int32_t max_active, pic_server **out) {
    if (!out || hbm_budget <= 0 || max_active < 1) return PIC_ERR_ARG;
    pic_server *srv = new pic_server();
    srv->budget = hbm_budget;
    srv->max_active = max_active;
    srv->scheduler = std::thread(schedule, srv);
    *out = srv;
    return PIC_OK;
}

int pic_server_destroy(pic_server *srv) {
    if (!srv) return PIC_ERR_ARG;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        for (auto &kv : srv->reqs) kv.second->cancel.store(true);
        srv->stop = true;
        srv->cv.notify_all();
    }
    srv->scheduler.join();
    {  // the queued requests never start
        std::lock_guard<std::mutex> g(srv->mu);
        for (int64_t id : srv->queue) {
            Request &q = *srv->reqs.at(id);
            q.state = PIC_REQ_CANCELLED;
            q.finish_s = seconds_since(srv->t0);
        }
        srv->queue.clear();
    }
    for (auto &kv : srv->reqs)
        if (kv.second->worker.joinable()) kv.second->worker.join();
    delete srv;
    return PIC_OK;
}

int pic_serve_submit(pic_server *srv, const pic_params *p, int32_t steps, int64_t *id) {
    if (!srv || !p || steps < 0 || !id) return PIC_ERR_ARG;
    int64_t need = 0;
    int rc = pic_estimate_bytes(p, &need);
    if (rc != PIC_OK) return rc;
    if (need > srv->budget) return PIC_ERR_CAPACITY;  // could never be admitted
    std::unique_ptr<Request> q(new Request());
    q->params = *p;
    q->steps = steps;
    q->need = need;
    std::lock_guard<std::mutex> g(srv->mu);
    if (srv->stop) return PIC_ERR_ARG;
    q->submit_s = seconds_since(srv->t0);
    *id = srv->next_id++;
    srv->queue.push_back(*id);
    srv->reqs[*id] = std::move(q);
    srv->cv.notify_all();
    return PIC_OK;
}

int pic_serve_poll(pic_server *srv, int64_t id, pic_serve_status *out) {
    if (!srv || !out) return PIC_ERR_ARG;
    std::lock_guard<std::mutex> g(srv->mu);
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return PIC_ERR_ARG;
    fill_status(*it->second, out);
    return PIC_OK;
}

int pic_serve_cancel(pic_server *srv, int64_t id) {
    if (!srv) return PIC_ERR_ARG;
    std::lock_guard<std::mutex> g(srv->mu);
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return PIC_ERR_ARG;
    it->second->cancel.store(true);
    srv->cv.notify_all();
    return PIC_OK;
}

int pic_serve_wait(pic_server *srv, int64_t id, double timeout_s, pic_serve_status *out) {
    if (!srv) return PIC_ERR_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return PIC_ERR_ARG;
    Request &q = *it->second;
    auto done = [&]() { return terminal(q.state); };
    if (timeout_s < 0)
        srv->cv.wait(lk, done);
    else
        srv->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
    if (out) fill_status(q, out);
    return terminal(q.state) ? PIC_OK : PIC_ERR_TIMEOUT;
}

// Advance n runs by `steps` fused steps each, all at once: one host thread per
// run (pic_advance on its own stream).  Returns the first error.
int pic_batch_advance(pic_handle *const *runs, int32_t n, int32_t steps) {
    if (!runs || n < 0 || steps < 0) return PIC_ERR_ARG;
    std::vector<int> rcs((size_t)n, PIC_OK);
    std::vector<std::thread> th;
    th.reserve((size_t)n);
    for (int32_t k = 0; k < n; k++) {
        th.emplace_back([&, k]() {
            int rc = pic_advance(runs[k], steps);
            if (rc == PIC_OK) rc = pic_sync(runs[k]);
            rcs[(size_t)k] = rc;
        });
    }
    for (std::thread &t : th) t.join();
    for (int rc : rcs)
        if (rc != PIC_OK) return rc;
    return PIC_OK;
}

int pic_serve(const pic_serve_request *reqs, int32_t n, int64_t hbm_budget, int32_t max_active,
              pic_serve_result *res, pic_serve_stats *stats) {
    if ((!reqs && n > 0) || n < 0 || (!res && n > 0) || !stats || hbm_budget <= 0 ||
        max_active < 1)
        return PIC_ERR_ARG;
    for (int32_t k = 0; k < n; k++) {
        int64_t need = 0;
        int rc = pic_estimate_bytes(&reqs[k].params, &need);
        if (rc != PIC_OK) return rc;
        if (need > hbm_budget) return PIC_ERR_CAPACITY;  // could never be admitted
        res[k] = pic_serve_result{};
    }
    pic_server *srv = nullptr;
    int rc = pic_server_create(hbm_budget, max_active, &srv);
    if (rc != PIC_OK) return rc;
    std::vector<int32_t> order((size_t)n);
    for (int32_t k = 0; k < n; k++) order[(size_t)k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return reqs[a].arrival_s < reqs[b].arrival_s;
    });
    // submit each request at its arrival time (seconds after the server started)
    std::vector<int64_t> ids((size_t)n, -1);
    for (int32_t k : order) {
        const double wait = reqs[k].arrival_s - seconds_since(srv->t0);
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        rc = pic_serve_submit(srv, &reqs[k].params, reqs[k].steps, &ids[(size_t)k]);
        if (rc != PIC_OK) break;
    }
    std::vector<double> all;
    int64_t ps = 0;
    int rc_first = rc;
    for (int32_t k = 0; k < n && rc == PIC_OK; k++) {
        pic_serve_status st{};
        pic_serve_wait(srv, ids[(size_t)k], -1.0, &st);
        res[k].admit_s = st.admit_s;
        res[k].finish_s = st.finish_s;
        res[k].particle_steps = st.particle_steps;
        res[k].step_p50_ms = st.step_p50_ms;
        res[k].step_p99_ms = st.step_p99_ms;
        res[k].status = st.error;
        res[k].phi_error = st.phi_error;
        res[k].e_error = st.e_error;
        ps += st.particle_steps;
        if (rc_first == PIC_OK && st.error != PIC_OK) rc_first = st.error;
        std::lock_guard<std::mutex> g(srv->mu);
        const Request &q = *srv->reqs.at(ids[(size_t)k]);
        all.insert(all.end(), q.lat_ms.begin(), q.lat_ms.end());
    }
    stats->wall_s = seconds_since(srv->t0);
    stats->particle_steps = ps;
    stats->throughput = stats->wall_s > 0 ? (double)ps / stats->wall_s : 0.0;
    stats->step_p50_ms = percentile(all, 0.50);
    stats->step_p99_ms = percentile(all, 0.99);
    {
        std::lock_guard<std::mutex> g(srv->mu);
        stats->peak_bytes = srv->peak;
        stats->max_concurrent = srv->max_seen;
    }
    pic_server_destroy(srv);
    return rc_first;
}

}  // extern "C"
