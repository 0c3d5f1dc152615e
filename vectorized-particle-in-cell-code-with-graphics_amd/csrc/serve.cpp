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
// The serving loop admits queued requests in arrival order whenever the
// estimated device memory of the active runs plus the next request fits the
// budget (and fewer than max_active runs are active), runs each admitted
// request to completion on a worker thread, and records per-step latency.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
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

extern "C" {

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
    std::vector<int64_t> need((size_t)n, 0);
    for (int32_t k = 0; k < n; k++) {
        int rc = pic_estimate_bytes(&reqs[k].params, &need[(size_t)k]);
        if (rc != PIC_OK) return rc;
        if (need[(size_t)k] > hbm_budget) return PIC_ERR_CAPACITY;  // could never be admitted
        res[k] = pic_serve_result{};
        res[k].status = PIC_OK;
    }
    std::vector<int32_t> order((size_t)n);
    for (int32_t k = 0; k < n; k++) order[(size_t)k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return reqs[a].arrival_s < reqs[b].arrival_s;
    });

    std::mutex mu;
    std::condition_variable cv;
    int64_t in_use = 0, peak = 0;
    int32_t active = 0, max_seen = 0;
    std::vector<std::vector<double>> lat((size_t)n);
    std::vector<std::thread> workers;
    workers.reserve((size_t)n);
    const Clock::time_point t0 = Clock::now();

    size_t next = 0;
    std::unique_lock<std::mutex> lk(mu);
    while (next < order.size()) {
        const int32_t k = order[next];
        const double now = seconds_since(t0);
        const bool arrived = reqs[k].arrival_s <= now;
        const bool fits = in_use + need[(size_t)k] <= hbm_budget && active < max_active;
        if (arrived && fits) {
            in_use += need[(size_t)k];
            peak = std::max(peak, in_use);
            active++;
            max_seen = std::max(max_seen, active);
            res[k].admit_s = now;
            next++;
            workers.emplace_back([&, k]() {
                int rc = run_request(reqs[k], res[k], lat[(size_t)k]);
                std::lock_guard<std::mutex> g(mu);
                res[k].status = rc;
                res[k].finish_s = seconds_since(t0);
                res[k].step_p50_ms = percentile(lat[(size_t)k], 0.50);
                res[k].step_p99_ms = percentile(lat[(size_t)k], 0.99);
                in_use -= need[(size_t)k];
                active--;
                cv.notify_all();
            });
            continue;
        }
        if (!arrived) {
            // sleep until the arrival, or until a run finishes
            cv.wait_for(lk, std::chrono::duration<double>(reqs[k].arrival_s - now));
        } else {
            cv.wait(lk);  // wait for memory or a slot
        }
    }
    lk.unlock();
    for (std::thread &t : workers) t.join();

    std::vector<double> all;
    int64_t ps = 0;
    int rc_first = PIC_OK;
    for (int32_t k = 0; k < n; k++) {
        all.insert(all.end(), lat[(size_t)k].begin(), lat[(size_t)k].end());
        ps += res[k].particle_steps;
        if (rc_first == PIC_OK && res[k].status != PIC_OK) rc_first = res[k].status;
    }
    stats->wall_s = seconds_since(t0);
    stats->particle_steps = ps;
    stats->throughput = stats->wall_s > 0 ? (double)ps / stats->wall_s : 0.0;
    stats->step_p50_ms = percentile(all, 0.50);
    stats->step_p99_ms = percentile(all, 0.99);
    stats->peak_bytes = peak;
    stats->max_concurrent = max_seen;
    return rc_first;
}

}  // extern "C"
