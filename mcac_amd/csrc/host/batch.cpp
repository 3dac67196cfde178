// Ensemble and serving runs: many independent boxes on one GPU.  Each worker
// thread owns one box at a time (its own HIP stream, through mcac_sim) and runs
// jobs to completion; the GPU interleaves the boxes' launches.  Jobs start in
// list order, not before their arrival time.  No counterpart in the reference,
// which runs one box per process (src/main.cpp).
#include <atomic>
#include <chrono>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/mcac_sim.h"

extern "C" int mcac_batch_run(const char *const *ini_texts, const double *arrival_s, int64_t n_jobs, int workers,
                              int device, int64_t max_steps, double *start_s, double *end_s, int64_t *steps,
                              int *status) {
    if (!ini_texts || n_jobs < 0 || workers < 1 || !start_s || !end_s || !steps || !status) return MCAC_E_ARGUMENT;
    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    auto now_s = [t0] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    std::atomic<int64_t> next{0};
    auto worker = [&] {
        for (int64_t j = next.fetch_add(1); j < n_jobs; j = next.fetch_add(1)) {
            const double at = arrival_s ? arrival_s[j] : 0.0;
            while (now_s() < at) std::this_thread::sleep_for(std::chrono::microseconds(200));
            start_s[j] = now_s();
            steps[j] = 0;
            char err[256] = {0};
            mcac_sim *sim = nullptr;
            int rc = mcac_sim_create(ini_texts[j], nullptr, device, &sim, err, static_cast<int>(sizeof err));
            if (rc == 0) rc = mcac_sim_init(sim);
            while (rc == 0 && !mcac_sim_finished(sim) && steps[j] < max_steps) {
                int64_t done = 0;
                rc = mcac_sim_run(sim, max_steps - steps[j], &done);
                steps[j] += done;
                if (done == 0) break;
            }
            if (sim) mcac_sim_destroy(sim);
            end_s[j] = now_s();
            status[j] = rc;
            if (rc) std::fprintf(stderr, "mcac_batch_run: job %lld status %d %s\n", static_cast<long long>(j), rc, err);
        }
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; ++w) pool.emplace_back(worker);
    for (auto &t : pool) t.join();
    for (int64_t j = 0; j < n_jobs; ++j)
        if (status[j]) return status[j];
    return 0;
}
