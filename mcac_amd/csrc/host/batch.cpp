// Ensemble and serving runs: many independent boxes on one GPU.  Each worker
// thread owns one box at a time (its own HIP stream, through mcac_sim) and runs
// jobs to completion; the GPU interleaves the boxes' launches.  Jobs start in
// list order, not before their arrival time.  No counterpart in the reference,
// which runs one box per process (src/main.cpp).
#include <algorithm>
#include <atomic>
#include <numeric>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/mcac_gpu.h"
#include "../../../include/mcac_sim.h"
#include "sim_group.hpp"

namespace {

// Admission by device memory: a job starts only when its box fits in what the
// running boxes leave of the budget (a job larger than the budget runs alone).
class HbmBudget {
  public:
    HbmBudget(int64_t cap, const char *const *ini_texts, int64_t n) : cap_(cap), bytes_(static_cast<size_t>(n), 0) {
        if (cap_ > 0)
            for (int64_t j = 0; j < n; ++j) bytes_[static_cast<size_t>(j)] = std::max<int64_t>(0, mcac_sim_device_bytes(ini_texts[j]));
    }
    bool try_take(int64_t job) {
        if (cap_ <= 0) return true;
        const int64_t b = bytes_[static_cast<size_t>(job)];
        int64_t u = used_.load();
        while (true) {
            if (u != 0 && u + b > cap_) return false;
            if (used_.compare_exchange_weak(u, u + b)) {
                int64_t pk = peak_.load();
                while (u + b > pk && !peak_.compare_exchange_weak(pk, u + b)) {
                }
                return true;
            }
        }
    }
    void give(int64_t job) {
        if (cap_ > 0) used_.fetch_sub(bytes_[static_cast<size_t>(job)]);
    }
    int64_t peak() const { return peak_.load(); }

  private:
    int64_t cap_;
    std::vector<int64_t> bytes_;
    std::atomic<int64_t> used_{0}, peak_{0};
};

// The order in which workers take the jobs of a grouped batch.  Jobs that
// arrive over time are taken in list order.  A batch that is all there at the
// start goes largest first; when the workers' box slots hold it all at once,
// it is dealt out ahead by the longest-processing-time rule: each job, largest
// first, to the worker with the least estimated work so far (work ~ N^2: N
// monomers take ~12 N steps whose windows grow with the box), so that the
// largest boxes, whose steps run one after another, spread over the workers
// instead of gathering on whichever took them first.  Sizes come from
// mcac_sim_device_bytes, which grows linearly with N.
class JobQueue {
  public:
    JobQueue(const char *const *ini_texts, int64_t n, const double *arrival_s, int workers, int64_t slots)
        : pos_(static_cast<size_t>(workers), 0) {
        order_.resize(static_cast<size_t>(n));
        std::iota(order_.begin(), order_.end(), int64_t(0));
        if (arrival_s || n == 0) return;
        std::vector<double> size(static_cast<size_t>(n));
        for (int64_t j = 0; j < n; ++j)
            size[static_cast<size_t>(j)] = static_cast<double>(std::max<int64_t>(1, mcac_sim_device_bytes(ini_texts[j])));
        std::stable_sort(order_.begin(), order_.end(), [&size](int64_t a, int64_t b) {
            return size[static_cast<size_t>(a)] > size[static_cast<size_t>(b)];
        });
        if (slots < n) return;  // largest first from the shared counter
        dealt_.resize(static_cast<size_t>(workers));
        std::vector<double> load(static_cast<size_t>(workers), 0.0);
        for (int64_t j : order_) {
            const size_t w = static_cast<size_t>(std::min_element(load.begin(), load.end()) - load.begin());
            dealt_[w].push_back(j);
            const double x = size[static_cast<size_t>(j)];
            load[w] += x * x;
        }
    }
    // The next job of worker w, or -1 when none is left for it.
    int64_t next(int w) {
        if (!dealt_.empty()) {
            size_t &p = pos_[static_cast<size_t>(w)];
            const auto &mine = dealt_[static_cast<size_t>(w)];
            return p < mine.size() ? mine[p++] : -1;
        }
        const int64_t k = next_.fetch_add(1);
        return k < static_cast<int64_t>(order_.size()) ? order_[static_cast<size_t>(k)] : -1;
    }

  private:
    std::vector<int64_t> order_;
    std::vector<std::vector<int64_t>> dealt_;
    std::vector<size_t> pos_;  // each touched by its own worker only
    std::atomic<int64_t> next_{0};
};

// Where a grouped worker's time went: picking the moves of a round, writing
// and launching them, reading the results (which includes waiting for the
// device: wait, summed from the finished runs' own counters).
struct WorkerTimes {
    double pick = 0, enqueue = 0, consume = 0, wait = 0;
    double g[4] = {0, 0, 0, 0};  // mcac_group_times: enqueue, rebuilds, launches, job-list waits
    int64_t rounds = 0, moves = 0;
};

// With MCAC_BATCH_PROGRESS=k (k >= 1) every k-th finished job prints a line
// to stderr (jobs finished, seconds since the call), so that a long batch shows
// its progress, and each grouped worker prints its WorkerTimes at its end;
// off by default.
class Progress {
  public:
    void worker_times(int w, const WorkerTimes &t) {
        if (every_ > 0)
            std::fprintf(stderr,
                         "batch: worker %d: %lld rounds, %lld moves, pick %.2f s, enqueue %.2f s (rebuilds %.2f s, "
                         "launches %.2f s, job-list waits %.2f s), consume %.2f s (waiting for the device %.2f s)\n",
                         w, static_cast<long long>(t.rounds), static_cast<long long>(t.moves), t.pick, t.enqueue,
                         t.g[1], t.g[2], t.g[3], t.consume, t.wait);
    }
    Progress(int64_t n_jobs) : n_(n_jobs) {
        if (const char *e = std::getenv("MCAC_BATCH_PROGRESS")) every_ = std::max(0, std::atoi(e));
    }
    void job_done(double t) {
        const int64_t d = done_.fetch_add(1) + 1;
        if (every_ > 0 && (d % every_ == 0 || d == n_))
            std::fprintf(stderr, "batch: %lld/%lld jobs done at %.1f s\n", static_cast<long long>(d),
                         static_cast<long long>(n_), t);
    }

  private:
    int64_t n_;
    int every_ = 0;
    std::atomic<int64_t> done_{0};
};

}  // namespace

extern "C" int mcac_batch_run(const char *const *ini_texts, const double *arrival_s, int64_t n_jobs, int workers,
                              int device, int64_t max_steps, int64_t hbm_budget, double *start_s, double *end_s,
                              int64_t *steps, int *status, int64_t *peak_bytes) {
    if (!ini_texts || n_jobs < 0 || workers < 1 || !start_s || !end_s || !steps || !status) return MCAC_E_ARGUMENT;
    HbmBudget budget(hbm_budget, ini_texts, n_jobs);
    Progress progress(n_jobs);
    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    auto now_s = [t0] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    std::atomic<int64_t> next{0};
    auto worker = [&] {
        for (int64_t j = next.fetch_add(1); j < n_jobs; j = next.fetch_add(1)) {
            const double at = arrival_s ? arrival_s[j] : 0.0;
            while (now_s() < at || !budget.try_take(j)) std::this_thread::sleep_for(std::chrono::microseconds(200));
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
            budget.give(j);
            end_s[j] = now_s();
            status[j] = rc;
            progress.job_done(end_s[j]);
            if (rc) std::fprintf(stderr, "mcac_batch_run: job %lld status %d %s\n", static_cast<long long>(j), rc, err);
        }
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; ++w) pool.emplace_back(worker);
    for (auto &t : pool) t.join();
    if (peak_bytes) *peak_bytes = budget.peak();
    for (int64_t j = 0; j < n_jobs; ++j)
        if (status[j]) return status[j];
    return 0;
}

// Lock-step form: worker w keeps up to G boxes in a group (one shared stream)
// and runs rounds: pick up to `depth` steps ahead for every box, enqueue them
// together (one launch; one block runs a box's steps in order, those behind a
// contact do nothing), read every box's results.  A box whose job ends leaves
// the round and the worker admits the next job that has arrived.  With two
// launches in flight (the default; MCAC_GROUP_OVERLAP=0 keeps one) a round
// enqueues the next launch before it reads the results of the one before, so
// the host's picks and reads overlap the device's steps: a box's new steps
// queue behind its steps still in flight and do nothing if one of those finds
// a contact, exactly as steps behind a contact in one launch.
extern "C" int mcac_batch_run_groups(const char *const *ini_texts, const double *arrival_s, int64_t n_jobs,
                                     int workers, int boxes_per_worker, int depth, int device, int64_t max_steps,
                                     int64_t hbm_budget, double *start_s, double *end_s, int64_t *steps, int *status,
                                     double *final_time, int64_t *peak_bytes) {
    if (!ini_texts || n_jobs < 0 || workers < 1 || boxes_per_worker < 1 || boxes_per_worker > 4096 || depth < 1 ||
        depth > 8 || !start_s || !end_s || !steps || !status)
        return MCAC_E_ARGUMENT;
    HbmBudget budget(hbm_budget, ini_texts, n_jobs);
    Progress progress(n_jobs);
    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    auto now_s = [t0] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    JobQueue queue(ini_texts, n_jobs, arrival_s, workers, static_cast<int64_t>(workers) * boxes_per_worker);
    std::atomic<int> worker_rc{0};
    int launches = 2;
    if (const char *e = std::getenv("MCAC_GROUP_OVERLAP")) launches = std::atoi(e) != 0 ? 2 : 1;
    auto worker = [&](int w) {
        const int G = boxes_per_worker;
        WorkerTimes prof;
        mcac_group *group = nullptr;
        int grc = mcac_group_create(device, G, &group);
        if (grc) {
            worker_rc.store(grc);
            return;
        }
        struct Active {
            mcac_sim *sim;
            int64_t job;
            size_t fresh;  // its moves in the newest launch
        };
        std::vector<Active> act;
        auto end_job = [&](size_t k, int rc) {
            const int64_t j = act[k].job;
            if (rc == 0) rc = mcac::sim_finalize(act[k].sim);
            if (final_time) final_time[j] = mcac_sim_scalar(act[k].sim, "time");
            mcac_sim_destroy(act[k].sim);
            budget.give(j);
            end_s[j] = now_s();
            status[j] = rc;
            progress.job_done(end_s[j]);
            if (rc) std::fprintf(stderr, "mcac_batch_run_groups: job %lld status %d\n", static_cast<long long>(j), rc);
            act[k] = act.back();
            act.pop_back();
        };
        int64_t pending = -1;
        bool exhausted = false;
        std::vector<mcac_box *> boxes;
        std::vector<int64_t> aggs;
        std::vector<double> dirs, lpms, maxrs;
        std::vector<uint64_t> tickets;
        std::vector<size_t> who;
        std::vector<int> step_rc;
        while (true) {
            // admit the jobs that have arrived while there is room
            while (static_cast<int>(act.size()) < G) {
                if (pending < 0 && !exhausted) {
                    pending = queue.next(w);
                    if (pending < 0) exhausted = true;
                }
                if (pending < 0) break;
                const double at = arrival_s ? arrival_s[pending] : 0.0;
                if (now_s() < at || !budget.try_take(pending)) {  // not arrived, or no memory for it yet
                    if (!act.empty()) break;
                    std::this_thread::sleep_for(std::chrono::microseconds(200));
                    continue;
                }
                const int64_t j = pending;
                pending = -1;
                start_s[j] = now_s();
                steps[j] = 0;
                char err[256] = {0};
                mcac_sim *sim = nullptr;
                int rc = mcac_sim_create(ini_texts[j], nullptr, device, &sim, err, static_cast<int>(sizeof err));
                if (rc == 0) rc = mcac::sim_join_group(sim, group, depth, launches);
                if (rc == 0) rc = mcac_sim_init(sim);
                if (rc) {
                    if (sim) mcac_sim_destroy(sim);
                    budget.give(j);
                    if (final_time) final_time[j] = 0.0;
                    end_s[j] = now_s();
                    status[j] = rc;
                    progress.job_done(end_s[j]);
                    std::fprintf(stderr, "mcac_batch_run_groups: job %lld status %d %s\n", static_cast<long long>(j),
                                 rc, err);
                    continue;
                }
                act.push_back({sim, j, 0});
            }
            if (act.empty()) {
                if (exhausted && pending < 0) break;
                continue;
            }
            // the next launch: up to `depth` steps of every box that has one
            // due (fewer while its earlier steps are in flight and a cell-list
            // rebuild is near, which would wait for them)
            const double r0 = now_s();
            boxes.clear();
            aggs.clear();
            dirs.clear();
            lpms.clear();
            maxrs.clear();
            who.clear();
            for (size_t k = 0; k < act.size(); ++k) {
                mcac::GroupMove m{};
                size_t limit = static_cast<size_t>(depth);
                if (mcac::sim_in_flight(act[k].sim) > 0)
                    limit = std::min<size_t>(limit, static_cast<size_t>(mcac_box_spare_moves(mcac::sim_box(act[k].sim))));
                size_t n = 0;
                while (n < limit && mcac::sim_next_move(act[k].sim, steps[act[k].job], max_steps, &m)) {
                    boxes.push_back(mcac::sim_box(act[k].sim));
                    aggs.push_back(m.agg);
                    dirs.insert(dirs.end(), m.dir, m.dir + 3);
                    lpms.push_back(m.lpm);
                    maxrs.push_back(m.maxradius);
                    who.push_back(k);
                    ++n;
                }
                act[k].fresh = n;
            }
            tickets.assign(boxes.size(), 0);
            const double r1 = now_s();
            int rc = boxes.empty() ? 0
                                   : mcac_move_enqueue_group(group, boxes.data(), aggs.data(), dirs.data(), lpms.data(),
                                                             maxrs.data(), static_cast<int>(boxes.size()),
                                                             tickets.data());
            if (rc) {
                // not expected; every box of the round ends with the status
                // (destroying a box waits for the group stream)
                for (size_t k = act.size(); k-- > 0;) end_job(k, rc);
                continue;
            }
            const double r2 = now_s();
            step_rc.assign(act.size(), 0);
            for (size_t q = 0; q < who.size(); ++q) mcac::sim_issued(act[who[q]].sim, tickets[q]);
            for (size_t k = 0; k < act.size(); ++k) {
                // the box's steps of the launch before (with one launch in
                // flight: of this one): run, or skipped behind a contact (then
                // its steps in this launch are skipped too, and read with them)
                const size_t keep = launches > 1 ? act[k].fresh : 0;
                while (step_rc[k] == 0 && mcac::sim_in_flight(act[k].sim) > keep) {
                    bool stepped = false;
                    step_rc[k] = mcac::sim_consume(act[k].sim, &stepped);
                    if (stepped) steps[act[k].job] += 1;
                }
            }
            const double r3 = now_s();
            prof.pick += r1 - r0;
            prof.enqueue += r2 - r1;
            prof.consume += r3 - r2;
            prof.rounds += 1;
            prof.moves += static_cast<int64_t>(boxes.size());
            // jobs that ended: finished, step budget spent, or an error (a
            // finished run has nothing in flight: a run ends at a merge, which
            // skips the steps behind it, or at its step budget, past which no
            // step is issued)
            for (size_t k = act.size(); k-- > 0;) {
                const bool over = steps[act[k].job] >= max_steps || mcac_sim_finished(act[k].sim);
                if (step_rc[k] || (over && mcac::sim_in_flight(act[k].sim) == 0)) {
                    prof.wait += mcac_sim_scalar(act[k].sim, "t_wait_ms") * 1e-3;
                    end_job(k, step_rc[k]);
                }
            }
        }
        mcac_group_times(group, prof.g);
        mcac_group_destroy(group);
        progress.worker_times(w, prof);
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; ++w) pool.emplace_back(worker, w);
    for (auto &t : pool) t.join();
    if (peak_bytes) *peak_bytes = budget.peak();
    if (worker_rc.load()) return worker_rc.load();
    for (int64_t j = 0; j < n_jobs; ++j)
        if (status[j]) return status[j];
    return 0;
}
