#include "pick_table.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace mcac {
namespace {

inline bool before(double ta, int ia, double tb, int ib) { return ta > tb || (ta == tb && ia < ib); }

inline uint64_t bits_of(double x) {
    uint64_t b;
    std::memcpy(&b, &x, sizeof b);
    return b;
}
inline double from_bits(uint64_t b) {
    double x;
    std::memcpy(&x, &b, sizeof x);
    return x;
}
// frexp's exponent: x = m 2^e with 0.5 <= |m| < 1
inline int binade(double x) {
    const int E = static_cast<int>((bits_of(x) >> 52) & 0x7ff);
    if (E == 0 || E == 0x7ff) {
        int e = 0;
        (void)std::frexp(x, &e);
        return e;
    }
    return E - 1022;
}
inline bool same_binade(double a, double b) { return a != 0.0 && b != 0.0 && binade(a) == binade(b); }
// half the grid spacing of the binade of x (normal x): 2^(binade(x) - 54)
inline double half_spacing(double x) {
    const int E = static_cast<int>((bits_of(x) >> 52) & 0x7ff);
    if (E > 53 && E < 0x7ff) return from_bits(static_cast<uint64_t>(E - 53) << 52);
    return std::ldexp(1.0, binade(x) - 54);
}

// Up to 8 host threads kept for the re-sum after a change of the largest time
// step and for the partitions of a tie replay; the calling thread takes part.
// Each call is a task whose jobs are claimed by an atomic counter; the call
// returns once every job has completed, without waiting for idle threads to
// check in (a worker that wakes late finds the counter exhausted, and its
// reference to the task keeps the task alive).  Workers spin for kSpin after
// a task before they sleep, and so does the caller before it sleeps on the
// completion count.
class Workers {
  public:
    static Workers &get() {
        static Workers w;
        return w;
    }
    // Runs fn(k) for k in [0, n); one caller at a time uses the pool, the
    // others (simulations on other host threads) run serially.
    template <class F> void run(size_t n, F &fn) {
        std::unique_lock<std::mutex> own(user_, std::try_to_lock);
        if (n < 2 || threads_.empty() || !own.owns_lock()) {
            for (size_t k = 0; k < n; ++k) fn(k);
            return;
        }
        auto task = std::make_shared<Task>();
        task->job = [&fn](size_t k) { fn(k); };
        task->n = n;
        {
            std::lock_guard<std::mutex> lk(m_);
            task_ = task;
            gen_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_all();
        work(*task);
        const auto t0 = std::chrono::steady_clock::now();
        while (task->done.load(std::memory_order_acquire) != n && std::chrono::steady_clock::now() - t0 < kSpin) {
        }
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return task->done.load() == n; });
        task_.reset();
    }
    ~Workers() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
            gen_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
    }

  private:
    struct Task {
        std::function<void(size_t)> job;
        size_t n = 0;
        std::atomic<size_t> next{0}, done{0};
    };
    static constexpr std::chrono::microseconds kSpin{300};
    Workers() {
        const size_t T = std::min<size_t>(16, std::max(1u, std::thread::hardware_concurrency()));
        for (size_t t = 1; t < T; ++t) threads_.emplace_back([this] { loop(); });
    }
    // claims and runs jobs of the task; the thread that completes the last one wakes the caller
    void work(Task &task) {
        for (size_t k = task.next.fetch_add(1); k < task.n; k = task.next.fetch_add(1)) {
            task.job(k);
            if (task.done.fetch_add(1) + 1 == task.n) {
                std::lock_guard<std::mutex> lk(m_);
                done_.notify_all();
            }
        }
    }
    void loop() {
        uint64_t seen = 0;
        while (true) {
            const auto t0 = std::chrono::steady_clock::now();
            while (gen_.load(std::memory_order_acquire) == seen && std::chrono::steady_clock::now() - t0 < kSpin) {
            }
            std::shared_ptr<Task> task;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return gen_.load() != seen; });
                seen = gen_.load();
                if (stop_) return;
                task = task_;
            }
            if (task) work(*task);
        }
    }
    std::vector<std::thread> threads_;
    std::mutex user_;  // held by the thread whose job the pool runs
    std::mutex m_;
    std::condition_variable cv_, done_;
    std::shared_ptr<Task> task_;
    std::atomic<uint64_t> gen_{0};
    bool stop_ = false;
};

// exact rounding error of s = fl(a + b) (TwoSum)
inline double two_sum_err(double a, double b, double s) {
    double bb = s - a;
    double aa = s - bb;
    return (a - aa) + (b - bb);
}

}  // namespace

void host_parallel_for(size_t n, const std::function<void(size_t)> &fn) {
    auto f = [&fn](size_t k) { fn(k); };
    Workers::get().run(n, f);
}

void PickTable::build(const std::vector<int> &ids, const std::vector<double> &ts_by_id, double factor) {
    // the previous order, when it holds the same ids, is nearly sorted after a
    // small change of every time step: insertion sort from it (the order is
    // strict, so any sort gives the same result)
    std::vector<int> order;
    order.reserve(ids.size());
    bool reuse = n_ > 0 && !ids.empty();
    if (reuse) {
        // the previous order restricted to the ids given, then the ids it lacks
        int top = 0;
        for (int id : ids) top = std::max(top, id);
        for (const Chunk &c : chunks_)
            for (int id : c.id) top = std::max(top, id);
        std::vector<char> given(static_cast<size_t>(top) + 1, 0), kept(static_cast<size_t>(top) + 1, 0);
        for (int id : ids) given[static_cast<size_t>(id)] = 1;
        for (const Chunk &c : chunks_)
            for (int id : c.id)
                if (given[static_cast<size_t>(id)]) {
                    order.push_back(id);
                    kept[static_cast<size_t>(id)] = 1;
                }
        for (int id : ids)
            if (!kept[static_cast<size_t>(id)]) order.push_back(id);
        reuse = order.size() == ids.size();
    }
    auto less = [&ts_by_id](int a, int b) {
        return before(ts_by_id[static_cast<size_t>(a)], a, ts_by_id[static_cast<size_t>(b)], b);
    };
    if (reuse) {
        size_t moves = 0;
        for (size_t i = 1; i < order.size() && reuse; ++i) {
            const int v = order[i];
            size_t j = i;
            while (j > 0 && less(v, order[j - 1])) {
                order[j] = order[j - 1];
                --j;
                if (++moves > 64 * order.size()) reuse = false;  // far from sorted: sort instead
            }
            order[j] = v;
        }
    }
    if (!reuse) {
        order.assign(ids.begin(), ids.end());
        std::sort(order.begin(), order.end(), less);
    }
    n_ = order.size();
    factor_ = factor > 0.0 ? factor : (n_ ? ts_by_id[static_cast<size_t>(order[0])] : 0.0);
    chunks_.clear();
    for (size_t i = 0; i < n_; i += chunk_) {
        Chunk c;
        size_t e = std::min(n_, i + chunk_);
        for (size_t j = i; j < e; ++j) {
            const double t = ts_by_id[static_cast<size_t>(order[j])];
            c.id.push_back(order[j]);
            c.ts.push_back(t);
            c.tp.push_back(factor_ / t);  // aggregat_list.cpp:114
        }
        chunks_.push_back(std::move(c));
    }
    reindex();
    walk(0);
}

void PickTable::reindex() {
    offset_.resize(chunks_.size());
    size_t o = 0;
    for (size_t k = 0; k < chunks_.size(); ++k) {
        offset_[k] = o;
        o += chunks_[k].ts.size();
    }
}

// Sequential sums of the chunk from A; positions where the shift argument needs
// an exact step: rounding ties and binade changes of L.
void PickTable::materialize(Chunk &c, double A) {
    const size_t m = c.ts.size();
    c.L.resize(m);
    c.flags.clear();
    c.A = A;
    double L = A;
    // biased exponents stand for binades (positive normal sums)
    int64_t e = (A != 0.0) ? static_cast<int64_t>(bits_of(A) >> 52) : -1;
    const double *tp = c.tp.data();
    double *out = c.L.data();
    for (size_t i = 0; i < m; ++i) {
        const double t = tp[i];
        const double x = L + t;
        // rounding error of x (Fast2Sum when |L| >= |t|, else TwoSum)
        const double err = (L >= t) ? t - (x - L) : two_sum_err(L, t, x);
        const uint64_t xb = bits_of(x);
        const int64_t ex = static_cast<int64_t>(xb >> 52);
        const uint64_t eb = bits_of(err) & 0x7fffffffffffffffull;
        // a tie leaves |err| equal to half the spacing of x's binade
        const bool tie = eb != 0 && (ex > 53 ? eb == (static_cast<uint64_t>(ex - 53) << 52)
                                             : std::fabs(err) == half_spacing(x));
        if (ex != e || tie) c.flags.push_back(static_cast<int>(i));
        e = ex;
        L = x;
        out[i] = L;
    }
    c.dirty = false;
}

// The sums from A again from position `from` on, after entries changed there:
// the positions before it keep their sums, their flags and the chunk's A.
void PickTable::resum_from(Chunk &c, size_t from) {
    const size_t m = c.ts.size();
    c.L.resize(m);
    c.flags.resize(static_cast<size_t>(std::lower_bound(c.flags.begin(), c.flags.end(), static_cast<int>(from)) -
                                       c.flags.begin()));
    double L = c.L[from - 1];
    int64_t e = static_cast<int64_t>(bits_of(L) >> 52);
    const double *tp = c.tp.data();
    double *out = c.L.data();
    for (size_t i = from; i < m; ++i) {
        const double t = tp[i];
        const double x = L + t;
        const double err = (L >= t) ? t - (x - L) : two_sum_err(L, t, x);
        const uint64_t xb = bits_of(x);
        const int64_t ex = static_cast<int64_t>(xb >> 52);
        const uint64_t eb = bits_of(err) & 0x7fffffffffffffffull;
        const bool tie = eb != 0 && (ex > 53 ? eb == (static_cast<uint64_t>(ex - 53) << 52)
                                             : std::fabs(err) == half_spacing(x));
        if (ex != e || tie) c.flags.push_back(static_cast<int>(i));
        e = ex;
        L = x;
        out[i] = L;
    }
    c.dirty = false;
}

// Offsets of the chunk for the true start S.
void PickTable::settle(Chunk &c, double S) {
    const size_t m = c.ts.size();
    if (m == 0) {
        c.fast = true;
        c.fast_d = 0.0;
        c.true_end = S;
        return;
    }
    if (S != c.A && !same_binade(S, c.A)) {
        // the stored start is in another binade: re-sum from the true start
        materialize(c, S);
    }
    const double d0 = S - c.A;
    const double L0 = c.L[0], L1 = c.L[m - 1];
    if (S == c.A || (c.flags.empty() && binade(L0) == binade(L1) && same_binade(L0 + d0, L0) &&
                     same_binade(L1 + d0, L1))) {
        // no tie and no binade change for L or L + d: the whole chunk is L + d
        c.fast = true;
        c.fast_d = d0;
        c.true_end = L1 + d0;
        return;
    }
    c.fast = false;
    c.seg.clear();
    c.segd.clear();
    c.segabs.clear();
    // bound on |d| over the chunk: the start offset plus a grid unit per exact step
    const double ulp_end = std::ldexp(1.0, binade(std::max(std::fabs(L1), std::fabs(S)) * 2.0) - 52);
    const double D = std::fabs(d0) + 64.0 * ulp_end;
    // positions stepped exactly: flags, and every position where L is within D of a
    // power of two lying between min(L, L + d) and max(L, L + d)
    std::vector<std::pair<size_t, size_t>> &ranges = ranges_;
    ranges.clear();
    for (int f : c.flags) ranges.push_back({static_cast<size_t>(f), static_cast<size_t>(f) + 1});
    const double vlo = std::min(std::min(L0, L0 + d0), std::min(S, c.A)) - D;
    const double vhi = std::max(L1, L1 + d0) + D;
    if (vhi > 0.0) {
        double p2 = std::ldexp(1.0, binade(std::max(vlo, std::numeric_limits<double>::min())) - 1);
        for (int guard = 0; p2 <= vhi && guard < 2100; ++guard, p2 *= 2.0) {
            if (p2 < vlo) continue;
            size_t lo = static_cast<size_t>(std::lower_bound(c.L.begin(), c.L.end(), p2 - D) - c.L.begin());
            size_t hi = static_cast<size_t>(std::lower_bound(c.L.begin(), c.L.end(), p2 + D) - c.L.begin()) + 1;
            if (lo >= m) lo = m - 1;
            if (hi > m) hi = m;
            ranges.push_back({lo, hi});
        }
    }
    std::sort(ranges.begin(), ranges.end());
    std::vector<std::pair<size_t, size_t>> &merged = merged_;
    merged.clear();
    for (auto &r : ranges) {
        if (!merged.empty() && r.first <= merged.back().second) merged.back().second = std::max(merged.back().second, r.second);
        else merged.push_back(r);
    }
    double d = d0;
    bool d_valid = true;
    bool ok = true;
    size_t pos = 0;
    for (size_t k = 0; k < merged.size() && ok; ++k) {
        size_t r0 = merged[k].first, r1 = merged[k].second;
        if (pos > r0) r0 = pos;
        if (r0 >= r1) continue;
        if (pos < r0) {
            if (!d_valid) {
                ok = false;
                break;
            }
            c.seg.push_back(static_cast<int>(pos));
            c.segd.push_back(d);
            c.segabs.push_back(0);
        }
        double s = (r0 == 0) ? S : c.L[r0 - 1] + d;
        size_t j = r0;
        while (true) {
            for (; j < r1; ++j) {
                s = s + c.tp[j];
                c.seg.push_back(static_cast<int>(j));
                c.segd.push_back(s);  // exact value of this position
                c.segabs.push_back(1);
                ++exact_steps_;
            }
            // past the range the shift holds only if s and L share a binade
            if (same_binade(s, c.L[r1 - 1]) || r1 >= m) break;
            r1 = std::min(m, r1 + 1);
        }
        d = s - c.L[r1 - 1];
        d_valid = same_binade(s, c.L[r1 - 1]);
        if (std::fabs(d) > D) ok = false;
        pos = r1;
        if (k + 1 < merged.size() && merged[k + 1].first < pos) merged[k + 1].first = pos;
    }
    if (ok && pos < m) {
        if (!d_valid) ok = false;
        else {
            c.seg.push_back(static_cast<int>(pos));
            c.segd.push_back(d);
            c.segabs.push_back(0);
        }
    }
    // many exact positions, or the offset left its bound: re-sum from the true start
    if (!ok || c.seg.size() > 24) {
        materialize(c, S);
        c.fast = true;
        c.fast_d = 0.0;
        c.true_end = c.L[m - 1];
        return;
    }
    c.true_end = c.segabs.back() ? c.segd.back() : c.L[m - 1] + c.segd.back();
}

void PickTable::walk(size_t from) {
    double S = (from == 0) ? 0.0 : chunks_[from - 1].true_end;
    for (size_t k = from; k < chunks_.size(); ++k) {
        Chunk &c = chunks_[k];
        if (c.dirty) {
            if (c.dirty_from > 0 && c.dirty_from < c.ts.size()) resum_from(c, c.dirty_from);
            else if (c.dirty_from == 0 || c.ts.empty()) materialize(c, S);
            else {  // only the last entries were removed: the sums before them hold
                c.L.resize(c.ts.size());
                c.flags.resize(static_cast<size_t>(
                    std::lower_bound(c.flags.begin(), c.flags.end(), static_cast<int>(c.ts.size())) -
                    c.flags.begin()));
                c.dirty = false;
            }
            c.dirty_from = 0;
        }
        settle(c, S);
        S = c.true_end;
    }
}

double PickTable::true_at(const Chunk &c, size_t i) const {
    if (c.fast) return c.L[i] + c.fast_d;
    size_t s = static_cast<size_t>(std::upper_bound(c.seg.begin(), c.seg.end(), static_cast<int>(i)) - c.seg.begin()) - 1;
    return c.segabs[s] ? c.segd[s] : c.L[i] + c.segd[s];
}

double PickTable::total() const { return chunks_.empty() ? 0.0 : chunks_.back().true_end; }

size_t PickTable::find_chunk(double t, int id) const {
    // first chunk whose last entry is not before (t, id)
    size_t lo = 0, hi = chunks_.size();
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        const Chunk &c = chunks_[mid];
        if (before(c.ts.back(), c.id.back(), t, id)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

void PickTable::erase_entry(double t, int id, size_t &first) {
    size_t k = find_chunk(t, id);
    if (k >= chunks_.size()) throw std::logic_error("pick table: entry not found");
    Chunk &c = chunks_[k];
    size_t lo = 0, hi = c.ts.size();
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (before(c.ts[mid], c.id[mid], t, id)) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= c.ts.size() || c.id[lo] != id) throw std::logic_error("pick table: entry not found");
    c.ts.erase(c.ts.begin() + static_cast<long>(lo));
    c.tp.erase(c.tp.begin() + static_cast<long>(lo));
    c.id.erase(c.id.begin() + static_cast<long>(lo));
    c.dirty_from = c.dirty ? std::min(c.dirty_from, lo) : lo;
    c.dirty = true;
    if (c.ts.empty()) {
        chunks_.erase(chunks_.begin() + static_cast<long>(k));
        if (k > 0) --k;
    }
    first = std::min(first, k);
}

void PickTable::insert_entry(double t, int id, size_t &first) {
    size_t k = find_chunk(t, id);
    if (k >= chunks_.size()) k = chunks_.size() - 1;  // after every entry: last chunk
    Chunk &c = chunks_[k];
    size_t lo = 0, hi = c.ts.size();
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (before(c.ts[mid], c.id[mid], t, id)) lo = mid + 1;
        else hi = mid;
    }
    c.ts.insert(c.ts.begin() + static_cast<long>(lo), t);
    c.tp.insert(c.tp.begin() + static_cast<long>(lo), factor_ / t);
    c.id.insert(c.id.begin() + static_cast<long>(lo), id);
    c.dirty_from = c.dirty ? std::min(c.dirty_from, lo) : lo;
    c.dirty = true;
    if (c.ts.size() > 2 * chunk_) {
        Chunk tail;
        const size_t h = c.ts.size() / 2;
        tail.ts.assign(c.ts.begin() + static_cast<long>(h), c.ts.end());
        tail.tp.assign(c.tp.begin() + static_cast<long>(h), c.tp.end());
        tail.id.assign(c.id.begin() + static_cast<long>(h), c.id.end());
        c.ts.resize(h);
        c.tp.resize(h);
        c.id.resize(h);
        tail.dirty = true;
        chunks_.insert(chunks_.begin() + static_cast<long>(k) + 1, std::move(tail));
    }
    first = std::min(first, k);
}

void PickTable::replace(int keep, double keep_old, double keep_new, int rem, double rem_ts) {
    size_t first = chunks_.size();
    erase_entry(keep_old, keep, first);
    erase_entry(rem_ts, rem, first);
    if (chunks_.empty()) {
        chunks_.push_back(Chunk());
        chunks_[0].ts.push_back(keep_new);
        chunks_[0].tp.push_back(factor_ / keep_new);
        chunks_[0].id.push_back(keep);
        first = 0;
    } else {
        insert_entry(keep_new, keep, first);
    }
    n_ -= 1;
    const double f = chunks_[0].ts[0];
    reindex();
    if (f != factor_) {
        // every tp changes: recompute them and re-sum every chunk in parallel
        // from a start scaled from the previous sums, then settle in order
        const double ratio = f / factor_;
        factor_ = f;
        ++rescales_;
        std::vector<double> start(chunks_.size());
        for (size_t k = 0; k < chunks_.size(); ++k) start[k] = (k == 0) ? 0.0 : chunks_[k - 1].true_end * ratio;
        auto job = [this, &start](size_t k) {
            Chunk &c = chunks_[k];
            const double fk = factor_;
            const double *t = c.ts.data();
            double *tp = c.tp.data();
            const size_t m = c.ts.size();
            for (size_t i = 0; i < m; ++i) tp[i] = fk / t[i];  // aggregat_list.cpp:114
            materialize(c, start[k]);
        };
        if (chunks_.size() >= 16) Workers::get().run(chunks_.size(), job);
        else
            for (size_t k = 0; k < chunks_.size(); ++k) job(k);
        first = 0;
    }
    walk(first);
}

int PickTable::pick_fast(double u, size_t *pos) {
    const double val = u * total();
    // first chunk whose last true sum is >= val, then lower_bound inside it
    size_t lo = 0, hi = chunks_.size();
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (chunks_[mid].true_end < val) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= chunks_.size()) lo = chunks_.size() - 1;
    const Chunk &c = chunks_[lo];
    size_t a = 0, b = c.ts.size();
    while (a < b) {
        size_t mid = (a + b) / 2;
        if (true_at(c, mid) < val) a = mid + 1;
        else b = mid;
    }
    if (a >= c.ts.size()) a = c.ts.size() - 1;
    const double tp = c.tp[a];
    bool t = false;
    if (a > 0) t = t || c.tp[a - 1] == tp;
    else if (lo > 0) t = t || chunks_[lo - 1].tp.back() == tp;
    if (a + 1 < c.ts.size()) t = t || c.tp[a + 1] == tp;
    else if (lo + 1 < chunks_.size()) t = t || chunks_[lo + 1].tp.front() == tp;
    *pos = offset_[lo] + a;
    return t ? -1 : c.id[a];
}

// Order that std::sort (libstdc++ introsort, bits/stl_algo.h: __introsort_loop,
// __unguarded_partition_pivot, __move_median_to_first, __unguarded_partition,
// then the final insertion sort) gives to equal keys of a key array.  The
// final insertion sort keeps the order of equal keys, so their order is their
// order after the partition phase.  Only the partitions holding a tracked key
// are carried out: O(n) work instead of a full O(n log n) sort.
namespace {

using KeyLabel = PickTable::KeyLabel;

inline void move_median_to_first(KeyLabel *a, size_t r, size_t x, size_t y, size_t z) {
    auto lt = [a](size_t u, size_t v) { return a[u].k < a[v].k; };
    size_t m;
    if (lt(x, y)) m = lt(y, z) ? y : (lt(x, z) ? z : x);
    else m = lt(x, z) ? x : (lt(y, z) ? z : y);
    std::swap(a[r], a[m]);
}

// partition of [f, l) around the median of three moved to f; returns the cut
inline size_t partition_pivot(KeyLabel *a, size_t f, size_t l) {
    move_median_to_first(a, f, f + 1, f + (l - f) / 2, l - 1);
    const double p = a[f].k;
    size_t lo = f + 1, hi = l;
    while (true) {
        while (a[lo].k < p) ++lo;
        --hi;
        while (p < a[hi].k) --hi;
        if (!(lo < hi)) return lo;
        std::swap(a[lo], a[hi]);
        ++lo;
    }
}

// The same partition computed by the worker pool.  With L_1 < L_2 < ... the
// positions in (f, l) whose key is not below the pivot p and R_1 > R_2 > ...
// the positions in [f, l) whose key is not above it, the loop above swaps
// L_t with R_t for every t before the first t* with L_t* >= R_t*, and returns
// min(L_t*, R_{t*-1}) (R_0 = l, L_t = +inf past the last L): while L_t < R_t
// the scans meet only positions no swap has touched yet, and the swapped-in
// elements stop them exactly where the sequential scans stop.  The L and R
// positions are not listed: one pass counts them per sub-block of kSub
// entries, L_t and R_t are found from those prefix counts and a scan of one
// sub-block, and each chunk of swaps starts from its first L and R position
// (found before any swap) and walks on by the keys, which no other chunk
// touches (every L_t with t < t* lies before every such R_t), gathering its
// positions first and swapping them after.
constexpr size_t kParallelPartition = size_t(1) << 16;
constexpr size_t kPartitionBlock = size_t(1) << 14;
constexpr size_t kSub = 256;  // entries per counted sub-block (kPartitionBlock is a multiple)

size_t partition_pivot_parallel(KeyLabel *a, size_t f, size_t l, std::vector<uint32_t> &cl,
                                std::vector<uint32_t> &cr) {
    move_median_to_first(a, f, f + 1, f + (l - f) / 2, l - 1);
    const double p = a[f].k;
    const size_t n = l - f;
    const size_t ns = (n + kSub - 1) / kSub;
    cl.assign(ns + 1, 0);  // cl[s + 1]: L positions in sub-block s, then prefix sums
    cr.assign(ns + 1, 0);
    auto count = [&](size_t b) {
        const size_t s0 = b * (kPartitionBlock / kSub), s1 = std::min(ns, s0 + kPartitionBlock / kSub);
        for (size_t s = s0; s < s1; ++s) {
            const size_t j0 = f + s * kSub, j1 = std::min(l, j0 + kSub);
            uint32_t x = 0, y = 0;
            for (size_t j = j0; j < j1; ++j) {
                x += (a[j].k < p) ? 0u : 1u;
                y += (p < a[j].k) ? 0u : 1u;
            }
            if (s == 0) x -= 1;  // the pivot's own position is not an L position
            cl[s + 1] = x;
            cr[s + 1] = y;
        }
    };
    Workers::get().run((n + kPartitionBlock - 1) / kPartitionBlock, count);
    for (size_t s = 0; s < ns; ++s) {
        cl[s + 1] += cl[s];
        cr[s + 1] += cr[s];
    }
    const size_t nl = cl[ns], nr = cr[ns];
    // u-th (1-based, from the left) L or R position, relative to f
    auto nth = [&](const std::vector<uint32_t> &c, size_t u, bool left) {
        const size_t s = static_cast<size_t>(std::upper_bound(c.begin(), c.end(), static_cast<uint32_t>(u - 1)) -
                                             c.begin()) - 1;
        size_t k = c[s];
        for (size_t j = s * kSub;; ++j) {
            const bool hit = left ? (j != 0 && !(a[f + j].k < p)) : !(p < a[f + j].k);
            if (hit && ++k == u) return j;
        }
    };
    auto L = [&](size_t t) { return nth(cl, t, true); };
    auto R = [&](size_t t) { return nth(cr, nr - t + 1, false); };
    auto crossed = [&](size_t t) { return t > nl || L(t) >= R(t); };
    size_t lo = 1, hi = nr;  // crossed(nr) holds: R_nr = f < any L
    while (lo < hi) {
        const size_t mid = lo + (hi - lo) / 2;
        if (crossed(mid)) hi = mid;
        else lo = mid + 1;
    }
    const size_t ts = lo;  // t*
    const size_t nsw = ts - 1;
    const size_t cut = std::min(ts > nl ? l : f + L(ts), ts == 1 ? l : f + R(ts - 1));
    const size_t nc = (nsw + kPartitionBlock - 1) / kPartitionBlock;
    std::vector<size_t> sl(nc), sr(nc);
    for (size_t c = 0; c < nc; ++c) {  // first L and R position of every chunk, before any swap
        sl[c] = L(c * kPartitionBlock + 1);
        sr[c] = R(c * kPartitionBlock + 1);
    }
    auto swaps = [&](size_t c) {  // the chunk's L and R positions gathered branch-free, then swapped
        thread_local std::vector<size_t> bl, br;
        const size_t m = std::min(nsw, (c + 1) * kPartitionBlock) - c * kPartitionBlock;
        bl.resize(m + 1);
        br.resize(m + 1);
        size_t x = 0, y = 0;
        for (size_t j = f + sl[c]; x < m; ++j) {
            bl[x] = j;
            x += (a[j].k < p) ? 0 : 1;
        }
        for (size_t j = f + sr[c]; y < m; --j) {
            br[y] = j;
            y += (p < a[j].k) ? 0 : 1;
        }
        for (size_t q = 0; q < m; ++q) std::swap(a[bl[q]], a[br[q]]);
    };
    Workers::get().run(nc, swaps);
    return cut;
}

int floor_log2(size_t n) {
    int r = 0;
    while (n > 1) {
        n >>= 1;
        ++r;
    }
    return r;
}

// Positions, after the partition phase, of the entries whose key equals v in [f, l).
void track_equal(KeyLabel *a, size_t f, size_t l, int depth, double v, std::vector<std::pair<size_t, int>> &out,
                 std::vector<uint32_t> &Lp, std::vector<uint32_t> &Rp) {
    auto count_in = [a, v](size_t x, size_t y) {
        size_t c = 0;
        for (size_t j = x; j < y; ++j) c += (a[j].k == v) ? 1 : 0;
        return c;
    };
    while (l - f > 16) {
        if (depth == 0) {
            std::partial_sort(a + f, a + l, a + l, [](const KeyLabel &x, const KeyLabel &y) { return x.k < y.k; });
            break;
        }
        --depth;
        const size_t cut =
            (l - f >= kParallelPartition) ? partition_pivot_parallel(a, f, l, Lp, Rp) : partition_pivot(a, f, l);
        // [f, cut) holds keys <= pivot and [cut, l) keys >= pivot (the pivot stays at f)
        const double p = a[f].k;
        const size_t left = (v < p) ? 1 : (p < v ? 0 : count_in(f, cut));
        const size_t right = (p < v) ? 1 : (v < p ? 0 : count_in(cut, l));
        if (right) track_equal(a, cut, l, depth, v, out, Lp, Rp);  // the right part is sorted first
        if (!left) return;
        l = cut;
    }
    for (size_t j = f; j < l; ++j)
        if (a[j].k == v) out.push_back({j, a[j].i});
}

}  // namespace

// Inside a run of equal keys the reference's std::sort permutation of the
// label-ordered keys decides (aggregat_list.cpp:109-123).
int PickTable::resolve_tie(size_t pos, const std::vector<int> &live, const std::vector<double> &ts_by_id) {
    // the run of equal keys around pos in the table order
    size_t k = 0;
    while (k + 1 < chunks_.size() && offset_[k + 1] <= pos) ++k;
    const double v = chunks_[k].tp[pos - offset_[k]];
    size_t p0 = pos;
    {
        size_t kk = k, j = pos - offset_[k];
        while (true) {
            if (j == 0) {
                if (kk == 0) break;
                --kk;
                j = chunks_[kk].tp.size();
            }
            if (chunks_[kk].tp[j - 1] != v) break;
            --j;
            --p0;
        }
    }
    std::vector<KeyLabel> &a = tie_keys_;
    a.resize(live.size());
    {
        // tp_t = factor / time_step in label order (aggregat_list.cpp:127-130)
        const size_t blk = 65536;
        auto fill = [this, &a, &live, &ts_by_id, blk](size_t b) {
            const size_t e = std::min(live.size(), (b + 1) * blk);
            for (size_t i = b * blk; i < e; ++i)
                a[i] = {factor_ / ts_by_id[static_cast<size_t>(live[i])], static_cast<int>(i)};
        };
        Workers::get().run((live.size() + blk - 1) / blk, fill);
    }
    std::vector<std::pair<size_t, int>> run;
    if (!a.empty()) track_equal(a.data(), 0, a.size(), 2 * floor_log2(a.size()), v, run, tie_l_, tie_r_);
    std::sort(run.begin(), run.end());
    if (pos - p0 >= run.size()) throw std::logic_error("pick table: tie run not found");
    return live[static_cast<size_t>(run[pos - p0].second)];
}

int PickTable::pick(double u, const std::vector<int> &live, const std::vector<double> &ts_by_id, bool *tie) {
    size_t pos = 0;
    int id = pick_fast(u, &pos);
    if (tie) *tie = id < 0;
    return id >= 0 ? id : resolve_tie(pos, live, ts_by_id);
}

void PickTable::dump(std::vector<double> &ts, std::vector<int> &id, std::vector<double> &cum) const {
    ts.clear();
    id.clear();
    cum.clear();
    for (const Chunk &c : chunks_)
        for (size_t i = 0; i < c.ts.size(); ++i) {
            ts.push_back(c.ts[i]);
            id.push_back(c.id[i]);
            cum.push_back(true_at(c, i));
        }
}

}  // namespace mcac
