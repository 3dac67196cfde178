// Time-step weighted pick table (a2): AggregatList::sort_time_steps,
// pick_random and get_time_step (aggregat_list.cpp:54-66, 109-141), with the
// same values as the reference and no O(N) work per merge.
//
// Live aggregates are kept in the order of descending time step (ids ascending
// inside equal steps); that is the reference's ascending order of
// tp_t = max/time_step, and because equal keys add equal values the reference's
// sequential sums cum[i] = cum[i-1] + tp[i] do not depend on how std::sort
// permutes a run of equal keys.  A pick that lands inside such a run is
// answered from the reference's own std::sort permutation.
//
// The order is split into chunks.  A chunk stores the sequential sums L of its
// entries taken from some start A.  With the true start S of the chunk,
// d = S - A is a multiple of the grid spacing of the binade both share, and a
// sequential FP64 sum shifted by such d stays shifted by exactly d, except at a
// step whose rounding was a tie, a step where L changes binade, or where L and
// L + d lie in different binades.  Those positions are stepped exactly; the
// rest of the chunk is L + d.  A merge edits at most three chunks, which are
// re-summed from their exact starts; the later chunks only get new offsets.
#pragma once
#include <cstddef>
#include <cstdint>
#include <functional>
#include <utility>
#include <vector>

namespace mcac {

// fn(k) for k in [0, n) on the host worker pool (up to 8 threads, the caller
// included); serially when another thread holds the pool.
void host_parallel_for(size_t n, const std::function<void(size_t)> &fn);

class PickTable {
  public:
    explicit PickTable(size_t chunk = 2048) : chunk_(chunk < 16 ? 16 : chunk) {}
    // ids in label order and time steps indexed by id; factor > 0: the keys
    // are factor / time step (the reference's max_time_step, which it
    // refreshes only at events), else the largest time step
    void build(const std::vector<int> &ids, const std::vector<double> &ts_by_id, double factor = 0.0);
    // keeper: old -> new time step; removed leaves the table
    void replace(int keep, double keep_old, double keep_new, int rem, double rem_ts);
    double factor() const { return factor_; }
    double total() const;
    size_t size() const { return n_; }
    // pick_random for the draw u; `live` holds ids in label order (for the tie fallback)
    int pick(double u, const std::vector<int> &live, const std::vector<double> &ts_by_id, bool *tie);
    // pick_random without the fallback: the id, or -1 with *pos set when the draw
    // lands in a run of equal keys (then resolve with resolve_tie)
    int pick_fast(double u, size_t *pos);
    int resolve_tie(size_t pos, const std::vector<int> &live, const std::vector<double> &ts_by_id);
    // full order and exact cumulative sums (tests)
    void dump(std::vector<double> &ts, std::vector<int> &id, std::vector<double> &cum) const;
    size_t chunks() const { return chunks_.size(); }
    uint64_t exact_steps() const { return exact_steps_; }
    uint64_t rescales() const { return rescales_; }  // merges that changed the largest time step

  private:
    struct Chunk {
        std::vector<double> ts, tp, L;
        std::vector<int> id;
        double A = 0.0;            // start of the stored sums
        std::vector<int> flags;    // local positions of ties and binade changes of L
        std::vector<int> seg;      // segment starts; true(i) = L[i] + segd[segment of i],
        std::vector<double> segd;  // or segd itself for an exactly stepped position
        std::vector<char> segabs;
        bool fast = true;          // true(i) = L[i] + fast_d over the whole chunk
        double fast_d = 0.0;
        double true_end = 0.0;
        bool dirty = true;         // entries changed: re-sum from the exact start
        size_t dirty_from = 0;     // first changed position (the sums before it from A still hold)
    };
    size_t chunk_;
    size_t n_ = 0;
    double factor_ = 0.0;
    std::vector<Chunk> chunks_;
    std::vector<size_t> offset_;   // first global position of each chunk
    uint64_t exact_steps_ = 0;
    uint64_t rescales_ = 0;
    std::vector<std::pair<size_t, size_t>> ranges_, merged_;  // scratch of settle()
  public:
    struct KeyLabel {
        double k;
        int i;
    };

  private:
    std::vector<KeyLabel> tie_keys_;  // scratch of resolve_tie()
    std::vector<uint32_t> tie_l_, tie_r_;  // scratch of its parallel partitions

    void materialize(Chunk &c, double A);
    void resum_from(Chunk &c, size_t from);
    void settle(Chunk &c, double S);
    void walk(size_t from);
    double true_at(const Chunk &c, size_t i) const;
    size_t find_chunk(double t, int id) const;
    void erase_entry(double t, int id, size_t &first_changed);
    void insert_entry(double t, int id, size_t &first_changed);
    void reindex();
};

}  // namespace mcac
