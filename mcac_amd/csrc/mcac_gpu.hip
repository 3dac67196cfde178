// MI355X (gfx950) implementation of the MCAC aggregate-move / collision-search
// path behind the C-ABI declared in include/mcac_gpu.h.
//
// HBM layout of one box (box_state.hpp):
//   spheres     SoA, index order: x y z r vol surf rx ry rz          (72 B/sphere)
//   aggregates  SoA, stable-id order: x y z rx ry rz rmax volume surface rg,
//               merge caches, member segment, reference Verlet cell (-1 dead)
//   members     pool of sphere indices; aggregate a owns [start, start+cap)
//   cell list   CSR by reference Verlet cell + dirty list (kernels_cells.hpp)
// One MC step is two launches (k_filter, k_solve; plus k_translate for movers
// of more than 1024 spheres) and no copy: the last block publishes a small
// record into pinned host memory, which the host polls by sequence number.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/mcac_gpu.h"
#include "kernels_cells.hpp"
#include "kernels_growth.hpp"
#include "kernels_merge.hpp"
#include "kernels_search.hpp"

using namespace mcac_dev;

struct mcac_box {
    mcac_box_params p{};
    hipStream_t stream = nullptr;
    bool own_stream = true;          // false once the box works on a group's stream
    mcac_group *group = nullptr;
    int64_t group_call = -1;         // last group enqueue that took a move of this box
    DevBox d{};
    int64_t n_spheres = 0, n_aggs = 0;
    int64_t cand_cap = 0, part_cap = 0, pool_cap = 0;
    int n_cells = 0, n_built = 0;
    int *h_cstart = nullptr;         // pinned copy of the cell starts of the current build (small grids)
    bool cstart_ok = false;
    bool verify_cells = false;       // MCAC_VERIFY_CELLS=1: host-counted cell starts checked against the device's
    bool grid_solve = true;          // large movers in one launch (k_step_grid); MCAC_GRID_SOLVE=0: k_filter + k_solve
    int grid_blocks = 32;            // its largest grid (MCAC_GRID_BLOCKS, 1..kFilterBlocks)
    int *h_alive = nullptr;          // pinned staging of the live ids uploaded by a rebuild
    int64_t alive_cap = 0;
    hipEvent_t alive_read = nullptr; // recorded after the upload that reads h_alive
    std::vector<int> cell_fill;      // host cell counts of a rebuild
    // host mirrors for allocation, launch sizing and the dirty list
    std::vector<int> count, start, cap;
    std::vector<int> alive_ids;  // ascending, may hold dead ids until rebuilt
    // host mirrors of the device aggregate state: position, relative CoM and
    // rmax are exact for every aggregate that no queued step moves (pend == 0);
    // the dirty flags and list length are those after every queued step that
    // will run (steps known to be skipped are taken out)
    std::vector<double> h_x, h_y, h_z, h_rx, h_ry, h_rz, h_rmax;
    std::vector<int> pend;
    std::vector<unsigned char> h_dirty;
    std::vector<int> h_dslot;        // dirty-list index of each dirty aggregate
    int64_t dirty_count = 0;         // device dirty-list length after the queued steps
    int64_t dirty_epoch = 0;         // cell-list builds so far
    bool gate_closed = false;        // a consumed queued step found a contact; not released yet
    unsigned long long released = 0; // gate values the steps enqueued from now on ignore
    unsigned long long last_gate = 0;  // seq of the last consumed step that closed the gate
    int automerge = 0;               // queued moves merge in the same launch on contact
    int fused_merge = 1;             // overlap-correction merges of <= kFusedMerge spheres in one launch
    bool inline_pending = false;     // a contact merged in its launch; mcac_merge not called yet
    StepRecord inline_rec{};
    int64_t inline_src = -1;
    int64_t inline_epoch = 0;
    int64_t n_inline = 0, n_host_merge = 0;
    int64_t dead_in_list = 0;
    int64_t pool_used = 0;
    // pinned records (host pointer, device pointer)
    static constexpr int kRing = 16;  // steps in flight (two group launches of 8 per box)
    StepRecord *h_step = nullptr, *d_step = nullptr;  // kRing records
    int inflight = 0;
    unsigned long long oldest = 1;   // oldest ticket not yet consumed
    int slot_apply[kRing] = {0}, slot_translate[kRing] = {0};
    int64_t slot_agg[kRing] = {0};
    SearchArgs slot_args[kRing]{};   // arguments of queued searches (second-launch solves)
    int slot_dirty_new[kRing] = {0}; // 1: the step's move appends its mover to the dirty list
    int64_t slot_epoch[kRing] = {0};
    int slot_void[kRing] = {0};      // known to be skipped: its predictions were taken out
    int slot_fused[kRing] = {0};
    int slot_solving[kRing] = {0};   // its k_solve was launched before a cell-list rebuild
    int max_count = 0;               // upper bound of the aggregate sizes
    int64_t n_handover = 0;          // k_step searches finished by k_solve
    int64_t n_fused = 0, n_split = 0;  // searches launched as k_step / k_filter + k_solve
    int64_t n_compactions = 0;
    MergeRecord *h_merge = nullptr, *d_merge = nullptr;
    // alphas: the merge state and the merged aggregate's neighbour-map walk
    // rows, copied back into pinned memory (grown to the largest merge seen)
    unsigned char *h_alpha = nullptr;
    size_t h_alpha_cap = 0;
    unsigned long long seq = 0;        // step tickets (searches)
    unsigned long long merge_seq = 0;  // merge records
    std::vector<void *> allocs;
    int timing = 0;                  // 0 off; n >= 1: every n-th search is timed
    unsigned long long timing_base = 1;  // first ticket after timing was set: timed
    int slot_timed[kRing] = {0};     // 1 timed, 2 timed and consumed, events not read yet
    hipEvent_t ev[4]{};
    hipEvent_t sev[kRing][4]{};      // per-slot events of pipelined steps
    double t_search = 0, t_translate = 0, t_merge = 0, t_filter = 0, t_solve = 0;
    double n_search = 0, n_translate = 0, n_merge = 0;
    int64_t n_rebuilds = 0;
    // persistent step engine (k_step_engine): steps posted to a ring in pinned
    // memory instead of one launch each, while eng_running
    int eng_blocks = 0;               // grid of the engine (0: one launch per step)
    int eng_grid = 0;                 // grid of the resident engine
    bool eng_running = false;         // an engine grid was launched and no exit job posted since
    unsigned long long eng_next = 1;  // next job index
    EngineJob *h_ring = nullptr, *dp_ring = nullptr;
    EngineCtl *h_ctl = nullptr, *dp_ctl = nullptr;
    unsigned long long eng_idle_ticks = 0;  // device clock ticks an idle grid waits before it ends
    double eng_clock_khz = 0;
    int64_t eng_launches = 0, eng_posted = 0, eng_relaunches = 0;
    // surface growth (mcac_box_grow): per-aggregate results, device and pinned
    // host copies (5 arrays of the live-list length), allocated at the first call
    double *d_grow = nullptr, *h_grow = nullptr;
    int64_t grow_cap = 0;
    bool cells_stale = false;    // positions changed everywhere: rebuild the cell list before the next step
    bool mirrors_exact = true;   // the host mirrors of positions and rmax hold every aggregate exactly
};

// Boxes that share one stream; their queued steps go into k_step_batch
// launches, one block per box.  Jobs and block segments are written to pinned
// memory, in turn into kBuffers job lists, so that a list is written while the
// launches that read the others are still queued.
struct mcac_group {
    static constexpr int kMovesPerBox = 8;  // = kMaxBlockJobs, steps of a box in one launch
    static constexpr int kBuffers = 3;      // job lists
    int device = 0;
    hipStream_t stream = nullptr;
    int cap = 0;                           // boxes
    int job_cap = 0;                       // jobs per list
    int parity = 0;                        // the list written next
    StepJob *h_jobs = nullptr, *d_jobs = nullptr;  // kBuffers * job_cap jobs
    int *h_seg = nullptr, *d_seg = nullptr;        // kBuffers * (job_cap + 1) segment starts
    hipEvent_t read[kBuffers]{};                    // recorded after the launch reading each list
    bool pending[kBuffers] = {};
    int64_t calls = 0;
    int64_t launches = 0, batched_steps = 0, single_steps = 0;
    // host seconds in mcac_move_enqueue_group: in all, in cell-list rebuilds
    // (with the waits they need), in launches, in waits for a list's reader
    double t_enqueue = 0, t_rebuild = 0, t_launch = 0, t_list = 0;
};

namespace {

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            std::fprintf(stderr, "mcac_gpu: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
            return MCAC_E_DEVICE;                                                             \
        }                                                                                     \
    } while (0)

template <class T> int dalloc(mcac_box *b, T **ptr, int64_t n) {
    void *p = nullptr;
    size_t bytes = static_cast<size_t>(n < 1 ? 1 : n) * sizeof(T);
    HIP_TRY(hipMalloc(&p, bytes));
    HIP_TRY(hipMemset(p, 0, bytes));
    b->allocs.push_back(p);
    *ptr = static_cast<T *>(p);
    return 0;
}

int engine_start(mcac_box *b, int grid);

// Waits until a kernel has published sequence number `seq` into pinned memory.
int wait_seq(mcac_box *b, const unsigned long long *seqp, unsigned long long seq) {
    auto t0 = std::chrono::steady_clock::now();
    unsigned spins = 0;
    while (__atomic_load_n(seqp, __ATOMIC_ACQUIRE) != seq) {
        if ((++spins & 1023u) == 0) {
            hipError_t e = hipStreamQuery(b->stream);
            if (e == hipSuccess && b->eng_running) {
                // an engine grid ended idle before its last job was posted: launch it again
                int rc = engine_start(b, b->eng_grid);
                if (rc) return rc;
                continue;
            }
            if (e == hipSuccess) {
                if (__atomic_load_n(seqp, __ATOMIC_ACQUIRE) == seq) break;
                unsigned int grid_fail = 0;
                (void)hipMemcpy(&grid_fail, &b->d.ss->grid_fail, sizeof grid_fail, hipMemcpyDeviceToHost);
                std::fprintf(stderr, "mcac_gpu: stream idle without a published record%s\n",
                             grid_fail ? " (a large-mover search gave up waiting at its grid-wide barrier)" : "");
                return MCAC_E_DEVICE;
            }
            if (e != hipErrorNotReady) {
                std::fprintf(stderr, "mcac_gpu: stream error: %s\n", hipGetErrorString(e));
                return MCAC_E_DEVICE;
            }
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(120)) {
                std::fprintf(stderr, "mcac_gpu: no record after 120 s\n");
                return MCAC_E_DEVICE;
            }
        }
        __builtin_ia32_pause();
    }
    return 0;
}

// Persistent step engine.  The host posts a job (one queued step) into the
// ring with its index written last; engine_stop posts the exit job, after
// which the grid ends, so that work enqueued later on the box stream runs
// after the engine's steps.  Nothing here waits for the device.
void engine_post(mcac_box *b, const SearchArgs *a, const CellTable *ct, int vg) {
    const unsigned long long k = b->eng_next++;
    // the slot's previous job, k - kEngineRing, was taken long ago: at most
    // kRing steps are in flight and a grid takes jobs in order
    EngineJob &j = b->h_ring[k % kEngineRing];
    if (a) {
        j.a = *a;
        j.cells.n = 0;
        if (ct && ct->n > 0) {
            j.cells.n = ct->n;
            std::memcpy(j.cells.c0, ct->c0, static_cast<size_t>(ct->n) * sizeof(int));
            std::memcpy(j.cells.sz, ct->sz, static_cast<size_t>(ct->n) * sizeof(int));
        }
        j.cmd = kEngineRun;
        j.vg = vg;
        b->eng_posted += 1;
    } else {
        j.cmd = kEngineExit;
    }
    __atomic_store_n(&j.seq, k, __ATOMIC_RELEASE);
}

void engine_stop(mcac_box *b) {
    if (!b->eng_running) return;
    engine_post(b, nullptr, nullptr, 0);
    b->eng_running = false;
}

// The box stream for work other than engine steps: a resident engine is told
// to end first.
hipStream_t bstream(mcac_box *b) {
    engine_stop(b);
    return b->stream;
}

// A resident grid that ended on its own (no job for its idle period) while the
// host still counts it as running: its first job not taken.
bool engine_ended(const mcac_box *b, unsigned long long *first) {
    if (!b->eng_running) return false;
    const unsigned long long ended = __atomic_load_n(&b->h_ctl->launches_ended, __ATOMIC_ACQUIRE);
    if (ended < static_cast<unsigned long long>(b->eng_launches)) return false;
    *first = b->h_ctl->stopped_at;
    return true;
}

// Makes sure an engine grid of `grid` blocks is resident, launching one (from
// the first job not taken) when none is or when it is smaller.
int engine_start(mcac_box *b, int grid) {
    unsigned long long first = b->eng_next;
    if (b->eng_running) {
        unsigned long long f = 0;
        if (engine_ended(b, &f)) {
            if (b->h_ctl->reason < 0) {
                std::fprintf(stderr, "mcac_gpu: engine grid ended on a wait that ran out\n");
                return MCAC_E_DEVICE;
            }
            first = f;  // jobs from there on were posted but not taken
            b->eng_running = false;
            b->eng_relaunches += 1;
        } else if (grid <= b->eng_grid) {
            return 0;
        } else {
            engine_stop(b);  // a larger grid takes over at the next job
            first = b->eng_next;
        }
    }
    HIP_TRY(hipMemsetAsync(&b->d.ss->eng_decision, 0, sizeof(unsigned long long), b->stream));
    b->eng_launches += 1;
    hipLaunchKernelGGL(k_step_engine, dim3(static_cast<unsigned>(grid)), dim3(kBlock), 0, b->stream, b->d, b->dp_ring,
                       b->dp_ctl, first, b->eng_idle_ticks, static_cast<unsigned long long>(b->eng_launches));
    HIP_TRY(hipGetLastError());
    b->eng_running = true;
    b->eng_grid = grid;
    return 0;
}

int rebuild_alive(mcac_box *b) {
    std::vector<int> ids;
    ids.reserve(b->alive_ids.size());
    for (int id : b->alive_ids)
        if (b->count[static_cast<size_t>(id)] > 0) ids.push_back(id);
    b->alive_ids.swap(ids);
    b->dead_in_list = 0;
    const size_t n = b->alive_ids.size();
    if (n == 0) return 0;
    if (b->h_alive && static_cast<int64_t>(n) <= b->alive_cap) {
        // through pinned staging, in stream order: no wait unless the last
        // upload has not read the staging yet
        HIP_TRY(hipEventSynchronize(b->alive_read));
        std::memcpy(b->h_alive, b->alive_ids.data(), n * sizeof(int));
        HIP_TRY(hipMemcpyAsync(b->d.alive, b->h_alive, n * sizeof(int), hipMemcpyHostToDevice, bstream(b)));
        HIP_TRY(hipEventRecord(b->alive_read, bstream(b)));
        return 0;
    }
    HIP_TRY(hipMemcpyAsync(b->d.alive, b->alive_ids.data(), n * sizeof(int), hipMemcpyHostToDevice, bstream(b)));
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    return 0;
}

// Cell starts of a build counted on the host from its position mirrors (the
// device files each live aggregate under ref_cell of its CoM, the same
// function on the same doubles).  Only when the mirrors hold every aggregate
// and no step is in flight.
void count_cells_on_host(mcac_box *b) {
    const int n = static_cast<int>(b->p.n_div);
    b->cell_fill.assign(static_cast<size_t>(b->n_cells), 0);
    for (int id : b->alive_ids) {
        const size_t a = static_cast<size_t>(id);
        if (b->count[a] <= 0) continue;
        b->cell_fill[static_cast<size_t>(ref_cell(b->h_x[a], b->h_y[a], b->h_z[a], n, b->p.box_length))] += 1;
    }
    int acc = 0;
    for (int c = 0; c < b->n_cells; ++c) {
        b->h_cstart[c] = acc;
        acc += b->cell_fill[static_cast<size_t>(c)];
    }
    b->h_cstart[b->n_cells] = acc;
}

// CSR cell list from the live ids; clears the dirty list (host and device).
int rebuild_cells(mcac_box *b) {
    if (b->dead_in_list > 0) {
        int rc = rebuild_alive(b);
        if (rc) return rc;
    }
    const int n_alive = static_cast<int>(b->alive_ids.size());
    if (b->n_cells <= kSmallCells && n_alive <= kSmallBuild) {
        hipLaunchKernelGGL(k_cell_build_small, dim3(1), dim3(kBuildThreads), 0, bstream(b), b->d, n_alive, b->n_cells);
    } else {
        const unsigned g =
            static_cast<unsigned>(std::min<int64_t>(2048, std::max<int64_t>(1, (n_alive + kBlock - 1) / kBlock)));
        const unsigned gc = static_cast<unsigned>(std::min<int64_t>(2048, (b->n_cells + kBlock) / kBlock));
        hipLaunchKernelGGL(k_cell_clear, dim3(gc), dim3(kBlock), 0, bstream(b), b->d, b->n_cells);
        hipLaunchKernelGGL(k_cell_count, dim3(g), dim3(kBlock), 0, bstream(b), b->d, n_alive);
        hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, bstream(b), b->d, b->n_cells);
        hipLaunchKernelGGL(k_cell_scatter, dim3(g), dim3(kBlock), 0, bstream(b), b->d, n_alive);
    }
    HIP_TRY(hipGetLastError());
    b->dirty_count = 0;
    std::fill(b->h_dirty.begin(), b->h_dirty.end(), 0);
    b->dirty_epoch += 1;
    b->n_built = n_alive;
    b->n_rebuilds++;
    b->cells_stale = false;
    b->cstart_ok = false;
    if (b->h_cstart) {
        // the host passes the visited cells' bounds with each step: counted
        // here when the mirrors allow, else read back from the build
        const bool on_host = b->mirrors_exact && b->inflight == 0;
        if (on_host) count_cells_on_host(b);
        if (!on_host || b->verify_cells) {
            std::vector<int> host;
            if (on_host) host.assign(b->h_cstart, b->h_cstart + b->n_cells + 1);
            HIP_TRY(hipMemcpyAsync(b->h_cstart, b->d.cstart, (static_cast<size_t>(b->n_cells) + 1) * sizeof(int),
                                   hipMemcpyDeviceToHost, bstream(b)));
            HIP_TRY(hipStreamSynchronize(bstream(b)));
            if (on_host && !std::equal(host.begin(), host.end(), b->h_cstart)) {
                std::fprintf(stderr, "mcac_gpu: cell starts counted on the host differ from the device's\n");
                return MCAC_UNKNOWN_ERROR;
            }
        }
        b->cstart_ok = true;
    }
    return 0;
}


int64_t rebuild_threshold(const mcac_box *b) {
    int64_t n = static_cast<int64_t>(b->alive_ids.size()) - b->dead_in_list;
    int64_t r = static_cast<int64_t>(std::sqrt(6.0 * static_cast<double>(std::max<int64_t>(n, 1))));
    return std::max<int64_t>(256, r);
}

int compact_pool(mcac_box *b, int64_t need) {
    b->n_compactions += 1;
    // new layout: every live aggregate gets capacity 2*count (at least 4)
    std::vector<int> ids, nstart;
    int64_t used = 0;
    for (size_t id = 0; id < b->count.size(); ++id) {
        if (b->count[id] <= 0) continue;
        ids.push_back(static_cast<int>(id));
        nstart.push_back(static_cast<int>(used));
        used += std::max<int64_t>(4, 2 * static_cast<int64_t>(b->count[id]));
    }
    int64_t cap = b->pool_cap;
    while (used + need > cap) cap *= 2;
    if (cap > INT32_MAX) return MCAC_E_CAPACITY;
    int *nm = nullptr, *d_ids = nullptr, *d_ns = nullptr, *d_os = nullptr;
    HIP_TRY(hipMalloc(&nm, static_cast<size_t>(cap) * sizeof(int)));
    HIP_TRY(hipMalloc(&d_ids, ids.size() * sizeof(int) + 4));
    HIP_TRY(hipMalloc(&d_ns, ids.size() * sizeof(int) + 4));
    HIP_TRY(hipMalloc(&d_os, b->start.size() * sizeof(int) + 4));
    HIP_TRY(hipMemcpy(d_ids, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_ns, nstart.data(), nstart.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_os, b->start.data(), b->start.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!ids.empty())
        hipLaunchKernelGGL(k_compact, dim3(static_cast<unsigned>(std::min<size_t>(ids.size(), 65535))), dim3(kBlock), 0,
                           bstream(b), b->d.members, nm, d_os, d_ns, b->d.acount, d_ids, static_cast<int>(ids.size()));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    for (size_t a = 0; a < ids.size(); ++a) {
        const size_t id = static_cast<size_t>(ids[a]);
        b->start[id] = nstart[a];
        b->cap[id] = static_cast<int>(std::max<int64_t>(4, 2 * static_cast<int64_t>(b->count[id])));
    }
    HIP_TRY(hipMemcpy(b->d.astart, b->start.data(), b->start.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->d.acap, b->cap.data(), b->cap.size() * sizeof(int), hipMemcpyHostToDevice));
    for (auto &p : b->allocs)
        if (p == b->d.members) p = nm;
    HIP_TRY(hipFree(b->d.members));
    b->d.members = nm;
    b->pool_cap = cap;
    b->pool_used = used;
    HIP_TRY(hipFree(d_ids));
    HIP_TRY(hipFree(d_ns));
    HIP_TRY(hipFree(d_os));
    return 0;
}

int harvest_timing(mcac_box *b, int slot);

// A step handed over to k_solve finishes (moves its aggregate, drops its cell
// entry and writes its dirty-list entry at the slot predicted for the current
// cell list) only when its k_solve runs, which the host launches when it reads
// the pending record.  A cell-list rebuild must not run in between: before a
// rebuild with steps in flight, every queued kernel is drained and the
// k_solve of a pending step (at most one: it closes the gate) is launched now.
int settle_handovers(mcac_box *b) {
    if (b->inflight == 0) return 0;
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    for (unsigned long long t = b->oldest; t <= b->seq; ++t) {
        const int slot = static_cast<int>(t % mcac_box::kRing);
        if (b->slot_void[slot] || b->slot_solving[slot]) continue;
        StepRecord &r = b->h_step[slot];
        if (__atomic_load_n(&r.seq, __ATOMIC_ACQUIRE) != t || r.skipped || r.status != kPendingSolve) continue;
        __atomic_store_n(&r.seq, 0ull, __ATOMIC_RELEASE);  // the k_solve publishes the final record
        hipLaunchKernelGGL(k_solve, dim3(kSolveBlocks), dim3(kBlock), 0, bstream(b), b->d, b->slot_args[slot]);
        HIP_TRY(hipGetLastError());
        b->slot_solving[slot] = 1;
        b->n_handover += 1;
    }
    return 0;
}

// A step that one block can take whole: the one-launch form (small mover, few
// sphere pairs possible), the mover moved by the finishing block, not timed.
bool one_block_step(const mcac_box *b, int n_mov, bool timed) {
    const bool fused = n_mov <= kStepMover && static_cast<long long>(n_mov) * b->max_count <= kStepPairs;
    return fused && n_mov <= 4 * kBlock && !timed;
}

// Entries a step may have to sweep (cells of a window for a mover of rmax at
// most `rmax`, at the mean occupancy, plus the dirty list), and the bound
// under which one workgroup takes a step in a group launch: with a step's
// entries one or a few per thread, one block is faster than a grid, whose
// cross-block ticket and cache fences cost more than the work they split.
constexpr double kBlockItems = 4096.0;
double window_items(const mcac_box *b, double reach) {
    const double cell = b->p.box_length / static_cast<double>(b->p.n_div);
    const double per_axis = std::min<double>(static_cast<double>(b->p.n_div), std::ceil(2.0 * reach / cell) + 2.0);
    const double per_cell = b->n_cells > 0 ? static_cast<double>(b->n_built) / b->n_cells : 0.0;
    return per_cell * per_axis * per_axis * per_axis + static_cast<double>(b->dirty_count) + 1.0;
}

// The next enqueued step rebuilds the cell list first (kernels on the stream).
bool rebuild_due(const mcac_box *b) { return b->dirty_count + 3 >= rebuild_threshold(b); }

bool next_step_timed(const mcac_box *b) {
    return b->timing > 0 && (b->seq + 1 - b->timing_base) % static_cast<unsigned long long>(b->timing) == 0;
}

// Picks the launch of one step and enqueues it.  With job_out (given only for
// a step that one block can take whole: one-launch form, mover moved by the
// finishing block, not timed, few window entries) the step is written there
// instead, for a k_step_batch launch by the caller (*batched = true).
int enqueue_search(mcac_box *b, int64_t agg, const double dir[3], double lpm, double maxradius, int apply, int gated,
                   unsigned long long *ticket, StepJob *job_out = nullptr, bool *batched = nullptr) {
    if (batched) *batched = false;
    if (agg < 0 || agg >= b->n_aggs || b->count[static_cast<size_t>(agg)] <= 0) return MCAC_E_ARGUMENT;
    if (b->inflight >= mcac_box::kRing) return MCAC_E_CAPACITY;
    if (b->inline_pending) return MCAC_E_ARGUMENT;  // the caller merges the contact first
    if (rebuild_due(b) || b->cells_stale) {
        int rc = settle_handovers(b);
        if (rc == 0) rc = rebuild_cells(b);
        if (rc) return rc;
    }
    const size_t A = static_cast<size_t>(agg);
    const int n_mov = b->count[A];
    const unsigned long long seq = ++b->seq;
    const int slot = static_cast<int>(seq % mcac_box::kRing);
    SearchArgs a{};
    a.src = static_cast<int>(agg);
    a.n_mov = n_mov;
    a.start = b->start[A];
    // a queued step behind a known closed gate does nothing: no predictions
    const bool doomed = gated && b->gate_closed;
    if (b->pend[A] == 0 && b->mirrors_exact) {
        a.known = 1;
        a.sx = b->h_x[A];
        a.sy = b->h_y[A];
        a.sz = b->h_z[A];
        a.srmax = b->h_rmax[A];
        a.srx = b->h_rx[A];
        a.sry = b->h_ry[A];
        a.srz = b->h_rz[A];
        const double src[3] = {a.sx, a.sy, a.sz};
        const double vec[3] = {lpm * dir[0], lpm * dir[1], lpm * dir[2]};
        verlet_window(src, vec, a.srmax + maxradius, static_cast<int>(b->p.n_div), b->p.box_length, a.b1, a.len);
        // the move without contact (finish_step: calcul.cpp:128-130, aggregat.cpp:148-153)
        for (int l = 0; l < 3; ++l) {
            a.nc_com[l] = periodic_position(src[l] + dir[l] * lpm, b->p.box_length);
            a.nc_ref[l] = a.nc_com[l] - (l == 0 ? a.srx : (l == 1 ? a.sry : a.srz));
        }
        a.nc_cell = ref_cell(a.nc_com[0], a.nc_com[1], a.nc_com[2], static_cast<int>(b->p.n_div), b->p.box_length);
        a.nc_ok = 1;
    }
    a.dirty_n = static_cast<int>(b->dirty_count);
    a.dirty_slot = -1;
    a.dirty_pos = 0;
    int dirty_new = 0;
    if (apply && !doomed) {
        if (!b->h_dirty[A]) {
            a.dirty_slot = static_cast<int>(b->dirty_count);
            b->h_dirty[A] = 1;
            b->h_dslot[A] = a.dirty_slot;
            b->dirty_count += 1;
            dirty_new = 1;
        }
        a.dirty_pos = b->h_dslot[A];
        b->pend[A] += 1;
    }
    a.n_div = static_cast<int>(b->p.n_div);
    a.cap = static_cast<int>(b->cand_cap);
    a.n_built = b->n_built;
    a.part_cap = b->part_cap;
    a.dir[0] = dir[0];
    a.dir[1] = dir[1];
    a.dir[2] = dir[2];
    a.lpm = lpm;
    a.maxradius = maxradius;
    a.L = b->p.box_length;
    a.apply = apply;
    a.translate_in_finish = n_mov <= 4 * kBlock;
    a.out = b->d_step + slot;
    a.seq = seq;
    a.gated = gated;
    a.released = b->released;
    a.automerge = (b->automerge && gated && apply) ? 1 : 0;
    // FP32 broad phase only where the minimum image decides (sweep + radii < 0.4 L)
    a.prec = (b->p.precision > 0 && a.known && lpm + a.srmax + maxradius < 0.4 * b->p.box_length)
                 ? static_cast<int>(std::min<int32_t>(b->p.precision, 2))
                 : 0;
    a.margin32 = static_cast<float>(1e-5 * b->p.box_length);
    a.pool_used = b->pool_used;
    a.pool_cap = b->pool_cap;
    if (b->slot_timed[slot] == 2) {
        int rc = harvest_timing(b, slot);
        if (rc) return rc;
    }
    __atomic_store_n(&b->h_step[slot].seq, 0ull, __ATOMIC_RELEASE);
    hipEvent_t *ev = b->sev[slot];
    const bool timed = b->timing > 0 && (seq - b->timing_base) % static_cast<unsigned long long>(b->timing) == 0;
    // one launch when the mover is small enough for the finishing block
    const bool fused = n_mov <= kStepMover && static_cast<long long>(n_mov) * b->max_count <= kStepPairs;
    // filter grid sized to the expected work: entries of the visited cells
    // (1.5x the mean cell occupancy) plus the dirty list, one per thread
    int grid = kFilterBlocks;
    if (a.known && b->n_cells > 0) {
        const double per_cell = static_cast<double>(b->n_built) / b->n_cells;
        const double items = 1.5 * per_cell * a.len[0] * a.len[1] * a.len[2] + a.dirty_n + 1;
        grid = static_cast<int>(std::min<double>(kFilterBlocks, std::ceil(items / kBlock)));
        grid = std::max(grid, 1);
    }
    // timed launches carry start/stop events that the runtime stamps with the
    // kernel's own execution interval; the queue is drained first so that the
    // interval does not include waiting behind the steps queued before it
    if (timed && b->inflight > 0) HIP_TRY(hipStreamSynchronize(bstream(b)));
    // the persistent engine takes one-launch steps whose mover the finishing
    // block moves (not timed: a timed step is its own launch between events)
    const bool engine = !job_out && b->eng_blocks > 0 && !b->group && fused && !timed && gated &&
                        (!apply || a.translate_in_finish);
    if (job_out || engine) {
        // queued step arguments sit in LDS: the host adds the bounds of the
        // visited cells, saving the kernel a dependent load (a launch of its
        // own reads them faster from cstart than from its arguments)
        const int V = a.len[0] * a.len[1] * a.len[2];
        CellTable ct_local;
        CellTable &ct = job_out ? job_out->cells : ct_local;
        ct.n = 0;
        if (a.known && b->cstart_ok && V <= 64) {
            const int n = static_cast<int>(b->p.n_div), lj = a.len[1], lk = a.len[2];
            for (int v = 0; v < V; ++v) {
                const int oi = v / (lj * lk), rem = v - oi * (lj * lk), oj = rem / lk, ok = rem - oj * lk;
                const int c = (window_offset(a.b1[0] + oi, 0, n) * n + window_offset(a.b1[1] + oj, 0, n)) * n +
                              window_offset(a.b1[2] + ok, 0, n);
                ct.c0[v] = b->h_cstart[c];
                ct.sz[v] = b->h_cstart[c + 1] - b->h_cstart[c];
            }
            ct.n = V;
        }
        if (job_out) {
            job_out->b = b->d;
            job_out->a = a;
            *batched = true;
        } else {
            // the resident grid has eng_blocks workgroups; the step takes the
            // share its window needs, as a launch of its own would
            int rc = engine_start(b, b->eng_blocks);
            if (rc) return rc;
            engine_post(b, &a, &ct, std::min(grid, b->eng_blocks));
        }
    } else if (fused) {
        if (timed) hipExtLaunchKernelGGL(k_step, dim3(grid), dim3(kBlock), 0, bstream(b), ev[0], ev[1], 0, b->d, a);
        else hipLaunchKernelGGL(k_step, dim3(grid), dim3(kBlock), 0, bstream(b), b->d, a);
    } else if (b->grid_solve) {
        // one launch: the filter grid, widened to the solve chunks that a
        // survivor as large as the largest aggregate would bring, at most
        // grid_blocks (a block of it holds 81 KB of LDS, one per CU, all of
        // them resident together for its barrier, whose arrivals are atomics
        // on one counter)
        const double chunks = std::ceil(static_cast<double>(n_mov) * static_cast<double>(b->max_count) / kChunk);
        const int g2 = std::min(b->grid_blocks, std::max(grid, static_cast<int>(std::min<double>(kFilterBlocks, chunks))));
        if (timed) hipExtLaunchKernelGGL(k_step_grid, dim3(g2), dim3(kBlock), 0, bstream(b), ev[0], ev[1], 0, b->d, a);
        else hipLaunchKernelGGL(k_step_grid, dim3(g2), dim3(kBlock), 0, bstream(b), b->d, a);
    } else if (timed) {
        hipExtLaunchKernelGGL(k_filter, dim3(grid), dim3(kBlock), 0, bstream(b), ev[0], ev[1], 0, b->d, a);
        hipExtLaunchKernelGGL(k_solve, dim3(kSolveBlocks), dim3(kBlock), 0, bstream(b), ev[2], ev[3], 0, b->d, a);
    } else {
        hipLaunchKernelGGL(k_filter, dim3(grid), dim3(kBlock), 0, bstream(b), b->d, a);
        hipLaunchKernelGGL(k_solve, dim3(kSolveBlocks), dim3(kBlock), 0, bstream(b), b->d, a);
    }
    if (apply && !a.translate_in_finish) {
        const int tb = std::min(1024, (n_mov + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_translate, dim3(tb), dim3(kBlock), 0, bstream(b), b->d, a.src, gated, seq, a.released);
    }
    b->slot_args[slot] = a;
    b->slot_fused[slot] = (fused || b->grid_solve) ? 1 : 0;  // one launch: its events are the first pair
    (fused ? b->n_fused : b->n_split) += 1;
    b->slot_timed[slot] = timed ? 1 : 0;
    HIP_TRY(hipGetLastError());
    b->slot_apply[slot] = apply;
    b->slot_dirty_new[slot] = dirty_new;
    b->slot_epoch[slot] = b->dirty_epoch;
    b->slot_void[slot] = doomed ? 1 : 0;
    b->slot_solving[slot] = 0;
    b->slot_agg[slot] = agg;
    b->slot_translate[slot] = apply && !a.translate_in_finish;
    b->inflight += 1;
    *ticket = seq;
    return 0;
}

// Takes the predictions of a queued step that will not run out of the mirrors.
void take_out(mcac_box *b, int slot) {
    if (b->slot_void[slot]) return;
    b->slot_void[slot] = 1;
    if (!b->slot_apply[slot]) return;
    const size_t A = static_cast<size_t>(b->slot_agg[slot]);
    b->pend[A] -= 1;
    if (b->slot_dirty_new[slot] && b->slot_epoch[slot] == b->dirty_epoch) {
        b->h_dirty[A] = 0;
        b->dirty_count -= 1;
    }
}

// Adds the event times of a consumed, timed step to the totals.
int harvest_timing(mcac_box *b, int slot) {
    hipEvent_t *ev = b->sev[slot];
    const bool fused = b->slot_fused[slot] != 0;
    HIP_TRY(hipEventSynchronize(ev[fused ? 1 : 3]));
    float t01 = 0, t23 = 0;
    HIP_TRY(hipEventElapsedTime(&t01, ev[0], ev[1]));
    if (!fused) HIP_TRY(hipEventElapsedTime(&t23, ev[2], ev[3]));
    b->t_filter += t01;
    b->t_solve += t23;
    b->t_search += t01 + t23;
    b->n_search += 1;
    b->slot_timed[slot] = 0;
    return 0;
}

int wait_search(mcac_box *b, unsigned long long ticket, int64_t agg, mcac_contact *out, int *skipped) {
    if (ticket != b->oldest || b->inflight <= 0) return MCAC_E_ARGUMENT;  // results are read in order
    const int slot = static_cast<int>(ticket % mcac_box::kRing);
    if (b->slot_void[slot]) {
        // queued behind a closed gate: the device skips it and writes no record
        b->inflight -= 1;
        b->oldest = ticket + 1;
        b->slot_timed[slot] = 0;
        if (skipped) *skipped = 1;
        if (out) {
            std::memset(out, 0, sizeof(*out));
            out->moving_agg = agg;
            out->moving_sphere = out->other_sphere = out->other_agg = -1;
        }
        return 0;
    }
    int rc = wait_seq(b, &b->h_step[slot].seq, ticket);
    if (rc) return rc;
    bool handed_over = b->slot_solving[slot] != 0;  // its k_solve already ran (settle_handovers)
    b->slot_solving[slot] = 0;
    if (!handed_over && !b->h_step[slot].skipped && b->h_step[slot].status == kPendingSolve) {
        handed_over = true;
        // k_step left its survivors to the multi-block solve; the gate it
        // closed keeps the queued steps behind it from running meanwhile
        __atomic_store_n(&b->h_step[slot].seq, 0ull, __ATOMIC_RELEASE);
        SearchArgs sa = b->slot_args[slot];
        hipLaunchKernelGGL(k_solve, dim3(kSolveBlocks), dim3(kBlock), 0, bstream(b), b->d, sa);
        HIP_TRY(hipGetLastError());
        b->n_handover += 1;
        rc = wait_seq(b, &b->h_step[slot].seq, ticket);
        if (rc) return rc;
    }
    b->inflight -= 1;
    b->oldest = ticket + 1;
    const StepRecord &s = b->h_step[slot];
    // events are read when the slot is reused or the totals are asked for,
    // so that the host never blocks on them here
    if (b->slot_timed[slot] == 1) b->slot_timed[slot] = s.skipped ? 0 : 2;
    if (skipped) *skipped = s.skipped ? 1 : ((handed_over && b->slot_args[slot].gated && !s.contact) ? 2 : 0);
    if (out) {
        out->distance = s.dist;
        out->move = s.move;
        out->moving_sphere = s.ms;
        out->other_sphere = s.os;
        out->moving_agg = agg;
        out->other_agg = s.oagg;
        out->contact = s.contact;
        out->status = s.status;
        out->n_query = static_cast<int64_t>(s.nquery);
        out->n_filtered = static_cast<int64_t>(s.nfilt);
        out->n_pairs = static_cast<int64_t>(s.npairs);
    }
    if (s.skipped) {
        take_out(b, slot);
    } else if (b->slot_apply[slot]) {
        // the mover's new CoM, as the finishing block computed it (aggregat.cpp:148-161)
        const size_t A = static_cast<size_t>(agg);
        const SearchArgs &sa = b->slot_args[slot];
        b->h_x[A] = periodic_position(b->h_x[A] + sa.dir[0] * s.move, b->p.box_length);
        b->h_y[A] = periodic_position(b->h_y[A] + sa.dir[1] * s.move, b->p.box_length);
        b->h_z[A] = periodic_position(b->h_z[A] + sa.dir[2] * s.move, b->p.box_length);
        b->pend[A] -= 1;
    }
    if (!s.skipped && b->slot_args[slot].gated && (s.contact || handed_over)) {
        // the steps queued behind this one will be skipped
        for (unsigned long long t = ticket + 1; t <= b->seq; ++t) take_out(b, static_cast<int>(t % mcac_box::kRing));
        if (s.contact) {
            b->gate_closed = true;
            b->last_gate = ticket;
        } else {
            b->released = ticket;  // solved in the second launch without contact: open again
        }
    }
    if (!s.skipped && s.contact && s.merged_inline) {
        b->inline_pending = true;
        b->inline_rec = s;
        b->inline_src = agg;
        b->inline_epoch = b->slot_epoch[slot];
    }
    return s.skipped ? 0 : s.status;
}

int search(mcac_box *b, int64_t agg, const double dir[3], double lpm, double maxradius, int apply, mcac_contact *out) {
    if (b->inflight != 0) return MCAC_E_ARGUMENT;  // pipelined steps pending
    unsigned long long ticket = 0;
    int rc = enqueue_search(b, agg, dir, lpm, maxradius, apply, 0, &ticket);
    if (rc) return rc;
    return wait_search(b, ticket, agg, out, nullptr);
}

}  // namespace

// Buffer sizes of a box (entries), from its parameters.
struct BoxSizes {
    int64_t Ns, Na, n_cells, cand_cap, part_cap, pool_cap;
};
BoxSizes box_sizes(const mcac_box_params &p) {
    BoxSizes z{};
    z.Ns = p.max_spheres;
    z.Na = p.max_aggregates;
    z.n_cells = p.n_div * p.n_div * p.n_div;
    z.cand_cap = p.candidate_capacity > 0 ? p.candidate_capacity : std::max<int64_t>(4096, std::min<int64_t>(z.Na, 1 << 20));
    z.part_cap = std::max<int64_t>(z.cand_cap + 1024, 1 << 20);
    z.pool_cap = p.member_pool > 0 ? p.member_pool : 4 * z.Ns + 4096;
    return z;
}

// Every device buffer of a box, in allocation order: f(pointer, entries) is
// called once per buffer (mcac_box_create allocates through it,
// mcac_box_device_bytes adds the sizes up).  Returns the first non-zero f.
template <class F> int box_buffers(DevBox &d, const BoxSizes &z, int volsurf, F &&f) {
    const int64_t Ns = z.Ns, Na = z.Na;
    int rc = 0;
#define ALLOC(ptr, n) \
    if ((rc = f(ptr, (n))) != 0) return rc
    ALLOC(d.sx, Ns); ALLOC(d.sy, Ns); ALLOC(d.sz, Ns); ALLOC(d.sr, Ns); ALLOC(d.svol, Ns); ALLOC(d.ssurf, Ns);
    ALLOC(d.srx, Ns); ALLOC(d.sry, Ns); ALLOC(d.srz, Ns);
    ALLOC(d.ax, Na); ALLOC(d.ay, Na); ALLOC(d.az, Na); ALLOC(d.arx, Na); ALLOC(d.ary, Na); ALLOC(d.arz, Na);
    ALLOC(d.armax, Na); ALLOC(d.avol, Na); ALLOC(d.asurf, Na); ALLOC(d.arg, Na);
    ALLOC(d.csvx, Na); ALLOC(d.csvy, Na); ALLOC(d.csvz, Na); ALLOC(d.csr, Na); ALLOC(d.csvr2, Na);
    ALLOC(d.astart, Na); ALLOC(d.acount, Na); ALLOC(d.acell, Na); ALLOC(d.acap, Na);
    ALLOC(d.members, z.pool_cap);
    ALLOC(d.alive, Na);
    ALLOC(d.cstart, z.n_cells + 1); ALLOC(d.citems, Na); ALLOC(d.ccursor, z.n_cells + 1);
    ALLOC(d.cpx, Na); ALLOC(d.cpy, Na); ALLOC(d.cpz, Na); ALLOC(d.cpr, Na); ALLOC(d.cslot, Na); ALLOC(d.cf, Na);
    ALLOC(d.dirty, Na); ALLOC(d.dirty_list, Na + 1);
    ALLOC(d.dpx, Na + 1); ALLOC(d.dpy, Na + 1); ALLOC(d.dpz, Na + 1); ALLOC(d.dpr, Na + 1);
    ALLOC(d.dpcell, Na + 1); ALLOC(d.dslot, Na);
    ALLOC(d.cand_db, z.cand_cap); ALLOC(d.cand_exact, z.cand_cap); ALLOC(d.cand_rank, z.cand_cap);
    ALLOC(d.cand_id, z.cand_cap); ALLOC(d.order, z.cand_cap); ALLOC(d.cand_pair, z.cand_cap);
    ALLOC(d.part_d, z.part_cap); ALLOC(d.part_p, z.part_cap);
    ALLOC(d.dc, Ns);
    ALLOC(d.row_cnt, Ns); ALLOC(d.row_j, Ns * kOverlapSlots); ALLOC(d.row_c, Ns * kOverlapSlots);
    if (volsurf == 1 || volsurf == 3) {
        ALLOC(d.row_d, Ns * kOverlapSlots); ALLOC(d.cvol, Ns); ALLOC(d.csurf, Ns);
    }
    if (volsurf == 2) {
        ALLOC(d.row_d, Ns * kOverlapSlots); ALLOC(d.sr3, Ns); ALLOC(d.walk, 3 * Ns * kOverlapSlots);
        ALLOC(d.walk_n, Ns);
    }
    ALLOC(d.ss, 1); ALLOC(d.ms, 1);
    ALLOC(d.eng_jobs, kEngineRing); ALLOC(d.eng_closed, kEngineRing);
#undef ALLOC
    return 0;
}

// Static staging of mcac_swept_sphere_batch_noalloc (13 doubles per input,
// 6.5 MiB).  In host code `scratch` names a host-side stand-in, not the
// device array: its device address comes from hipGetSymbolAddress, once per
// device with that device current, and only that address is used.
constexpr int64_t kNoAllocInputs = int64_t(1) << 16;
static __device__ __align__(256) double scratch[13 * kNoAllocInputs];

namespace {
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of `scratch` on `device` (current), looked up on first use.
int scratch_address(int device, double **out) {
    if (device < 0 || device >= kMaxDevices) return MCAC_E_ARGUMENT;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return MCAC_E_DEVICE;
        g_scratch[device] = static_cast<double *>(p);
    }
    *out = g_scratch[device];
    return 0;
}
}  // namespace

extern "C" {

int mcac_swept_sphere_batch(int64_t n, const double *p1, const double *r1, const double *p2, const double *r2,
                            const double *dir, const double *dist, double box, double *out, int device) {
    if (n < 0 || (n > 0 && (!p1 || !r1 || !p2 || !r2 || !dir || !dist || !out)) || box <= 0) return MCAC_E_ARGUMENT;
    if (n == 0) return 0;
    HIP_TRY(hipSetDevice(device));
    const size_t N = static_cast<size_t>(n);
    double *d = nullptr;
    const size_t total = N * (3 + 1 + 3 + 1 + 3 + 1 + 1);
    HIP_TRY(hipMalloc(&d, total * sizeof(double)));
    double *dp1 = d, *dr1 = dp1 + 3 * N, *dp2 = dr1 + N, *dr2 = dp2 + 3 * N, *ddir = dr2 + N, *ddist = ddir + 3 * N,
           *dout = ddist + N;
    int rc = 0;
    if (hipMemcpy(dp1, p1, 3 * N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dr1, r1, N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dp2, p2, 3 * N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dr2, r2, N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ddir, dir, 3 * N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ddist, dist, N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        rc = MCAC_E_DEVICE;
    if (!rc) {
        unsigned blocks = static_cast<unsigned>(std::min<int64_t>(4096, (n + kBlock - 1) / kBlock));
        hipLaunchKernelGGL(k_swept_batch, dim3(blocks), dim3(kBlock), 0, nullptr, static_cast<long long>(n), dp1, dr1,
                           dp2, dr2, ddir, ddist, box, dout);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(out, dout, N * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            rc = MCAC_E_DEVICE;
    }
    (void)hipFree(d);
    return rc;
}

// mcac_swept_sphere_batch without device allocation: the same staging
// (inputs, then outputs, 13 doubles per input) lives in `scratch`, a
// file-scope device array of kNoAllocInputs inputs that each GPU holds until
// the process ends.  Every element a call reads is written earlier in the
// same call, so nothing needs clearing; two calls that overlap on one GPU
// share the array and race.
int mcac_swept_sphere_batch_noalloc(int64_t n, const double *p1, const double *r1, const double *p2,
                                    const double *r2, const double *dir, const double *dist, double box, double *out,
                                    int device) {
    if (n < 0 || (n > 0 && (!p1 || !r1 || !p2 || !r2 || !dir || !dist || !out)) || box <= 0) return MCAC_E_ARGUMENT;
    if (n > kNoAllocInputs) {
        std::fprintf(stderr, "mcac_swept_sphere_batch_noalloc: %lld inputs exceed the static staging of %lld\n",
                     static_cast<long long>(n), static_cast<long long>(kNoAllocInputs));
        return MCAC_E_CAPACITY;
    }
    if (n == 0) return 0;
    HIP_TRY(hipSetDevice(device));
    double *d = nullptr;
    int rc = scratch_address(device, &d);
    if (rc) return rc;
    const size_t N = static_cast<size_t>(n);
    double *dp1 = d, *dr1 = dp1 + 3 * N, *dp2 = dr1 + N, *dr2 = dp2 + 3 * N, *ddir = dr2 + N, *ddist = ddir + 3 * N,
           *dout = ddist + N;
    HIP_TRY(hipMemcpy(dp1, p1, 3 * N * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dr1, r1, N * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dp2, p2, 3 * N * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dr2, r2, N * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ddir, dir, 3 * N * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ddist, dist, N * sizeof(double), hipMemcpyHostToDevice));
    const unsigned blocks = static_cast<unsigned>((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_swept_batch, dim3(blocks), dim3(kBlock), 0, nullptr, static_cast<long long>(n), dp1, dr1, dp2,
                       dr2, ddir, ddist, box, dout);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, dout, N * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

const char *mcac_version(void) { return "mcac_gpu 0.2 (gfx950, FP64 collision path)"; }

int mcac_device_memory(int device, int64_t *free_bytes, int64_t *total_bytes) {
    if (!free_bytes || !total_bytes) return MCAC_E_ARGUMENT;
    HIP_TRY(hipSetDevice(device));
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    *free_bytes = static_cast<int64_t>(f);
    *total_bytes = static_cast<int64_t>(t);
    return 0;
}

int64_t mcac_box_device_bytes(const mcac_box_params *params) {
    if (!params || params->max_spheres < 1 || params->max_aggregates < 1 || params->n_div < 1 ||
        params->n_div > 1290 || params->volsurf < 0 || params->volsurf > 3)
        return -1;
    DevBox d{};
    int64_t bytes = 0;
    box_buffers(d, box_sizes(*params), params->volsurf, [&](auto *&ptr, int64_t n) {
        bytes += (n < 1 ? 1 : n) * static_cast<int64_t>(sizeof(*ptr));
        return 0;
    });
    return bytes;
}

int mcac_box_create(const mcac_box_params *params, mcac_box **out) {
    if (!params || !out) return MCAC_E_ARGUMENT;
    if (params->box_length <= 0 || params->n_div < 1 || params->n_div > 1290 || params->max_spheres < 1 ||
        params->max_aggregates < 1 || params->max_spheres > INT32_MAX / 2 || params->max_aggregates > INT32_MAX / 2 ||
        params->volsurf < 0 || params->volsurf > 3)
        return MCAC_E_ARGUMENT;
    auto *b = new (std::nothrow) mcac_box();
    if (!b) return MCAC_E_CAPACITY;
    b->p = *params;
    if (const char *e = std::getenv("MCAC_FUSED_MERGE")) b->fused_merge = std::atoi(e) != 0;  // 0: the kernel sequence
    if (const char *e = std::getenv("MCAC_VERIFY_CELLS")) b->verify_cells = std::atoi(e) != 0;
    if (const char *e = std::getenv("MCAC_GRID_SOLVE")) b->grid_solve = std::atoi(e) != 0;
    if (const char *e = std::getenv("MCAC_GRID_BLOCKS")) b->grid_blocks = std::min(kFilterBlocks, std::max(1, std::atoi(e)));
    int rc = 0;
    auto fail = [&](int code) {
        mcac_box_destroy(b);
        return code;
    };
    if (hipSetDevice(params->device) != hipSuccess) return fail(MCAC_E_DEVICE);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(MCAC_E_DEVICE);
    const BoxSizes z = box_sizes(*params);
    b->n_cells = static_cast<int>(z.n_cells);
    b->cand_cap = z.cand_cap;
    b->part_cap = z.part_cap;
    b->pool_cap = z.pool_cap;
    rc = box_buffers(b->d, z, params->volsurf, [&](auto *&ptr, int64_t n) { return dalloc(b, &ptr, n); });
    if (rc) return fail(rc);
    if (b->n_cells + 1 <= (1 << 20) &&
        hipHostMalloc(reinterpret_cast<void **>(&b->h_cstart), (static_cast<size_t>(b->n_cells) + 1) * sizeof(int),
                      hipHostMallocDefault) != hipSuccess)
        return fail(MCAC_E_DEVICE);
    b->alive_cap = z.Na;
    if (hipHostMalloc(reinterpret_cast<void **>(&b->h_alive), static_cast<size_t>(z.Na) * sizeof(int),
                      hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&b->alive_read, hipEventDisableTiming) != hipSuccess)
        return fail(MCAC_E_DEVICE);
    if (hipHostMalloc(reinterpret_cast<void **>(&b->h_step), mcac_box::kRing * sizeof(StepRecord),
                      hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&b->h_merge), sizeof(MergeRecord), hipHostMallocCoherent) != hipSuccess)
        return fail(MCAC_E_DEVICE);
    std::memset(b->h_step, 0, mcac_box::kRing * sizeof(StepRecord));
    std::memset(b->h_merge, 0, sizeof(MergeRecord));
    if (hipHostGetDevicePointer(reinterpret_cast<void **>(&b->d_step), b->h_step, 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void **>(&b->d_merge), b->h_merge, 0) != hipSuccess)
        return fail(MCAC_E_DEVICE);
    if (hipHostMalloc(reinterpret_cast<void **>(&b->h_ring), kEngineRing * sizeof(EngineJob), hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&b->h_ctl), sizeof(EngineCtl), hipHostMallocCoherent) != hipSuccess)
        return fail(MCAC_E_DEVICE);
    std::memset(b->h_ring, 0, kEngineRing * sizeof(EngineJob));
    std::memset(b->h_ctl, 0, sizeof(EngineCtl));
    if (hipHostGetDevicePointer(reinterpret_cast<void **>(&b->dp_ring), b->h_ring, 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void **>(&b->dp_ctl), b->h_ctl, 0) != hipSuccess)
        return fail(MCAC_E_DEVICE);
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, params->device) != hipSuccess || khz <= 0)
            khz = 100000;  // the MI300-series constant clock, 100 MHz
        b->eng_clock_khz = khz;
        b->eng_idle_ticks = static_cast<unsigned long long>(khz) * 500ull;  // an idle grid ends after 0.5 s
    }
    for (auto &e : b->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(MCAC_E_DEVICE);
    *out = b;
    return 0;
}

int mcac_box_destroy(mcac_box *b) {
    if (!b) return 0;
    if (b->stream) (void)hipStreamSynchronize(bstream(b));  // a resident engine ends first
    for (void *p : b->allocs) (void)hipFree(p);
    if (b->h_step) (void)hipHostFree(b->h_step);
    if (b->h_merge) (void)hipHostFree(b->h_merge);
    if (b->h_alpha) (void)hipHostFree(b->h_alpha);
    if (b->h_cstart) (void)hipHostFree(b->h_cstart);
    if (b->h_alive) (void)hipHostFree(b->h_alive);
    if (b->alive_read) (void)hipEventDestroy(b->alive_read);
    if (b->h_ring) (void)hipHostFree(b->h_ring);
    if (b->h_grow) (void)hipHostFree(b->h_grow);
    if (b->d_grow) (void)hipFree(b->d_grow);
    if (b->h_ctl) (void)hipHostFree(b->h_ctl);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &row : b->sev)
        for (auto &e : row)
            if (e) (void)hipEventDestroy(e);
    if (b->stream && b->own_stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int mcac_box_upload(mcac_box *b, int64_t ns, const double *x, const double *y, const double *z, const double *r,
                    const double *vol, const double *surf, const double *rx, const double *ry, const double *rz,
                    int64_t na, const double *pos3, const double *rcm3, const double *rmax, const double *volume,
                    const double *surface, const int64_t *count, const int64_t *members) {
    if (!b || ns < 0 || na < 0 || ns > b->p.max_spheres || na > b->p.max_aggregates) return MCAC_E_ARGUMENT;
    const size_t S = static_cast<size_t>(ns), A = static_cast<size_t>(na);
    int rc = 0;
    auto up = [&](double *dst, const double *src, size_t n) -> int {
        if (n) HIP_TRY(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice));
        return 0;
    };
    std::vector<double> hx(A), hy(A), hz(A), hrx(A), hry(A), hrz(A), c0(A), c1(A), c2(A), csr(A), cvr2(A);
    std::vector<int> cell(A, -1), cnt(A, 0), st(A, 0);
    std::vector<int> alive;
    std::vector<int> pool;
    pool.reserve(S);
    const double L = b->p.box_length;
    const int n = static_cast<int>(b->p.n_div);
    int64_t off = 0;
    for (size_t a = 0; a < A; ++a) {
        hx[a] = pos3[3 * a];
        hy[a] = pos3[3 * a + 1];
        hz[a] = pos3[3 * a + 2];
        hrx[a] = rcm3[3 * a];
        hry[a] = rcm3[3 * a + 1];
        hrz[a] = rcm3[3 * a + 2];
        const int64_t c = count[a];
        if (c < 0 || off + c > ns) return MCAC_E_ARGUMENT;
        cnt[a] = static_cast<int>(c);
        st[a] = static_cast<int>(pool.size());
        // list-order sums a merge continues from (aggregat.cpp:437-441, 257-260, 474-478)
        double svx = 0., svy = 0., svz = 0., sr = 0., svr2 = 0.;
        for (int64_t i = 0; i < c; ++i) {
            const int64_t s = members[off + i];
            if (s < 0 || s >= ns) return MCAC_E_ARGUMENT;
            pool.push_back(static_cast<int>(s));
            svx += rx[s] * vol[s];
            svy += ry[s] * vol[s];
            svz += rz[s] * vol[s];
            sr += r[s];
            svr2 = svr2 + vol[s] * (r[s] * r[s]);
        }
        off += c;
        c0[a] = svx;
        c1[a] = svy;
        c2[a] = svz;
        csr[a] = sr;
        cvr2[a] = svr2;
        if (c > 0) {
            alive.push_back(static_cast<int>(a));
            const double step = static_cast<double>(n) / L;
            const double p[3] = {hx[a], hy[a], hz[a]};
            int ci[3];
            for (int l = 0; l < 3; ++l) {
                int v = static_cast<int>(std::floor(p[l] * step));
                ci[l] = v < 0 ? 0 : (v >= n ? n - 1 : v);
            }
            cell[a] = (ci[0] * n + ci[1]) * n + ci[2];
        }
    }
    if (static_cast<int64_t>(pool.size()) > b->pool_cap) return MCAC_E_CAPACITY;
    if (b->p.volsurf == 2) {  // pow(r, 3) of the alphas sums (aggregat.cpp:402-403), host libm as the reference
        std::vector<double> r3(S);
        for (size_t s = 0; s < S; ++s) r3[s] = std::pow(r[s], 3.0);
        if ((rc = up(b->d.sr3, r3.data(), S))) return rc;
    }
    if ((rc = up(b->d.sx, x, S)) || (rc = up(b->d.sy, y, S)) || (rc = up(b->d.sz, z, S)) || (rc = up(b->d.sr, r, S)) ||
        (rc = up(b->d.svol, vol, S)) || (rc = up(b->d.ssurf, surf, S)) || (rc = up(b->d.srx, rx, S)) ||
        (rc = up(b->d.sry, ry, S)) || (rc = up(b->d.srz, rz, S)))
        return rc;
    if ((rc = up(b->d.ax, hx.data(), A)) || (rc = up(b->d.ay, hy.data(), A)) || (rc = up(b->d.az, hz.data(), A)) ||
        (rc = up(b->d.arx, hrx.data(), A)) || (rc = up(b->d.ary, hry.data(), A)) || (rc = up(b->d.arz, hrz.data(), A)) ||
        (rc = up(b->d.armax, rmax, A)) || (rc = up(b->d.avol, volume, A)) || (rc = up(b->d.asurf, surface, A)) ||
        (rc = up(b->d.csvx, c0.data(), A)) || (rc = up(b->d.csvy, c1.data(), A)) || (rc = up(b->d.csvz, c2.data(), A)) ||
        (rc = up(b->d.csr, csr.data(), A)) || (rc = up(b->d.csvr2, cvr2.data(), A)))
        return rc;
    if (A) {
        HIP_TRY(hipMemcpy(b->d.astart, st.data(), A * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b->d.acount, cnt.data(), A * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b->d.acell, cell.data(), A * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b->d.acap, cnt.data(), A * sizeof(int), hipMemcpyHostToDevice));
    }
    if (!pool.empty()) HIP_TRY(hipMemcpy(b->d.members, pool.data(), pool.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!alive.empty()) HIP_TRY(hipMemcpy(b->d.alive, alive.data(), alive.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(b->d.ss, 0, sizeof(SearchState)));
    HIP_TRY(hipDeviceSynchronize());
    b->n_spheres = ns;
    b->n_aggs = na;
    b->count = cnt;
    b->max_count = 0;
    for (int c : b->count) b->max_count = std::max(b->max_count, c);
    b->start = st;
    b->cap = cnt;
    b->alive_ids = alive;
    b->dead_in_list = 0;
    b->pool_used = static_cast<int64_t>(pool.size());
    b->inflight = 0;
    b->oldest = b->seq + 1;
    b->h_x = hx;
    b->h_y = hy;
    b->h_z = hz;
    b->h_rx = hrx;
    b->h_ry = hry;
    b->h_rz = hrz;
    b->h_rmax.assign(rmax, rmax + A);
    b->pend.assign(A, 0);
    b->h_dirty.assign(A, 0);
    b->h_dslot.assign(A, 0);
    b->dirty_count = 0;
    b->gate_closed = false;
    b->released = b->seq;
    b->last_gate = 0;
    b->inline_pending = false;
    rc = rebuild_cells(b);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    return 0;
}

int mcac_next_contact(mcac_box *b, int64_t agg, const double dir[3], double lpm, double maxradius,
                      mcac_contact *out) {
    if (!b || !dir) return MCAC_E_ARGUMENT;
    return search(b, agg, dir, lpm, maxradius, 0, out);
}

int mcac_move(mcac_box *b, int64_t agg, const double dir[3], double lpm, double maxradius, mcac_contact *out) {
    if (!b || !dir) return MCAC_E_ARGUMENT;
    return search(b, agg, dir, lpm, maxradius, 1, out);
}

int mcac_translate(mcac_box *b, int64_t agg, const double vec[3]) {
    if (!b || !vec || agg < 0 || agg >= b->n_aggs || b->count[static_cast<size_t>(agg)] <= 0 || b->inflight)
        return MCAC_E_ARGUMENT;
    if (b->dirty_count >= rebuild_threshold(b)) {
        int rc = rebuild_cells(b);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_set_move, dim3(1), dim3(1), 0, bstream(b), b->d, static_cast<int>(agg), vec[0], vec[1], vec[2],
                       static_cast<int>(b->p.n_div), b->p.box_length);
    {
        const size_t A = static_cast<size_t>(agg);
        b->h_x[A] = periodic_position(b->h_x[A] + vec[0], b->p.box_length);
        b->h_y[A] = periodic_position(b->h_y[A] + vec[1], b->p.box_length);
        b->h_z[A] = periodic_position(b->h_z[A] + vec[2], b->p.box_length);
        if (!b->h_dirty[A]) {
            b->h_dirty[A] = 1;
            b->h_dslot[A] = static_cast<int>(b->dirty_count);
            b->dirty_count += 1;
        }
    }
    const int n = b->count[static_cast<size_t>(agg)];
    const int tb = std::min(1024, std::max(1, (n + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_translate, dim3(tb), dim3(kBlock), 0, bstream(b), b->d, static_cast<int>(agg), 0, 0ull, 0ull);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    return 0;
}

// Host bookkeeping after a merge ran on the device (kernel or inline).
static int merge_done(mcac_box *b, size_t K, size_t R, int new_start, int new_cap, bool dirty_now,
                      mcac_merge_result *out) {
    if (!out->merged) {
        out->keeper = static_cast<int64_t>(K);
        out->removed = static_cast<int64_t>(R);
        return 0;
    }
    if (new_start != b->start[K]) {
        b->pool_used += new_cap;
        b->start[K] = new_start;
        b->cap[K] = new_cap;
    }
    const int n = static_cast<int>(out->n);
    b->count[K] = n;
    b->max_count = std::max(b->max_count, n);
    b->count[R] = 0;
    b->dead_in_list += 1;
    // store_merged marks both aggregates as moved since the cell-list build
    // (unless a build ran after the merge on the device)
    if (dirty_now)
        for (size_t id : {K, R})
            if (!b->h_dirty[id]) {
                b->h_dirty[id] = 1;
                b->h_dslot[id] = static_cast<int>(b->dirty_count);
                b->dirty_count += 1;
            }
    b->h_x[K] = out->com[0];
    b->h_y[K] = out->com[1];
    b->h_z[K] = out->com[2];
    b->h_rx[K] = out->rcm[0];
    b->h_ry[K] = out->rcm[1];
    b->h_rz[K] = out->rcm[2];
    b->h_rmax[K] = out->rmax;
    return out->status;
}

// Alphas correction of a merged aggregate (aggregat.cpp:289-319, 392-422) from
// its neighbour-map walk rows (c_ij, vp1 + vp2, sp1 + sp2) in walk order and
// its list-order volume V and surface S.  Host FP64 with the reference's
// libm calls: pow(x, 2) is x * x in the reference's build, pow(x, 3) and
// pow(x, 1.5) are glibc's.
struct AlphaFactors {
    double volume, surface, overlapping, coordination;
    bool any;
};
AlphaFactors alpha_factors(int n, const int *walk_n, const double *walk, double V, double S) {
    double overlapping = 0.0, c_v30 = 0.0, c_v20 = 0.0, c_s10 = 0.0, vp_sum = 0.0, sp_sum = 0.0;
    size_t intersections = 0;
    for (int i = 0; i < n; ++i) {
        for (int y = 0; y < walk_n[i]; ++y) {
            const double *w = walk + 3 * (static_cast<size_t>(i) * kOverlapSlots + static_cast<size_t>(y));
            const double c_ij = w[0], vp = w[1], sp = w[2];
            vp_sum += vp;
            sp_sum += sp;
            overlapping += c_ij;
            c_s10 += c_ij * sp;
            c_v20 += (c_ij * c_ij) * vp;
            c_v30 += std::pow(c_ij, 3.0) * vp;
        }
        intersections += static_cast<size_t>(walk_n[i]);
    }
    AlphaFactors f{V, S, overlapping, 0.0, intersections > 0};
    if (intersections == 0) return f;
    c_s10 /= sp_sum;
    c_v20 /= vp_sum;
    c_v30 /= vp_sum;
    const double min_coordination_number = 2 * (1.0 - 1.0 / static_cast<double>(n));
    const double alpha_vs_extreme = 1.0 / static_cast<double>(n);  // aggregat.cpp:539
    f.overlapping = overlapping / static_cast<double>(intersections);
    f.coordination = static_cast<double>(intersections) / static_cast<double>(n);
    const double diff = std::abs(f.coordination - min_coordination_number);
    // volume_alpha_correction :289-304
    double corr = 0.25 * (3.0 * c_v20 - c_v30) * f.coordination - c_v30 * diff * 0.62741833 -
                  std::pow(diff, 1.5) * 0.00332425;
    if (corr < 0.0) corr = 1.0;
    corr = smin(corr, 1.0);
    const double alpha_v = smax(1.0 - corr, alpha_vs_extreme);
    // surface_alpha_correction :305-319
    double corr_s = 0.5 * c_s10 * f.coordination - (c_s10 * c_s10) * diff * 0.70132500 - (diff * diff) * 0.00450000;
    if (corr_s < 0.0) corr_s = 1.0;
    corr_s = smin(corr_s, 1.0);
    const double alpha_s = smax(1.0 - corr_s, alpha_vs_extreme);
    f.volume = V * alpha_v;
    f.surface = S * alpha_s;
    return f;
}

int mcac_merge(mcac_box *b, const mcac_contact *c, mcac_merge_result *out) {
    if (!b || !c || !out) return MCAC_E_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    if (b->inline_pending) {
        // the contact step already merged in its own launch
        const StepRecord &r = b->inline_rec;
        if (c->moving_agg != b->inline_src || c->other_agg != r.oagg || c->moving_sphere != r.ms ||
            c->other_sphere != r.os)
            return MCAC_E_ARGUMENT;
        b->inline_pending = false;
        b->n_inline += 1;
        *out = r.merge;
        const size_t K = static_cast<size_t>(std::min(c->moving_agg, c->other_agg));
        const size_t R = static_cast<size_t>(std::max(c->moving_agg, c->other_agg));
        return merge_done(b, K, R, r.m_new_start, r.m_new_cap, b->inline_epoch == b->dirty_epoch, out);
    }
    if (c->moving_sphere < 0 || c->other_sphere < 0 || c->moving_sphere >= b->n_spheres ||
        c->other_sphere >= b->n_spheres || c->moving_agg < 0 || c->other_agg < 0 || c->moving_agg >= b->n_aggs ||
        c->other_agg >= b->n_aggs || c->moving_agg == c->other_agg)
        return MCAC_E_ARGUMENT;
    const int keep = static_cast<int>(std::min(c->moving_agg, c->other_agg));
    const int rem = static_cast<int>(std::max(c->moving_agg, c->other_agg));
    const size_t K = static_cast<size_t>(keep), R = static_cast<size_t>(rem);
    if (b->count[K] <= 0 || b->count[R] <= 0) return MCAC_E_ARGUMENT;
    if (b->dirty_count + 2 >= rebuild_threshold(b)) {
        int rc = rebuild_cells(b);
        if (rc) return rc;
    }
    const int nk = b->count[K], nr = b->count[R];
    // member segment: append in place when the keeper's segment has room, else
    // a fresh segment of twice the merged size
    int new_start = b->start[K];
    int new_cap = b->cap[K];
    if (nk + nr > b->cap[K]) {
        new_cap = std::max(4, 2 * (nk + nr));
        if (b->pool_used + new_cap > b->pool_cap) {
            int rc = compact_pool(b, new_cap);
            if (rc) return rc;
        }
        new_start = static_cast<int>(b->pool_used);
    }
    MergeArgs a{};
    a.ms = c->moving_sphere;
    a.os = c->other_sphere;
    a.magg = c->moving_agg;
    a.oagg = c->other_agg;
    a.keep = keep;
    a.rem = rem;
    a.n_k = nk;
    a.n_r = nr;
    a.new_start = new_start;
    a.start_k = b->start[K];
    a.start_r = b->start[R];
    a.cap_k = new_cap;
    a.n_div = static_cast<int>(b->p.n_div);
    a.L = b->p.box_length;
    a.out = b->d_merge;
    a.seq = ++b->merge_seq;
    a.volsurf = b->p.volsurf;
    const int n = nk + nr;
    if (b->timing) HIP_TRY(hipEventRecord(b->ev[0], bstream(b)));
    const bool fused = b->fused_merge && a.volsurf != 0 && n <= kFusedMerge;
    if (fused && a.volsurf != 2) {
        hipLaunchKernelGGL(k_merge_fused, dim3(1), dim3(kBlock), 0, bstream(b), b->d, a);
    } else if (a.volsurf == 1 || a.volsurf == 3) {
        // spherical caps or arvo: the two-way contact rows first, then the
        // per-sphere parts, the full ordered sums, dc, rmax and Rg (publishes)
        const int nb = std::min(4096, (n + kBlock - 1) / kBlock);
        const unsigned tiles = static_cast<unsigned>((n + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_merge_prep, dim3(1), dim3(64), 0, bstream(b), b->d, a);
        hipLaunchKernelGGL(k_merge_lists, dim3(nb), dim3(kBlock), 0, bstream(b), b->d);
        hipLaunchKernelGGL(k_overlap_pairs, dim3(tiles, tiles), dim3(kBlock), 0, bstream(b), b->d, 1);
        hipLaunchKernelGGL(k_overlap_sum, dim3(1), dim3(kBlock), 0, bstream(b), b->d, a);
        if (a.volsurf == 1) hipLaunchKernelGGL(k_caps_volumes, dim3(nb), dim3(kBlock), 0, bstream(b), b->d);
        else hipLaunchKernelGGL(k_arvo_volumes, dim3(nb), dim3(kBlock), 0, bstream(b), b->d);
        hipLaunchKernelGGL(k_caps_sums, dim3(1), dim3(kBlock), 0, bstream(b), b->d, a);
        hipLaunchKernelGGL(k_merge_dc, dim3(nb), dim3(kBlock), 0, bstream(b), b->d);
        hipLaunchKernelGGL(k_merge_rg, dim3(1), dim3(kBlock), 0, bstream(b), b->d, a);
    } else if (a.volsurf == 2) {
        // alphas: contact rows, the neighbour-map walks and the list-order sums
        // on the device; the correction factors on the host (glibc pow, as
        // the reference); then the centre of mass, dc, rmax and Rg (publishes)
        const int nb = std::min(4096, (n + kBlock - 1) / kBlock);
        const unsigned tiles = static_cast<unsigned>((n + kBlock - 1) / kBlock);
        const size_t off_n = (sizeof(MergeState) + 63) & ~static_cast<size_t>(63);
        const size_t off_w = off_n + ((static_cast<size_t>(n) * sizeof(int) + 63) & ~static_cast<size_t>(63));
        const size_t walk_bytes = 3 * static_cast<size_t>(n) * kOverlapSlots * sizeof(double);
        if (off_w + walk_bytes > b->h_alpha_cap) {
            const size_t cap = std::max(off_w + walk_bytes, 2 * b->h_alpha_cap);
            if (b->h_alpha) HIP_TRY(hipHostFree(b->h_alpha));
            b->h_alpha = nullptr;
            b->h_alpha_cap = 0;
            HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&b->h_alpha), cap, hipHostMallocDefault));
            b->h_alpha_cap = cap;
        }
        if (fused) {
            hipLaunchKernelGGL(k_merge_fused, dim3(1), dim3(kBlock), 0, bstream(b), b->d, a);
        } else {
            hipLaunchKernelGGL(k_merge_prep, dim3(1), dim3(64), 0, bstream(b), b->d, a);
            hipLaunchKernelGGL(k_merge_lists, dim3(nb), dim3(kBlock), 0, bstream(b), b->d);
            hipLaunchKernelGGL(k_overlap_pairs, dim3(tiles, tiles), dim3(kBlock), 0, bstream(b), b->d, 1);
            hipLaunchKernelGGL(k_overlap_sum, dim3(1), dim3(kBlock), 0, bstream(b), b->d, a);
            hipLaunchKernelGGL(k_alpha_walk, dim3(nb), dim3(kBlock), 0, bstream(b), b->d);
            hipLaunchKernelGGL(k_caps_sums, dim3(1), dim3(kBlock), 0, bstream(b), b->d, a);
        }
        HIP_TRY(hipGetLastError());
        // the state and the walk rows in one wait (the rows are unused when the re-check failed)
        HIP_TRY(hipMemcpyAsync(b->h_alpha, b->d.ms, sizeof(MergeState), hipMemcpyDeviceToHost, bstream(b)));
        HIP_TRY(hipMemcpyAsync(b->h_alpha + off_n, b->d.walk_n, static_cast<size_t>(n) * sizeof(int),
                               hipMemcpyDeviceToHost, bstream(b)));
        HIP_TRY(hipMemcpyAsync(b->h_alpha + off_w, b->d.walk, walk_bytes, hipMemcpyDeviceToHost, bstream(b)));
        HIP_TRY(hipStreamSynchronize(bstream(b)));
        MergeState ms{};
        std::memcpy(&ms, b->h_alpha, sizeof(MergeState));
        if (ms.ok) {
            const AlphaFactors f = alpha_factors(n, reinterpret_cast<const int *>(b->h_alpha + off_n),
                                                 reinterpret_cast<const double *>(b->h_alpha + off_w), ms.sums[0],
                                                 ms.sums[1]);
            if (fused) {
                hipLaunchKernelGGL(k_alpha_finish, dim3(1), dim3(kBlock), 0, bstream(b), b->d, a, f.volume,
                                   f.surface, f.overlapping, f.coordination, f.any ? 1 : 0);
            } else {
                hipLaunchKernelGGL(k_alpha_store, dim3(1), dim3(64), 0, bstream(b), b->d, a, f.volume, f.surface,
                                   f.overlapping, f.coordination, f.any ? 1 : 0);
                hipLaunchKernelGGL(k_merge_dc, dim3(nb), dim3(kBlock), 0, bstream(b), b->d);
                hipLaunchKernelGGL(k_merge_rg, dim3(1), dim3(kBlock), 0, bstream(b), b->d, a);
            }
        }
    } else if (n <= kSmallMerge) {
        hipLaunchKernelGGL(k_merge_small, dim3(1), dim3(kBlock), 0, bstream(b), b->d, a);
    } else {
        const int nb = std::min(4096, (n + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_merge_prep, dim3(1), dim3(64), 0, bstream(b), b->d, a);
        hipLaunchKernelGGL(k_merge_lists, dim3(nb), dim3(kBlock), 0, bstream(b), b->d);
        hipLaunchKernelGGL(k_merge_seq, dim3(1), dim3(kBlock), 0, bstream(b), b->d, a);
        hipLaunchKernelGGL(k_merge_dc, dim3(nb), dim3(kBlock), 0, bstream(b), b->d);
        hipLaunchKernelGGL(k_merge_rg, dim3(1), dim3(kBlock), 0, bstream(b), b->d, a);
        const unsigned tiles = static_cast<unsigned>((n + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_overlap_pairs, dim3(tiles, tiles), dim3(kBlock), 0, bstream(b), b->d, 0);
        hipLaunchKernelGGL(k_overlap_sum, dim3(1), dim3(kBlock), 0, bstream(b), b->d, a);
    }
    if (b->timing) HIP_TRY(hipEventRecord(b->ev[1], bstream(b)));
    HIP_TRY(hipGetLastError());
    int rc = wait_seq(b, &b->h_merge->seq, a.seq);
    if (rc) return rc;
    if (b->timing) {
        HIP_TRY(hipEventSynchronize(b->ev[1]));
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
        b->t_merge += t;
        b->n_merge += 1;
    }
    *out = b->h_merge->res;
    b->n_host_merge += 1;
    return merge_done(b, K, R, new_start, new_cap, true, out);
}

int mcac_move_enqueue(mcac_box *b, int64_t agg, const double dir[3], double lpm, double maxradius, uint64_t *ticket) {
    if (!b || !dir || !ticket) return MCAC_E_ARGUMENT;
    unsigned long long t = 0;
    int rc = enqueue_search(b, agg, dir, lpm, maxradius, 1, 1, &t);
    *ticket = t;
    return rc;
}

int mcac_group_create(int device, int max_boxes, mcac_group **out) {
    if (!out || max_boxes < 1 || max_boxes > 65535) return MCAC_E_ARGUMENT;
    auto *g = new (std::nothrow) mcac_group();
    if (!g) return MCAC_E_CAPACITY;
    g->device = device;
    g->cap = max_boxes;
    g->job_cap = max_boxes * mcac_group::kMovesPerBox;
    const size_t bytes = mcac_group::kBuffers * static_cast<size_t>(g->job_cap) * sizeof(StepJob);
    const size_t seg_bytes = mcac_group::kBuffers * (static_cast<size_t>(g->job_cap) + 1) * sizeof(int);
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&g->h_jobs), bytes, hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void **>(&g->d_jobs), g->h_jobs, 0) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&g->h_seg), seg_bytes, hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void **>(&g->d_seg), g->h_seg, 0) != hipSuccess ||
        hipEventCreateWithFlags(&g->read[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->read[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->read[2], hipEventDisableTiming) != hipSuccess) {
        mcac_group_destroy(g);
        return MCAC_E_DEVICE;
    }
    std::memset(g->h_jobs, 0, bytes);
    std::memset(g->h_seg, 0, seg_bytes);
    *out = g;
    return 0;
}

int mcac_group_destroy(mcac_group *g) {
    if (!g) return 0;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->h_jobs) (void)hipHostFree(g->h_jobs);
    if (g->h_seg) (void)hipHostFree(g->h_seg);
    for (auto &e : g->read)
        if (e) (void)hipEventDestroy(e);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return 0;
}

int mcac_group_add(mcac_group *g, mcac_box *b) {
    if (!g || !b || b->group || b->inflight || b->p.device != g->device) return MCAC_E_ARGUMENT;
    b->eng_blocks = 0;  // the group's launches take the box's steps
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    if (b->own_stream) HIP_TRY(hipStreamDestroy(b->stream));
    b->stream = g->stream;
    b->own_stream = false;
    b->group = g;
    return 0;
}

namespace {

// Job list being written: jobs[0..n), block segments seg[0..blocks]; opened
// when its first job is written (open = false: none yet).
struct JobList {
    StepJob *jobs = nullptr;
    int *seg = nullptr;
    int n = 0, blocks = 0;
    bool open = false;
    const mcac_box *last = nullptr;
};

int group_open(mcac_group *g, JobList &l) {
    const int q = g->parity;
    if (g->pending[q]) {  // the launch that read this list has copied its jobs
        const auto t0 = std::chrono::steady_clock::now();
        HIP_TRY(hipEventSynchronize(g->read[q]));
        g->t_list += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        g->pending[q] = false;
    }
    l.jobs = g->h_jobs + static_cast<size_t>(q) * static_cast<size_t>(g->job_cap);
    l.seg = g->h_seg + static_cast<size_t>(q) * (static_cast<size_t>(g->job_cap) + 1);
    l.n = l.blocks = 0;
    l.open = true;
    l.last = nullptr;
    return 0;
}

int group_flush(mcac_group *g, JobList &l) {
    if (!l.open || l.n == 0) return 0;
    const int q = g->parity;
    l.seg[l.blocks] = l.n;
    const auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(k_step_batch, dim3(static_cast<unsigned>(l.blocks)), dim3(kBlock), 0, g->stream,
                       g->d_jobs + static_cast<size_t>(q) * static_cast<size_t>(g->job_cap),
                       g->d_seg + static_cast<size_t>(q) * (static_cast<size_t>(g->job_cap) + 1));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(g->read[q], g->stream));
    g->t_launch += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    g->pending[q] = true;
    g->parity = (q + 1) % mcac_group::kBuffers;
    g->launches += 1;
    g->batched_steps += l.n;
    l.open = false;
    l.last = nullptr;
    return 0;
}

}  // namespace

int mcac_move_enqueue_group(mcac_group *g, mcac_box *const *boxes, const int64_t *aggs, const double *dirs,
                            const double *lpm, const double *maxradius, int n, uint64_t *tickets) {
    if (!g || n < 0 || n > g->job_cap || (n > 0 && (!boxes || !aggs || !dirs || !lpm || !maxradius || !tickets)))
        return MCAC_E_ARGUMENT;
    // the moves of one box are contiguous and in the order they are to run
    const auto t_in = std::chrono::steady_clock::now();
    const int64_t call = ++g->calls;
    for (int i = 0; i < n; ++i) {
        mcac_box *b = boxes[i];
        if (!b || b->group != g) return MCAC_E_ARGUMENT;
        if (i > 0 && boxes[i - 1] == b) continue;
        if (b->group_call == call) return MCAC_E_ARGUMENT;
        b->group_call = call;
    }
    JobList l;
    int rc = 0;
    for (int i = 0; i < n; ++i) {
        mcac_box *b = boxes[i];
        const int64_t agg = aggs[i];
        const bool valid = agg >= 0 && agg < b->n_aggs;
        const double reach = valid ? b->h_rmax[static_cast<size_t>(agg)] + maxradius[i] + lpm[i] : 0.0;
        const bool batch = valid && one_block_step(b, b->count[static_cast<size_t>(agg)], next_step_timed(b)) &&
                           window_items(b, reach) <= kBlockItems;
        // a step launched on its own, or the cell-list rebuild that the
        // enqueue launches, runs after every job already listed; a block
        // runs at most kMaxBlockJobs jobs (the batch runner gives a box at
        // most that many moves per call, so this never splits them there)
        if ((i == 0 || boxes[i - 1] != b) && b->inflight == 0 && !b->inline_pending &&
            (b->cells_stale || b->dirty_count + mcac_group::kMovesPerBox + 3 >= rebuild_threshold(b))) {
            // a rebuild that this call's moves would make due, run before the
            // box's first move: nothing of the box is in flight, so it needs no
            // drain, and its kernels run ahead of the box's jobs on the stream
            const auto t_r = std::chrono::steady_clock::now();
            rc = rebuild_cells(b);
            g->t_rebuild += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_r).count();
            if (rc) break;
        }
        const bool full = batch && l.last == b && l.n - l.seg[l.blocks - 1] >= kMaxBlockJobs;
        const bool rebuild = rebuild_due(b) || b->cells_stale;
        if ((!batch || full || (rebuild && l.last == b)) && (rc = group_flush(g, l)) != 0) break;
        if (batch && !l.open && (rc = group_open(g, l)) != 0) break;
        unsigned long long t = 0;
        bool batched = false;
        const auto t_s = std::chrono::steady_clock::now();
        rc = enqueue_search(b, agg, dirs + 3 * i, lpm[i], maxradius[i], 1, 1, &t, batch ? l.jobs + l.n : nullptr,
                            batch ? &batched : nullptr);
        if (rebuild) g->t_rebuild += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_s).count();
        if (rc) break;
        if (batched != batch) {
            rc = MCAC_UNKNOWN_ERROR;  // the two decisions must agree
            break;
        }
        tickets[i] = t;
        if (batched) {
            if (l.last != b) l.seg[l.blocks++] = l.n;  // a new block: the box's first job of this launch
            l.last = b;
            l.n += 1;
        } else {
            g->single_steps += 1;
            l.last = nullptr;
        }
    }
    // the steps enqueued before a failing one still run: their results are read as usual
    const int frc = group_flush(g, l);
    g->t_enqueue += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_in).count();
    return rc ? rc : frc;
}

int mcac_box_spare_moves(mcac_box *b) {
    if (!b) return 0;
    int spare = mcac_box::kRing - b->inflight;
    if (b->inflight > 0) {
        // with steps in flight a due rebuild would drain the stream first:
        // stop short of it (a box with nothing in flight rebuilds early instead)
        if (b->cells_stale) return 0;
        const int64_t room = rebuild_threshold(b) - 4 - b->dirty_count;
        spare = static_cast<int>(std::min<int64_t>(spare, std::max<int64_t>(0, room)));
    }
    return std::max(0, spare);
}

int mcac_group_times(mcac_group *g, double out[4]) {
    if (!g || !out) return MCAC_E_ARGUMENT;
    out[0] = g->t_enqueue;
    out[1] = g->t_rebuild;
    out[2] = g->t_launch;
    out[3] = g->t_list;
    return 0;
}

int mcac_box_block_steps(mcac_box *b, double reach) {
    if (!b) return MCAC_E_ARGUMENT;
    return window_items(b, reach) <= kBlockItems ? 1 : 0;
}

int mcac_group_stats(mcac_group *g, int64_t out[3]) {
    if (!g || !out) return MCAC_E_ARGUMENT;
    out[0] = g->launches;
    out[1] = g->batched_steps;
    out[2] = g->single_steps;
    return 0;
}

int mcac_move_result(mcac_box *b, uint64_t ticket, mcac_contact *out, int *skipped) {
    if (!b) return MCAC_E_ARGUMENT;
    return wait_search(b, ticket, b->slot_agg[ticket % mcac_box::kRing], out, skipped);
}

int mcac_release(mcac_box *b) {
    if (!b) return MCAC_E_ARGUMENT;
    // the steps enqueued from now on pass the gate value of the contact
    b->gate_closed = false;
    b->released = std::max(b->released, b->last_gate);
    return 0;
}

int mcac_box_set_auto_merge(mcac_box *b, int enable) {
    if (!b || b->inflight) return MCAC_E_ARGUMENT;
    b->automerge = (enable && b->p.volsurf == 0) ? 1 : 0;  // caps merges run their own kernel sequence
    return 0;
}

int mcac_arvo_part(int k, const double *x, const double *y, const double *z, const double *r, double out[2]) {
    if (!x || !y || !z || !r || !out || k < 0 || k > kArvoBalls) return MCAC_E_ARGUMENT;
    double bx[kArvoBalls + 1], by[kArvoBalls + 1], bz[kArvoBalls + 1], br[kArvoBalls + 1];
    for (int j = 0; j <= k; ++j) {
        bx[j] = x[j];
        by[j] = y[j];
        bz[j] = z[j];
        br[j] = r[j];
    }
    arvo_ball(bx, by, bz, br, k, &out[0], &out[1]);
    return 0;
}

int mcac_neighbour_order(int prev, const int64_t *keys, int m, int64_t *out) {
    if (!keys || !out || prev < 0 || m < 0 || m > kMapKeys) return MCAC_E_ARGUMENT;
    for (int c = 1; c < m; ++c)
        if (keys[c] <= keys[c - 1]) return MCAC_E_ARGUMENT;
    return neighbour_order(prev, keys, m, out) == 0 ? 0 : MCAC_E_ARGUMENT;
}

int mcac_box_download_spheres(mcac_box *b, double *x, double *y, double *z, double *rx, double *ry, double *rz) {
    if (!b) return MCAC_E_ARGUMENT;
    const size_t n = static_cast<size_t>(b->n_spheres) * sizeof(double);
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    if (x) HIP_TRY(hipMemcpy(x, b->d.sx, n, hipMemcpyDeviceToHost));
    if (y) HIP_TRY(hipMemcpy(y, b->d.sy, n, hipMemcpyDeviceToHost));
    if (z) HIP_TRY(hipMemcpy(z, b->d.sz, n, hipMemcpyDeviceToHost));
    if (rx) HIP_TRY(hipMemcpy(rx, b->d.srx, n, hipMemcpyDeviceToHost));
    if (ry) HIP_TRY(hipMemcpy(ry, b->d.sry, n, hipMemcpyDeviceToHost));
    if (rz) HIP_TRY(hipMemcpy(rz, b->d.srz, n, hipMemcpyDeviceToHost));
    return 0;
}

int mcac_box_download_radii(mcac_box *b, double *r, double *vol, double *surf) {
    if (!b) return MCAC_E_ARGUMENT;
    const size_t n = static_cast<size_t>(b->n_spheres) * sizeof(double);
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    if (r) HIP_TRY(hipMemcpy(r, b->d.sr, n, hipMemcpyDeviceToHost));
    if (vol) HIP_TRY(hipMemcpy(vol, b->d.svol, n, hipMemcpyDeviceToHost));
    if (surf) HIP_TRY(hipMemcpy(surf, b->d.ssurf, n, hipMemcpyDeviceToHost));
    return 0;
}

int mcac_box_download_aggs(mcac_box *b, double *pos3, double *rcm3, double *rmax, double *volume, double *surface,
                           double *rg, int64_t *count) {
    if (!b) return MCAC_E_ARGUMENT;
    const size_t A = static_cast<size_t>(b->n_aggs);
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    std::vector<double> t0(A), t1(A), t2(A);
    if (pos3) {
        HIP_TRY(hipMemcpy(t0.data(), b->d.ax, A * sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(t1.data(), b->d.ay, A * sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(t2.data(), b->d.az, A * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t a = 0; a < A; ++a) {
            pos3[3 * a] = t0[a];
            pos3[3 * a + 1] = t1[a];
            pos3[3 * a + 2] = t2[a];
        }
    }
    if (rcm3) {
        HIP_TRY(hipMemcpy(t0.data(), b->d.arx, A * sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(t1.data(), b->d.ary, A * sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(t2.data(), b->d.arz, A * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t a = 0; a < A; ++a) {
            rcm3[3 * a] = t0[a];
            rcm3[3 * a + 1] = t1[a];
            rcm3[3 * a + 2] = t2[a];
        }
    }
    if (rmax) HIP_TRY(hipMemcpy(rmax, b->d.armax, A * sizeof(double), hipMemcpyDeviceToHost));
    if (volume) HIP_TRY(hipMemcpy(volume, b->d.avol, A * sizeof(double), hipMemcpyDeviceToHost));
    if (surface) HIP_TRY(hipMemcpy(surface, b->d.asurf, A * sizeof(double), hipMemcpyDeviceToHost));
    if (rg) HIP_TRY(hipMemcpy(rg, b->d.arg, A * sizeof(double), hipMemcpyDeviceToHost));
    if (count)
        for (size_t a = 0; a < A; ++a) count[a] = b->count[a];
    return 0;
}

int mcac_box_members(mcac_box *b, int64_t agg, int64_t *out, int64_t max, int64_t *n) {
    if (!b || agg < 0 || agg >= b->n_aggs || !n) return MCAC_E_ARGUMENT;
    int cnt = b->count[static_cast<size_t>(agg)];
    *n = cnt;
    if (!out || cnt == 0) return 0;
    int *tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, static_cast<size_t>(cnt) * sizeof(int)));
    hipLaunchKernelGGL(k_gather_members, dim3(std::min(1024, (cnt + kBlock - 1) / kBlock)), dim3(kBlock), 0, bstream(b),
                       b->d, static_cast<int>(agg), tmp, cnt);
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    std::vector<int> h(static_cast<size_t>(cnt));
    HIP_TRY(hipMemcpy(h.data(), tmp, h.size() * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(hipFree(tmp));
    for (int64_t i = 0; i < std::min<int64_t>(max, cnt); ++i) out[i] = h[static_cast<size_t>(i)];
    return 0;
}

int mcac_box_download_members(mcac_box *b, int64_t *count, int64_t *members) {
    if (!b || !count || !members || b->inflight) return MCAC_E_ARGUMENT;
    const size_t A = static_cast<size_t>(b->n_aggs);
    std::vector<int> st(A), cnt(A), pool(static_cast<size_t>(std::max<int64_t>(b->pool_used, 1)));
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    if (A) {
        HIP_TRY(hipMemcpy(st.data(), b->d.astart, A * sizeof(int), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(cnt.data(), b->d.acount, A * sizeof(int), hipMemcpyDeviceToHost));
    }
    HIP_TRY(hipMemcpy(pool.data(), b->d.members, pool.size() * sizeof(int), hipMemcpyDeviceToHost));
    int64_t off = 0;
    for (size_t a = 0; a < A; ++a) {
        const int c = cnt[a] > 0 ? cnt[a] : 0;
        count[a] = c;
        if (c > 0 && (st[a] < 0 || static_cast<size_t>(st[a]) + static_cast<size_t>(c) > pool.size()))
            return MCAC_E_DEVICE;
        for (int i = 0; i < c; ++i) members[off + i] = pool[static_cast<size_t>(st[a]) + static_cast<size_t>(i)];
        off += c;
    }
    return 0;
}

int mcac_box_sphere_owners(mcac_box *b, int64_t *owner) {
    if (!b || !owner || b->inflight) return MCAC_E_ARGUMENT;
    std::vector<int> ids;
    for (size_t id = 0; id < b->count.size(); ++id)
        if (b->count[id] > 0) ids.push_back(static_cast<int>(id));
    const size_t S = static_cast<size_t>(b->n_spheres);
    int *d_ids = nullptr, *d_own = nullptr;
    HIP_TRY(hipMalloc(&d_ids, ids.size() * sizeof(int) + 4));
    HIP_TRY(hipMalloc(&d_own, S * sizeof(int) + 4));
    HIP_TRY(hipMemset(d_own, 0xff, S * sizeof(int) + 4));
    HIP_TRY(hipMemcpy(d_ids, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!ids.empty())
        hipLaunchKernelGGL(k_sphere_owner, dim3(static_cast<unsigned>(std::min<size_t>(ids.size(), 65535))),
                           dim3(kBlock), 0, bstream(b), b->d, d_ids, static_cast<int>(ids.size()), d_own);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    std::vector<int> h(S);
    if (S) HIP_TRY(hipMemcpy(h.data(), d_own, S * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(hipFree(d_ids));
    HIP_TRY(hipFree(d_own));
    for (size_t i = 0; i < S; ++i) owner[i] = h[i];
    return 0;
}

int mcac_box_grow_ids(mcac_box *b, const int64_t *ids, int64_t n_ids, double dr, int grow, double rp_min,
                      int check_split, double *volume, double *surface, double *radius_sum, double *rg, double *rmax,
                      int32_t *flags) {
    if (!b || !volume || !surface || !radius_sum || !rg || !rmax || b->inflight || b->inline_pending ||
        !(dr == dr) || (ids && n_ids < 0))
        return MCAC_E_ARGUMENT;
    const int64_t n = ids ? n_ids : static_cast<int64_t>(b->alive_ids.size());
    if (n == 0) return 0;
    if (ids)
        for (int64_t i = 0; i < n; ++i)
            if (ids[i] < 0 || ids[i] >= b->n_aggs || b->count[static_cast<size_t>(ids[i])] <= 0) return MCAC_E_ARGUMENT;
    // layout: five outputs per aggregate (doubles), its flags, the ids (ints), the
    // count of aggregates that left their cell
    const size_t N = static_cast<size_t>(n);
    auto bytes_for = [](size_t m) { return 5 * m * sizeof(double) + 2 * m * sizeof(int) + sizeof(double); };
    if (b->grow_cap < n) {
        if (b->h_grow) HIP_TRY(hipHostFree(b->h_grow));
        if (b->d_grow) HIP_TRY(hipFree(b->d_grow));
        b->h_grow = nullptr;
        b->d_grow = nullptr;
        b->grow_cap = 0;
        HIP_TRY(hipMalloc(reinterpret_cast<void **>(&b->d_grow), bytes_for(N)));
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&b->h_grow), bytes_for(N), hipHostMallocDefault));
        b->grow_cap = n;
    }
    double *g = b->d_grow;
    int *d_flag = reinterpret_cast<int *>(g + 5 * N);
    int *d_ids = d_flag + N;
    int *left = d_ids + N;
    hipStream_t st = bstream(b);
    const int *idp = b->d.alive;
    if (ids) {
        int *h_ids = reinterpret_cast<int *>(b->h_grow + 5 * N) + N;
        for (size_t i = 0; i < N; ++i) h_ids[i] = static_cast<int>(ids[i]);
        HIP_TRY(hipMemcpyAsync(d_ids, h_ids, N * sizeof(int), hipMemcpyHostToDevice, st));
        idp = d_ids;
    }
    HIP_TRY(hipMemsetAsync(left, 0, sizeof(int), st));
    const unsigned blocks = static_cast<unsigned>((n * 64 + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_grow_update, dim3(blocks), dim3(kBlock), 0, st, b->d, idp, static_cast<int>(n), dr,
                       grow ? 1 : 0, rp_min, b->p.box_length, static_cast<int>(b->p.n_div), g, g + N, g + 2 * N,
                       g + 3 * N, g + 4 * N, left, d_flag);
    if (check_split)
        hipLaunchKernelGGL(k_connected, dim3(blocks), dim3(kBlock), 0, st, b->d, idp, static_cast<int>(n),
                           b->p.box_length, d_flag);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(b->h_grow, g, 5 * N * sizeof(double) + N * sizeof(int), hipMemcpyDeviceToHost, st));
    int n_left = 0;
    HIP_TRY(hipMemcpyAsync(&n_left, left, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const int *h_flag = reinterpret_cast<const int *>(b->h_grow + 5 * N);
    for (size_t w = 0; w < N; ++w) {
        const int id = ids ? static_cast<int>(ids[w]) : b->alive_ids[w];
        if (b->count[static_cast<size_t>(id)] <= 0) continue;
        volume[id] = b->h_grow[w];
        surface[id] = b->h_grow[N + w];
        radius_sum[id] = b->h_grow[2 * N + w];
        rg[id] = b->h_grow[3 * N + w];
        rmax[id] = b->h_grow[4 * N + w];
        if (flags) flags[id] = h_flag[w];
    }
    if (n_left > 0) b->cells_stale = true;  // else the search's copies were refreshed in place
    b->mirrors_exact = false;
    return 0;
}

int mcac_box_grow(mcac_box *b, double dr, double *volume, double *surface, double *radius_sum, double *rg,
                  double *rmax) {
    if (!(dr >= 0.0)) return MCAC_E_ARGUMENT;
    return mcac_box_grow_ids(b, nullptr, 0, dr, 1, -1.0, 0, volume, surface, radius_sum, rg, rmax, nullptr);
}

int mcac_box_set_engine(mcac_box *b, int blocks) {
    if (!b || blocks < 0 || blocks > 255 || (blocks > 0 && b->group)) return MCAC_E_ARGUMENT;
    engine_stop(b);
    b->eng_blocks = blocks;
    return 0;
}

int mcac_box_engine_stats(mcac_box *b, double out[8]) {
    if (!b || !out) return MCAC_E_ARGUMENT;
    SearchState s{};
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    HIP_TRY(hipMemcpy(&s, b->d.ss, sizeof(SearchState), hipMemcpyDeviceToHost));
    out[0] = static_cast<double>(b->eng_launches);
    out[1] = static_cast<double>(b->eng_posted);
    out[2] = static_cast<double>(s.eng_steps);
    out[3] = static_cast<double>(s.eng_busy) / b->eng_clock_khz;  // ms
    out[4] = static_cast<double>(b->eng_relaunches);
    out[5] = static_cast<double>(b->h_ctl->reason);
    out[6] = static_cast<double>(s.eng_idle) / b->eng_clock_khz;  // ms
    out[7] = static_cast<double>(b->eng_grid);
    return 0;
}

int mcac_box_set_timing(mcac_box *b, int enable) {
    if (!b) return MCAC_E_ARGUMENT;
    if (enable > 0)  // the per-slot events of timed steps, made on first use
        for (auto &row : b->sev)
            for (auto &e : row)
                if (!e) HIP_TRY(hipEventCreate(&e));
    b->timing = enable < 0 ? 0 : enable;
    b->timing_base = b->seq + 1;  // the next search is the first timed one
    for (int k = 0; k < mcac_box::kRing; ++k) b->slot_timed[k] = 0;
    b->t_search = b->t_translate = b->t_merge = b->t_filter = b->t_solve = 0;
    b->n_search = b->n_translate = b->n_merge = 0;
    return 0;
}

int mcac_box_stats(mcac_box *b, int64_t out[8]) {
    if (!b || !out) return MCAC_E_ARGUMENT;
    out[6] = b->n_inline;
    out[7] = b->n_host_merge;
    out[0] = b->n_fused;
    out[1] = b->n_split;
    out[2] = b->n_handover;
    out[3] = b->n_rebuilds;
    out[4] = b->n_compactions;
    out[5] = b->pool_used;
    return 0;
}

int mcac_box_prune_stats(mcac_box *b, int64_t out[3]) {
    if (!b || !out) return MCAC_E_ARGUMENT;
    SearchState s{};
    HIP_TRY(hipStreamSynchronize(bstream(b)));
    HIP_TRY(hipMemcpy(&s, b->d.ss, sizeof(SearchState), hipMemcpyDeviceToHost));
    out[0] = static_cast<int64_t>(s.fp32_tested);
    out[1] = static_cast<int64_t>(s.fp32_kept);
    out[2] = static_cast<int64_t>(s.fp32_missed);
    return 0;
}

int mcac_box_timing(mcac_box *b, double out[8]) {
    if (!b || !out) return MCAC_E_ARGUMENT;
    for (int k = 0; k < mcac_box::kRing; ++k)
        if (b->slot_timed[k] == 2) {
            int rc = harvest_timing(b, k);
            if (rc) return rc;
        }
    out[0] = b->t_search;
    out[1] = b->t_translate;
    out[2] = b->t_merge;
    out[3] = b->n_search;
    out[4] = b->n_translate;
    out[5] = b->n_merge;
    out[6] = b->t_filter;
    out[7] = b->t_solve;
    return 0;
}

}  // extern "C"
