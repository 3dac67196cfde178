// Device-resident state of one aggregation box and the small records that
// kernels publish to pinned host memory.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/mcac_gpu.h"

namespace mcac_dev {

constexpr int kBlock = 256;            // threads per block (4 waves of 64)
constexpr int kFilterBlocks = 128;     // contact-search filter grid
constexpr int kSolveBlocks = 256;      // pair-solve grid (one block per CU)
constexpr long long kChunk = 4096;     // sphere pairs per block the host sizes the grid form's grid by
constexpr long long kMinChunk = 256;   // minimum sphere pairs per solve chunk (one per thread)
constexpr int kMaxVisit = 2048;        // visited cells handled with an LDS prefix
constexpr int kMaxPlan = 4096;         // filtered candidates planned in LDS
constexpr int kOverlapSlots = 16;      // contacts kept per sphere row (multi-kernel merge)
constexpr int kSmallMerge = 256;       // merged aggregates up to this size use one fused kernel
constexpr int kFusedMerge = 512;       // overlap-correction merges up to this size run in one workgroup
constexpr int kStepCands = 32;         // k_step solves up to this many survivors itself
constexpr long long kStepPairs = 16384;  // ... holding up to this many sphere pairs
constexpr int kStepMover = 1024;       // movers up to this size may take k_step
constexpr int kThreadPairs = 4;        // candidates with this few sphere pairs are solved by one thread
constexpr int kPendingSolve = -1;      // StepRecord.status: pairs left for k_solve

// Counters and results of the contact search that stay on the device.
struct SearchState {
    unsigned int nfilt;        // filtered candidates appended by k_filter
    unsigned int overflow;     // candidate buffer overflow flag
    unsigned int done_solve;   // ticket of the participating solve blocks
    unsigned int dirty_count;  // entries of the dirty list
    unsigned long long nquery; // aggregates of the visited cells
    double newcom[3];          // position of the mover after the move
    double refpos[3];          // new CoM - relative CoM (sphere 0 position)
    int newcell;
    int pad;
    unsigned long long gate;   // seq of a gated step that found a contact (0: open)
    unsigned long long tick;   // k_step blocks: finished << 56 | appended << 32 | visited
    unsigned long long fp32_tested, fp32_kept, fp32_missed;  // FP32 broad-phase accounting
    // persistent step engine (k_step_engine)
    unsigned long long eng_done;      // 2 * last job index completed + 1 if that step closed the gate
    unsigned long long eng_decision;  // job index * 4 + outcome (kEngineRun, kEngineQuit, kEngineExit), block 0
    unsigned long long eng_busy;      // clock ticks of completed steps (finishing block: done wait to done)
    unsigned long long eng_steps;     // steps completed by the engine
    unsigned long long eng_idle;      // clock ticks block 0 waited for posted jobs
    // k_step_grid: blocks past the filter (grid-wide barrier), and a flag set
    // when a block gave up waiting at it
    unsigned int grid_bar, grid_fail;
};

// Record published to pinned host memory by the last block of a search.
struct StepRecord {
    double dist, move;
    long long ms, os, oagg;
    int contact, status;
    unsigned long long nquery, nfilt, npairs;
    int skipped, dirty_count;  // step skipped behind a closed gate; dirty list length after it
    mcac_merge_result merge;   // merge done in the same launch (merged_inline)
    int merged_inline, m_new_start, m_new_cap, pad;
    unsigned long long seq;    // written last (system-scope release)
};

// Record published by the last merge kernel.
struct MergeRecord {
    mcac_merge_result res;
    unsigned long long seq;
};

// Merge work shared between the kernels of the multi-kernel merge.
struct MergeState {
    int ok, keep, rem, n_k, n_r, new_start, start_k, start_r;
    double diffpos[3], refpos[3], r[3];
    unsigned long long rmax_bits;
    unsigned int row_overflow, pad;
    // list-order sums of the first alphas phase: V, S, Σ rel·vol (3), Σ r, Σ vol·r², Σ vol
    double sums[8];
    mcac_merge_result res;
};

struct DevBox {
    // spheres, index order
    double *sx, *sy, *sz, *sr, *svol, *ssurf, *srx, *sry, *srz;
    // aggregates, stable-id order
    double *ax, *ay, *az, *arx, *ary, *arz, *armax, *avol, *asurf, *arg;
    double *csvx, *csvy, *csvz, *csr, *csvr2;  // list-order partial sums the merge continues
    int *astart, *acount, *acell;              // member segment, size, reference Verlet cell (-1 dead)
    int *acap;                                 // member segment capacity
    int *members;
    int *alive;                                // live ids, ascending (may hold dead ids)
    // cell list: CSR by reference cell, built from the live ids
    int *cstart, *citems, *ccursor;
    double *cpx, *cpy, *cpz, *cpr;             // CoM and rmax in CSR order (rmax < 0: entry dropped)
    int *cslot;                                // CSR position of each aggregate (-1: none)
    float4 *cf;                                // FP32 copy of (CoM, rmax) in CSR order (w < 0: dropped)
    unsigned char *dirty;                      // moved, merged or removed since the build
    int *dirty_list;                           // ids in the order they became dirty
    double *dpx, *dpy, *dpz, *dpr;             // their current CoM and rmax, by dirty-list index
    int *dpcell;                               // their current reference cell (-1: removed)
    int *dslot;                                // dirty-list index of each dirty aggregate
    // search scratch
    double *cand_db, *cand_exact;
    int *cand_rank, *cand_id, *order;
    long long *cand_pair;
    double *part_d;
    long long *part_p;
    // merge scratch
    double *dc;
    int *row_cnt, *row_j;
    double *row_c;
    double *row_d;             // pair distances of the rows (overlap corrections only)
    double *cvol, *csurf;      // per-position corrected sphere volume and surface (spherical caps, arvo)
    double *sr3;               // pow(r, 3) per sphere, from the host's libm (alphas only)
    double *walk;              // per-position neighbour-map walk rows of (c_ij, vp1 + vp2, sp1 + sp2) (alphas only)
    int *walk_n;               // entries of each walk row (alphas only)
    SearchState *ss;
    MergeState *ms;
    struct EngineJob *eng_jobs;        // persistent engine: block 0's copy of each queued job, by ring slot
    unsigned long long *eng_closed;    // persistent engine: eng_done's word of each completed job, by ring slot
};

struct SearchArgs {
    int src;
    int n_mov, start;          // mover size and member segment (host mirror)
    int known;                 // host supplies the mover's CoM, rmax, relative CoM and window
    double sx, sy, sz, srmax, srx, sry, srz;
    int b1[3], len[3];
    int dirty_n;               // dirty-list length when this step runs
    int dirty_slot;            // dirty-list index this step's move appends, -1 when already listed
    int dirty_pos;             // dirty-list index of the mover after this step
    int n_div;
    int cap;
    int n_built;               // items in the CSR cell list
    long long part_cap;
    double dir[3];
    double lpm;
    double maxradius;
    double L;
    int apply;                 // translate by the resulting move
    int translate_in_finish;   // mover small enough for the finishing block
    StepRecord *out;           // pinned host record
    unsigned long long seq;
    int gated;                 // pipelined step: closes the gate on contact, skipped behind it
    unsigned long long released;  // gate values up to this one are open for this step
    int automerge;             // merge in the same launch on contact when small enough
    int prec;                  // 0 FP64; 1 FP32 capsule prune of cell entries; 2 as 1 plus miss count
    float margin32;            // FP32 prune margin added to the radius sum [m]
    long long pool_used, pool_cap;  // member pool state when this step runs
    int nc_ok;                 // the host computed the move without contact (known mover):
    int nc_cell;               //   its reference cell,
    double nc_com[3], nc_ref[3];  //   new CoM and CoM - relative CoM (aggregat.cpp:148-161)
};

// Bounds of the visited cells of a batched step, supplied by the host from
// its copy of the cell starts (n = 0: the kernel reads them from cstart).
struct CellTable {
    int n;
    int c0[64], sz[64];  // first CSR entry and entry count of visited cell v
    int pad;
};

// Persistent step engine (k_step_engine): queued steps of one box in a ring
// of pinned host memory, taken in order by a resident grid.
constexpr int kEngineRing = 64;  // ring slots (more than the steps in flight, 8, many times over)
constexpr int kEngineRun = 1, kEngineQuit = 2, kEngineExit = 3;
struct EngineJob {
    SearchArgs a;
    CellTable cells;           // bounds of the visited cells (n = 0: the kernel reads cstart)
    int cmd;                   // kEngineRun, or kEngineExit: the grid ends at this job
    int vg;                    // workgroups that take part in the step (the first vg of the grid)
    unsigned long long seq;    // job index, written last by the host (release)
};
// Written by the engine into pinned host memory when its grid ends.
struct EngineCtl {
    unsigned long long stopped_at;  // first job index this launch did not take
    unsigned long long launches_ended;  // number of the last launch that ended (launches count from 1)
    int reason;                     // kEngineQuit (idle), kEngineExit (host), or -1 (a wait ran out)
    int pad;
};

struct MergeArgs {
    long long ms, os, magg, oagg;
    int keep, rem, n_k, n_r, new_start, start_k, start_r, n_div;
    int cap_k;                 // keeper's segment capacity after the merge
    int volsurf;               // 0: sphere sums (method none); 1: spherical caps; 2: alphas; 3: arvo
    double L;
    MergeRecord *out;          // pinned host record
    unsigned long long seq;
};

}  // namespace mcac_dev
