// Device cell list of the contact search (a4/a5): live aggregates bucketed by
// their reference Verlet cell (aggregat.cpp:699-705) in CSR form.  Moves and
// merges do not edit the CSR; they drop the aggregate's packed entry (CoM and
// rmax copied in CSR order) and append it to a dirty list, which the search
// scans in full.  The host rebuilds the CSR when
// the dirty list grows past ~sqrt(6 N) entries (rebuild cost ~3 N reads vs
// ~half a dirty list per step).
#pragma once
#include "kernels_common.hpp"

namespace mcac_dev {

__global__ __launch_bounds__(kBlock) void k_cell_clear(DevBox b, int n_cells) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= n_cells; c += gridDim.x * blockDim.x) b.ccursor[c] = 0;
}

__global__ __launch_bounds__(kBlock) void k_cell_count(DevBox b, int n_alive) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n_alive; t += gridDim.x * blockDim.x) {
        int id = b.alive[t];
        int c = b.acell[id];
        if (c >= 0) atomicAdd(&b.ccursor[c], 1);
    }
}

// Exclusive scan of the counts into cstart and ccursor; resets the dirty list.
__global__ __launch_bounds__(1024) void k_cell_scan(DevBox b, int n_cells) {
    __shared__ int part[1024];
    const int per = (n_cells + 1023) / 1024;
    const int lo = threadIdx.x * per;
    const int hi = (lo + per < n_cells) ? lo + per : n_cells;
    int s = 0;
    for (int i = lo; i < hi; ++i) s += b.ccursor[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int i = 0; i < 1024; ++i) {
            int v = part[i];
            part[i] = acc;
            acc += v;
        }
        b.cstart[n_cells] = acc;
        b.ss->dirty_count = 0;
    }
    __syncthreads();
    int acc = part[threadIdx.x];
    for (int i = lo; i < hi; ++i) {
        int v = b.ccursor[i];
        b.cstart[i] = acc;
        b.ccursor[i] = acc;
        acc += v;
    }
}

__global__ __launch_bounds__(kBlock) void k_cell_scatter(DevBox b, int n_alive) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n_alive; t += gridDim.x * blockDim.x) {
        int id = b.alive[t];
        int c = b.acell[id];
        b.dirty[id] = (c >= 0) ? 0 : 1;
        int pos = -1;
        if (c >= 0) {
            pos = atomicAdd(&b.ccursor[c], 1);
            b.citems[pos] = id;
            b.cpx[pos] = b.ax[id];
            b.cpy[pos] = b.ay[id];
            b.cpz[pos] = b.az[id];
            b.cpr[pos] = b.armax[id];
            b.cf[pos] = make_float4(static_cast<float>(b.ax[id]), static_cast<float>(b.ay[id]),
                                    static_cast<float>(b.az[id]), static_cast<float>(b.armax[id]));
        }
        b.cslot[id] = pos;
    }
}

// The four kernels above in one workgroup, for grids of at most kSmallCells
// cells and at most kSmallBuild live aggregates: counts, scan and cursors in
// LDS, one launch instead of four.  Same result: cstart, the CSR entries (their
// order inside a cell is the order of the cursor atomics, as above) and every
// live id's slot and dirty flag.  Its time grows with the live aggregates
// (108 us for 13,000 on one workgroup), past kSmallBuild the grid form is
// faster.
constexpr int kSmallCells = 8192;
constexpr int kSmallBuild = 4096;
constexpr int kBuildThreads = 1024;
__global__ __launch_bounds__(kBuildThreads) void k_cell_build_small(DevBox b, int n_alive, int n_cells) {
    __shared__ int cnt[kSmallCells];
    __shared__ int part[kBuildThreads];
    const int tid = threadIdx.x;
    for (int c = tid; c < n_cells; c += kBuildThreads) cnt[c] = 0;
    __syncthreads();
    for (int t = tid; t < n_alive; t += kBuildThreads) {
        const int c = b.acell[b.alive[t]];
        if (c >= 0) atomicAdd(&cnt[c], 1);
    }
    __syncthreads();
    // exclusive scan: each thread sums a contiguous run of cells, the run
    // totals are scanned in LDS (Hillis-Steele), then each run is written
    const int per = (n_cells + kBuildThreads - 1) / kBuildThreads;
    const int lo = tid * per < n_cells ? tid * per : n_cells;
    const int hi = lo + per < n_cells ? lo + per : n_cells;
    int s = 0;
    for (int i = lo; i < hi; ++i) s += cnt[i];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < kBuildThreads; off <<= 1) {
        const int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int acc = part[tid] - s;
    for (int i = lo; i < hi; ++i) {
        const int v = cnt[i];
        cnt[i] = acc;
        b.cstart[i] = acc;
        acc += v;
    }
    if (tid == kBuildThreads - 1) {
        b.cstart[n_cells] = part[kBuildThreads - 1];
        b.ss->dirty_count = 0;
    }
    __syncthreads();
    for (int t = tid; t < n_alive; t += kBuildThreads) {
        const int id = b.alive[t];
        const int c = b.acell[id];
        b.dirty[id] = (c >= 0) ? 0 : 1;
        int pos = -1;
        if (c >= 0) {
            pos = atomicAdd(&cnt[c], 1);
            b.citems[pos] = id;
            b.cpx[pos] = b.ax[id];
            b.cpy[pos] = b.ay[id];
            b.cpz[pos] = b.az[id];
            b.cpr[pos] = b.armax[id];
            b.cf[pos] = make_float4(static_cast<float>(b.ax[id]), static_cast<float>(b.ay[id]),
                                    static_cast<float>(b.az[id]), static_cast<float>(b.armax[id]));
        }
        b.cslot[id] = pos;
    }
}

}  // namespace mcac_dev
