// Block-level helpers shared by the kernels (wave64, kBlock threads).
#pragma once
#include "box_state.hpp"
#include "device_math.hpp"

namespace mcac_dev {

__device__ __forceinline__ void lexmin(double &d, long long &p, double od, long long op) {
    if (od < d || (od == d && op < p)) {
        d = od;
        p = op;
    }
}

// Lexicographic (d, p) minimum over the block; result valid in thread 0.
__device__ __forceinline__ void block_lexmin(double &d, long long &p) {
    __shared__ double sd[kBlock / 64];
    __shared__ long long sp[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        double od = __shfl_down(d, off, 64);
        long long op = __shfl_down(p, off, 64);
        lexmin(d, p, od, op);
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        sd[w] = d;
        sp[w] = p;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 1; i < kBlock / 64; ++i) lexmin(d, p, sd[i], sp[i]);
    __syncthreads();
}

// Sum over the block; result valid in thread 0.
template <class T> __device__ __forceinline__ T block_sum(T v) {
    __shared__ T s[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) s[w] = v;
    __syncthreads();
    T t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) t += s[i];
    __syncthreads();
    return t;
}

// Barrier for data exchanged through LDS only: waits for this wave's LDS
// operations but not for its global stores, which __syncthreads also drains.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Inclusive prefix sum over the 64 lanes of a wave.
template <class T> __device__ __forceinline__ T wave_inclusive_scan(T v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// In-place exclusive prefix sum of a[0..n) in LDS; writes the total to a[n].
// Each thread sums a contiguous run, the runs are scanned across the block
// with wave shuffles.  All threads of the block must call it.
template <class T> __device__ __forceinline__ void block_exclusive_scan(T *a, int n) {
    __shared__ T wsum[kBlock / 64];
    const int per = (n + kBlock - 1) / kBlock;
    const int lo = threadIdx.x * per;
    const int hi = (lo + per < n) ? lo + per : n;
    T s = 0;
    for (int i = lo; i < hi; ++i) s += a[i];
    const T incl = wave_inclusive_scan(s);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[w] = incl;
    lds_barrier();
    T off = 0;
    for (int k = 0; k < w; ++k) off += wsum[k];
    T acc = off + incl - s;
    if (threadIdx.x == kBlock - 1) a[n] = off + incl;
    for (int i = lo; i < hi; ++i) {
        T v = a[i];
        a[i] = acc;
        acc += v;
    }
    lds_barrier();
}

// aggregat.cpp:699-705 — reference Verlet cell of an aggregate position.
__host__ __device__ __forceinline__ int ref_cell(double x, double y, double z, int n, double L) {
    const double step = static_cast<double>(n) / L;
    const double p[3] = {x, y, z};
    int c[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        int v = static_cast<int>(__builtin_floor(p[l] * step));
        c[l] = v < 0 ? 0 : (v >= n ? n - 1 : v);
    }
    return (c[0] * n + c[1]) * n + c[2];
}

// verlet.cpp:57-78 — first visited cell and number of cells per axis.
__host__ __device__ __forceinline__ void verlet_window(const double src[3], const double vec[3], double mindist, int n,
                                              double W, int b1[3], int len[3]) {
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        double p = src[l] + mindist + smax(vec[l], 0.);
        double m = src[l] - mindist + smin(vec[l], 0.);
        int lo = static_cast<int>(__builtin_floor(static_cast<double>(n) * m / W));
        int hi = static_cast<int>(__builtin_floor(static_cast<double>(n) * p / W) + 1);
        if (hi - lo >= n) {
            lo = 0;
            hi = n - 1;
        }
        b1[l] = lo;
        len[l] = hi - lo + 1;
    }
}

// Offset of cell coordinate c in a window starting at b1 (n cells per axis).
__host__ __device__ __forceinline__ int window_offset(int c, int b1, int n) {
    int o = c - b1;
    if (o < 0) o += n;
    else if (o >= n) o -= n;
    return o;
}

// Marks an aggregate as moved since the cell-list build (single thread): its
// packed cell entry is dropped and it joins the dirty list.
__device__ __forceinline__ void mark_dirty(const DevBox &b, int id) {
    if (!b.dirty[id]) {
        const int cs = b.cslot[id];
        if (cs >= 0) {
            b.cpr[cs] = -1.0;
            b.cf[cs].w = -1.f;
        }
        b.dirty[id] = 1;
        unsigned int k = b.ss->dirty_count;
        b.dirty_list[k] = id;
        b.dslot[id] = static_cast<int>(k);
        b.ss->dirty_count = k + 1;
    }
}

// Current CoM, rmax and cell of a dirty aggregate, read by the search.
__device__ __forceinline__ void set_dirty_entry(const DevBox &b, int slot, double x, double y, double z, double r,
                                                int cell) {
    b.dpx[slot] = x;
    b.dpy[slot] = y;
    b.dpz[slot] = z;
    b.dpr[slot] = r;
    b.dpcell[slot] = cell;
}

}  // namespace mcac_dev
