// Contact search of one MC step (a4 + a6 + a7 + a8 + a9 + a13).
//
// k_filter   gathers the aggregates of the reference's visited Verlet cells
//            (CSR cells, skipping dirty entries, plus the whole dirty list),
//            runs the bounding-sphere sweep (sphere_contact.cpp:47-125 on
//            (CoM, rmax), aggregat_list.cpp:510-532) and appends the survivors
//            with their multimap key (bounding distance, visit rank, label).
// k_solve    splits the sphere pairs of all survivors into chunks over the
//            participating blocks (lexicographic (d, i, j) minimum per chunk,
//            aggregat_distance.cpp:24-44); the last participating block
//            reduces per candidate, orders the candidates like the multimap,
//            runs the reference early-exit scan (aggregat_list.cpp:459-482),
//            moves the aggregate (aggregat.cpp:148-161) when it is small, and
//            publishes the result to pinned host memory.
// k_step     both in one launch for small movers with few survivors.
// k_translate moves the spheres of a large mover.
#pragma once
#include "kernels_common.hpp"
#include "kernels_merge.hpp"

namespace mcac_dev {

__device__ __forceinline__ void append_candidate(const DevBox &b, const SearchArgs &a, int id, int rank, double d) {
    unsigned int slot = atomicAdd(&b.ss->nfilt, 1u);
    if (slot < static_cast<unsigned int>(a.cap)) {
        b.cand_db[slot] = d;
        b.cand_rank[slot] = rank;
        b.cand_id[slot] = id;
    } else {
        atomicOr(&b.ss->overflow, 1u);
    }
}

// FP32 broad phase: can the candidate's bounding sphere (centre c, radius rc)
// come within the mover's radius rs of the sweep segment p1 + t dir, t in
// [0, dist]?  Every swept-sphere contact within dist passes (the contact point
// lies on the segment); with dist + rs + rc < L/2 only the minimum image can
// be that close, and `margin` (far above FP32 rounding of box-sized
// coordinates) keeps the test conservative.
__device__ __forceinline__ bool capsule_pass32(float p1x, float p1y, float p1z, float rs, float4 c, float dx_,
                                               float dy_, float dz_, float dist, float L, float margin) {
    const float half = 0.5f * L;
    float d[3] = {c.x - p1x, c.y - p1y, c.z - p1z};
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        if (d[l] < -half) d[l] += L;
        else if (d[l] >= half) d[l] -= L;
    }
    float t = d[0] * dx_ + d[1] * dy_ + d[2] * dz_;
    t = t < 0.f ? 0.f : (t > dist ? dist : t);
    const float ex = d[0] - t * dx_, ey = d[1] - t * dy_, ez = d[2] - t * dz_;
    const float R = rs + c.w + margin;
    return ex * ex + ey * ey + ez * ez <= R * R;
}

// Per-thread counts of the filter: appended candidates << 32 | aggregates in the window.
constexpr unsigned long long kAppendOne = 1ull << 32;
constexpr unsigned long long kTicketOne = 1ull << 56;

// A gated step enqueued behind a step that found a contact (or handed its
// solve over) does nothing; the host releases a gate value by passing it as
// `released` to the steps it enqueues afterwards.
__device__ __forceinline__ bool behind_gate(const DevBox &b, const SearchArgs &a) {
    if (!a.gated) return false;
    const unsigned long long g = b.ss->gate;
    return g != 0ull && g < a.seq && g > a.released;
}

// Bounds of the visited cells, loaded by lane v of wave 0 for cell v of a
// host-supplied window of at most 64 cells (valid = 0: filter_phase loads them).
struct CellBounds {
    int valid, c0, sz;
};

__device__ __forceinline__ CellBounds load_cell_bounds(const DevBox &b, const SearchArgs &a, const CellTable *ct) {
    CellBounds cb{0, 0, 0};
    const int lj = a.len[1], lk = a.len[2];
    const int V = a.len[0] * lj * lk;
    if (!a.known || V > 64) return cb;
    if (ct && ct->n == V) {  // the host's table (in LDS): lane v takes cell v's pair
        cb.valid = 1;
        const int v = static_cast<int>(threadIdx.x);
        if (v < V) {
            cb.c0 = ct->c0[v];
            cb.sz = ct->sz[v];
        }
        return cb;
    }
    cb.valid = 1;
    const int v = static_cast<int>(threadIdx.x), n = a.n_div;
    if (v < V) {
        const int oi = v / (lj * lk), rem = v - oi * (lj * lk), oj = rem / lk, ok = rem - oj * lk;
        const int c = (window_offset(a.b1[0] + oi, 0, n) * n + window_offset(a.b1[1] + oj, 0, n)) * n +
                      window_offset(a.b1[2] + ok, 0, n);
        cb.c0 = b.cstart[c];
        cb.sz = b.cstart[c + 1] - cb.c0;
    }
    return cb;
}

// Filter work of block vb of the vg blocks on this step: the aggregates of the
// reference's visited Verlet cells and the dirty list, bounding sweep,
// candidate append (shared by k_filter, k_step and k_step_batch).  With skip
// the work is read-only (a step behind a closed gate appends nothing).
__device__ __forceinline__ unsigned long long filter_phase(const DevBox &b, const SearchArgs &a, unsigned vb,
                                                           unsigned vg, const CellBounds &cb, bool skip) {
    __shared__ int win[6];
    __shared__ double q[4];
    __shared__ int pre[kMaxVisit + 1];
    __shared__ int cbeg[kMaxVisit];
    __shared__ int nvisit;
    const int n = a.n_div;
    int b1i, b1j, b1k, li, lj, lk, V;
    double sx, sy, sz, sr;
    if (a.known) {
        // window and mover data from the host mirror (verlet.cpp:57-78)
        b1i = a.b1[0];
        b1j = a.b1[1];
        b1k = a.b1[2];
        li = a.len[0];
        lj = a.len[1];
        lk = a.len[2];
        sx = a.sx;
        sy = a.sy;
        sz = a.sz;
        sr = a.srmax;
        V = li * lj * lk;
    } else {
        if (threadIdx.x == 0) {
            double src[3] = {b.ax[a.src], b.ay[a.src], b.az[a.src]};
            double vec[3] = {a.lpm * a.dir[0], a.lpm * a.dir[1], a.lpm * a.dir[2]};
            double mindist = b.armax[a.src] + a.maxradius;  // aggregat_list.cpp:492
            int b1[3], len[3];
            verlet_window(src, vec, mindist, n, a.L, b1, len);
            for (int l = 0; l < 3; ++l) {
                win[l] = b1[l];
                win[3 + l] = len[l];
                q[l] = src[l];
            }
            q[3] = b.armax[a.src];
            nvisit = len[0] * len[1] * len[2];
        }
        __syncthreads();
        b1i = win[0];
        b1j = win[1];
        b1k = win[2];
        li = win[3];
        lj = win[4];
        lk = win[5];
        sx = q[0];
        sy = q[1];
        sz = q[2];
        sr = q[3];
        V = nvisit;
    }
    const int gtid = static_cast<int>(vb * blockDim.x + threadIdx.x);
    const int gstride = static_cast<int>(vg * blockDim.x);
    unsigned long long cnt = 0;
    int total;  // entries of the visited cells (or of the whole list for a very wide window)
    if (V <= 64) {
        // one wave: lane v loads cell v's start and size and the prefix runs in
        // registers (wave shuffles); one barrier publishes it
        if (threadIdx.x < 64) {
            const int v = threadIdx.x;
            int sz = 0, c0 = 0;
            if (v < V && cb.valid) {
                c0 = cb.c0;
                sz = cb.sz;
            } else if (v < V) {
                int oi = v / (lj * lk), rem = v - oi * (lj * lk), oj = rem / lk, ok = rem - oj * lk;
                int ci = window_offset(b1i + oi, 0, n), cj = window_offset(b1j + oj, 0, n),
                    ck = window_offset(b1k + ok, 0, n);
                int c = (ci * n + cj) * n + ck;
                c0 = b.cstart[c];
                sz = b.cstart[c + 1] - c0;
            }
            const int incl = wave_inclusive_scan(sz);
            if (v < V) {
                cbeg[v] = c0;
                pre[v] = incl - sz;
            }
            if (v == 63) pre[V] = incl;
        }
        lds_barrier();
        total = pre[V];
    } else if (V <= kMaxVisit) {
        // sizes and starts of the visited cells in visit order, then their prefix
        for (int v = threadIdx.x; v < V; v += blockDim.x) {
            int oi = v / (lj * lk), rem = v - oi * (lj * lk), oj = rem / lk, ok = rem - oj * lk;
            int ci = window_offset(b1i + oi, 0, n), cj = window_offset(b1j + oj, 0, n),
                ck = window_offset(b1k + ok, 0, n);
            int c = (ci * n + cj) * n + ck;
            const int c0 = b.cstart[c];
            cbeg[v] = c0;
            pre[v] = b.cstart[c + 1] - c0;
        }
        lds_barrier();
        block_exclusive_scan(pre, V);
        total = pre[V];
    } else {
        total = b.cstart[n * n * n];  // every entry of the build
    }
    // one index space: cell entries, then the aggregates moved since the build
    // (at their current cell), so that both kinds run side by side
    const int nd = a.dirty_n;
    unsigned long long n32 = 0, kept32 = 0, miss32 = 0;
    const float fx = static_cast<float>(sx), fy = static_cast<float>(sy), fz = static_cast<float>(sz);
    const float fr = static_cast<float>(sr), fdx = static_cast<float>(a.dir[0]), fdy = static_cast<float>(a.dir[1]);
    const float fdz = static_cast<float>(a.dir[2]), flpm = static_cast<float>(a.lpm), fL = static_cast<float>(a.L);
    if (a.prec == 0 && V <= kMaxVisit && total > gstride) {
        // FP64 sweep with several entries per thread (a block that takes a
        // window alone): kSweep entries per pass have their rank searches and
        // row loads in flight together instead of one load latency per entry
        constexpr int kSweep = 4;
        int top = 1;
        while (2 * top < V) top *= 2;
        for (int w0 = gtid; w0 < total; w0 += kSweep * gstride) {
            int rk[kSweep] = {}, idv[kSweep];
            double xv[kSweep], yv[kSweep], zv[kSweep], rv[kSweep];
            for (int s = top; s > 0; s >>= 1) {  // largest v with pre[v] <= w, as above
#pragma unroll
                for (int u = 0; u < kSweep; ++u) {
                    const int m = rk[u] + s;
                    if (m < V && pre[m] <= w0 + u * gstride) rk[u] = m;
                }
            }
#pragma unroll
            for (int u = 0; u < kSweep; ++u) {
                const int w = w0 + u * gstride;
                idv[u] = a.src;
                xv[u] = yv[u] = zv[u] = 0.0;
                rv[u] = -1.0;
                if (w < total) {
                    const int pos = cbeg[rk[u]] + (w - pre[rk[u]]);
                    idv[u] = b.citems[pos];
                    xv[u] = b.cpx[pos];
                    yv[u] = b.cpy[pos];
                    zv[u] = b.cpz[pos];
                    rv[u] = b.cpr[pos];
                }
            }
#pragma unroll
            for (int u = 0; u < kSweep; ++u) {
                if (rv[u] < 0.0 || idv[u] == a.src) continue;  // past the end, moved since the build, or the mover
                cnt += 1;
                const double d = swept_sphere(sx, sy, sz, sr, xv[u], yv[u], zv[u], rv[u], a.dir[0], a.dir[1],
                                              a.dir[2], a.lpm, a.L);
                if (d < a.lpm && !skip) {
                    append_candidate(b, a, idv[u], rk[u], d);
                    cnt += kAppendOne;
                }
            }
        }
        for (int i = gtid; i < nd; i += gstride) {  // the dirty list, as below
            const int id = b.dirty_list[i];
            const int c = b.dpcell[i];
            if (c < 0 || id == a.src) continue;
            int oi = window_offset(c / (n * n), b1i, n), oj = window_offset((c / n) % n, b1j, n),
                ok = window_offset(c % n, b1k, n);
            if (oi >= li || oj >= lj || ok >= lk) continue;
            cnt += 1;
            const double d = swept_sphere(sx, sy, sz, sr, b.dpx[i], b.dpy[i], b.dpz[i], b.dpr[i], a.dir[0], a.dir[1],
                                          a.dir[2], a.lpm, a.L);
            if (d < a.lpm && !skip) {
                append_candidate(b, a, id, (oi * lj + oj) * lk + ok, d);
                cnt += kAppendOne;
            }
        }
        return cnt;
    }
    for (int w = gtid; w < total + nd; w += gstride) {
        int id, rank;
        double x, y, z, r;
        if (w < total) {
            int pos;
            if (V <= kMaxVisit) {
                int lo = 0, hi = V;  // largest v with pre[v] <= w
                while (hi - lo > 1) {
                    int mid = (lo + hi) >> 1;
                    if (pre[mid] <= w) lo = mid;
                    else hi = mid;
                }
                rank = lo;
                pos = cbeg[lo] + (w - pre[lo]);
            } else {
                pos = w;
            }
            if (a.prec > 0) {
                // FP32 broad phase: 16 B per entry; only its survivors read the FP64 row
                const float4 f = b.cf[pos];
                if (f.w < 0.f) continue;  // moved since the build
                const bool pass = capsule_pass32(fx, fy, fz, fr, f, fdx, fdy, fdz, flpm, fL, a.margin32);
                n32 += 1;
                if (!pass) {
                    if (a.prec == 2) {  // would the FP64 filter have kept it?
                        const int id2 = b.citems[pos];
                        const double d2 = swept_sphere(sx, sy, sz, sr, b.cpx[pos], b.cpy[pos], b.cpz[pos],
                                                       b.cpr[pos], a.dir[0], a.dir[1], a.dir[2], a.lpm, a.L);
                        if (id2 != a.src && d2 < a.lpm) miss32 += 1;
                    }
                    continue;
                }
                kept32 += 1;
            }
            id = b.citems[pos];
            x = b.cpx[pos];
            y = b.cpy[pos];
            z = b.cpz[pos];
            r = b.cpr[pos];
            if (r < 0.0 || id == a.src) continue;  // moved since the build, or the mover
            if (V > kMaxVisit) {
                const int c = ref_cell(x, y, z, n, a.L);
                int oi = window_offset(c / (n * n), b1i, n), oj = window_offset((c / n) % n, b1j, n),
                    ok = window_offset(c % n, b1k, n);
                if (oi >= li || oj >= lj || ok >= lk) continue;
                rank = (oi * lj + oj) * lk + ok;
            }
        } else {
            const int i = w - total;
            id = b.dirty_list[i];
            const int c = b.dpcell[i];
            x = b.dpx[i];
            y = b.dpy[i];
            z = b.dpz[i];
            r = b.dpr[i];
            if (c < 0 || id == a.src) continue;
            int oi = window_offset(c / (n * n), b1i, n), oj = window_offset((c / n) % n, b1j, n),
                ok = window_offset(c % n, b1k, n);
            if (oi >= li || oj >= lj || ok >= lk) continue;
            rank = (oi * lj + oj) * lk + ok;
        }
        cnt += 1;
        const double d = swept_sphere(sx, sy, sz, sr, x, y, z, r, a.dir[0], a.dir[1], a.dir[2], a.lpm, a.L);
        if (d < a.lpm && !skip) {
            append_candidate(b, a, id, rank, d);
            cnt += kAppendOne;
        }
    }
    if (a.prec == 2) {  // FP32 broad-phase accounting, one atomic per block and counter
        n32 = block_sum(n32);
        kept32 = block_sum(kept32);
        miss32 = block_sum(miss32);
        if (threadIdx.x == 0) {
            atomicAdd(&b.ss->fp32_tested, n32);
            atomicAdd(&b.ss->fp32_kept, kept32);
            if (miss32) atomicAdd(&b.ss->fp32_missed, miss32);
        }
    }
    return cnt;
}

__global__ __launch_bounds__(kBlock) void k_filter(DevBox b, SearchArgs a) {
    if (behind_gate(b, a)) return;
    const unsigned long long tot =
        block_sum(filter_phase(b, a, blockIdx.x, gridDim.x, CellBounds{0, 0, 0}, false)) & 0xffffffffull;
    if (threadIdx.x == 0 && tot) atomicAdd(&b.ss->nquery, tot);
}

// What the last block of a step knows besides the candidates.
struct StepTotals {
    unsigned long long nquery;  // aggregates of the visited cells tested
    unsigned int nfilt;         // candidates appended (may exceed the capacity)
    int overflow;
    int src_slot;               // packed cell entry of the mover (-1: none)
    long long tpairs;           // sphere pairs of the survivors
};

// Survivors of a one-launch step (at most kStepCands), kept in LDS by the
// finishing block: bounding distance, visit rank, label, size and member
// segment, exact (d, pair) minimum, and the multimap order.
struct StepCands {
    double db[kStepCands], ex[kStepCands];
    long long pair[kStepCands];
    int rank[kStepCands], id[kStepCands], nk[kStepCands], sk[kStepCands], order[kStepCands];
};

// Sphere of the mover a thread places in the finishing block, loaded early
// when the mover has at most one sphere per thread.
struct MoverSphere {
    int loaded;  // the block loaded the mover (uniform)
    int sp;      // this thread's sphere, -1: none
    double rx, ry, rz;
};

__device__ __forceinline__ MoverSphere load_mover_sphere(const DevBox &b, const SearchArgs &a) {
    MoverSphere m{a.apply && a.n_mov <= kBlock, -1, 0., 0., 0.};
    if (m.loaded && static_cast<int>(threadIdx.x) < a.n_mov) {
        m.sp = b.members[a.start + threadIdx.x];
        m.rx = b.srx[m.sp];
        m.ry = b.sry[m.sp];
        m.rz = b.srz[m.sp];
    }
    return m;
}

// What a step loads first that no other step changes while it is queued
// behind them without a contact: the mover's packed-entry slot (thread 0) and
// its spheres with their relative positions.  The persistent engine loads it
// while the previous step finishes.
struct StepPrefetch {
    int src_slot;
    MoverSphere mover;
};

__device__ __forceinline__ StepPrefetch prefetch_step(const DevBox &b, const SearchArgs &a) {
    StepPrefetch p;
    p.src_slot = (threadIdx.x == 0) ? b.cslot[a.src] : -1;
    p.mover = load_mover_sphere(b, a);
    return p;
}

__device__ __forceinline__ long long chunk_count(long long pairs, long long chunk) {
    return (pairs + chunk - 1) / chunk;
}

// Last block of a step: orders the candidates like the reference's multimap,
// runs its early-exit scan over the exact per-candidate distances (from LDS
// when lc is given, else cand_exact / cand_pair), moves the aggregate,
// writes the record (publishing it unless the caller publishes later) and
// resets the search counters.  Returns (in thread 0) whether the step closed
// the gate for the steps queued behind it.
__device__ __forceinline__ bool finish_step(const DevBox &b, const SearchArgs &a, unsigned int F, long long nsrc,
                                            int ssrc, const StepTotals &t, const MoverSphere &m, StepCands *lc,
                                            bool publish) {
    bool closed = false;
    // std::multimap order: (bounding distance, visit rank, label)
    if (lc) {
        for (unsigned int c = threadIdx.x; c < F; c += blockDim.x) {
            const double dc = lc->db[c];
            const int rc = lc->rank[c], ic = lc->id[c];
            unsigned int pos = 0;
            for (unsigned int e = 0; e < F; ++e) {
                const double de = lc->db[e];
                const int re = lc->rank[e];
                pos += (de < dc || (de == dc && (re < rc || (re == rc && lc->id[e] < ic)))) ? 1u : 0u;
            }
            lc->order[pos] = static_cast<int>(c);
        }
        if (F > 0) lds_barrier();
    } else {
        for (unsigned int c = threadIdx.x; c < F; c += blockDim.x) {
            const double dc = b.cand_db[c];
            const int rc = b.cand_rank[c], ic = b.cand_id[c];
            unsigned int pos = 0;
            for (unsigned int e = 0; e < F; ++e) {
                double de = b.cand_db[e];
                int re = b.cand_rank[e];
                pos += (de < dc || (de == dc && (re < rc || (re == rc && b.cand_id[e] < ic)))) ? 1u : 0u;
            }
            b.order[pos] = static_cast<int>(c);
        }
        if (F > 0) __syncthreads();  // order and exact distances come from other threads
    }
    __shared__ double s_ref[3];
    __shared__ StepRecord srec;  // assembled here, copied to host memory by one wave
    SearchState *s = b.ss;
    if (threadIdx.x == 0) {
        double closest = kInf;
        int best = -1;
        for (unsigned int pos = 0; pos < F; ++pos) {
            const int c = lc ? lc->order[pos] : b.order[pos];
            if (closest <= 0.) break;
            if (closest < (lc ? lc->db[c] : b.cand_db[c])) break;
            const double ex = lc ? lc->ex[c] : b.cand_exact[c];
            if (ex < closest) {
                closest = ex;
                best = c;
            }
        }
        // every load of this thread comes before its stores (no reloads through global memory)
        long long ms = -1, os = -1, oagg = -1;
        if (best >= 0) {
            const long long p = lc ? lc->pair[best] : b.cand_pair[best];
            const int k = lc ? lc->id[best] : b.cand_id[best];
            const long long nk = lc ? lc->nk[best] : b.acount[k];
            const int sk = lc ? lc->sk[best] : b.astart[k];
            const long long i = p / nk, j = p - i * nk;
            ms = b.members[ssrc + i];
            os = b.members[sk + j];
            oagg = k;
        }
        // calcul.cpp:128-130 and aggregat.cpp:148-153
        const int contact = closest <= a.lpm;
        const double move = contact ? closest : a.lpm;
        double old[3], rc3[3], rmax_src;
        if (a.known) {
            old[0] = a.sx, old[1] = a.sy, old[2] = a.sz;
            rc3[0] = a.srx, rc3[1] = a.sry, rc3[2] = a.srz;
            rmax_src = a.srmax;
        } else {
            old[0] = b.ax[a.src], old[1] = b.ay[a.src], old[2] = b.az[a.src];
            rc3[0] = b.arx[a.src], rc3[1] = b.ary[a.src], rc3[2] = b.arz[a.src];
            rmax_src = b.armax[a.src];
        }
        double ncom[3], nref[3];
        int ncell;
        if (!contact && a.nc_ok) {  // the full move, computed by the host with the same operations
            for (int l = 0; l < 3; ++l) {
                ncom[l] = a.nc_com[l];
                nref[l] = a.nc_ref[l];
                s_ref[l] = nref[l];
            }
            ncell = a.nc_cell;
        } else {
            for (int l = 0; l < 3; ++l) {
                ncom[l] = periodic_position(old[l] + a.dir[l] * move, a.L);
                nref[l] = ncom[l] - rc3[l];
                s_ref[l] = nref[l];
            }
            ncell = ref_cell(ncom[0], ncom[1], ncom[2], a.n_div, a.L);
        }
        StepRecord *o = &srec;
        o->dist = closest;
        o->move = move;
        o->ms = ms;
        o->os = os;
        o->oagg = oagg;
        o->contact = contact;
        o->status = t.overflow ? MCAC_E_CAPACITY : 0;
        o->nquery = t.nquery;
        o->nfilt = t.nfilt;
        o->npairs = static_cast<unsigned long long>(t.tpairs);
        o->skipped = 0;
        o->merged_inline = 0;
        o->m_new_start = 0;
        o->m_new_cap = 0;
        o->pad = 0;
        o->dirty_count = a.dirty_n + ((a.apply && a.dirty_slot >= 0) ? 1 : 0);
        for (int l = 0; l < 3; ++l) {
            s->newcom[l] = ncom[l];
            s->refpos[l] = nref[l];
        }
        s->newcell = ncell;
        if (a.apply) {
            b.ax[a.src] = ncom[0];
            b.ay[a.src] = ncom[1];
            b.az[a.src] = ncom[2];
            b.acell[a.src] = ncell;
            if (a.dirty_slot >= 0) {  // first move since the cell-list build
                b.dirty[a.src] = 1;
                b.dirty_list[a.dirty_slot] = a.src;
                b.dslot[a.src] = a.dirty_slot;
                s->dirty_count = static_cast<unsigned int>(a.dirty_slot + 1);
                if (t.src_slot >= 0) {
                    b.cpr[t.src_slot] = -1.0;
                    b.cf[t.src_slot].w = -1.f;
                }
            }
            set_dirty_entry(b, a.dirty_pos, ncom[0], ncom[1], ncom[2], rmax_src, ncell);
        }
        if (a.gated && contact) s->gate = a.seq;  // the steps behind wait for the merge
        closed = a.gated && contact;
        s->nfilt = 0;
        s->nquery = 0;
        s->overflow = 0;
        s->done_solve = 0;
        s->tick = 0ull;
        s->grid_bar = 0;
    }
    lds_barrier();  // s_ref
    if (a.apply && a.translate_in_finish && m.loaded) {
        if (m.sp >= 0) {
            b.sx[m.sp] = s_ref[0] + m.rx;
            b.sy[m.sp] = s_ref[1] + m.ry;
            b.sz[m.sp] = s_ref[2] + m.rz;
        }
    } else if (a.apply && a.translate_in_finish) {
        const double r0 = s_ref[0], r1 = s_ref[1], r2 = s_ref[2];
        for (int t = threadIdx.x; t < nsrc; t += blockDim.x) {
            int sp = b.members[ssrc + t];
            b.sx[sp] = r0 + b.srx[sp];
            b.sy[sp] = r1 + b.sry[sp];
            b.sz[sp] = r2 + b.srz[sp];
        }
    }
    // merge in the same launch (AggregatList::merge, aggregat_list.cpp:367-410)
    // when the merged aggregate is small and its member segment fits
    if (a.automerge && a.apply && a.translate_in_finish) {
        __shared__ MergeArgs s_ma;
        __shared__ int s_inline;
        if (threadIdx.x == 0) {
            s_inline = 0;
            if (srec.contact) {
                const int oagg = static_cast<int>(srec.oagg);
                const int keep = a.src < oagg ? a.src : oagg, rem = a.src < oagg ? oagg : a.src;
                const int nk = (keep == a.src) ? a.n_mov : b.acount[keep];
                const int nr = (rem == a.src) ? a.n_mov : b.acount[rem];
                const int start_k = (keep == a.src) ? a.start : b.astart[keep];
                const int start_r = (rem == a.src) ? a.start : b.astart[rem];
                const int n = nk + nr;
                int new_start = start_k, new_cap = b.acap[keep];
                bool fits = n <= kSmallMerge;
                if (fits && n > new_cap) {  // a fresh segment of twice the merged size
                    new_cap = (2 * n > 4) ? 2 * n : 4;
                    fits = a.pool_used + new_cap <= a.pool_cap;
                    new_start = static_cast<int>(a.pool_used);
                }
                if (fits) {
                    MergeArgs m{};
                    m.ms = srec.ms;
                    m.os = srec.os;
                    m.magg = a.src;
                    m.oagg = oagg;
                    m.keep = keep;
                    m.rem = rem;
                    m.n_k = nk;
                    m.n_r = nr;
                    m.new_start = new_start;
                    m.start_k = start_k;
                    m.start_r = start_r;
                    m.n_div = a.n_div;
                    m.cap_k = new_cap;
                    m.L = a.L;
                    s_ma = m;
                    s_inline = 1;
                    srec.m_new_start = new_start;
                    srec.m_new_cap = new_cap;
                }
            }
        }
        lds_barrier();
        if (s_inline) {
            // the merge takes the mover's sphere positions from s_ref, and the
            // thread that placed a sphere of the mover also remaps it when the
            // mover is removed: no wait for the placements above
            merge_small_block(b, s_ma, srec.merge, s_ref);
            if (threadIdx.x == 0) srec.merged_inline = 1;  // thread 0 also copies the record
        }
    }
    // the record body in one coalesced write, then the sequence number
    constexpr int kWords = static_cast<int>(offsetof(StepRecord, seq) / sizeof(unsigned long long));
    static_assert(offsetof(StepRecord, seq) % sizeof(unsigned long long) == 0, "record body in 8-byte words");
    if (threadIdx.x < 64) {
        if (static_cast<int>(threadIdx.x) < kWords)
            reinterpret_cast<unsigned long long *>(a.out)[threadIdx.x] =
                reinterpret_cast<const unsigned long long *>(&srec)[threadIdx.x];
        if (threadIdx.x == 0 && publish) {
            __threadfence_system();
            __hip_atomic_store(&a.out->seq, a.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    return closed;
}

// Lexicographic (d, pair) minimum of the swept-sphere distances over pairs
// p = first, first + step, ... below end of the mover (spheres members[ssrc +
// i]) against a candidate of nk spheres (members[sk + j]), p = i nk + j
// (aggregat_distance.cpp:24-44).  kPairsInFlight pairs have their index and
// position loads issued together, instead of one dependent chain per pair; the
// minimum does not depend on the order pairs are taken in.
constexpr int kPairsInFlight = 4;
__device__ __forceinline__ void pair_scan(const DevBox &b, const SearchArgs &a, int ssrc, int sk, long long nk,
                                          long long first, long long end, long long step, double &bd,
                                          long long &bp) {
    for (long long p = first; p < end; p += kPairsInFlight * step) {
        int si[kPairsInFlight], sj[kPairsInFlight];
#pragma unroll
        for (int u = 0; u < kPairsInFlight; ++u) {
            const long long q = p + u * step;
            si[u] = sj[u] = -1;
            if (q < end) {
                const long long i = q / nk;
                si[u] = b.members[ssrc + i];
                sj[u] = b.members[sk + (q - i * nk)];
            }
        }
        double xi[kPairsInFlight], yi[kPairsInFlight], zi[kPairsInFlight], ri[kPairsInFlight];
        double xj[kPairsInFlight], yj[kPairsInFlight], zj[kPairsInFlight], rj[kPairsInFlight];
#pragma unroll
        for (int u = 0; u < kPairsInFlight; ++u) {
            const int i = si[u] >= 0 ? si[u] : 0, j = si[u] >= 0 ? sj[u] : 0;
            xi[u] = b.sx[i];
            yi[u] = b.sy[i];
            zi[u] = b.sz[i];
            ri[u] = b.sr[i];
            xj[u] = b.sx[j];
            yj[u] = b.sy[j];
            zj[u] = b.sz[j];
            rj[u] = b.sr[j];
        }
#pragma unroll
        for (int u = 0; u < kPairsInFlight; ++u) {
            if (si[u] < 0) continue;
            const double d = swept_sphere(xi[u], yi[u], zi[u], ri[u], xj[u], yj[u], zj[u], rj[u], a.dir[0], a.dir[1],
                                          a.dir[2], a.lpm, a.L);
            lexmin(bd, bp, d, p + u * step);
        }
    }
}

// Pair solve of the survivors filtered for step a (aggregat_distance.cpp:24-44):
// chunks of each candidate's sphere-pair space over the blocks of the grid,
// per-chunk lexicographic (d, pair) minima, and the last block to finish
// reduces them per candidate and finishes the step.
__device__ __forceinline__ void solve_phase(const DevBox &b, const SearchArgs &a) {
    __shared__ long long cpre[kMaxPlan + 1];
    __shared__ long long s_chunk;
    __shared__ bool last;
    unsigned int F = b.ss->nfilt;
    if (F > static_cast<unsigned int>(a.cap)) F = static_cast<unsigned int>(a.cap);
    const long long nsrc = a.n_mov;
    const int ssrc = a.start;
    // a few candidates' labels, sizes and member segments are loaded once,
    // into LDS, for the plan and the chunks
    constexpr unsigned int kStaged = 256;
    __shared__ int s_cid[kStaged], s_cnk[kStaged], s_csk[kStaged];
    const bool staged = F <= kStaged;
    // plan: chunks of the pair space of each candidate
    long long tpairs = 0;
    for (unsigned int c = threadIdx.x; c < F; c += blockDim.x) {
        const int k = b.cand_id[c];
        const int nk = b.acount[k];
        if (staged) {
            s_cid[c] = k;
            s_cnk[c] = nk;
            s_csk[c] = b.astart[k];
        }
        tpairs += nsrc * nk;
    }
    tpairs = block_sum(tpairs);  // its barrier also publishes the staged candidates
    if (threadIdx.x == 0) {
        // chunks sized to spread the pairs over the grid (at least kMinChunk
        // pairs each), fewer when the partials would not fit
        long long chunk = (tpairs + gridDim.x - 1) / gridDim.x;
        if (chunk < kMinChunk) chunk = kMinChunk;
        if (tpairs / chunk + F > a.part_cap) chunk = tpairs / (a.part_cap - F) + 1;
        s_chunk = chunk;
    }
    __syncthreads();
    const long long chunk = s_chunk;
    const bool planned = F <= static_cast<unsigned int>(kMaxPlan);
    long long nchunks = 0;
    if (planned) {
        for (unsigned int c = threadIdx.x; c < F; c += blockDim.x)
            cpre[c] = chunk_count(nsrc * (staged ? s_cnk[c] : b.acount[b.cand_id[c]]), chunk);
        __syncthreads();
        block_exclusive_scan(cpre, static_cast<int>(F));
        nchunks = cpre[F];
    } else {
        long long t = 0;
        for (unsigned int c = threadIdx.x; c < F; c += blockDim.x) t += chunk_count(nsrc * b.acount[b.cand_id[c]], chunk);
        nchunks = block_sum(t);
        __shared__ long long s_n;
        if (threadIdx.x == 0) s_n = nchunks;
        __syncthreads();
        nchunks = s_n;
    }
    const long long P = (nchunks < gridDim.x) ? (nchunks > 0 ? nchunks : 1) : gridDim.x;
    if (blockIdx.x >= P) return;
    for (long long g = blockIdx.x; g < nchunks; g += P) {
        unsigned int c = 0;
        long long c0 = 0;
        if (planned) {
            unsigned int lo = 0, hi = F;
            while (hi - lo > 1) {
                unsigned int mid = (lo + hi) >> 1;
                if (cpre[mid] <= g) lo = mid;
                else hi = mid;
            }
            c = lo;
            c0 = cpre[c];
        } else {
            long long acc = 0;
            for (c = 0; c < F; ++c) {
                long long k = chunk_count(nsrc * b.acount[b.cand_id[c]], chunk);
                if (g < acc + k) break;
                acc += k;
            }
            c0 = acc;
        }
        const int k = staged ? s_cid[c] : b.cand_id[c];
        const long long nk = staged ? s_cnk[c] : b.acount[k];
        const long long Ptot = nsrc * nk;
        const long long p0 = (g - c0) * chunk;
        const long long p1 = (p0 + chunk < Ptot) ? p0 + chunk : Ptot;
        const int sk = staged ? s_csk[c] : b.astart[k];
        double bd = kInf;
        long long bp = 0x7fffffffffffffffLL;
        pair_scan(b, a, ssrc, sk, nk, p0 + threadIdx.x, p1, blockDim.x, bd, bp);
        block_lexmin(bd, bp);
        if (threadIdx.x == 0) {
            b.part_d[g] = bd;
            b.part_p[g] = bp;
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned int ticket = atomicAdd(&b.ss->done_solve, 1u);
        last = (ticket == static_cast<unsigned int>(P) - 1u);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    // per-candidate reduction of the chunk partials: a candidate of a few
    // chunks by one thread, one of many (planned) by the whole block
    constexpr long long kThreadChunks = 8;
    for (unsigned int c = threadIdx.x; c < F; c += blockDim.x) {
        long long g0, g1;
        if (planned) {
            g0 = cpre[c];
            g1 = cpre[c + 1];
            if (g1 - g0 > kThreadChunks) continue;
        } else {
            g0 = 0;
            for (unsigned int e = 0; e < c; ++e) g0 += chunk_count(nsrc * b.acount[b.cand_id[e]], chunk);
            g1 = g0 + chunk_count(nsrc * b.acount[b.cand_id[c]], chunk);
        }
        double bd = kInf;
        long long bp = 0x7fffffffffffffffLL;
        for (long long g = g0; g < g1; ++g) lexmin(bd, bp, b.part_d[g], b.part_p[g]);
        b.cand_exact[c] = bd;
        b.cand_pair[c] = (bd < kInf) ? bp : -1;
    }
    if (planned) {
        for (unsigned int c = 0; c < F; ++c) {
            const long long g0 = cpre[c], g1 = cpre[c + 1];
            if (g1 - g0 <= kThreadChunks) continue;  // uniform over the block
            double bd = kInf;
            long long bp = 0x7fffffffffffffffLL;
            for (long long g = g0 + threadIdx.x; g < g1; g += blockDim.x) lexmin(bd, bp, b.part_d[g], b.part_p[g]);
            block_lexmin(bd, bp);
            if (threadIdx.x == 0) {
                b.cand_exact[c] = bd;
                b.cand_pair[c] = (bd < kInf) ? bp : -1;
            }
        }
    }
    __syncthreads();  // every candidate's minimum before the ordered scan reads them
    StepTotals t;
    t.nquery = b.ss->nquery;
    t.nfilt = b.ss->nfilt;
    t.overflow = b.ss->overflow != 0u;
    t.src_slot = b.cslot[a.src];
    t.tpairs = tpairs;
    finish_step(b, a, F, nsrc, ssrc, t, MoverSphere{0, -1, 0., 0., 0.}, nullptr, true);
}

__global__ __launch_bounds__(kBlock) void k_solve(DevBox b, SearchArgs a) {
    if (behind_gate(b, a)) return;  // the host knows it is skipped: no record
    solve_phase(b, a);
}

// One launch for a large mover: k_filter's work, a grid-wide barrier in place
// of the boundary between k_filter and k_solve, then k_solve's work by the same
// blocks; the same arithmetic in the same order, so the same step.  The grid
// is small enough to be resident at once (at most kFilterBlocks blocks, half
// the CUs; one box stream, launches in order); a block that waits past kGridWaitTicks at the
// barrier gives up and sets grid_fail, the step then publishes no record and
// the host reports a device error instead of waiting without end.
constexpr unsigned long long kGridWaitTicks = 2000000;  // 20 ms of the 100 MHz constant clock
__global__ __launch_bounds__(kBlock) void k_step_grid(DevBox b, SearchArgs a) {
    if (behind_gate(b, a)) return;  // every block reads the same gate: the host knows it is skipped
    const unsigned long long tot =
        block_sum(filter_phase(b, a, blockIdx.x, gridDim.x, CellBounds{0, 0, 0}, false)) & 0xffffffffull;
    __shared__ int s_ok;
    if (threadIdx.x == 0) {
        if (tot) atomicAdd(&b.ss->nquery, tot);
        __threadfence();  // this block's candidates and counts before its arrival
        atomicAdd(&b.ss->grid_bar, 1u);
        const unsigned long long t0 = wall_clock64();
        int ok = 0;
        while (wall_clock64() - t0 <= kGridWaitTicks) {
            if (__hip_atomic_load(&b.ss->grid_bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >= gridDim.x) {
                ok = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (!ok) atomicExch(&b.ss->grid_fail, 1u);
        s_ok = ok;
    }
    __syncthreads();
    if (!s_ok) return;
    __threadfence();  // every block's candidates are visible to this one
    solve_phase(b, a);
}

// One MC step for a small mover, worked on by vg blocks (vb is this one):
// every block filters its share of the Verlet window; the last block to finish
// (ticket) solves the sphere pairs of the survivors itself, one candidate at a
// time, and finishes the step.  When the survivors hold more pairs than one
// block should take, it leaves them in place, closes the gate for the steps
// queued behind and publishes a pending record: the host then launches k_solve
// for the same step.  With vg == 1 the block is the last one by construction.
// ct: bounds of the visited cells from the host (in LDS), or null.
// publish: publish the record (else wave 0 writes its body and the caller
// publishes the sequence number later, after a system-scope fence).
// gate_open: the caller knows the gate is open for this step; otherwise the
// gate is read with the first loads and a step behind it does no more than
// read-only filter work (the host knows it is skipped: no record).  Returns (in
// thread 0) whether the step closed the gate for the steps behind it.
// role (optional, LDS, set by thread 0): 0 the step was skipped behind the
// gate, 1 this block took part, 2 this block finished the step.
__device__ __forceinline__ bool step_body(const DevBox &b, const SearchArgs &a, unsigned vb, unsigned vg,
                                          bool gate_open, const CellTable *ct, bool publish, int *role = nullptr,
                                          const StepPrefetch *pre = nullptr) {
    // independent loads first, so that their latencies overlap: the gate, the
    // mover's cell entry and sphere indices, the bounds of the visited cells
    const unsigned long long gate = (!gate_open && a.gated) ? b.ss->gate : 0ull;
    const bool mload = a.apply && a.n_mov <= kBlock;
    int src_slot = -1, msp = -1;
    if (pre) {
        src_slot = pre->src_slot;
        msp = pre->mover.sp;
    } else {
        src_slot = (threadIdx.x == 0) ? b.cslot[a.src] : -1;
        msp = (mload && static_cast<int>(threadIdx.x) < a.n_mov) ? b.members[a.start + threadIdx.x] : -1;
    }
    const CellBounds cb = load_cell_bounds(b, a, ct);
    const bool skip = gate != 0ull && gate < a.seq && gate > a.released;
    MoverSphere mover{mload, msp, 0., 0., 0.};
    if (pre) {
        mover = pre->mover;
    } else if (msp >= 0) {
        mover.rx = b.srx[msp];
        mover.ry = b.sry[msp];
        mover.rz = b.srz[msp];
    }
    const unsigned long long cnt = block_sum(filter_phase(b, a, vb, vg, cb, skip));
    if (role && threadIdx.x == 0) *role = skip ? 0 : 1;
    if (skip) return false;
    // one atomic per block: finished blocks, appended candidates, visited aggregates
    __shared__ bool last;
    __shared__ unsigned long long s_total;
    if (vg > 1) {
        if (threadIdx.x == 0) {
            const unsigned long long mine = kTicketOne + cnt;
            __threadfence();
            const unsigned long long old = atomicAdd(&b.ss->tick, mine);
            last = (old >> 56) == vg - 1u;
            s_total = old + mine;
        }
        __syncthreads();
        if (!last) return false;
        __threadfence();
        if (role && threadIdx.x == 0) *role = 2;
    } else {
        if (role && threadIdx.x == 0) *role = 2;
        if (threadIdx.x == 0) s_total = cnt;
        __syncthreads();  // the block's appended candidates are visible to all its threads
    }
    SearchState *s = b.ss;
    const unsigned long long total = s_total;
    StepTotals t;
    t.nquery = total & 0xffffffffull;
    t.nfilt = static_cast<unsigned int>((total >> 32) & 0xffffffull);
    t.overflow = t.nfilt > static_cast<unsigned int>(a.cap);
    t.src_slot = src_slot;
    const unsigned int F = t.overflow ? static_cast<unsigned int>(a.cap) : t.nfilt;
    const long long nsrc = a.n_mov;
    const int ssrc = a.start;
    __shared__ StepCands sc;
    __shared__ int s_done[kStepCands];
    t.tpairs = 0;
    if (F > 0) {
        // survivors into LDS: label, size, member segment, bounding key
        long long tp = 0;
        for (unsigned int c = threadIdx.x; c < F; c += blockDim.x) {
            const int k = b.cand_id[c];
            const int nk = b.acount[k];
            tp += nsrc * nk;
            if (c < static_cast<unsigned int>(kStepCands)) {
                sc.id[c] = k;
                sc.nk[c] = nk;
                sc.sk[c] = b.astart[k];
                sc.db[c] = b.cand_db[c];
                sc.rank[c] = b.cand_rank[c];
            }
        }
        const long long tpairs = block_sum(tp);
        t.tpairs = tpairs;
        __shared__ bool handover;
        if (threadIdx.x == 0) handover = F > static_cast<unsigned int>(kStepCands) || tpairs > kStepPairs || t.overflow;
        lds_barrier();
        if (handover) {
            if (threadIdx.x == 0) {
                s->tick = 0ull;
                s->nquery = t.nquery;  // k_solve reports it
                if (a.gated) s->gate = a.seq;
                StepRecord *o = a.out;
                o->status = kPendingSolve;
                o->contact = 0;
                o->skipped = 0;
                if (publish) {
                    __threadfence_system();
                    __hip_atomic_store(&o->seq, a.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
                }
            }
            return a.gated != 0;
        }
        // exact sphere-pair distances, lexicographic (d, i, j) minimum per
        // candidate: a candidate with at most kThreadPairs pairs is walked by
        // one thread in (i, j) order (strict < keeps the first minimum), the
        // others by the whole block
        if (threadIdx.x < F) {
            const unsigned int c = threadIdx.x;
            const int nk = sc.nk[c];  // this thread's own entries
            const int sk = sc.sk[c];
            int done = 0;
            if (nsrc * nk <= kThreadPairs) {
                double bd = kInf;
                long long bp = 0x7fffffffffffffffLL;
                for (long long p = 0; p < nsrc * nk; ++p) {
                    const long long i = p / nk, j = p - i * nk;
                    const int si = b.members[ssrc + i], sj = b.members[sk + j];
                    const double d = swept_sphere(b.sx[si], b.sy[si], b.sz[si], b.sr[si], b.sx[sj], b.sy[sj],
                                                  b.sz[sj], b.sr[sj], a.dir[0], a.dir[1], a.dir[2], a.lpm, a.L);
                    lexmin(bd, bp, d, p);
                }
                sc.ex[c] = bd;
                sc.pair[c] = (bd < kInf) ? bp : -1;
                done = 1;
            }
            s_done[c] = done;
        }
        lds_barrier();
        for (unsigned int c = 0; c < F; ++c) {
            if (s_done[c]) continue;
            const long long nk = sc.nk[c];
            const int sk = sc.sk[c];
            double bd = kInf;
            long long bp = 0x7fffffffffffffffLL;
            pair_scan(b, a, ssrc, sk, nk, threadIdx.x, nsrc * nk, blockDim.x, bd, bp);
            block_lexmin(bd, bp);
            if (threadIdx.x == 0) {
                sc.ex[c] = bd;
                sc.pair[c] = (bd < kInf) ? bp : -1;
            }
        }
    }
    return finish_step(b, a, F, nsrc, ssrc, t, mover, &sc, publish);
}

__global__ __launch_bounds__(kBlock) void k_step(DevBox b, SearchArgs a) {
    step_body(b, a, blockIdx.x, gridDim.x, false, nullptr, true);
}

// Steps of many boxes in one launch: block i runs jobs seg[i] .. seg[i+1]-1,
// consecutive queued steps of one box (at most kMaxBlockJobs), in order; jobs
// are in pinned host memory.  Small boxes are bound by launch and memory
// latency, so their steps share launches instead of each taking one.  A step
// behind a contact (or a hand-over to k_solve) of an earlier job of the block
// does nothing, as it would in separate launches behind the gate.
struct StepJob {
    DevBox b;
    SearchArgs a;
    CellTable cells;
};
constexpr int kMaxBlockJobs = 8;
static_assert(sizeof(DevBox) % sizeof(unsigned long long) == 0 && sizeof(SearchArgs) % sizeof(unsigned long long) == 0 &&
                  sizeof(CellTable) % sizeof(unsigned long long) == 0,
              "job copied in 8-byte words");

__global__ __launch_bounds__(kBlock) void k_step_batch(const StepJob *jobs, const int *seg) {
    __shared__ DevBox box;
    __shared__ SearchArgs args[kMaxBlockJobs];
    __shared__ CellTable cells[kMaxBlockJobs];
    __shared__ int s_closed;
    constexpr int kBoxWords = static_cast<int>(sizeof(DevBox) / sizeof(unsigned long long));
    constexpr int kArgWords = static_cast<int>(sizeof(SearchArgs) / sizeof(unsigned long long));
    constexpr int kCellWords = static_cast<int>(sizeof(CellTable) / sizeof(unsigned long long));
    const int j0 = seg[blockIdx.x];
    int nj = seg[blockIdx.x + 1] - j0;
    nj = nj < 0 ? 0 : (nj > kMaxBlockJobs ? kMaxBlockJobs : nj);  // the host never lists more
    // one round trip: the box (the same for every job of the block), every
    // job's arguments and cell table
    unsigned long long *bw = reinterpret_cast<unsigned long long *>(&box);
    unsigned long long *aw = reinterpret_cast<unsigned long long *>(args);
    unsigned long long *cw = reinterpret_cast<unsigned long long *>(cells);
    const int na = nj * kArgWords;
    for (int i = threadIdx.x; i < kBoxWords + na + nj * kCellWords; i += blockDim.x) {
        if (i < kBoxWords) {
            bw[i] = reinterpret_cast<const unsigned long long *>(&jobs[j0].b)[i];
        } else if (i < kBoxWords + na) {
            const int q = (i - kBoxWords) / kArgWords, w = (i - kBoxWords) - q * kArgWords;
            aw[q * kArgWords + w] = reinterpret_cast<const unsigned long long *>(&jobs[j0 + q].a)[w];
        } else {
            const int q = (i - kBoxWords - na) / kCellWords, w = (i - kBoxWords - na) - q * kCellWords;
            cw[q * kCellWords + w] = reinterpret_cast<const unsigned long long *>(&jobs[j0 + q].cells)[w];
        }
    }
    if (threadIdx.x == 0) s_closed = 0;
    __syncthreads();
    if (nj == 0 || behind_gate(box, args[0])) return;  // every job of the block is behind the same gate
    int ran = 0;
    // the mover's entry slot and spheres of the next job load while a step
    // runs: no queued step changes them unless it closes the gate (the slot
    // holds until the next cell-list build, a move keeps the relative
    // positions, a merge is a contact)
    StepPrefetch next = prefetch_step(box, args[0]);
    for (int q = 0; q < nj; ++q) {
        // the previous step's writes (positions, lists) and its use of the
        // shared state are complete before the next step starts
        if (q > 0) __syncthreads();
        if (s_closed) break;
        const StepPrefetch cur = next;
        if (q + 1 < nj) next = prefetch_step(box, args[q + 1]);
        const bool closed = step_body(box, args[q], 0u, 1u, true, &cells[q], false, nullptr, &cur);
        ran = q + 1;
        if (threadIdx.x == 0 && closed) s_closed = 1;
    }
    // the records of the steps that ran, written by wave 0: one system-scope
    // fence, then their sequence numbers (the host reads them after the launch)
    if (threadIdx.x == 0 && ran > 0) {
        __threadfence_system();
        for (int q = 0; q < ran; ++q)
            __hip_atomic_store(&args[q].out->seq, args[q].seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Persistent form of k_step for one box: a resident grid takes the queued
// steps from a ring in pinned host memory, in order, instead of one launch
// per step.  Block 0 alone polls the ring; it copies job k into device memory
// (the box's eng_jobs slot) and publishes its decision for k (run, exit, or
// quit after idle_ticks of clock without a job) in eng_decision; the other
// blocks wait for that decision, which usually comes while the previous step
// still runs.  Each block then loads what the step reads of the mover (no
// queued step changes it unless it closes the gate) and waits until the last
// step that ran is done: its finishing block writes eng_done = 2 k + (1 if it
// closed the gate) after the step's device writes and before it publishes the
// record to the host, so the next step starts one host round trip earlier.
// Every engine step is gated and its released floor never decreases, so a
// step is behind the gate exactly when the last step that ran closed it and
// was not released by it; the first step of a launch reads the gate itself.
// Blocks beyond a job's share of the grid (vg) follow the steps without work.
// Every wait is bounded (a wait that runs out ends the grid and reports reason
// -1), so the grid always ends.
constexpr unsigned long long kEngineWaitFactor = 8;  // a block's wait for the others: 8 idle periods

__device__ __forceinline__ void copy_words(unsigned long long *dst, const unsigned long long *src, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_step_engine(DevBox b, const EngineJob *ring, EngineCtl *ctl,
                                                        unsigned long long first, unsigned long long idle_ticks,
                                                        unsigned long long launch_no) {
    __shared__ EngineJob job;
    __shared__ int s_cmd, s_run;
    constexpr int kJobWords = static_cast<int>(offsetof(EngineJob, seq) / sizeof(unsigned long long));
    static_assert(offsetof(EngineJob, seq) % sizeof(unsigned long long) == 0, "job body in 8-byte words");
    SearchState *s = b.ss;
    unsigned long long last = 0;      // the last job that ran in this launch (0: none yet)
    unsigned long long last_seq = 0;  // its step's sequence number
    unsigned long long busy = 0, idle = 0, steps = 0;
    unsigned long long k = first;
    int reason = kEngineExit;
    for (;; ++k) {
        const int slot = static_cast<int>(k % kEngineRing);
        if (threadIdx.x == 0) {
            const unsigned long long t0 = wall_clock64();
            int cmd = -1;
            if (blockIdx.x == 0) {
                cmd = kEngineQuit;
                while (wall_clock64() - t0 <= idle_ticks) {
                    if (__hip_atomic_load(&ring[slot].seq, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) == k) {
                        cmd = ring[slot].cmd == kEngineExit ? kEngineExit : kEngineRun;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                idle += wall_clock64() - t0;
            } else {
                const unsigned long long limit = kEngineWaitFactor * idle_ticks;
                while (wall_clock64() - t0 <= limit) {
                    const unsigned long long d =
                        __hip_atomic_load(&s->eng_decision, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                    if ((d >> 2) > k) {  // block 0 has gone on: job k ran
                        cmd = kEngineRun;
                        break;
                    }
                    if ((d >> 2) == k && (d & 3ull) != 0ull) {
                        cmd = static_cast<int>(d & 3ull);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            s_cmd = cmd;
        }
        __syncthreads();
        const int cmd = s_cmd;
        if (blockIdx.x == 0) {
            if (cmd == kEngineRun) {
                copy_words(reinterpret_cast<unsigned long long *>(&job),
                           reinterpret_cast<const unsigned long long *>(&ring[slot]), kJobWords);
                __syncthreads();
                copy_words(reinterpret_cast<unsigned long long *>(&b.eng_jobs[slot]),
                           reinterpret_cast<const unsigned long long *>(&job), kJobWords);
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                __threadfence();  // the job copy before the decision
                __hip_atomic_store(&s->eng_decision, k * 4ull + static_cast<unsigned long long>(cmd),
                                   __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else if (cmd == kEngineRun) {
            copy_words(reinterpret_cast<unsigned long long *>(&job),
                       reinterpret_cast<const unsigned long long *>(&b.eng_jobs[slot]), kJobWords);
            __syncthreads();
        }
        if (cmd != kEngineRun) {
            reason = cmd;
            break;
        }
        const unsigned vg = job.vg < 1 ? 1u : (job.vg > static_cast<int>(gridDim.x) ? gridDim.x : static_cast<unsigned>(job.vg));
        const bool works = blockIdx.x < vg;
        const SearchArgs a = job.a;  // in registers for the step, as a launch's kernel arguments are
        StepPrefetch pre{};
        if (works) pre = prefetch_step(b, a);  // in flight during the wait below
        unsigned long long t_start = 0;
        if (threadIdx.x == 0) {
            int run = 1;
            if (last == 0) {
                run = behind_gate(b, a) ? 0 : 1;
            } else {
                const unsigned long long t0 = wall_clock64(), limit = kEngineWaitFactor * idle_ticks;
                unsigned long long w = 0;
                run = -1;
                while (wall_clock64() - t0 <= limit) {
                    w = __hip_atomic_load(&s->eng_done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                    if ((w >> 1) >= last) {
                        run = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                if (run == 1 && (w >> 1) > last) {
                    // later steps ran without this block: the word of `last` is in its slot
                    w = __hip_atomic_load(&b.eng_closed[last % kEngineRing], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                    if ((w >> 1) != last) run = -1;
                }
                if (run == 1) {
                    // behind_gate with the gate closed by `last` (or open)
                    const bool closed = (w & 1ull) != 0ull;
                    run = (a.gated && closed && last_seq > a.released) ? 0 : 1;
                }
            }
            t_start = wall_clock64();
            s_run = run;
        }
        __syncthreads();
        const int run = s_run;
        if (run < 0) {
            reason = -1;
            break;
        }
        if (!run) continue;  // behind the gate: no block works on it, no record
        last = k;
        last_seq = a.seq;
        if (!works) continue;
        const CellTable *ct = job.cells.n > 0 ? &job.cells : nullptr;
        int role = 0;
        __shared__ int s_role;
        const bool closed = step_body(b, a, blockIdx.x, vg, true, ct, false, &s_role, &pre);
        __syncthreads();
        role = s_role;
        if (role == 2 && threadIdx.x == 0) {
            // the step's device writes, then eng_done for the other blocks,
            // then the record for the host
            const unsigned long long word = (k << 1) | (closed ? 1ull : 0ull);
            b.eng_closed[slot] = word;
            __threadfence();
            __hip_atomic_store(&s->eng_done, word, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence_system();
            __hip_atomic_store(&a.out->seq, a.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            busy += wall_clock64() - t_start;
            steps += 1;
        }
    }
    if (threadIdx.x == 0) {
        if (steps) {
            atomicAdd(&s->eng_steps, steps);
            atomicAdd(&s->eng_busy, busy);
        }
        if (blockIdx.x == 0) atomicAdd(&s->eng_idle, idle);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->stopped_at = k;
        ctl->reason = reason;
        __threadfence_system();
        __hip_atomic_store(&ctl->launches_ended, launch_no, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    } else if (reason == -1 && threadIdx.x == 0) {
        ctl->reason = -1;
        __threadfence_system();
    }
}

// a13 for a large mover: sphere positions from the refpos of the last search.
__global__ __launch_bounds__(kBlock) void k_translate(DevBox b, int src, int gated, unsigned long long seq,
                                                      unsigned long long released) {
    const SearchState *s = b.ss;
    const unsigned long long g = s->gate;
    if (gated && g != 0ull && g < seq && g > released) return;  // the step was skipped
    const int n = b.acount[src];
    const int st = b.astart[src];
    const double r0 = s->refpos[0], r1 = s->refpos[1], r2 = s->refpos[2];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        int sp = b.members[st + t];
        b.sx[sp] = r0 + b.srx[sp];
        b.sy[sp] = r1 + b.sry[sp];
        b.sz[sp] = r2 + b.srz[sp];
    }
}

// Aggregate::translate by a given vector: new position, cell and refpos.
__global__ void k_set_move(DevBox b, int src, double v0, double v1, double v2, int n_div, double L) {
    SearchState *s = b.ss;
    const double v[3] = {v0, v1, v2};
    const double old[3] = {b.ax[src], b.ay[src], b.az[src]};
    const double rc3[3] = {b.arx[src], b.ary[src], b.arz[src]};
    for (int l = 0; l < 3; ++l) {
        s->newcom[l] = periodic_position(old[l] + v[l], L);
        s->refpos[l] = s->newcom[l] - rc3[l];
    }
    s->newcell = ref_cell(s->newcom[0], s->newcom[1], s->newcom[2], n_div, L);
    b.ax[src] = s->newcom[0];
    b.ay[src] = s->newcom[1];
    b.az[src] = s->newcom[2];
    b.acell[src] = s->newcell;
    mark_dirty(b, src);
    set_dirty_entry(b, b.dslot[src], s->newcom[0], s->newcom[1], s->newcom[2], b.armax[src], s->newcell);
}

__global__ void k_swept_batch(long long n, const double *p1, const double *r1, const double *p2, const double *r2,
                              const double *dir, const double *dist, double L, double *out) {
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        out[i] = swept_sphere(p1[3 * i], p1[3 * i + 1], p1[3 * i + 2], r1[i], p2[3 * i], p2[3 * i + 1], p2[3 * i + 2],
                              r2[i], dir[3 * i], dir[3 * i + 1], dir[3 * i + 2], dist[i], L);
}

}  // namespace mcac_dev
