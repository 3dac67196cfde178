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
// k_translate moves the spheres of a large mover.
#pragma once
#include "kernels_common.hpp"

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

// A gated step enqueued behind a step that found a contact does nothing.
__device__ __forceinline__ bool behind_gate(const DevBox &b, const SearchArgs &a) {
    if (!a.gated) return false;
    const unsigned long long g = b.ss->gate;
    return g != 0ull && g < a.seq;
}

__global__ __launch_bounds__(kBlock) void k_filter(DevBox b, SearchArgs a) {
    if (behind_gate(b, a)) return;
    __shared__ int win[6];
    __shared__ double q[4];
    __shared__ int pre[kMaxVisit + 1];
    __shared__ int nvisit;
    const int n = a.n_div;
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
    const int b1i = win[0], b1j = win[1], b1k = win[2];
    const int li = win[3], lj = win[4], lk = win[5];
    const double sx = q[0], sy = q[1], sz = q[2], sr = q[3];
    const int V = nvisit;
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int gstride = gridDim.x * blockDim.x;
    unsigned long long inwin = 0;
    if (V <= kMaxVisit) {
        // sizes of the visited cells in visit order, then their prefix
        for (int v = threadIdx.x; v < V; v += blockDim.x) {
            int oi = v / (lj * lk), rem = v - oi * (lj * lk), oj = rem / lk, ok = rem - oj * lk;
            int ci = window_offset(b1i + oi, 0, n), cj = window_offset(b1j + oj, 0, n),
                ck = window_offset(b1k + ok, 0, n);
            int c = (ci * n + cj) * n + ck;
            pre[v] = b.cstart[c + 1] - b.cstart[c];
        }
        __syncthreads();
        block_exclusive_scan(pre, V);
        const int total = pre[V];
        for (int w = gtid; w < total; w += gstride) {
            int lo = 0, hi = V;  // largest v with pre[v] <= w
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (pre[mid] <= w) lo = mid;
                else hi = mid;
            }
            const int v = lo;
            int oi = v / (lj * lk), rem = v - oi * (lj * lk), oj = rem / lk, ok = rem - oj * lk;
            int ci = window_offset(b1i + oi, 0, n), cj = window_offset(b1j + oj, 0, n),
                ck = window_offset(b1k + ok, 0, n);
            int c = (ci * n + cj) * n + ck;
            int id = b.citems[b.cstart[c] + (w - pre[v])];
            if (b.dirty[id] || id == a.src) continue;
            ++inwin;
            double d = swept_sphere(sx, sy, sz, sr, b.ax[id], b.ay[id], b.az[id], b.armax[id], a.dir[0], a.dir[1],
                                    a.dir[2], a.lpm, a.L);
            if (d < a.lpm) append_candidate(b, a, id, v, d);
        }
    } else {
        // very wide window: scan every built entry
        for (int w = gtid; w < a.n_built; w += gstride) {
            int id = b.citems[w];
            if (b.dirty[id] || id == a.src) continue;
            int c = b.acell[id];
            int oi = window_offset(c / (n * n), b1i, n), oj = window_offset((c / n) % n, b1j, n),
                ok = window_offset(c % n, b1k, n);
            if (oi >= li || oj >= lj || ok >= lk) continue;
            ++inwin;
            double d = swept_sphere(sx, sy, sz, sr, b.ax[id], b.ay[id], b.az[id], b.armax[id], a.dir[0], a.dir[1],
                                    a.dir[2], a.lpm, a.L);
            if (d < a.lpm) append_candidate(b, a, id, (oi * lj + oj) * lk + ok, d);
        }
    }
    // aggregates moved since the build, at their current cell
    const int nd = static_cast<int>(b.ss->dirty_count);
    for (int w = gtid; w < nd; w += gstride) {
        int id = b.dirty_list[w];
        int c = b.acell[id];
        if (c < 0 || id == a.src) continue;
        int oi = window_offset(c / (n * n), b1i, n), oj = window_offset((c / n) % n, b1j, n),
            ok = window_offset(c % n, b1k, n);
        if (oi >= li || oj >= lj || ok >= lk) continue;
        ++inwin;
        double d = swept_sphere(sx, sy, sz, sr, b.ax[id], b.ay[id], b.az[id], b.armax[id], a.dir[0], a.dir[1], a.dir[2],
                                a.lpm, a.L);
        if (d < a.lpm) append_candidate(b, a, id, (oi * lj + oj) * lk + ok, d);
    }
    unsigned long long tot = block_sum(inwin);
    if (threadIdx.x == 0 && tot) atomicAdd(&b.ss->nquery, tot);
}

__device__ __forceinline__ long long chunk_count(long long pairs, long long chunk) {
    return (pairs + chunk - 1) / chunk;
}

__global__ __launch_bounds__(kBlock) void k_solve(DevBox b, SearchArgs a) {
    if (behind_gate(b, a)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.out->skipped = 1;
            a.out->contact = 0;
            a.out->status = 0;
            __threadfence_system();
            __hip_atomic_store(&a.out->seq, a.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        return;
    }
    __shared__ long long cpre[kMaxPlan + 1];
    __shared__ long long s_chunk;
    __shared__ bool last;
    unsigned int F = b.ss->nfilt;
    if (F > static_cast<unsigned int>(a.cap)) F = static_cast<unsigned int>(a.cap);
    const long long nsrc = b.acount[a.src];
    const int ssrc = b.astart[a.src];
    // plan: chunks of the pair space of each candidate
    long long tpairs = 0;
    for (unsigned int c = threadIdx.x; c < F; c += blockDim.x) tpairs += nsrc * b.acount[b.cand_id[c]];
    tpairs = block_sum(tpairs);
    if (threadIdx.x == 0) {
        long long chunk = kChunk;
        if (tpairs / chunk + F > a.part_cap) chunk = tpairs / (a.part_cap - F) + 1;
        s_chunk = chunk;
    }
    __syncthreads();
    const long long chunk = s_chunk;
    const bool planned = F <= static_cast<unsigned int>(kMaxPlan);
    long long nchunks = 0;
    if (planned) {
        for (unsigned int c = threadIdx.x; c < F; c += blockDim.x)
            cpre[c] = chunk_count(nsrc * b.acount[b.cand_id[c]], chunk);
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
        const int k = b.cand_id[c];
        const long long nk = b.acount[k];
        const long long Ptot = nsrc * nk;
        const long long p0 = (g - c0) * chunk;
        const long long p1 = (p0 + chunk < Ptot) ? p0 + chunk : Ptot;
        const int sk = b.astart[k];
        double bd = kInf;
        long long bp = 0x7fffffffffffffffLL;
        for (long long p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
            long long i = p / nk;
            long long j = p - i * nk;
            int si = b.members[ssrc + i];
            int sj = b.members[sk + j];
            double d = swept_sphere(b.sx[si], b.sy[si], b.sz[si], b.sr[si], b.sx[sj], b.sy[sj], b.sz[sj], b.sr[sj],
                                    a.dir[0], a.dir[1], a.dir[2], a.lpm, a.L);
            lexmin(bd, bp, d, p);
        }
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
    // per-candidate reduction of the chunk partials
    for (unsigned int c = threadIdx.x; c < F; c += blockDim.x) {
        long long g0, g1;
        if (planned) {
            g0 = cpre[c];
            g1 = cpre[c + 1];
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
    // std::multimap order: (bounding distance, visit rank, label)
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
    __syncthreads();
    __shared__ double s_ref[3];
    SearchState *s = b.ss;
    if (threadIdx.x == 0) {
        double closest = kInf;
        int best = -1;
        for (unsigned int pos = 0; pos < F; ++pos) {
            const int c = b.order[pos];
            if (closest <= 0.) break;
            if (closest < b.cand_db[c]) break;
            if (b.cand_exact[c] < closest) {
                closest = b.cand_exact[c];
                best = c;
            }
        }
        long long ms = -1, os = -1, oagg = -1;
        if (best >= 0) {
            long long p = b.cand_pair[best];
            int k = b.cand_id[best];
            long long nk = b.acount[k];
            long long i = p / nk, j = p - i * nk;
            ms = b.members[ssrc + i];
            os = b.members[b.astart[k] + j];
            oagg = k;
        }
        // calcul.cpp:128-130 and aggregat.cpp:148-153
        const int contact = closest <= a.lpm;
        const double move = contact ? closest : a.lpm;
        const double old[3] = {b.ax[a.src], b.ay[a.src], b.az[a.src]};
        const double rc3[3] = {b.arx[a.src], b.ary[a.src], b.arz[a.src]};
        for (int l = 0; l < 3; ++l) {
            s->newcom[l] = periodic_position(old[l] + a.dir[l] * move, a.L);
            s->refpos[l] = s->newcom[l] - rc3[l];
            s_ref[l] = s->refpos[l];
        }
        s->newcell = ref_cell(s->newcom[0], s->newcom[1], s->newcom[2], a.n_div, a.L);
        StepRecord *o = a.out;
        o->dist = closest;
        o->move = move;
        o->ms = ms;
        o->os = os;
        o->oagg = oagg;
        o->contact = contact;
        o->status = s->overflow ? MCAC_E_CAPACITY : 0;
        o->nquery = s->nquery;
        o->nfilt = s->nfilt;
        unsigned long long tp = 0;
        for (unsigned int c = 0; c < F; ++c) tp += static_cast<unsigned long long>(nsrc) * b.acount[b.cand_id[c]];
        o->npairs = tp;
        if (a.apply) {
            b.ax[a.src] = s->newcom[0];
            b.ay[a.src] = s->newcom[1];
            b.az[a.src] = s->newcom[2];
            b.acell[a.src] = s->newcell;
            mark_dirty(b, a.src);
        }
        if (a.gated && contact) s->gate = a.seq;
        o->skipped = 0;
        o->dirty_count = static_cast<int>(s->dirty_count);
        s->nfilt = 0;
        s->nquery = 0;
        s->overflow = 0;
        s->done_solve = 0;
    }
    __syncthreads();
    if (a.apply && a.translate_in_finish) {
        const double r0 = s_ref[0], r1 = s_ref[1], r2 = s_ref[2];
        for (int t = threadIdx.x; t < nsrc; t += blockDim.x) {
            int sp = b.members[ssrc + t];
            b.sx[sp] = r0 + b.srx[sp];
            b.sy[sp] = r1 + b.sry[sp];
            b.sz[sp] = r2 + b.srz[sp];
        }
    }
    if (threadIdx.x == 0) {
        __threadfence_system();
        __hip_atomic_store(&a.out->seq, a.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ void k_release(DevBox b) { b.ss->gate = 0ull; }

// a13 for a large mover: sphere positions from the refpos of the last search.
__global__ __launch_bounds__(kBlock) void k_translate(DevBox b, int src, int gated, unsigned long long seq) {
    const SearchState *s = b.ss;
    if (gated && s->gate != 0ull && s->gate < seq) return;
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
}

__global__ void k_swept_batch(long long n, const double *p1, const double *r1, const double *p2, const double *r2,
                              const double *dir, const double *dist, double L, double *out) {
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        out[i] = swept_sphere(p1[3 * i], p1[3 * i + 1], p1[3 * i + 2], r1[i], p2[3 * i], p2[3 * i + 1], p2[3 * i + 2],
                              r2[i], dir[3 * i], dir[3 * i + 1], dir[3 * i + 2], dist[i], L);
}

}  // namespace mcac_dev
