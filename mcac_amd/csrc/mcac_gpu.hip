// MI355X (gfx950) implementation of the MCAC aggregate-move / collision-search
// path behind the C-ABI declared in include/mcac_gpu.h.
//
// HBM layout (one box):
//   spheres     SoA, index order: x y z r vol surf rx ry rz          (72 B/sphere)
//   aggregates  SoA, stable-id order: x y z rx ry rz rmax volume surface rg,
//               merge caches (sum vol*rel xyz, sum r, sum vol*r^2), start,
//               count, packed reference Verlet cell (-1 = dead)
//   members     pool of sphere indices; aggregate a owns [start, start+cap)
//   alive       ascending live ids (label order), rebuilt lazily
// A contact search is three launches on the box stream with no host round trip
// in between: k_filter (Verlet window + bounding sweep, one thread per live
// aggregate, ticketed last block plans the exact solve), k_solve (sphere pairs
// in fixed-size chunks spread over all CUs, ticketed last block orders the
// candidates like the reference multimap and runs its early-exit scan) and
// k_translate.  The host reads back one small result record per step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/mcac_gpu.h"
#include "device_math.hpp"

using namespace mcac_dev;

namespace {

constexpr int kBlock = 256;
constexpr int kSolveBlocks = 256;       // one block per CU for the pair solve
constexpr int64_t kChunk = 4096;        // minimum sphere pairs per solve chunk
constexpr int kOverlapSlots = 16;       // contacts kept per sphere row (kissing number < 16 on this path)

struct SearchState {
    // filter
    unsigned int nfilt;
    unsigned int overflow;
    unsigned int done_filter;
    unsigned int done_solve;
    unsigned long long nquery;
    unsigned long long npairs;
    unsigned long long nchunks;
    unsigned long long chunk_size;
    // result
    double dist, move;
    double newcom[3];
    double refpos[3];
    long long ms, os, oagg;
    int contact, newcell, status, pad;
};

struct MergeState {
    int ok, keep, rem, n_k, n_r, new_start, start_k, start_r;
    double diffpos[3], refpos[3], r[3];
    unsigned long long rmax_bits;
    unsigned int row_overflow, pad;
    // result mirrored to the host
    mcac_merge_result res;
};

struct DevBox {
    double *sx, *sy, *sz, *sr, *svol, *ssurf, *srx, *sry, *srz;
    double *ax, *ay, *az, *arx, *ary, *arz, *armax, *avol, *asurf, *arg;
    double *csvx, *csvy, *csvz, *csr, *csvr2;
    int *astart, *acount, *acell;
    int *members;
    int *alive;
    // search scratch
    double *cand_db, *cand_exact;
    int *cand_rank, *cand_id, *order;
    long long *cand_pair, *cand_chunk0;
    double *part_d;
    long long *part_p;
    // merge scratch
    double *dc;
    int *row_cnt, *row_j;
    double *row_c;
    SearchState *ss;
    MergeState *ms;
};

struct SearchArgs {
    int src;
    int n_alive;
    int n_div;
    int cap;
    long long part_cap;
    double dir[3];
    double lpm;
    double maxradius;
    double L;
    int apply;
    int precision;
};

// ---------------------------------------------------------------------------------------------
// wave / block reductions (wave64)
__device__ __forceinline__ void lexmin(double &d, long long &p, double od, long long op) {
    if (od < d || (od == d && op < p)) {
        d = od;
        p = op;
    }
}
__device__ __forceinline__ void block_lexmin(double &d, long long &p) {
    __shared__ double sd[kBlock / 64];
    __shared__ long long sp[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        double od = __shfl_down(d, off, 64);
        long long op = __shfl_down(p, off, 64);
        lexmin(d, p, od, op);
    }
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        sd[w] = d;
        sp[w] = p;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kBlock / 64; ++i) lexmin(d, p, sd[i], sp[i]);
    }
    __syncthreads();
}
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v) {
    __shared__ unsigned long long s[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) s[w] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) t += s[i];
    __syncthreads();
    return t;
}

// aggregat.cpp:699-705 — reference Verlet cell of a position (clamped to the grid)
__device__ __forceinline__ int ref_cell(double x, double y, double z, int n, double L) {
    double step = static_cast<double>(n) / L;
    int c[3];
    double p[3] = {x, y, z};
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        double f = __builtin_floor(p[l] * step);
        int v = static_cast<int>(f);
        c[l] = v < 0 ? 0 : (v >= n ? n - 1 : v);
    }
    return (c[0] * n + c[1]) * n + c[2];
}

// verlet.cpp:57-78 — visited window (first cell and length per axis)
__device__ __forceinline__ void verlet_window(const double src[3], const double vec[3], double mindist, int n,
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

// ---------------------------------------------------------------------------------------------
// a4 + a6: Verlet window membership and bounding-sphere sweep for every live aggregate.
// Last block: plan of the exact solve (pair chunks per filtered candidate).
__global__ __launch_bounds__(kBlock) void k_filter(DevBox b, SearchArgs a) {
    __shared__ int win[6];
    __shared__ double q[7];
    const int n = a.n_div;
    if (threadIdx.x == 0) {
        double src[3] = {b.ax[a.src], b.ay[a.src], b.az[a.src]};
        double vec[3] = {a.lpm * a.dir[0], a.lpm * a.dir[1], a.lpm * a.dir[2]};
        double mindist = b.armax[a.src] + a.maxradius;
        int b1[3], len[3];
        verlet_window(src, vec, mindist, n, a.L, b1, len);
        for (int l = 0; l < 3; ++l) {
            win[l] = b1[l];
            win[3 + l] = len[l];
            q[l] = src[l];
        }
        q[3] = b.armax[a.src];
    }
    __syncthreads();
    const int b1i = win[0], b1j = win[1], b1k = win[2];
    const int li = win[3], lj = win[4], lk = win[5];
    const double sx = q[0], sy = q[1], sz = q[2], sr = q[3];
    unsigned long long inwin = 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < a.n_alive; t += gridDim.x * blockDim.x) {
        int id = b.alive[t];
        int c = b.acell[id];
        if (c < 0 || id == a.src) continue;
        int ci = c / (n * n), cj = (c / n) % n, ck = c % n;
        int oi = ((ci - b1i) % n + n) % n;
        int oj = ((cj - b1j) % n + n) % n;
        int ok = ((ck - b1k) % n + n) % n;
        if (oi >= li || oj >= lj || ok >= lk) continue;
        ++inwin;
        double d = swept_sphere(sx, sy, sz, sr, b.ax[id], b.ay[id], b.az[id], b.armax[id], a.dir[0], a.dir[1],
                                a.dir[2], a.lpm, a.L);
        if (d < a.lpm) {
            unsigned int slot = atomicAdd(&b.ss->nfilt, 1u);
            if (slot < static_cast<unsigned int>(a.cap)) {
                b.cand_db[slot] = d;
                b.cand_rank[slot] = (oi * lj + oj) * lk + ok;
                b.cand_id[slot] = id;
            } else {
                atomicOr(&b.ss->overflow, 1u);
            }
        }
    }
    unsigned long long tot = block_sum_u64(inwin);
    __shared__ bool last;
    if (threadIdx.x == 0) {
        if (tot) atomicAdd(&b.ss->nquery, tot);
        __threadfence();
        unsigned int ticket = atomicAdd(&b.ss->done_filter, 1u);
        last = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    // plan: chunks of the pair space of each candidate
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned int F = atomicAdd(&b.ss->nfilt, 0u);
        if (F > static_cast<unsigned int>(a.cap)) F = static_cast<unsigned int>(a.cap);
        long long nsrc = b.acount[a.src];
        unsigned long long total = 0;
        for (unsigned int c = 0; c < F; ++c) total += static_cast<unsigned long long>(nsrc) * b.acount[b.cand_id[c]];
        unsigned long long chunk = kChunk;
        if (total / chunk + F > static_cast<unsigned long long>(a.part_cap))
            chunk = total / (static_cast<unsigned long long>(a.part_cap) - F) + 1;
        unsigned long long acc = 0;
        for (unsigned int c = 0; c < F; ++c) {
            unsigned long long P = static_cast<unsigned long long>(nsrc) * b.acount[b.cand_id[c]];
            b.cand_chunk0[c] = static_cast<long long>(acc);
            acc += (P + chunk - 1) / chunk;
        }
        b.cand_chunk0[F] = static_cast<long long>(acc);
        b.ss->nchunks = acc;
        b.ss->chunk_size = chunk;
        b.ss->npairs = total;
        b.ss->done_filter = 0;
    }
}

// a8 + a9 + a7: exact sphere-pair sweep per candidate chunk, then (last block)
// reference candidate order and early-exit scan, and the move.
__global__ __launch_bounds__(kBlock) void k_solve(DevBox b, SearchArgs a) {
    unsigned int F = b.ss->nfilt;
    if (F > static_cast<unsigned int>(a.cap)) F = static_cast<unsigned int>(a.cap);
    const unsigned long long nchunks = b.ss->nchunks;
    const unsigned long long chunk = b.ss->chunk_size;
    const int ssrc = b.astart[a.src];
    for (unsigned long long g = blockIdx.x; g < nchunks; g += gridDim.x) {
        // candidate owning chunk g (binary search over cand_chunk0)
        unsigned int lo = 0, hi = F;
        while (hi - lo > 1) {
            unsigned int mid = (lo + hi) >> 1;
            if (static_cast<unsigned long long>(b.cand_chunk0[mid]) <= g) lo = mid;
            else hi = mid;
        }
        const int c = static_cast<int>(lo);
        const int k = b.cand_id[c];
        const long long nk = b.acount[k];
        const long long P = static_cast<long long>(b.acount[a.src]) * nk;
        const long long p0 = static_cast<long long>((g - b.cand_chunk0[c]) * chunk);
        const long long p1 = (p0 + static_cast<long long>(chunk) < P) ? p0 + static_cast<long long>(chunk) : P;
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
    __shared__ bool last;
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned int ticket = atomicAdd(&b.ss->done_solve, 1u);
        last = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    // per-candidate reduction of chunk partials (lexicographic (d, i, j) minimum)
    for (unsigned int c = threadIdx.x; c < F; c += blockDim.x) {
        double bd = kInf;
        long long bp = 0x7fffffffffffffffLL;
        for (long long g = b.cand_chunk0[c]; g < b.cand_chunk0[c + 1]; ++g) lexmin(bd, bp, b.part_d[g], b.part_p[g]);
        b.cand_exact[c] = bd;
        b.cand_pair[c] = (bd < kInf) ? bp : -1;
    }
    // multimap order: (bounding distance, Verlet visit rank, label)
    for (unsigned int c = threadIdx.x; c < F; c += blockDim.x) {
        double dc = b.cand_db[c];
        int rc = b.cand_rank[c], ic = b.cand_id[c];
        unsigned int pos = 0;
        for (unsigned int e = 0; e < F; ++e) {
            double de = b.cand_db[e];
            int re = b.cand_rank[e];
            bool less = de < dc || (de == dc && (re < rc || (re == rc && b.cand_id[e] < ic)));
            pos += less ? 1u : 0u;
        }
        b.order[pos] = static_cast<int>(c);
    }
    __syncthreads();
    __threadfence_block();
    if (threadIdx.x == 0) {
        // aggregat_list.cpp:459-482
        double closest = kInf;
        int best = -1;
        for (unsigned int pos = 0; pos < F; ++pos) {
            int c = b.order[pos];
            if (closest <= 0.) break;
            if (closest < b.cand_db[c]) break;
            if (b.cand_exact[c] < closest) {
                closest = b.cand_exact[c];
                best = c;
            }
        }
        SearchState *s = b.ss;
        s->dist = closest;
        s->ms = s->os = s->oagg = -1;
        if (best >= 0) {
            long long p = b.cand_pair[best];
            int k = b.cand_id[best];
            long long nk = b.acount[k];
            long long i = p / nk, j = p - (p / nk) * nk;
            s->ms = b.members[ssrc + i];
            s->os = b.members[b.astart[k] + j];
            s->oagg = k;
        }
        // calcul.cpp:128-144 and aggregat.cpp:148-161
        int contact = closest <= a.lpm;
        double move = contact ? closest : a.lpm;
        s->contact = contact;
        s->move = move;
        double old[3] = {b.ax[a.src], b.ay[a.src], b.az[a.src]};
        double rc3[3] = {b.arx[a.src], b.ary[a.src], b.arz[a.src]};
        for (int l = 0; l < 3; ++l) {
            double v = a.dir[l] * move;
            s->newcom[l] = periodic_position(old[l] + v, a.L);
            s->refpos[l] = s->newcom[l] - rc3[l];
        }
        s->newcell = ref_cell(s->newcom[0], s->newcom[1], s->newcom[2], a.n_div, a.L);
        s->status = (b.ss->overflow != 0) ? MCAC_E_CAPACITY : 0;
        b.ss->done_solve = 0;
    }
}

// a13: translate the moving aggregate by the move of the last search.
__global__ __launch_bounds__(kBlock) void k_translate(DevBox b, int src) {
    const SearchState *s = b.ss;
    const int n = b.acount[src];
    const int st = b.astart[src];
    const double r0 = s->refpos[0], r1 = s->refpos[1], r2 = s->refpos[2];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        int sp = b.members[st + t];
        b.sx[sp] = r0 + b.srx[sp];
        b.sy[sp] = r1 + b.sry[sp];
        b.sz[sp] = r2 + b.srz[sp];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        b.ax[src] = s->newcom[0];
        b.ay[src] = s->newcom[1];
        b.az[src] = s->newcom[2];
        b.acell[src] = s->newcell;
    }
}

// Explicit translate by a host vector (Aggregate::translate with a given vector).
__global__ void k_set_move(DevBox b, int src, double v0, double v1, double v2, int n_div, double L) {
    SearchState *s = b.ss;
    double v[3] = {v0, v1, v2};
    double old[3] = {b.ax[src], b.ay[src], b.az[src]};
    double rc3[3] = {b.arx[src], b.ary[src], b.arz[src]};
    for (int l = 0; l < 3; ++l) {
        s->newcom[l] = periodic_position(old[l] + v[l], L);
        s->refpos[l] = s->newcom[l] - rc3[l];
    }
    s->newcell = ref_cell(s->newcom[0], s->newcom[1], s->newcom[2], n_div, L);
}

// ---------------------------------------------------------------------------------------------
// a10: merge preparation — contact re-check, keeper/removed, relative shift.
struct MergeArgs {
    long long ms, os, magg, oagg;
    int keep, rem, n_k, n_r, new_start, start_k, start_r, n_div;
    double L;
};
__global__ void k_merge_prep(DevBox b, MergeArgs a) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    MergeState *m = b.ms;
    m->ok = 0;
    m->rmax_bits = 0ull;
    m->row_overflow = 0u;
    m->res.merged = 0;
    m->res.status = 0;
    const int ms = static_cast<int>(a.ms), os = static_cast<int>(a.os);
    // aggregat_list.cpp:375-378
    if (!contact(b.sx[ms], b.sy[ms], b.sz[ms], b.sr[ms], b.sx[os], b.sy[os], b.sz[os], b.sr[os], a.L)) return;
    // aggregat.cpp:496-535
    int mysphere, othersphere;
    if (a.keep == a.magg) {
        mysphere = ms;
        othersphere = os;
    } else {
        mysphere = os;
        othersphere = ms;
    }
    const int s0 = b.members[a.start_k];
    double ref[3] = {b.sx[s0], b.sy[s0], b.sz[s0]};
    double mpos[3] = {b.sx[mysphere], b.sy[mysphere], b.sz[mysphere]};
    double opos[3] = {b.sx[othersphere], b.sy[othersphere], b.sz[othersphere]};
    double mrel[3] = {b.srx[mysphere], b.sry[mysphere], b.srz[mysphere]};
    double orel[3] = {b.srx[othersphere], b.sry[othersphere], b.srz[othersphere]};
    for (int l = 0; l < 3; ++l) {
        double dcontact = periodic_distance(opos[l] - mpos[l], a.L);
        m->diffpos[l] = (mrel[l] + dcontact) - orel[l];
        m->refpos[l] = ref[l];
    }
    m->keep = a.keep;
    m->rem = a.rem;
    m->n_k = a.n_k;
    m->n_r = a.n_r;
    m->new_start = a.new_start;
    m->start_k = a.start_k;
    m->start_r = a.start_r;
    m->ok = 1;
}

// Member lists of the merged aggregate (keeper's list then the removed one's)
// and the removed spheres' new relative and absolute positions.
__global__ __launch_bounds__(kBlock) void k_merge_lists(DevBox b) {
    const MergeState *m = b.ms;
    if (!m->ok) return;
    const int nk = m->n_k, nr = m->n_r;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nk + nr; t += gridDim.x * blockDim.x) {
        if (t < nk) {
            if (m->new_start != m->start_k) b.members[m->new_start + t] = b.members[m->start_k + t];
        } else {
            int s = b.members[m->start_r + (t - nk)];
            b.members[m->new_start + t] = s;
            double rx = b.srx[s] + m->diffpos[0];
            double ry = b.sry[s] + m->diffpos[1];
            double rz = b.srz[s] + m->diffpos[2];
            b.srx[s] = rx;
            b.sry[s] = ry;
            b.srz[s] = rz;
            b.sx[s] = rx + m->refpos[0];
            b.sy[s] = ry + m->refpos[1];
            b.sz[s] = rz + m->refpos[2];
        }
    }
}

// a11 (ordered sums): continue the keeper's list-order sums over the removed
// spheres; centre of mass; aggregate position and cell.  One block: a tile of
// spheres is staged in LDS by all threads, thread 0 accumulates in order.
__global__ __launch_bounds__(kBlock) void k_merge_seq(DevBox b, int n_div, double L) {
    MergeState *m = b.ms;
    if (!m->ok) return;
    __shared__ double tv[kBlock], ts[kBlock], tx[kBlock], ty[kBlock], tz[kBlock], tr[kBlock];
    const int k = m->keep, nk = m->n_k, nr = m->n_r, base = m->new_start + nk;
    double V = b.avol[k], S = b.asurf[k], vx = b.csvx[k], vy = b.csvy[k], vz = b.csvz[k];
    double sr = b.csr[k], svr2 = b.csvr2[k];
    for (int t0 = 0; t0 < nr; t0 += kBlock) {
        int cnt = (nr - t0 < kBlock) ? nr - t0 : kBlock;
        if (threadIdx.x < cnt) {
            int s = b.members[base + t0 + threadIdx.x];
            tv[threadIdx.x] = b.svol[s];
            ts[threadIdx.x] = b.ssurf[s];
            tx[threadIdx.x] = b.srx[s];
            ty[threadIdx.x] = b.sry[s];
            tz[threadIdx.x] = b.srz[s];
            tr[threadIdx.x] = b.sr[s];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = 0; i < cnt; ++i) {
                double v = tv[i];
                V = V + v;                       // aggregat.cpp:388
                S = S + ts[i];                   // aggregat.cpp:389
                vx += tx[i] * v;                 // aggregat.cpp:439
                vy += ty[i] * v;
                vz += tz[i] * v;
                sr += tr[i];                     // aggregat.cpp:259
                svr2 = svr2 + v * (tr[i] * tr[i]);  // aggregat.cpp:477
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    int rem = m->rem;
    if (V <= 0 || S <= 0) m->res.status = MCAC_VOL_SURF_ERROR;  // aggregat.cpp:427-429
    double r[3] = {vx / V, vy / V, vz / V};                      // aggregat.cpp:442
    const int s0 = b.members[m->new_start];
    double com[3];
    com[0] = periodic_position(b.sx[s0] + r[0], L);               // aggregat.cpp:448
    com[1] = periodic_position(b.sy[s0] + r[1], L);
    com[2] = periodic_position(b.sz[s0] + r[2], L);
    b.avol[k] = V;
    b.asurf[k] = S;
    b.csvx[k] = vx;
    b.csvy[k] = vy;
    b.csvz[k] = vz;
    b.csr[k] = sr;
    b.csvr2[k] = svr2;
    b.arx[k] = r[0];
    b.ary[k] = r[1];
    b.arz[k] = r[2];
    b.ax[k] = com[0];
    b.ay[k] = com[1];
    b.az[k] = com[2];
    b.acell[k] = ref_cell(com[0], com[1], com[2], n_div, L);
    b.astart[k] = m->new_start;
    b.acount[k] = nk + nr;
    b.acell[rem] = -1;
    b.acount[rem] = 0;
    for (int l = 0; l < 3; ++l) m->r[l] = r[l];
    mcac_merge_result &res = m->res;
    res.keeper = k;
    res.removed = rem;
    res.n = nk + nr;
    res.merged = 1;
    res.volume = V;
    res.surface = S;
    res.radius_sum = sr;
    for (int l = 0; l < 3; ++l) {
        res.com[l] = com[l];
        res.rcm[l] = r[l];
    }
}

// a11 (parallel part): distance of every sphere to the new centre of mass and rmax.
__global__ __launch_bounds__(kBlock) void k_merge_dc(DevBox b) {
    MergeState *m = b.ms;
    if (!m->ok) return;
    const int k = m->keep;
    const int n = m->n_k + m->n_r, st = m->new_start;
    const double r0 = m->r[0], r1 = m->r[1], r2 = m->r[2];
    double mx = 0.0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        int s = b.members[st + t];
        double d0 = b.srx[s] - r0, d1 = b.sry[s] - r1, d2 = b.srz[s] - r2;
        double dc = __builtin_sqrt(d0 * d0 + d1 * d1 + d2 * d2);  // aggregat.cpp:446
        b.dc[t] = dc;
        mx = smax(mx, b.sr[s] + dc);                               // aggregat.cpp:462
    }
    (void)k;
    // non-negative doubles order like their bit patterns
    for (int off = 32; off > 0; off >>= 1) mx = smax(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&m->rmax_bits, static_cast<unsigned long long>(__double_as_longlong(mx)));
}

// a11 (ordered sum): Rg from the list-order sum of vol*dc^2.
__global__ __launch_bounds__(kBlock) void k_merge_rg(DevBox b) {
    MergeState *m = b.ms;
    if (!m->ok) return;
    __shared__ double tv[kBlock], td[kBlock];
    const int k = m->keep, n = m->n_k + m->n_r, st = m->new_start;
    double arg = 0.;
    for (int t0 = 0; t0 < n; t0 += kBlock) {
        int cnt = (n - t0 < kBlock) ? n - t0 : kBlock;
        if (threadIdx.x < cnt) {
            int s = b.members[st + t0 + threadIdx.x];
            tv[threadIdx.x] = b.svol[s];
            td[threadIdx.x] = b.dc[t0 + threadIdx.x];
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < cnt; ++i) arg = arg + tv[i] * (td[i] * td[i]);  // aggregat.cpp:476
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double rmax = __longlong_as_double(static_cast<long long>(m->rmax_bits));
    double brg = b.csvr2[k];
    double V = b.avol[k];
    double rg = __builtin_sqrt(__builtin_fabs((arg + 3. / 5. * brg) / V));  // aggregat.cpp:481
    b.armax[k] = rmax;
    b.arg[k] = rg;
    m->res.rmax = rmax;
    m->res.rg = rg;
}

// a11 (pair pass): contacts i<j inside the merged aggregate (aggregat.cpp:736-759).
// Tiles of kBlock x kBlock spheres; the j tile is staged in LDS.
__global__ __launch_bounds__(kBlock) void k_overlap_pairs(DevBox b) {
    const MergeState *m = b.ms;
    if (!m->ok) return;
    const int n = m->n_k + m->n_r, st = m->new_start;
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (tj < ti) return;
    __shared__ double jx[kBlock], jy[kBlock], jz[kBlock], jr[kBlock];
    int jbase = tj * kBlock;
    int jt = jbase + threadIdx.x;
    if (jt < n) {
        int s = b.members[st + jt];
        jx[threadIdx.x] = b.srx[s];
        jy[threadIdx.x] = b.sry[s];
        jz[threadIdx.x] = b.srz[s];
        jr[threadIdx.x] = b.sr[s];
    }
    __syncthreads();
    int i = ti * kBlock + threadIdx.x;
    if (i >= n) return;
    int si = b.members[st + i];
    const double ix = b.srx[si], iy = b.sry[si], iz = b.srz[si], ir = b.sr[si];
    int jend = (jbase + kBlock < n) ? jbase + kBlock : n;
    for (int j = (jbase > i + 1 ? jbase : i + 1); j < jend; ++j) {
        int l = j - jbase;
        double d0 = ix - jx[l], d1 = iy - jy[l], d2 = iz - jz[l];
        double s2 = ir + jr[l];
        double q = d0 * d0 + d1 * d1 + d2 * d2;
        // cheap reject well outside the contact tolerance before the exact test
        if (q > 1.0001 * (s2 * s2)) continue;
        double dist = __builtin_sqrt(q);
        if (dist <= (1. + kCoordinationEpsilon) * s2) {
            double c = (s2 - dist) / s2;
            int slot = atomicAdd(&b.row_cnt[i], 1);
            if (slot < kOverlapSlots) {
                b.row_j[static_cast<long long>(i) * kOverlapSlots + slot] = j;
                b.row_c[static_cast<long long>(i) * kOverlapSlots + slot] = c;
            } else {
                atomicOr(&b.ms->row_overflow, 1u);
            }
        }
    }
}

// Ordered sum of the contacts in (i, j) order; clears the row counters.
__global__ __launch_bounds__(kBlock) void k_overlap_sum(DevBox b) {
    MergeState *m = b.ms;
    if (!m->ok) return;
    const int n = m->n_k + m->n_r;
    __shared__ int tcnt[kBlock];
    double overlapping = 0.0;
    long long inter = 0;
    for (int t0 = 0; t0 < n; t0 += kBlock) {
        int cnt = (n - t0 < kBlock) ? n - t0 : kBlock;
        int i = t0 + threadIdx.x;
        int rc = 0;
        if (threadIdx.x < cnt) {
            rc = b.row_cnt[i];
            if (rc > kOverlapSlots) rc = kOverlapSlots;
            // sort the row by j (insertion sort, at most kOverlapSlots entries)
            long long base = static_cast<long long>(i) * kOverlapSlots;
            for (int x = 1; x < rc; ++x) {
                int kj = b.row_j[base + x];
                double kc = b.row_c[base + x];
                int y = x - 1;
                while (y >= 0 && b.row_j[base + y] > kj) {
                    b.row_j[base + y + 1] = b.row_j[base + y];
                    b.row_c[base + y + 1] = b.row_c[base + y];
                    --y;
                }
                b.row_j[base + y + 1] = kj;
                b.row_c[base + y + 1] = kc;
            }
            b.row_cnt[i] = 0;
            tcnt[threadIdx.x] = rc;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int x = 0; x < cnt; ++x) {
                long long base = static_cast<long long>(t0 + x) * kOverlapSlots;
                for (int y = 0; y < tcnt[x]; ++y) {
                    overlapping += 2.0 * b.row_c[base + y];  // aggregat.cpp:755
                    inter += 2;
                }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double coord = 0.0;
    if (inter > 0) {
        overlapping /= static_cast<double>(inter);
        coord = static_cast<double>(inter) / static_cast<double>(n);
    }
    m->res.overlapping = overlapping;
    m->res.coordination = coord;
    if (m->row_overflow) m->res.status = MCAC_E_CAPACITY;
}

__global__ void k_gather_members(DevBox b, int agg, int *out, int n) {
    int st = b.astart[agg];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) out[t] = b.members[st + t];
}

// Pool compaction: copy each live aggregate's members to its new start.
__global__ void k_compact(const int *old_members, int *new_members, const int *old_start, const int *new_start,
                          const int *count, const int *ids, int n_ids) {
    for (int a = blockIdx.x; a < n_ids; a += gridDim.x) {
        int id = ids[a];
        int n = count[id], so = old_start[id], sn = new_start[a];
        for (int t = threadIdx.x; t < n; t += blockDim.x) new_members[sn + t] = old_members[so + t];
    }
}

__global__ void k_swept_batch(long long n, const double *p1, const double *r1, const double *p2, const double *r2,
                              const double *dir, const double *dist, double L, double *out) {
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        out[i] = swept_sphere(p1[3 * i], p1[3 * i + 1], p1[3 * i + 2], r1[i], p2[3 * i], p2[3 * i + 1], p2[3 * i + 2],
                              r2[i], dir[3 * i], dir[3 * i + 1], dir[3 * i + 2], dist[i], L);
}

}  // namespace

// =============================================================================================
// Host side of the box.
struct mcac_box {
    mcac_box_params p{};
    hipStream_t stream = nullptr;
    DevBox d{};
    int64_t n_spheres = 0, n_aggs = 0;
    int64_t cand_cap = 0, part_cap = 0, pool_cap = 0;
    // host mirrors for allocation and launch sizing
    std::vector<int> count, start, cap;
    std::vector<int> alive_ids;     // ascending, may hold dead ids until rebuilt
    int64_t dead_in_list = 0;
    int64_t pool_used = 0;
    SearchState *h_ss = nullptr;    // pinned
    MergeState *h_ms = nullptr;     // pinned
    std::vector<void *> allocs;
    bool timing = false;
    hipEvent_t ev[4]{};
    double t_search = 0, t_translate = 0, t_merge = 0, t_filter = 0, t_solve = 0;
    double n_search = 0, n_translate = 0, n_merge = 0;
    char last_error[256] = {0};
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

int rebuild_alive(mcac_box *b) {
    std::vector<int> ids;
    ids.reserve(b->alive_ids.size());
    for (int id : b->alive_ids)
        if (b->count[static_cast<size_t>(id)] > 0) ids.push_back(id);
    b->alive_ids.swap(ids);
    b->dead_in_list = 0;
    if (!b->alive_ids.empty())
        HIP_TRY(hipMemcpyAsync(b->d.alive, b->alive_ids.data(), b->alive_ids.size() * sizeof(int),
                               hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

int compact_pool(mcac_box *b, int64_t need) {
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
                           b->stream, b->d.members, nm, d_os, d_ns, b->d.acount, d_ids, static_cast<int>(ids.size()));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (size_t a = 0; a < ids.size(); ++a) {
        b->start[static_cast<size_t>(ids[a])] = nstart[a];
        b->cap[static_cast<size_t>(ids[a])] = static_cast<int>(std::max<int64_t>(4, 2 * static_cast<int64_t>(b->count[static_cast<size_t>(ids[a])])));
    }
    HIP_TRY(hipMemcpy(b->d.astart, b->start.data(), b->start.size() * sizeof(int), hipMemcpyHostToDevice));
    // replace the pool
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

int launch_search(mcac_box *b, int64_t agg, const double dir[3], double lpm, double maxradius, int apply) {
    if (agg < 0 || agg >= b->n_aggs || b->count[static_cast<size_t>(agg)] <= 0) return MCAC_E_ARGUMENT;
    if (b->dead_in_list > 1024 && b->dead_in_list * 4 > static_cast<int64_t>(b->alive_ids.size())) {
        int rc = rebuild_alive(b);
        if (rc) return rc;
    }
    SearchArgs a{};
    a.src = static_cast<int>(agg);
    a.n_alive = static_cast<int>(b->alive_ids.size());
    a.n_div = static_cast<int>(b->p.n_div);
    a.cap = static_cast<int>(b->cand_cap);
    a.part_cap = b->part_cap;
    a.dir[0] = dir[0];
    a.dir[1] = dir[1];
    a.dir[2] = dir[2];
    a.lpm = lpm;
    a.maxradius = maxradius;
    a.L = b->p.box_length;
    a.apply = apply;
    a.precision = b->p.precision;
    int fblocks = static_cast<int>(std::min<int64_t>(4096, (a.n_alive + kBlock - 1) / kBlock));
    if (fblocks < 1) fblocks = 1;
    if (b->timing) HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    hipLaunchKernelGGL(k_filter, dim3(fblocks), dim3(kBlock), 0, b->stream, b->d, a);
    if (b->timing) HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    hipLaunchKernelGGL(k_solve, dim3(kSolveBlocks), dim3(kBlock), 0, b->stream, b->d, a);
    if (b->timing) HIP_TRY(hipEventRecord(b->ev[2], b->stream));
    if (apply) {
        int n = b->count[static_cast<size_t>(agg)];
        int tb = std::min(1024, std::max(1, (n + kBlock - 1) / kBlock));
        hipLaunchKernelGGL(k_translate, dim3(tb), dim3(kBlock), 0, b->stream, b->d, static_cast<int>(agg));
    }
    if (b->timing) HIP_TRY(hipEventRecord(b->ev[3], b->stream));
    HIP_TRY(hipGetLastError());
    return 0;
}

int finish_search(mcac_box *b, int64_t agg, int apply, mcac_contact *out) {
    // read back the result record and reset the per-search counters
    HIP_TRY(hipMemcpyAsync(b->h_ss, b->d.ss, sizeof(SearchState), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipMemsetAsync(b->d.ss, 0, offsetof(SearchState, nchunks), b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    const SearchState &s = *b->h_ss;
    if (b->timing) {
        float t01 = 0, t12 = 0, t23 = 0;
        HIP_TRY(hipEventElapsedTime(&t01, b->ev[0], b->ev[1]));
        HIP_TRY(hipEventElapsedTime(&t12, b->ev[1], b->ev[2]));
        HIP_TRY(hipEventElapsedTime(&t23, b->ev[2], b->ev[3]));
        b->t_filter += t01;
        b->t_solve += t12;
        b->t_search += t01 + t12;
        b->n_search += 1;
        if (apply) {
            b->t_translate += t23;
            b->n_translate += 1;
        }
    }
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
    if (s.status) return s.status;
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

const char *mcac_version(void) { return "mcac_gpu 0.1 (gfx950, FP64 collision path)"; }

int mcac_box_create(const mcac_box_params *params, mcac_box **out) {
    if (!params || !out) return MCAC_E_ARGUMENT;
    if (params->box_length <= 0 || params->n_div < 1 || params->n_div > 1290 || params->max_spheres < 1 ||
        params->max_aggregates < 1 || params->max_spheres > INT32_MAX / 2 || params->max_aggregates > INT32_MAX / 2)
        return MCAC_E_ARGUMENT;
    auto *b = new (std::nothrow) mcac_box();
    if (!b) return MCAC_E_CAPACITY;
    b->p = *params;
    int rc = 0;
    auto fail = [&](int code) {
        mcac_box_destroy(b);
        return code;
    };
    if (hipSetDevice(params->device) != hipSuccess) return fail(MCAC_E_DEVICE);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(MCAC_E_DEVICE);
    const int64_t Ns = params->max_spheres, Na = params->max_aggregates;
    b->cand_cap = params->candidate_capacity > 0 ? params->candidate_capacity : std::max<int64_t>(4096, std::min<int64_t>(Na, 1 << 20));
    b->part_cap = std::max<int64_t>(b->cand_cap + 1024, 1 << 20);
    b->pool_cap = params->member_pool > 0 ? params->member_pool : 4 * Ns + 4096;
    DevBox &d = b->d;
#define ALLOC(ptr, n) \
    if ((rc = dalloc(b, &(ptr), (n))) != 0) return fail(rc)
    ALLOC(d.sx, Ns); ALLOC(d.sy, Ns); ALLOC(d.sz, Ns); ALLOC(d.sr, Ns); ALLOC(d.svol, Ns); ALLOC(d.ssurf, Ns);
    ALLOC(d.srx, Ns); ALLOC(d.sry, Ns); ALLOC(d.srz, Ns);
    ALLOC(d.ax, Na); ALLOC(d.ay, Na); ALLOC(d.az, Na); ALLOC(d.arx, Na); ALLOC(d.ary, Na); ALLOC(d.arz, Na);
    ALLOC(d.armax, Na); ALLOC(d.avol, Na); ALLOC(d.asurf, Na); ALLOC(d.arg, Na);
    ALLOC(d.csvx, Na); ALLOC(d.csvy, Na); ALLOC(d.csvz, Na); ALLOC(d.csr, Na); ALLOC(d.csvr2, Na);
    ALLOC(d.astart, Na); ALLOC(d.acount, Na); ALLOC(d.acell, Na);
    ALLOC(d.members, b->pool_cap);
    ALLOC(d.alive, Na);
    ALLOC(d.cand_db, b->cand_cap); ALLOC(d.cand_exact, b->cand_cap); ALLOC(d.cand_rank, b->cand_cap);
    ALLOC(d.cand_id, b->cand_cap); ALLOC(d.order, b->cand_cap); ALLOC(d.cand_pair, b->cand_cap);
    ALLOC(d.cand_chunk0, b->cand_cap + 1);
    ALLOC(d.part_d, b->part_cap); ALLOC(d.part_p, b->part_cap);
    ALLOC(d.dc, Ns);
    ALLOC(d.row_cnt, Ns); ALLOC(d.row_j, Ns * kOverlapSlots); ALLOC(d.row_c, Ns * kOverlapSlots);
    ALLOC(d.ss, 1); ALLOC(d.ms, 1);
#undef ALLOC
    if (hipHostMalloc(reinterpret_cast<void **>(&b->h_ss), sizeof(SearchState), hipHostMallocDefault) != hipSuccess)
        return fail(MCAC_E_DEVICE);
    if (hipHostMalloc(reinterpret_cast<void **>(&b->h_ms), sizeof(MergeState), hipHostMallocDefault) != hipSuccess)
        return fail(MCAC_E_DEVICE);
    for (auto &e : b->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(MCAC_E_DEVICE);
    *out = b;
    return 0;
}

int mcac_box_destroy(mcac_box *b) {
    if (!b) return 0;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (void *p : b->allocs) (void)hipFree(p);
    if (b->h_ss) (void)hipHostFree(b->h_ss);
    if (b->h_ms) (void)hipHostFree(b->h_ms);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int mcac_box_upload(mcac_box *b, int64_t ns, const double *x, const double *y, const double *z, const double *r,
                    const double *vol, const double *surf, const double *rx, const double *ry, const double *rz,
                    int64_t na, const double *pos3, const double *rcm3, const double *rmax, const double *volume,
                    const double *surface, const int64_t *count, const int64_t *members) {
    if (!b || ns < 0 || na < 0 || ns > b->p.max_spheres || na > b->p.max_aggregates) return MCAC_E_ARGUMENT;
    b->n_spheres = ns;
    b->n_aggs = na;
    const size_t S = static_cast<size_t>(ns), A = static_cast<size_t>(na);
    auto up = [&](double *dst, const double *src, size_t n) -> int {
        if (n) HIP_TRY(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice));
        return 0;
    };
    int rc = 0;
    if ((rc = up(b->d.sx, x, S)) || (rc = up(b->d.sy, y, S)) || (rc = up(b->d.sz, z, S)) || (rc = up(b->d.sr, r, S)) ||
        (rc = up(b->d.svol, vol, S)) || (rc = up(b->d.ssurf, surf, S)) || (rc = up(b->d.srx, rx, S)) ||
        (rc = up(b->d.sry, ry, S)) || (rc = up(b->d.srz, rz, S)))
        return rc;
    std::vector<double> hx(A), hy(A), hz(A), hrx(A), hry(A), hrz(A), c0(A), c1(A), c2(A), csr(A), cvr2(A);
    std::vector<int> cell(A, -1);
    b->count.assign(A, 0);
    b->start.assign(A, 0);
    b->cap.assign(A, 0);
    b->alive_ids.clear();
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
        int64_t c = count[a];
        if (c < 0 || off + c > ns) return MCAC_E_ARGUMENT;
        b->count[a] = static_cast<int>(c);
        b->start[a] = static_cast<int>(pool.size());
        b->cap[a] = static_cast<int>(c);
        // list-order sums the merge continues from (aggregat.cpp:437-441, 257-260, 474-478)
        double svx = 0., svy = 0., svz = 0., sr = 0., svr2 = 0.;
        for (int64_t i = 0; i < c; ++i) {
            int64_t s = members[off + i];
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
            b->alive_ids.push_back(static_cast<int>(a));
            double step = static_cast<double>(n) / L;
            int ci[3];
            double p[3] = {hx[a], hy[a], hz[a]};
            for (int l = 0; l < 3; ++l) {
                int v = static_cast<int>(std::floor(p[l] * step));
                ci[l] = v < 0 ? 0 : (v >= n ? n - 1 : v);
            }
            cell[a] = (ci[0] * n + ci[1]) * n + ci[2];
        }
    }
    if (static_cast<int64_t>(pool.size()) > b->pool_cap) return MCAC_E_CAPACITY;
    b->pool_used = static_cast<int64_t>(pool.size());
    if ((rc = up(b->d.ax, hx.data(), A)) || (rc = up(b->d.ay, hy.data(), A)) || (rc = up(b->d.az, hz.data(), A)) ||
        (rc = up(b->d.arx, hrx.data(), A)) || (rc = up(b->d.ary, hry.data(), A)) || (rc = up(b->d.arz, hrz.data(), A)) ||
        (rc = up(b->d.armax, rmax, A)) || (rc = up(b->d.avol, volume, A)) || (rc = up(b->d.asurf, surface, A)) ||
        (rc = up(b->d.csvx, c0.data(), A)) || (rc = up(b->d.csvy, c1.data(), A)) || (rc = up(b->d.csvz, c2.data(), A)) ||
        (rc = up(b->d.csr, csr.data(), A)) || (rc = up(b->d.csvr2, cvr2.data(), A)))
        return rc;
    if (A) {
        HIP_TRY(hipMemcpy(b->d.astart, b->start.data(), A * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b->d.acount, b->count.data(), A * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b->d.acell, cell.data(), A * sizeof(int), hipMemcpyHostToDevice));
    }
    if (!pool.empty()) HIP_TRY(hipMemcpy(b->d.members, pool.data(), pool.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!b->alive_ids.empty())
        HIP_TRY(hipMemcpy(b->d.alive, b->alive_ids.data(), b->alive_ids.size() * sizeof(int), hipMemcpyHostToDevice));
    b->dead_in_list = 0;
    HIP_TRY(hipMemset(b->d.ss, 0, sizeof(SearchState)));
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}

int mcac_next_contact(mcac_box *b, int64_t agg, const double dir[3], double lpm, double maxradius,
                      mcac_contact *out) {
    if (!b || !dir) return MCAC_E_ARGUMENT;
    int rc = launch_search(b, agg, dir, lpm, maxradius, 0);
    if (rc) return rc;
    return finish_search(b, agg, 0, out);
}

int mcac_move(mcac_box *b, int64_t agg, const double dir[3], double lpm, double maxradius, mcac_contact *out) {
    if (!b || !dir) return MCAC_E_ARGUMENT;
    int rc = launch_search(b, agg, dir, lpm, maxradius, 1);
    if (rc) return rc;
    return finish_search(b, agg, 1, out);
}

int mcac_translate(mcac_box *b, int64_t agg, const double vec[3]) {
    if (!b || !vec || agg < 0 || agg >= b->n_aggs || b->count[static_cast<size_t>(agg)] <= 0) return MCAC_E_ARGUMENT;
    hipLaunchKernelGGL(k_set_move, dim3(1), dim3(1), 0, b->stream, b->d, static_cast<int>(agg), vec[0], vec[1], vec[2],
                       static_cast<int>(b->p.n_div), b->p.box_length);
    int n = b->count[static_cast<size_t>(agg)];
    int tb = std::min(1024, std::max(1, (n + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_translate, dim3(tb), dim3(kBlock), 0, b->stream, b->d, static_cast<int>(agg));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

int mcac_merge(mcac_box *b, const mcac_contact *c, mcac_merge_result *out) {
    if (!b || !c || !out) return MCAC_E_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    if (c->moving_sphere < 0 || c->other_sphere < 0 || c->moving_sphere >= b->n_spheres ||
        c->other_sphere >= b->n_spheres || c->moving_agg < 0 || c->other_agg < 0 || c->moving_agg >= b->n_aggs ||
        c->other_agg >= b->n_aggs || c->moving_agg == c->other_agg)
        return MCAC_E_ARGUMENT;
    const int keep = static_cast<int>(std::min(c->moving_agg, c->other_agg));
    const int rem = static_cast<int>(std::max(c->moving_agg, c->other_agg));
    const size_t K = static_cast<size_t>(keep), R = static_cast<size_t>(rem);
    if (b->count[K] <= 0 || b->count[R] <= 0) return MCAC_E_ARGUMENT;
    const int nk = b->count[K], nr = b->count[R];
    // member segment of the merged aggregate: append in place when the keeper's
    // segment has room, else a fresh segment with doubled capacity
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
    a.n_div = static_cast<int>(b->p.n_div);
    a.L = b->p.box_length;
    const int n = nk + nr;
    const int nb = std::min(4096, std::max(1, (n + kBlock - 1) / kBlock));
    if (b->timing) HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    hipLaunchKernelGGL(k_merge_prep, dim3(1), dim3(64), 0, b->stream, b->d, a);
    hipLaunchKernelGGL(k_merge_lists, dim3(nb), dim3(kBlock), 0, b->stream, b->d);
    hipLaunchKernelGGL(k_merge_seq, dim3(1), dim3(kBlock), 0, b->stream, b->d, a.n_div, a.L);
    hipLaunchKernelGGL(k_merge_dc, dim3(nb), dim3(kBlock), 0, b->stream, b->d);
    hipLaunchKernelGGL(k_merge_rg, dim3(1), dim3(kBlock), 0, b->stream, b->d);
    const unsigned tiles = static_cast<unsigned>((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_overlap_pairs, dim3(tiles, tiles), dim3(kBlock), 0, b->stream, b->d);
    hipLaunchKernelGGL(k_overlap_sum, dim3(1), dim3(kBlock), 0, b->stream, b->d);
    if (b->timing) HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(b->h_ms, b->d.ms, sizeof(MergeState), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->timing) {
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
        b->t_merge += t;
        b->n_merge += 1;
    }
    *out = b->h_ms->res;
    if (!b->h_ms->ok) {
        out->merged = 0;
        out->keeper = keep;
        out->removed = rem;
        return 0;
    }
    if (new_start != b->start[K]) {
        b->pool_used += new_cap;
        b->start[K] = new_start;
        b->cap[K] = new_cap;
    }
    b->count[K] = n;
    b->count[R] = 0;
    b->dead_in_list += 1;
    return out->status;
}

int mcac_box_download_spheres(mcac_box *b, double *x, double *y, double *z, double *rx, double *ry, double *rz) {
    if (!b) return MCAC_E_ARGUMENT;
    const size_t n = static_cast<size_t>(b->n_spheres) * sizeof(double);
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (x) HIP_TRY(hipMemcpy(x, b->d.sx, n, hipMemcpyDeviceToHost));
    if (y) HIP_TRY(hipMemcpy(y, b->d.sy, n, hipMemcpyDeviceToHost));
    if (z) HIP_TRY(hipMemcpy(z, b->d.sz, n, hipMemcpyDeviceToHost));
    if (rx) HIP_TRY(hipMemcpy(rx, b->d.srx, n, hipMemcpyDeviceToHost));
    if (ry) HIP_TRY(hipMemcpy(ry, b->d.sry, n, hipMemcpyDeviceToHost));
    if (rz) HIP_TRY(hipMemcpy(rz, b->d.srz, n, hipMemcpyDeviceToHost));
    return 0;
}

int mcac_box_download_aggs(mcac_box *b, double *pos3, double *rcm3, double *rmax, double *volume, double *surface,
                           double *rg, int64_t *count) {
    if (!b) return MCAC_E_ARGUMENT;
    const size_t A = static_cast<size_t>(b->n_aggs);
    HIP_TRY(hipStreamSynchronize(b->stream));
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
    hipLaunchKernelGGL(k_gather_members, dim3(std::min(1024, (cnt + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream,
                       b->d, static_cast<int>(agg), tmp, cnt);
    HIP_TRY(hipStreamSynchronize(b->stream));
    std::vector<int> h(static_cast<size_t>(cnt));
    HIP_TRY(hipMemcpy(h.data(), tmp, h.size() * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(hipFree(tmp));
    for (int64_t i = 0; i < std::min<int64_t>(max, cnt); ++i) out[i] = h[static_cast<size_t>(i)];
    return 0;
}

int mcac_box_set_timing(mcac_box *b, int enable) {
    if (!b) return MCAC_E_ARGUMENT;
    b->timing = enable != 0;
    b->t_search = b->t_translate = b->t_merge = b->t_filter = b->t_solve = 0;
    b->n_search = b->n_translate = b->n_merge = 0;
    return 0;
}

int mcac_box_timing(mcac_box *b, double out[8]) {
    if (!b || !out) return MCAC_E_ARGUMENT;
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
