// Merge of two aggregates and the post-merge update (a10 + a11):
//   AggregatList::merge     aggregat_list.cpp:367-410
//   Aggregate::merge        aggregat.cpp:486-544
//   Aggregate::update       aggregat.cpp:284-288 -> update_distances_and_overlapping
//                           (:719-764), compute_volume_surface (:321-430, method
//                           none), compute_mass_center / max_radius /
//                           giration_radius (:431-483)
// Every order-dependent FP64 sum runs sequentially in the reference's list
// order.  The merged list is the keeper's list followed by the removed one's,
// so the keeper's partial sums are its cached totals and only the removed
// spheres are added; Σ vol·dc² and the overlap sum need the whole list.
// Up to kSmallMerge spheres one fused block does everything; larger merges run
// the multi-kernel sequence below.  Both publish the result to pinned memory.
#pragma once
#include "kernels_common.hpp"
#include "arvo.hpp"
#include "volsurf.hpp"

namespace mcac_dev {

// aggregat_list.cpp:375-378 and aggregat.cpp:496-535 (single thread).
__device__ __forceinline__ bool merge_prepare(const DevBox &b, const MergeArgs &a, double diffpos[3],
                                              double refpos[3]) {
    const int ms = static_cast<int>(a.ms), os = static_cast<int>(a.os);
    if (!contact(b.sx[ms], b.sy[ms], b.sz[ms], b.sr[ms], b.sx[os], b.sy[os], b.sz[os], b.sr[os], a.L)) return false;
    int mysphere = ms, othersphere = os;
    if (a.keep != a.magg) {
        mysphere = os;
        othersphere = ms;
    }
    const int s0 = b.members[a.start_k];
    const double ref[3] = {b.sx[s0], b.sy[s0], b.sz[s0]};
    const double mpos[3] = {b.sx[mysphere], b.sy[mysphere], b.sz[mysphere]};
    const double opos[3] = {b.sx[othersphere], b.sy[othersphere], b.sz[othersphere]};
    const double mrel[3] = {b.srx[mysphere], b.sry[mysphere], b.srz[mysphere]};
    const double orel[3] = {b.srx[othersphere], b.sry[othersphere], b.srz[othersphere]};
    for (int l = 0; l < 3; ++l) {
        double dcontact = periodic_distance(opos[l] - mpos[l], a.L);
        diffpos[l] = (mrel[l] + dcontact) - orel[l];
        refpos[l] = ref[l];
    }
    return true;
}

// Relabel/remap one sphere of the removed aggregate (aggregat.cpp:524-535).
__device__ __forceinline__ void remap_sphere(const DevBox &b, int s, const double diffpos[3], const double refpos[3]) {
    double rx = b.srx[s] + diffpos[0];
    double ry = b.sry[s] + diffpos[1];
    double rz = b.srz[s] + diffpos[2];
    b.srx[s] = rx;
    b.sry[s] = ry;
    b.srz[s] = rz;
    b.sx[s] = rx + refpos[0];
    b.sy[s] = ry + refpos[1];
    b.sz[s] = rz + refpos[2];
}

// Stores the merged aggregate's geometry and marks keeper/removed (single thread).
__device__ __forceinline__ void store_merged(const DevBox &b, const MergeArgs &a, double V, double S, double vx, double vy,
                                             double vz, double sr, double svr2, const double r[3], const double com[3]) {
    const int k = a.keep;
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
    b.acell[k] = ref_cell(com[0], com[1], com[2], a.n_div, a.L);
    b.astart[k] = a.new_start;
    b.acap[k] = a.cap_k;
    b.acount[k] = a.n_k + a.n_r;
    b.acell[a.rem] = -1;
    b.acount[a.rem] = 0;
    mark_dirty(b, k);
    mark_dirty(b, a.rem);
    b.dpx[b.dslot[k]] = com[0];  // rmax follows once computed
    b.dpy[b.dslot[k]] = com[1];
    b.dpz[b.dslot[k]] = com[2];
    b.dpcell[b.dslot[k]] = b.acell[k];
    b.dpcell[b.dslot[a.rem]] = -1;
    b.cslot[a.rem] = -1;  // removed for good: no entry of a later build is its
}

__device__ __forceinline__ void publish_merge(const MergeArgs &a, const mcac_merge_result &res) {
    a.out->res = res;
    __threadfence_system();
    __hip_atomic_store(&a.out->seq, a.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ------------------------------------------------------------------ fused merge (n <= kSmallMerge)
// One block merges; thread 0 ends with the result in res_out.  Used by
// k_merge_small and by the finishing block of a step that found a contact.
// Global loads are issued in two batches (member indices with the keeper's
// caches and flags, then the sphere rows); everything after works from LDS,
// so the block synchronises through LDS only (no thread reads another
// thread's global stores here).  mover_ref (LDS, or null): the finishing
// block has just moved aggregate a.magg so that its spheres sit at
// mover_ref + relative position, and thread q placed its q-th sphere; the
// merge takes those positions from mover_ref instead of reading them back,
// and thread q also rewrites that sphere when the mover is the one removed.
__device__ __forceinline__ void merge_small_block(const DevBox &b, const MergeArgs &a, mcac_merge_result &res_out,
                                                  const double *mover_ref) {
    __shared__ double diffpos[3], refpos[3], r[3];
    __shared__ int ok;
    __shared__ int sidx[kSmallMerge];
    __shared__ double dcs[kSmallMerge], vols[kSmallMerge], rads[kSmallMerge], surfs[kSmallMerge];
    __shared__ double rlx[kSmallMerge], rly[kSmallMerge], rlz[kSmallMerge];
    __shared__ int nrow[kSmallMerge];
    __shared__ double crow[kSmallMerge * 8];
    __shared__ double s_V, s_svr2;
    __shared__ mcac_merge_result s_res;
    const int n = a.n_k + a.n_r;
    const int k = a.keep;
    const bool mover_keeps = a.keep == a.magg;
    // batch 1: member indices; thread 0 also takes the contact spheres, the
    // keeper's cached sums and the dirty-list state
    double cV = 0, cS = 0, cvx = 0, cvy = 0, cvz = 0, csr = 0, csvr2 = 0;
    int dk = 0, dr = 0, csk = -1, csr_slot = -1, dsk = 0, dsr = 0, s0 = 0;
    unsigned int dcount = 0;
    double m4[4] = {0, 0, 0, 0}, o4[4] = {0, 0, 0, 0}, mrel[3] = {0, 0, 0}, orel[3] = {0, 0, 0};
    for (int t = threadIdx.x; t < n; t += blockDim.x)
        sidx[t] = (t < a.n_k) ? b.members[a.start_k + t] : b.members[a.start_r + (t - a.n_k)];
    if (threadIdx.x == 0) {
        // the moving aggregate's contact sphere (ms) and the other's (os); "mine"
        // is the keeper's side (aggregat.cpp:496-509)
        const int ms = static_cast<int>(a.ms), os = static_cast<int>(a.os);
        double mv[4], ot[4], mvr[3], otr[3];
        mvr[0] = b.srx[ms], mvr[1] = b.sry[ms], mvr[2] = b.srz[ms];
        otr[0] = b.srx[os], otr[1] = b.sry[os], otr[2] = b.srz[os];
        mv[3] = b.sr[ms];
        ot[0] = b.sx[os], ot[1] = b.sy[os], ot[2] = b.sz[os], ot[3] = b.sr[os];
        if (mover_ref) {
            for (int l = 0; l < 3; ++l) mv[l] = mover_ref[l] + mvr[l];
        } else {
            mv[0] = b.sx[ms], mv[1] = b.sy[ms], mv[2] = b.sz[ms];
        }
        for (int l = 0; l < 4; ++l) {
            m4[l] = mover_keeps ? mv[l] : ot[l];
            o4[l] = mover_keeps ? ot[l] : mv[l];
        }
        for (int l = 0; l < 3; ++l) {
            mrel[l] = mover_keeps ? mvr[l] : otr[l];
            orel[l] = mover_keeps ? otr[l] : mvr[l];
        }
        cV = b.avol[k], cS = b.asurf[k], cvx = b.csvx[k], cvy = b.csvy[k], cvz = b.csvz[k];
        csr = b.csr[k], csvr2 = b.csvr2[k];
        dk = b.dirty[k], dr = b.dirty[a.rem];
        csk = b.cslot[k], csr_slot = b.cslot[a.rem];
        dsk = b.dslot[k], dsr = b.dslot[a.rem];
        dcount = b.ss->dirty_count;
        s0 = b.members[a.start_k];  // the keeper's first sphere
    }
    lds_barrier();
    // batch 2: the sphere rows (relative positions; absolute ones are not needed)
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int sp = sidx[t];
        vols[t] = b.svol[sp];
        surfs[t] = b.ssurf[sp];
        rads[t] = b.sr[sp];
        rlx[t] = b.srx[sp];
        rly[t] = b.sry[sp];
        rlz[t] = b.srz[sp];
    }
    double k0[3] = {0, 0, 0};
    if (threadIdx.x == 0 && !(mover_ref && mover_keeps)) k0[0] = b.sx[s0], k0[1] = b.sy[s0], k0[2] = b.sz[s0];
    if (threadIdx.x == 0) {
        // re-check of the contact and the removed spheres' shift
        // (aggregat_list.cpp:375-378, aggregat.cpp:496-535)
        // contact(moving sphere, other sphere) in the reference's argument order
        const double *p1 = mover_keeps ? m4 : o4, *p2 = mover_keeps ? o4 : m4;
        ok = contact(p1[0], p1[1], p1[2], p1[3], p2[0], p2[1], p2[2], p2[3], a.L) ? 1 : 0;
        const double mpos[3] = {m4[0], m4[1], m4[2]}, opos[3] = {o4[0], o4[1], o4[2]};
        for (int l = 0; l < 3; ++l) {
            const double dcontact = periodic_distance(opos[l] - mpos[l], a.L);
            diffpos[l] = (mrel[l] + dcontact) - orel[l];
        }
    }
    lds_barrier();
    if (!ok) {
        if (threadIdx.x == 0) {
            mcac_merge_result res{};
            res.keeper = a.keep;
            res.removed = a.rem;
            res.merged = 0;
            res_out = res;
        }
        return;
    }
    if (threadIdx.x == 0) {
        // the keeper's first sphere (rlx[0] is its relative position)
        if (mover_ref && mover_keeps) {
            refpos[0] = mover_ref[0] + rlx[0];
            refpos[1] = mover_ref[1] + rly[0];
            refpos[2] = mover_ref[2] + rlz[0];
        } else {
            refpos[0] = k0[0];
            refpos[1] = k0[1];
            refpos[2] = k0[2];
        }
    }
    lds_barrier();
    // member list, and the removed spheres moved into the keeper's frame (the
    // removed one's q-th sphere by thread q)
    if (a.new_start != a.start_k)
        for (int t = threadIdx.x; t < a.n_k; t += blockDim.x) b.members[a.new_start + t] = sidx[t];
    for (int q = threadIdx.x; q < a.n_r; q += blockDim.x) {
        const int t = a.n_k + q;
        const int sp = sidx[t];
        b.members[a.new_start + t] = sp;
        const double rx = rlx[t] + diffpos[0], ry = rly[t] + diffpos[1], rz = rlz[t] + diffpos[2];
        rlx[t] = rx;
        rly[t] = ry;
        rlz[t] = rz;
        b.srx[sp] = rx;
        b.sry[sp] = ry;
        b.srz[sp] = rz;
        b.sx[sp] = rx + refpos[0];
        b.sy[sp] = ry + refpos[1];
        b.sz[sp] = rz + refpos[2];
    }
    lds_barrier();
    if (threadIdx.x == 0) {
        double V = cV, S = cS, vx = cvx, vy = cvy, vz = cvz, sr = csr, svr2 = csvr2;
        for (int i = a.n_k; i < n; ++i) {
            const double v = vols[i];
            V = V + v;                              // aggregat.cpp:388
            vx += rlx[i] * v;                       // aggregat.cpp:439
            vy += rly[i] * v;
            vz += rlz[i] * v;
            sr += rads[i];                          // aggregat.cpp:259
            svr2 = svr2 + v * (rads[i] * rads[i]);  // aggregat.cpp:477
        }
        for (int i = a.n_k; i < n; ++i) S = S + surfs[i];  // aggregat.cpp:389
        double rr[3] = {vx / V, vy / V, vz / V};           // aggregat.cpp:442
        double com[3] = {periodic_position(refpos[0] + rr[0], a.L), periodic_position(refpos[1] + rr[1], a.L),
                         periodic_position(refpos[2] + rr[2], a.L)};  // aggregat.cpp:448
        const int cell = ref_cell(com[0], com[1], com[2], a.n_div, a.L);
        b.avol[k] = V;
        b.asurf[k] = S;
        b.csvx[k] = vx;
        b.csvy[k] = vy;
        b.csvz[k] = vz;
        b.csr[k] = sr;
        b.csvr2[k] = svr2;
        b.arx[k] = rr[0];
        b.ary[k] = rr[1];
        b.arz[k] = rr[2];
        b.ax[k] = com[0];
        b.ay[k] = com[1];
        b.az[k] = com[2];
        b.acell[k] = cell;
        b.astart[k] = a.new_start;
        b.acap[k] = a.cap_k;
        b.acount[k] = n;
        b.acell[a.rem] = -1;
        b.acount[a.rem] = 0;
        // both join the dirty list (keeper first) and leave the packed cell list
        unsigned int cnt = dcount;
        if (!dk) {
            if (csk >= 0) {
                b.cpr[csk] = -1.0;
                b.cf[csk].w = -1.f;
            }
            b.dirty[k] = 1;
            b.dirty_list[cnt] = k;
            b.dslot[k] = static_cast<int>(cnt);
            dsk = static_cast<int>(cnt);
            ++cnt;
        }
        if (!dr) {
            if (csr_slot >= 0) {
                b.cpr[csr_slot] = -1.0;
                b.cf[csr_slot].w = -1.f;
            }
            b.dirty[a.rem] = 1;
            b.dirty_list[cnt] = a.rem;
            b.dslot[a.rem] = static_cast<int>(cnt);
            dsr = static_cast<int>(cnt);
            ++cnt;
        }
        b.ss->dirty_count = cnt;
        b.dpx[dsk] = com[0];
        b.dpy[dsk] = com[1];
        b.dpz[dsk] = com[2];
        b.dpcell[dsk] = cell;
        b.dpcell[dsr] = -1;
        b.cslot[a.rem] = -1;
        for (int l = 0; l < 3; ++l) r[l] = rr[l];
        s_V = V;
        s_svr2 = svr2;
        mcac_merge_result &res = s_res;
        res.keeper = k;
        res.removed = a.rem;
        res.n = n;
        res.merged = 1;
        res.status = (V <= 0 || S <= 0) ? MCAC_VOL_SURF_ERROR : 0;  // aggregat.cpp:427-429
        res.volume = V;
        res.surface = S;
        res.sphere_volume = V;  // the sphere sum is the volume for method none
        res.radius_sum = sr;
        for (int l = 0; l < 3; ++l) {
            res.com[l] = com[l];
            res.rcm[l] = rr[l];
        }
    }
    lds_barrier();
    // distances to the centre of mass (aggregat.cpp:444-447) and contacts per row
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        double d0 = rlx[t] - r[0], d1 = rly[t] - r[1], d2 = rlz[t] - r[2];
        dcs[t] = __builtin_sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    }
    // aggregat.cpp:736-759: row i in ascending j, so each row is already in reference order
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int cnt = 0;
        for (int j = i + 1; j < n; ++j) {
            double d0 = rlx[i] - rlx[j], d1 = rly[i] - rly[j], d2 = rlz[i] - rlz[j];
            double s2 = rads[i] + rads[j];
            double q = d0 * d0 + d1 * d1 + d2 * d2;
            if (q > 1.0001 * (s2 * s2)) continue;
            double dist = __builtin_sqrt(q);
            if (dist <= (1. + kCoordinationEpsilon) * s2) {
                if (cnt < 8) crow[i * 8 + cnt] = (s2 - dist) / s2;  // a longer row is summed again in order
                ++cnt;
            }
        }
        nrow[i] = cnt;
    }
    lds_barrier();
    if (threadIdx.x == 0) {
        double rmax = 0.0, arg = 0.;
        for (int i = 0; i < n; ++i) {
            rmax = smax(rmax, rads[i] + dcs[i]);       // aggregat.cpp:462
            arg = arg + vols[i] * (dcs[i] * dcs[i]);   // aggregat.cpp:476
        }
        const double V = s_V;
        const double rg = __builtin_sqrt(__builtin_fabs((arg + 3. / 5. * s_svr2) / V));  // aggregat.cpp:481
        b.armax[k] = rmax;
        b.dpr[dsk] = rmax;
        b.arg[k] = rg;
        double overlapping = 0.0;
        long long inter = 0;
        for (int i = 0; i < n; ++i) {
            if (nrow[i] <= 8) {
                for (int c = 0; c < nrow[i]; ++c) {
                    overlapping += 2.0 * crow[i * 8 + c];  // aggregat.cpp:755
                    inter += 2;
                }
                continue;
            }
            // more contacts than the row holds (spheres grown into many
            // neighbours): the row again, in the same order and arithmetic
            for (int j = i + 1; j < n; ++j) {
                double d0 = rlx[i] - rlx[j], d1 = rly[i] - rly[j], d2 = rlz[i] - rlz[j];
                double s2 = rads[i] + rads[j];
                double q = d0 * d0 + d1 * d1 + d2 * d2;
                if (q > 1.0001 * (s2 * s2)) continue;
                double dist = __builtin_sqrt(q);
                if (dist <= (1. + kCoordinationEpsilon) * s2) {
                    overlapping += 2.0 * ((s2 - dist) / s2);
                    inter += 2;
                }
            }
        }
        double coord = 0.0;
        if (inter > 0) {
            overlapping /= static_cast<double>(inter);
            coord = static_cast<double>(inter) / static_cast<double>(n);
        }
        mcac_merge_result res = s_res;
        res.rmax = rmax;
        res.rg = rg;
        res.overlapping = overlapping;
        res.coordination = coord;
        res_out = res;
    }
}

__global__ __launch_bounds__(kBlock) void k_merge_small(DevBox b, MergeArgs a) {
    __shared__ mcac_merge_result res;
    merge_small_block(b, a, res, nullptr);
    if (threadIdx.x == 0) publish_merge(a, res);
}

// ------------------------------------------------------------------ multi-kernel merge
__device__ __forceinline__ void merge_prep_one(const DevBox &b, const MergeArgs &a) {
    if (threadIdx.x != 0) return;
    MergeState *m = b.ms;
    m->ok = 0;
    m->rmax_bits = 0ull;
    m->row_overflow = 0u;
    m->res = mcac_merge_result{};
    m->res.keeper = a.keep;
    m->res.removed = a.rem;
    double dp[3], rp[3];
    if (!merge_prepare(b, a, dp, rp)) {
        publish_merge(a, m->res);
        return;
    }
    for (int l = 0; l < 3; ++l) {
        m->diffpos[l] = dp[l];
        m->refpos[l] = rp[l];
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

__global__ void k_merge_prep(DevBox b, MergeArgs a) {
    if (blockIdx.x == 0) merge_prep_one(b, a);
}

// Grid-stride phases take their first position and stride, so that one
// workgroup can run them (first = threadIdx.x, stride = blockDim.x).
__device__ __forceinline__ void merge_lists_part(const DevBox &b, int first, int stride) {
    const MergeState *m = b.ms;
    if (!m->ok) return;
    const int nk = m->n_k, nr = m->n_r;
    for (int t = first; t < nk + nr; t += stride) {
        if (t < nk) {
            if (m->new_start != m->start_k) b.members[m->new_start + t] = b.members[m->start_k + t];
        } else {
            int s = b.members[m->start_r + (t - nk)];
            b.members[m->new_start + t] = s;
            remap_sphere(b, s, m->diffpos, m->refpos);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_merge_lists(DevBox b) {
    merge_lists_part(b, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Ordered sums over the removed spheres: one block stages tiles in LDS,
// thread 0 accumulates them in list order.
__global__ __launch_bounds__(kBlock) void k_merge_seq(DevBox b, MergeArgs a) {
    MergeState *m = b.ms;
    if (!m->ok) return;
    __shared__ double tv[kBlock], ts[kBlock], tx[kBlock], ty[kBlock], tz[kBlock], tr[kBlock];
    const int k = m->keep, nk = m->n_k, nr = m->n_r, base = m->new_start + nk;
    double V = b.avol[k], S = b.asurf[k], vx = b.csvx[k], vy = b.csvy[k], vz = b.csvz[k];
    double sr = b.csr[k], svr2 = b.csvr2[k];
    for (int t0 = 0; t0 < nr; t0 += kBlock) {
        const int cnt = (nr - t0 < kBlock) ? nr - t0 : kBlock;
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
                const double v = tv[i];
                V = V + v;
                S = S + ts[i];
                vx += tx[i] * v;
                vy += ty[i] * v;
                vz += tz[i] * v;
                sr += tr[i];
                svr2 = svr2 + v * (tr[i] * tr[i]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double r[3] = {vx / V, vy / V, vz / V};
    const int s0 = b.members[m->new_start];
    double com[3] = {periodic_position(b.sx[s0] + r[0], a.L), periodic_position(b.sy[s0] + r[1], a.L),
                     periodic_position(b.sz[s0] + r[2], a.L)};
    store_merged(b, a, V, S, vx, vy, vz, sr, svr2, r, com);
    for (int l = 0; l < 3; ++l) m->r[l] = r[l];
    mcac_merge_result &res = m->res;
    res.keeper = k;
    res.removed = a.rem;
    res.n = nk + nr;
    res.merged = 1;
    res.status = (V <= 0 || S <= 0) ? MCAC_VOL_SURF_ERROR : 0;
    res.volume = V;
    res.surface = S;
    res.sphere_volume = V;
    res.radius_sum = sr;
    for (int l = 0; l < 3; ++l) {
        res.com[l] = com[l];
        res.rcm[l] = r[l];
    }
}

__device__ __forceinline__ void merge_dc_part(const DevBox &b, int first, int stride) {
    MergeState *m = b.ms;
    if (!m->ok) return;
    const int n = m->n_k + m->n_r, st = m->new_start;
    const double r0 = m->r[0], r1 = m->r[1], r2 = m->r[2];
    double mx = 0.0;
    for (int t = first; t < n; t += stride) {
        int s = b.members[st + t];
        double d0 = b.srx[s] - r0, d1 = b.sry[s] - r1, d2 = b.srz[s] - r2;
        double dc = __builtin_sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        b.dc[t] = dc;
        mx = smax(mx, b.sr[s] + dc);
    }
    for (int off = 32; off > 0; off >>= 1) mx = smax(mx, __shfl_down(mx, off, 64));
    // non-negative doubles order like their bit patterns
    if ((threadIdx.x & 63) == 0) atomicMax(&m->rmax_bits, static_cast<unsigned long long>(__double_as_longlong(mx)));
}

__global__ __launch_bounds__(kBlock) void k_merge_dc(DevBox b) {
    merge_dc_part(b, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// With spherical caps the weights are the corrected volumes by list position
// and this kernel ends the sequence (publishes).
__device__ __forceinline__ void merge_rg_block(const DevBox &b, const MergeArgs &a) {
    MergeState *m = b.ms;
    if (!m->ok) return;  // the prep step published the failed re-check
    __shared__ double tv[kBlock], td[kBlock];
    const int k = m->keep, n = m->n_k + m->n_r, st = m->new_start;
    double arg = 0.;
    for (int t0 = 0; t0 < n; t0 += kBlock) {
        const int cnt = (n - t0 < kBlock) ? n - t0 : kBlock;
        if (threadIdx.x < cnt) {
            int s = b.members[st + t0 + threadIdx.x];
            tv[threadIdx.x] = (a.volsurf == 1 || a.volsurf == 3) ? b.cvol[t0 + threadIdx.x] : b.svol[s];
            td[threadIdx.x] = b.dc[t0 + threadIdx.x];
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < cnt; ++i) arg = arg + tv[i] * (td[i] * td[i]);
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double rmax = __longlong_as_double(static_cast<long long>(m->rmax_bits));
    const double rg = __builtin_sqrt(__builtin_fabs((arg + 3. / 5. * b.csvr2[k]) / b.avol[k]));
    b.armax[k] = rmax;
    b.dpr[b.dslot[k]] = rmax;
    b.arg[k] = rg;
    m->res.rmax = rmax;
    m->res.rg = rg;
    if (a.volsurf) publish_merge(a, m->res);
}

__global__ __launch_bounds__(kBlock) void k_merge_rg(DevBox b, MergeArgs a) { merge_rg_block(b, a); }

// Pair pass of the merged aggregate in kBlock x kBlock tiles, j tile in LDS.
// With both (spherical caps) a contact goes into both spheres' rows, with
// its distance, so that each row holds the sphere's whole neighbour map.
// One (i tile, j tile) pair with tj >= ti; every thread of the block calls it
// (one barrier inside).
__device__ __forceinline__ void overlap_pairs_tile(const DevBox &b, int ti, int tj, int both) {
    const MergeState *m = b.ms;
    const int n = m->n_k + m->n_r, st = m->new_start;
    __shared__ double jx[kBlock], jy[kBlock], jz[kBlock], jr[kBlock];
    const int jbase = tj * kBlock;
    const int jt = jbase + threadIdx.x;
    if (jt < n) {
        int s = b.members[st + jt];
        jx[threadIdx.x] = b.srx[s];
        jy[threadIdx.x] = b.sry[s];
        jz[threadIdx.x] = b.srz[s];
        jr[threadIdx.x] = b.sr[s];
    }
    __syncthreads();
    const int i = ti * kBlock + threadIdx.x;
    if (i >= n) return;  // after the block's only barrier
    const int si = b.members[st + i];
    const double ix = b.srx[si], iy = b.sry[si], iz = b.srz[si], ir = b.sr[si];
    const int jend = (jbase + kBlock < n) ? jbase + kBlock : n;
    for (int j = (jbase > i + 1 ? jbase : i + 1); j < jend; ++j) {
        const int l = j - jbase;
        double d0 = ix - jx[l], d1 = iy - jy[l], d2 = iz - jz[l];
        double s2 = ir + jr[l];
        double q = d0 * d0 + d1 * d1 + d2 * d2;
        if (q > 1.0001 * (s2 * s2)) continue;  // well outside the contact tolerance
        double dist = __builtin_sqrt(q);
        if (dist <= (1. + kCoordinationEpsilon) * s2) {
            double c = (s2 - dist) / s2;
            int slot = atomicAdd(&b.row_cnt[i], 1);
            if (slot < kOverlapSlots) {
                b.row_j[static_cast<long long>(i) * kOverlapSlots + slot] = j;
                b.row_c[static_cast<long long>(i) * kOverlapSlots + slot] = c;
                if (both) b.row_d[static_cast<long long>(i) * kOverlapSlots + slot] = dist;
            } else {
                atomicOr(&b.ms->row_overflow, 1u);
            }
            if (both) {
                slot = atomicAdd(&b.row_cnt[j], 1);
                if (slot < kOverlapSlots) {
                    b.row_j[static_cast<long long>(j) * kOverlapSlots + slot] = i;
                    b.row_c[static_cast<long long>(j) * kOverlapSlots + slot] = c;
                    b.row_d[static_cast<long long>(j) * kOverlapSlots + slot] = dist;
                } else {
                    atomicOr(&b.ms->row_overflow, 1u);
                }
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_overlap_pairs(DevBox b, int both) {
    if (!b.ms->ok || static_cast<int>(blockIdx.x) < static_cast<int>(blockIdx.y)) return;
    overlap_pairs_tile(b, blockIdx.y, blockIdx.x, both);
}

// Row i of the merged aggregate's contacts toward higher j, summed into the
// overlap sum in order (aggregat.cpp:736-759), read from global memory: for
// a row with more contacts than its slots (single thread).
__device__ __forceinline__ void overlapping_row_again(const DevBox &b, const MergeState *m, int i, int n,
                                                      double &overlapping, long long &inter) {
    const int st = m->new_start;
    const int si = b.members[st + i];
    const double ix = b.srx[si], iy = b.sry[si], iz = b.srz[si], ir = b.sr[si];
    for (int j = i + 1; j < n; ++j) {
        const int sj = b.members[st + j];
        double d0 = ix - b.srx[sj], d1 = iy - b.sry[sj], d2 = iz - b.srz[sj];
        double s2 = ir + b.sr[sj];
        double q = d0 * d0 + d1 * d1 + d2 * d2;
        if (q > 1.0001 * (s2 * s2)) continue;
        double dist = __builtin_sqrt(q);
        if (dist <= (1. + kCoordinationEpsilon) * s2) {
            overlapping += 2.0 * ((s2 - dist) / s2);
            inter += 2;
        }
    }
}

// Ordered overlap sum in (i, j) order; clears the row counters; publishes.
// Each thread sorts its row by j in registers and stages its coefficients
// toward higher j in LDS, from which thread 0 adds them in order.  With an
// overlap correction (volsurf != 0) the rows hold both directions (pairs with
// j > i are summed) and stay, sorted, for the per-sphere kernel, which clears
// them; no publication.
__device__ __forceinline__ void overlap_sum_block(const DevBox &b, const MergeArgs &a) {
    MergeState *m = b.ms;
    if (!m->ok) return;
    const int n = m->n_k + m->n_r;
    __shared__ double tc[kBlock * kOverlapSlots];  // each row's coefficients toward higher j, in order
    __shared__ int tcnt[kBlock];
    double overlapping = 0.0;
    long long inter = 0;
    for (int t0 = 0; t0 < n; t0 += kBlock) {
        const int cnt = (n - t0 < kBlock) ? n - t0 : kBlock;
        const int i = t0 + threadIdx.x;
        if (threadIdx.x < cnt) {
            const int rc_all = b.row_cnt[i];
            const int rc = rc_all > kOverlapSlots ? kOverlapSlots : rc_all;
            const long long base = static_cast<long long>(i) * kOverlapSlots;
            // the row in registers (unused slots sort last), a bitonic network by j
            int kj[kOverlapSlots];
            double kc[kOverlapSlots], kd[kOverlapSlots];
#pragma unroll
            for (int x = 0; x < kOverlapSlots; ++x) {
                kj[x] = x < rc ? b.row_j[base + x] : 0x7fffffff;
                kc[x] = x < rc ? b.row_c[base + x] : 0.0;
                kd[x] = (x < rc && a.volsurf) ? b.row_d[base + x] : 0.0;
            }
#pragma unroll
            for (int k = 2; k <= kOverlapSlots; k <<= 1) {
#pragma unroll
                for (int s = k >> 1; s > 0; s >>= 1) {
#pragma unroll
                    for (int x = 0; x < kOverlapSlots; ++x) {
                        const int y = x ^ s;
                        if (y > x && ((kj[x] > kj[y]) == ((x & k) == 0))) {
                            const int tj = kj[x];
                            kj[x] = kj[y];
                            kj[y] = tj;
                            const double tcv = kc[x];
                            kc[x] = kc[y];
                            kc[y] = tcv;
                            const double td = kd[x];
                            kd[x] = kd[y];
                            kd[y] = td;
                        }
                    }
                }
            }
            int hi = 0;
#pragma unroll
            for (int x = 0; x < kOverlapSlots; ++x) {
                if (x < rc && a.volsurf) {  // the sorted two-way row for the per-sphere kernels
                    b.row_j[base + x] = kj[x];
                    b.row_c[base + x] = kc[x];
                    b.row_d[base + x] = kd[x];
                }
                if (x < rc && kj[x] > i) tc[threadIdx.x * kOverlapSlots + hi++] = kc[x];  // j > i only
            }
            if (!a.volsurf) b.row_cnt[i] = 0;
            // a row longer than its slots (method none only: the corrections
            // need whole rows and report the overflow) is summed again below
            tcnt[threadIdx.x] = (!a.volsurf && rc_all > kOverlapSlots) ? -1 : hi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int x = 0; x < cnt; ++x) {
                if (tcnt[x] < 0) {
                    overlapping_row_again(b, m, t0 + x, n, overlapping, inter);
                    continue;
                }
                for (int y = 0; y < tcnt[x]; ++y) {
                    overlapping += 2.0 * tc[x * kOverlapSlots + y];
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
    if (m->row_overflow && a.volsurf) m->res.status = MCAC_E_CAPACITY;  // method none sums long rows again
    if (!a.volsurf) publish_merge(a, m->res);
}

__global__ __launch_bounds__(kBlock) void k_overlap_sum(DevBox b, MergeArgs a) { overlap_sum_block(b, a); }

// ------------------------------------------------------------------ spherical caps (aggregat.cpp:354-386)
// Corrected volume and surface of every sphere of the merged aggregate, one
// thread per list position t, from its two-way row (neighbours ascending,
// with distances): the caps of each lower neighbour in ascending order, then
// those of each higher neighbour in the order of the sphere's neighbour map,
// then the clamp at zero.  The map's previous size is the count of the
// sphere's neighbours inside the keeper for a keeper sphere (the keeper's last
// update saw the same rigid geometry) and 0 for a sphere of the removed
// aggregate (a new map).  Clears the row counters.
__device__ __forceinline__ void caps_volumes_part(const DevBox &b, int first, int stride) {
    const MergeState *m = b.ms;
    if (!m->ok) return;
    const int n = m->n_k + m->n_r, st = m->new_start, nk = m->n_k;
    for (int t = first; t < n; t += stride) {
        int rc = b.row_cnt[t];
        if (rc > kOverlapSlots) rc = kOverlapSlots;
        const long long base = static_cast<long long>(t) * kOverlapSlots;
        int64_t keys[kOverlapSlots], order[kOverlapSlots];
        int prev = 0;
        for (int x = 0; x < rc; ++x) {
            keys[x] = b.row_j[base + x];
            if (t < nk && keys[x] < nk) ++prev;
        }
        neighbour_order(prev, keys, rc, order);
        const int s = b.members[st + t];
        const double rt = b.sr[s], vt = b.svol[s], at = b.ssurf[s];
        double vol = vt, surf = at;
        for (int x = 0; x < rc && keys[x] < t; ++x) {  // lower neighbours: this sphere is the second
            const int sj = b.members[st + keys[x]];
            double v1, s1, v2, s2;
            caps_intersection(b.sr[sj], b.svol[sj], b.ssurf[sj], rt, vt, at, b.row_d[base + x], &v1, &s1, &v2, &s2);
            vol = vol - v2;
            surf = surf - s2;
        }
        for (int y = 0; y < rc; ++y) {  // higher neighbours in map order: this sphere is the first
            const int64_t j = order[y];
            if (j <= t) continue;
            int x = 0;
            while (keys[x] != j) ++x;
            const int sj = b.members[st + static_cast<int>(j)];
            double v1, s1, v2, s2;
            caps_intersection(rt, vt, at, b.sr[sj], b.svol[sj], b.ssurf[sj], b.row_d[base + x], &v1, &s1, &v2, &s2);
            vol = vol - v1;
            surf = surf - s1;
        }
        b.cvol[t] = smax(vol, 0.0);
        b.csurf[t] = smax(surf, 0.0);
        b.row_cnt[t] = 0;
    }
}

__global__ __launch_bounds__(kBlock) void k_caps_volumes(DevBox b) {
    caps_volumes_part(b, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Thread 0: centre of mass from the list-order sums and the aggregate volume V
// (aggregat.cpp:442-451), the merged aggregate's row, and the result fields.
__device__ __forceinline__ void merge_store_result(const DevBox &b, const MergeArgs &a, int k, int n, double V,
                                                   double S, double vx, double vy, double vz, double sr,
                                                   double svr2, double raw) {
    MergeState *m = b.ms;
    const int st = m->new_start;
    double r[3] = {vx / V, vy / V, vz / V};
    const int s0 = b.members[st];
    double com[3] = {periodic_position(b.sx[s0] + r[0], a.L), periodic_position(b.sy[s0] + r[1], a.L),
                     periodic_position(b.sz[s0] + r[2], a.L)};
    store_merged(b, a, V, S, vx, vy, vz, sr, svr2, r, com);
    for (int l = 0; l < 3; ++l) m->r[l] = r[l];
    mcac_merge_result &res = m->res;
    res.keeper = k;
    res.removed = a.rem;
    res.n = n;
    res.merged = 1;
    res.status = (V <= 0 || S <= 0) ? MCAC_VOL_SURF_ERROR : (m->row_overflow ? MCAC_E_CAPACITY : 0);
    res.volume = V;
    res.surface = S;
    res.sphere_volume = raw;
    res.radius_sum = sr;
    for (int l = 0; l < 3; ++l) {
        res.com[l] = com[l];
        res.rcm[l] = r[l];
    }
}

// ARVO parts of every sphere of the merged aggregate (aggregat.cpp:337-347
// with the restated arvo.cpp), one thread per list position t: the
// neighbours are the sphere's map entries in the map's order that overlap it
// and are not inside it; a sphere inside a neighbour carries nothing.  Writes
// the parts (m^3, m^2) for the ordered sums; clears the row counters.
__device__ __forceinline__ void arvo_volumes_part(const DevBox &b, int first, int stride) {
    const MergeState *m = b.ms;
    if (!m->ok) return;
    const int n = m->n_k + m->n_r, st = m->new_start, nk = m->n_k;
    const double nm = 1e+09;
    for (int t = first; t < n; t += stride) {
        int rc = b.row_cnt[t];
        if (rc > kOverlapSlots) rc = kOverlapSlots;
        const long long base = static_cast<long long>(t) * kOverlapSlots;
        int64_t keys[kOverlapSlots], order[kOverlapSlots];
        int prev = 0;
        for (int x = 0; x < rc; ++x) {
            keys[x] = b.row_j[base + x];
            if (t < nk && keys[x] < nk) ++prev;
        }
        neighbour_order(prev, keys, rc, order);
        const int s = b.members[st + t];
        const double ri = b.sr[s];
        double x[kArvoBalls + 1], y[kArvoBalls + 1], z[kArvoBalls + 1], r[kArvoBalls + 1];
        x[0] = b.srx[s] * nm;
        y[0] = b.sry[s] * nm;
        z[0] = b.srz[s] * nm;
        r[0] = ri * nm;
        int k = 0;
        bool inside = false;
        for (int q = 0; q < rc; ++q) {
            const int j = static_cast<int>(order[q]);
            int e = 0;
            while (keys[e] != order[q]) ++e;
            const double d = b.row_d[base + e];
            const int sj = b.members[st + j];
            const double rj = b.sr[sj];
            if (d >= (1. - kCoordinationEpsilon) * (ri + rj)) continue;       // touching within the tolerance
            if (d + ri < rj || (d + ri == rj && j < t)) inside = true;        // inside the neighbour
            if (d + rj < ri || (d + rj == ri && t < j)) continue;             // the neighbour is inside
            ++k;
            x[k] = b.srx[sj] * nm;
            y[k] = b.sry[sj] * nm;
            z[k] = b.srz[sj] * nm;
            r[k] = rj * nm;
        }
        double vol = 0.0, area = 0.0;
        if (!inside) arvo_ball(x, y, z, r, k, &vol, &area);
        b.cvol[t] = vol * 1e-27;
        b.csurf[t] = area * 1e-18;
        b.row_cnt[t] = 0;
    }
}

__global__ __launch_bounds__(kBlock) void k_arvo_volumes(DevBox b) {
    arvo_volumes_part(b, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Ordered sums of the merged aggregate, all positions from the start
// (aggregat.cpp:387-390, 437-441, 257-260, 474-478), with the corrected
// volumes (caps: stores the aggregate) or the spheres' own (alphas: the sums
// wait in MergeState for the host's correction factors): one block stages
// tiles in LDS, thread 0 accumulates in list order.
__device__ __forceinline__ void caps_sums_block(const DevBox &b, const MergeArgs &a) {
    MergeState *m = b.ms;
    if (!m->ok) return;
    const bool alphas = a.volsurf == 2;
    __shared__ double tv[kBlock], ts[kBlock], tx[kBlock], ty[kBlock], tz[kBlock], tr[kBlock], tw[kBlock];
    const int k = m->keep, n = m->n_k + m->n_r, st = m->new_start;
    double V = 0.0, S = 0.0, vx = 0.0, vy = 0.0, vz = 0.0, sr = 0.0, svr2 = 0.0, raw = 0.0;
    for (int t0 = 0; t0 < n; t0 += kBlock) {
        const int cnt = (n - t0 < kBlock) ? n - t0 : kBlock;
        if (threadIdx.x < cnt) {
            const int t = t0 + threadIdx.x;
            const int s = b.members[st + t];
            tv[threadIdx.x] = alphas ? b.svol[s] : b.cvol[t];
            ts[threadIdx.x] = alphas ? b.ssurf[s] : b.csurf[t];
            tw[threadIdx.x] = b.svol[s];
            tx[threadIdx.x] = b.srx[s];
            ty[threadIdx.x] = b.sry[s];
            tz[threadIdx.x] = b.srz[s];
            tr[threadIdx.x] = b.sr[s];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = 0; i < cnt; ++i) {
                const double v = tv[i];
                V = V + v;
                S = S + ts[i];
                vx += tx[i] * v;
                vy += ty[i] * v;
                vz += tz[i] * v;
                sr += tr[i];
                svr2 = svr2 + v * (tr[i] * tr[i]);
                raw += tw[i];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (alphas) {
        const double s8[8] = {V, S, vx, vy, vz, sr, svr2, raw};
        for (int l = 0; l < 8; ++l) m->sums[l] = s8[l];
        return;
    }
    merge_store_result(b, a, k, n, V, S, vx, vy, vz, sr, svr2, raw);
}

__global__ __launch_bounds__(kBlock) void k_caps_sums(DevBox b, MergeArgs a) { caps_sums_block(b, a); }

// ------------------------------------------------------------------ alphas (aggregat.cpp:392-422)
// Each sphere's walk of its neighbour map in the map's order (all
// neighbours, as the reference's loop over distances[i]): per entry the
// overlap coefficient c_ij = (r_i + r_j - d) / (r_i + r_j), pow(r_i, 3) +
// pow(r_j, 3) (host libm values) and r_i² + r_j².  The host sums the rows in
// order (the cube of c_ij is glibc pow) and derives the correction factors.
// Moves the row counters to walk_n and clears them.
__device__ __forceinline__ void alpha_walk_part(const DevBox &b, int first, int stride) {
    const MergeState *m = b.ms;
    if (!m->ok) return;
    const int n = m->n_k + m->n_r, st = m->new_start, nk = m->n_k;
    for (int t = first; t < n; t += stride) {
        int rc = b.row_cnt[t];
        if (rc > kOverlapSlots) rc = kOverlapSlots;
        const long long base = static_cast<long long>(t) * kOverlapSlots;
        int64_t keys[kOverlapSlots], order[kOverlapSlots];
        int prev = 0;
        for (int x = 0; x < rc; ++x) {
            keys[x] = b.row_j[base + x];
            if (t < nk && keys[x] < nk) ++prev;
        }
        neighbour_order(prev, keys, rc, order);
        const int s = b.members[st + t];
        const double ri = b.sr[s], vi = b.sr3[s];
        for (int y = 0; y < rc; ++y) {
            int x = 0;
            while (keys[x] != order[y]) ++x;
            const int sj = b.members[st + static_cast<int>(order[y])];
            const double rj = b.sr[sj];
            const double c = (ri + rj - b.row_d[base + x]) / (ri + rj);
            double *w = b.walk + 3 * (base + y);
            w[0] = c;
            w[1] = vi + b.sr3[sj];
            w[2] = ri * ri + rj * rj;
        }
        b.walk_n[t] = rc;
        b.row_cnt[t] = 0;
    }
}

__global__ __launch_bounds__(kBlock) void k_alpha_walk(DevBox b) {
    alpha_walk_part(b, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Second alphas phase (thread 0): the corrected volume and surface, and the
// walk's overlap and coordination (aggregat.cpp:417-421) when it had entries.
__device__ __forceinline__ void alpha_store_one(const DevBox &b, const MergeArgs &a, double V, double S,
                                                double overlapping, double coordination, int set_overlap) {
    MergeState *m = b.ms;
    if (!m->ok || threadIdx.x != 0) return;
    const double *s8 = m->sums;
    merge_store_result(b, a, m->keep, m->n_k + m->n_r, V, S, s8[2], s8[3], s8[4], s8[5], s8[6], s8[7]);
    if (set_overlap) {
        m->res.overlapping = overlapping;
        m->res.coordination = coordination;
    }
}

__global__ void k_alpha_store(DevBox b, MergeArgs a, double V, double S, double overlapping, double coordination,
                              int set_overlap) {
    if (blockIdx.x == 0) alpha_store_one(b, a, V, S, overlapping, coordination, set_overlap);
}

// ------------------------------------------------------------------ one-workgroup overlap-correction merge
// The phases of the multi-kernel sequence above, in the same order and with
// the same per-position work, run by one workgroup with a barrier between
// phases (after a barrier the block sees every global write and atomic its
// threads made before it), for merged aggregates of at most kFusedMerge
// spheres: one launch per merge instead of six to eight.  Caps and arvo end
// with the publication; alphas stops after the list-order sums, which the
// host turns into correction factors for k_alpha_finish.
__global__ __launch_bounds__(kBlock) void k_merge_fused(DevBox b, MergeArgs a) {
    merge_prep_one(b, a);
    __syncthreads();
    if (!b.ms->ok) return;  // the prep step published the failed re-check
    const int n = a.n_k + a.n_r;
    merge_lists_part(b, threadIdx.x, kBlock);
    __syncthreads();
    const int tiles = (n + kBlock - 1) / kBlock;
    for (int ti = 0; ti < tiles; ++ti) {
        for (int tj = ti; tj < tiles; ++tj) {
            overlap_pairs_tile(b, ti, tj, 1);
            __syncthreads();  // the j tile in LDS is reloaded next
        }
    }
    overlap_sum_block(b, a);
    __syncthreads();
    if (a.volsurf == 1) caps_volumes_part(b, threadIdx.x, kBlock);
    else if (a.volsurf == 3) arvo_volumes_part(b, threadIdx.x, kBlock);
    else alpha_walk_part(b, threadIdx.x, kBlock);
    __syncthreads();
    caps_sums_block(b, a);
    if (a.volsurf == 2) return;
    __syncthreads();
    merge_dc_part(b, threadIdx.x, kBlock);
    __syncthreads();
    merge_rg_block(b, a);
}

// Second alphas phase in one workgroup: the corrected sums stored, then dc,
// rmax and Rg (publishes).
__global__ __launch_bounds__(kBlock) void k_alpha_finish(DevBox b, MergeArgs a, double V, double S,
                                                         double overlapping, double coordination, int set_overlap) {
    alpha_store_one(b, a, V, S, overlapping, coordination, set_overlap);
    __syncthreads();
    merge_dc_part(b, threadIdx.x, kBlock);
    __syncthreads();
    merge_rg_block(b, a);
}

__global__ void k_gather_members(DevBox b, int agg, int *out, int n) {
    const int st = b.astart[agg];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) out[t] = b.members[st + t];
}

// Aggregate id of every sphere: one block per live aggregate scatters its id.
__global__ void k_sphere_owner(DevBox b, const int *ids, int n_ids, int *owner) {
    for (int a = blockIdx.x; a < n_ids; a += gridDim.x) {
        const int id = ids[a];
        const int n = b.acount[id], st = b.astart[id];
        for (int t = threadIdx.x; t < n; t += blockDim.x) owner[b.members[st + t]] = id;
    }
}

// Pool compaction: copy each live aggregate's members to its new start.
__global__ void k_compact(const int *old_members, int *new_members, const int *old_start, const int *new_start,
                          const int *count, const int *ids, int n_ids) {
    for (int a = blockIdx.x; a < n_ids; a += gridDim.x) {
        const int id = ids[a];
        const int n = count[id], so = old_start[id], sn = new_start[a];
        for (int t = threadIdx.x; t < n; t += blockDim.x) new_members[sn + t] = old_members[so + t];
    }
}

}  // namespace mcac_dev
