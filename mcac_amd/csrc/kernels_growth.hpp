// Surface growth of every sphere and the whole-box aggregate update that
// follows it (calcul.cpp:155-205 with individual_surf_reactions off and
// full_aggregate_update_frequency 1, volsurf_method none):
//   Sphere::croissance_surface   sphere.cpp:113-120 (r + u_sg dt, 4/3 pi r^3, 4 pi r^2)
//   Aggregate::update            aggregat.cpp:284-288: compute_volume_surface
//                                (:321-390, sphere sums), compute_mass_center
//                                (:431-452), compute_max_radius (:453-464),
//                                compute_giration_radius (:465-483)
// One wave per live aggregate walks its member list in 64-sphere chunks; every
// lane adds the chunk's terms in list order (broadcast by shuffles), so each
// sum is the reference's sequential FP64 sum.  The merge caches (sums the next
// merge continues) are rewritten from the grown spheres.
#pragma once
#include "kernels_common.hpp"

namespace mcac_dev {

constexpr double kVolumeFactorD = 4 * kPi / 3;  // constants.hpp:114
constexpr double kSurfaceFactorD = 4 * kPi;     // constants.hpp:115

// ids[0..n_ids) may hold dead ids (count 0): skipped.  Outputs by position in
// ids: volume, surface, list-order radius sum, Rg and rmax of each live aggregate.
// The contact search's copies of CoM and rmax follow in place: a dirty
// aggregate's dirty-list entry, or the packed cell entry of one still in the
// cell it was built in; *o_left counts the aggregates whose growth took them
// out of their built cell (their entries are stale: the cell list must be
// rebuilt before the next search).
// With grow = 0 the spheres keep their radii, volumes and surfaces (the update
// alone); o_flag[wave] gets 1 when a sphere of the aggregate is at or below
// rp_min after growing (oxidation removes it).
__global__ __launch_bounds__(kBlock) void k_grow_update(DevBox b, const int *ids, int n_ids, double dr, int grow,
                                                       double rp_min, double L, int n_div, double *o_vol,
                                                       double *o_surf, double *o_rsum, double *o_rg, double *o_rmax,
                                                       int *o_left, int *o_flag) {
    const int wave = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = static_cast<int>(threadIdx.x & 63);
    if (wave >= n_ids) return;
    const int id = ids[wave];
    const int n = b.acount[id];
    if (n <= 0) return;
    const int st = b.astart[id];
    // grow, then volume, surface, sum of vol * relative position, radius sum
    // and sum of vol * r^2 in list order
    double V = 0.0, S = 0.0, vx = 0.0, vy = 0.0, vz = 0.0, rs = 0.0, brg = 0.0;
    int small = 0;
    for (int c0 = 0; c0 < n; c0 += 64) {
        const int i = c0 + lane;
        double r = 0.0, vol = 0.0, surf = 0.0, rx = 0.0, ry = 0.0, rz = 0.0;
        if (i < n) {
            const int s = b.members[st + i];
            if (grow) {
                r = b.sr[s] + dr;  // PhysicalModel::grow: r + u_sg * dt, dr = u_sg * dt
                const double r2 = r * r;
                const double r3 = r2 * r;
                vol = kVolumeFactorD * r3;
                surf = kSurfaceFactorD * r2;
                b.sr[s] = r;
                b.svol[s] = vol;
                b.ssurf[s] = surf;
            } else {
                r = b.sr[s];
                vol = b.svol[s];
                surf = b.ssurf[s];
            }
            small |= (r <= rp_min) ? 1 : 0;
            rx = b.srx[s];
            ry = b.sry[s];
            rz = b.srz[s];
        }
        const int m = (n - c0 < 64) ? n - c0 : 64;
        for (int j = 0; j < m; ++j) {
            const double vj = __shfl(vol, j, 64), rj = __shfl(r, j, 64);
            V = V + vj;
            S = S + __shfl(surf, j, 64);
            vx = vx + __shfl(rx, j, 64) * vj;
            vy = vy + __shfl(ry, j, 64) * vj;
            vz = vz + __shfl(rz, j, 64) * vj;
            rs = rs + rj;
            brg = brg + vj * (rj * rj);
        }
    }
    const double cx = vx / V, cy = vy / V, cz = vz / V;  // r /= agregat_volume
    // distances to the centre of mass, rmax and sum of vol * dc^2 in list order
    double rmax = 0.0, arg = 0.0;
    for (int c0 = 0; c0 < n; c0 += 64) {
        const int i = c0 + lane;
        double dc = 0.0, r = 0.0, vol = 0.0;
        if (i < n) {
            const int s = b.members[st + i];
            const double dx = b.srx[s] - cx, dy = b.sry[s] - cy, dz = b.srz[s] - cz;
            dc = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
            r = b.sr[s];  // written above by this lane
            vol = b.svol[s];
        }
        const int m = (n - c0 < 64) ? n - c0 : 64;
        for (int j = 0; j < m; ++j) {
            const double dj = __shfl(dc, j, 64);
            rmax = smax(rmax, __shfl(r, j, 64) + dj);
            arg = arg + __shfl(vol, j, 64) * (dj * dj);
        }
    }
    const double rg = __builtin_sqrt(__builtin_fabs((arg + 3. / 5. * brg) / V));
    if (lane == 0) {
        // set_position(sphere 0 + r_cm) wraps into the box (aggregat.cpp:109-118)
        const int s0 = b.members[st];
        const double px = periodic_position(b.sx[s0] + cx, L);
        const double py = periodic_position(b.sy[s0] + cy, L);
        const double pz = periodic_position(b.sz[s0] + cz, L);
        b.avol[id] = V;
        b.asurf[id] = S;
        b.csvx[id] = vx;
        b.csvy[id] = vy;
        b.csvz[id] = vz;
        b.csr[id] = rs;
        b.csvr2[id] = brg;
        b.arx[id] = cx;
        b.ary[id] = cy;
        b.arz[id] = cz;
        b.ax[id] = px;
        b.ay[id] = py;
        b.az[id] = pz;
        const int cell = ref_cell(px, py, pz, n_div, L);
        const int built = b.acell[id];
        b.acell[id] = cell;
        b.armax[id] = rmax;
        if (b.dirty[id]) {
            set_dirty_entry(b, b.dslot[id], px, py, pz, rmax, cell);
        } else {
            const int cs = b.cslot[id];
            if (cs >= 0 && cell == built) {
                b.cpx[cs] = px;
                b.cpy[cs] = py;
                b.cpz[cs] = pz;
                b.cpr[cs] = rmax;
                b.cf[cs] = make_float4(static_cast<float>(px), static_cast<float>(py), static_cast<float>(pz),
                                       static_cast<float>(rmax));
            } else {
                atomicAdd(o_left, 1);
            }
        }
        b.arg[id] = rg;
        o_vol[wave] = V;
        o_surf[wave] = S;
        o_rsum[wave] = rs;
        o_rg[wave] = rg;
        o_rmax[wave] = rmax;
    }
    const bool any_small = __any(small);
    if (lane == 0) o_flag[wave] = any_small ? 1 : 0;
}

// Aggregate::split's question (aggregat.cpp:553-584) for each aggregate of
// ids: are its spheres one connected set under the contact test
// (sphere_distances.cpp:81-90)?  One wave per aggregate of at most 64
// spheres: lane i builds sphere i's contact bits, and the set reached from
// sphere 0 grows by wave-wide ORs until it stops.  o_flag[wave] |= 2 when the
// aggregate falls apart, or when it has more than 64 spheres (left to the host).
__global__ __launch_bounds__(kBlock) void k_connected(DevBox b, const int *ids, int n_ids, double L, int *o_flag) {
    const int wave = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = static_cast<int>(threadIdx.x & 63);
    if (wave >= n_ids) return;
    const int id = ids[wave];
    const int n = b.acount[id];
    if (n <= 1) return;
    if (n > 64) {
        if (lane == 0) o_flag[wave] |= 2;
        return;
    }
    const int st = b.astart[id];
    double x = 0.0, y = 0.0, z = 0.0, r = 0.0;
    if (lane < n) {
        const int s = b.members[st + lane];
        x = b.sx[s];
        y = b.sy[s];
        z = b.sz[s];
        r = b.sr[s];
    }
    unsigned long long adj = 0ull;
    for (int j = 0; j < n; ++j) {
        const double xj = __shfl(x, j, 64), yj = __shfl(y, j, 64), zj = __shfl(z, j, 64), rj = __shfl(r, j, 64);
        if (lane < n && contact(x, y, z, r, xj, yj, zj, rj, L)) adj |= 1ull << j;
    }
    const unsigned long long all = (n == 64) ? ~0ull : ((1ull << n) - 1ull);
    unsigned long long reached = 1ull;
    for (int it = 0; it < n; ++it) {
        unsigned long long mine = ((reached >> lane) & 1ull) ? adj : 0ull;
        unsigned lo = static_cast<unsigned>(mine), hi = static_cast<unsigned>(mine >> 32);
        for (int m = 1; m < 64; m <<= 1) {
            lo |= __shfl_xor(lo, m, 64);
            hi |= __shfl_xor(hi, m, 64);
        }
        const unsigned long long next = reached | (static_cast<unsigned long long>(hi) << 32) | lo;
        if (next == reached) break;
        reached = next;
    }
    if (lane == 0 && (reached & all) != all) o_flag[wave] |= 2;
}

}  // namespace mcac_dev
