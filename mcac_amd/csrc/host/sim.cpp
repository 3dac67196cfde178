// Host driver of one aggregation run: the reference MC loop (src/calcul.cpp)
// on the host CPU, the collision path on the device through mcac_gpu.h.
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <limits>
#include <memory>
#include <numeric>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mcac_sim.h"
#include "physics.hpp"
#include "pick_table.hpp"
#include "sim_group.hpp"
#include "xdmf_writer.hpp"

namespace mcac {
namespace {

using Clock = std::chrono::steady_clock;
inline double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}
inline double smax(double a, double b) { return (a < b) ? b : a; }

inline double periodic_position(double p, double L) {
    while (p < 0) p += L;
    while (p >= L) p -= L;
    return p;
}
inline double periodic_distance(double d, double L) {
    double half = 0.5 * L;
    while (d < -half) d += L;
    while (d >= half) d -= L;
    return d;
}
inline double distance_2(const Vec3 &a, const Vec3 &b, double L) {
    double dx = periodic_distance(a[0] - b[0], L);
    double dy = periodic_distance(a[1] - b[1], L);
    double dz = periodic_distance(a[2] - b[2], L);
    return dx * dx + dy * dy + dz * dz;
}
inline bool contact(const Vec3 &a, double ra, const Vec3 &b, double rb, double L) {
    double s = ra + rb;
    return distance_2(a, b, L) - s * s <= kContactEpsilon;
}

// Live aggregate ids: flags plus a Fenwick tree of counts, so that removal and
// the label of an id (its rank among live ids, aggregat_list_storage.cpp:37-44)
// cost O(log N); the label-ordered list is built only where it is read.
class LiveSet {
  public:
    void reset(size_t n) {
        alive_.assign(n, 1);
        tree_.assign(n + 1, 0);
        for (size_t i = 1; i <= n; ++i) {
            tree_[i] += 1;
            size_t j = i + (i & (~i + 1));
            if (j <= n) tree_[j] += tree_[i];
        }
        count_ = n;
        cache_valid_ = false;
    }
    size_t size() const { return count_; }
    bool alive(size_t id) const { return alive_[id] != 0; }
    void erase(size_t id) {
        if (!alive_[id]) return;
        alive_[id] = 0;
        for (size_t i = id + 1; i < tree_.size(); i += i & (~i + 1)) tree_[i] -= 1;
        --count_;
        cache_valid_ = false;
    }
    int64_t rank(size_t id) const {  // live ids before id
        int64_t r = 0;
        for (size_t i = id; i > 0; i -= i & (~i + 1)) r += tree_[i];
        return r;
    }
    const std::vector<int> &ids() {
        if (!cache_valid_) {
            cache_.clear();
            cache_.reserve(count_);
            for (size_t i = 0; i < alive_.size(); ++i)
                if (alive_[i]) cache_.push_back(static_cast<int>(i));
            cache_valid_ = true;
        }
        return cache_;
    }

  private:
    std::vector<char> alive_;
    std::vector<int64_t> tree_;
    size_t count_ = 0;
    std::vector<int> cache_;
    bool cache_valid_ = false;
};

}  // namespace

struct Simulation {
    PhysicalModel pm;
    std::string ini_text;
    std::string out_dir;
    RandomStream rng{0};
    int device = 0;
    int precision = 0;  // box contact-search precision (mcac_box_params)
    mcac_box *box = nullptr;
    mcac_group *group = nullptr;  // lock-step form: the box works on the group's stream
    mcac_group *own_group = nullptr;  // one-box group for rounds of steps in one launch
    bool rounds = false;              // issue rounds through own_group when the steps fit one workgroup
    int engine_blocks = 0;            // persistent step engine grid (mcac_box_set_engine), 0: off
    std::vector<int64_t> r_aggs;      // scratch of issue_round
    std::vector<double> r_dirs, r_lpm, r_maxr;
    std::vector<uint64_t> r_tickets;
    std::vector<mcac_box *> r_boxes;
    // spheres (index order) — host copy of the initial state only
    size_t N = 0;              // spheres in the box
    size_t n_ids = 0;          // stable ids handed out (per-id host arrays hold this many)
    size_t dup_threshold = 0;  // calcul.cpp:58: duplication when an event leaves at most this many aggregates
    int n_duplications = 0;
    // aggregates by stable id
    std::vector<double> ts, lpm, f_agg, d_m, dp, rg, rmax, vol, surf, ptime, dg_over_dp, overlapping, coord;
    std::vector<Vec3> com, rcm;
    std::vector<int> nsph;
    LiveSet live;  // live ids (label order = ascending id)
    double maxradius = 0, max_time_step = 0, avg_npp = 1;
    PickTable pick;
    bool event = true;
    bool pm_stale = false;
    size_t total_events = 0;
    int64_t step_count = 0, merges = 0;
    bool trace_steps = false, trace_merges = false;
    std::vector<mcac_step_rec> steps;
    std::vector<mcac_merge_rec> merge_recs;
    double q_query = 0, q_filtered = 0, q_pairs = 0, n_ties = 0;
    double t_pick = 0, t_merge_host = 0, t_device = 0, t_init = 0;
    double t_tie = 0, t_enqueue = 0, t_wait = 0, t_merge_call = 0, t_drain = 0;
    bool initialized = false;
    bool final_written = false;
    std::vector<double> sphere_r;                            // radii (grow with surface reactions)
    // surface growth: per-id outputs of mcac_box_grow
    std::vector<double> g_vol, g_surf, g_rsum, g_rg, g_rmax;
    double t_grow = 0;
    std::unique_ptr<SnapshotWriter> out_spheres, out_aggs;  // snapshot files when out_dir is set

    ~Simulation() {
        if (box) mcac_box_destroy(box);
        if (own_group) mcac_group_destroy(own_group);  // after the box that works on its stream
    }

    // Volume and surface a one-sphere aggregate carries (compute_volume_surface,
    // aggregat.cpp:321-430): the sphere's own, or with ARVO (method 3) the
    // closed form arvo.cpp computes for a ball without neighbours, in
    // nanometres (VOLUME_FACTOR r^3, SURFACE_FACTOR r^2) and back to SI.
    double single_volume(double r, double v) const {
        if (pm.volsurf != 3) return v;
        const double rn = r * 1e+09;
        return kVolumeFactor * rn * rn * rn * 1e-27;
    }
    double single_surface(double r, double s) const {
        if (pm.volsurf != 3) return s;
        const double rn = r * 1e+09;
        return kSurfaceFactor * rn * rn * 1e-18;
    }

    // Monomer update() (aggregat.cpp:284-288) for a single sphere of radius r
    // (own volume v and surface s).
    void monomer_update(size_t i, double r, double v, double s) {
        const double part = single_volume(r, v);
        vol[i] = 0.0 + part;
        surf[i] = 0.0 + single_surface(r, s);
        rcm[i] = {0., 0., 0.};
        rmax[i] = smax(0.0, r + 0.0);
        double brg = 0. + part * (r * r);
        double arg = 0. + part * (0.0 * 0.0);
        rg[i] = std::sqrt(std::abs((arg + 3. / 5. * brg) / vol[i]));
        AggregatePhysics ph = aggregate_physics(pm, 1, vol[i], 0.0 + v, 0. + r, rg[i]);
        dp[i] = ph.dp;
        f_agg[i] = ph.f_agg;
        d_m[i] = ph.d_m;
        ts[i] = ph.time_step;
        lpm[i] = ph.lpm;
        dg_over_dp[i] = ph.dg_over_dp;
        overlapping[i] = 0.0;
        coord[i] = 0.0;
        if (rmax[i] > maxradius) maxradius = rmax[i];
    }

    // AggregatList constructor (aggregat_list_storage.cpp:52-88): random
    // non-overlapping monomers (aggregat.cpp:162-229), then the volume
    // fraction rescale.  test_free_space (aggregat_list.cpp:534-548) asks
    // whether any placed monomer in the reference Verlet window touches the
    // new sphere; a private hash grid finds the candidates and the window
    // test keeps exactly the reference's set.
    int place(std::vector<double> &sx, std::vector<double> &sy, std::vector<double> &sz, std::vector<double> &sr,
              std::vector<double> &svol, std::vector<double> &ssurf, std::vector<int> &vcell) {
        const double L = pm.box_length;
        const int nd = static_cast<int>(pm.n_verlet_divisions);
        const double vstep = static_cast<double>(nd) / L;
        const double target = 2e-9 * pm.mean_diameter;
        const int hn = static_cast<int>(std::min<double>(128, std::max<double>(3, std::floor(L / target))));
        const double h = L / hn;
        std::vector<int> head(static_cast<size_t>(hn) * hn * hn, -1), next(N, -1);
        for (size_t i = 0; i < N; ++i) {
            ptime[i] = pm.time;
            double diameter = pm.random_diameter(rng);
            const double radius = diameter * 0.5;
            bool placed = false;
            for (size_t attempt = 0; attempt < N && !placed; ++attempt) {
                Vec3 p;
                p[0] = rng.uniform() * L;
                p[1] = rng.uniform() * L;
                p[2] = rng.uniform() * L;
                // reference window of get_neighborhood(pos, radius + maxradius) (verlet.cpp:52-78)
                const double reach = radius + maxradius;
                int b1[3], len[3];
                for (int l = 0; l < 3; ++l) {
                    double hi = p[l] + reach + 0.0;
                    double lo = p[l] - reach + 0.0;
                    int a = static_cast<int>(std::floor(static_cast<double>(nd) * lo / L));
                    int c = static_cast<int>(std::floor(static_cast<double>(nd) * hi / L) + 1);
                    if (c - a >= nd) {
                        a = 0;
                        c = nd - 1;
                    }
                    b1[l] = a;
                    len[l] = c - a + 1;
                }
                int span = static_cast<int>(std::ceil(reach / h)) + 1;
                int lim = std::min(span, (hn - 1) / 2);
                int extra = span > (hn - 1) / 2 ? hn - 1 - 2 * lim : 0;
                int c0[3];
                for (int l = 0; l < 3; ++l) c0[l] = static_cast<int>(std::floor(periodic_position(p[l], L) / h));
                bool free = true;
                for (int di = -lim; di <= lim + extra && free; ++di)
                    for (int dj = -lim; dj <= lim + extra && free; ++dj)
                        for (int dk = -lim; dk <= lim + extra && free; ++dk) {
                            int ci = ((c0[0] + di) % hn + hn) % hn;
                            int cj = ((c0[1] + dj) % hn + hn) % hn;
                            int ck = ((c0[2] + dk) % hn + hn) % hn;
                            for (int j = head[(static_cast<size_t>(ci) * hn + cj) * hn + ck]; j >= 0 && free;
                                 j = next[static_cast<size_t>(j)]) {
                                const size_t J = static_cast<size_t>(j);
                                int vc = vcell[J];
                                int v[3] = {vc / (nd * nd), (vc / nd) % nd, vc % nd};
                                bool inside = true;
                                for (int l = 0; l < 3; ++l)
                                    inside = inside && (((v[l] - b1[l]) % nd + nd) % nd) < len[l];
                                if (!inside) continue;
                                if (!contact(p, radius, com[J], rmax[J], L)) continue;
                                if (contact(p, radius, {sx[J], sy[J], sz[J]}, sr[J], L)) free = false;
                            }
                        }
                if (!free) continue;
                // Aggregate::init (aggregat.cpp:194-229)
                com[i] = {periodic_position(p[0], L), periodic_position(p[1], L), periodic_position(p[2], L)};
                int vi[3];
                for (int l = 0; l < 3; ++l) {
                    int v = static_cast<int>(std::floor(com[i][static_cast<size_t>(l)] * vstep));
                    vi[l] = v < 0 ? 0 : (v >= nd ? nd - 1 : v);
                }
                vcell[i] = (vi[0] * nd + vi[1]) * nd + vi[2];
                sx[i] = p[0];
                sy[i] = p[1];
                sz[i] = p[2];
                sr[i] = radius;
                svol[i] = kVolumeFactor * std::pow(sr[i], 3);
                ssurf[i] = kSurfaceFactor * std::pow(sr[i], 2);
                nsph[i] = 1;
                monomer_update(i, sr[i], svol[i], ssurf[i]);
                int hc[3];
                for (int l = 0; l < 3; ++l)
                    hc[l] = std::min(hn - 1, static_cast<int>(com[i][static_cast<size_t>(l)] / h));
                size_t cell = (static_cast<size_t>(hc[0]) * hn + hc[1]) * hn + hc[2];
                next[i] = head[cell];
                head[cell] = static_cast<int>(i);
                placed = true;
            }
            if (!placed) return 6;  // TooDenseError (aggregat.cpp:192)
        }
        return 0;
    }

    // The device box for the current side, spheres and ids; it joins the
    // group when the run is one of a group's.
    int create_box() {
        mcac_box_params bp{};
        bp.box_length = pm.box_length;
        bp.n_div = static_cast<int64_t>(pm.n_verlet_divisions);
        bp.max_spheres = std::max<int64_t>(1, static_cast<int64_t>(N));
        bp.max_aggregates = std::max<int64_t>(1, static_cast<int64_t>(n_ids));
        bp.device = device;
        bp.precision = precision;
        bp.volsurf = pm.volsurf;
        int rc = mcac_box_create(&bp, &box);
        if (rc) return rc;
        if (group) rc = mcac_group_add(group, box);
        return rc;
    }
    // Launch modes of an uploaded box: merges in the contact step's launch
    // (calcul.cpp:156-160 always merges on contact) unless a growth step must
    // come first, and the persistent engine when asked for.
    int box_modes() {
        int rc = mcac_box_set_auto_merge(box, pm.with_surface_reactions ? 0 : 1);
        if (rc) return rc;
        if (pm.with_surface_reactions)
            for (auto *v : {&g_vol, &g_surf, &g_rsum, &g_rg, &g_rmax}) v->assign(n_ids, 0.0);
        if (!group && !own_group && engine_blocks > 0) rc = mcac_box_set_engine(box, engine_blocks);
        return rc;
    }

    // AggregatList::duplication (aggregat_list.cpp:142-190), at the top of the
    // loop after an event left at most N0/8 aggregates (calcul.cpp:72-81): the
    // box side doubles and every live aggregate, in label order, gets seven
    // copies, (i, j, k) in {0,1}^3 minus the origin in loop order, each a copy
    // of all its fields and spheres (aggregat_storage.cpp:122-164) translated by
    // (i, j, k) times the old side (Aggregate::translate, aggregat.cpp:148-161,
    // wrapping into the new box).  Copies take the next stable ids and their
    // spheres the next sphere indices, in that order.  The device box is
    // rebuilt at the new size from the state read back.
    bool duplication_due() const {
        return event && pm.with_domain_duplication && live.size() <= dup_threshold && !(pm.u_sg < 0.0) &&
               flight.empty();
    }
    int duplicate() {
        // the copies of the overlap corrections' neighbour maps would need the
        // maps' own copy history replayed: duplication runs with method none
        if (pm.volsurf != 0) {
            std::fprintf(stderr, "mcac_sim: domain duplication with volsurf_method other than none is not offered\n");
            return MCAC_INPUT_ERROR;
        }
        const size_t S = N, A = n_ids;
        std::vector<double> x(S), y(S), z(S), rx(S), ry(S), rz(S), r(S), sv(S), ss(S);
        int rc = mcac_box_download_spheres(box, x.data(), y.data(), z.data(), rx.data(), ry.data(), rz.data());
        if (rc == 0) rc = mcac_box_download_radii(box, r.data(), sv.data(), ss.data());
        std::vector<int64_t> cnt(A), mem(S);
        if (rc == 0) rc = mcac_box_download_members(box, cnt.data(), mem.data());
        // the host keeps positions only from placement and merges: the moves
        // since then are on the device
        std::vector<double> pos_d(3 * A), rcm_d(3 * A);
        if (rc == 0)
            rc = mcac_box_download_aggs(box, pos_d.data(), rcm_d.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
        if (rc) return rc;
        for (size_t a = 0; a < A; ++a)
            for (size_t l = 0; l < 3; ++l) {
                com[a][l] = pos_d[3 * a + l];
                rcm[a][l] = rcm_d[3 * a + l];
            }
        std::vector<int64_t> off(A + 1, 0);
        for (size_t a = 0; a < A; ++a) off[a + 1] = off[a] + cnt[a];
        const std::vector<int> ids(live.ids());
        const size_t n_live = ids.size();
        const double old_l = pm.box_length;
        pm.box_length *= 2;
        pm.n_monomeres *= 8;
        pm.box_volume = std::pow(pm.box_length, 3);
        const double L = pm.box_length;
        const size_t S2 = 8 * S, A2 = A + 7 * n_live;
        for (auto *v : {&x, &y, &z, &rx, &ry, &rz, &r, &sv, &ss}) v->reserve(S2);
        for (auto *v : {&ts, &lpm, &f_agg, &d_m, &dp, &rg, &rmax, &vol, &surf, &ptime, &dg_over_dp, &overlapping, &coord})
            v->resize(A2, 0.0);
        com.resize(A2, {0, 0, 0});
        rcm.resize(A2, {0, 0, 0});
        nsph.resize(A2, 0);
        cnt.resize(A2, 0);
        std::vector<int64_t> mem2(mem.begin(), mem.end());
        mem2.reserve(S2);
        size_t next = A;
        for (int id : ids) {
            const size_t a = static_cast<size_t>(id);
            for (int i = 0; i <= 1; ++i)
                for (int j = 0; j <= 1; ++j)
                    for (int k = 0; k <= 1; ++k) {
                        if (i == 0 && j == 0 && k == 0) continue;
                        const size_t c = next++;
                        for (auto *v : {&ts, &lpm, &f_agg, &d_m, &dp, &rg, &rmax, &vol, &surf, &ptime, &dg_over_dp,
                                        &overlapping, &coord})
                            (*v)[c] = (*v)[a];
                        rcm[c] = rcm[a];
                        nsph[c] = nsph[a];
                        cnt[c] = cnt[a];
                        const double vec[3] = {static_cast<double>(i) * old_l, static_cast<double>(j) * old_l,
                                               static_cast<double>(k) * old_l};
                        Vec3 pos;
                        for (int l = 0; l < 3; ++l)
                            pos[static_cast<size_t>(l)] =
                                periodic_position(com[a][static_cast<size_t>(l)] + vec[l], L);
                        com[c] = pos;
                        const Vec3 ref{pos[0] - rcm[c][0], pos[1] - rcm[c][1], pos[2] - rcm[c][2]};
                        for (int64_t q = off[a]; q < off[a + 1]; ++q) {
                            const size_t sp = static_cast<size_t>(mem[static_cast<size_t>(q)]);
                            mem2.push_back(static_cast<int64_t>(x.size()));
                            x.push_back(ref[0] + rx[sp]);
                            y.push_back(ref[1] + ry[sp]);
                            z.push_back(ref[2] + rz[sp]);
                            rx.push_back(rx[sp]);
                            ry.push_back(ry[sp]);
                            rz.push_back(rz[sp]);
                            r.push_back(r[sp]);
                            sv.push_back(sv[sp]);
                            ss.push_back(ss[sp]);
                        }
                    }
        }
        // the device state at the new size
        mcac_box_destroy(box);
        box = nullptr;
        N = S2;
        n_ids = A2;
        rc = create_box();
        if (rc) return rc;
        std::vector<double> pos3(3 * A2), rcm3(3 * A2);
        for (size_t a = 0; a < A2; ++a)
            for (size_t l = 0; l < 3; ++l) {
                pos3[3 * a + l] = com[a][l];
                rcm3[3 * a + l] = rcm[a][l];
            }
        rc = mcac_box_upload(box, static_cast<int64_t>(N), x.data(), y.data(), z.data(), r.data(), sv.data(),
                             ss.data(), rx.data(), ry.data(), rz.data(), static_cast<int64_t>(A2), pos3.data(),
                             rcm3.data(), rmax.data(), vol.data(), surf.data(), cnt.data(), mem2.data());
        if (rc) return rc;
        rc = box_modes();
        if (rc) return rc;
        live.reset(A2);
        for (size_t a = 0; a < A; ++a)
            if (cnt[a] == 0) live.erase(a);
        if (!sphere_r.empty()) sphere_r = r;
        pm_stale = true;
        n_duplications += 1;
        pick.build(live.ids(), ts, max_time_step);  // sort_time_steps(max) at the event
        return 0;
    }

    int init() {
        auto t0 = Clock::now();
        N = pm.n_monomeres;
        rng.reseed(static_cast<unsigned>(pm.random_seed));
        for (auto *v : {&ts, &lpm, &f_agg, &d_m, &dp, &rg, &rmax, &vol, &surf, &ptime, &dg_over_dp, &overlapping, &coord})
            v->assign(N, 0.0);
        com.assign(N, {0, 0, 0});
        rcm.assign(N, {0, 0, 0});
        nsph.assign(N, 0);
        std::vector<double> sx(N), sy(N), sz(N), sr(N), svol(N), ssurf(N), zero(N, 0.0);
        std::vector<int> vcell(N, 0);
        int rc = place(sx, sy, sz, sr, svol, ssurf, vcell);
        if (rc) return rc;
        // refresh (aggregat_list.cpp:100-108)
        max_time_step = ts[0];
        for (size_t i = 0; i < N; ++i) max_time_step = smax(ts[i], max_time_step);
        avg_npp = 1.0;
        if (pm.enforce_volume_fraction) {
            double current = 0.0;
            for (size_t i = 0; i < N; ++i) current += vol[i];
            double prescribed = pm.volume_fraction * std::pow(pm.box_length, 3);
            double correction = std::pow(prescribed / current, 1. / 3.);
            for (size_t s = 0; s < N; ++s) {
                sr[s] = sr[s] * correction;
                svol[s] = kVolumeFactor * std::pow(sr[s], 3);
                ssurf[s] = kSurfaceFactor * std::pow(sr[s], 2);
            }
            for (size_t i = 0; i < N; ++i) {  // compute_volume_surface only (aggregat_list_storage.cpp:84-86)
                vol[i] = 0.0 + single_volume(sr[i], svol[i]);
                surf[i] = 0.0 + single_surface(sr[i], ssurf[i]);
            }
        }
        live.reset(N);
        n_ids = N;
        dup_threshold = N / 8;
        // device state
        rc = create_box();
        if (rc) return rc;
        if (group) {
            depth = group_depth;  // steps picked ahead per box over its group launches
        } else if (rounds && out_dir.empty() && depth > 1) {
            rc = mcac_group_create(device, 1, &own_group);
            if (rc == 0) rc = mcac_group_add(own_group, box);
            if (rc) return rc;
        }
        std::vector<double> pos3(3 * N), rcm3(3 * N, 0.0);
        std::vector<int64_t> cnt(N, 1), mem(N);
        for (size_t i = 0; i < N; ++i) {
            for (int l = 0; l < 3; ++l) pos3[3 * i + l] = com[i][static_cast<size_t>(l)];
            mem[i] = static_cast<int64_t>(i);
        }
        rc = mcac_box_upload(box, static_cast<int64_t>(N), sx.data(), sy.data(), sz.data(), sr.data(), svol.data(),
                             ssurf.data(), zero.data(), zero.data(), zero.data(), static_cast<int64_t>(N), pos3.data(),
                             rcm3.data(), rmax.data(), vol.data(), surf.data(), cnt.data(), mem.data());
        if (rc) return rc;
        if (pm.with_surface_reactions) {
            // every step grows the whole box and updates every aggregate, so a
            // step depends on the one before it: no look-ahead, and a contact
            // merges after the growth (calcul.cpp:164-181)
            depth = 1;
            engine_blocks = 0;
        }
        rc = box_modes();
        if (rc) return rc;
        pick.build(live.ids(), ts);  // first sort_time_steps (calcul.cpp:99-101)
        event = true;
        initialized = true;
        t_init = ms_since(t0);
        if (!out_dir.empty()) {
            mkdir(out_dir.c_str(), 0755);
            std::ofstream os(out_dir + "/params.ini");
            os << IniFile(ini_text).dump();
            // sphere_list_storage.cpp:74, aggregat_list_storage.cpp:62
            out_spheres = std::make_unique<SnapshotWriter>(out_dir + "/Spheres", N, pm.n_time_per_file);
            out_aggs = std::make_unique<SnapshotWriter>(out_dir + "/Aggregats", N, pm.n_time_per_file);
            sphere_r = sr;
            depth = 1;  // a snapshot is the state at the top of the loop: no step ahead of it
        }
        return 0;
    }

    // physical_model.cpp (io) xmf_write: the "Physics" information, written as
    // `stream << value` does (six significant digits)
    std::vector<std::pair<std::string, std::string>> physics_info() {
        refresh_pm();
        auto f = [](double v) {
            std::ostringstream s;
            s << v;
            return s.str();
        };
        return {{"flux_surfgrowth", f(pm.flux_surfgrowth)}, {"u_sg", f(pm.u_sg)},
                {"dfe", f(pm.fractal_dimension)},           {"kfe", f(pm.fractal_prefactor)},
                {"lambda", f(pm.gaz_mean_free_path)},       {"rpeqmass", f(pm.mean_massic_radius)},
                {"gamma_", f(pm.friction_exponnant)},       {"P [Pa]", f(pm.pressure)},
                {"T [K]", f(pm.temperature)},               {"Mu", f(pm.viscosity)},
                {"Rho [kg/m3]", f(pm.density)},             {"Dpm [nm]", f(pm.mean_diameter)},
                {"sigmaDpm [nm]", f(pm.dispersion_diameter)}, {"FV [ppt]", f(pm.volume_fraction)},
                {"L", f(pm.box_length)},                    {"N []", std::to_string(static_cast<int>(pm.n_monomeres))}};
    }

    // SphereList::save and AggregatList::save (src/io/sphere_list.cpp:26-56,
    // src/io/aggregat_list.cpp:26-60): one grid each, in list order
    int save_snapshot() {
        if (!out_spheres || !out_aggs) return 0;
        const auto phys = physics_info();
        const size_t S = N;
        std::vector<double> x(S), y(S), z(S);
        int rc = mcac_box_download_spheres(box, x.data(), y.data(), z.data(), nullptr, nullptr, nullptr);
        if (rc) return rc;
        std::vector<int64_t> owner(S);
        rc = mcac_box_sphere_owners(box, owner.data());
        if (rc) return rc;
        if (pm.with_surface_reactions) {  // the radii have grown since the placement
            rc = mcac_box_download_radii(box, sphere_r.data(), nullptr, nullptr);
            if (rc == 0) rc = refresh_overlaps();
            if (rc) return rc;
        }
        std::vector<double> pos3(3 * S);
        std::vector<int64_t> slabel(S);
        std::vector<int32_t> zeros_s(S, 0);
        for (size_t i = 0; i < S; ++i) {
            pos3[3 * i] = x[i];
            pos3[3 * i + 1] = y[i];
            pos3[3 * i + 2] = z[i];
            slabel[i] = owner[i] >= 0 ? label_of(static_cast<int>(owner[i])) : -1;
        }
        out_spheres->write("Spheres", pm.time, pos3,
                           {{"electric_charge", 1, zeros_s.data(), S},
                            {"Radius", 0, sphere_r.data(), S},
                            {"Label", 2, slabel.data(), S}},
                           phys);
        // aggregates in label order (ascending live id)
        std::vector<double> apos(3 * n_ids), arcm(3 * n_ids), armax(n_ids), avol(n_ids), asurf(n_ids), arg(n_ids);
        std::vector<int64_t> acount(n_ids);
        rc = mcac_box_download_aggs(box, apos.data(), arcm.data(), armax.data(), avol.data(), asurf.data(), arg.data(),
                                    acount.data());
        if (rc) return rc;
        const std::vector<int> &ids = live.ids();
        const size_t A = ids.size();
        std::vector<double> p3(3 * A), c_rg(A), c_fagg(A), c_lpm(A), c_dt(A), c_rmax(A), c_vol(A), c_surf(A), c_pt(A),
            c_coord(A), c_ovl(A), c_dm(A);
        std::vector<int64_t> c_np(A), c_label(A);
        std::vector<int32_t> zeros_a(A, 0);
        for (size_t l = 0; l < A; ++l) {
            const size_t id = static_cast<size_t>(ids[l]);
            for (int k = 0; k < 3; ++k) p3[3 * l + static_cast<size_t>(k)] = apos[3 * id + static_cast<size_t>(k)];
            c_rg[l] = rg[id];
            c_np[l] = nsph[id];
            c_fagg[l] = f_agg[id];
            c_lpm[l] = lpm[id];
            c_dt[l] = ts[id];
            c_rmax[l] = rmax[id];
            c_vol[l] = vol[id];
            c_surf[l] = surf[id];
            c_pt[l] = ptime[id];
            c_coord[l] = coord[id];
            c_ovl[l] = overlapping[id];
            c_dm[l] = d_m[id];
            c_label[l] = static_cast<int64_t>(l);
        }
        out_aggs->write("Aggregats", pm.time, p3,
                        {{"Rg", 0, c_rg.data(), A},           {"Np", 2, c_np.data(), A},
                         {"f_agg", 0, c_fagg.data(), A},      {"lpm", 0, c_lpm.data(), A},
                         {"Deltat", 0, c_dt.data(), A},       {"Rmax", 0, c_rmax.data(), A},
                         {"Volume", 0, c_vol.data(), A},      {"Surface", 0, c_surf.data(), A},
                         {"proper_time", 0, c_pt.data(), A},  {"coordination_number", 0, c_coord.data(), A},
                         {"overlapping", 0, c_ovl.data(), A}, {"electric_charge", 1, zeros_a.data(), A},
                         {"d_m", 0, c_dm.data(), A},          {"Label", 2, c_label.data(), A}},
                        phys);
        return 0;
    }

    int64_t label_of(int id) const {
        return live.rank(static_cast<size_t>(id));
    }

    // calcul.cpp:29-43
    void save_advancement() {
        if (out_dir.empty()) return;
        refresh_pm();
        std::ofstream f(out_dir + "/advancement.dat", std::ios_base::app);
        f << pm.time;
        f << " " << pm.aggregate_concentration;
        f << " " << pm.volume_fraction;
        f << " " << avg_npp;
        f << " " << pm.temperature;
        f << " " << pm.box_volume;
        f << " " << pm.monomer_concentration;
        f << " " << pm.u_sg;
        f << " " << pm.flux_nucleation;
        f << std::endl;
    }
    // PhysicalModel::update at the last event (calcul.cpp:272-277), evaluated
    // when its outputs are read; the aggregate volumes have not changed since.
    void refresh_pm() {
        if (!pm_stale) return;
        double tv = 0.0, tsf = 0.0;
        for (int id : live.ids()) tv += vol[static_cast<size_t>(id)];
        for (int id : live.ids()) tsf += surf[static_cast<size_t>(id)];
        pm.update(live.size(), N, tv, tsf);
        pm_stale = false;
    }

    bool finished() const {
        size_t n = live.size();
        if (n < 1) return true;
        if (n <= pm.number_of_aggregates_limit) return true;
        if (pm.n_iter_without_event_limit > 0 &&
            pm.n_iter_without_event >= static_cast<size_t>(pm.n_iter_without_event_limit))
            return true;
        if (pm.physical_time_limit > 0 && pm.time >= pm.physical_time_limit) return true;
        if (pm.mean_monomere_per_aggregate_limit > 0 && avg_npp >= pm.mean_monomere_per_aggregate_limit) return true;
        if (!out_dir.empty() && (step_count & 4095) == 0) {
            struct stat st;
            if (stat((out_dir + "/STOPCODE").c_str(), &st) == 0) return true;
        }
        return false;
    }

    // The MC loop of calcul.cpp:66-281 (collisions on; no surface reactions,
    // nucleation, potentials or flame coupling), pipelined: up to `depth` steps
    // are picked and enqueued ahead assuming no merge in between; a step that
    // finds a contact closes the device gate, the steps behind it do nothing,
    // and they are issued again with the same draws once the merge is done.
    struct Draw {
        double u;
        Vec3 dir;
    };
    struct Pending {
        uint64_t ticket;
        int num;
        bool tie;
        double full;
        Draw draw;
    };
    std::deque<Draw> draws;      // taken from the stream, not yet consumed by a step
    std::deque<Pending> flight;  // enqueued on the device, in order
    size_t unticketed = 0;       // newest entries of `flight` picked for a group launch, not enqueued yet
    size_t group_depth = 1;      // steps of a box in flight over its group launches
    size_t depth = 3;

    // the k-th draw after the last consumed step (calcul.cpp:102, tools.cpp:83-84)
    const Draw &draw_at(size_t k) {
        while (draws.size() <= k) {
            Draw d;
            d.u = rng.uniform();
            d.dir = rng.direction();
            draws.push_back(d);
        }
        return draws[k];
    }

    bool can_issue(int64_t consumed, int64_t max_steps) const {
        if (consumed + static_cast<int64_t>(flight.size()) >= max_steps) return false;
        if (flight.size() >= depth) return false;
        if (pm.n_iter_without_event_limit > 0 &&
            pm.n_iter_without_event + flight.size() >= static_cast<size_t>(pm.n_iter_without_event_limit))
            return flight.empty();
        if (pm.physical_time_limit > 0) return flight.empty();  // time-driven stop: no look-ahead
        return true;
    }

    // The mover and move of the next step to issue (calcul.cpp:99-127).
    Pending pick_move() {
        auto t0 = Clock::now();
        const Draw &d = draw_at(flight.size());
        size_t tie_pos = 0;
        int num = pick.pick_fast(d.u, &tie_pos);  // aggregat_list.cpp:59-66
        bool tie = false;
        if (num < 0) {
            tie = true;
            auto tt0 = Clock::now();
            num = pick.resolve_tie(tie_pos, live.ids(), ts);
            t_tie += ms_since(tt0);
        }
        const double full = lpm[static_cast<size_t>(num)];
        t_pick += ms_since(t0);
        return {0, num, tie, full, d};
    }

    int issue() {
        Pending P = pick_move();
        auto te0 = Clock::now();
        int rc = mcac_move_enqueue(box, P.num, P.draw.dir.data(), P.full, maxradius, &P.ticket);
        t_enqueue += ms_since(te0);
        if (rc) return rc;
        flight.push_back(P);
        return 0;
    }

    // Steps picked ahead in rounds: when the next move's cell window is small
    // enough for one workgroup, up to kRoundDepth moves go into one group
    // launch (a workgroup runs them in order; those behind a contact do
    // nothing, as behind the gate); otherwise the 3-deep one-launch pipeline.
    // A new round is picked once at most kRoundRefill steps remain in flight.
    static constexpr size_t kRoundDepth = 8, kRoundRefill = 3;

    int issue_more(int64_t consumed, int64_t max_steps) {
        if (duplication_due() && !finished()) {
            int rc = duplicate();
            if (rc) return rc;
        }
        if (!own_group) {
            while (!finished() && can_issue(consumed, max_steps)) {
                int rc = issue();
                if (rc) return rc;
            }
            return 0;
        }
        if (flight.size() > kRoundRefill || finished() || !can_issue(consumed, max_steps)) return 0;
        Pending first = pick_move();
        const size_t U = static_cast<size_t>(first.num);
        if (mcac_box_block_steps(box, rmax[U] + maxradius + first.full) != 1) {
            // one launch per step for this one and up to the pipeline depth
            auto te0 = Clock::now();
            int rc = mcac_move_enqueue(box, first.num, first.draw.dir.data(), first.full, maxradius, &first.ticket);
            t_enqueue += ms_since(te0);
            if (rc) return rc;
            flight.push_back(first);
            while (!finished() && can_issue(consumed, max_steps)) {
                rc = issue();
                if (rc) return rc;
            }
            return 0;
        }
        const size_t saved = depth;
        depth = kRoundDepth;
        flight.push_back(first);
        size_t k = 1;
        while (!finished() && can_issue(consumed, max_steps)) {
            flight.push_back(pick_move());
            ++k;
        }
        depth = saved;
        r_boxes.assign(k, box);
        r_aggs.resize(k);
        r_dirs.resize(3 * k);
        r_lpm.resize(k);
        r_maxr.assign(k, maxradius);
        r_tickets.assign(k, 0);
        for (size_t q = 0; q < k; ++q) {
            const Pending &P = flight[flight.size() - k + q];
            r_aggs[q] = P.num;
            for (int l = 0; l < 3; ++l) r_dirs[3 * q + static_cast<size_t>(l)] = P.draw.dir[static_cast<size_t>(l)];
            r_lpm[q] = P.full;
        }
        auto te0 = Clock::now();
        int rc = mcac_move_enqueue_group(own_group, r_boxes.data(), r_aggs.data(), r_dirs.data(), r_lpm.data(),
                                         r_maxr.data(), static_cast<int>(k), r_tickets.data());
        t_enqueue += ms_since(te0);
        if (rc) return rc;
        for (size_t q = 0; q < k; ++q) flight[flight.size() - k + q].ticket = r_tickets[q];
        return 0;
    }

    // Surface reactions of one step and the updates they call for
    // (calcul.cpp:155-205).  The device grows (or, with a negative u_sg,
    // shrinks) the spheres of every live aggregate, or of the moved one alone
    // with individual_surf_reactions, and updates those aggregates; the
    // physics (update_partial, aggregat.cpp:254-281) runs here with the host's
    // libm.  Oxidation's removals and splits, which the device flags, are
    // carried out on the state read back (remove_and_split).
    std::vector<int32_t> g_flag;
    int react(double dt, int only, bool *split, bool *disappear) {
        auto t0 = Clock::now();
        *split = *disappear = false;
        const double dr = pm.u_sg * dt;  // PhysicalModel::grow, physical_model.cpp:587-590
        const bool oxid = pm.u_sg < 0.0;
        g_flag.assign(n_ids, 0);
        const int64_t one = only;
        int rc = mcac_box_grow_ids(box, only >= 0 ? &one : nullptr, only >= 0 ? 1 : 0, dr, 1, pm.rp_min_oxid,
                                   oxid ? 1 : 0, g_vol.data(), g_surf.data(), g_rsum.data(), g_rg.data(),
                                   g_rmax.data(), g_flag.data());
        if (rc) return rc;
        std::vector<int> grown;
        if (only >= 0) grown.push_back(only);
        else grown = live.ids();
        bool events = false;
        for (int id : grown) events = events || g_flag[static_cast<size_t>(id)] != 0;
        take_physics(grown, true);
        if (events) {
            rc = remove_and_split(grown, split, disappear);
            if (rc) return rc;
        }
        pm_stale = true;
        t_grow += ms_since(t0);
        return 0;
    }

    // Aggregate::update of every live aggregate without growth: the whole-box
    // update after an event (calcul.cpp:186-191, 197-202), which re-derives
    // each position from its first sphere.
    int update_all() {
        auto t0 = Clock::now();
        g_flag.assign(n_ids, 0);
        int rc = mcac_box_grow_ids(box, nullptr, 0, 0.0, 0, -1.0, 0, g_vol.data(), g_surf.data(), g_rsum.data(),
                                   g_rg.data(), g_rmax.data(), g_flag.data());
        if (rc) return rc;
        take_physics(live.ids(), false);
        t_grow += ms_since(t0);
        return 0;
    }

    // The device's update outputs of `ids` into the per-id arrays and the
    // physics; aggregates flagged for removal or split are left to
    // remove_and_split (their state before it was never updated in the
    // reference, so it does not raise maxradius).
    void take_physics(const std::vector<int> &ids, bool skip_flagged) {
        constexpr size_t kPer = 512;
        host_parallel_for((ids.size() + kPer - 1) / kPer, [&](size_t k) {
            const size_t e = std::min(ids.size(), (k + 1) * kPer);
            for (size_t l = k * kPer; l < e; ++l) {
                const size_t id = static_cast<size_t>(ids[l]);
                if (skip_flagged && g_flag[id]) continue;
                vol[id] = g_vol[id];
                surf[id] = g_surf[id];
                rg[id] = g_rg[id];
                rmax[id] = g_rmax[id];
                set_physics(id, g_rsum[id]);
            }
        });
        for (int id : ids)
            if (!(skip_flagged && g_flag[static_cast<size_t>(id)]) && rmax[static_cast<size_t>(id)] > maxradius)
                maxradius = rmax[static_cast<size_t>(id)];
    }
    // update_partial's physics of one aggregate (method none: the list-order
    // sum of the spheres' volumes is the volume)
    void set_physics(size_t id, double radius_sum) {
        const AggregatePhysics ph =
            aggregate_physics(pm, static_cast<size_t>(nsph[id]), vol[id], vol[id], radius_sum, rg[id]);
        dp[id] = ph.dp;
        f_agg[id] = ph.f_agg;
        d_m[id] = ph.d_m;
        ts[id] = ph.time_step;
        lpm[id] = ph.lpm;
        dg_over_dp[id] = ph.dg_over_dp;
    }

    // Host copy of the box for the removals and splits of a step.
    struct HostState {
        std::vector<double> x, y, z, rx, ry, rz, r, sv, ss;
        std::vector<std::vector<int64_t>> mem;  // member lists by id
    };
    // The contact's spheres followed through removals (-1: removed), so that
    // the merge sees them as they are after the reactions.
    int64_t track_ms = -1, track_os = -1;
    bool track_touch = false;  // they touch after the reactions (sphere_distances.cpp:81-90)

    // update_distances_and_overlapping (aggregat.cpp:719-764) of one aggregate:
    // mean overlap and coordination number of its sphere pairs in contact
    void overlaps(const HostState &h, size_t id) {
        const std::vector<int64_t> &m = h.mem[id];
        double ovl = 0.0;
        int inter = 0;
        for (size_t i = 0; i < m.size(); ++i)
            for (size_t j = i + 1; j < m.size(); ++j) {
                const size_t a = static_cast<size_t>(m[i]), b = static_cast<size_t>(m[j]);
                const double ex = h.rx[a] - h.rx[b], ey = h.ry[a] - h.ry[b], ez = h.rz[a] - h.rz[b];
                const double dist = std::sqrt(ex * ex + ey * ey + ez * ez);
                if (dist <= (1. + kCoordinationEpsilon) * (h.r[a] + h.r[b])) {
                    ovl += 2.0 * ((h.r[a] + h.r[b] - dist) / (h.r[a] + h.r[b]));
                    inter += 2;
                }
            }
        overlapping[id] = inter > 0 ? ovl / static_cast<double>(inter) : 0.0;
        coord[id] = inter > 0 ? static_cast<double>(inter) / static_cast<double>(m.size()) : 0.0;
    }
    // With surface reactions every step's whole update changes the overlaps,
    // which only the snapshots read: they are computed when one is written.
    int refresh_overlaps() {
        const size_t S = N, A = n_ids;
        HostState h;
        for (auto *v : {&h.rx, &h.ry, &h.rz, &h.r}) v->resize(S);
        int rc = mcac_box_download_spheres(box, nullptr, nullptr, nullptr, h.rx.data(), h.ry.data(), h.rz.data());
        if (rc == 0) rc = mcac_box_download_radii(box, h.r.data(), nullptr, nullptr);
        std::vector<int64_t> cnt(A), flat(S);
        if (rc == 0) rc = mcac_box_download_members(box, cnt.data(), flat.data());
        if (rc) return rc;
        h.mem.resize(A);
        size_t off = 0;
        for (size_t a = 0; a < A; ++a) {
            h.mem[a].assign(flat.begin() + static_cast<long>(off), flat.begin() + static_cast<long>(off + cnt[a]));
            off += static_cast<size_t>(cnt[a]);
        }
        for (int id : live.ids()) overlaps(h, static_cast<size_t>(id));
        return 0;
    }

    // Aggregate::update (aggregat.cpp:284-288, 321-483, method none) of one
    // aggregate on the host state: the same list-order sums and operations as
    // the device update, plus update_distances_and_overlapping (:719-764).
    void host_update(HostState &h, size_t id) {
        const std::vector<int64_t> &m = h.mem[id];
        double V = 0.0, S = 0.0, vx = 0.0, vy = 0.0, vz = 0.0, rs = 0.0, brg = 0.0;
        for (int64_t s64 : m) {
            const size_t s = static_cast<size_t>(s64);
            const double v = h.sv[s], r = h.r[s];
            V = V + v;
            S = S + h.ss[s];
            vx = vx + h.rx[s] * v;
            vy = vy + h.ry[s] * v;
            vz = vz + h.rz[s] * v;
            rs = rs + r;
            brg = brg + v * (r * r);
        }
        const double cx = vx / V, cy = vy / V, cz = vz / V;
        double rm = 0.0, arg = 0.0;
        for (int64_t s64 : m) {
            const size_t s = static_cast<size_t>(s64);
            const double dx = h.rx[s] - cx, dy = h.ry[s] - cy, dz = h.rz[s] - cz;
            const double dc = std::sqrt(dx * dx + dy * dy + dz * dz);
            rm = smax(rm, h.r[s] + dc);
            arg = arg + h.sv[s] * (dc * dc);
        }
        overlaps(h, id);
        const size_t s0 = static_cast<size_t>(m[0]);
        const double L = pm.box_length;
        com[id] = {periodic_position(h.x[s0] + cx, L), periodic_position(h.y[s0] + cy, L),
                   periodic_position(h.z[s0] + cz, L)};
        rcm[id] = {cx, cy, cz};
        vol[id] = V;
        surf[id] = S;
        rmax[id] = rm;
        rg[id] = std::sqrt(std::fabs((arg + 3. / 5. * brg) / V));
        nsph[id] = static_cast<int>(m.size());
        set_physics(id, rs);
        if (rmax[id] > maxradius) maxradius = rmax[id];  // aggregat.cpp:277-281
    }
    // Relative positions re-based on the first sphere, the update, and the
    // spheres placed again from the new position (aggregat.cpp:639-652, 606-617).
    void rebase_update_place(HostState &h, size_t id) {
        const std::vector<int64_t> &m = h.mem[id];
        const size_t s0 = static_cast<size_t>(m[0]);
        const double r0x = h.x[s0], r0y = h.y[s0], r0z = h.z[s0];
        for (int64_t s64 : m) {
            const size_t s = static_cast<size_t>(s64);
            h.rx[s] = h.x[s] - r0x;
            h.ry[s] = h.y[s] - r0y;
            h.rz[s] = h.z[s] - r0z;
        }
        host_update(h, id);
        const double px = com[id][0] - rcm[id][0], py = com[id][1] - rcm[id][1], pz = com[id][2] - rcm[id][2];
        for (int64_t s64 : m) {
            const size_t s = static_cast<size_t>(s64);
            h.x[s] = px + h.rx[s];
            h.y[s] = py + h.ry[s];
            h.z[s] = pz + h.rz[s];
        }
    }

    // Oxidation's events of the step on the state read back: spheres at or
    // below rp_min leave their aggregates in list order (Aggregate::
    // croissance_surface and remove_sphere, aggregat.cpp:230-243, 629-653; an
    // aggregate left empty is removed), then aggregates that fell apart split
    // into their connected parts (Aggregate::split, aggregat.cpp:553-628,
    // AggregatList::split, aggregat_list.cpp:413-444): each part, grown from
    // the first unvisited sphere by a stack of discovered spheres scanning the
    // unvisited ones in order, becomes a copy of the aggregate with the next
    // stable id.  The spheres that remain keep their order (SphereList::remove)
    // and the device box is rebuilt from the result.
    int remove_and_split(const std::vector<int> &grown, bool *split, bool *disappear) {
        const size_t S = N, A = n_ids;
        HostState h;
        for (auto *v : {&h.x, &h.y, &h.z, &h.rx, &h.ry, &h.rz, &h.r, &h.sv, &h.ss}) v->resize(S);
        int rc = mcac_box_download_spheres(box, h.x.data(), h.y.data(), h.z.data(), h.rx.data(), h.ry.data(),
                                           h.rz.data());
        if (rc == 0) rc = mcac_box_download_radii(box, h.r.data(), h.sv.data(), h.ss.data());
        std::vector<int64_t> cnt(A), flat(S);
        if (rc == 0) rc = mcac_box_download_members(box, cnt.data(), flat.data());
        std::vector<double> pos_d(3 * A), rcm_d(3 * A);
        if (rc == 0)
            rc = mcac_box_download_aggs(box, pos_d.data(), rcm_d.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
        if (rc) return rc;
        h.mem.resize(A);
        size_t off = 0;
        for (size_t a = 0; a < A; ++a) {
            h.mem[a].assign(flat.begin() + static_cast<long>(off), flat.begin() + static_cast<long>(off + cnt[a]));
            off += static_cast<size_t>(cnt[a]);
            com[a] = {pos_d[3 * a], pos_d[3 * a + 1], pos_d[3 * a + 2]};
            rcm[a] = {rcm_d[3 * a], rcm_d[3 * a + 1], rcm_d[3 * a + 2]};
        }
        std::vector<char> gone(S, 0);
        std::vector<int> checked;
        // removals, aggregate by aggregate in label order
        for (int id : grown) {
            const size_t a = static_cast<size_t>(id);
            if (!(g_flag[a] & 1)) {
                if (g_flag[a] & 2) checked.push_back(id);
                continue;
            }
            std::vector<int64_t> &m = h.mem[a];
            for (size_t i = 0; i < m.size();) {
                const size_t s = static_cast<size_t>(m[i]);
                if (h.r[s] <= pm.rp_min_oxid) {
                    gone[s] = 1;
                    for (int64_t *t : {&track_ms, &track_os})
                        if (*t == static_cast<int64_t>(s)) *t = -1;
                    m.erase(m.begin() + static_cast<long>(i));
                    if (!m.empty()) rebase_update_place(h, a);
                } else {
                    ++i;
                }
            }
            if (m.empty()) *disappear = true;  // AggregatList::croissance_surface removes it
            else checked.push_back(id);
        }
        // splits (only with oxidation), in label order; the parts take the next ids
        const double L = pm.box_length;
        for (int id : (pm.u_sg < 0.0 ? checked : std::vector<int>{})) {
            const size_t a = static_cast<size_t>(id);
            const std::vector<int64_t> m = h.mem[a];
            const size_t n = m.size();
            std::vector<size_t> unvisited(n);
            std::iota(unvisited.begin(), unvisited.end(), size_t(0));
            std::vector<std::vector<size_t>> comps;
            while (!unvisited.empty()) {
                std::vector<size_t> comp{unvisited.front()};
                unvisited.erase(unvisited.begin());
                std::vector<size_t> discovered = comp;
                while (!discovered.empty()) {
                    const size_t s1 = static_cast<size_t>(m[discovered.back()]);
                    discovered.pop_back();
                    for (size_t idx = 0; idx < unvisited.size();) {
                        const size_t s2 = static_cast<size_t>(m[unvisited[idx]]);
                        if (contact({h.x[s1], h.y[s1], h.z[s1]}, h.r[s1], {h.x[s2], h.y[s2], h.z[s2]}, h.r[s2], L)) {
                            comp.push_back(unvisited[idx]);
                            discovered.push_back(unvisited[idx]);
                            unvisited.erase(unvisited.begin() + static_cast<long>(idx));
                        } else {
                            ++idx;
                        }
                    }
                }
                comps.push_back(std::move(comp));
            }
            if (comps.size() < 2) {
                if (g_flag[a] & 1) continue;  // updated after its removals above
                host_update(h, a);            // the device update stands, read back for the physics
                continue;
            }
            *split = true;
            for (const auto &comp : comps) {
                const size_t c = h.mem.size();
                h.mem.emplace_back();
                for (size_t i : comp) h.mem[c].push_back(m[i]);
                grow_ids_arrays(c + 1);
                copy_fields(a, c);  // Aggregate(const Aggregate &): the fields, then its own spheres
                rebase_update_place(h, c);
            }
            h.mem[a].clear();
        }
        return rebuild_after_events(h, gone);
    }

    void grow_ids_arrays(size_t n) {
        if (n <= n_ids) return;
        for (auto *v : {&ts, &lpm, &f_agg, &d_m, &dp, &rg, &rmax, &vol, &surf, &ptime, &dg_over_dp, &overlapping, &coord,
                        &g_vol, &g_surf, &g_rsum, &g_rg, &g_rmax})
            v->resize(n, 0.0);
        com.resize(n, {0, 0, 0});
        rcm.resize(n, {0, 0, 0});
        nsph.resize(n, 0);
        g_flag.resize(n, 0);
        n_ids = n;
    }
    void copy_fields(size_t from, size_t to) {
        for (auto *v : {&ts, &lpm, &f_agg, &d_m, &dp, &rg, &rmax, &vol, &surf, &ptime, &dg_over_dp, &overlapping, &coord})
            (*v)[to] = (*v)[from];
        com[to] = com[from];
        rcm[to] = rcm[from];
        nsph[to] = nsph[from];
    }

    // The device box for the state after removals and splits: the remaining
    // spheres in their order, every id's member list remapped.
    int rebuild_after_events(HostState &h, const std::vector<char> &gone) {
        const size_t S = gone.size();
        std::vector<int64_t> remap(S, -1);
        size_t k = 0;
        for (size_t s = 0; s < S; ++s)
            if (!gone[s]) remap[s] = static_cast<int64_t>(k++);
        const size_t S2 = k;
        HostState c;
        for (auto *v : {&c.x, &c.y, &c.z, &c.rx, &c.ry, &c.rz, &c.r, &c.sv, &c.ss}) v->resize(S2);
        for (size_t s = 0; s < S; ++s) {
            if (gone[s]) continue;
            const size_t t = static_cast<size_t>(remap[s]);
            c.x[t] = h.x[s], c.y[t] = h.y[s], c.z[t] = h.z[s];
            c.rx[t] = h.rx[s], c.ry[t] = h.ry[s], c.rz[t] = h.rz[s];
            c.r[t] = h.r[s], c.sv[t] = h.sv[s], c.ss[t] = h.ss[s];
        }
        for (int64_t *t : {&track_ms, &track_os})
            if (*t >= 0) *t = remap[static_cast<size_t>(*t)];
        track_touch = track_ms >= 0 && track_os >= 0 &&
                      contact({c.x[static_cast<size_t>(track_ms)], c.y[static_cast<size_t>(track_ms)],
                               c.z[static_cast<size_t>(track_ms)]},
                              c.r[static_cast<size_t>(track_ms)],
                              {c.x[static_cast<size_t>(track_os)], c.y[static_cast<size_t>(track_os)],
                               c.z[static_cast<size_t>(track_os)]},
                              c.r[static_cast<size_t>(track_os)], pm.box_length);
        const size_t A2 = h.mem.size();
        std::vector<int64_t> cnt(A2, 0), flat;
        flat.reserve(S2);
        for (size_t a = 0; a < A2; ++a) {
            for (int64_t s : h.mem[a]) flat.push_back(remap[static_cast<size_t>(s)]);
            cnt[a] = static_cast<int64_t>(h.mem[a].size());
            nsph[a] = static_cast<int>(cnt[a]);
        }
        mcac_box_destroy(box);
        box = nullptr;
        N = S2;
        n_ids = A2;
        int rc = create_box();
        if (rc) return rc;
        std::vector<double> pos3(3 * A2), rcm3(3 * A2);
        for (size_t a = 0; a < A2; ++a)
            for (size_t l = 0; l < 3; ++l) {
                pos3[3 * a + l] = com[a][l];
                rcm3[3 * a + l] = rcm[a][l];
            }
        rc = mcac_box_upload(box, static_cast<int64_t>(N), c.x.data(), c.y.data(), c.z.data(), c.r.data(),
                             c.sv.data(), c.ss.data(), c.rx.data(), c.ry.data(), c.rz.data(), static_cast<int64_t>(A2),
                             pos3.data(), rcm3.data(), rmax.data(), vol.data(), surf.data(), cnt.data(), flat.data());
        if (rc) return rc;
        rc = box_modes();
        if (rc) return rc;
        live.reset(A2);
        for (size_t a = 0; a < A2; ++a)
            if (cnt[a] == 0) live.erase(a);
        if (!sphere_r.empty()) sphere_r = c.r;
        return 0;
    }

    // Outputs at the end of a run (calcul.cpp:282-284), once.
    int finalize() {
        if (!finished() || final_written) return 0;
        save_advancement();
        int rc = save_snapshot();
        if (out_spheres) out_spheres->close();
        if (out_aggs) out_aggs->close();
        final_written = true;
        return rc;
    }

    // Device state after a contact: the steps behind it were skipped; forget them
    // (their draws stay queued) and reopen the gate after the merge.
    int drain_skipped() {
        while (!flight.empty()) {
            mcac_contact c{};
            int skipped = 0;
            int rc = mcac_move_result(box, flight.front().ticket, &c, &skipped);
            if (rc) return rc;
            if (!skipped) return MCAC_UNKNOWN_ERROR;  // a step ran past a contact
            flight.pop_front();
        }
        return 0;
    }

    // Reads the oldest queued step.  *done is false when the device skipped it
    // (queued behind a step that closed the gate): then every step still in
    // flight was skipped too, and their draws are issued again.
    int consume(bool *done) {
        *done = false;
        const Pending P = flight.front();
        mcac_contact c{};
        int skipped = 0;
        auto td0 = Clock::now();
        int rc = mcac_move_result(box, P.ticket, &c, &skipped);
        t_device += ms_since(td0);
        t_wait += ms_since(td0);
        if (rc) return rc;
        flight.pop_front();
        if (skipped == 1) return drain_skipped();
        if (skipped == 2) {
            rc = drain_skipped();  // the steps behind it did nothing; issue them again
            if (rc) return rc;
        }
        if (pm.time_to_write(total_events)) {  // calcul.cpp:67-71
            rc = save_snapshot();
            if (rc) return rc;
            save_advancement();
        }
        *done = true;
        const int num = P.num;
        const size_t U = static_cast<size_t>(num);
        const double full = P.full;
        const Vec3 dir = P.draw.dir;
        n_ties += P.tie ? 1 : 0;
        const int64_t num_label = trace_steps ? label_of(num) : -1;
        double deltatemps = max_time_step / pick.total();  // aggregat_list.cpp:54-58
        draws.pop_front();
        q_query += static_cast<double>(c.n_query);
        q_filtered += static_cast<double>(c.n_filtered);
        q_pairs += static_cast<double>(c.n_pairs);
        const bool contact_flag = c.contact != 0;
        const double move = c.move;
        // calcul.cpp:147-153
        deltatemps = deltatemps * (move / full + static_cast<double>(1 - 1));
        ptime[U] += deltatemps;
        pm.time = pm.time + deltatemps;
        // calcul.cpp:155-172: the reactions of every aggregate, or of the moved
        // one over its own time step scaled like the global one (calcul.cpp:108, 148)
        bool split_ev = false, disappear = false;
        const bool indiv = pm.individual_surf_reactions;
        if (pm.with_surface_reactions) {
            const double dt_indiv = ts[U] * (move / full + static_cast<double>(1 - 1));
            track_ms = c.moving_sphere;
            track_os = c.other_sphere;
            track_touch = true;
            rc = react(indiv ? dt_indiv : deltatemps, indiv ? num : -1, &split_ev, &disappear);
            if (rc) return rc;
        }
        bool merged = false;
        mcac_contact cm = c;  // the contact as the merge sees it after the reactions
        bool try_merge = contact_flag && (!indiv || !(split_ev || disappear));  // calcul.cpp:175-181
        if (try_merge && (split_ev || disappear)) {
            cm.moving_sphere = track_ms;
            cm.other_sphere = track_os;
            const bool both = live.alive(static_cast<size_t>(c.moving_agg)) && live.alive(static_cast<size_t>(c.other_agg));
            if (track_ms < 0 || track_os < 0 || !track_touch) {
                try_merge = false;  // AggregatList::merge returns false (aggregat_list.cpp:369-378)
            } else if (!both) {
                // Aggregate::merge refuses a contact whose aggregate is gone and
                // AggregatList::merge throws (aggregat_list.cpp:390-393)
                std::fprintf(stderr, "mcac_sim: MergeError at step %lld\n", static_cast<long long>(step_count));
                return MCAC_MERGE_ERROR;
            }
        }
        if (try_merge) {
            auto tdr0 = Clock::now();
            rc = drain_skipped();
            t_drain += ms_since(tdr0);
            if (rc) return rc;
            mcac_merge_result mr{};
            auto tm0 = Clock::now();
            rc = mcac_merge(box, &cm, &mr);
            t_device += ms_since(tm0);
            t_merge_call += ms_since(tm0);
            if (rc) return rc;
            if (mr.merged) {
                merged = true;
                auto th0 = Clock::now();
                const size_t K = static_cast<size_t>(mr.keeper), R = static_cast<size_t>(mr.removed);
                double newtime = ptime[K] + ptime[R] - pm.time;  // aggregat_list.cpp:387
                double keep_old_ts = ts[K], rem_ts = ts[R];
                nsph[K] = static_cast<int>(mr.n);
                nsph[R] = 0;
                vol[K] = mr.volume;
                surf[K] = mr.surface;
                rg[K] = mr.rg;
                rmax[K] = mr.rmax;
                com[K] = {mr.com[0], mr.com[1], mr.com[2]};
                rcm[K] = {mr.rcm[0], mr.rcm[1], mr.rcm[2]};
                overlapping[K] = mr.overlapping;
                coord[K] = mr.coordination;
                AggregatePhysics ph = aggregate_physics(pm, static_cast<size_t>(mr.n), mr.volume, mr.sphere_volume,
                                                        mr.radius_sum, mr.rg);
                dp[K] = ph.dp;
                f_agg[K] = ph.f_agg;
                d_m[K] = ph.d_m;
                ts[K] = ph.time_step;
                lpm[K] = ph.lpm;
                dg_over_dp[K] = ph.dg_over_dp;
                if (rmax[K] > maxradius) maxradius = rmax[K];  // aggregat.cpp:277-281
                live.erase(R);
                ptime[K] = newtime;
                if (!pm.with_surface_reactions)  // with growth the table is rebuilt below
                    pick.replace(static_cast<int>(K), keep_old_ts, ts[K], static_cast<int>(R), rem_ts);
                ++merges;
                t_merge_host += ms_since(th0);
                if (trace_merges) {
                    mcac_merge_rec m{};
                    m.step = step_count;
                    m.keeper = static_cast<int64_t>(K);
                    m.removed = static_cast<int64_t>(R);
                    m.n = mr.n;
                    m.volume = vol[K];
                    m.surface = surf[K];
                    m.rg = rg[K];
                    m.rmax = rmax[K];
                    m.dp = dp[K];
                    m.f_agg = f_agg[K];
                    m.lpm = lpm[K];
                    m.time_step = ts[K];
                    m.d_m = d_m[K];
                    m.proper_time = ptime[K];
                    m.overlapping = overlapping[K];
                    m.coordination = coord[K];
                    for (int l = 0; l < 3; ++l) {
                        m.com[l] = com[K][static_cast<size_t>(l)];
                        m.rcm[l] = rcm[K][static_cast<size_t>(l)];
                    }
                    merge_recs.push_back(m);
                }
            }
            rc = mcac_release(box);
            if (rc) return rc;
        }
        // calcul.cpp:183-205: after an event every aggregate gets the whole update
        // (with the reactions of every aggregate, those not touched by the event
        // were updated with their growth)
        if (pm.with_surface_reactions && (split_ev || disappear || (indiv && merged)) && live.size() > 0) {
            rc = update_all();
            if (rc) return rc;
        }
        event = merged || split_ev || disappear;  // calcul.cpp:222
        if (event) {
            pm.n_iter_without_event = 0;
            total_events++;
        } else {
            pm.n_iter_without_event++;
        }
        if (event && live.size() > 0) {
            // refresh (aggregat_list.cpp:100-108): the largest time step heads the pick table
            if (pm.with_surface_reactions) {
                const std::vector<int> &ids = live.ids();
                max_time_step = ts[static_cast<size_t>(ids[0])];
                for (int id : ids) max_time_step = smax(ts[static_cast<size_t>(id)], max_time_step);
            } else {
                max_time_step = pick.factor();
            }
            avg_npp = static_cast<double>(N) / static_cast<double>(live.size());
            pm_stale = true;
        }
        if (pm.with_surface_reactions && live.size() > 0) {
            // sort_time_steps(max) at the top of every step (calcul.cpp:99-101),
            // with max refreshed only at events
            auto tp0 = Clock::now();
            pick.build(live.ids(), ts, max_time_step);
            t_pick += ms_since(tp0);
        }
        if (trace_steps) {
            mcac_step_rec r{};
            r.step = step_count;
            r.agg = num;
            r.label = num_label;
            r.dir[0] = dir[0];
            r.dir[1] = dir[1];
            r.dir[2] = dir[2];
            r.lpm = full;
            r.dist = c.distance;
            r.moving_sphere = c.moving_sphere;
            r.other_sphere = c.other_sphere;
            r.other_agg = c.other_agg;
            r.contact = contact_flag;
            r.merged = merged;
            r.dt = deltatemps;
            r.time = pm.time;
            steps.push_back(r);
        }
        step_count++;
        return 0;
    }

    // Runs at most max_steps steps; returns the number completed in *done.
    int run(int64_t max_steps, int64_t *done) {
        int64_t n = 0;
        while (n < max_steps) {
            if (flight.empty() && finished()) break;
            int rc = issue_more(n, max_steps);
            if (rc) return rc;
            if (flight.empty()) break;
            bool stepped = false;
            rc = consume(&stepped);
            if (rc) return rc;
            if (stepped) ++n;
            // a stop met with steps still in flight (STOPCODE) lets those
            // steps, which already ran on the device, complete first
        }
        *done = n;
        // the resident engine grid ends with the call, so that the device is
        // idle for the caller (it is launched again by the next step)
        if (!group && !own_group && engine_blocks > 0) return mcac_box_set_engine(box, engine_blocks);
        return 0;
    }
};

}  // namespace mcac

struct mcac_sim {
    mcac::Simulation s;
};

namespace mcac {

int sim_join_group(mcac_sim *sim, mcac_group *group, int depth, int launches) {
    if (!sim || !group || sim->s.initialized || depth < 1 || depth > 8 || launches < 1 || launches > 2)
        return MCAC_E_ARGUMENT;
    sim->s.group = group;
    sim->s.group_depth = static_cast<size_t>(depth) * static_cast<size_t>(launches);  // steps in flight
    return 0;
}

bool sim_next_move(mcac_sim *sim, int64_t steps_done, int64_t max_steps, GroupMove *move) {
    Simulation &S = sim->s;
    if (!S.initialized || S.finished() || !S.can_issue(steps_done, max_steps)) return false;
    if (S.duplication_due() && S.duplicate() != 0) return false;  // the next consume reports nothing: stop
    Simulation::Pending P = S.pick_move();  // the draw after the ones already in flight
    S.flight.push_back(P);
    S.unticketed += 1;
    move->agg = P.num;
    for (int l = 0; l < 3; ++l) move->dir[l] = P.draw.dir[static_cast<size_t>(l)];
    move->lpm = P.full;
    move->maxradius = S.maxradius;
    return true;
}

mcac_box *sim_box(mcac_sim *sim) { return sim->s.box; }

void sim_issued(mcac_sim *sim, uint64_t ticket) {
    Simulation &S = sim->s;
    S.flight[S.flight.size() - S.unticketed].ticket = ticket;
    S.unticketed -= 1;
}

size_t sim_in_flight(mcac_sim *sim) { return sim->s.flight.size(); }

int sim_consume(mcac_sim *sim, bool *stepped) {
    try {
        return sim->s.consume(stepped);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "mcac_sim: %s\n", e.what());
        return MCAC_UNKNOWN_ERROR;
    }
}

int sim_finalize(mcac_sim *sim) { return sim->s.finalize(); }

}  // namespace mcac

extern "C" {

int mcac_sim_create(const char *ini_text, const char *output_dir, int device, mcac_sim **out, char *err, int errlen) {
    if (!ini_text || !out) return MCAC_E_ARGUMENT;
    auto *h = new mcac_sim();
    try {
        h->s.ini_text = ini_text;
        h->s.pm.load(mcac::IniFile(ini_text));
        if (!h->s.pm.unsupported.empty()) throw std::invalid_argument("unsupported option: " + h->s.pm.unsupported);
    } catch (const std::exception &e) {
        if (err && errlen > 0) std::snprintf(err, static_cast<size_t>(errlen), "%s", e.what());
        delete h;
        return MCAC_INPUT_ERROR;
    }
    h->s.out_dir = output_dir ? output_dir : "";
    h->s.device = device;
    if (const char *e = std::getenv("MCAC_PRECISION")) {  // 0 FP64, 1 FP32 broad phase, 2 with miss count
        const long v = std::strtol(e, nullptr, 10);
        h->s.precision = static_cast<int>(v < 0 ? 0 : (v > 2 ? 2 : v));
    }
    if (const char *e = std::getenv("MCAC_PIPELINE_DEPTH")) {  // steps kept in flight, 1..6
        const long v = std::strtol(e, nullptr, 10);
        h->s.depth = static_cast<size_t>(v < 1 ? 1 : (v > 6 ? 6 : v));
    }
    if (const char *e = std::getenv("MCAC_ROUNDS")) {  // 0: always one launch per step
        h->s.rounds = std::strtol(e, nullptr, 10) != 0;
    }
    if (const char *e = std::getenv("MCAC_ENGINE")) {  // 0: one launch per step; n: resident grid of n blocks
        const long v = std::strtol(e, nullptr, 10);
        h->s.engine_blocks = static_cast<int>(v < 0 ? 0 : (v > 255 ? 255 : v));
    }
    *out = h;
    return 0;
}

int mcac_sim_destroy(mcac_sim *sim) {
    delete sim;
    return 0;
}

int64_t mcac_sim_device_bytes(const char *ini_text) {
    if (!ini_text) return -1;
    try {
        mcac::PhysicalModel pm;
        pm.load(mcac::IniFile(ini_text));
        if (!pm.unsupported.empty()) return -1;
        // the box Simulation::init creates (one sphere and one id per monomer),
        // or with domain duplication the largest one the run can reach: the
        // k-th duplication finds at least 8^k monomers per aggregate, so it
        // happens only below a mean-size limit above 8^k (without that limit,
        // the initial box)
        int64_t dups = 0, scale = 1;
        if (pm.with_domain_duplication && pm.mean_monomere_per_aggregate_limit > 0)
            for (double npp = 8.0; npp < pm.mean_monomere_per_aggregate_limit && dups < 6; npp *= 8.0) {
                dups += 1;
                scale *= 8;
            }
        const int64_t n0 = static_cast<int64_t>(pm.n_monomeres);
        mcac_box_params bp{};
        bp.box_length = (pm.box_length > 0 ? pm.box_length : 1.0) * static_cast<double>(int64_t(1) << dups);
        bp.n_div = static_cast<int64_t>(pm.n_verlet_divisions);
        bp.max_spheres = n0 * scale;
        bp.max_aggregates = n0 + 7 * dups * (n0 / 8);
        bp.volsurf = pm.volsurf;
        return mcac_box_device_bytes(&bp);
    } catch (const std::exception &) {
        return -1;
    }
}

int mcac_sim_init(mcac_sim *sim) {
    if (!sim) return MCAC_E_ARGUMENT;
    try {
        return sim->s.init();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "mcac_sim_init: %s\n", e.what());
        return MCAC_UNKNOWN_ERROR;
    }
}

int mcac_sim_run(mcac_sim *sim, int64_t max_steps, int64_t *done) {
    if (!sim || !sim->s.initialized) return MCAC_E_ARGUMENT;
    int64_t n = 0;
    int rc = 0;
    try {
        rc = sim->s.run(max_steps, &n);
        if (rc == 0) rc = sim->s.finalize();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "mcac_sim_run: %s\n", e.what());
        rc = MCAC_UNKNOWN_ERROR;
    }
    if (done) *done = n;
    return rc;
}

int mcac_sim_finished(mcac_sim *sim) { return sim && sim->s.finished() ? 1 : 0; }

int mcac_sim_set_trace(mcac_sim *sim, int steps, int merges) {
    if (!sim) return MCAC_E_ARGUMENT;
    sim->s.trace_steps = steps != 0;
    sim->s.trace_merges = merges != 0;
    return 0;
}

int64_t mcac_sim_n_steps(mcac_sim *sim) { return sim ? static_cast<int64_t>(sim->s.steps.size()) : 0; }
int64_t mcac_sim_get_steps(mcac_sim *sim, mcac_step_rec *out, int64_t max) {
    int64_t n = std::min<int64_t>(max, static_cast<int64_t>(sim->s.steps.size()));
    std::memcpy(out, sim->s.steps.data(), static_cast<size_t>(n) * sizeof(mcac_step_rec));
    return n;
}
int64_t mcac_sim_n_merges(mcac_sim *sim) { return sim ? static_cast<int64_t>(sim->s.merge_recs.size()) : 0; }
int64_t mcac_sim_get_merges(mcac_sim *sim, mcac_merge_rec *out, int64_t max) {
    int64_t n = std::min<int64_t>(max, static_cast<int64_t>(sim->s.merge_recs.size()));
    std::memcpy(out, sim->s.merge_recs.data(), static_cast<size_t>(n) * sizeof(mcac_merge_rec));
    return n;
}

double mcac_sim_scalar(mcac_sim *sim, const char *name) {
    if (!sim || !name) return std::numeric_limits<double>::quiet_NaN();
    auto &s = sim->s;
    std::string n(name);
    if (n == "box_length") return s.pm.box_length;
    if (n == "time") return s.pm.time;
    if (n == "maxradius") return s.maxradius;
    if (n == "n_agg") return static_cast<double>(s.live.size());
    if (n == "pick_exact_steps") return static_cast<double>(s.pick.exact_steps());
    if (n == "pick_chunks") return static_cast<double>(s.pick.chunks());
    if (n == "pick_rescales") return static_cast<double>(s.pick.rescales());
    if (n == "n_sph") return static_cast<double>(s.N);
    if (n == "avg_npp") return s.avg_npp;
    if (n == "max_time_step") return s.max_time_step;
    if (n == "steps") return static_cast<double>(s.step_count);
    if (n == "merges") return static_cast<double>(s.merges);
    if (n == "events") return static_cast<double>(s.total_events);
    if (n == "query") return s.q_query;
    if (n == "filtered") return s.q_filtered;
    if (n == "pairs") return s.q_pairs;
    if (n == "ties") return s.n_ties;
    if (n == "t_host_pick_ms") return s.t_pick;
    if (n == "t_host_merge_ms") return s.t_merge_host;
    if (n == "t_device_ms") return s.t_device;
    if (n == "t_tie_ms") return s.t_tie;
    if (n == "t_enqueue_ms") return s.t_enqueue;
    if (n == "t_wait_ms") return s.t_wait;
    if (n == "t_merge_call_ms") return s.t_merge_call;
    if (n == "t_drain_ms") return s.t_drain;
    if (n == "t_grow_ms") return s.t_grow;
    if (n == "init_ms") return s.t_init;
    if (n == "total_volume") {
        double t = 0.0;
        for (int id : s.live.ids()) t += s.vol[static_cast<size_t>(id)];
        return t;
    }
    if (n == "volume_fraction") {
        s.refresh_pm();
        return s.pm.volume_fraction;
    }
    return std::numeric_limits<double>::quiet_NaN();
}

int64_t mcac_sim_n_aggs(mcac_sim *sim) { return sim ? static_cast<int64_t>(sim->s.live.size()) : 0; }

int mcac_sim_export_aggs(mcac_sim *sim, int64_t *uid, double *lpm, double *time_step, double *proper_time,
                         double *volume, double *rg, double *rmax, int64_t *n) {
    if (!sim) return MCAC_E_ARGUMENT;
    auto &s = sim->s;
    const std::vector<int> &ids = s.live.ids();
    for (size_t l = 0; l < ids.size(); ++l) {
        size_t id = static_cast<size_t>(ids[l]);
        if (uid) uid[l] = static_cast<int64_t>(id);
        if (lpm) lpm[l] = s.lpm[id];
        if (time_step) time_step[l] = s.ts[id];
        if (proper_time) proper_time[l] = s.ptime[id];
        if (volume) volume[l] = s.vol[id];
        if (rg) rg[l] = s.rg[id];
        if (rmax) rmax[l] = s.rmax[id];
        if (n) n[l] = s.nsph[id];
    }
    return 0;
}

mcac_box *mcac_sim_box(mcac_sim *sim) { return sim ? sim->s.box : nullptr; }

int mcac_write_test_snapshots(const char *prefix, int64_t n, int64_t grids, int64_t per_file) {
    // synthetic grids through the snapshot writer: position k of grid g is
    // (g + k, 2k, 3k) / 8, Radius k / 4, Label k, electric_charge 0
    if (!prefix || n < 1 || grids < 0 || per_file < 1) return MCAC_E_ARGUMENT;
    try {
        mcac::SnapshotWriter w(prefix, static_cast<size_t>(n), static_cast<size_t>(per_file));
        const size_t S = static_cast<size_t>(n);
        std::vector<double> pos3(3 * S), rad(S);
        std::vector<int64_t> lab(S);
        std::vector<int32_t> q(S, 0);
        for (int64_t g = 0; g < grids; ++g) {
            for (size_t k = 0; k < S; ++k) {
                pos3[3 * k] = static_cast<double>(g + static_cast<int64_t>(k)) / 8.0;
                pos3[3 * k + 1] = 2.0 * static_cast<double>(k) / 8.0;
                pos3[3 * k + 2] = 3.0 * static_cast<double>(k) / 8.0;
                rad[k] = static_cast<double>(k) / 4.0;
                lab[k] = static_cast<int64_t>(k);
            }
            w.write("Spheres", 0.5 * static_cast<double>(g), pos3,
                    {{"electric_charge", 1, q.data(), S}, {"Radius", 0, rad.data(), S}, {"Label", 2, lab.data(), S}},
                    {{"L", "1e-06"}, {"N []", std::to_string(n)}});
        }
        w.close();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "mcac_write_test_snapshots: %s\n", e.what());
        return MCAC_IO_ERROR;
    }
    return 0;
}

static int64_t selftest_ties = 0;

int64_t mcac_pick_table_selftest_ties() { return selftest_ties; }

int64_t mcac_pick_table_selftest(int64_t n, int64_t events, int64_t chunk, uint64_t seed, int mode) {
    // random time steps with repeated values, random merges; every cumulative
    // sum and every pick compared with a plain re-sort and sequential sum
    std::mt19937_64 rng(seed);
    std::lognormal_distribution<double> logn(0.0, 0.5);
    std::vector<double> ts(static_cast<size_t>(n));
    for (auto &t : ts) {
        if (mode == 1) t = 1e-7 * static_cast<double>(1 + rng() % 12);  // a few large runs of equal keys
        else if (mode == 2) t = 1e-7 * (1.0 + static_cast<double>(rng() % static_cast<uint64_t>(n)) / n);  // pairs
        else t = (rng() % 7 == 0) ? 1.0 : logn(rng) * 1e-7;  // some exact ties
    }
    std::vector<int> live(static_cast<size_t>(n));
    std::iota(live.begin(), live.end(), 0);
    mcac::PickTable table(static_cast<size_t>(chunk));
    table.build(live, ts);
    int64_t bad = 0;
    auto check = [&]() {
        std::vector<int> order(live);
        std::sort(order.begin(), order.end(), [&ts](int a, int b) {
            return ts[static_cast<size_t>(a)] > ts[static_cast<size_t>(b)] ||
                   (ts[static_cast<size_t>(a)] == ts[static_cast<size_t>(b)] && a < b);
        });
        const double f = ts[static_cast<size_t>(order[0])];
        std::vector<double> cum(order.size());
        for (size_t i = 0; i < order.size(); ++i)
            cum[i] = (i == 0 ? 0.0 : cum[i - 1]) + f / ts[static_cast<size_t>(order[i])];
        cum[0] = f / ts[static_cast<size_t>(order[0])];
        std::vector<double> dts, dcum;
        std::vector<int> did;
        table.dump(dts, did, dcum);
        if (table.factor() != f || table.total() != cum.back() || dcum.size() != cum.size()) return ++bad;
        for (size_t i = 0; i < cum.size(); ++i)
            if (dcum[i] != cum[i] || did[i] != order[i]) return ++bad;
        // reference pick on the label-ordered keys
        std::vector<double> tp_t(live.size());
        for (size_t i = 0; i < live.size(); ++i) tp_t[i] = f / ts[static_cast<size_t>(live[i])];
        std::vector<size_t> idx(live.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::sort(idx.begin(), idx.end(), [&tp_t](size_t i1, size_t i2) { return tp_t[i1] < tp_t[i2]; });
        for (int q = 0; q < 8; ++q) {
            double u = static_cast<double>(rng() % 2147483648ull) / 2147483647.0;
            double val = u * cum.back();
            size_t k = static_cast<size_t>(std::lower_bound(cum.begin(), cum.end(), val) - cum.begin());
            bool tie = false;
            int got = table.pick(u, live, ts, &tie);
            selftest_ties += tie ? 1 : 0;
            if (got != live[idx[k]]) return ++bad;
        }
        return bad;
    };
    check();
    for (int64_t e = 0; e < events && live.size() > 2; ++e) {
        size_t a = rng() % live.size(), b = rng() % live.size();
        if (a == b) continue;
        int keep = std::min(live[a], live[b]), rem = std::max(live[a], live[b]);
        double old_k = ts[static_cast<size_t>(keep)], old_r = ts[static_cast<size_t>(rem)];
        double nt = (rng() % 5 == 0 || mode == 1) ? old_k : old_k * (1.0 + logn(rng) * 0.3);
        if (rng() % 50 == 0) nt = ts[static_cast<size_t>(live[rng() % live.size()])];  // equal to another key
        ts[static_cast<size_t>(keep)] = nt;
        live.erase(std::find(live.begin(), live.end(), rem));
        table.replace(keep, old_k, nt, rem, old_r);
        check();
    }
    return bad;
}

int mcac_physics_scalars(const char *ini_text, double out[8]) {
    if (!ini_text || !out) return MCAC_E_ARGUMENT;
    try {
        mcac::PhysicalModel pm;
        pm.load(mcac::IniFile(ini_text));
        double v[8] = {pm.box_length, pm.gaz_mean_free_path, pm.viscosity, pm.mean_massic_radius,
                       pm.friction_exponnant, pm.u_sg, pm.box_volume, pm.aggregate_concentration};
        for (int i = 0; i < 8; ++i) out[i] = v[i];
        return 0;
    } catch (const std::exception &) {
        return MCAC_INPUT_ERROR;
    }
}

}  // extern "C"
