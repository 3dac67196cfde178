/*
 * ORACLE — test infrastructure only (see mcac_oracle.h).
 *
 * CPU restatement of the reference MC loop for the aggregate-move /
 * collision-search path, keeping the reference's data-structure semantics
 * (label-indexed aggregate rows erased on merge, Verlet cells as ordered
 * sets of labels, multimap candidate order, libstdc++ std::sort pick table,
 * sequential FP64 sums in list order) so that its results, and its cost
 * profile, follow the reference.  Build: -O2 -ffp-contract=off.
 *
 * Pinning (what checks each part; see tests/test_oracle_*.py):
 *   RNG stream, random_direction, inverf/inverfc  -> golden vectors written by
 *       oracle/_ref (reference tools.cpp compiled as-is).
 *   Verlet query order (a4)                       -> golden vectors from
 *       oracle/_ref (reference verlet.cpp compiled as-is).
 *   PhysicalModel derived scalars                 -> known answers of
 *       pymcac/tests/test_read.py:27-52 (L, lambda, Mu, rpeqmass, gamma_, u_sg).
 *   swept-sphere primitive (a9)                   -> closed forms (head-on,
 *       grazing, miss, touching, behind, periodic images, end contact) and a
 *       brute-force ray march.
 *   aggregate-pair argmin (a8), ordered scan (a7) -> brute force on tiny inputs.
 *   merge / update (a10, a11)                     -> closed forms (two touching
 *       spheres: CoM at midpoint, rmax = 2r, Rg^2 = r^2 + 3/5 r^2, V = 2 v).
 *   overlap corrections (a11 caps / alphas)       -> lens volume and cap areas;
 *       the alphas fits against the exact cap fractions of two equal spheres
 *       (its loop-only coefficients: transcription, partly unpinned); the
 *       contact maps are real std::unordered_map objects.
 *   ARVO (a12, restated with fixes F1-F5)         -> lens / caps closed forms,
 *       contained and disjoint spheres, Monte Carlo union volume and areas.
 *   full loop (a1)                                -> invariants (sphere count,
 *       total volume, merge count) and the product's bit-exact agreement.
 * The parts of the reference that need inipp.h / <gsl/gsl> / libunwind.h
 * (physical_model.cpp, the sphere/aggregate classes) cannot be compiled in
 * this image, so they are pinned by the checks above, not by _ref.
 */
#include "mcac_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <list>
#include <map>
#include <numeric>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

// constants.hpp:113-128
const double kPi = std::atan(1.0) * 4;
const double kVolumeFactor = 4 * kPi / 3;
const double kSurfaceFactor = 4 * kPi;
const double kBoltzmann = 1.38066E-23;
const double kLambdaRef = 66.5E-9;
const double kTemperatureRef = 293.15;
const double kSutherland = 110;
const double kPressureRef = 101300;
const double kViscosityRef = 18.203E-6;
const double kContactEpsilon = 1e-28;
const double kCoordinationEpsilon = 1e-10;
const double kInf = std::numeric_limits<double>::infinity();

using V3 = std::array<double, 3>;
inline V3 operator+(const V3 &a, const V3 &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 operator-(const V3 &a, const V3 &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 operator*(const V3 &a, double f) { return {a[0] * f, a[1] * f, a[2] * f}; }
inline V3 operator*(double f, const V3 &a) { return {f * a[0], f * a[1], f * a[2]}; }

// std::min / std::max semantics (libstdc++: min returns (b<a)?b:a, max (a<b)?b:a)
inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double smax(double a, double b) { return (a < b) ? b : a; }

// physical_model.hpp:99-109
inline double periodic_distance(double dist, double dim) {
    double p = dist;
    double half = 0.5 * dim;
    while (p < -half) p += dim;
    while (p >= half) p -= dim;
    return p;
}
// physical_model.hpp:113-122
inline double periodic_position(double pos, double dim) {
    double p = pos;
    while (p < 0) p += dim;
    while (p >= dim) p -= dim;
    return p;
}
inline V3 periodic_position(const V3 &p, double dim) {
    return {periodic_position(p[0], dim), periodic_position(p[1], dim), periodic_position(p[2], dim)};
}

// sphere_distances.cpp:74-81 (distance_2 with periodicity)
inline double distance_2(const V3 &p1, const V3 &p2, double box) {
    V3 d = p1 - p2;
    double dx = periodic_distance(d[0], box);
    double dy = periodic_distance(d[1], box);
    double dz = periodic_distance(d[2], box);
    return dx * dx + dy * dy + dz * dz;
}
// sphere_distances.cpp:82-86
inline double relative_distance_2(const V3 &p1, const V3 &p2) {
    V3 d = p1 - p2;
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}
// sphere_distances.cpp:87-96
inline bool contact(const V3 &p1, double r1, const V3 &p2, double r2, double box) {
    double distance = distance_2(p1, p2, box);
    double s = r1 + r2;
    double dist_contact = s * s;
    return (distance - dist_contact <= kContactEpsilon);
}

// Aggregate::split's component search (aggregat.cpp:553-577): each component
// starts with the first unvisited sphere; a stack of discovered spheres is
// popped from its back, and each popped sphere scans the unvisited ones in
// order, appending every one it touches to the component and to the stack.
template <class Touch> std::vector<std::vector<size_t>> contact_components(size_t n, Touch touch) {
    std::vector<size_t> unvisited(n);
    std::iota(unvisited.begin(), unvisited.end(), size_t(0));
    std::vector<std::vector<size_t>> comps;
    while (!unvisited.empty()) {
        std::vector<size_t> comp{unvisited.front()};
        unvisited.erase(unvisited.begin());
        std::vector<size_t> discovered = comp;
        while (!discovered.empty()) {
            const size_t a1 = discovered.back();
            discovered.pop_back();
            for (size_t idx = 0; idx < unvisited.size();) {
                if (touch(a1, unvisited[idx])) {
                    comp.push_back(unvisited[idx]);
                    discovered.push_back(unvisited[idx]);
                    unvisited.erase(unvisited.begin() + long(idx));
                } else {
                    ++idx;
                }
            }
        }
        comps.push_back(std::move(comp));
    }
    return comps;
}

// sphere_contact.cpp:47-125 — swept-sphere distance to contact
double swept_sphere(const V3 &pos1, double r1, V3 pos2, double r2, const V3 &dir, double dist,
                    double box) {
    double dist_contact = r1 + r2;
    double dist_contact_2 = dist_contact * dist_contact;
    if (distance_2(pos1, pos2, box) <= dist_contact_2) {
        return 0.;
    }
    V3 total = dir * dist;
    int nper[3] = {0, 0, 0};
    double zone[3] = {0., 0., 0.};
    for (int l = 0; l < 3; ++l) {
        double base = smin(pos1[l], pos1[l] + total[l]) - dist_contact;
        double end = smax(pos1[l], pos1[l] + total[l]) + dist_contact;
        zone[l] = end - base;
        pos2[l] = std::fmod((pos2[l] - base), box);
        if (pos2[l] < 0) {
            pos2[l] += box;
        }
        pos2[l] += base;
        nper[l] = static_cast<int>(std::floor(zone[l] / box));
    }
    double res = kInf;
    for (int i = 0; i <= nper[0]; i++) {
        for (int j = 0; j <= nper[1]; j++) {
            for (int k = 0; k <= nper[2]; k++) {
                V3 pos3{pos2[0] + i * box, pos2[1] + j * box, pos2[2] + k * box};
                V3 diff = pos3 - pos1;
                if (std::abs(diff[0]) > zone[0]) continue;
                if (std::abs(diff[1]) > zone[1]) continue;
                if (std::abs(diff[2]) > zone[2]) continue;
                double proj = diff[0] * dir[0] + diff[1] * dir[1] + diff[2] * dir[2];
                if (proj < 0) continue;
                bool end_contact = relative_distance_2(pos1 + total, pos3) <= dist_contact_2;
                if ((!end_contact) && dist < proj) continue;
                V3 cross{diff[1] * dir[2] - diff[2] * dir[1], diff[2] * dir[0] - diff[0] * dir[2],
                         diff[0] * dir[1] - diff[1] * dir[0]};
                double c2 = cross[0] * cross[0] + cross[1] * cross[1] + cross[2] * cross[2];
                if (c2 > dist_contact_2) continue;
                res = smin(res, proj - std::sqrt(dist_contact_2 - c2));
            }
        }
    }
    return res;
}

// glibc TYPE_3 additive feedback generator, private stream (tools.cpp:41-55 uses
// srand/rand; random_r with a 128-byte state is the same generator).
struct Rng {
    struct random_data rd;
    char state[128];
    // random_data points into state[]: never copy an Rng, re-seed it in place.
    explicit Rng(unsigned seed) { reseed(seed); }
    Rng(const Rng &) = delete;
    Rng &operator=(const Rng &) = delete;
    void reseed(unsigned seed) {
        std::memset(&rd, 0, sizeof(rd));
        std::memset(state, 0, sizeof(state));
        initstate_r(seed, state, sizeof(state), &rd);
        srandom_r(seed, &rd);
    }
    int next() {
        int32_t v;
        random_r(&rd, &v);
        return v;
    }
    // tools.cpp:51-55
    double uniform() {
        double v = next();
        v = v / RAND_MAX;
        return v;
    }
    // tools.cpp:82-89
    V3 direction() {
        double theta = uniform() * 2 * kPi;
        double phi = std::acos(1 - 2 * uniform());
        return {std::sin(phi) * std::cos(theta), std::sin(phi) * std::sin(theta), std::cos(phi)};
    }
};

// tools.cpp:56-74
double inverfc(double p) {
    if (p >= 2.) return -100.;
    if (p <= 0.0) return 100.;
    double pp = (p < 1.0) ? p : 2. - p;
    double t = std::sqrt(-2. * std::log(pp / 2.));
    double x = -0.70711 * ((2.30753 + t * 0.27061) / (1. + t * (0.99229 + t * 0.04481)) - t);
    for (int j = 0; j < 2; j++) {
        double err = std::erfc(x) - pp;
        x += err / (1.12837916709551257 * std::exp(-(x * x)) - x * err);
    }
    return (p < 1.0 ? x : -x);
}
// tools.cpp:75-77
double inverf(double p) { return inverfc(1. - p); }

// ---------------------------------------------------------------- INI (inipp semantics)
using Ini = std::map<std::string, std::map<std::string, std::string>>;

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) a++;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) b--;
    return s.substr(a, b - a);
}
Ini parse_ini(const std::string &text) {
    Ini ini;
    std::istringstream is(text);
    std::string line, section;
    while (std::getline(is, line)) {
        line = trim(line);
        if (line.empty() || line[0] == ';' || line[0] == '#') continue;
        if (line.front() == '[' && line.back() == ']') {
            section = trim(line.substr(1, line.size() - 2));
            continue;
        }
        auto eq = line.find('=');
        if (eq == std::string::npos) continue;
        ini[section][trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
    }
    return ini;
}
// inipp::extract: leaves dst untouched on missing key or parse failure
template <class T> void extract(const Ini &ini, const char *sec, const char *key, T &dst) {
    auto s = ini.find(sec);
    if (s == ini.end()) return;
    auto k = s->second.find(key);
    if (k == s->second.end()) return;
    std::istringstream is(k->second);
    T result;
    char c;
    if ((is >> std::boolalpha >> result) && !(is >> c)) dst = result;
}
void extract(const Ini &ini, const char *sec, const char *key, std::string &dst) {
    auto s = ini.find(sec);
    if (s == ini.end()) return;
    auto k = s->second.find(key);
    if (k == s->second.end()) return;
    dst = k->second;
}

enum { LOGNORMAL = 0, NORMAL = 1 };

// PhysicalModel (physical_model.cpp:32-287, 536-617), the parts this path uses.
struct Physics {
    double fractal_dimension = 1.4, fractal_prefactor = 1.8;
    double flux_surfgrowth = 0., u_sg = 0.;
    double pressure = kPressureRef, temperature = kTemperatureRef;
    double gaz_mean_free_path = kLambdaRef, viscosity = kViscosityRef, density = 1800.;
    double mean_diameter = 30., dispersion_diameter = 1.0;
    double mean_massic_radius = 0., friction_exponnant = 0.;
    double time = 0.;
    double volume_fraction = 1e-3, box_length = 0., box_volume = 0.;
    double aggregate_concentration = 0., monomer_concentration = 0.;
    double total_surface_concent = 0., total_volume_concent = 0.;
    size_t n_verlet_divisions = 10;
    size_t n_monomeres = 2500;
    int init_mode = LOGNORMAL;
    int volsurf = 0;  // volsurf_method: 0 none, 1 caps, 2 alphas, 3 arvo (constants.hpp:103-111)
    int pick_random = 1;
    size_t n_iter_without_event = 0;
    double physical_time_limit = -1;
    int mean_monomere_per_aggregate_limit = -1;
    size_t number_of_aggregates_limit = 1;
    int n_iter_without_event_limit = -1;
    int random_seed = -1;
    bool with_domain_duplication = true, with_domain_reduction = false, with_nucleation = false;
    bool with_collisions = true, with_surface_reactions = false, with_flame_coupling = false;
    bool enforce_volume_fraction = true, with_potentials = false;
    // surface reactions (physical_model.cpp:59, 79, 91, 127-159, 264)
    size_t full_aggregate_update_frequency = 1;
    bool individual_surf_reactions = false;
    double rp_min_oxid = 0.166e-09;
    std::string unsupported;

    void load(const Ini &ini) {
        std::string s;
        extract(ini, "monomers", "number", n_monomeres);
        extract(ini, "monomers", "density", density);
        extract(ini, "monomers", "dispersion_diameter", dispersion_diameter);
        extract(ini, "monomers", "mean_diameter", mean_diameter);
        extract(ini, "monomers", "initialisation_mode", s);
        if (s == "normal") init_mode = NORMAL;
        else if (s == "lognormal" || s.empty()) init_mode = LOGNORMAL;
        else throw std::runtime_error("Monomere initialisation mode unknown: " + s);
        extract(ini, "environment", "initial_time", time);
        extract(ini, "environment", "volume_fraction", volume_fraction);
        extract(ini, "environment", "temperature", temperature);
        extract(ini, "environment", "pressure", pressure);
        extract(ini, "environment", "fractal_prefactor", fractal_prefactor);
        extract(ini, "environment", "fractal_dimension", fractal_dimension);
        extract(ini, "surface_growth", "with_surface_reactions", with_surface_reactions);
        extract(ini, "surface_growth", "flux_surfgrowth", flux_surfgrowth);
        extract(ini, "surface_growth", "full_aggregate_update_frequency", full_aggregate_update_frequency);
        extract(ini, "oxidation", "rp_min", rp_min_oxid);
        extract(ini, "numerics", "individual_surf_reactions", individual_surf_reactions);
        s.clear();
        extract(ini, "surface_growth", "volsurf_method", s);
        if (s == "caps") volsurf = 1;
        else if (s == "alphas") volsurf = 2;
        else if (s == "arvo") volsurf = 3;
        else if (!s.empty() && s != "none") unsupported = "volsurf_method=" + s;
        extract(ini, "nucleation", "with_nucleation", with_nucleation);
        extract(ini, "limits", "number_of_aggregates", number_of_aggregates_limit);
        extract(ini, "limits", "n_iter_without_event", n_iter_without_event_limit);
        extract(ini, "limits", "physical_time", physical_time_limit);
        extract(ini, "limits", "mean_monomere_per_aggregate", mean_monomere_per_aggregate_limit);
        extract(ini, "numerics", "with_domain_duplication", with_domain_duplication);
        extract(ini, "numerics", "with_domain_reduction", with_domain_reduction);
        extract(ini, "numerics", "with_collisions", with_collisions);
        extract(ini, "numerics", "enforce_volume_fraction", enforce_volume_fraction);
        extract(ini, "numerics", "n_verlet_divisions", n_verlet_divisions);
        s.clear();
        extract(ini, "numerics", "pick_method", s);
        if (!s.empty() && s != "random") unsupported = "pick_method=" + s;
        extract(ini, "numerics", "random_seed", random_seed);
        extract(ini, "inter_potential", "with_potentials", with_potentials);
        extract(ini, "flame_coupling", "with_flame_coupling", with_flame_coupling);
        // surface growth (flux >= 0) with the whole-box update every step and no
        // overlap correction; oxidation (flux < 0: sphere removal and split),
        // per-aggregate reactions and sparser full updates are not restated
        if (with_surface_reactions && full_aggregate_update_frequency < 1)
            unsupported = "full_aggregate_update_frequency<1";
        if (with_surface_reactions && volsurf != 0) unsupported = "volsurf_method with surface reactions";
        if (with_nucleation) unsupported = "with_nucleation";
        if (with_potentials) unsupported = "with_potentials";
        if (with_flame_coupling) unsupported = "with_flame_coupling";
        if (with_domain_reduction) unsupported = "with_domain_reduction";
        if (random_seed < 0) unsupported = "random_seed<0 (clock seeded)";
        derive();
    }
    // physical_model.cpp:228-269
    void derive() {
        if (init_mode == NORMAL) {
            box_length = mean_diameter * 1E-9 *
                         std::pow(static_cast<double>(n_monomeres) * kPi / 6. / volume_fraction *
                                      (1. + 3. * std::pow(dispersion_diameter / mean_diameter, 2)),
                                  1. / 3.);
            mean_massic_radius =
                0.5 * 1E-9 *
                (std::pow(mean_diameter, 4) + 6 * std::pow(mean_diameter, 2) * std::pow(dispersion_diameter, 2) +
                 3 * std::pow(dispersion_diameter, 4)) /
                (std::pow(mean_diameter, 3) + 3 * mean_diameter * std::pow(dispersion_diameter, 2));
        } else {
            box_length = mean_diameter * 1E-9 *
                         std::pow(static_cast<double>(n_monomeres) * kPi / 6. / volume_fraction *
                                      std::exp(9. / 2. * std::pow(std::log(dispersion_diameter), 2)),
                                  1. / 3.);
            mean_massic_radius =
                0.5 * mean_diameter * 1E-9 * std::exp(1.5 * std::pow(std::log(dispersion_diameter), 2));
        }
        box_volume = std::pow(box_length, 3);
        update_temperature(temperature);
        u_sg = flux_surfgrowth / density;
        aggregate_concentration = static_cast<double>(n_monomeres) / box_volume;
        monomer_concentration = aggregate_concentration;
    }
    // physical_model.cpp:536-546
    void update_temperature(double t) {
        temperature = t;
        viscosity = kViscosityRef * (kSutherland + kTemperatureRef) / (kSutherland + temperature) *
                    std::pow(temperature / kTemperatureRef, 1.5);
        gaz_mean_free_path = kLambdaRef * (kPressureRef / pressure) * (temperature / kTemperatureRef) *
                             (1. + kSutherland / kTemperatureRef) / (1. + kSutherland / temperature);
        friction_exponnant = 0.689 * (1. + std::erf(((gaz_mean_free_path / mean_massic_radius) + 4.454) / 10.628));
    }
    // physical_model.cpp:550-555
    double cunningham(double r) const {
        double a = 1.142, b = 0.558, c = 0.999;
        return 1.0 + a * gaz_mean_free_path / r + b * gaz_mean_free_path / r * std::exp(-c * r / gaz_mean_free_path);
    }
    // physical_model.cpp:591-593
    double friction_exponent(double r) const {
        return 0.689 * (1. + std::erf(((gaz_mean_free_path / r) + 4.454) / 10.628));
    }
    // physical_model.cpp:594-600
    double friction_coeff(double av, double sv, double sr) const {
        double fe = friction_exponent(sr);
        double cc = cunningham(sr);
        return (6. * kPi * viscosity * sr / cc) * std::pow(av / sv, fe / fractal_dimension);
    }
    // physical_model.cpp:607-617
    double mobility_diameter(double av, double sv, double sr) const {
        double fe = friction_exponent(sr);
        double ar = sr * std::pow(av / sv, fe / fractal_dimension / 2.0);
        double cc_pp = cunningham(sr);
        double cc_a = cunningham(ar);
        return (cc_a / cc_pp) * 2.0 * sr * std::pow(av / sv, fe / fractal_dimension);
    }
    // physical_model.cpp:557-578 (random_normal: tools.cpp:78-81)
    double random_diameter(Rng &rng) const {
        double d = 0;
        if (init_mode == NORMAL) {
            d = mean_diameter + std::sqrt(2.) * dispersion_diameter * inverf(2. * rng.uniform() - 1.0);
        } else {
            if (dispersion_diameter < 1.0) throw std::runtime_error("dispersion_diameter cannot be lower than 1");
            d = mean_diameter * std::pow(dispersion_diameter, std::sqrt(2.) * inverf(2. * rng.uniform() - 1.0));
        }
        if (d <= 0) d = mean_diameter;
        return d * 1E-9;
    }
    // physical_model.cpp:489-498
    void update(size_t n_agg, size_t n_mono, double tv, double ts) {
        total_volume_concent = tv / box_volume;
        total_surface_concent = ts / box_volume;
        aggregate_concentration = static_cast<double>(n_agg) / box_volume;
        monomer_concentration = static_cast<double>(n_mono) / box_volume;
        volume_fraction = tv / box_volume;
    }
};

// aggregat.cpp:289-304
double volume_alpha_correction(const double coordination_number, const double c_20, const double c_30,
                               const double min_coordination_number, const double alpha_vs_extreme) {
    double difference_coordination = std::abs(coordination_number - min_coordination_number);
    double correction = 0.25 * (3.0 * c_20 - c_30) * coordination_number -
                        c_30 * difference_coordination * 0.62741833 -
                        pow(difference_coordination, 1.5) * 0.00332425;
    if (correction < 0.0) correction = 1.0;
    correction = std::min(correction, 1.0);
    double alpha_v = 1.0 - correction;
    alpha_v = std::max(alpha_v, alpha_vs_extreme);
    return alpha_v;
}
// aggregat.cpp:305-319
double surface_alpha_correction(const double coordination_number, const double c_10,
                                const double min_coordination_number, const double alpha_vs_extreme) {
    double difference_coordination = std::abs(coordination_number - min_coordination_number);
    double correction = 0.5 * c_10 * coordination_number - std::pow(c_10, 2) * difference_coordination * 0.70132500 -
                        std::pow(difference_coordination, 2) * 0.00450000;
    if (correction < 0.0) correction = 1.0;
    correction = std::min(correction, 1.0);
    double alpha_s = 1.0 - correction;
    alpha_s = std::max(alpha_s, alpha_vs_extreme);
    return alpha_s;
}

// sphere_intersection.cpp:29-66 (spherical caps, Morán et al. 2018)
struct Intersection {
    double volume_1 = 0., volume_2 = 0., surface_1 = 0., surface_2 = 0.;
};
Intersection sphere_intersection(double radius_1, double vol_1, double surf_1, double radius_2, double vol_2,
                                 double surf_2, double dist) {
    Intersection x;
    if (dist <= 0) return x;
    if (dist < radius_1 + radius_2) {
        if (dist >= std::fdim(radius_1, radius_2)) {
            double h_1 = (std::pow(radius_2, 2) - std::pow((radius_1 - dist), 2)) / (2. * dist);
            double h_2 = (std::pow(radius_1, 2) - std::pow((radius_2 - dist), 2)) / (2. * dist);
            x.volume_1 = kPi * std::pow(h_1, 2) * (3 * radius_1 - h_1) / 3.;
            x.volume_2 = kPi * std::pow(h_2, 2) * (3 * radius_2 - h_2) / 3.;
            x.surface_1 = 2 * kPi * radius_1 * h_1;
            x.surface_2 = 2 * kPi * radius_2 * h_2;
        } else if (radius_1 < radius_2) {
            x.volume_1 = vol_1;
            x.surface_1 = surf_1;
        } else {
            x.volume_2 = vol_2;
            x.surface_2 = surf_2;
        }
    }
    return x;
}

// ------------------------------------------------------------------ ARVO (a12)
// Restatement of the ARVO method the reference carries in arvo.cpp (Busa et
// al., Comput. Phys. Commun. 165 (2005) 59-96, C version 2.0): the part of the
// volume and area of a union of balls carried by one ball, from the
// stereographic images (north pole -> plane tangent at the south pole) of its
// neighbours' intersection circles, the boundary arcs of its exposed region,
// and closed-form arc integrals (Green / divergence theorem with the field
// (0, 0, z)).  The parts sum to the union's volume and area.  Fixes to the
// reference (whose copy is never compiled, src/CMakeLists.txt:108-111):
//  F1 positions relative to the aggregate's sphere 0, not wrapped absolute
//     ones (arvo_mcac_call.cpp:28: the z-field split depends on the origin);
//  F2 results in m^3 and m^2 (arvo_mcac_call.cpp:29-32 scales the input to nm
//     and returns nm units that aggregat.cpp then sums as SI);
//  F3 the sphere count is known before the spheres are loaded (arvo.cpp:903-904
//     sets it after LoadProtein_external read it);
//  F4 a ball inside a neighbour carries nothing and a neighbour inside the
//     ball is not a neighbour (MakeNeighbors, arvo.cpp:858-876, flags neither);
//     of two coincident equal balls the lower index keeps its part;
//  F5 map entries that do not overlap by more than the contact tolerance
//     (distance >= (1 - 1e-10)(r_i + r_j); the map keeps pairs up to
//     (1 + 1e-10)(r_i + r_j)) are not neighbours: their image circle
//     degenerates (a square root of a rounding-level negative, NaN) and their
//     caps are far below one ulp of the ball's volume.
// As in the reference the computation runs in nanometres, so that its
// absolute tolerances (north-pole distance, angle merging) keep their scale.
namespace arvo {
const double kPiA = 3.14159265358979323846264;  // arvo.cpp:37
const double kEpsNorthPole = 0.0001, kNorthPoleReduce = 0.9999;  // arvo.cpp:38-39
const double kEpsAngle = 1e-12, kEpsTwoPi = 1e-12;               // arvo.cpp:41-42
const double kVolumeFactor = 4 * kPiA / 3, kSurfaceFactor = 4 * kPiA;

struct Ball {
    double x, y, z, r;
};
struct ImageCircle {
    double t, s, r, o;  // centre, radius, orientation (+1: the north pole is inside the neighbour)
};

// A neighbour whose surface passes within kEpsNorthPole of the ball's north
// pole is shrunk until it does not (arvo.cpp:690-707).
void north_pole_fix(std::vector<Ball> &b) {
    const double nx = b[0].x, ny = b[0].y, nz = b[0].z + b[0].r;
    for (size_t j = 1; j < b.size(); j++) {
        for (;;) {
            const double d = std::fabs(std::sqrt((nx - b[j].x) * (nx - b[j].x) + (ny - b[j].y) * (ny - b[j].y) +
                                                 (nz - b[j].z) * (nz - b[j].z)) - b[j].r);
            if (d >= kEpsNorthPole) break;
            b[j].r = b[j].r * kNorthPoleReduce;
        }
    }
}

// Image of neighbour j's intersection circle (arvo.cpp:651-678), in plane
// coordinates centred under the ball.
ImageCircle image_circle(const Ball &c, const Ball &n) {
    const double r1 = c.r, dx = c.x - n.x, dy = c.y - n.y;
    const double zn = c.z + r1 - n.z, zs = c.z - r1 - n.z;
    const double a = dx * dx + dy * dy + zn * zn - n.r * n.r;
    const double b = 8.0 * r1 * r1 * dx, cc = 8.0 * r1 * r1 * dy;
    const double d = 4.0 * r1 * r1 * (dx * dx + dy * dy + zs * zs - n.r * n.r);
    return {-b / (2.0 * a), -cc / (2.0 * a), std::sqrt((b * b + cc * cc - 4.0 * a * d) / (4.0 * a * a)),
            a > 0 ? -1.0 : 1.0};
}

// Is plane point (t, s) exposed with respect to circle m: inside a positive
// circle or outside a negative one (arvo.cpp:298-315)?
bool exposed(double t, double s, const ImageCircle &m) {
    const double d = std::sqrt((t - m.t) * (t - m.t) + (s - m.s) * (s - m.s));
    return d < m.r ? m.o > 0 : m.o < 0;
}

// Angles on circle a of its two intersection points with circle b, in [0, 2 pi).
void crossing_angles(const ImageCircle &a, const ImageCircle &b, double &p, double &q) {
    const double dx = b.t - a.t, dy = b.s - a.s;
    const double d = std::sqrt(dx * dx + dy * dy);
    const double x = (a.r * a.r - b.r * b.r + d * d) / (2.0 * d);  // along the centre line
    const double h = std::sqrt(std::max(a.r * a.r - x * x, 0.0));
    const double base = std::atan2(dy, dx), w = std::atan2(h, x);
    p = base + w;
    q = base - w;
    while (p < 0) p += 2.0 * kPiA;
    while (q < 0) q += 2.0 * kPiA;
    while (p >= 2.0 * kPiA) p -= 2.0 * kPiA;
    while (q >= 2.0 * kPiA) q -= 2.0 * kPiA;
}

// Closed-form arc integrals (arvo.cpp:149-217): volume and area parts of one
// arc of circle c from angle `start` over the oriented angle `delta`.
void arc_integral(const ImageCircle &c, double start, double delta, double r1, double z1, double &dv, double &da) {
    const double A = (4.0 * r1 * r1 + c.t * c.t + c.s * c.s + c.r * c.r) / 2.0;
    const double B = c.t * c.r, C = c.s * c.r;
    const double S = std::sqrt(A * A - B * B - C * C);
    const double rr = c.r * c.r - A;
    double I1, I2, I3, J1;
    if (std::fabs(std::fabs(delta) - 2.0 * kPiA) < kEpsTwoPi) {
        I1 = 2.0 * kPiA / S;
        I2 = 2.0 * kPiA * A / std::pow(S, 3.);
        I3 = kPiA * (2.0 * A * A + B * B + C * C) / std::pow(S, 5.);
        J1 = kPiA + rr / 2.0 * I1;
    } else {
        const double al = delta < 0 ? start + delta : start;
        const double be = delta < 0 ? start : start + delta;
        auto f = [&](double ang, double k) {
            return (-B * std::sin(ang) + C * std::cos(ang)) / std::pow(A + B * std::cos(ang) + C * std::sin(ang), k);
        };
        I1 = 2.0 * (kPiA / 2.0 - std::atan((A * std::cos((be - al) / 2.0) + B * std::cos((al + be) / 2.0) +
                                            C * std::sin((al + be) / 2.0)) / (S * std::sin((be - al) / 2.0)))) / S;
        I2 = (f(be, 1) - f(al, 1) + A * I1) / (S * S);
        I3 = (f(be, 2) - f(al, 2) + (f(be, 1) - f(al, 1)) / A + (2.0 * A * A + B * B + C * C) * I2 / A) / (2.0 * S * S);
        J1 = ((be - al) + rr * I1) / 2.0;
    }
    const double J2 = (I1 + rr * I2) / 4.0, J3 = (I2 + rr * I3) / 8.0;
    dv = (128.0 * J3 * std::pow(r1, 7.) + 8.0 * J2 * std::pow(r1, 5.) + 2.0 * J1 * std::pow(r1, 3.)) / 3.0 -
         8.0 * std::pow(r1, 4.) * J2 * (z1 + r1);
    da = 2.0 * J1 * r1 * r1;
    if (delta < 0) {
        dv = -dv;
        da = -da;
    }
}

// Part of ball b[0] given its neighbours b[1..] (nm; arvo.cpp:728-788 with
// CirclesToArcs / NewArcs, :497-641, summing each arc as it is found).
void ball_part(std::vector<Ball> b, double &vol, double &area) {
    const double r1 = b[0].r;
    vol = 0.0;
    area = 0.0;
    if (b.size() == 1) {
        vol = kVolumeFactor * r1 * r1 * r1;
        area = kSurfaceFactor * r1 * r1;
        return;
    }
    north_pole_fix(b);
    const size_t k = b.size() - 1;
    std::vector<ImageCircle> c(k);
    for (size_t j = 0; j < k; j++) c[j] = image_circle(b[0], b[j + 1]);
    const double z1 = b[0].z;
    double pv = 0.0, pa = 0.0, dv, da;
    auto add = [&](size_t i, double start, double delta) {
        arc_integral(c[i], start, delta, r1, z1, dv, da);
        pv = pv + dv;
        pa = pa + da;
    };
    for (size_t i = 0; i < k; i++) {
        std::vector<double> ang;
        for (size_t m = 0; m < k; m++) {
            if (m == i) continue;
            const double d = std::sqrt((c[i].t - c[m].t) * (c[i].t - c[m].t) + (c[i].s - c[m].s) * (c[i].s - c[m].s));
            if (d < c[m].r + c[i].r && std::fabs(c[m].r - c[i].r) < d) {
                double p, q;
                crossing_angles(c[i], c[m], p, q);
                ang.push_back(p);
                ang.push_back(q);
            }
        }
        if (ang.empty()) {  // the whole circle, if it lies in every other exposed region (arvo.cpp:526-548)
            bool all = true;
            for (size_t m = 0; m < k && all; m++) {
                if (m == i) continue;
                double pt = c[i].t + c[i].r, ps = c[i].s;
                const double d = std::sqrt((pt - c[m].t) * (pt - c[m].t) + (ps - c[m].s) * (ps - c[m].s));
                if (d == c[m].r) pt = c[i].t;  // that point lies on circle m: test the centre
                all = exposed(pt, ps, c[m]);
            }
            if (all) add(i, 0.0, 2.0 * kPiA * c[i].o);
            continue;
        }
        if (c[i].o > 0) std::sort(ang.begin(), ang.end());
        else std::sort(ang.begin(), ang.end(), std::greater<double>());
        std::vector<double> u{ang[0]};  // merge angles closer than kEpsAngle (arvo.cpp:219-242)
        for (size_t q = 1; q < ang.size(); q++)
            if (std::fabs(ang[q] - u.back()) > kEpsAngle) u.push_back(ang[q]);
        auto mid_exposed = [&](double mid) {
            const double pt = c[i].t + c[i].r * std::cos(mid), ps = c[i].s + c[i].r * std::sin(mid);
            for (size_t m = 0; m < k; m++)
                if (m != i && !exposed(pt, ps, c[m])) return false;
            return true;
        };
        for (size_t q = 0; q + 1 < u.size(); q++)
            if (mid_exposed((u[q] + u[q + 1]) / 2.0)) add(i, u[q], u[q + 1] - u[q]);
        if (mid_exposed((u[0] + 2.0 * kPiA + u.back()) / 2.0))
            add(i, u.back(), u[0] + c[i].o * 2.0 * kPiA - u.back());
    }
    bool positive = false;
    for (const auto &ci : c) positive = positive || ci.o > 0;
    vol = pv;
    area = pa;
    if (!positive) {  // every circle negative: the arcs bound the covered part
        vol = vol + kVolumeFactor * r1 * r1 * r1;
        area = area + kSurfaceFactor * r1 * r1;
    }
}

// Parts of every ball of an aggregate (positions relative to its sphere 0,
// metres) from its contact map, neighbours in the map's order (F1-F5).
void aggregate_parts(const std::vector<V3> &rel, const std::vector<double> &rad,
                     const std::vector<std::unordered_map<size_t, double>> &dist, std::vector<double> &vol,
                     std::vector<double> &area) {
    const size_t n = rel.size();
    vol.assign(n, 0.0);
    area.assign(n, 0.0);
    const double nm = 1e+09;
    for (size_t i = 0; i < n; i++) {
        std::vector<Ball> b{{rel[i][0] * nm, rel[i][1] * nm, rel[i][2] * nm, rad[i] * nm}};
        bool inside = false;
        for (const auto &[j, d] : dist[i]) {
            if (d >= (1. - kCoordinationEpsilon) * (rad[i] + rad[j])) continue;          // F5
            if (d + rad[i] < rad[j] || (d + rad[i] == rad[j] && j < i)) inside = true;  // F4
            if (d + rad[j] < rad[i] || (d + rad[j] == rad[i] && i < j)) continue;
            b.push_back({rel[j][0] * nm, rel[j][1] * nm, rel[j][2] * nm, rad[j] * nm});
        }
        if (inside) continue;
        double v, a;
        ball_part(b, v, a);
        vol[i] = v * 1e-27;  // F2
        area[i] = a * 1e-18;
    }
}
}  // namespace arvo

struct Agg {
    double rg = 0, f_agg = 0, lpm = 0, time_step = 0, rmax = 0, volume = 0, surface = 0;
    V3 pos{0, 0, 0}, rcm{0, 0, 0};
    double proper_time = 0, dp = 0, dg_over_dp = 0, overlapping = 0, coordination = 0, d_m = 0;
    double bulk_density = 0;
    std::vector<size_t> sph;           // myspheres, list order
    std::vector<double> dcenter;       // distances_center
    std::vector<std::unordered_map<size_t, double>> distances;  // aggregat.hpp:61-65 (contacts only)
    std::vector<double> volumes, surfaces;                       // aggregat.hpp:67-68
    std::array<size_t, 3> vidx{0, 0, 0};
    int64_t uid = 0;
};

struct Contact {
    double distance = kInf;
    int64_t ms = -1, os = -1;        // moving / other sphere
    int64_t magg = -1, oagg = -1;    // moving / other aggregate, stable ids
};

struct Sim {
    Physics pm;
    Rng rng{0};
    // SphereList storage (constants.hpp:34-45)
    std::vector<double> sx, sy, sz, sr, svol, ssurf, srx, sry, srz;
    std::vector<long> slabel;
    std::vector<Agg> aggs;  // label order
    // Verlet (verlet.hpp:28)
    size_t n_div = 10;
    double width = 0;
    std::vector<std::set<size_t>> grid;
    double maxradius = 0, avg_npp = 1, max_time_step = 0;
    std::vector<size_t> index_sorted;
    std::vector<double> cumulative;
    // loop state (calcul.cpp:57-61)
    bool event = true;
    size_t total_events = 0;
    size_t dup_threshold = 0;  // duplication when an event leaves at most this many aggregates
    int64_t next_uid = 0;      // stable id of the next aggregate created
    int64_t step_count = 0;
    bool finished_flag = false;
    bool trace_steps = true, trace_merges = true;
    std::vector<ora_step_rec> steps;
    std::vector<ora_merge_rec> merges;
    // counters of the last query (work accounting)
    int64_t q_query = 0, q_filtered = 0, q_solved = 0, q_pairs = 0;

    std::set<size_t> &cell(const std::array<size_t, 3> &i) { return grid[(i[0] * n_div + i[1]) * n_div + i[2]]; }
    const std::set<size_t> &cell(size_t i, size_t j, size_t k) const { return grid[(i * n_div + j) * n_div + k]; }
    V3 spos(size_t s) const { return {sx[s], sy[s], sz[s]}; }
    V3 srel(size_t s) const { return {srx[s], sry[s], srz[s]}; }
    void set_spos(size_t s, const V3 &p) { sx[s] = p[0]; sy[s] = p[1]; sz[s] = p[2]; }

    // aggregat.cpp:699-705
    std::array<size_t, 3> compute_index_verlet(const Agg &a) const {
        double step = double(n_div) / pm.box_length;
        return {size_t(std::floor(a.pos[0] * step)), size_t(std::floor(a.pos[1] * step)),
                size_t(std::floor(a.pos[2] * step))};
    }
    // aggregat.cpp:109-118 + 706-717
    void set_position(size_t label, const V3 &p, bool in_verlet = true) {
        Agg &a = aggs[label];
        a.pos = periodic_position(p, pm.box_length);
        if (in_verlet) {
            auto ni = compute_index_verlet(a);
            if (ni != a.vidx) {
                cell(a.vidx).erase(label);
                a.vidx = ni;
                cell(ni).insert(label);
            }
        }
    }

    // verlet.cpp:52-98
    std::vector<size_t> verlet_query(const V3 &src, const V3 &vec, double distance) const {
        double xp{src[0] + distance + smax(vec[0], 0.)};
        double xm{src[0] - distance + smin(vec[0], 0.)};
        double yp{src[1] + distance + smax(vec[1], 0.)};
        double ym{src[1] - distance + smin(vec[1], 0.)};
        double zp{src[2] + distance + smax(vec[2], 0.)};
        double zm{src[2] - distance + smin(vec[2], 0.)};
        int n = static_cast<int>(n_div);
        int bi1 = static_cast<int>(std::floor(static_cast<double>(n_div) * xm / width));
        int bi2 = static_cast<int>(std::floor(static_cast<double>(n_div) * xp / width) + 1);
        int bj1 = static_cast<int>(std::floor(static_cast<double>(n_div) * ym / width));
        int bj2 = static_cast<int>(std::floor(static_cast<double>(n_div) * yp / width) + 1);
        int bk1 = static_cast<int>(std::floor(static_cast<double>(n_div) * zm / width));
        int bk2 = static_cast<int>(std::floor(static_cast<double>(n_div) * zp / width) + 1);
        if (bi2 - bi1 >= n) { bi1 = 0; bi2 = n - 1; }
        if (bj2 - bj1 >= n) { bj1 = 0; bj2 = n - 1; }
        if (bk2 - bk1 >= n) { bk1 = 0; bk2 = n - 1; }
        std::list<size_t> tmp;
        for (int i = bi1; i <= bi2; i++) {
            for (int j = bj1; j <= bj2; j++) {
                for (int k = bk1; k <= bk2; k++) {
                    auto ii = static_cast<size_t>(periodic_position(i, static_cast<int>(n_div)));
                    auto jj = static_cast<size_t>(periodic_position(j, static_cast<int>(n_div)));
                    auto kk = static_cast<size_t>(periodic_position(k, static_cast<int>(n_div)));
                    const auto &c = cell(ii, jj, kk);
                    tmp.insert(tmp.end(), c.begin(), c.end());
                }
            }
        }
        return std::vector<size_t>(tmp.begin(), tmp.end());
    }

    // ------------------------------------------------------------ aggregate update
    // aggregat.cpp:719-764, with the contact map kept as the reference keeps it
    // (a std::unordered_map per sphere, cleared and re-reserved at its old size)
    // because spherical caps walk it in its iteration order
    void update_distances_and_overlapping(Agg &a) {
        a.overlapping = a.coordination = 0.0;
        int intersections = 0;
        const size_t n = a.sph.size();
        a.distances.resize(n);
        a.dcenter.resize(n);
        for (size_t i = 0; i < n; i++) {
            size_t old_size = a.distances[i].size();
            a.distances[i].clear();
            a.distances[i].reserve(old_size);
        }
        for (size_t i = 0; i < n; i++) {
            size_t si = a.sph[i];
            for (size_t j = i + 1; j < n; j++) {
                size_t sj = a.sph[j];
                double dist = std::sqrt(relative_distance_2(srel(si), srel(sj)));
                if (dist <= (1. + kCoordinationEpsilon) * (sr[si] + sr[sj])) {
                    a.distances[i][j] = dist;
                    a.distances[j][i] = dist;
                    double c_ij = (sr[si] + sr[sj] - dist) / (sr[si] + sr[sj]);
                    a.overlapping += 2.0 * c_ij;
                    intersections += 2;
                }
            }
        }
        if (intersections > 0) {
            a.overlapping /= static_cast<double>(intersections);
            a.coordination = static_cast<double>(intersections) / static_cast<double>(n);
        }
    }
    Intersection intersection(size_t s1, size_t s2, double dist) const {
        return sphere_intersection(sr[s1], svol[s1], ssurf[s1], sr[s2], svol[s2], ssurf[s2], dist);
    }
    // aggregat.cpp:321-430 (methods none and caps)
    void compute_volume_surface(Agg &a) {
        a.volume = a.surface = 0.0;
        const size_t n = a.sph.size();
        a.volumes.resize(n);
        a.surfaces.resize(n);
        if (pm.volsurf == 3) {  // :337-347, the restated ARVO (fixes F1-F5 above)
            std::vector<V3> rel(n);
            std::vector<double> rad(n);
            for (size_t i = 0; i < n; i++) {
                rel[i] = srel(a.sph[i]);
                rad[i] = sr[a.sph[i]];
            }
            arvo::aggregate_parts(rel, rad, a.distances, a.volumes, a.surfaces);
        } else {
            for (size_t i = 0; i < n; i++) {
                a.volumes[i] = svol[a.sph[i]];
                a.surfaces[i] = ssurf[a.sph[i]];
            }
        }
        if (pm.volsurf == 1) {  // :354-386, the map walked in its own iteration order
            for (size_t i = 0; i < n; i++) {
                for (const auto &[j, dist] : a.distances[i]) {
                    if (j <= i) continue;
                    Intersection x = intersection(a.sph[i], a.sph[j], dist);
                    a.volumes[i] = a.volumes[i] - x.volume_1;
                    a.surfaces[i] = a.surfaces[i] - x.surface_1;
                    a.volumes[j] = a.volumes[j] - x.volume_2;
                    a.surfaces[j] = a.surfaces[j] - x.surface_2;
                }
                a.volumes[i] = std::max(a.volumes[i], 0.0);
                a.surfaces[i] = std::max(a.surfaces[i], 0.0);
            }
        }
        for (size_t i = 0; i < n; i++) {
            a.volume = a.volume + a.volumes[i];
            a.surface = a.surface + a.surfaces[i];
        }
        if (pm.volsurf == 2) {  // :391-426
            a.overlapping = 0.0;
            double c_v30(0.0), c_v20(0.0), c_s10(0.0);
            double vp_sum(0.0), sp_sum(0.0);
            size_t intersections(0);
            for (size_t i = 0; i < n; i++) {
                for (const auto &it : a.distances[i]) {
                    double radius_1 = sr[a.sph[i]];
                    double radius_2 = sr[a.sph[it.first]];
                    double c_ij = (radius_1 + radius_2 - it.second) / (radius_1 + radius_2);
                    double vp1 = std::pow(radius_1, 3);
                    double vp2 = std::pow(radius_2, 3);
                    double sp1 = std::pow(radius_1, 2);
                    double sp2 = std::pow(radius_2, 2);
                    vp_sum += (vp1 + vp2);
                    sp_sum += (sp1 + sp2);
                    a.overlapping += c_ij;
                    c_s10 += c_ij * (sp1 + sp2);
                    c_v20 += std::pow(c_ij, 2) * (vp1 + vp2);
                    c_v30 += std::pow(c_ij, 3) * (vp1 + vp2);
                }
                intersections += a.distances[i].size();
            }
            if (intersections > 0) {
                c_s10 /= sp_sum;
                c_v20 /= vp_sum;
                c_v30 /= vp_sum;
                double alpha_vs_extreme = 1.0 / static_cast<double>(n);  // aggregat.cpp:539
                double min_coordination_number = 2 * (1.0 - 1.0 / static_cast<double>(n));
                a.overlapping /= static_cast<double>(intersections);
                a.coordination = static_cast<double>(intersections) / static_cast<double>(n);
                a.volume *= volume_alpha_correction(a.coordination, c_v20, c_v30, min_coordination_number,
                                                    alpha_vs_extreme);
                a.surface *= surface_alpha_correction(a.coordination, c_s10, min_coordination_number,
                                                      alpha_vs_extreme);
            }
        }
        if (a.volume <= 0 || a.surface <= 0) throw std::runtime_error("VolSurfError");
    }
    // aggregat.cpp:247-282 (+ 431-483)
    // compute_mass_center / max_radius / giration_radius (aggregat.cpp:431-483)
    // without the position update; returns the centre of mass relative to sphere 0
    V3 geometry(Agg &a) {
        const size_t n = a.sph.size();
        V3 r{0., 0., 0.};
        for (size_t i = 0; i < n; i++) {
            V3 t = srel(a.sph[i]) * a.volumes[i];
            r[0] += t[0];
            r[1] += t[1];
            r[2] += t[2];
        }
        r[0] /= a.volume;
        r[1] /= a.volume;
        r[2] /= a.volume;
        a.dcenter.resize(n);
        for (size_t i = 0; i < n; i++) {
            V3 diff = srel(a.sph[i]) - r;
            a.dcenter[i] = std::sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
        }
        a.rcm = r;
        // compute_max_radius :453-464
        a.rmax = 0.0;
        for (size_t i = 0; i < n; i++) a.rmax = smax(a.rmax, sr[a.sph[i]] + a.dcenter[i]);
        // compute_giration_radius :465-483
        double arg = 0., brg = 0.;
        for (size_t i = 0; i < n; i++) {
            double v = a.volumes[i];
            arg = arg + v * (a.dcenter[i] * a.dcenter[i]);
            brg = brg + v * (sr[a.sph[i]] * sr[a.sph[i]]);
        }
        a.rg = std::sqrt(std::abs((arg + 3. / 5. * brg) / a.volume));
        return r;
    }
    void update_partial(size_t label, bool in_verlet) {
        Agg &a = aggs[label];
        const size_t n = a.sph.size();
        V3 r = geometry(a);
        set_position(label, spos(a.sph[0]) + r, in_verlet);  // :448
        // :254-276
        a.dp = 0.;
        double vol_pp = 0.0;
        for (size_t i = 0; i < n; i++) {
            vol_pp += svol[a.sph[i]];
            a.dp += sr[a.sph[i]];
        }
        a.dp = 2 * a.dp / static_cast<double>(n);
        vol_pp = vol_pp / static_cast<double>(n);
        a.bulk_density = pm.density;  // set_bulk_density without maturity :139-141
        a.f_agg = pm.friction_coeff(a.volume, vol_pp, 0.5 * a.dp);
        a.d_m = pm.mobility_diameter(a.volume, vol_pp, 0.5 * a.dp);
        double masse = a.bulk_density * a.volume;
        double relax_time = masse / a.f_agg;
        a.time_step = 3. * relax_time;
        double diffusivity = kBoltzmann * pm.temperature / a.f_agg;
        a.lpm = std::sqrt(6. * diffusivity * a.time_step);
        a.dg_over_dp = 2 * a.rg / a.dp;
        if (a.rmax > maxradius) maxradius = a.rmax;
    }
    // aggregat.cpp:284-288
    void update(size_t label, bool in_verlet = true) {
        update_distances_and_overlapping(aggs[label]);
        compute_volume_surface(aggs[label]);
        update_partial(label, in_verlet);
    }

    // ------------------------------------------------------------ init
    // aggregat_list.cpp:534-548 + aggregat_distance.cpp:45-58
    bool test_free_space(const V3 &pos, double radius) const {
        std::vector<size_t> nb = verlet_query(pos, {0, 0, 0}, radius + maxradius);
        for (size_t suspect : nb) {
            const Agg &a = aggs[suspect];
            if (!contact(pos, radius, a.pos, a.rmax, pm.box_length)) continue;
            for (size_t s : a.sph)
                if (contact(pos, radius, spos(s), sr[s], pm.box_length)) return false;
        }
        return true;
    }
    // Same predicate with a private hash grid of the placed spheres; only valid
    // while every aggregate is a monomer (initial placement).
    struct HashGrid {
        size_t n = 1;
        double h = 1;
        std::vector<std::vector<size_t>> cells;
    } hg;
    bool test_free_space_fast(const V3 &pos, double radius) const {
        double reach = radius + maxradius;
        double L = pm.box_length;
        // reference Verlet window for exact membership (verlet.cpp:57-69)
        double d = reach;
        int n = static_cast<int>(n_div);
        int b1[3], b2[3];
        for (int l = 0; l < 3; l++) {
            double p = pos[l] + d + 0.0;
            double m = pos[l] - d + 0.0;
            b1[l] = static_cast<int>(std::floor(static_cast<double>(n_div) * m / width));
            b2[l] = static_cast<int>(std::floor(static_cast<double>(n_div) * p / width) + 1);
            if (b2[l] - b1[l] >= n) { b1[l] = 0; b2[l] = n - 1; }
        }
        int span = static_cast<int>(std::ceil(reach / hg.h)) + 1;
        int c0[3];
        for (int l = 0; l < 3; l++) c0[l] = static_cast<int>(std::floor(periodic_position(pos[l], L) / hg.h));
        int hn = static_cast<int>(hg.n);
        int lim = std::min(span, (hn - 1) / 2);
        bool full = span > (hn - 1) / 2;
        for (int di = -lim; di <= lim + (full ? (hn - 1 - 2 * lim) : 0); di++)
            for (int dj = -lim; dj <= lim + (full ? (hn - 1 - 2 * lim) : 0); dj++)
                for (int dk = -lim; dk <= lim + (full ? (hn - 1 - 2 * lim) : 0); dk++) {
                    size_t ci = static_cast<size_t>(((c0[0] + di) % hn + hn) % hn);
                    size_t cj = static_cast<size_t>(((c0[1] + dj) % hn + hn) % hn);
                    size_t ck = static_cast<size_t>(((c0[2] + dk) % hn + hn) % hn);
                    for (size_t label : hg.cells[(ci * hg.n + cj) * hg.n + ck]) {
                        const Agg &a = aggs[label];
                        bool inside = true;
                        for (int l = 0; l < 3 && inside; l++) {
                            int c = static_cast<int>(a.vidx[l]);
                            int off = ((c - b1[l]) % n + n) % n;
                            inside = off <= b2[l] - b1[l];
                        }
                        if (!inside) continue;
                        if (!contact(pos, radius, a.pos, a.rmax, L)) continue;
                        for (size_t s : a.sph)
                            if (contact(pos, radius, spos(s), sr[s], L)) return false;
                    }
                }
        return true;
    }

    // aggregat_list_storage.cpp:52-88 + aggregat.cpp:162-229
    int init(bool fast) {
        const size_t N = pm.n_monomeres;
        n_div = pm.n_verlet_divisions;
        width = pm.box_length;
        grid.assign(n_div * n_div * n_div, {});
        sx.assign(N, 0.); sy.assign(N, 0.); sz.assign(N, 0.); sr.assign(N, 0.);
        svol.assign(N, 0.); ssurf.assign(N, 0.); srx.assign(N, 0.); sry.assign(N, 0.); srz.assign(N, 0.);
        slabel.assign(N, -1);
        aggs.assign(N, Agg());
        rng.reseed(static_cast<unsigned>(pm.random_seed));
        if (fast) {
            double target = 2e-9 * pm.mean_diameter;  // two mean diameters
            hg.n = std::min<size_t>(128, std::max<size_t>(3, static_cast<size_t>(pm.box_length / target)));
            hg.h = pm.box_length / static_cast<double>(hg.n);
            hg.cells.assign(hg.n * hg.n * hg.n, {});
        }
        for (size_t i = 0; i < N; i++) {
            Agg &a = aggs[i];
            a.uid = static_cast<int64_t>(i);
            a.proper_time = pm.time;
            double diameter = pm.random_diameter(rng);
            bool placed = false;
            for (size_t n_try = 0; n_try < N; n_try++) {
                V3 newpos;
                newpos[0] = rng.uniform() * pm.box_length;
                newpos[1] = rng.uniform() * pm.box_length;
                newpos[2] = rng.uniform() * pm.box_length;
                bool free = fast ? test_free_space_fast(newpos, diameter * 0.5)
                                 : test_free_space(newpos, diameter * 0.5);
                if (free) {
                    // aggregat.cpp:194-229
                    a.pos = periodic_position(newpos, pm.box_length);
                    a.proper_time = pm.time;
                    a.vidx = compute_index_verlet(a);
                    cell(a.vidx).insert(i);
                    slabel[i] = static_cast<long>(i);
                    set_spos(i, newpos);
                    sr[i] = diameter * 0.5;
                    srx[i] = sry[i] = srz[i] = 0.;
                    svol[i] = kVolumeFactor * std::pow(sr[i], 3);
                    ssurf[i] = kSurfaceFactor * std::pow(sr[i], 2);
                    a.sph.assign(1, i);
                    a.d_m = diameter;
                    update(i);
                    if (fast) {
                        size_t c[3];
                        for (int l = 0; l < 3; l++)
                            c[l] = std::min(hg.n - 1, static_cast<size_t>(a.pos[l] / hg.h));
                        hg.cells[(c[0] * hg.n + c[1]) * hg.n + c[2]].push_back(i);
                    }
                    placed = true;
                    break;
                }
            }
            if (!placed) return 6;  // TooDenseError
        }
        refresh();
        if (pm.enforce_volume_fraction) {
            double current = get_total_volume();
            double prescribed = pm.volume_fraction * std::pow(pm.box_length, 3);
            double correction = std::pow(prescribed / current, 1. / 3.);
            for (size_t s = 0; s < N; s++) {
                sr[s] = sr[s] * correction;
                if (slabel[s] > -1) {
                    svol[s] = kVolumeFactor * std::pow(sr[s], 3);
                    ssurf[s] = kSurfaceFactor * std::pow(sr[s], 2);
                }
            }
            for (auto &a : aggs) compute_volume_surface(a);
        }
        hg.cells.clear();
        dup_threshold = aggs.size() / 8;  // calcul.cpp:58, fixed for the run
        next_uid = static_cast<int64_t>(N);
        return 0;
    }

    // AggregatList::duplication (aggregat_list.cpp:142-190): the box doubles
    // per axis and every aggregate, in label order, gets seven copies (all its
    // fields and spheres, aggregat_storage.cpp:122-164, list_storage_methods.hpp:
    // 123-131) appended in (i, j, k) order and translated by (i, j, k) times
    // the old side; the Verlet grid is rebuilt for the new side.  Copies take
    // the next stable ids, in that order.
    void duplication() {
        const size_t old_n_agg = aggs.size();
        const double old_l = pm.box_length;
        pm.box_length *= 2;
        pm.n_monomeres *= 8;
        pm.box_volume = std::pow(pm.box_length, 3);
        for (size_t iagg = 0; iagg < old_n_agg; iagg++)
            for (int i = 0; i <= 1; i++)
                for (int j = 0; j <= 1; j++)
                    for (int k = 0; k <= 1; k++) {
                        if (i == 0 && j == 0 && k == 0) continue;
                        Agg copy = aggs[iagg];
                        const size_t label = aggs.size();
                        copy.uid = next_uid++;
                        copy.sph.clear();
                        for (size_t sp : aggs[iagg].sph) {
                            const size_t ns = sx.size();
                            sx.push_back(sx[sp]); sy.push_back(sy[sp]); sz.push_back(sz[sp]);
                            sr.push_back(sr[sp]); svol.push_back(svol[sp]); ssurf.push_back(ssurf[sp]);
                            srx.push_back(srx[sp]); sry.push_back(sry[sp]); srz.push_back(srz[sp]);
                            slabel.push_back(static_cast<long>(label));
                            copy.sph.push_back(ns);
                        }
                        aggs.push_back(std::move(copy));
                        // Aggregate::translate with no Verlet grid attached (unset_verlet)
                        translate(label, {static_cast<double>(i) * old_l, static_cast<double>(j) * old_l,
                                          static_cast<double>(k) * old_l}, false);
                    }
        width = pm.box_length;  // Verlet(n_verlet_divisions, box_length), set_verlet for every aggregate
        grid.assign(n_div * n_div * n_div, {});
        for (size_t l = 0; l < aggs.size(); l++) {
            aggs[l].vidx = compute_index_verlet(aggs[l]);
            cell(aggs[l].vidx).insert(l);
        }
        pm.update(aggs.size(), sx.size(), get_total_volume(), get_total_surface());
    }

    // aggregat_list.cpp:28-36
    double get_total_volume() const {
        double t = 0.0;
        for (const auto &a : aggs) t += a.volume;
        return t;
    }
    double get_total_surface() const {
        double t = 0.0;
        for (const auto &a : aggs) t += a.surface;
        return t;
    }
    // aggregat_list.cpp:100-108
    void refresh() {
        max_time_step = aggs[0].time_step;
        for (const auto &a : aggs) max_time_step = smax(a.time_step, max_time_step);
        avg_npp = static_cast<double>(sx.size()) / static_cast<double>(aggs.size());
    }
    // aggregat_list.cpp:109-141
    void sort_time_steps(double factor) {
        std::vector<double> tp_t(aggs.size());
        for (size_t i = 0; i < aggs.size(); i++) tp_t[i] = factor / aggs[i].time_step;
        std::vector<size_t> idx(tp_t.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::sort(idx.begin(), idx.end(), [&tp_t](size_t i1, size_t i2) { return tp_t[i1] < tp_t[i2]; });
        index_sorted = std::move(idx);
        cumulative.resize(aggs.size());
        cumulative[0] = tp_t[index_sorted[0]];
        for (size_t i = 1; i < aggs.size(); i++) cumulative[i] = cumulative[i - 1] + tp_t[index_sorted[i]];
    }
    // aggregat_list.cpp:59-66
    size_t pick_random() {
        double val = rng.uniform() * cumulative[aggs.size() - 1];
        long n = std::lower_bound(cumulative.begin(), cumulative.end(), val) - cumulative.begin();
        return index_sorted[static_cast<size_t>(n)];
    }

    // ------------------------------------------------------------ contact search
    // aggregat_distance.cpp:24-44 + sphere_contact.cpp:126-139
    Contact agg_distance(size_t l1, size_t l2, const V3 &dir, double dist) {
        Contact best;
        const Agg &a = aggs[l1];
        const Agg &b = aggs[l2];
        for (size_t i = 0; i < a.sph.size(); i++) {
            size_t si = a.sph[i];
            double hd = kInf;
            int64_t hj = -1;
            for (size_t j = 0; j < b.sph.size(); j++) {
                size_t sj = b.sph[j];
                double d = swept_sphere(spos(si), sr[si], spos(sj), sr[sj], dir, dist, pm.box_length);
                if (d < hd) {
                    hd = d;
                    hj = static_cast<int64_t>(sj);
                }
            }
            if (hd < best.distance) {
                best.distance = hd;
                best.os = hj;
                best.ms = static_cast<int64_t>(si);
                best.magg = a.uid;
                best.oagg = b.uid;
            }
        }
        q_pairs += static_cast<int64_t>(a.sph.size() * b.sph.size());
        return best;
    }
    // aggregat_list.cpp:447-532
    Contact next_contact(size_t source, const V3 &direction, double distance) {
        // get_neighborhood :485-508
        double mindist = aggs[source].rmax + maxradius;
        V3 vec = distance * direction;
        std::vector<size_t> nb = verlet_query(aggs[source].pos, vec, mindist);
        bool found = false;
        for (size_t i = 0; i < nb.size(); i++) {
            if (nb[i] == source) {
                nb.erase(nb.begin() + long(i));
                found = true;
                break;
            }
        }
        if (!found) throw std::runtime_error("VerletError");
        q_query = static_cast<int64_t>(nb.size());
        // filter_neighborhood :510-532
        std::multimap<double, size_t> sorted;
        for (size_t other : nb) {
            double d = swept_sphere(aggs[source].pos, aggs[source].rmax, aggs[other].pos, aggs[other].rmax,
                                    direction, distance, pm.box_length);
            if (d < distance) sorted.insert(std::pair<double, size_t>(d, other));
        }
        q_filtered = static_cast<int64_t>(sorted.size());
        q_solved = 0;
        q_pairs = 0;
        Contact closest;
        for (auto &suspect : sorted) {
            auto [sd, id] = suspect;
            if (closest.distance <= 0.) break;
            if (closest.distance < sd) break;
            Contact pc = agg_distance(source, id, direction, distance);
            q_solved++;
            if (pc.distance < closest.distance) closest = pc;
        }
        return closest;
    }
    // aggregat.cpp:148-161
    void translate(size_t label, const V3 &v, bool in_verlet = true) {
        Agg &a = aggs[label];
        set_position(label, a.pos + v, in_verlet);
        V3 refpos = a.pos - a.rcm;
        set_spos(a.sph[0], refpos);
        for (size_t s : a.sph) set_spos(s, refpos + srel(s));
    }

    // ------------------------------------------------------------ merge
    // aggs stays sorted by uid: erasing a row keeps the others' order
    // (aggregat_list_storage.cpp:37-44), so a label is found by bisection
    size_t label_of_uid(int64_t uid) const {
        auto it = std::lower_bound(aggs.begin(), aggs.end(), uid, [](const Agg &a, int64_t u) { return a.uid < u; });
        if (it == aggs.end() || it->uid != uid) throw std::runtime_error("uid not found");
        return static_cast<size_t>(it - aggs.begin());
    }
    // aggregat_list.cpp:367-410
    int64_t merged_keeper = -1, merged_removed = -1;  // stable ids of the last merge
    bool merge(const Contact &ci) {
        if (ci.ms < 0 || ci.os < 0) return false;
        size_t ms = static_cast<size_t>(ci.ms), os = static_cast<size_t>(ci.os);
        if (!contact(spos(ms), sr[ms], spos(os), sr[os], pm.box_length)) return false;
        // Aggregate::merge (aggregat.cpp:488-493) refuses when an aggregate of the
        // contact is gone (split or vanished since), and AggregatList::merge throws
        auto alive = [this](int64_t uid) {
            auto it = std::lower_bound(aggs.begin(), aggs.end(), uid,
                                       [](const Agg &a, int64_t u) { return a.uid < u; });
            return it != aggs.end() && it->uid == uid;
        };
        if (!alive(ci.magg) || !alive(ci.oagg)) throw std::runtime_error("MergeError");
        size_t keep = static_cast<size_t>(std::min(slabel[ms], slabel[os]));
        size_t rem = static_cast<size_t>(std::max(slabel[ms], slabel[os]));
        merged_keeper = aggs[keep].uid;
        merged_removed = aggs[rem].uid;
        double newtime = aggs[keep].proper_time + aggs[rem].proper_time - pm.time;
        // Aggregate::merge aggregat.cpp:486-544
        {
            Agg &k = aggs[keep];
            size_t mysphere, othersphere;
            if (k.uid == ci.magg) {
                mysphere = ms;
                othersphere = os;
            } else if (k.uid == ci.oagg) {
                mysphere = os;
                othersphere = ms;
            } else {
                throw std::runtime_error("MergeError");
            }
            V3 refpos = spos(k.sph[0]);
            V3 ref_root_to_contact = srel(mysphere);
            V3 dc = spos(othersphere) - spos(mysphere);
            V3 diffcontact{periodic_distance(dc[0], pm.box_length), periodic_distance(dc[1], pm.box_length),
                           periodic_distance(dc[2], pm.box_length)};
            V3 other_root_to_contact = srel(othersphere);
            V3 diffpos = ref_root_to_contact + diffcontact - other_root_to_contact;
            for (size_t s : aggs[rem].sph) {
                slabel[s] = static_cast<long>(keep);
                srx[s] += diffpos[0];
                sry[s] += diffpos[1];
                srz[s] += diffpos[2];
                V3 newpos = srel(s);
                newpos[0] += refpos[0];
                newpos[1] += refpos[1];
                newpos[2] += refpos[2];
                set_spos(s, newpos);
            }
            k.sph.insert(k.sph.end(), aggs[rem].sph.begin(), aggs[rem].sph.end());
            update(keep);
        }
        remove_aggregate(rem);
        aggs[keep].proper_time = newtime;
        return true;
    }

    // AggregatList::remove (aggregat_list_storage.cpp:37-44; ~Aggregate unsets
    // its Verlet entry first): later labels shift down by one, in the Verlet
    // cells and in the spheres' labels
    void remove_aggregate(size_t rem) {
        cell(aggs[rem].vidx).erase(rem);
        aggs.erase(aggs.begin() + long(rem));
        for (size_t i = rem; i < aggs.size(); i++) {
            auto &c = cell(aggs[i].vidx);
            c.erase(i + 1);
            for (size_t s : aggs[i].sph) slabel[s]--;
            c.insert(i);
        }
    }

    // ------------------------------------------------------------ surface reactions
    // spheres of the pending contact, followed through sphere removals (-1: removed)
    int64_t track_ms = -1, track_os = -1;

    // The aggregate's relative positions re-based on its first sphere, the
    // full update, and the spheres placed again from the new position
    // (aggregat.cpp:639-652, 606-617)
    void rebase_update_place(size_t label) {
        const V3 ref0 = spos(aggs[label].sph[0]);
        for (size_t x : aggs[label].sph) {
            const V3 r = spos(x) - ref0;
            srx[x] = r[0];
            sry[x] = r[1];
            srz[x] = r[2];
        }
        update(label);
        const Agg &a = aggs[label];
        const V3 refpos = a.pos - a.rcm;
        for (size_t x : a.sph) set_spos(x, refpos + srel(x));
    }

    // Aggregate::remove_sphere (aggregat.cpp:629-653) with SphereList::remove:
    // the sphere leaves its aggregate and the storage, where later sphere
    // indices shift down by one
    void remove_sphere(size_t label, size_t s) {
        Agg &a = aggs[label];
        a.sph.erase(std::find(a.sph.begin(), a.sph.end(), s));
        for (auto *v : {&sx, &sy, &sz, &sr, &svol, &ssurf, &srx, &sry, &srz}) v->erase(v->begin() + long(s));
        slabel.erase(slabel.begin() + long(s));
        for (auto &b : aggs)
            for (auto &x : b.sph)
                if (x > s) --x;
        for (int64_t *t : {&track_ms, &track_os}) {
            if (*t == static_cast<int64_t>(s)) *t = -1;
            else if (*t > static_cast<int64_t>(s)) --*t;
        }
        if (!aggs[label].sph.empty()) rebase_update_place(label);
    }

    // AggregatList::croissance_surface(dt, index) (aggregat_list.cpp:567-579)
    // and Aggregate::croissance_surface (aggregat.cpp:230-243): the aggregate's
    // spheres grow by u_sg dt (Sphere::croissance_surface, sphere.cpp:113-120;
    // PhysicalModel::grow, physical_model.cpp:587-590), then each sphere at or
    // below rp_min_oxid is removed in list order; true when none is left (the
    // aggregate is removed)
    bool croissance_surface(double dt, size_t l) {
        for (size_t sp : aggs[l].sph) {
            const double new_r = sr[sp] + pm.u_sg * dt;
            const double new_r_2 = new_r * new_r;
            const double new_r_3 = new_r_2 * new_r;
            sr[sp] = new_r;
            svol[sp] = kVolumeFactor * new_r_3;
            ssurf[sp] = kSurfaceFactor * new_r_2;
        }
        for (size_t i = 0; i < aggs[l].sph.size();) {
            const size_t sp = aggs[l].sph[i];
            if (sr[sp] <= pm.rp_min_oxid) remove_sphere(l, sp);
            else ++i;
        }
        if (aggs[l].sph.empty()) {
            remove_aggregate(l);
            return true;
        }
        return false;
    }
    // AggregatList::croissance_surface(dt) (aggregat_list.cpp:549-566)
    bool croissance_surface_all(double dt) {
        bool removed = false;
        for (size_t l = 0; l < aggs.size();) {
            if (croissance_surface(dt, l)) removed = true;  // the next aggregate took label l
            else ++l;
        }
        return removed;
    }

    // Aggregate::split (aggregat.cpp:553-628): the connected components of the
    // spheres' contact graph (sphere_distances.cpp:81-90), each grown from the
    // first unvisited sphere by a stack of discovered spheres that scans the
    // unvisited ones in order; with two or more, each becomes a copy of the
    // aggregate holding its spheres, appended to the list with the next stable
    // id, re-based, updated and placed; AggregatList::split (aggregat_list.cpp:
    // 433-444) then removes the original
    bool split(size_t l) {
        const auto comps = contact_components(aggs[l].sph.size(), [&](size_t a1, size_t a2) {
            const size_t s1 = aggs[l].sph[a1], s2 = aggs[l].sph[a2];
            return contact(spos(s1), sr[s1], spos(s2), sr[s2], pm.box_length);
        });
        if (comps.size() < 2) return false;
        for (const auto &comp : comps) {
            Agg copy = aggs[l];
            copy.uid = next_uid++;
            copy.sph.clear();
            for (size_t i : comp) copy.sph.push_back(aggs[l].sph[i]);
            const size_t label = aggs.size();
            copy.vidx = compute_index_verlet(copy);  // set_verlet at the copied position
            aggs.push_back(std::move(copy));
            cell(aggs[label].vidx).insert(label);
            for (size_t x : aggs[label].sph) slabel[x] = static_cast<long>(label);
            rebase_update_place(label);
        }
        remove_aggregate(l);
        return true;
    }
    // AggregatList::split() (aggregat_list.cpp:413-432)
    bool split_all() {
        bool any = false;
        for (size_t l = 0; l < aggs.size();) {
            if (split(l)) any = true;  // its parts were appended; the next aggregate took label l
            else ++l;
        }
        return any;
    }

    // physical_model.cpp:288-337 (STOPCODE / cpu limits are host-environment checks, not modelled)
    bool finished() const {
        size_t n = aggs.size();
        if (n < 1) return true;
        if (n <= pm.number_of_aggregates_limit) return true;
        if (pm.n_iter_without_event_limit > 0 &&
            pm.n_iter_without_event >= static_cast<size_t>(pm.n_iter_without_event_limit))
            return true;
        if (pm.physical_time_limit > 0 && pm.time >= pm.physical_time_limit) return true;
        if (pm.mean_monomere_per_aggregate_limit > 0 && avg_npp >= pm.mean_monomere_per_aggregate_limit)
            return true;
        return false;
    }

    // calcul.cpp:66-281 (one iteration; surface reactions as calcul.cpp:155-205:
    // growth or oxidation with sphere removal and splits, for every aggregate
    // or the moved one alone, whole or partial updates; no nucleation,
    // potentials or flame coupling)
    int step() {
        // calcul.cpp:72-81
        if (event && pm.with_domain_duplication && aggs.size() <= dup_threshold && !(pm.u_sg < 0.0)) duplication();
        double max = max_time_step;
        if (event || pm.with_surface_reactions) sort_time_steps(max);
        size_t num = pick_random();
        double deltatemps = max / cumulative[aggs.size() - 1];
        double full = aggs[num].lpm;
        int64_t num_uid = aggs[num].uid;
        V3 dir = rng.direction();
        Contact nc;
        bool contact_flag = false;
        double move = full;
        if (pm.with_collisions) {
            nc = next_contact(num, dir, full);
            contact_flag = nc.distance <= full;
            if (contact_flag) move = nc.distance;
        }
        translate(num, dir * move);
        double deltatemps_indiv = aggs[num].time_step;  // calcul.cpp:108
        deltatemps = deltatemps * (move / full + static_cast<double>(1 - 1));
        deltatemps_indiv = deltatemps_indiv * (move / full + static_cast<double>(1 - 1));
        aggs[num].proper_time += deltatemps;
        pm.time = pm.time + deltatemps;
        // calcul.cpp:155-172
        bool split_flag = false, disappear = false;
        track_ms = nc.ms;
        track_os = nc.os;
        if (pm.with_surface_reactions) {
            if (pm.individual_surf_reactions) {
                disappear = croissance_surface(deltatemps_indiv, num);
                if (pm.u_sg < 0.0 && !disappear) split_flag = split(num);
            } else {
                disappear = croissance_surface_all(deltatemps);
                if (pm.u_sg < 0.0) split_flag = split_all();
            }
        }
        // calcul.cpp:175-181 (the contact's spheres as they are now)
        bool merged = false;
        if (contact_flag && (!pm.individual_surf_reactions || !(split_flag || disappear))) {
            Contact now = nc;
            now.ms = track_ms;
            now.os = track_os;
            merged = merge(now);
        }
        // calcul.cpp:183-205
        if (pm.with_surface_reactions) {
            const bool whole = pm.n_iter_without_event % pm.full_aggregate_update_frequency == 0;
            if (pm.individual_surf_reactions && !merged && !split_flag && !disappear) {
                if (whole) update(num);
                else update_partial(num, true);
            } else {
                for (size_t l = 0; l < aggs.size(); l++) {
                    if (whole) update(l);
                    else update_partial(l, true);
                }
            }
        }
        event = merged || split_flag || disappear;  // calcul.cpp:222
        if (event) {
            pm.n_iter_without_event = 0;
            total_events++;
        } else {
            pm.n_iter_without_event++;
        }
        if (event) refresh();
        if (event || pm.with_surface_reactions)
            pm.update(aggs.size(), sx.size(), get_total_volume(), get_total_surface());
        if (trace_steps) {
            ora_step_rec r{};
            r.step = step_count;
            r.agg = num_uid;
            r.label = static_cast<int64_t>(num);
            r.dir[0] = dir[0]; r.dir[1] = dir[1]; r.dir[2] = dir[2];
            r.lpm = full;
            r.dist = nc.distance;
            r.moving_sphere = nc.ms;
            r.other_sphere = nc.os;
            r.other_agg = nc.oagg;
            r.contact = contact_flag;
            r.merged = merged;
            r.dt = deltatemps;
            r.time = pm.time;
            steps.push_back(r);
        }
        if (merged && trace_merges) {
            const Agg &a = aggs[label_of_uid(merged_keeper)];
            ora_merge_rec m{};
            m.step = step_count;
            m.keeper = a.uid;
            m.removed = merged_removed;
            m.n = static_cast<int64_t>(a.sph.size());
            m.volume = a.volume; m.surface = a.surface; m.rg = a.rg; m.rmax = a.rmax; m.dp = a.dp;
            m.f_agg = a.f_agg; m.lpm = a.lpm; m.time_step = a.time_step; m.d_m = a.d_m;
            m.proper_time = a.proper_time; m.overlapping = a.overlapping; m.coordination = a.coordination;
            for (int l = 0; l < 3; l++) { m.com[l] = a.pos[l]; m.rcm[l] = a.rcm[l]; }
            merges.push_back(m);
        }
        step_count++;
        return 0;
    }
};

}  // namespace

extern "C" {

void *ora_create(const char *ini_text, char *err, int errlen) {
    try {
        auto *s = new Sim();
        s->pm.load(parse_ini(ini_text));
        if (!s->pm.unsupported.empty()) {
            std::string m = "unsupported option: " + s->pm.unsupported;
            if (err && errlen > 0) std::snprintf(err, static_cast<size_t>(errlen), "%s", m.c_str());
            delete s;
            return nullptr;
        }
        return s;
    } catch (const std::exception &e) {
        if (err && errlen > 0) std::snprintf(err, static_cast<size_t>(errlen), "%s", e.what());
        return nullptr;
    }
}
void ora_destroy(void *sim) { delete static_cast<Sim *>(sim); }
int ora_init(void *sim, int fast) {
    try {
        return static_cast<Sim *>(sim)->init(fast != 0);
    } catch (const std::exception &) {
        return 1;
    }
}
int64_t ora_run(void *sim, int64_t max_steps) {
    auto *s = static_cast<Sim *>(sim);
    int64_t done = 0;
    try {
        while (done < max_steps) {
            if (s->finished()) {
                s->finished_flag = true;
                break;
            }
            int rc = s->step();
            if (rc != 0) return -rc;
            done++;
        }
        if (s->finished()) s->finished_flag = true;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "oracle: %s at step %lld\n", e.what(), static_cast<long long>(s->step_count));
        return -1;
    }
    return done;
}
int ora_finished(void *sim) { return static_cast<Sim *>(sim)->finished() ? 1 : 0; }
void ora_set_trace(void *sim, int steps, int merges) {
    auto *s = static_cast<Sim *>(sim);
    s->trace_steps = steps != 0;
    s->trace_merges = merges != 0;
}
int64_t ora_n_steps(void *sim) { return static_cast<int64_t>(static_cast<Sim *>(sim)->steps.size()); }
int64_t ora_get_steps(void *sim, ora_step_rec *out, int64_t max) {
    auto *s = static_cast<Sim *>(sim);
    int64_t n = std::min<int64_t>(max, static_cast<int64_t>(s->steps.size()));
    std::memcpy(out, s->steps.data(), static_cast<size_t>(n) * sizeof(ora_step_rec));
    return n;
}
int64_t ora_n_merges(void *sim) { return static_cast<int64_t>(static_cast<Sim *>(sim)->merges.size()); }
int64_t ora_get_merges(void *sim, ora_merge_rec *out, int64_t max) {
    auto *s = static_cast<Sim *>(sim);
    int64_t n = std::min<int64_t>(max, static_cast<int64_t>(s->merges.size()));
    std::memcpy(out, s->merges.data(), static_cast<size_t>(n) * sizeof(ora_merge_rec));
    return n;
}
double ora_get_scalar(void *sim, const char *name) {
    auto *s = static_cast<Sim *>(sim);
    std::string n(name);
    if (n == "box_length") return s->pm.box_length;
    if (n == "viscosity") return s->pm.viscosity;
    if (n == "lambda") return s->pm.gaz_mean_free_path;
    if (n == "rpeqmass") return s->pm.mean_massic_radius;
    if (n == "gamma") return s->pm.friction_exponnant;
    if (n == "u_sg") return s->pm.u_sg;
    if (n == "time") return s->pm.time;
    if (n == "maxradius") return s->maxradius;
    if (n == "n_agg") return static_cast<double>(s->aggs.size());
    if (n == "n_sph") return static_cast<double>(s->sx.size());
    if (n == "avg_npp") return s->avg_npp;
    if (n == "max_time_step") return s->max_time_step;
    if (n == "volume_fraction") return s->pm.volume_fraction;
    if (n == "total_volume") return s->get_total_volume();
    if (n == "steps") return static_cast<double>(s->step_count);
    if (n == "n_iter_without_event") return static_cast<double>(s->pm.n_iter_without_event);
    return std::nan("");
}
int64_t ora_n_spheres(void *sim) { return static_cast<int64_t>(static_cast<Sim *>(sim)->sx.size()); }
int64_t ora_n_aggs(void *sim) { return static_cast<int64_t>(static_cast<Sim *>(sim)->aggs.size()); }
void ora_export_spheres(void *sim, double *x, double *y, double *z, double *r, double *vol, double *surf,
                        double *rx, double *ry, double *rz, int64_t *label) {
    auto *s = static_cast<Sim *>(sim);
    size_t n = s->sx.size();
    for (size_t i = 0; i < n; i++) {
        x[i] = s->sx[i]; y[i] = s->sy[i]; z[i] = s->sz[i]; r[i] = s->sr[i];
        vol[i] = s->svol[i]; surf[i] = s->ssurf[i];
        rx[i] = s->srx[i]; ry[i] = s->sry[i]; rz[i] = s->srz[i];
        label[i] = s->slabel[i];
    }
}
void ora_export_aggs(void *sim, double *pos, double *rcm, double *rmax, double *volume, double *surface,
                     double *rg, double *lpm, double *time_step, double *proper_time, int64_t *n,
                     int64_t *uid, int64_t *offset, int64_t *members) {
    auto *s = static_cast<Sim *>(sim);
    int64_t off = 0;
    for (size_t l = 0; l < s->aggs.size(); l++) {
        const Agg &a = s->aggs[l];
        for (int c = 0; c < 3; c++) {
            pos[3 * l + c] = a.pos[c];
            rcm[3 * l + c] = a.rcm[c];
        }
        rmax[l] = a.rmax; volume[l] = a.volume; surface[l] = a.surface; rg[l] = a.rg;
        lpm[l] = a.lpm; time_step[l] = a.time_step; proper_time[l] = a.proper_time;
        n[l] = static_cast<int64_t>(a.sph.size());
        uid[l] = a.uid;
        offset[l] = off;
        for (size_t q : a.sph) members[off++] = static_cast<int64_t>(q);
    }
}
double ora_next_contact(void *sim, int64_t label, const double dir[3], double lpm, int64_t *moving_sphere,
                        int64_t *other_sphere, int64_t *other_label, int64_t *n_query, int64_t *n_filtered,
                        int64_t *n_solved, int64_t *n_pairs) {
    auto *s = static_cast<Sim *>(sim);
    Contact c = s->next_contact(static_cast<size_t>(label), {dir[0], dir[1], dir[2]}, lpm);
    *moving_sphere = c.ms;
    *other_sphere = c.os;
    *other_label = c.os >= 0 ? s->slabel[static_cast<size_t>(c.os)] : -1;
    if (n_query) *n_query = s->q_query;
    if (n_filtered) *n_filtered = s->q_filtered;
    if (n_solved) *n_solved = s->q_solved;
    if (n_pairs) *n_pairs = s->q_pairs;
    return c.distance;
}
int ora_physics(const char *ini_text, double out[8]) {
    try {
        Physics pm;
        pm.load(parse_ini(ini_text));
        double v[8] = {pm.box_length, pm.gaz_mean_free_path, pm.viscosity, pm.mean_massic_radius,
                       pm.friction_exponnant, pm.u_sg, pm.box_volume, pm.aggregate_concentration};
        for (int i = 0; i < 8; ++i) out[i] = v[i];
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}
double ora_swept_sphere(const double p1[3], double r1, const double p2[3], double r2, const double dir[3],
                        double dist, double box) {
    return swept_sphere({p1[0], p1[1], p1[2]}, r1, {p2[0], p2[1], p2[2]}, r2, {dir[0], dir[1], dir[2]}, dist,
                        box);
}
int ora_contact(const double p1[3], double r1, const double p2[3], double r2, double box) {
    return contact({p1[0], p1[1], p1[2]}, r1, {p2[0], p2[1], p2[2]}, r2, box) ? 1 : 0;
}
// The split's components of n spheres: `order` receives them one after the
// other, `starts` each one's first position (n + 1 entries at most, the last
// one n); returns the number of components.
int64_t ora_contact_components(int64_t n, const double *x, const double *y, const double *z, const double *r,
                               double box, int64_t *order, int64_t *starts) {
    const auto comps = contact_components(static_cast<size_t>(n), [&](size_t a, size_t b) {
        return contact({x[a], y[a], z[a]}, r[a], {x[b], y[b], z[b]}, r[b], box);
    });
    int64_t k = 0;
    for (size_t c = 0; c < comps.size(); ++c) {
        starts[c] = k;
        for (size_t i : comps[c]) order[k++] = static_cast<int64_t>(i);
    }
    starts[comps.size()] = k;
    return static_cast<int64_t>(comps.size());
}
void *ora_rng_create(int seed) { return new Rng(static_cast<unsigned>(seed)); }
void ora_rng_destroy(void *rng) { delete static_cast<Rng *>(rng); }
double ora_rng_uniform(void *rng) { return static_cast<Rng *>(rng)->uniform(); }
void ora_rng_direction(void *rng, double dir[3]) {
    V3 d = static_cast<Rng *>(rng)->direction();
    dir[0] = d[0]; dir[1] = d[1]; dir[2] = d[2];
}
double ora_inverf(double p) { return inverf(p); }
double ora_inverfc(double p) { return inverfc(p); }
int64_t ora_verlet_query(int64_t n_div, double width, const int64_t *cells, int64_t n_labels, const double src[3],
                         const double vec[3], double dist, int64_t *out, int64_t max) {
    Sim s;
    s.n_div = static_cast<size_t>(n_div);
    s.width = width;
    s.grid.assign(s.n_div * s.n_div * s.n_div, {});
    for (int64_t l = 0; l < n_labels; l++) {
        std::array<size_t, 3> c{static_cast<size_t>(cells[3 * l]), static_cast<size_t>(cells[3 * l + 1]),
                                static_cast<size_t>(cells[3 * l + 2])};
        s.cell(c).insert(static_cast<size_t>(l));
    }
    auto v = s.verlet_query({src[0], src[1], src[2]}, {vec[0], vec[1], vec[2]}, dist);
    int64_t n = std::min<int64_t>(max, static_cast<int64_t>(v.size()));
    for (int64_t i = 0; i < n; i++) out[i] = static_cast<int64_t>(v[static_cast<size_t>(i)]);
    return static_cast<int64_t>(v.size());
}

void ora_caps_intersection(double r1, double vol1, double surf1, double r2, double vol2, double surf2, double dist,
                           double out[4]) {
    Intersection x = sphere_intersection(r1, vol1, surf1, r2, vol2, surf2, dist);
    out[0] = x.volume_1; out[1] = x.surface_1; out[2] = x.volume_2; out[3] = x.surface_2;
}
int ora_merge_geometry(int volsurf, int64_t nk, int64_t n, const double *rx, const double *ry, const double *rz,
                       const double *r, const double *vol, const double *surf, double out[11]) {
    if (nk < 1 || n < nk) return -1;
    Sim s;
    s.pm.volsurf = volsurf;
    const size_t N = static_cast<size_t>(n);
    s.srx.assign(rx, rx + N); s.sry.assign(ry, ry + N); s.srz.assign(rz, rz + N);
    s.sx = s.srx; s.sy = s.sry; s.sz = s.srz;
    s.sr.assign(r, r + N); s.svol.assign(vol, vol + N); s.ssurf.assign(surf, surf + N);
    s.slabel.assign(N, 0);
    Agg a;
    try {
        // the keeper's previous update on its own spheres, then the merged update
        for (int64_t q = 0; q < nk; q++) a.sph.push_back(static_cast<size_t>(q));
        s.update_distances_and_overlapping(a);
        s.compute_volume_surface(a);
        for (int64_t q = nk; q < n; q++) a.sph.push_back(static_cast<size_t>(q));
        s.update_distances_and_overlapping(a);
        s.compute_volume_surface(a);
    } catch (const std::exception &) {
        return -2;
    }
    V3 c = s.geometry(a);
    double raw = 0.0, rs = 0.0;
    for (size_t q = 0; q < N; q++) {
        raw += s.svol[q];
        rs += s.sr[q];
    }
    const double v[11] = {a.volume, a.surface, a.rg, a.rmax, c[0], c[1], c[2], a.overlapping, a.coordination, raw, rs};
    for (int q = 0; q < 11; q++) out[q] = v[q];
    return 0;
}
void ora_arvo_part(int k, const double *x, const double *y, const double *z, const double *r, double out[2]) {
    std::vector<arvo::Ball> b(static_cast<size_t>(k) + 1);
    for (int j = 0; j <= k; j++) b[static_cast<size_t>(j)] = {x[j], y[j], z[j], r[j]};
    arvo::ball_part(b, out[0], out[1]);
}
int ora_neighbour_order(int prev, const int64_t *keys, int m, int64_t *out) {
    // aggregat.cpp:730-751 on one map: the previous update's prev entries,
    // then clear(), reserve(old size) and the new keys in ascending order
    std::unordered_map<size_t, double> mp;
    for (int q = 0; q < prev; q++) mp[static_cast<size_t>(7 * q + 3)] = 0.0;
    size_t old_size = mp.size();
    mp.clear();
    mp.reserve(old_size);
    for (int c = 0; c < m; c++) mp[static_cast<size_t>(keys[c])] = 1.0;
    int c = 0;
    for (const auto &kv : mp) out[c++] = static_cast<int64_t>(kv.first);
    return c;
}

}  // extern "C"
