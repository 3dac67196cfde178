#include "physics.hpp"

#include <cctype>
#include <cmath>
#include <cstring>
#include <sstream>
#include <stdexcept>

namespace mcac {

const double kPi = std::atan(1.0) * 4;          // constants.hpp:113
const double kVolumeFactor = 4 * kPi / 3;       // constants.hpp:114
const double kSurfaceFactor = 4 * kPi;          // constants.hpp:115

// ------------------------------------------------------------------ INI
static std::string strip(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

IniFile::IniFile(const std::string &text) {
    std::istringstream in(text);
    std::string line, section;
    while (std::getline(in, line)) {
        line = strip(line);
        if (line.empty() || line[0] == ';' || line[0] == '#') continue;
        if (line.front() == '[' && line.back() == ']') {
            section = strip(line.substr(1, line.size() - 2));
            continue;
        }
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        sections_[section][strip(line.substr(0, eq))] = strip(line.substr(eq + 1));
    }
}

bool IniFile::get(const std::string &sec, const std::string &key, std::string &out) const {
    auto s = sections_.find(sec);
    if (s == sections_.end()) return false;
    auto k = s->second.find(key);
    if (k == s->second.end()) return false;
    out = k->second;
    return true;
}

template <class T> void IniFile::extract(const char *sec, const char *key, T &dst) const {
    std::string v;
    if (!get(sec, key, v)) return;
    std::istringstream is(v);
    T result;
    char c;
    if ((is >> std::boolalpha >> result) && !(is >> c)) dst = result;
}
template <> void IniFile::extract<std::string>(const char *sec, const char *key, std::string &dst) const {
    std::string v;
    if (get(sec, key, v)) dst = v;
}

std::string IniFile::dump() const {
    std::ostringstream os;
    for (const auto &s : sections_) {
        os << "[" << s.first << "]\n";
        for (const auto &kv : s.second) os << kv.first << " = " << kv.second << "\n";
        os << "\n";
    }
    return os.str();
}

// ------------------------------------------------------------------ RNG
void RandomStream::reseed(unsigned seed) {
    std::memset(&rd_, 0, sizeof(rd_));
    std::memset(state_, 0, sizeof(state_));
    initstate_r(seed, state_, sizeof(state_), &rd_);
    srandom_r(seed, &rd_);
}
int32_t RandomStream::next() {
    int32_t v = 0;
    random_r(&rd_, &v);
    return v;
}
double RandomStream::uniform() {
    double v = next();
    v = v / RAND_MAX;
    return v;
}
Vec3 RandomStream::direction() {
    double theta = uniform() * 2 * kPi;
    double phi = std::acos(1 - 2 * uniform());
    return {std::sin(phi) * std::cos(theta), std::sin(phi) * std::sin(theta), std::cos(phi)};
}

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
double inverf(double p) { return inverfc(1. - p); }

// ------------------------------------------------------------------ PhysicalModel
void PhysicalModel::load(const IniFile &ini) {
    std::string s;
    ini.extract("monomers", "number", n_monomeres);
    ini.extract("monomers", "density", density);
    ini.extract("monomers", "dispersion_diameter", dispersion_diameter);
    ini.extract("monomers", "mean_diameter", mean_diameter);
    if (ini.get("monomers", "initialisation_mode", s) && !s.empty()) {
        if (s == "normal") init_mode = InitMode::Normal;
        else if (s == "lognormal") init_mode = InitMode::LogNormal;
        else throw std::invalid_argument("Monomere initialisation mode unknown: " + s);
    }
    ini.extract("environment", "initial_time", time);
    ini.extract("environment", "volume_fraction", volume_fraction);
    ini.extract("environment", "temperature", temperature);
    ini.extract("environment", "pressure", pressure);
    ini.extract("environment", "fractal_prefactor", fractal_prefactor);
    ini.extract("environment", "fractal_dimension", fractal_dimension);
    bool flag = false;
    ini.extract("surface_growth", "with_surface_reactions", with_surface_reactions);
    ini.extract("surface_growth", "flux_surfgrowth", flux_surfgrowth);
    ini.extract("surface_growth", "full_aggregate_update_frequency", full_aggregate_update_frequency);
    ini.extract("numerics", "individual_surf_reactions", individual_surf_reactions);
    ini.extract("oxidation", "rp_min", rp_min_oxid);
    s.clear();
    if (ini.get("surface_growth", "volsurf_method", s) && !s.empty()) {  // constants.hpp:103-111
        if (s == "caps") volsurf = 1;
        else if (s == "alphas") volsurf = 2;
        else if (s == "arvo") volsurf = 3;
        else if (s != "none") unsupported = "volsurf_method=" + s;
    }
    // surface reactions (calcul.cpp:155-205): growth or oxidation (sphere
    // removal, splits), for every aggregate or the moved one, with the whole
    // update every step and no overlap correction; sparser whole updates are
    // not offered
    if (with_surface_reactions && full_aggregate_update_frequency != 1)
        unsupported = "full_aggregate_update_frequency!=1";
    if (with_surface_reactions && volsurf != 0) unsupported = "volsurf_method with surface reactions";
    flag = false;
    ini.extract("nucleation", "with_nucleation", flag);
    if (flag) unsupported = "with_nucleation";
    ini.extract("limits", "number_of_aggregates", number_of_aggregates_limit);
    ini.extract("limits", "n_iter_without_event", n_iter_without_event_limit);
    ini.extract("limits", "physical_time", physical_time_limit);
    ini.extract("limits", "mean_monomere_per_aggregate", mean_monomere_per_aggregate_limit);
    ini.extract("numerics", "with_domain_duplication", with_domain_duplication);
    flag = false;
    ini.extract("numerics", "with_domain_reduction", flag);
    if (flag) unsupported = "with_domain_reduction";
    ini.extract("numerics", "with_collisions", with_collisions);
    ini.extract("numerics", "enforce_volume_fraction", enforce_volume_fraction);
    ini.extract("numerics", "n_verlet_divisions", n_verlet_divisions);
    s.clear();
    if (ini.get("numerics", "pick_method", s) && !s.empty() && s != "random") unsupported = "pick_method=" + s;
    ini.extract("numerics", "random_seed", random_seed);
    if (random_seed < 0) unsupported = "random_seed<0 (clock seeded)";
    flag = false;
    ini.extract("inter_potential", "with_potentials", flag);
    if (flag) unsupported = "with_potentials";
    flag = false;
    ini.extract("flame_coupling", "with_flame_coupling", flag);
    if (flag) unsupported = "with_flame_coupling";
    ini.extract("output", "output_dir", output_dir);
    ini.extract("output", "write_between_event_frequency", write_between_event_frequency);
    ini.extract("output", "write_events_frequency", write_events_frequency);
    ini.extract("output", "write_Delta_t", write_Delta_t);
    ini.extract("output", "n_time_per_file", n_time_per_file);
    derive();
}

// physical_model.cpp:228-269
void PhysicalModel::derive() {
    if (init_mode == InitMode::Normal) {
        box_length = mean_diameter * 1E-9 *
                     std::pow(static_cast<double>(n_monomeres) * kPi / 6. / volume_fraction *
                                  (1. + 3. * std::pow(dispersion_diameter / mean_diameter, 2)),
                              1. / 3.);
        mean_massic_radius = 0.5 * 1E-9 *
                             (std::pow(mean_diameter, 4) +
                              6 * std::pow(mean_diameter, 2) * std::pow(dispersion_diameter, 2) +
                              3 * std::pow(dispersion_diameter, 4)) /
                             (std::pow(mean_diameter, 3) + 3 * mean_diameter * std::pow(dispersion_diameter, 2));
    } else {
        box_length = mean_diameter * 1E-9 *
                     std::pow(static_cast<double>(n_monomeres) * kPi / 6. / volume_fraction *
                                  std::exp(9. / 2. * std::pow(std::log(dispersion_diameter), 2)),
                              1. / 3.);
        mean_massic_radius = 0.5 * mean_diameter * 1E-9 * std::exp(1.5 * std::pow(std::log(dispersion_diameter), 2));
    }
    box_volume = std::pow(box_length, 3);
    update_temperature(temperature);
    u_sg = flux_surfgrowth / density;
    aggregate_concentration = static_cast<double>(n_monomeres) / box_volume;
    monomer_concentration = aggregate_concentration;
    double tot_v = 0, tot_s = 0;
    if (init_mode == InitMode::Normal) {
        double mr = 0.5 * mean_diameter * 1E-9, dr = 0.5 * dispersion_diameter * 1E-9;
        tot_v = static_cast<double>(n_monomeres) * (4.0 * kPi / 3.0) * (mr) * (std::pow(mr, 2) + 3.0 * std::pow(dr, 2));
        tot_s = static_cast<double>(n_monomeres) * (4.0 * kPi) * (std::pow(mr, 2) + std::pow(dr, 2));
    } else {
        double mr = 0.5 * mean_diameter * 1E-9;
        tot_v = static_cast<double>(n_monomeres) * (4.0 * kPi / 3.0) * std::pow(mr, 3) *
                std::exp(4.5 * std::pow(std::log(dispersion_diameter), 2));
        tot_s = static_cast<double>(n_monomeres) * (4.0 * kPi) * std::pow(mr, 2) *
                std::exp(2 * std::pow(std::log(dispersion_diameter), 2));
    }
    total_surface_concent = tot_s / box_volume;
    total_volume_concent = tot_v / box_volume;
}

// physical_model.cpp:536-546
void PhysicalModel::update_temperature(double t) {
    temperature = t;
    viscosity = kViscosityRef * (kSutherland + kTemperatureRef) / (kSutherland + temperature) *
                std::pow(temperature / kTemperatureRef, 1.5);
    gaz_mean_free_path = kLambdaRef * (kPressureRef / pressure) * (temperature / kTemperatureRef) *
                         (1. + kSutherland / kTemperatureRef) / (1. + kSutherland / temperature);
    friction_exponnant = 0.689 * (1. + std::erf(((gaz_mean_free_path / mean_massic_radius) + 4.454) / 10.628));
}
// physical_model.cpp:550-555
double PhysicalModel::cunningham(double r) const {
    double a = 1.142, b = 0.558, c = 0.999;
    return 1.0 + a * gaz_mean_free_path / r + b * gaz_mean_free_path / r * std::exp(-c * r / gaz_mean_free_path);
}
// physical_model.cpp:591-593
double PhysicalModel::friction_exponent(double r) const {
    return 0.689 * (1. + std::erf(((gaz_mean_free_path / r) + 4.454) / 10.628));
}
// physical_model.cpp:594-600
double PhysicalModel::friction_coeff(double av, double sv, double sr) const {
    double fe = friction_exponent(sr);
    double cc = cunningham(sr);
    return (6. * kPi * viscosity * sr / cc) * std::pow(av / sv, fe / fractal_dimension);
}
// friction_coeff and mobility_diameter of one aggregate with their common
// terms (friction exponent, Cunningham factor of the monomer radius, the
// volume-ratio power) evaluated once: the same operations on the same
// arguments, so the same results, with 5 libm calls instead of 8.
void PhysicalModel::friction_and_mobility(double av, double sv, double sr, double *f, double *dm) const {
    const double fe = friction_exponent(sr);
    const double cc = cunningham(sr);
    const double ratio_pow = std::pow(av / sv, fe / fractal_dimension);
    *f = (6. * kPi * viscosity * sr / cc) * ratio_pow;
    const double ar = sr * std::pow(av / sv, fe / fractal_dimension / 2.0);
    const double cc_a = cunningham(ar);
    *dm = (cc_a / cc) * 2.0 * sr * ratio_pow;
}
// physical_model.cpp:607-617
double PhysicalModel::mobility_diameter(double av, double sv, double sr) const {
    double fe = friction_exponent(sr);
    double ar = sr * std::pow(av / sv, fe / fractal_dimension / 2.0);
    double cc_pp = cunningham(sr);
    double cc_a = cunningham(ar);
    return (cc_a / cc_pp) * 2.0 * sr * std::pow(av / sv, fe / fractal_dimension);
}
// physical_model.cpp:557-578 with random_normal (tools.cpp:78-81)
double PhysicalModel::random_diameter(RandomStream &rng) const {
    double d = 0;
    if (init_mode == InitMode::Normal) {
        d = mean_diameter + std::sqrt(2.) * dispersion_diameter * inverf(2. * rng.uniform() - 1.0);
    } else {
        if (dispersion_diameter < 1.0) throw std::invalid_argument("dispersion_diameter cannot be lower than 1");
        d = mean_diameter * std::pow(dispersion_diameter, std::sqrt(2.) * inverf(2. * rng.uniform() - 1.0));
    }
    if (d <= 0) d = mean_diameter;
    return d * 1E-9;
}
// physical_model.cpp:489-498
void PhysicalModel::update(size_t n_agg, size_t n_mono, double tv, double ts) {
    total_volume_concent = tv / box_volume;
    total_surface_concent = ts / box_volume;
    aggregate_concentration = static_cast<double>(n_agg) / box_volume;
    monomer_concentration = static_cast<double>(n_mono) / box_volume;
    volume_fraction = tv / box_volume;
}
// physical_model.cpp:338-356
bool PhysicalModel::time_to_write(size_t total_events) {
    if (write_Delta_t > 0) {
        size_t timestep = static_cast<size_t>(std::floor(time / write_Delta_t));
        if (timestep > last_timestep_written) {
            last_timestep_written = timestep;
            return true;
        }
    }
    if (n_iter_without_event == 0 && total_events % write_events_frequency == 0) return true;
    if (n_iter_without_event > 0 && n_iter_without_event % write_between_event_frequency == 0) return true;
    return false;
}

// aggregat.cpp:261-276
AggregatePhysics aggregate_physics(const PhysicalModel &pm, size_t n, double volume, double sphere_volume,
                                   double radius_sum, double rg) {
    AggregatePhysics o{};
    o.dp = 2 * radius_sum / static_cast<double>(n);
    double vol_pp = sphere_volume / static_cast<double>(n);  // list-order sum of the spheres' own volumes
    pm.friction_and_mobility(volume, vol_pp, 0.5 * o.dp, &o.f_agg, &o.d_m);
    double masse = pm.density * volume;
    double relax_time = masse / o.f_agg;
    o.time_step = 3. * relax_time;
    double diffusivity = pm.diffusivity(o.f_agg);
    o.lpm = std::sqrt(6. * diffusivity * o.time_step);
    o.dg_over_dp = 2 * rg / o.dp;
    return o;
}

}  // namespace mcac
