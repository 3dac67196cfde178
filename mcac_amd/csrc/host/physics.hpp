// Host side of the MI355X path: MCAC parameter file, physical model and the
// glibc random stream.  These stay on the host CPU by design (libm pow, erf,
// exp, acos, sin and cos define the reference's values); the device only does
// + - * / sqrt fmod floor.
#pragma once
#include <stdlib.h>

#include <array>
#include <cstdint>
#include <map>
#include <string>

namespace mcac {

// constants.hpp:113-128
extern const double kPi;
extern const double kVolumeFactor;
extern const double kSurfaceFactor;
constexpr double kBoltzmann = 1.38066E-23;
constexpr double kLambdaRef = 66.5E-9;
constexpr double kTemperatureRef = 293.15;
constexpr double kSutherland = 110;
constexpr double kPressureRef = 101300;
constexpr double kViscosityRef = 18.203E-6;
constexpr double kContactEpsilon = 1e-28;
constexpr double kCoordinationEpsilon = 1e-10;  // constants.hpp:128

using Vec3 = std::array<double, 3>;

// MCAC INI file: [section] key = value, ';' comments; a missing key or an
// unparsable value leaves the default (inipp::extract semantics).
class IniFile {
  public:
    explicit IniFile(const std::string &text);
    bool get(const std::string &sec, const std::string &key, std::string &out) const;
    template <class T> void extract(const char *sec, const char *key, T &dst) const;
    std::string dump() const;

  private:
    std::map<std::string, std::map<std::string, std::string>> sections_;
};

// glibc TYPE_3 generator on a private state (same stream as srand(seed)/rand()).
class RandomStream {
  public:
    explicit RandomStream(unsigned seed = 0) { reseed(seed); }
    RandomStream(const RandomStream &) = delete;
    RandomStream &operator=(const RandomStream &) = delete;
    void reseed(unsigned seed);
    int32_t next();
    double uniform();     // tools.cpp:51-55
    Vec3 direction();     // tools.cpp:82-89
  private:
    struct random_data rd_;
    char state_[128];
};

double inverfc(double p);  // tools.cpp:56-74
double inverf(double p);   // tools.cpp:75-77

enum class InitMode { LogNormal, Normal };

// PhysicalModel (physical_model.cpp:32-287, 489-617)
struct PhysicalModel {
    double fractal_dimension = 1.4, fractal_prefactor = 1.8;
    double flux_surfgrowth = 0., u_sg = 0., flux_nucleation = 0.;
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
    InitMode init_mode = InitMode::LogNormal;
    size_t n_iter_without_event = 0;
    size_t last_timestep_written = 0;
    double physical_time_limit = -1, write_Delta_t = -1;
    int mean_monomere_per_aggregate_limit = -1;
    size_t number_of_aggregates_limit = 1;
    int n_iter_without_event_limit = -1;
    int random_seed = -1;
    size_t write_events_frequency = 1, write_between_event_frequency = 100;
    size_t n_time_per_file = 10;
    bool with_domain_duplication = true, with_collisions = true, enforce_volume_fraction = true;
    std::string output_dir = "MCAC_output";
    int volsurf = 0;          // volsurf_method: 0 none, 1 caps (spherical caps), 2 alphas, 3 arvo
    bool with_surface_reactions = false;  // surface growth (physical_model.cpp:79, 91, 127-159)
    size_t full_aggregate_update_frequency = 1;
    bool individual_surf_reactions = false;
    double rp_min_oxid = 0.166e-09;  // physical_model.cpp:59, [oxidation] rp_min
    std::string unsupported;  // first option this path does not implement

    void load(const IniFile &ini);
    void derive();
    void update_temperature(double t);
    double cunningham(double r) const;
    double friction_exponent(double r) const;
    double friction_coeff(double aggregate_volume, double sphere_volume, double sphere_radius) const;
    double mobility_diameter(double aggregate_volume, double sphere_volume, double sphere_radius) const;
    void friction_and_mobility(double aggregate_volume, double sphere_volume, double sphere_radius, double *friction,
                               double *mobility_diam) const;
    double diffusivity(double f_agg) const { return kBoltzmann * temperature / f_agg; }
    double random_diameter(RandomStream &rng) const;
    void update(size_t n_aggregates, size_t n_monomers, double total_volume, double total_surface);
    bool time_to_write(size_t total_events);
};

// Per-aggregate physics of Aggregate::update_partial (aggregat.cpp:254-276)
// from the geometry returned by the device.
struct AggregatePhysics {
    double dp, f_agg, d_m, time_step, lpm, dg_over_dp;
};
AggregatePhysics aggregate_physics(const PhysicalModel &pm, size_t n, double volume, double sphere_volume,
                                   double radius_sum, double rg);

}  // namespace mcac
