// Device-side FP64 primitives of the collision path, written so that every
// operation is the same IEEE-754 operation, in the same order, as the
// reference C++ (compiled with -ffp-contract=off on both sides):
//   + - * / sqrt are correctly rounded on gfx950 for FP64 without fast-math;
//   fmod is computed exactly (its result is always representable);
//   std::min / std::max keep libstdc++'s comparison order.
#pragma once
#include <hip/hip_runtime.h>

namespace mcac_dev {

constexpr double kInf = __builtin_huge_val();
constexpr double kContactEpsilon = 1e-28;      // constants.hpp:127
constexpr double kCoordinationEpsilon = 1e-10; // constants.hpp:128
constexpr double kPi = 3.141592653589793;      // atan(1.0) * 4, constants.hpp:113 (the same double)
// Loops that fold a coordinate into the box stop after this many turns; the
// inputs of this path stay within a few box lengths, so the cap is never hit
// for finite data and only guarantees termination on corrupted input.
constexpr int kWrapCap = 64;

__host__ __device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// physical_model.hpp:99-109
__host__ __device__ __forceinline__ double periodic_distance(double d, double L) {
    double p = d;
    double half = 0.5 * L;
    for (int it = 0; it < kWrapCap && p < -half; ++it) p += L;
    for (int it = 0; it < kWrapCap && p >= half; ++it) p -= L;
    return p;
}
// physical_model.hpp:113-122
__host__ __device__ __forceinline__ double periodic_position(double x, double L) {
    double p = x;
    for (int it = 0; it < kWrapCap && p < 0; ++it) p += L;
    for (int it = 0; it < kWrapCap && p >= L; ++it) p -= L;
    return p;
}
// Exact fmod(x, y) for y > 0: x - q*y with q = trunc(x/y) corrected by one
// step when the rounded quotient is off; fma gives the exact difference.
__device__ __forceinline__ double exact_fmod(double x, double y) {
    if (!(x == x) || !(y == y) || y == 0.0 || __builtin_isinf(x)) return __builtin_nan("");
    if (__builtin_fabs(x) < y) return x;
    double q = __builtin_trunc(x / y);
    double r = __builtin_fma(-q, y, x);
    if (x >= 0.0) {
        if (r < 0.0) r = __builtin_fma(-(q - 1.0), y, x);
        else if (r >= y) r = __builtin_fma(-(q + 1.0), y, x);
    } else {
        if (r > 0.0) r = __builtin_fma(-(q + 1.0), y, x);
        else if (r <= -y) r = __builtin_fma(-(q - 1.0), y, x);
    }
    return r;
}

// sphere_distances.cpp:74-81
__device__ __forceinline__ double distance_2(double ax, double ay, double az, double bx, double by, double bz,
                                             double L) {
    double dx = periodic_distance(ax - bx, L);
    double dy = periodic_distance(ay - by, L);
    double dz = periodic_distance(az - bz, L);
    return dx * dx + dy * dy + dz * dz;
}
// sphere_distances.cpp:87-96
__device__ __forceinline__ bool contact(double ax, double ay, double az, double ra, double bx, double by,
                                        double bz, double rb, double L) {
    double d2 = distance_2(ax, ay, az, bx, by, bz, L);
    double s = ra + rb;
    return (d2 - s * s <= kContactEpsilon);
}

// distance_to_contact(Sphere, Sphere, dir, dist) — sphere_contact.cpp:47-125.
// Sphere 1 sweeps along dir for dist; returns the travel to first contact with
// any periodic image of sphere 2 (0 when already touching, +inf when none).
__device__ __forceinline__ double swept_sphere(double p1x, double p1y, double p1z, double r1, double p2x,
                                               double p2y, double p2z, double r2, double dx, double dy,
                                               double dz, double dist, double L) {
    double dc = r1 + r2;
    double dc2 = dc * dc;
    if (distance_2(p1x, p1y, p1z, p2x, p2y, p2z, L) <= dc2) return 0.;
    const double p1[3] = {p1x, p1y, p1z};
    const double dir[3] = {dx, dy, dz};
    double p2[3] = {p2x, p2y, p2z};
    double tot[3], zone[3];
    int nper[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        tot[l] = dir[l] * dist;
        double base = smin(p1[l], p1[l] + tot[l]) - dc;
        double end = smax(p1[l], p1[l] + tot[l]) + dc;
        zone[l] = end - base;
        double q = exact_fmod(p2[l] - base, L);
        if (q < 0) q += L;
        p2[l] = q + base;
        double np = __builtin_floor(zone[l] / L);
        // images beyond 7 per axis would need a sweep of 7 box lengths
        nper[l] = np > 7.0 ? 7 : static_cast<int>(np);
    }
    double res = kInf;
    for (int i = 0; i <= nper[0]; ++i) {
        for (int j = 0; j <= nper[1]; ++j) {
            for (int k = 0; k <= nper[2]; ++k) {
                double q0 = p2[0] + static_cast<double>(i) * L;
                double q1 = p2[1] + static_cast<double>(j) * L;
                double q2 = p2[2] + static_cast<double>(k) * L;
                double d0 = q0 - p1[0];
                double d1 = q1 - p1[1];
                double d2 = q2 - p1[2];
                if (__builtin_fabs(d0) > zone[0]) continue;
                if (__builtin_fabs(d1) > zone[1]) continue;
                if (__builtin_fabs(d2) > zone[2]) continue;
                double proj = d0 * dir[0] + d1 * dir[1] + d2 * dir[2];
                if (proj < 0) continue;
                double e0 = (p1[0] + tot[0]) - q0;
                double e1 = (p1[1] + tot[1]) - q1;
                double e2 = (p1[2] + tot[2]) - q2;
                bool end_contact = (e0 * e0 + e1 * e1 + e2 * e2) <= dc2;
                if ((!end_contact) && dist < proj) continue;
                double c0 = d1 * dir[2] - d2 * dir[1];
                double c1 = d2 * dir[0] - d0 * dir[2];
                double c2 = d0 * dir[1] - d1 * dir[0];
                double dist_to_axis = c0 * c0 + c1 * c1 + c2 * c2;
                if (dist_to_axis > dc2) continue;
                res = smin(res, proj - __builtin_sqrt(dc2 - dist_to_axis));
            }
        }
    }
    return res;
}

}  // namespace mcac_dev
