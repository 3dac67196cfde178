// Sphere-overlap corrections of an aggregate's volume and surface (a11,
// volsurf_method = caps), shared by the device kernels and host checks.
//
// Spherical caps, aggregat.cpp:354-386: sphere i walks its neighbour map
// distances[i] (std::unordered_map<size_t, double>, aggregat.hpp:61-65) and,
// for every neighbour j > i, subtracts the caps of the pair from both
// spheres; sphere i's running values are clamped at zero after its own walk.
// The subtractions into one sphere therefore come in this order: one per
// lower neighbour, in ascending index (the outer loop), then one per higher
// neighbour in the map's iteration order.  That order is libstdc++'s:
// neighbour_order() replays it (see below).
#pragma once
#include <stdint.h>

#include "device_math.hpp"

namespace mcac_dev {

constexpr int kMapKeys = 32;  // neighbours one sphere's map may hold here

// Bucket count libstdc++ picks for at least n buckets (_Prime_rehash_policy::
// _M_next_bkt): a fixed table below 14, else the first prime of its list not
// below n.  The primes listed cover every count a map of kMapKeys keys
// reaches; 0 beyond them.
__host__ __device__ inline unsigned umap_next_bkt(unsigned n) {
    const unsigned char fast[14] = {2, 2, 2, 3, 5, 5, 7, 7, 11, 11, 11, 11, 13, 13};
    if (n < 14) return n == 0 ? 1u : fast[n];
    const unsigned short primes[] = {17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 103, 109, 113};
    for (unsigned p : primes)
        if (p >= n) return p;
    return 0u;
}

// Inserts key at the front of its bucket's run in list l of len keys (a new
// bucket goes to the front of the list): _Hashtable::_M_insert_bucket_begin,
// also the order _M_rehash_aux rebuilds the list in.
__host__ __device__ inline void umap_link(int64_t *l, int len, int64_t key, unsigned nb) {
    const int64_t b = key % static_cast<int64_t>(nb);
    int p = 0;
    while (p < len && l[p] % static_cast<int64_t>(nb) != b) ++p;
    if (p == len) p = 0;
    for (int q = len; q > p; --q) l[q] = l[q - 1];
    l[p] = key;
}

// Iteration order of one sphere's neighbour map after
// update_distances_and_overlapping (aggregat.cpp:730-751): clear(), then
// reserve(prev) with prev the map's size after the previous update (0 for a
// new map), then the m neighbours inserted in ascending index (keys, sorted).
// libstdc++ with the identity hash and max load factor 1: reserve sets
// next_bkt(max(1, prev)) buckets; inserting the (c+1)-th key into c keys and
// nb buckets with c + 1 > nb first rehashes to next_bkt(max(c + 2, 2 nb)).
// out receives the m keys; returns 0, or -1 when m > kMapKeys.
__host__ __device__ inline int neighbour_order(int prev, const int64_t *keys, int m, int64_t *out) {
    if (m > kMapKeys || m < 0) return -1;
    unsigned nb = umap_next_bkt(prev > 1 ? static_cast<unsigned>(prev) : 1u);
    int64_t tmp[kMapKeys];
    for (int c = 0; c < m; ++c) {
        if (static_cast<unsigned>(c + 1) > nb) {
            const unsigned want = static_cast<unsigned>(c + 2) > 2 * nb ? static_cast<unsigned>(c + 2) : 2 * nb;
            const unsigned nb2 = umap_next_bkt(want);
            for (int q = 0; q < c; ++q) tmp[q] = out[q];
            for (int q = 0; q < c; ++q) umap_link(out, q, tmp[q], nb2);
            nb = nb2;
        }
        umap_link(out, c, keys[c], nb);
    }
    return 0;
}

// Spherical caps of two spheres at distance dist (sphere_intersection.cpp:29-66,
// Morán et al. 2018): volume and surface of sphere 1 covered by sphere 2
// (v1, s1) and the converse (v2, s2).  pow(x, 2) is x * x here as in the
// reference's build (GCC folds it); vol/surf are the spheres' own values for
// the contained case.  Operation order as the reference's expressions.
__host__ __device__ inline void caps_intersection(double r1, double vol1, double surf1, double r2, double vol2, double surf2,
                                      double dist, double *v1, double *s1, double *v2, double *s2) {
    *v1 = *v2 = *s1 = *s2 = 0.0;
    if (dist <= 0) return;
    if (dist < r1 + r2) {
        const double fd = r1 > r2 ? r1 - r2 : 0.0;  // std::fdim
        if (dist >= fd) {
            const double a = r1 - dist, c = r2 - dist;
            const double h1 = (r2 * r2 - a * a) / (2. * dist);
            const double h2 = (r1 * r1 - c * c) / (2. * dist);
            *v1 = kPi * (h1 * h1) * (3 * r1 - h1) / 3.;
            *v2 = kPi * (h2 * h2) * (3 * r2 - h2) / 3.;
            *s1 = 2 * kPi * r1 * h1;
            *s2 = 2 * kPi * r2 * h2;
        } else if (r1 < r2) {
            *v1 = vol1;
            *s1 = surf1;
        } else {
            *v2 = vol2;
            *s2 = surf2;
        }
    }
}

}  // namespace mcac_dev
