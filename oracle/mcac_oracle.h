/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of MCAC's aggregate-move / collision-search path (the
 * reference `calcul` loop and the functions it calls), written from the
 * reference's semantics for use as the parity checker of the HIP product.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library; the product (mcac_amd/) never links or calls it.
 *
 * Every function in mcac_oracle.cpp cites the reference file:line it follows.
 * Pinning of each part (golden vectors from oracle/_ref, closed forms, brute
 * force) is listed in the header comment of mcac_oracle.cpp.
 */
#ifndef MCAC_ORACLE_H
#define MCAC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One MC step (calcul.cpp:66-281). Layout shared with the product's trace. */
typedef struct ora_step_rec {
    int64_t step;          /* 0-based MC step index                                   */
    int64_t agg;           /* picked aggregate, stable id (creation index)            */
    int64_t label;         /* picked aggregate, label (index in label order)          */
    double dir[3];         /* random_direction()                                      */
    double lpm;            /* full_distance                                           */
    double dist;           /* next_contact.distance (+inf when none)                  */
    int64_t moving_sphere; /* sphere index, -1 when none                              */
    int64_t other_sphere;  /* sphere index, -1 when none                              */
    int64_t other_agg;     /* stable id, -1 when none                                 */
    int32_t contact;       /* next_contact <= full_distance                           */
    int32_t merged;        /* AggregatList::merge returned true                       */
    double dt;             /* deltatemps after the move correction (calcul.cpp:147)   */
    double time;           /* physicalmodel.time after the step                       */
} ora_step_rec;

/* One merge (aggregat_list.cpp:367-410), recorded after Aggregate::update. */
typedef struct ora_merge_rec {
    int64_t step;
    int64_t keeper;        /* stable id of the kept aggregate (min label)             */
    int64_t removed;       /* stable id of the removed aggregate (max label)          */
    int64_t n;             /* spheres in the merged aggregate                         */
    double volume, surface, rg, rmax, dp, f_agg, lpm, time_step, d_m, proper_time;
    double overlapping, coordination;
    double com[3];         /* aggregate position (wrapped)                            */
    double rcm[3];         /* centre of mass relative to sphere 0                     */
} ora_merge_rec;

/* Simulation lifetime. ini_text is the content of an MCAC parameter file. */
void *ora_create(const char *ini_text, char *err, int errlen);
void ora_destroy(void *sim);
/* AggregatList constructor equivalent (aggregat_list_storage.cpp:52-88).
 * fast != 0 answers test_free_space with a private hash grid instead of the
 * Verlet sets; the answer (any contact or none) is the same.  Returns the
 * reference ErrorCodes value (0 = ok, 6 = TooDenseError). */
int ora_init(void *sim, int fast);
/* Runs at most max_steps iterations of the calcul loop; returns the number run,
 * or -(error code) on error. Stops early when PhysicalModel::finished. */
int64_t ora_run(void *sim, int64_t max_steps);
int ora_finished(void *sim);
void ora_set_trace(void *sim, int steps, int merges);
int64_t ora_n_steps(void *sim);
int64_t ora_get_steps(void *sim, ora_step_rec *out, int64_t max);
int64_t ora_n_merges(void *sim);
int64_t ora_get_merges(void *sim, ora_merge_rec *out, int64_t max);

/* Scalars: box_length, viscosity, lambda, rpeqmass, gamma, u_sg, time,
 * maxradius, n_agg, n_sph, avg_npp, max_time_step.  name -> value. */
double ora_get_scalar(void *sim, const char *name);

/* State export (for driving the HIP kernels from the same state). Arrays are
 * filled in sphere-index order and label order respectively. */
int64_t ora_n_spheres(void *sim);
int64_t ora_n_aggs(void *sim);
void ora_export_spheres(void *sim, double *x, double *y, double *z, double *r,
                        double *vol, double *surf, double *rx, double *ry, double *rz,
                        int64_t *label);
/* Per aggregate (label order): position, relative CoM, rmax, volume, surface,
 * rg, lpm, time_step, proper_time, n, uid, member offset into members[] */
void ora_export_aggs(void *sim, double *pos /*3n*/, double *rcm /*3n*/, double *rmax,
                     double *volume, double *surface, double *rg, double *lpm,
                     double *time_step, double *proper_time, int64_t *n, int64_t *uid,
                     int64_t *offset, int64_t *members);

/* AggregatList::distance_to_next_contact on the current state
 * (aggregat_list.cpp:447-484). Outputs the contact distance, the moving and
 * other sphere indices and the other aggregate's label (-1 when none), and
 * the number of candidates (Verlet query, after the bounding filter, and
 * exact aggregate-pair solves) for work accounting. */
double ora_next_contact(void *sim, int64_t label, const double dir[3], double lpm,
                        int64_t *moving_sphere, int64_t *other_sphere, int64_t *other_label,
                        int64_t *n_query, int64_t *n_filtered, int64_t *n_solved,
                        int64_t *n_pairs);

/* PhysicalModel derived scalars for a parameter file, without running:
 * out = [box_length, lambda, viscosity, rpeqmass, gamma, u_sg, box_volume,
 * aggregate_concentration]. Returns 0 on success. */
int ora_physics(const char *ini_text, double out[8]);

/* Primitive: distance_to_contact(Sphere,Sphere,dir,dist) (sphere_contact.cpp:47-125). */
double ora_swept_sphere(const double p1[3], double r1, const double p2[3], double r2,
                        const double dir[3], double dist, double box);
/* Primitive: contact(Sphere,Sphere) (sphere_distances.cpp:81-90). */
int ora_contact(const double p1[3], double r1, const double p2[3], double r2, double box);

/* RNG / draws (tools.cpp:41-89, physical_model.cpp:557-578) on a private
 * glibc TYPE_3 stream seeded like srand(seed). */
void *ora_rng_create(int seed);
void ora_rng_destroy(void *rng);
double ora_rng_uniform(void *rng);
void ora_rng_direction(void *rng, double dir[3]);
double ora_inverf(double p);
double ora_inverfc(double p);

/* Verlet::get_neighborhood over an explicit cell membership
 * (verlet.cpp:52-98): returns visited labels in visit order. cells[] holds
 * for each label its (i,j,k) cell; n_labels labels. out receives up to max. */
int64_t ora_verlet_query(int64_t n_div, double width, const int64_t *cells, int64_t n_labels,
                         const double src[3], const double vec[3], double dist,
                         int64_t *out, int64_t max);

/* Spherical caps of two spheres (sphere_intersection.cpp:29-66):
 * out = {volume_1, surface_1, volume_2, surface_2}. */
void ora_caps_intersection(double r1, double vol1, double surf1, double r2, double vol2, double surf2, double dist,
                           double out[4]);
/* Geometry of Aggregate::update (aggregat.cpp:284-288, 431-483) for n spheres
 * given by their positions relative to sphere 0, of which the first nk were
 * last updated on their own (the keeper of a merge): volsurf 0 none, 1 caps,
 * 2 alphas, 3 arvo.
 * out = {volume, surface, rg, rmax, rcm x, y, z, overlapping, coordination,
 * sum of sphere volumes, sum of radii}; returns 0, or < 0 (bad sizes, VolSurfError). */
int ora_merge_geometry(int volsurf, int64_t nk, int64_t n, const double *rx, const double *ry, const double *rz,
                       const double *r, const double *vol, const double *surf, double out[11]);
/* ARVO part (arvo.cpp:728-788 restated) of ball 0 among k + 1 balls (nm):
 * out = {volume, area}. */
void ora_arvo_part(int k, const double *x, const double *y, const double *z, const double *r, double out[2]);
/* Iteration order of a std::unordered_map<size_t,double> that held prev
 * entries, was cleared, reserved at its old size and given keys[0..m) in
 * order (aggregat.cpp:730-751); returns the number of keys written. */
int ora_neighbour_order(int prev, const int64_t *keys, int m, int64_t *out);

#ifdef __cplusplus
}
#endif
#endif
