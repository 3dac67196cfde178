/*
 * mcac_gpu.h — C-ABI of the MI355X aggregate-move / collision-search path.
 *
 * One `mcac_box` holds one periodic aggregation box resident in HBM:
 * structure-of-arrays sphere and aggregate fields, per-aggregate member
 * lists, and the scratch of the contact search.  Aggregates are named by a
 * stable id (their creation index); the reference's label of an aggregate is
 * its rank among the live ids, so ascending id order is label order and no
 * relabelling is ever needed.  Spheres are named by their index, which never
 * changes within a box (oxidation's sphere removals are carried out by the
 * caller on the state read back, and a new box holds the remaining spheres in
 * their order, as SphereList::remove leaves them).
 *
 * The entry points replace these reference operators (MCAC, C++):
 *   mcac_next_contact  AggregatList::distance_to_next_contact
 *                      (include/aggregats/aggregat_list.hpp:69-71,
 *                       src/aggregats/aggregat_list.cpp:447-532, with
 *                       src/verlet/verlet.cpp:52-98 and
 *                       src/spheres/sphere_contact.cpp:47-139)
 *   mcac_translate     Aggregate::translate (include/aggregats/aggregat.hpp:108,
 *                       src/aggregats/aggregat.cpp:148-161)
 *   mcac_move          the pair of calls at src/calcul.cpp:127-144 fused:
 *                       contact search, then translate by min(contact, lpm)
 *   mcac_merge         AggregatList::merge (aggregat_list.hpp:63,
 *                       aggregat_list.cpp:367-410) with Aggregate::merge and
 *                       Aggregate::update (aggregat.cpp:284-288, 486-544,
 *                       719-764, 321-483) minus the host physics
 *                       (friction, mobility, lpm, time step), which the caller
 *                       computes from the returned geometry.
 *   mcac_box_upload /  the SoA storage the reference keeps in
 *   mcac_box_download  ListStorage (include/list_storage/list_storage.hpp:25-73)
 *
 * Every function returns an int status whose values equal the reference's
 * ErrorCodes (include/constants.hpp:70-83) where one applies, or one of the
 * MCAC_E_* codes below.  No C++ exception crosses this boundary.
 */
#ifndef MCAC_GPU_H
#define MCAC_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. 0..11 mirror the reference ErrorCodes. */
#define MCAC_NO_ERROR 0
#define MCAC_UNKNOWN_ERROR 1
#define MCAC_IO_ERROR 2
#define MCAC_VERLET_ERROR 3
#define MCAC_INPUT_ERROR 4
#define MCAC_VOL_SURF_ERROR 8
#define MCAC_MERGE_ERROR 9
#define MCAC_E_DEVICE 32     /* a HIP runtime call failed                     */
#define MCAC_E_CAPACITY 33   /* a fixed-capacity buffer would overflow         */
#define MCAC_E_ARGUMENT 34   /* invalid argument (id out of range, dead id...) */

typedef struct mcac_box mcac_box;
typedef struct mcac_group mcac_group;

typedef struct mcac_box_params {
    double box_length;          /* periodic box side L [m]                          */
    int64_t n_div;              /* reference Verlet divisions per axis (candidate   */
                                /* visit order; physical_model.cpp:162)             */
    int64_t max_spheres;        /* sphere capacity                                  */
    int64_t max_aggregates;     /* stable-id capacity                               */
    int64_t candidate_capacity; /* filtered-candidate buffer per search (0: auto)   */
    int64_t member_pool;        /* member-list pool entries (0: auto)               */
    int32_t device;             /* HIP device ordinal                               */
    int32_t precision;          /* 0: FP64 throughout; 1: FP32 broad phase, FP64    */
                                /*    final solve (results identical to 0); 2: as   */
                                /*    1, and every entry the FP32 test drops is     */
                                /*    also tested in FP64 to count misses           */
    int32_t volsurf;            /* overlap correction of the merge update           */
                                /* (compute_volume_surface, aggregat.cpp:321-430):  */
                                /* 0 none (sphere sums); 1 spherical caps           */
                                /* (SPHERICAL_CAPS, :354-386); 2 alphas (ALPHAS,    */
                                /* :392-422, correction factors on the host between */
                                /* two device phases); 3 arvo (EXACT_ARVO, :337-347,*/
                                /* arvo.cpp restated with fixes, per-sphere parts   */
                                /* within 1e-10 of the oracle).  With 1 to 3 no     */
                                /* merge runs inside the contact step (auto-merge   */
                                /* stays off)                                       */
    int32_t reserved;           /* 0                                                */
} mcac_box_params;

/* Result of a contact search (AggregateContactInfo, tools/contact_info.hpp:73-85). */
typedef struct mcac_contact {
    double distance;            /* +inf when no contact is found                    */
    double move;                /* contact ? distance : lpm                         */
    int64_t moving_sphere;      /* -1 when none                                     */
    int64_t other_sphere;       /* -1 when none                                     */
    int64_t moving_agg;         /* stable id of the searching aggregate             */
    int64_t other_agg;          /* stable id, -1 when none                          */
    int32_t contact;            /* distance <= lpm                                  */
    int32_t status;
    int64_t n_query;            /* aggregates in the visited Verlet cells (C_q)     */
    int64_t n_filtered;         /* candidates passing the bounding filter (F)       */
    int64_t n_pairs;            /* exact sphere-pair tests performed                */
} mcac_contact;

/* Geometry of the aggregate after a merge (Aggregate::update minus physics). */
typedef struct mcac_merge_result {
    int64_t keeper;             /* stable id kept (min label)                       */
    int64_t removed;            /* stable id removed (max label)                    */
    int64_t n;                  /* spheres in the merged aggregate                  */
    int32_t merged;             /* 0 when the contact re-check failed               */
    int32_t status;
    double volume, surface, rg, rmax;
    double radius_sum;          /* sum of sphere radii in list order (dp = 2 s / n) */
    double com[3];              /* aggregate position (periodic)                    */
    double rcm[3];              /* centre of mass relative to sphere 0              */
    double overlapping, coordination;
    double sphere_volume;       /* sum of the spheres' own volumes in list order    */
                                /* (vol_pp = this / n, aggregat.cpp:255-262)        */
} mcac_merge_result;

int mcac_box_create(const mcac_box_params *params, mcac_box **out);
/* Device memory (bytes) mcac_box_create allocates for these parameters: the
 * sum of its buffer sizes, before the allocator's rounding (host computation,
 * no GPU call); -1 for invalid parameters.  No reference counterpart: it lets
 * a driver admit boxes by an HBM budget. */
int64_t mcac_box_device_bytes(const mcac_box_params *params);
int mcac_box_destroy(mcac_box *box);

/* Uploads a full state.  Spheres in index order (positions, radius, volume,
 * surface, position relative to their aggregate's sphere 0).  Aggregates in
 * stable-id order: position, relative CoM, rmax, volume, surface, sphere count
 * (0 marks a dead id) and members (concatenated sphere indices, list order). */
int mcac_box_upload(mcac_box *box, int64_t n_spheres, const double *x, const double *y, const double *z,
                    const double *r, const double *vol, const double *surf, const double *rx,
                    const double *ry, const double *rz, int64_t n_aggs, const double *pos3,
                    const double *rcm3, const double *rmax, const double *volume, const double *surface,
                    const int64_t *count, const int64_t *members);

int mcac_next_contact(mcac_box *box, int64_t agg, const double dir[3], double lpm, double maxradius,
                      mcac_contact *out);
int mcac_translate(mcac_box *box, int64_t agg, const double vec[3]);
int mcac_move(mcac_box *box, int64_t agg, const double dir[3], double lpm, double maxradius,
              mcac_contact *out);
int mcac_merge(mcac_box *box, const mcac_contact *contact, mcac_merge_result *out);

/* Pipelined form of mcac_move (calcul.cpp:119-144 issued ahead of time): the
 * search and move are enqueued on the box stream and a ticket returned at once,
 * so the caller can prepare and enqueue the next step while this one runs.  A
 * step that finds a contact closes a gate: steps enqueued after it do nothing
 * and report skipped = 1 from mcac_move_result, until mcac_release is called
 * (after the caller has merged).  At most 16 steps may be in flight. */
int mcac_move_enqueue(mcac_box *box, int64_t agg, const double dir[3], double lpm, double maxradius,
                      uint64_t *ticket);
/* Waits for an enqueued step, in ticket order.  *skipped is 0 when the step
 * ran, 1 when it did nothing behind a closed gate (issue it again), and 2 when
 * it ran without a contact but needed a second launch, during which the steps
 * enqueued after it were skipped (they report 1; the gate is open again). */
int mcac_move_result(mcac_box *box, uint64_t ticket, mcac_contact *out, int *skipped);
/* Reopens the gate closed by a step that found a contact. */
int mcac_release(mcac_box *box);
/* With auto-merge on, a queued move that finds a contact also merges the two
 * aggregates in the same launch when the merged aggregate has at most 256
 * spheres and its member segment fits (calcul.cpp:156-160 always merges on
 * contact); mcac_merge then returns that result without a launch.  The
 * caller must call mcac_merge for such a contact before enqueueing further
 * moves.  Off by default; set while no move is queued.  A box with
 * spherical caps (volsurf 1) keeps it off. */
int mcac_box_set_auto_merge(mcac_box *box, int enable);
/* Host-only check of the spherical-caps walk order (no GPU call): the
 * iteration order of one sphere's neighbour map (std::unordered_map<size_t,
 * double>, aggregat.hpp:61-65) after Aggregate::update_distances_and_overlapping
 * (aggregat.cpp:730-751) cleared it, reserved prev entries (its size after the
 * previous update) and inserted the m keys (ascending, m <= 32).  out gets the
 * m keys in that order.  No counterpart in the reference: the reference walks
 * the map itself; the device kernels walk this order. */
int mcac_neighbour_order(int prev, const int64_t *keys, int m, int64_t *out);
/* Host-only check of the ARVO part (no GPU call): ball 0 of the k + 1 balls
 * (x, y, z, r in nm, k <= 16 neighbours, each overlapping ball 0 and not
 * inside it) carries out[0] of the union's volume and out[1] of its area
 * (arvo.cpp:728-788 restated).  No counterpart in the reference: it is the
 * per-sphere step the device kernel runs, exposed for host checks. */
int mcac_arvo_part(int k, const double *x, const double *y, const double *z, const double *r, double out[2]);

/* Groups of boxes: one queued move of each of many boxes in a single launch
 * (one workgroup per box).  No counterpart in the reference, which runs one box
 * per process (src/main.cpp); for ensembles of small boxes, whose steps are
 * bound by launch and memory latency rather than by work.  A box joins a group
 * with no step in flight and then works on the group's stream (its own stream
 * is released); it leaves the group only when destroyed, which must happen
 * before the group is destroyed. */
int mcac_group_create(int device, int max_boxes, mcac_group **out);
int mcac_group_destroy(mcac_group *group);
int mcac_group_add(mcac_group *group, mcac_box *box);
/* mcac_move_enqueue for n moves of boxes of the group: the moves of one box
 * are contiguous, in the order they are to run (at most 16 in flight per
 * box; past 8 of a box in one call the list goes to a further launch),
 * n <= 8 * max_boxes.  One launch takes every move that fits one workgroup,
 * one workgroup running each box's moves in order; the others take their
 * usual launches on the group stream, in order with the rest.  dirs holds 3
 * doubles per move.  Results are read per box with mcac_move_result, contacts
 * and skipped moves handled as for mcac_move_enqueue.  On an error the moves
 * enqueued before the failing one keep their tickets and run. */
int mcac_move_enqueue_group(mcac_group *group, mcac_box *const *boxes, const int64_t *aggs, const double *dirs,
                            const double *lpm, const double *maxradius, int n, uint64_t *tickets);
/* [group launches, moves in group launches, moves launched on their own] */
int mcac_group_stats(mcac_group *group, int64_t out[3]);
/* Host seconds spent in mcac_move_enqueue_group on this group: [in all, in
 * the cell-list rebuilds it ran (with the waits they need), in kernel
 * launches, waiting for the launch that read a job half to have read it].  No
 * counterpart in the reference; for tuning a driver loop. */
int mcac_group_times(mcac_group *group, double out[4]);
/* Moves the box can take in its next group launch without waiting for the
 * device: its free in-flight slots and, while steps are in flight, the moves
 * left before a cell-list rebuild falls due (a rebuild with steps in flight
 * drains the stream first).  No counterpart in the reference; lets a driver
 * keep two group launches in flight. */
int mcac_box_spare_moves(mcac_box *box);
/* 1 when a move of the box whose sweep reaches `reach` (the mover's rmax +
 * maxradius + lpm) around its centre would currently run in one workgroup of
 * a group launch (its cell window holds few enough aggregates), else 0.  No
 * counterpart in the reference; lets a driver choose between group launches
 * and one launch per move. */
int mcac_box_block_steps(mcac_box *box, double reach);

int mcac_box_download_spheres(mcac_box *box, double *x, double *y, double *z, double *rx, double *ry,
                              double *rz);
/* Radius, volume and surface of every sphere (index order; NULL skips an
 * array): the SphereList fields that surface growth changes
 * (sphere.cpp:113-120). */
int mcac_box_download_radii(mcac_box *box, double *r, double *vol, double *surf);
int mcac_box_download_aggs(mcac_box *box, double *pos3, double *rcm3, double *rmax, double *volume,
                           double *surface, double *rg, int64_t *count);
/* Stable id of the aggregate owning each sphere (n_spheres entries; -1 for a
 * sphere no live aggregate owns): the spheres' aggregate labels written by
 * SphereList::save (src/io/sphere_list.cpp:44-46) once mapped to labels. */
int mcac_box_sphere_owners(mcac_box *box, int64_t *owner);
/* Member lists of every aggregate: count[id] for each id below the box's
 * aggregate count (0 for a dead id), and members the concatenated lists in
 * id order, each in list order (SphereList order of Aggregate::myspheres).
 * members must hold the box's live sphere count. */
int mcac_box_download_members(mcac_box *box, int64_t *count, int64_t *members);
/* Members of one aggregate in list order; returns the count via *n. */
int mcac_box_members(mcac_box *box, int64_t agg, int64_t *out, int64_t max, int64_t *n);

/* Surface growth of the whole box and the aggregate update that follows it
 * (calcul.cpp:155-205 with individual_surf_reactions off and
 * full_aggregate_update_frequency 1, volsurf none): every sphere of a live
 * aggregate grows by dr = u_sg * dt (Sphere::croissance_surface,
 * sphere.cpp:113-120, PhysicalModel::grow, physical_model.cpp:587-590), then
 * every live aggregate gets Aggregate::update (aggregat.cpp:284-288, 321-483:
 * volume, surface, centre of mass and position, rmax, Rg).  The outputs are
 * indexed by stable id (arrays of the box's aggregate capacity; dead ids are
 * not written): volume, surface, list-order radius sum (dp = 2 sum / n), Rg,
 * rmax.  The caller then merges a contact (its geometry comes from the grown
 * spheres, as AggregatList::merge runs after the growth) and computes the
 * physics.  No step may be in flight and no inline merge pending. */
int mcac_box_grow(mcac_box *box, double dr, double *volume, double *surface, double *radius_sum, double *rg,
                  double *rmax);
/* mcac_box_grow for the surface reactions in general (calcul.cpp:155-205):
 * the aggregates of ids[0..n_ids) (ids NULL: every live aggregate), spheres
 * grown by dr (negative: oxidation) when grow != 0, else the update alone
 * (radii, volumes and surfaces as stored: the whole-box update after an event
 * with individual_surf_reactions, calcul.cpp:186-191).  flags (by id, NULL to
 * skip) gets bit 0 when a sphere of the aggregate is at or below rp_min after
 * growing (Aggregate::croissance_surface, aggregat.cpp:230-243, removes it)
 * and, with check_split, bit 1 when its spheres are no longer one connected
 * set under the contact test (Aggregate::split, aggregat.cpp:553-584) or it
 * has more than 64 spheres (not checked).  The caller carries out removals
 * and splits on the state it reads back, and uploads the result to a new box. */
int mcac_box_grow_ids(mcac_box *box, const int64_t *ids, int64_t n_ids, double dr, int grow, double rp_min,
                      int check_split, double *volume, double *surface, double *radius_sum, double *rg, double *rmax,
                      int32_t *flags);

/* Persistent step engine: with blocks >= 1 (at most 255), queued one-launch
 * steps (mcac_move_enqueue of a mover the step's own finishing block moves,
 * not timed) are posted to a ring in pinned host memory and taken in order by
 * a resident grid of that many workgroups (fewer while the steps' windows
 * are small), instead of one kernel launch each; results, tickets, gate and
 * skipped steps are exactly as with launches.  Every other operation on the
 * box first ends the resident grid, which is launched again at the next
 * step; a grid with no step for 0.5 s ends by itself.  Call it again (with
 * the same count) to end the resident grid, so that the device is idle (for
 * a device-wide synchronize), or with 0 to turn the engine off.  Not for a
 * box in a group.  No counterpart in the reference, which calls its search
 * once per step (src/calcul.cpp:127). */
int mcac_box_set_engine(mcac_box *box, int blocks);
/* [grids launched, steps posted, steps completed by the engine, device ms of
 * those steps (each from when it could start: its job taken and the previous
 * step done, to its record), grids launched again after an idle end, the last
 * grid's end reason (2 idle, 3 ended by the host, -1 a wait ran out), ms the
 * grid waited for the host's next job between steps, workgroups of the last
 * grid]. */
int mcac_box_engine_stats(mcac_box *box, double out[8]);

/* Kernel timing on the box's stream.  enable = n >= 1 times the first search
 * enqueued after the call and every n-th one after it, with start/stop events stamped by the runtime at the kernel's own
 * start and end (hipExtLaunchKernel), and every merge launch with events
 * around it; 0 turns timing off.  The host never blocks on the events of a
 * search.  mcac_box_timing returns accumulated milliseconds and counts:
 * [search_ms, translate_ms (not timed, 0), merge_ms, n_search, n_translate,
 * n_merge, first-launch ms (k_step or k_filter), second-launch ms (k_solve)]. */
int mcac_box_set_timing(mcac_box *box, int enable);
int mcac_box_timing(mcac_box *box, double out[8]);
/* Launch statistics: [searches of small movers (k_step, one launch),
 * searches of large movers (k_step_grid, one launch with a grid-wide barrier
 * between filter and pair solve; with MCAC_GRID_SOLVE=0 two launches, k_filter
 * + k_solve), small-mover searches finished by a second launch, cell-list
 * rebuilds, member-pool compactions, member-pool slots in use, merges done
 * inside a queued move's launch, merges in their own launch].  No reference
 * counterpart (diagnostics). */
int mcac_box_stats(mcac_box *box, int64_t out[8]);
/* FP32 broad phase, counted with precision 2: [cell entries tested in FP32,
 * entries passed on to the FP64 filter, entries dropped in FP32 that the FP64
 * filter keeps (must be 0)].  No reference counterpart. */
int mcac_box_prune_stats(mcac_box *box, int64_t out[3]);

/* distance_to_contact(Sphere, Sphere, dir, dist) (src/spheres/sphere_contact.cpp:47-125)
 * evaluated on the device for n independent inputs: p1, p2, dir are (n,3)
 * row-major, r1, r2, dist are (n,), box is the period.  out[i] receives the
 * travel to contact (0 when touching, +inf when none).  Host buffers. */
int mcac_swept_sphere_batch(int64_t n, const double *p1, const double *r1, const double *p2, const double *r2,
                            const double *dir, const double *dist, double box, double *out, int device);

/* mcac_swept_sphere_batch without any device allocation: no counterpart in
 * the reference; for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results; the device staging is a static array that each GPU holds until
 * the process ends, so n is limited to 65,536 inputs (larger n: returns
 * MCAC_E_CAPACITY and writes nothing).  Calls that overlap on one GPU share
 * that array and race. */
int mcac_swept_sphere_batch_noalloc(int64_t n, const double *p1, const double *r1, const double *p2,
                                    const double *r2, const double *dir, const double *dist, double box, double *out,
                                    int device);

/* Library self-description for loaders and tests. */
const char *mcac_version(void);
/* Free and total device memory of a GPU (hipMemGetInfo; no reference
 * counterpart), for drivers that size HBM budgets and for tests. */
int mcac_device_memory(int device, int64_t *free_bytes, int64_t *total_bytes);

#ifdef __cplusplus
}
#endif
#endif
