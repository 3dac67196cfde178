/*
 * mcac_sim.h — host driver of one MCAC aggregation run on the MI355X path.
 *
 * The driver keeps the reference's MC loop on the host CPU (MCAC parameter
 * file, physical model, glibc random stream, time-step weighted pick and
 * direction draw, time bookkeeping, advancement.dat) and calls the device
 * through include/mcac_gpu.h.  It mirrors:
 *   mcac_sim_create / mcac_sim_init   PhysicalModel(const std::string&) and
 *                                     AggregatList(PhysicalModel*)
 *                                     (physical_model.cpp:32-287,
 *                                      aggregat_list_storage.cpp:52-88)
 *   mcac_sim_run                      mcac::calcul (src/calcul.cpp:55-290)
 */
#ifndef MCAC_SIM_H
#define MCAC_SIM_H
#include <stdint.h>

#include "mcac_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mcac_sim mcac_sim;

/* Per-step and per-merge trace records (same layout as the oracle's). */
typedef struct mcac_step_rec {
    int64_t step, agg, label;
    double dir[3];
    double lpm, dist;
    int64_t moving_sphere, other_sphere, other_agg;
    int32_t contact, merged;
    double dt, time;
} mcac_step_rec;

typedef struct mcac_merge_rec {
    int64_t step, keeper, removed, n;
    double volume, surface, rg, rmax, dp, f_agg, lpm, time_step, d_m, proper_time;
    double overlapping, coordination;
    double com[3], rcm[3];
} mcac_merge_rec;

/* ini_text: content of an MCAC parameter file.  output_dir: directory for
 * advancement.dat and params.ini (NULL or "" disables file output).
 * Options the path runs: collisions and merges with every volsurf_method but
 * sbl; surface reactions with volsurf_method none (growth, or oxidation with
 * sphere removal and splits; for every aggregate or the moved one with
 * individual_surf_reactions; full_aggregate_update_frequency 1); domain
 * duplication with volsurf_method none.  Others (nucleation, potentials,
 * flame coupling, domain reduction, pick_method last, a clock seed, sparser
 * whole updates) make mcac_sim_create fail with MCAC_INPUT_ERROR and a
 * message naming the option.
 * Environment: MCAC_PRECISION (0, 1, 2) selects the box's contact-search
 * precision (mcac_box_params.precision; results are identical);
 * MCAC_PIPELINE_DEPTH (1..6, default 3) the MC steps kept in flight;
 * MCAC_ROUNDS (default 0; 1 on): when the next move's cell window is small
 * enough for one workgroup, up to 8 moves are picked ahead and run in one
 * group launch (mcac_move_enqueue_group); depth 1 or an output directory
 * turns this off.  Off by default: on the N=1e6 box it is no faster than the
 * one-launch pipeline, and slower where windows hold a few thousand entries.
 * MCAC_FUSED_MERGE (default 1; 0 off, read by mcac_box_create): merges with
 * an overlap correction of at most 512 spheres run in one workgroup launch
 * instead of the method's kernel sequence (results are identical).
 * MCAC_ENGINE (default 0, off; n >= 1 on): queued one-launch steps go to a resident
 * grid of up to that many workgroups (mcac_box_set_engine) instead of one
 * launch each; the grid ends when mcac_sim_run returns (results are
 * identical). */
int mcac_sim_create(const char *ini_text, const char *output_dir, int device, mcac_sim **out, char *err,
                    int errlen);
int mcac_sim_destroy(mcac_sim *sim);
/* Initial placement (aggregat.cpp:162-229, host) and upload to the device. */
int mcac_sim_init(mcac_sim *sim);
/* Runs at most max_steps MC steps; *done receives the number run. */
int mcac_sim_run(mcac_sim *sim, int64_t max_steps, int64_t *done);
int mcac_sim_finished(mcac_sim *sim);
int mcac_sim_set_trace(mcac_sim *sim, int steps, int merges);
int64_t mcac_sim_n_steps(mcac_sim *sim);
int64_t mcac_sim_get_steps(mcac_sim *sim, mcac_step_rec *out, int64_t max);
int64_t mcac_sim_n_merges(mcac_sim *sim);
int64_t mcac_sim_get_merges(mcac_sim *sim, mcac_merge_rec *out, int64_t max);
/* Named scalars: box_length, time, maxradius, n_agg, n_sph, avg_npp,
 * max_time_step, steps, merges, events, total_volume, volume_fraction, query,
 * filtered, pairs, ties, t_host_pick_ms, t_host_merge_ms, t_device_ms, t_tie_ms,
 * t_enqueue_ms, t_wait_ms, t_merge_call_ms, t_drain_ms, init_ms,
 * pick_exact_steps, pick_chunks, pick_rescales. NaN for unknown names. */
double mcac_sim_scalar(mcac_sim *sim, const char *name);
/* Host view of the live aggregates in label order (stable ids). */
int64_t mcac_sim_n_aggs(mcac_sim *sim);
int mcac_sim_export_aggs(mcac_sim *sim, int64_t *uid, double *lpm, double *time_step, double *proper_time,
                         double *volume, double *rg, double *rmax, int64_t *n);
mcac_box *mcac_sim_box(mcac_sim *sim);

/* Derived scalars of PhysicalModel for a parameter file, without running
 * (physical_model.cpp:228-269, 536-546): out = [box_length, lambda, viscosity,
 * rpeqmass, gamma, u_sg, box_volume, aggregate_concentration]. Options this
 * path does not run are accepted here. */
int mcac_physics_scalars(const char *ini_text, double out[8]);

/* Self-check of the pick table (aggregat_list.cpp:54-66, 109-141): random time
 * steps and merges on n entries; after every merge each cumulative sum and
 * eight picks are compared with a full re-sort and a plain sequential sum.
 * mode 0: mostly distinct steps, 1: a dozen distinct values (most picks land
 * in a run of equal keys), 2: many pairs of equal steps.  Returns the number
 * of mismatching checks (0 expected).  Host only.  The second call returns the
 * picks so far that landed in a run of equal keys. */
int64_t mcac_pick_table_selftest(int64_t n, int64_t events, int64_t chunk, uint64_t seed, int mode);
int64_t mcac_pick_table_selftest_ties(void);
/* Snapshot-writer check (host only): `grids` synthetic sphere grids of n
 * spheres written to <prefix>_<number>.{h5,xmf}, per_file grids per file;
 * position k of grid g = ((g + k) / 8, 2k / 8, 3k / 8), Radius k / 4, Label k. */
int mcac_write_test_snapshots(const char *prefix, int64_t n, int64_t grids, int64_t per_file);

/* Device memory (bytes) the box of this parameter file takes
 * (mcac_box_device_bytes of the parameters Simulation uses: one sphere and
 * one aggregate id per monomer; with domain duplication, the box after the
 * last duplication a mean-size limit allows, 8^k times the spheres where
 * 8^k < mean_monomere_per_aggregate); -1 when the text does not parse or asks
 * for an option this path does not run.  No reference counterpart: admission
 * of runs by an HBM budget. */
int64_t mcac_sim_device_bytes(const char *ini_text);

/* Ensemble / serving runs (no reference counterpart: the reference runs one
 * box per process, src/main.cpp): n_jobs parameter files run to completion
 * (or max_steps MC steps) on `workers` host threads sharing one GPU, each
 * worker with its own box and stream.  Job j starts in list order, not before
 * arrival_s[j] seconds after the call (NULL: all at once), and, with
 * hbm_budget > 0, not before its box (mcac_sim_device_bytes) fits in what the
 * running boxes leave of hbm_budget bytes (a job larger than the budget runs
 * alone).  Per job: start and end time (s since the call), MC steps, status;
 * *peak_bytes (may be NULL) the most device memory the admitted boxes held
 * at once (0 without a budget).  Returns the first non-zero job status, else 0. */
int mcac_batch_run(const char *const *ini_texts, const double *arrival_s, int64_t n_jobs, int workers, int device,
                   int64_t max_steps, int64_t hbm_budget, double *start_s, double *end_s, int64_t *steps, int *status,
                   int64_t *peak_bytes);
/* The same runs in lock step: each worker drives up to boxes_per_worker boxes
 * that share one stream (mcac_group); per round it picks up to `depth` (1..8)
 * steps ahead for every box, as the pipelined loop does, and enqueues them all
 * in one launch, one workgroup running each box's steps in order.  Same
 * outputs, job order and admission as mcac_batch_run, plus each job's final
 * physical time (final_time may be NULL); each box's steps and results are
 * those of a run on its own. */
int mcac_batch_run_groups(const char *const *ini_texts, const double *arrival_s, int64_t n_jobs, int workers,
                          int boxes_per_worker, int depth, int device, int64_t max_steps, int64_t hbm_budget,
                          double *start_s, double *end_s, int64_t *steps, int *status, double *final_time,
                          int64_t *peak_bytes);

#ifdef __cplusplus
}
#endif
#endif
