/*
 * kmcg.h -- C ABI of the MI355X kMC engine (libkmcgpu.so).
 *
 * The engine runs exact serial-trajectory lattice kMC (kmos's local_smart
 * and lat_int step loops) for many independent replicas at once on one GPU.
 * Every entry point returns an int status (KMCG_OK or a negative error);
 * per-replica trajectory outcomes are reported separately as replica status
 * codes (KMCG_REPLICA_*).  No entry point aborts the process.
 *
 * Numbering follows the reference Fortran: processes 1..P, sites 1..V,
 * species 0..S-1 in name order, null species -1 (one layer) or S (several
 * layers: the sites of layers that are not populated,
 * kmos/fortran_src/proclist_constants.mpy:61-68).
 *
 * Reference interfaces these replace (paths relative to the kmos tree):
 *   f2py module kmc_model.{base,lattice,proclist}  kmos/utils/__init__.py:406-521
 *   WASM C bindings kmc_*                          kmos/utils/__init__.py:1181-1417
 *   the generated step loop                        kmos/fortran_src (.mpy templates)
 *
 * Model blob (int32 words, written by kmos_amd/compiler.py to_blob()):
 *   header[32]: 0 magic 'KMCG', 1 version, 2 backend (0 local_smart,
 *   1 lat_int), 3 nr_of_proc, 4 nr_of_species, 5 null_species, 6 spuck,
 *   7 dimension, 8 off_site_default, 9 off_proc_table, 10 n_routines,
 *   11 off_routine_table, 12 off_touchup, 13 n_groups, 14 off_group_table,
 *   15 min_size (cells per periodic dimension), 16 total words,
 *   17 off_site_init (per site type: 1 = initialize_state touches it up,
 *   the sites of the initial layer).  Section layouts are documented in
 *   kmos_amd/compiler.py (_to_blob).
 */
#ifndef KMCG_H
#define KMCG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes of entry points */
#define KMCG_OK 0
#define KMCG_EINVAL (-1)        /* bad argument */
#define KMCG_ENOMEM (-2)        /* device allocation failed */
#define KMCG_EHIP (-3)          /* HIP runtime error */
#define KMCG_EUNSUPPORTED (-4)  /* model outside the engine's limits */

/* replica status codes (per trajectory) */
#define KMCG_REPLICA_OK 0
#define KMCG_REPLICA_ABSORBING 1        /* total rate is zero: no process possible */
#define KMCG_REPLICA_SPECIES_MISMATCH 2 /* replace_species found another species */
#define KMCG_REPLICA_DEL_ABSENT 3       /* lat_int del_proc of an unregistered process */
#define KMCG_REPLICA_CORRUPT 4          /* a process list would exceed the volume */

typedef struct kmcg_model kmcg_model;
typedef struct kmcg_batch kmcg_batch;

/* Text of the last error in this thread. */
const char *kmcg_last_error(void);

/* Number of visible GPUs (0 when none). */
int kmcg_device_count(int32_t *count);

/* Load a compiled model (replaces building the generated Fortran,
 * kmos/utils/__init__.py:406-521). */
int kmcg_model_load(const int32_t *blob, int64_t nwords, kmcg_model **out);
void kmcg_model_free(kmcg_model *m);
/* nr_of_proc, nr_of_species, spuck, backend, dimension */
int kmcg_model_info(const kmcg_model *m, int32_t *info5);
/* The step kernel's forms this model can run: *li = 1 when a lat_int
 * group's two segments are posted in one pass (kmc_steps_kernel_li: every
 * nli group tabulated, at most 64 processes).  No counterpart in the
 * reference (a profiling aid). */
int kmcg_model_forms(const kmcg_model *m, int32_t *li);

/* Device bytes a batch of n replicas of sizes[n][3] needs. */
int kmcg_batch_footprint(const kmcg_model *m, int32_t n, const int32_t *sizes, int64_t *bytes);

/* Allocate n replicas (system_size per replica in sizes[n][3]) on `device`;
 * allocate_system (kmos/fortran_src/lattice.mpy:212-317, base.mpy:648-744). */
int kmcg_batch_create(const kmcg_model *m, int32_t n, const int32_t *sizes,
                      int32_t device, kmcg_batch **out);
void kmcg_batch_free(kmcg_batch *b);

/* proclist.init / initialize_state for replicas [r0, r0+n): the default
 * species of every site (or, if configs is non-NULL, the given lattices,
 * one after the other, V values each in site order), empty bookkeeping and
 * the touchups in k,j,i order, clocks, procstat and integ_rates reset, and
 * the uniform stream keyed (seeds[i], streams[i]) (streams may be NULL)
 * (proclist_generic_subroutines.mpy:166-304).  A configuration holding an
 * unknown species fails the call before any replica is touched. */
int kmcg_init(kmcg_batch *b, int32_t r0, int32_t n, const uint32_t *seeds,
              const uint32_t *streams, const int8_t *configs);

/* base.set_rate_const for every process of replicas [r0, r0+n); rates is
 * [n][P], finite and >= 0 (base.mpy:581-600). */
int kmcg_set_rates(kmcg_batch *b, int32_t r0, int32_t n, const double *rates);
int kmcg_get_rates(kmcg_batch *b, int32_t r, double *rates);

/* KMC_Model._set_configuration + _adjust_database for replicas
 * [r0, r0+n): species per site (config holds the n lattices one after the
 * other, each V values in site order), then touchups with x outermost
 * (kmos/run/__init__.py:1408-1457).  A configuration holding an unknown
 * species fails the call before any lattice is replaced. */
int kmcg_set_config(kmcg_batch *b, int32_t r0, int32_t n, const int8_t *config);

/* proclist.do_kmc_steps(steps) for the listed replicas
 * (proclist_generic_subroutines.mpy:1-44).  All replicas advance in one
 * launch; a replica stops early when its status becomes nonzero.  If
 * `budget` is non-NULL, replica replicas[i] runs min(steps, budget[i])
 * steps and budget[i] is decreased by the steps it executed.  Returns once
 * the work is queued on the batch's stream (after waiting for the batch's
 * previous run, if any); kmcg_sync waits for it and only then writes the
 * decreased budgets back, so `budget` must stay valid until kmcg_sync (or
 * any other call on the batch, all of which wait for the run first).
 * Batches have their own streams: runs of different batches overlap.
 * `replicas` names each replica at most once (KMCG_EINVAL otherwise, before
 * anything is queued). */
int kmcg_run(kmcg_batch *b, const int32_t *replicas, int32_t n, int64_t steps,
             int64_t *budget);
int kmcg_sync(kmcg_batch *b);
/* Launch configuration for kmcg_run: lattice in LDS (1) or HBM (0, slower,
 * no size limit), -1 = automatic; replica-waves per workgroup sharing one
 * copy of the op program (1..4, 0 = as many as LDS allows). */
int kmcg_set_launch(kmcg_batch *b, int32_t lattice_in_lds, int32_t waves_per_group);
/* Step kernel form for kmcg_run with one process slot per lane (P <= 64)
 * and no tracking: 1 = the next step's accumulate and search run while the
 * replay's HBM reads are in flight, 0 = after them, -1 = automatic (on for
 * local_smart models of more than 16 processes in runs of at most 8
 * replicas per compute unit).  Both forms
 * give identical trajectories.  No counterpart in the reference (a launch
 * setting of the batched engine). */
int kmcg_set_overlap(kmcg_batch *b, int32_t mode);
/* Runs through HIP graphs (1) or direct launches (0, the default).  With
 * graphs, the stream work of a kmcg_run / kmcg_run_recorded (replica list
 * and budget copies, the step kernel, copies back) is captured once per
 * distinct set of launch values and replayed while runs repeat them (same
 * replica count, steps, launch shape, buffers); the copies read the pinned
 * staging buffers, so each replay uses that run's replica list and budgets.
 * Up to 8 graphs are kept per batch; 0 frees them.  No counterpart in the
 * reference (do_kmc_steps is a host loop, proclist_generic_subroutines.mpy:
 * 1-80); it is for callers that repeat short runs. */
int kmcg_set_graph(kmcg_batch *b, int32_t enable);
/* Milliseconds the last kmcg_run took on the device (HIP events around the
 * kernel on the batch's stream). */
int kmcg_last_run_ms(kmcg_batch *b, float *ms);
/* Phase profile of the step loop, for replicas r0..r0+n-1: out[n][8] holds
 * the shader-clock ticks each replica's wave spent selecting (accumulate,
 * clock, search), reading its site, posting the process's ops, replaying
 * the lists and on the rest of its steps, then the steps counted, the
 * replay blocks (batches of list reads) and the replays; the counters are
 * cleared after the read.  Only a library built with
 * -DKMCG_PHASES=1 (libkmcgpu_phases.so) counts; the plain library returns
 * KMCG_EUNSUPPORTED.  No counterpart in the reference (a profiling aid). */
int kmcg_phase_profile(kmcg_batch *b, int32_t r0, int32_t n, uint64_t *out);

/* proclist.get_next_kmc_step: draws one step's uniforms and returns the
 * process and site it would pick, passing ran_time as the site uniform
 * (proclist_generic_subroutines.mpy:85-110). */
int kmcg_get_next_step(kmcg_batch *b, int32_t r, int32_t *proc, int32_t *site);
/* proclist.run_proc_nr(proc, site) for replica r. */
int kmcg_run_proc_nr(kmcg_batch *b, int32_t r, int32_t proc, int32_t site);
/* touchup of one site (local_smart: touchup_<layer>_<site>) or cell
 * (lat_int: touchup_cell); x, y, z 0-based, n 1-based site type. */
int kmcg_touchup(kmcg_batch *b, int32_t r, int32_t x, int32_t y, int32_t z, int32_t n);
/* base.replace_species without bookkeeping (lattice.replace_species) */
int kmcg_replace_species(kmcg_batch *b, int32_t r, int32_t site, int32_t old_sp, int32_t new_sp);

/* scalars: kmc_time, kmc_time_step (f64), kmc_step, rng counter (i64), status */
int kmcg_get_scalars(kmcg_batch *b, int32_t r, double *time2, int64_t *step2, int32_t *status);
int kmcg_get_scalars_all(kmcg_batch *b, double *kmc_time, int64_t *kmc_step, int32_t *status);
int kmcg_set_kmc_time(kmcg_batch *b, int32_t r, double t);
int kmcg_get_nr_of_sites(kmcg_batch *b, int32_t r, int32_t *out);
int kmcg_get_procstat(kmcg_batch *b, int32_t r, int64_t *out);
int kmcg_get_integ_rates(kmcg_batch *b, int32_t r, double *out);
/* update_integ_rate: integ_rates += nr_of_sites * rates * kmc_time_step
 * (base.mpy:626-645) */
int kmcg_update_integ_rate(kmcg_batch *b, int32_t r);
/* update_accum_rate then accum_rates (base.mpy:603-623) */
int kmcg_get_accum_rates(kmcg_batch *b, int32_t r, double *out);
int kmcg_get_lattice(kmcg_batch *b, int32_t r, int8_t *out);
/* avail_sites(:,:,1) as sites[P][V] and avail_sites(:,:,2) as back[P][V] */
int kmcg_get_avail(kmcg_batch *b, int32_t r, int32_t *sites, int32_t *back);
/* Displacement tracking for mean squared displacements (kmos/run/acf.py
 * initialize_msd / do_kmc_steps_displacement / get_displacement /
 * get_trajectory / set_displacement_to_zero over base_acf.mpy:217-361,
 * 825-891 and proclist_generic_subroutines_acf.mpy:143-219).
 *
 * kmcg_msd_init numbers the sites holding `trace_species` 1.. in site order
 * (initialize_mean_squared_displacement) for replicas [r0, r0+n) and zeroes
 * their displacements.  `hops` [P][9] is get_diff_sites_displacement as a
 * table (kmos/io/acf.py:443-671): per process a flag (1: the process moves
 * one particle), then the site it leaves and the site it reaches, each as
 * a cell offset dx, dy, dz from the chosen site's cell and a site type n
 * (1-based); `disp` [P][3] is each hop's displacement vector. */
int kmcg_msd_init(kmcg_batch *b, int32_t r0, int32_t n, int32_t trace_species, const int32_t *hops,
                  const double *disp);
/* allocate_trajectory: record the next nr_of_steps steps of every replica
 * run with traj_on (one event per step: particle moved, 0 if none, and its
 * new site). */
int kmcg_msd_trajectory(kmcg_batch *b, int64_t nr_of_steps);
/* do_kmc_steps_displacement(n, traj_on): kmcg_run with tracking. */
int kmcg_run_displacement(kmcg_batch *b, const int32_t *replicas, int32_t n, int64_t steps, int64_t *budget,
                          int32_t traj_on);
/* set_displacement_to_zero for replicas [r0, r0+n) */
int kmcg_msd_zero(kmcg_batch *b, int32_t r0, int32_t n);
/* nr_of_ions, site_arr[V+1], id_arr[V+1], displacement[V+1][3] (row 0 and
 * rows past nr_of_ions included), events recorded and the events [2*n]
 * (any pointer may be NULL) */
int kmcg_msd_get(kmcg_batch *b, int32_t r, int32_t *nr_of_ions, int32_t *site_arr, int32_t *id_arr,
                 double *displacement, int64_t *n_events, int32_t *events);
/* Autocorrelation of a site property along the traced particles' paths
 * (kmos/run/acf.py allocate_acf / set_property_acf / calc_product_property /
 * initialize_acf / do_kmc_steps_acf / get_acf / set_acf_to_zero over
 * base_acf.mpy:382-552,1053-1201 and proclist_generic_subroutines_acf.mpy:
 * 1-98).
 *
 * kmcg_acf_alloc: allocate_tracing_arr(nr_of_types) and
 * allocate_config_bin_acf(t_bin, t_f, safety, extending) for every replica
 * (0 = the reference's default factor 2).  kmcg_acf_set_properties:
 * property_acf per site for replicas [r0, r0+n) (V+1 values each, entry 0
 * unused) and, when non-NULL, product_property [(T+1)][(T+1)]; products must
 * be integers (|p| <= 2^20): sum(buffer_acf) is then exact as a running sum.
 * kmcg_acf_init: initialize_acf(trace_species) with the hop table of
 * kmcg_msd_init.  kmcg_run_acf: do_kmc_steps_acf(steps, traj_on).
 * kmcg_acf_get: config_bin[extended bins], counter_write_in_bin,
 * d4 = {sum(buffer_acf), t_o, wrap_count, t_bin}, i5 = {bin_index,
 * nr_of_bins, extended bins, nr_of_ions_before_kmc_step_acf, nr_of_ions},
 * buffer_acf[V+1], property_o[V+1] (any pointer may be NULL).  A step that
 * would run past the last extended bin stops there (the reference writes
 * past its array). */
int kmcg_acf_alloc(kmcg_batch *b, int32_t nr_of_types, double t_bin, double t_f, int32_t safety_factor,
                   int32_t extending_factor);
int kmcg_acf_set_properties(kmcg_batch *b, int32_t r0, int32_t n, const int32_t *property_acf,
                            const double *product_property);
int kmcg_acf_init(kmcg_batch *b, int32_t r0, int32_t n, int32_t trace_species, const int32_t *hops);
int kmcg_run_acf(kmcg_batch *b, const int32_t *replicas, int32_t n, int64_t steps, int64_t *budget,
                 int32_t traj_on);
/* set_acf_to_zero for replicas [r0, r0+n) */
int kmcg_acf_zero(kmcg_batch *b, int32_t r0, int32_t n);
int kmcg_acf_get(kmcg_batch *b, int32_t r, double *config_bin, int32_t *counter_write_in_bin, double *d4,
                 int32_t *i5, double *buffer_acf, int32_t *property_o);
/* Overwrite the whole state of replica r: the restore half of a .reload
 * snapshot (base.mpy:363-514 reload_system; its writer is save_system,
 * base.mpy:517-578).  time2 = kmc_time, kmc_time_step; step2 = kmc_step,
 * rng counter; lattice[V] species; nr_of_sites[P]; procstat[P];
 * integ_rates[P] (NULL: zero, as reload_system leaves them); sites[P][V],
 * back[P][V] laid out as kmcg_get_avail returns them.  The lists must hold
 * the invariant add_proc/del_proc keep (base.mpy:211-301): for every
 * process, sites[1..nr] distinct sites in 1..V with back[site] = position,
 * every other slot and back entry 0; anything else is rejected with
 * KMCG_EINVAL before any byte reaches the device.  The replica's status
 * is reset to running; its seed and stream are kept. */
int kmcg_set_state(kmcg_batch *b, int32_t r, const double *time2, const int64_t *step2,
                   const int8_t *lattice, const int32_t *nr_of_sites, const int64_t *procstat,
                   const double *integ_rates, const int32_t *sites, const int32_t *back);
/* Debug check (a slow path, for tests and suspicion, not for the step
 * loop): for replicas [r0, r0+n) counts in bad[i] the violations of the
 * list invariant add_proc/del_proc keep (base.mpy:211-301) and the
 * (process, site) pairs whose registration disagrees with the process's
 * conditions on the current lattice (the set the reference's touchups
 * build, kmos/io/__init__.py:2411-2443; lat_int: the nli case trees).  0
 * everywhere for a consistent state.  Synchronous. */
int kmcg_check_lists(kmcg_batch *b, int32_t r0, int32_t n, int32_t *bad);
/* proclist.get_occupation for replicas [r0, r0+n): out[n][S][spuck] */
int kmcg_get_occupation(kmcg_batch *b, int32_t r0, int32_t n, double *out);
/* procstat and integ_rates of replicas [r0, r0+n): [n][P] each */
int kmcg_get_counters(kmcg_batch *b, int32_t r0, int32_t n, int64_t *procstat, double *integ);

/* One sampling record per listed replica, KMCG_SAMPLE_WIDTH(S, spuck, P)
 * doubles each, in list order: kmc_time, kmc_time_step, kmc_step, status,
 * occupation[S][spuck] (proclist.get_occupation,
 * proclist_generic_subroutines.mpy:113-158), integ_rates[P], procstat[P]:
 * what KMC_Model.get_atoms(geometry=False) reads (kmos/run/__init__.py:
 * 680-850), gathered on the device and copied in one transfer.  Waits for
 * the batch's stream. */
#define KMCG_SAMPLE_WIDTH(S, spuck, P) (4 + (S) * (spuck) + 2 * (P))
int kmcg_sample(kmcg_batch *b, const int32_t *replicas, int32_t n, double *out);

/* kmcg_run followed, on the same stream, by kmcg_sample of the same
 * replicas: one launch and one wait per sampling interval.  The records
 * (KMCG_SAMPLE_WIDTH doubles per replica, list order) and the budgets are
 * written to `samples` / `budget` when kmcg_sync (or any other call on the
 * batch) waits for the run; both must stay valid until then. */
int kmcg_run_sampled(kmcg_batch *b, const int32_t *replicas, int32_t n, int64_t steps,
                     int64_t *budget, double *samples);

/* *done = 1 when the batch's stream has no queued work left (kmcg_sync would
 * not wait), else 0; never waits. */
int kmcg_query(kmcg_batch *b, int32_t *done);

/* kmcg_run that also records, inside the step loop, one sample record
 * (KMCG_SAMPLE_WIDTH doubles, the layout of kmcg_sample) after every
 * `sample_every` steps of each listed replica and one for its final state:
 * record j of replicas[i] at records[(i * max_records + j) * width], for
 * j < ceil(executed / sample_every) (one record if it executed no step).
 * steps <= sample_every * max_records.  This is what
 * KMC_Model.get_std_sampled_data reads per sample (kmos/run/__init__.py:
 * 865-966), without a host round trip per sample.  `records` and `budget`
 * are written when kmcg_sync (or any other call on the batch) waits. */
int kmcg_run_recorded(kmcg_batch *b, const int32_t *replicas, int32_t n, int64_t steps,
                      int64_t *budget, int32_t sample_every, int32_t max_records, double *records);

#ifdef __cplusplus
}
#endif
#endif /* KMCG_H */
