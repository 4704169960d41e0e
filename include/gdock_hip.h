/*
 * gdock_hip.h — C ABI of the MI355X (gfx950) population-scoring and FCC
 * clustering library.  Plain pointers and sizes only; no HIP or torch types
 * cross this boundary ("stream" arguments are hipStream_t passed as void*).
 *
 * Every entry point returns a gdock_status; on failure the context's
 * gdock_last_error() holds a message.  Nothing aborts across the ABI.
 * Host arrays passed in are copied; output arrays are caller-allocated.
 * A context is not thread safe: use one per host thread (or lock around it).
 *
 * Reference interfaces replaced (gdock 2.0.0-rc.2, Rust):
 *   Population::eval_fitness ............ src/population.rs:48-61
 *   Chromosome::fitness / apply_genes ... src/chromosome.rs:79-128
 *   fitness::{vdw,elec,desolv,air}_energy, satisfaction_ratio
 *                                         src/fitness.rs:86,122,205,251,284
 *   restraints::create_restraints_from_pairs  src/restraints.rs:44-77
 *   clustering::{calculate_contacts, cluster_structures}
 *                                         src/clustering.rs:62,265
 *   structure::read_pdb (+ toppar)  ..... src/structure.rs:187-300
 *   utils::position_ligand .............. src/utils.rs:77-95
 */
#ifndef GDOCK_HIP_H
#define GDOCK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gdock_status {
  GDOCK_OK = 0,
  GDOCK_ERR_INVALID_ARG = 1, /* bad size, null pointer, NaN gene, unknown handle */
  GDOCK_ERR_OOM = 2,         /* device allocation failed / over the memory budget */
  GDOCK_ERR_DEVICE = 3,      /* HIP runtime error */
  GDOCK_ERR_BUSY = 4,        /* admission refused: no slot or memory right now */
  GDOCK_ERR_NOT_FOUND = 5    /* unknown job id */
} gdock_status;

/* One parameterized PDB ATOM record; mirrors structure::Atom
 * (src/structure.rs:28-49).  Strings are trimmed and NUL-terminated. */
typedef struct gdock_atom {
  double x, y, z;     /* Angstrom, FP64 */
  double charge;      /* e */
  double vdw_radius;  /* Angstrom (H 1.2, C 1.7, N 1.55, O 1.52, else 1.0) */
  double epsilon;     /* kcal/mol, <= 0 */
  double rmin2;       /* Angstrom */
  float occupancy;    /* PDB columns 55-60 (1.0 when absent) */
  float tempfactor;   /* PDB columns 61-66 (0.0 when absent) */
  int32_t serial;
  int16_t resseq;
  char chainid, altloc, icode;
  char name[5];       /* atom name, e.g. "CA" */
  char resname[5];
  char element[5];    /* inferred from name[0] when the PDB column is blank */
} gdock_atom;

/* constants::EnergyWeights (src/constants.rs:28-34); defaults 0.4/0.05/3.4/1.0 */
typedef struct gdock_weights { double vdw, elec, desolv, air; } gdock_weights;

/* chromosome::Chromosome (src/chromosome.rs:12-24): genes = alpha, beta,
 * gamma (rad), dx, dy, dz (Angstrom).  eval_fitness fills the energy fields. */
typedef struct gdock_chromosome {
  double genes[6];
  double fitness, vdw, elec, desolv, air, restraint_penalty, rmsd, fnat;
} gdock_chromosome;

/* clustering::ClusteringConfig (src/clustering.rs:15-35); defaults 5.0/0.60/0.75/4 */
typedef struct gdock_clustering_config {
  double contact_distance, fcc_cutoff, strictness;
  int64_t min_cluster_size;
} gdock_clustering_config;

/* clustering::ClusterResult (src/clustering.rs:39-46); members live in a
 * caller array at [members_offset, members_offset + size). */
typedef struct gdock_cluster_result {
  int64_t center_idx, size, members_offset;
} gdock_cluster_result;

typedef struct gdock_ctx gdock_ctx;         /* one device + stream + scratch */
typedef struct gdock_complex gdock_complex; /* receptor + positioned ligand, device resident */

const char* gdock_version(void);
gdock_status gdock_ctx_create(int device, gdock_ctx** out);
void gdock_ctx_destroy(gdock_ctx* ctx);
const char* gdock_last_error(const gdock_ctx* ctx);

/* ---- input formats (host) ------------------------------------------------ */
/* structure::read_pdb over text (src/structure.rs:187-300): parses ATOM
 * records, infers elements and radii and looks up the OPLS parameters
 * (src/toppar.rs).  *atoms holds every model back to back; model m spans
 * [(*model_offsets)[m], (*model_offsets)[m+1]).  Free both with gdock_free. */
gdock_status gdock_read_pdb(const char* text, size_t len, gdock_atom** atoms,
                            int32_t** model_offsets, int32_t* n_models);
void gdock_free(void* p);
/* utils::position_ligand (src/utils.rs:77-95), in place on lig. */
gdock_status gdock_position_ligand(const gdock_atom* rec, int32_t n_rec, gdock_atom* lig,
                                   int32_t n_lig);

/* ---- complexes ------------------------------------------------------------- */
/* Uploads receptor and (already positioned) ligand and resolves restraint
 * residue pairs (receptor resseq, ligand resseq) to their first "CA" atoms as
 * restraints::create_restraints_from_pairs does; unresolved pairs are dropped
 * (src/restraints.rs:68-73).  n_restraints() reports how many were kept. */
gdock_status gdock_complex_create(gdock_ctx* ctx, const gdock_atom* rec, int32_t n_rec,
                                  const gdock_atom* lig, int32_t n_lig,
                                  const int32_t* restraint_pairs, int32_t n_pairs,
                                  gdock_complex** out);
void gdock_complex_destroy(gdock_ctx* ctx, gdock_complex* cx);
int32_t gdock_complex_n_restraints(const gdock_complex* cx);

/* ---- population scoring (the drop-in for Population::eval_fitness) ------- */
/* Scores every chromosome in place: vdw, elec, desolv, air, restraint_penalty
 * and fitness, as Chromosome::fitness (src/chromosome.rs:85-128, normal mode).
 * Synchronous. */
gdock_status gdock_eval_fitness(gdock_ctx* ctx, gdock_complex* cx, gdock_chromosome* chromosomes,
                                int32_t n, const gdock_weights* w);

/* Flat form.  genes: n x 6.  out: n x 6 = {vdw, elec, desolv, air,
 * restraint_penalty, fitness}.  n_satisfied (optional): satisfied restraints. */
gdock_status gdock_score_poses(gdock_ctx* ctx, gdock_complex* cx, const double* genes, int32_t n,
                               const gdock_weights* w, double* out, int32_t* n_satisfied);

/* commands::score (src/commands/score.rs:32-49): scores ligand coordinates as
 * given (no transform) for n models of the complex's ligand topology.
 * lig_xyz: n x n_lig x 3 FP64; restraints resolve as for the complex. */
gdock_status gdock_score_coords(gdock_ctx* ctx, gdock_complex* cx, const double* lig_xyz, int32_t n,
                                const gdock_weights* w, double* out, int32_t* n_satisfied);

/* Device-resident form for callers that keep populations in HBM.  A pose
 * record is 16 doubles: the 6 genes, then cos/sin of alpha, beta, gamma as
 * gdock_prepare_poses computes them on the host (glibc, like the reference),
 * then the rotation's four pure trigonometric products of structure.rs:96-101
 * (-sin a cos g + cos a sin b sin g, sin a sin g + cos a sin b cos g,
 * cos a cos g + sin a sin b sin g, -cos a sin g + sin a sin b cos g), FP64
 * in that operation order.
 * d_poses, d_out (n x 6) and d_nsat (n) are device pointers; the launch is
 * queued on `stream` (hipStream_t, 0 = the context's stream) and not waited on. */
gdock_status gdock_prepare_poses(const double* genes, int32_t n, double* poses);
gdock_status gdock_score_poses_device(gdock_ctx* ctx, gdock_complex* cx, const double* d_poses,
                                      int32_t n, const gdock_weights* w, double* d_out,
                                      int32_t* d_nsat, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): gdock_score_poses_device with the same kernels and
 * results, but it allocates nothing.  Its per-pose scratch is a static device
 * array of 256 KiB per GPU, resolved when a context is created on that GPU, so
 * n is at most 65536 (GDOCK_ERR_INVALID_ARG beyond).  Calls that overlap on one
 * GPU share that scratch and race: serialise them (one stream, or events). */
gdock_status gdock_score_poses_device_noalloc(gdock_ctx* ctx, gdock_complex* cx, const double* d_poses,
                                              int32_t n, const gdock_weights* w, double* d_out,
                                              int32_t* d_nsat, void* stream);

/* ---- batched scoring (many complexes, one persistent launch) ------------- */
/* Population::eval_fitness (src/population.rs:48-61) for several complexes at
 * once: one generation of each complex's population, called where
 * src/commands/run.rs:382,554 call eval_fitness, for many jobs together.
 * The batch fixes the complexes and their pose counts; complex i owns pose
 * rows [off_i, off_i + poses_per_complex[i]) in complex order.  At creation
 * the library measures, per complex shape, how many ligand atoms each lane
 * holds (1-8); every launch is one replay of a captured HIP graph, which
 * plans the launch's work from the poses themselves (an expensive pose is
 * split by receptor range, the work runs most expensive first).  A pose's
 * result depends only on the pose and its complex.  Complexes must outlive
 * the batch.  Results equal gdock_score_poses within the scoring tolerance. */
typedef struct gdock_batch gdock_batch;
gdock_status gdock_batch_create(gdock_ctx* ctx, gdock_complex* const* complexes, int32_t n,
                                const int32_t* poses_per_complex, gdock_batch** out);
void gdock_batch_destroy(gdock_ctx* ctx, gdock_batch* b);
int32_t gdock_batch_n_poses(const gdock_batch* b);
/* Chosen ligand atoms per lane (n entries), the most workgroups one pose may
 * be split over (n entries, at most 8; a function of the complex's shape),
 * persistent grid and LDS bytes.  Any output may be null. */
gdock_status gdock_batch_plan(const gdock_batch* b, int32_t* lanes_k, int32_t* parts, int32_t* grid,
                              int32_t* lds_bytes);
/* Work items of the last launch (one per pose plus one per extra part of
 * each split pose); waits for the batch's work on its context stream. */
gdock_status gdock_batch_last_items(const gdock_batch* b, int32_t* n_items);
/* Synchronous: genes (n_poses x 6) in, out (n_poses x 6) and optional
 * n_satisfied out, as gdock_score_poses. */
gdock_status gdock_batch_score(gdock_ctx* ctx, gdock_batch* b, const double* genes, const gdock_weights* w,
                               double* out, int32_t* n_satisfied);
/* Device-resident form: the batch's own HBM buffers (pose records of 16
 * doubles as gdock_prepare_poses writes them, outputs n_poses x 6, satisfied
 * counts) and an asynchronous graph replay on `stream`. */
gdock_status gdock_batch_device_buffers(gdock_batch* b, double** d_poses, double** d_out, int32_t** d_nsat);
gdock_status gdock_batch_launch(gdock_ctx* ctx, gdock_batch* b, const gdock_weights* w, void* stream);
/* Ligand atoms per lane the library uses for a complex (a pure function of
 * its ligand and receptor sizes, read from the library's committed table of
 * measured choices), and the autotuner's measurement: the mean time (ms) of
 * one plain scoring launch of n_poses fixed probe poses with lanes_k ligand
 * atoms per lane (1-8), over reps launches after one warm-up.  No
 * counterpart in the reference (its CPU loop has no such parameter). */
int32_t gdock_choose_lanes(const gdock_complex* cx);
gdock_status gdock_time_lanes(gdock_ctx* ctx, gdock_complex* cx, int32_t lanes_k, int32_t n_poses, int32_t reps,
                              float* ms);

/* ---- GA docking job (host driver, device scoring) ------------------------ */
/* commands::run (src/commands/run.rs:227-733) from the positioned complex on:
 * initial population (src/chromosome.rs:27-54), per generation
 * Population::eval_fitness on the device, Hall of Fame (run.rs:77-216),
 * early stopping (run.rs:480-531), Population::evolve (src/population.rs:
 * 74-158, src/ga.rs:7-39, src/chromosome.rs:56-77), then FCC clustering of
 * the Hall of Fame (run.rs:665-693) and model selection (run.rs:694-733).
 * The random stream is this library's (SplitMix64-seeded xoshiro256**,
 * Box-Muller normals), not rand's StdRng: runs are reproducible per seed. */
typedef struct gdock_ga_params {
  int32_t population_size;   /* 150 (constants.rs:2) */
  int32_t max_generations;   /* 250 */
  int32_t elitism;           /* 5 */
  int32_t tournament_size;   /* 3 */
  int32_t convergence_window;/* 10 */
  int32_t early_stopping;    /* 1 */
  int32_t cluster;           /* 1: FCC-cluster the Hall of Fame and select models */
  int32_t keep_trace;        /* 1: keep every generation's genes and fitness */
  double mutation_rate;      /* 0.1 */
  double crossover_rate;     /* 0.6 */
  double convergence_threshold; /* 0.001 */
  double max_displacement;   /* 50.0 */
} gdock_ga_params;
typedef struct gdock_ga_result gdock_ga_result;
gdock_status gdock_ga_default_params(gdock_ga_params* p);
/* One job, one plain scoring launch per generation (the drop-in path).  A
 * population under 512 poses is split over receptor ranges so that each
 * launch holds 512 workgroups: its energies then sum in another order than
 * gdock_score_poses's for the same pose (the same tolerance; a run gives the
 * same bits every time, and gdock_ga_run_device splits alike). */
gdock_status gdock_ga_run(gdock_ctx* ctx, gdock_complex* cx, const gdock_weights* w, uint64_t seed,
                          const gdock_ga_params* params, gdock_ga_result** out);
/* The same job with every generation after the first drawn on the device:
 * the GA kernel takes the scores, updates the Hall of Fame, tests
 * convergence, breeds the next population with the same stream and writes
 * its pose records, so the host only watches for the end.  The result equals
 * gdock_ga_run's bit for bit.  population_size at most 1024. */
gdock_status gdock_ga_run_device(gdock_ctx* ctx, gdock_complex* cx, const gdock_weights* w, uint64_t seed,
                                 const gdock_ga_params* params, gdock_ga_result** out);
/* Debug mode (src/chromosome.rs:97-108, `--debug` with a reference): the
 * fitness is -DockQ from gdock_calc_metrics, the energy terms are zero. */
typedef struct gdock_evaluator gdock_evaluator; /* see the DockQ evaluator section */
gdock_status gdock_ga_run_dockq(gdock_ctx* ctx, gdock_complex* cx, gdock_evaluator* ev, uint64_t seed,
                                const gdock_ga_params* params, gdock_ga_result** out);
void gdock_ga_free(gdock_ga_result* r);
gdock_status gdock_ga_info(const gdock_ga_result* r, int32_t* generations, int32_t* converged_at,
                           int32_t* hof_size, int32_t* n_clusters, int32_t* n_selected);
/* With keep_trace: generations x population genes (x6) and fitness, and the
 * index of each generation's best. */
gdock_status gdock_ga_trace(const gdock_ga_result* r, double* genes, double* fitness, int32_t* best_idx);
int32_t gdock_ga_hof(const gdock_ga_result* r, double* genes, double* fitness, int32_t* generation);
int32_t gdock_ga_clusters(const gdock_ga_result* r, int64_t* center, int64_t* size);
int32_t gdock_ga_selected(const gdock_ga_result* r, int64_t* hof_index, int64_t* cluster_size);
/* The population the generation loop of src/commands/run.rs:381-531 leaves
 * behind, which `--no-clust` scores (run.rs:553-554): after an early stop the
 * last population scored; after the last generation the loop has evolved
 * once more (run.rs:530), and that unscored population is bred on the first
 * call.  genes: population_size x 6 (may be null); returns the population
 * size, or -1 for a null or unfinished result. */
int32_t gdock_ga_final_population(gdock_ga_result* r, double* genes);
/* The first n outputs of the GA's random stream for a seed. */
void gdock_rng_stream(uint64_t seed, int32_t n, uint64_t* out);
/* The first n normal deviates (mean + sd z) the GA's stream draws for a seed. */
void gdock_rng_normals(uint64_t seed, int32_t n, double mean, double sd, double* out);
/* The stream's own functions behind its normals, element-wise: which 0: ln(x)
 * (x > 0, normal); 1: cos(2 pi x) (0 <= x < 1); 2: the correctly rounded
 * square root. */
void gdock_stream_math(const double* x, int32_t n, int32_t which, double* out);

/* ---- output stage ------------------------------------------------------------ */
/* Atom::to_pdb_string over n atoms, concatenated as write_pdb writes them
 * (src/structure.rs:159-176,303-310): one "ATOM  " record per atom with the
 * reference's field widths, alignment and rounding.  *len receives the text
 * length; with out non-null the text and a terminating NUL are written when
 * cap > *len (else GDOCK_ERR_INVALID_ARG). */
gdock_status gdock_format_pdb(const gdock_atom* atoms, int32_t n, char* out, size_t cap, size_t* len);
/* Chromosome::apply_genes (src/chromosome.rs:79-83): the ligand rotated about
 * its centre of mass by genes[0..3] and displaced by genes[3..6], FP64 in the
 * reference's operation order with glibc cos/sin; every other field copied. */
gdock_status gdock_apply_genes(const gdock_atom* lig, int32_t n, const double* genes, gdock_atom* out);
/* The files commands::run writes after clustering (src/commands/run.rs:
 * 738-880) into out_dir (created if missing): model_<i>.pdb for each selected
 * model (Hall of Fame index, cluster size, as gdock_ga_selected returns them),
 * ranked_<i>.pdb for the five best Hall of Fame entries by fitness, each the
 * complex's receptor followed by its positioned ligand moved by the entry's
 * genes, and metrics.tsv (model, cluster_size, score; with an evaluator also
 * dockq, rmsd, irmsd, fnat), every number with four decimals. */
gdock_status gdock_write_run_outputs(gdock_ctx* ctx, gdock_complex* cx, gdock_evaluator* ev, const double* hof_genes,
                                     const double* hof_fitness, int32_t n_hof, const int64_t* selected_hof_index,
                                     const int64_t* selected_cluster_size, int32_t n_selected, const char* out_dir);
/* The `--no-clust` outputs (src/commands/run.rs:549-637) into out_dir: the
 * n poses (the final population, gdock_ga_final_population) are scored with
 * w; best_by_score.pdb is the first pose of least fitness.  With an
 * evaluator also best_by_dockq.pdb (the last pose of greatest DockQ, as
 * Iterator::max_by picks) and metrics.tsv (model, dockq, rmsd, irmsd, fnat,
 * score for the two models, four decimals). */
gdock_status gdock_write_best_outputs(gdock_ctx* ctx, gdock_complex* cx, gdock_evaluator* ev, const gdock_weights* w,
                                      const double* genes, int32_t n, const char* out_dir);
/* The same outputs in the reference's debug mode (`--debug --no-clust`,
 * src/commands/run.rs:549-637 with the debug fitness of
 * src/chromosome.rs:97-108): each pose's fitness is -DockQ (energies zero),
 * so best_by_score.pdb is the first pose of greatest DockQ and the score
 * column of metrics.tsv is -DockQ.  The evaluator is required. */
gdock_status gdock_write_best_outputs_dockq(gdock_ctx* ctx, gdock_complex* cx, gdock_evaluator* ev,
                                            const double* genes, int32_t n, const char* out_dir);

/* ---- DockQ evaluator --------------------------------------------------------- */
/* Evaluator::new (src/evaluator.rs:50-66) for the complex's receptor and a
 * reference (native) ligand as read: interface residues within 5 A, the
 * interface atoms and the native contacts.  Models are the complex's
 * positioned ligand moved by genes, as Population::eval_metrics
 * (src/population.rs:63-71) builds them; the reference must list the ligand's
 * atoms in the same order (the reference zips them). */
gdock_status gdock_evaluator_create(gdock_ctx* ctx, gdock_complex* cx, const gdock_atom* reference, int32_t n_ref,
                                    gdock_evaluator** out);
void gdock_evaluator_destroy(gdock_ctx* ctx, gdock_evaluator* ev);
gdock_status gdock_evaluator_info(const gdock_evaluator* ev, int32_t* n_receptor_interface_res,
                                  int32_t* n_reference_interface_res, int32_t* n_native_contacts);
/* Evaluator::calc_metrics (src/evaluator.rs:68-145) for n poses: metrics is
 * n x 4 = {rmsd (L-RMSD), irmsd, fnat, dockq}. */
gdock_status gdock_calc_metrics(gdock_ctx* ctx, gdock_evaluator* ev, const double* genes, int32_t n, double* metrics);

/* ---- docking server (continuous batching) -------------------------------- */
/* Many commands::run jobs (src/commands/run.rs:227-733) served concurrently
 * on one GPU.  gdock_server_submit queues a job (receptor and raw ligand
 * atoms as structure::read_pdb gives them, restraint residue pairs, seed, GA
 * parameters); a worker thread admits queued jobs as slots and device memory
 * free up, keeps each admitted job's complex resident in HBM, scores the
 * current generation of every active job in one persistent launch per step
 * (one graph, captured at creation), advances each job's GA, and clusters a
 * finished job's Hall of Fame.  A job's result is the same whichever jobs
 * share its launches.  The server owns its own context on `device`. */
typedef struct gdock_server_config {
  int32_t max_jobs;            /* active job slots (16) */
  int32_t max_population;      /* largest population a job may use (150; at most 65536, 1024 with
                                  ga_on_device; max_jobs x max_population x 8 at most 2^30) */
  int32_t max_receptor_atoms;  /* sizes the kernel's LDS receptor table (8192) */
  int32_t ga_on_device;        /* 0: the GAs advance on host threads between steps (default, faster);
                                  1: on the device (gdock_ga_run_device's kernel), population state
                                  resident in HBM, no genes or scores cross to the host */
  int64_t memory_budget;       /* device bytes for resident jobs; 0: half the free memory */
  gdock_weights weights;       /* one energy weighting for every job */
} gdock_server_config;
typedef struct gdock_job_stats {
  double submit_s, admit_s, finish_s;  /* seconds since the server started */
  int32_t generations, slot;
} gdock_job_stats;
typedef struct gdock_server_stats_t {
  int64_t steps, poses_scored, jobs_finished, jobs_queued;
  int32_t peak_active, reserved;
  double busy_s;               /* host time spent in scoring steps */
  int64_t bytes_in_use, memory_budget;
  double kernel_s;             /* device time of the scoring graph replays (HIP events) */
  double admit_s;              /* time preparing admitted jobs (positioning, complex, lanes), on a thread of its own */
  double ga_s;                 /* time advancing the GAs after each step: host time on the scoring
                                  thread, or device time of the GA launches (ga_on_device) */
  double finish_s;             /* time finishing jobs (FCC of the Hall of Fame, model selection), on a thread of its own */
} gdock_server_stats_t;
typedef struct gdock_server gdock_server;
gdock_status gdock_server_default_config(gdock_server_config* c);
gdock_status gdock_server_create(int device, const gdock_server_config* config, gdock_server** out);
/* Stops the server's threads.  Jobs still queued or running end as failed
 * (GDOCK_ERR_NOT_FOUND): threads blocked in gdock_server_wait return that
 * status, and destroy returns once the last of them has left.  No call may
 * name the server after destroy returns. */
void gdock_server_destroy(gdock_server* s);
/* Thread safe.  The ligand is positioned on the receptor (utils::
 * position_ligand) at admission.  params: NULL for the defaults. */
gdock_status gdock_server_submit(gdock_server* s, const gdock_atom* rec, int32_t n_rec, const gdock_atom* lig,
                                 int32_t n_lig, const int32_t* restraint_pairs, int32_t n_pairs, uint64_t seed,
                                 const gdock_ga_params* params, int64_t* job_id);
/* Blocks until the job ends (timeout_s < 0: no limit; GDOCK_ERR_BUSY on
 * timeout).  On success *result (free with gdock_ga_free) holds the job's GA
 * and clusters; the job id is released either way once it has ended. */
gdock_status gdock_server_wait(gdock_server* s, int64_t job_id, double timeout_s, gdock_ga_result** result,
                               gdock_job_stats* stats);
gdock_status gdock_server_stats(gdock_server* s, gdock_server_stats_t* out);

/* Diagnostic (no counterpart in the reference; for measurement): runs the
 * counting instantiation of the plain scoring kernel over the poses of n gene
 * vectors and returns counts[0] = receptor x ligand lane-pairs the kernel
 * steps through after its culling (64 per wave step, padding lanes
 * included) and counts[1] = those within 15 A, against the brute-force
 * n_rec x n_lig x n.  Scores are not returned. */
gdock_status gdock_count_pairs(gdock_ctx* ctx, gdock_complex* cx, const double* genes, int32_t n, int64_t* counts);

/* ---- FCC clustering ---------------------------------------------------------- */
/* Contact bitsets (clustering::calculate_contacts, src/clustering.rs:62-115) of
 * the complexes receptor + (ligand moved by genes i), for any chain layout:
 * residues are (chain, resseq) keys over the heavy atoms of both molecules,
 * and every key pair of different chains is a candidate (receptor-receptor
 * and ligand-ligand pairs included; a chain ID shared by receptor and
 * ligand merges their residues of equal number, as the reference does).
 * Bit b of a pose's set stands for the residue pair
 * gdock_contact_pair(cx, contact_distance, b).  The bit layout depends on the
 * cutoff: gdock_contact_words gives the uint64 words per pose (-1: cutoff not
 * supported), gdock_contact_candidates the bits.  bitsets: n x words uint64;
 * set_sizes (optional): n.  With a ligand of several chains, poses whose
 * translation exceeds a bound (about 10^6 A for a 5 A cutoff) are rejected:
 * beyond it the rounding of the pose transform could change a ligand-ligand
 * decision that the library takes once per cutoff. */
int32_t gdock_contact_words(gdock_complex* cx, double contact_distance);
int32_t gdock_contact_candidates(gdock_complex* cx, double contact_distance);
gdock_status gdock_contact_pair(gdock_complex* cx, double contact_distance, int32_t bit, char* chain_a,
                                int16_t* res_a, char* chain_b, int16_t* res_b);
gdock_status gdock_contacts(gdock_ctx* ctx, gdock_complex* cx, const double* genes, int32_t n,
                            double contact_distance, uint64_t* bitsets, int32_t* set_sizes);

/* clustering::cluster_structures (src/clustering.rs:265-293) over the
 * complexes receptor + (ligand moved by genes i), i < n.  out: up to n
 * results in the reference's order; members: up to n indices. */
gdock_status gdock_cluster_structures(gdock_ctx* ctx, gdock_complex* cx, const double* genes, int32_t n,
                                      const gdock_clustering_config* cfg, gdock_cluster_result* out,
                                      int64_t* members, int32_t* n_clusters);

#ifdef __cplusplus
}
#endif
#endif /* GDOCK_HIP_H */
