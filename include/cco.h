/*
 * cco.h -- C ABI of the MI355X (gfx950) stealthy collective-coordinate library.
 *
 * Plain pointers and sizes only.  Every function returns 0 on success or one
 * of the CCO_ERR_* codes; cco_last_error() gives the message of the last
 * failure on the calling thread.  Host pointers ("host") are C-contiguous
 * float64/int64 buffers borrowed for the duration of the call; device
 * pointers ("dev") are device allocations on the context's GPU and calls on
 * them are asynchronous on the context stream.
 *
 * Each declaration names the reference interface it replaces
 * (paths are in the sdawley1/xla-cco tree).
 */
#ifndef CCO_H
#define CCO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CCO_OK 0
#define CCO_ERR_ARG 1       /* bad argument or shape  -> ValueError   */
#define CCO_ERR_DTYPE 2     /* not float64            -> TypeError    */
#define CCO_ERR_DEVICE 3    /* HIP runtime failure    -> RuntimeError */
#define CCO_ERR_BOUND 4     /* relax round bound hit  -> RuntimeError */
#define CCO_ERR_STATE 5     /* call out of sequence   -> RuntimeError */

#define CCO_METHOD_DESCENT 0
#define CCO_METHOD_LBFGS 1
#define CCO_METHOD_FLOW 2
#define CCO_METHOD_VERLET 3

typedef struct cco_ctx cco_ctx;
typedef struct cco_energy cco_energy;
typedef struct cco_relax cco_relax;

/* Library version (major*10000 + minor*100 + patch). */
int cco_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* cco_last_error(void);

/* Number of visible GPUs.  Replaces the device probe of
 * src/xla_cco/_device.py:137 (require_gpu); no CPU fallback exists. */
int cco_device_count(int* count);

/* One HIP stream on one device; calls on a context are serialised and a
 * context is not thread-safe. */
int cco_ctx_create(int device, cco_ctx** out);
int cco_ctx_destroy(cco_ctx* ctx);
int cco_ctx_synchronize(cco_ctx* ctx);

/* Device memory helpers for the *_dev entry points. */
int cco_dev_alloc(cco_ctx* ctx, size_t bytes, void** out);
int cco_dev_free(cco_ctx* ctx, void* ptr);
int cco_memcpy_h2d(cco_ctx* ctx, void* dst_dev, const void* src_host, size_t bytes);
int cco_memcpy_d2h(cco_ctx* ctx, void* dst_host, const void* src_dev, size_t bytes);

/* Upload a half-space k-set: n (M x d int64 host), weights w (M host),
 * box volume, padding multiple.  Replaces make_stealthy_energy
 * (src/xla_cco/energy.py:361-414) and the StealthyEnergy fields (:158-162).
 * M == 0 -> CCO_ERR_ARG ("empty k-set"). */
int cco_energy_create(cco_ctx* ctx, const int64_t* n, const double* w, int M, int d,
                      double volume, int pad_to, cco_energy** out);
int cco_energy_destroy(cco_energy* e);
/* M, M_pad, d, volume, max |n_a| per axis. */
int cco_energy_info(const cco_energy* e, int* M, int* M_pad, int* d, double* volume,
                    int* nmax3);

/* Batched Phi, dPhi/ds, Re rho, Im rho for B configurations of N particles.
 * s: B x N x d; phi: B; grad: B x N x d (NULL: energy only); re, im:
 * B x M_pad (NULL: not returned; padding rows give re = N, im = 0).
 * Replaces StealthyEnergy.value_grad_density / value_and_grad / __call__
 * (src/xla_cco/energy.py:186-225) and collective_density (:327-358). */
int cco_value_grad_density(cco_energy* e, const double* s_host, int B, int N,
                           double* phi_host, double* grad_host, double* re_host,
                           double* im_host);
int cco_value_grad_density_dev(cco_energy* e, const double* s_dev, int B, int N,
                               double* phi_dev, double* grad_dev, double* re_dev,
                               double* im_dev);

/* Tile knob of an energy's FP64 density kernel (no reference counterpart; the
 * autotuner's).  CCO_EOPT_RHO_CHUNK: particles whose phase tables a rho
 * workgroup stages in LDS at a time, 16, 32 or 64 (default: the longest whose
 * tables fit 80 KB).  Each thread sums its particles in the same order for any
 * chunk, so results are bit-identical; shorter chunks leave room for more
 * workgroups per CU.  Applies to later evaluations and sessions. */
#define CCO_EOPT_RHO_CHUNK 1
/* CCO_EOPT_MP_LANE: 1 (default where the k-set fits) evaluates split-precision
 * lanes of up to about 4M column-row-particle products in one workgroup per
 * lane (rho, Phi, C and the forces in one launch, nothing staged in HBM);
 * 0 keeps them on the tiled split-precision kernels.  The two differ in
 * summation order (within the split-precision error bound, not bitwise). */
#define CCO_EOPT_MP_LANE 2
/* CCO_EOPT_MP_LAYOUT (d = 3): 1 (default where it fits) runs the tiled
 * split-precision kernels on the pair layout (columns (n1, n2), rows n3), about
 * half the operand products of the column layout (columns (n2, n3), rows n1);
 * 0 keeps the column layout.  Summation orders differ (within the
 * split-precision error bound, not bitwise); get returns the layout in use. */
#define CCO_EOPT_MP_LAYOUT 3
/* CCO_EOPT_RHO_SPLIT: particles per density split of a uniform batch, 0 (the
 * default: 128 up to N = 2048, else the multiple of 128 giving at most 16
 * splits) or a multiple of 128 up to 65536.  A value longer than the default
 * gives fewer, longer splits (less partial-sum traffic, fewer workgroups); a
 * shorter one acts as the default, whose split count the partial-sum buffers
 * hold.  The splits' sums are added in a different grouping, so results
 * change at rounding level (not bitwise).  Get returns the requested value. */
#define CCO_EOPT_RHO_SPLIT 4
int cco_energy_set_option(cco_energy* e, int key, int64_t value);
int cco_energy_get_option(const cco_energy* e, int key, int64_t* value);

/* Allocation-free variant of cco_value_grad_density_dev.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results as
 * cco_value_grad_density_dev (FP64 only: an energy set to split precision
 * gives CCO_ERR_STATE); the caller passes every input and output, and the
 * partial sums live in a fixed device array of cco_noalloc_scratch_bytes()
 * bytes that the library holds on each GPU until the process ends.  A batch
 * needing more (about 16 B x N/128 x box entries for the rho partials) gives
 * CCO_ERR_ARG before any work.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (two contexts, two threads) race. */
int cco_value_grad_density_dev_noalloc(cco_energy* e, const double* s_dev, int B, int N,
                                       double* phi_dev, double* grad_dev, double* re_dev,
                                       double* im_dev);
size_t cco_noalloc_scratch_bytes(void);

/* Settings of the two minimizers (src/xla_cco/minimize/descent.py:52-63,
 * src/xla_cco/minimize/lbfgs.py:55-67). */
typedef struct {
  double step_init, step_min, shrink, grow, rel_tol, energy_tol;
  int64_t max_evals;
  int32_t record_every, max_records;
} cco_descent_settings;

typedef struct {
  int32_t memory;
  int32_t n_rescales;
  double rescales[8];
  int64_t max_evals_per_pass;
  double energy_tol, grad_tol;
  int32_t stagnation_patience, max_linesearch_steps;
  int32_t record_every, max_records;
} cco_lbfgs_settings;

/* A batched relaxation of B configurations (N particles, d dims) held on the
 * device.  e == NULL selects an external energy: the caller evaluates the
 * trial points itself (cco_relax_pending / cco_relax_submit).  Replaces
 * relax (src/xla_cco/minimize/__init__.py:102-151) with init_batch /
 * chunk_batch / finalize_batch of descent.py:204-215 and lbfgs.py:336-350. */
int cco_relax_create(cco_ctx* ctx, cco_energy* e, int B, int N, int d, int method,
                     const cco_descent_settings* ds, const cco_lbfgs_settings* ls,
                     cco_relax** out);
/* Uneven batch (the batching row of the target design; no reference
 * counterpart: the reference relaxes one (N, d, k-set) at a time and its CLI
 * loops over them, cli/relax.py:339-358).  Lane i relaxes N = lane_n[i]
 * particles under energies[lane_energy[i]] (its own d, box volume and
 * k-set); every lane shares the settings.  Positions are packed lane after
 * lane (N_i d_i doubles each) in cco_relax_init, cco_relax_result and the
 * trajectory (record_cap rows of N_i d_i per lane).  One launch per kernel
 * covers every lane; descent and L-BFGS (two-loop variants) only; the
 * energies must outlive the session. */
int cco_relax_uneven_create(cco_ctx* ctx, int n_energies, cco_energy* const* energies, int B,
                            const int32_t* lane_energy, const int32_t* lane_n, int method,
                            const cco_descent_settings* ds, const cco_lbfgs_settings* ls,
                            cco_relax** out);
int cco_relax_destroy(cco_relax* r);
/* Load starting points (B x N x d) and mark every lane for its first evaluation. */
int cco_relax_init(cco_relax* r, const double* s0_host);
int cco_relax_init_dev(cco_relax* r, const double* s0_dev);
/* Built-in energy: run until no lane is running, syncing with the host every
 * sync_every evaluations; returns CCO_ERR_BOUND if max_evals evaluations
 * pass first (the reference's round bound).  Steps are queued at most 16
 * ahead of the device, which reports its progress through mapped host
 * memory, so no step is queued once every lane has stopped. */
int cco_relax_run(cco_relax* r, int64_t max_evals, int sync_every);
/* Same, stopping after at most n_evals evaluations; *n_active = lanes left. */
int cco_relax_step(cco_relax* r, int64_t n_evals, int* n_active);
/* External energy: lanes awaiting an evaluation and their trial points
 * (lanes: n_active ints, s_trial: n_active x N x d), then submit phi and
 * grad (dPhi/ds) for those lanes in the same order. */
int cco_relax_pending(cco_relax* r, int* n_active, int32_t* lanes_host, double* s_trial_host);
int cco_relax_submit(cco_relax* r, const double* phi_host, const double* grad_host);
/* Results in input order: wrapped s (B x N x d), energy, grad_norm, n_evals,
 * n_iters, status (codes of src/xla_cco/minimize/result.py:35-42). */
int cco_relax_result(cco_relax* r, double* s_host, double* energy_host, double* grad_norm_host,
                     int64_t* n_evals_host, int64_t* n_iters_host, int64_t* status_host);
int cco_relax_result_dev(cco_relax* r, double* s_dev, double* energy_dev,
                         double* grad_norm_dev, int64_t* n_evals_dev, int64_t* n_iters_dev,
                         int64_t* status_dev);
/* Recorded rows per lane (count), then iteration, s (wrapped), energy, step
 * for B x max_records rows (src/xla_cco/minimize/_loop.py:189-248). */
int cco_relax_trajectory(cco_relax* r, int64_t* count_host, int64_t* iteration_host,
                         double* s_host, double* energy_host, double* step_host);
/* Total energy evaluations issued so far (sum over lanes of evaluation steps). */
int cco_relax_stats(cco_relax* r, int64_t* steps, int64_t* lane_evals);
/* Sums over lanes: energy evaluations (descent: iterations + 1), iterations,
 * two-loop (s, y) pairs visited (L-BFGS); then the evaluation steps queued
 * over all shards since the last init (the launches of each step kernel). */
int cco_relax_counters(cco_relax* r, int64_t* out4);
/* Tuning knobs (no reference counterpart).  CCO_OPT_ADVANCE_VARIANT: 0 auto,
 * 1 L-BFGS vectors in global memory, 2 register-resident (N*d <= 8192),
 * 3 Gram form (memory <= 128).  Takes effect at the next cco_relax_init. */
#define CCO_OPT_ADVANCE_VARIANT 1
/* CCO_OPT_SHARDS: lanes split into this many shards (0 auto, for
 * built-in-energy descent and L-BFGS sessions: one shard per 32 lanes, at
 * most 3; at most 4 when set), each run on its own stream so that one
 * shard's HBM-bound L-BFGS update overlaps another's energy evaluation.  A
 * lane's results do not depend on the shard count.  Takes effect at the next
 * init. */
#define CCO_OPT_SHARDS 2
/* CCO_OPT_GRAPH: built-in-energy descent and L-BFGS sessions queue each
 * shard's evaluation steps as replays of a HIP graph holding this many steps
 * (0: plain launches, the default; at most 64).  A graph is instantiated per
 * shard and launch bound (rounded up to a power of two) and kept until the
 * next init; profiled steps and a step count that is not a multiple left in
 * a cco_relax_step call use plain launches.  Results are bit-identical to
 * plain launches.  Takes effect at the next step. */
#define CCO_OPT_GRAPH 3
/* Read-only (cco_relax_get_option): graph replays since the session was
 * created. */
#define CCO_OPT_GRAPH_LAUNCHES 4
/* CCO_OPT_SPLIT_HISTORY: L-BFGS sessions in split precision on the
 * register-resident advance: 0 (default) keeps every (s, y) pair in FP64;
 * 1 stores a split-precision lane's pairs in FP32 (the two-loop arithmetic
 * stays FP64), which halves the history the two-loop streams, and widens
 * them to FP64 when the lane turns FP64 (a rescue empties them instead).
 * FP64 lanes are unaffected.  Takes effect at the next init. */
#define CCO_OPT_SPLIT_HISTORY 5
int cco_relax_set_option(cco_relax* r, int key, int64_t value);
/* The value in force: for CCO_OPT_ADVANCE_VARIANT the variant latched by the
 * last cco_relax_init (1, 2 or 3; 0 before the first init); for
 * CCO_OPT_SHARDS the shards of the last init; for CCO_OPT_GRAPH the steps per
 * graph. */
int cco_relax_get_option(cco_relax* r, int key, int64_t* value);
/* Split-precision evaluations (no reference counterpart; the precision row of
 * the target design).  mode 1: rho and the forces run as complex GEMMs on the
 * FP16 matrix cores with every operand split hi + lo (about 22 significant
 * bits) and FP32 accumulation; the density partials, Phi and every minimizer
 * and integrator update stay FP64.  L-BFGS lanes start split and turn FP64 for
 * good at the first iteration whose energy is <= switch_phi (1e-8 by default;
 * 0 keeps them split); DOPRI5 gradient-flow lanes likewise at the first step
 * that starts at energy <= switch_phi (ROS2 flows stay FP64); MD lanes stay
 * split.  mode 0: FP64 (the default).
 * Takes effect at the next init; a k-set that does not fit the MFMA kernels
 * (more than 48 rows per column) runs FP64. */
int cco_relax_set_precision(cco_relax* r, int mode, double switch_phi);
/* Whether the last init enabled split precision, and how many lanes are
 * still split now (synchronises). */
int cco_relax_precision_state(cco_relax* r, int* active_mode, int* lanes_split);
/* Split-precision cco_value_grad_density* calls on this energy (mode 1) or
 * FP64 (0, default); *applied reports whether the k-set allows it. */
int cco_energy_set_precision(cco_energy* e, int mode, int* applied);

/* Closed-form Hessians d^2 Phi / ds_ja ds_lb of B configurations (s:
 * B x N x d), each flattened row-major to (N d) x (N d) in h.  Replaces
 * StealthyEnergy.hessian (src/xla_cco/energy.py:272-299). */
int cco_hessian(cco_energy* e, const double* s_host, int B, int N, double* h_host);
int cco_hessian_dev(cco_energy* e, const double* s_dev, int B, int N, double* h_dev);

/* Gradient flow dr/dt = -grad_r Phi in the Cartesian metric of basis
 * (d x d host, rows = lattice vectors), Dormand-Prince 5(4) with error
 * control on the RMS Cartesian step and on the clocks (t, arclength).
 * Replaces gradient_flow (src/xla_cco/minimize/flow.py:258-323, DOPRI5,
 * vmapped by the caller) and FlowSettings (:93-125; arclength = 1 selects
 * parametrization "arclength").  e == NULL: external energy through
 * cco_relax_pending / cco_relax_submit.  Run with cco_relax_run/_step. */
typedef struct {
  double rtol, atol, dt_init, dt_min, dt_max, t_max, energy_tol, grad_tol;
  int64_t max_steps;
  int32_t arclength;
  int32_t record_every, max_records;
} cco_flow_settings;

int cco_flow_create(cco_ctx* ctx, cco_energy* e, int B, int N, int d,
                    const cco_flow_settings* fs, const double* basis_host, cco_relax** out);
/* Integration method before init: 0 Dormand-Prince 5(4) (default), 1 ROS2
 * (flow.py:367-398: L-stable Rosenbrock with the exact Hessian, time
 * parametrization only, built-in energy only: each step factors
 * I - gamma h J per lane on the GPU). */
int cco_flow_set_method(cco_relax* r, int method);
/* s0: B x N x d; dt0: B first steps (NULL: dt_init), as gradient_flow's dt0. */
int cco_flow_init(cco_relax* r, const double* s0_host, const double* dt0_host);
/* FlowResult fields (flow.py:138-154): unwrapped s, energy, t, length, next
 * dt, accepted steps, evaluations, status. */
int cco_flow_result(cco_relax* r, double* s_host, double* energy_host, double* t_host,
                    double* length_host, double* dt_host, int64_t* n_steps_host,
                    int64_t* n_evals_host, int64_t* status_host);
/* Recorded rows (count per lane), s (unwrapped), t, length, energy of
 * B x max_records rows (the final state is appended by the caller). */
int cco_flow_trajectory(cco_relax* r, int64_t* count_host, double* s_host, double* t_host,
                        double* length_host, double* energy_host);

/* A molecular-dynamics session of B lanes of N particles in d dimensions
 * (mass 1, Cartesian positions r, forces at s = r inv(B), F = -g inv(B)^T:
 * make_force_fn, src/xla_cco/dynamics/run.py:229-256).  dt is the step,
 * drift_tol bounds |Phi + KE - e0| by drift_tol d (N - 1) T / 2 for the
 * verlet scheme (status 7 and the lane stops; inf disables the guard).
 * cco_md_configure sets the scheme and the tracked waves, cco_md_put
 * uploads the lanes, cco_md_run advances them, cco_md_get reads them back;
 * release with cco_relax_destroy. */
int cco_md_create(cco_ctx* ctx, cco_energy* e, int B, int N, int d, const double* basis_host,
                  double dt, double drift_tol, cco_relax** out);
/* n_steps steps of every running lane (replaces the chunk loop of
 * simulate, run.py:599-654); the host waits for the device every
 * sync_every steps; n_active = lanes still running. */
int cco_md_run(cco_relax* r, int64_t n_steps, int sync_every, int* n_active);

/* Schemes, thermostats and tracked density waves of an MD session
 * (src/xla_cco/dynamics/state.py:57-88, run.py:157-537, integrators.py:
 * 99-176, observables.py).  scheme 0 verlet, 1 andersen (a velocity reset
 * after each Verlet step: andersen_mode 0 count, 1 bernoulli), 2 baoab
 * (Langevin, friction gamma); the drift guard applies to verlet only.  The
 * n_waves tracked wavevectors (kvec, n_waves x d, Cartesian) start with the
 * n_constrained rows that are the energy's first k rows (their rho comes from
 * the force pass); the rest are summed every wind_every steps.  Each step
 * updates rho_prev (COM-corrected unless verlet), the winding sums (only
 * where S = |rho|^2 / N exceeds phase_floor at both ends), the S sums, a
 * snapshot at every step in record_steps (strictly increasing, at most 256)
 * and the rho trace every record_rho_every steps (max_rho_records rows).
 * Random draws: Philox-4x32-10 with keys split per step (the host module
 * xla_cco/dynamics/rng.py defines the same draws). */
typedef struct {
  double andersen_p, gamma, phase_floor;
  int32_t scheme, andersen_mode, wind_every, record_rho_every, max_rho_records, n_record_steps;
} cco_md_settings;
int cco_md_configure(cco_relax* r, const cco_md_settings* ms, int n_waves, int n_constrained,
                     const double* kvec_host, const int64_t* record_steps_host);
/* The complete lane state (MDState, state.py:129-165) in host arrays with a
 * leading lane axis.  lane_f64 B x 10: phi, temperature, time, reset_acc, e0,
 * max_drift, sum_kinetic, sum_phi, sum_phi2, kinetic; lane_i64 B x 5: step,
 * n_sum, status, snapshot count, trace count; key B x 2.  Per wave (w):
 * rho_prev B x w x 2, wind_abs, wind_net, n_valid, sum_s B x w; snapshots
 * (nr rows): step, time, phi, kinetic B x nr, wind_abs, wind_net B x nr x w,
 * rho B x nr x w x 2, r B x nr x N x d; trace (q rows): rho B x q x w x 2,
 * time B x q.  cco_md_put replaces the device state (lanes whose status is
 * not running stay frozen); cco_md_get reads it. */
typedef struct {
  double *r, *v, *f, *lane_f64;
  int64_t* lane_i64;
  uint32_t* key;
  double *rho_prev, *wind_abs, *wind_net;
  int64_t* n_valid;
  double* sum_s;
  int64_t* snap_step;
  double *snap_time, *snap_phi, *snap_kinetic, *snap_wind_abs, *snap_wind_net, *snap_rho, *snap_r;
  double *trace_rho, *trace_time;
} cco_md_state_io;
int cco_md_put(cco_relax* r, const cco_md_state_io* s);
int cco_md_get(cco_relax* r, cco_md_state_io* s);

/* Isotropic structure-factor moments: for the D directions n_c (D x d
 * int64) with L_c multiples each, k = m n_c (m = 1..L_c, directions in
 * order, multiples contiguous), tot[k] = sum over the B configurations of
 * S(k) = |rho(k)|^2 / N and tot2[k] = sum of S(k)^2.  Replaces the device
 * part of isotropic_structure_factor (src/xla_cco/analysis/structure.py:
 * 55-173: _sk_sums / _sk_moments); the binning stays on the host. */
int cco_sk_moments(cco_ctx* ctx, const double* s_host, int B, int N, int d,
                   const int64_t* dir_n_host, const int32_t* dir_L_host, int D,
                   double* tot_host, double* tot2_host);
/* Minimum-image pair statistics over the 3^d neighbouring images: the
 * pair-distance histogram summed over the B configurations (each unordered
 * pair once; np.histogram bins, 2..1024 edges) and, if nn_host is not NULL,
 * every particle's nearest-neighbour distance (B x N).  Replaces
 * pair_distance_histogram and nearest_neighbor_distances
 * (src/xla_cco/analysis/pairs.py:42-104). */
int cco_pair_stats(cco_ctx* ctx, const double* s_host, int B, int N, int d,
                   const double* basis_host, const double* edges_host, int n_edges,
                   int64_t* counts_host, double* nn_host);

/* Measurement (no reference counterpart; the reference times with
 * block_until_ready, benchmarks/bench_energy.py:104-115).  Per-stage device
 * time from HIP events recorded on a shard's stream around the launches of
 * every `every`-th evaluation step of that shard (0 turns it off; events
 * between launches cost device time, so a period above 1 keeps them from
 * slowing the run they measure): stages rho, reduce, force, advance (or MD
 * finish; the line-search kernel of the Gram form), compact, MD prepare,
 * Gram dot products, Gram recursion, Gram direction, MD tracked-wave rho +
 * reduce (ms summed, timed launches counted; CCO_PROFILE_STAGES entries
 * each). */
int cco_relax_set_profiling(cco_relax* r, int every);
#define CCO_PROFILE_STAGES 10
int cco_relax_profile(cco_relax* r, double* ms_sum, int64_t* launches);
/* reps timed evaluations of B x N on device data: avg ms of the rho, reduce
 * and force kernels and of the whole evaluation (avg_ms[0..3]). */
int cco_energy_profile(cco_energy* e, const double* s_dev, int B, int N, int reps,
                       double* avg_ms4);
/* Plain device copy kernel (calibration of the HBM traffic counters). */
int cco_copy_dev(cco_ctx* ctx, void* dst_dev, const void* src_dev, size_t bytes, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* CCO_H */
