// Device-side batched minimizers: shared structures.
#pragma once

#include "cco_internal.h"

namespace cco {

constexpr int kAdvBlock = 256;
constexpr int kMaxRescales = 8;

// MODE_REEVAL (L-BFGS, split precision): the lane's last search failed or stagnated
// on split-precision energies; it turned FP64 and its current point is being
// evaluated again before its iteration restarts.
enum LaneMode : int { MODE_INIT = 0, MODE_EVAL = 1, MODE_REEVAL = 2 };

enum Status : int {
  ST_RUNNING = 0,
  ST_ENERGY_TOL = 1,
  ST_STEP_FLOOR = 2,
  ST_BUDGET = 3,
  ST_STAGNATED = 4,
  ST_LINESEARCH_FAIL = 5,
  ST_GRAD_TOL = 6,
  ST_DRIFT = 7,
};

struct DescentCfg {
  double step_init, step_min, shrink, grow, rel_tol, energy_tol;
  long long max_iters;  // max_evals // 2
  int record_every, record_cap;
};

struct LbfgsCfg {
  double rescales[kMaxRescales];
  int n_rescales;
  int memory;
  long long max_evals_per_pass;
  double energy_tol, grad_tol;
  int stagnation_patience, max_linesearch_steps;
  int record_every, record_cap;
  double mixed_switch;  // split-precision lanes turn FP64 once f <= this (<= 0: never)
};

// Zoom line-search scalars of one lane (optax 0.2.8 state, restated).
struct LsState {
  double f0, slope0, a;
  double prev_a, prev_v, prev_s;
  double low_a, low_v, low_s;
  double high_a, high_v, high_s;
  double cref_a, cref_v;
  double safe_a, safe_v;
  double last_dec, delta;
  int count, interval_found, pad0, pad1;
};

struct DescentLane {
  double f, step, norm;
  long long it, n_evals;
  int mode, rec_count;
};

struct LbfgsLane {
  double f, last_step;
  long long it, n_evals, pass_evals, stag;
  int pass_idx, opt_count, mode, rec_count;
  long long pairs;  // two-loop (s, y) pairs visited so far (traffic accounting)
  LsState ls;
};

// MD settings (src/xla_cco/dynamics/state.py:57-88).  scheme 0 verlet, 1
// andersen (verlet then a velocity reset), 2 baoab (Langevin); andersen_mode
// 0 count, 1 bernoulli.  Waves: w tracked wavevectors kvec [w][d]
// (Cartesian), the first nc of which are the energy's first k rows (their
// rho comes from the force pass); the rest are summed every wind_every steps.
struct MdCfg {
  double dt, half;
  double drift_tol;   // +inf disables the drift guard
  double inv[9];      // inv(B), row-major d x d (s = r inv(B))
  double inv_t[9];    // inv(B)^T (F = -g inv(B)^T)
  int d, n;
  int scheme, andersen_mode;
  double andersen_p, ou_c, ou_var;  // BAOAB: c = exp(-gamma dt), 1 - c^2
  double phase_floor;
  int w, nc, wind_every, mbox;      // mbox: row stride of the force pass's o.rbox
  int n_rec, rho_every, rho_cap, pad2;
  const double* kvec;               // [w][d]
  const long long* rec_steps;       // [n_rec], strictly increasing
  // waves nc .. w-1 on the reciprocal lattice: their rho comes from a rho +
  // reduce pass over their own k-set, run for the lanes whose flag md_prepare
  // clears (cadence steps); null: md_finish sums them particle by particle
  int* trk_skip;                    // [B]
  int trk_mbox, pad3;               // row stride of o.trk_rbox
};

struct MdLane {
  double phi, e0, max_drift, sum_kinetic, sum_phi, sum_phi2, time, temperature, kinetic;
  double reset_acc;
  long long step, n_sum, snap_count, trace_count;
  unsigned int key[2];
  int mode, pad;
};

// Per-lane observables buffers of an MD session (null w: none).
struct MdObs {
  double* rho_prev;   // [B][w][2]
  double* wind_abs;   // [B][w]
  double* wind_net;   // [B][w]
  long long* n_valid; // [B][w]
  double* sum_s;      // [B][w]
  // densities per box entry (reduce_kernel, coalesced) and each wave's box
  // entry (kbox: ~j = conjugate of entry j): the force pass's for waves < nc
  const double2* rbox;  // [B][mbox]
  const int* kbox;      // [nc]
  const double2* trk_rbox;  // [B][trk_mbox] tracked pass (waves nc .. w-1; null: none)
  const int* trk_kbox;      // [w - nc]
  double* cm;         // [B][3] centre of mass at the step (md_finish -> md_waves)
  int* rows;          // [B][2] the step's snapshot row and trace row, -1: none
  long long* snap_step;  // [B][n_rec]
  double* snap_time;     // [B][n_rec]
  double* snap_phi;      // [B][n_rec]
  double* snap_kin;      // [B][n_rec]
  double* snap_wabs;     // [B][n_rec][w]
  double* snap_wnet;     // [B][n_rec][w]
  double* snap_rho;      // [B][n_rec][w][2]
  double* snap_r;        // [B][n_rec][N d]
  double* tr_rho;        // [B][rho_cap][w][2]
  double* tr_time;       // [B][rho_cap]
};

// Dormand-Prince 5(4) gradient flow settings (src/xla_cco/minimize/flow.py:93-125).
struct FlowCfg {
  double rtol, atol, dt_init, dt_min, dt_max, t_max, energy_tol, grad_tol;
  long long max_steps;
  double metric_inv[9];  // inv(B B^T): v = -g metric_inv
  double basis[9];       // B (Cartesian error norms)
  double inv_t[9];       // inv(B)^T (|grad_r Phi|)
  int d, arclength;
  int record_every, record_cap;
  const double* dt0;     // optional per-lane first step (nullptr: dt_init)
  double mixed_switch;   // split-precision lanes turn FP64 when a step starts at Phi <= this (0: never)
};

struct FlowLane {
  double f, gnorm;           // at the current point
  double rate[2];            // d(t, length)/d(var) at the current point
  double var, t, length, dt, dt_try;
  double srate[7][2];        // stage rates of the attempted step
  double f_new, gnorm_new;   // at the candidate point (stage 6)
  long long n_steps, n_tries, n_evals;
  int stage, mode, rec_count;
  int gzero, gzero_new;      // gradient exactly zero at the current / candidate point
  int pad;
};

// Session arrays (device).  nd = N * d, stride = lane stride in doubles.
struct RelaxDev {
  int B, nd, stride;
  // uneven batches (null: every lane has nd = N d and stride `stride`): per
  // lane vector offset and length, and offset of its (s, y) history
  const long long* loff;
  const int* lnd;
  const long long* lhoff;
  int* status;    // [B]
  int* prec;      // [B] 1 = split-precision evaluations (null: all FP64)
  const int* prec_ok;  // [B] lanes that may start split (null: every lane)
  int* active;    // [B]
  int* n_active;  // [1]
  // mapped host word (null: none) that compact_kernel sets to
  // (completed steps << 32) | running lanes after every step
  unsigned long long* progress;
  // device step counter of the shard (null: the host passes each step's tag):
  // compact_kernel increments it and tags the progress word with the new
  // value, so a step launched from a replayed HIP graph carries its own number
  unsigned* step_ctr;
  double* s_eval; // [B][stride] trial point (energy input, s coordinates)
  double* g_eval; // [B][stride] energy gradient at the trial point
  double* phi_eval;  // [B]
  double* s;      // [B][stride] current point (s coordinates)
  double* g;      // [B][stride] descent: gradient at s; L-BFGS: gx
  double* x;      // L-BFGS x = R s
  double* px;     // L-BFGS previous x
  double* pg;     // L-BFGS previous gx
  double* u;      // L-BFGS direction
  double* safe_g; // L-BFGS gradient at the safe step
  double* S;      // [B][m][stride]
  double* Y;      // [B][m][stride]
  double* rho;    // [B][m]
  // FP32 (s, y) rows of split-precision lanes, same offsets as S and Y (null:
  // every lane keeps FP64 rows); register-resident two-loop only
  float* S32;
  float* Y32;
  // L-BFGS in Gram form: per-lane start flag {1, count} and dot-product tables
  int2* dir;       // [B]
  double* gram;    // [B][gram_stride]: SY (m x m), YY (m x m), Sg, Yg, cs, cy (m), misc (8)
  int gram_stride;
  DescentLane* dl;
  LbfgsLane* ll;
  // MD: r = positions (Cartesian), v = velocities, f = forces
  double* r;
  double* v;
  double* f;
  MdLane* ml;
  // flow: k = 7 stage velocities [B][7][stride], g = gradient at s
  double* k;
  FlowLane* fl;
  // trajectory [B][cap][stride], [B][cap]
  double* tr_s;
  double* tr_e;
  double* tr_step;     // descent / L-BFGS: step; flow: flow time t
  long long* tr_it;
  double* tr_len;      // flow: arclength
};

// Per-lane views of the session arrays (uniform or uneven batches).
__device__ __forceinline__ size_t rv_off(const RelaxDev& r, int lane) {
  return r.loff != nullptr ? (size_t)r.loff[lane] : (size_t)lane * r.stride;
}
__device__ __forceinline__ int rv_nd(const RelaxDev& r, int lane) {
  return r.lnd != nullptr ? r.lnd[lane] : r.nd;
}
__device__ __forceinline__ size_t rv_hist(const RelaxDev& r, int lane, int m) {
  return r.lhoff != nullptr ? (size_t)r.lhoff[lane] : (size_t)lane * m * r.stride;
}
// trajectory row `row` of a lane: rows of nd doubles, record_cap rows per lane
__device__ __forceinline__ size_t rv_trow(const RelaxDev& r, int lane, int cap, int row) {
  return (size_t)cap * rv_off(r, lane) + (size_t)row * rv_nd(r, lane);
}

struct RelaxOutDev {
  double* s;          // [B][stride] wrapped
  double* energy;     // [B]
  double* grad_norm;  // [B]
  long long* n_evals;
  long long* n_iters;
  long long* status;
};

void launch_descent_advance(const RelaxDev& r, const DescentCfg& c, int slots_ub, hipStream_t st);
// variant: 1 global-memory vectors (any size), 2 register-resident vectors
// (nd <= 8192, one 1024-thread workgroup per lane), 3 Gram form (first kernel of four)
void launch_lbfgs_advance(const RelaxDev& r, const LbfgsCfg& c, int slots_ub, hipStream_t st,
                          int variant);
// Gram form, kernels 2..4: part 1 dot products, 2 recursion, 3 direction and trial point
constexpr int kGramMaxMem = 128;
inline int gram_stride_for(int m) { return 2 * m * m + 4 * m + 8; }
void launch_lbfgs_gram(const RelaxDev& r, const LbfgsCfg& c, int slots_ub, int part,
                       hipStream_t st);
int lbfgs_register_ept(int nd);  // elements per thread of the register variant, 0 if too large
void launch_compact(const RelaxDev& r, hipStream_t st, int lo, int count, unsigned step_tag);
void launch_md_prepare(const RelaxDev& r, const MdCfg& c, int slots_ub, hipStream_t st);
void launch_md_finish(const RelaxDev& r, const MdCfg& c, const MdObs& o, int slots_ub,
                      hipStream_t st);
void launch_flow_advance(const RelaxDev& r, const FlowCfg& c, int slots_ub, hipStream_t st);
// ROS2 flow round kernels (dynamics_kernels.hip)
void launch_ros_gather(const RelaxDev& r, const FlowCfg& c, int slots, double* S, double* dts,
                       hipStream_t st);
void launch_ros_build(double* H, const double* dts, const int* n_active, const FlowCfg& c,
                      int nd, int slots, hipStream_t st);
void launch_ros_rhs(const RelaxDev& r, const FlowCfg& c, double* rhs, const double* dts, int mode,
                    int slots, hipStream_t st);
void launch_ros_point(const RelaxDev& r, const double* rhs, const double* dts, int mode, int slots,
                      hipStream_t st);
void launch_ros_finish(const RelaxDev& r, const FlowCfg& c, const double* rhs, const double* dts,
                       int slots, hipStream_t st);
void launch_finalize(const RelaxDev& r, int method, const LbfgsCfg& lc, const RelaxOutDev& o,
                     hipStream_t st);

}  // namespace cco
