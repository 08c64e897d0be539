// Device-side batched minimizers: shared structures.
#pragma once

#include "cco_internal.h"

namespace cco {

constexpr int kAdvBlock = 256;
constexpr int kMaxRescales = 8;

enum LaneMode : int { MODE_INIT = 0, MODE_EVAL = 1 };

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

// Session arrays (device).  nd = N * d, stride = lane stride in doubles.
struct RelaxDev {
  int B, nd, stride;
  int* status;    // [B]
  int* active;    // [B]
  int* n_active;  // [1]
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
  DescentLane* dl;
  LbfgsLane* ll;
  // trajectory [B][cap][stride], [B][cap]
  double* tr_s;
  double* tr_e;
  double* tr_step;
  long long* tr_it;
};

struct RelaxOutDev {
  double* s;          // [B][stride] wrapped
  double* energy;     // [B]
  double* grad_norm;  // [B]
  long long* n_evals;
  long long* n_iters;
  long long* status;
};

void launch_descent_advance(const RelaxDev& r, const DescentCfg& c, int slots_ub, hipStream_t st);
// variant: 0 auto, 1 global-memory vectors (any size), 2 register-resident vectors
// (nd <= 8192, one 1024-thread workgroup per lane)
void launch_lbfgs_advance(const RelaxDev& r, const LbfgsCfg& c, int slots_ub, hipStream_t st,
                          int variant = 0);
int lbfgs_register_ept(int nd);  // elements per thread of the register variant, 0 if too large
void launch_compact(const RelaxDev& r, hipStream_t st);
void launch_finalize(const RelaxDev& r, int method, const LbfgsCfg& lc, const RelaxOutDev& o,
                     hipStream_t st);

}  // namespace cco
