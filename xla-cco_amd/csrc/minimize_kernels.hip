// Batched relax on the device: adaptive descent and three-pass L-BFGS.
//
// Each lane is a state machine advanced once per energy evaluation, so all
// active lanes share every energy launch whatever stage of a line search
// they are in.  One workgroup per active lane runs the vector work (norms,
// memory update, two-loop recursion) with deterministic block reductions;
// thread 0 runs the scalar line-search logic.  A lane whose status leaves
// RUNNING is dropped from the active list by compact_kernel and is never
// written again (bit-exact freeze, src/xla_cco/minimize/_loop.py:142-152).
//
// Restated algorithms:
//   descent  : src/xla_cco/minimize/descent.py:92-197
//   L-BFGS   : src/xla_cco/minimize/lbfgs.py:163-329 over optax 0.2.8
//              lbfgs(memory, scale_by_zoom_linesearch(max_steps, "one"))
// Floating-point contraction is off in this file so the minimizer arithmetic
// rounds exactly like its NumPy statement (x + a*u, q - a*y, ...).
#pragma clang fp contract(off)

#include <type_traits>

#include "minimize_kernels.h"

namespace cco {

__device__ __forceinline__ double nanmax(double a, double b) {
  if (a != a || b != b) return __longlong_as_double(0x7ff8000000000000LL);
  return a > b ? a : b;
}
__device__ __forceinline__ double nanmin(double a, double b) {
  if (a != a || b != b) return __longlong_as_double(0x7ff8000000000000LL);
  return a < b ? a : b;
}

__device__ __forceinline__ double wrap01(double v) {
  const double w = v - floor(v);
  const double below_one = 0x1.fffffffffffffp-1;
  return w < below_one ? w : below_one;
}

// ---------------------------------------------------------------------------
// compaction: active list = lanes lo .. lo + count - 1 with status RUNNING,
// in lane order (one list per shard of a session)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void compact_kernel(const int* __restrict__ status,
                                                       int* __restrict__ active,
                                                       int* __restrict__ n_active, int lo,
                                                       int count, unsigned long long* progress,
                                                       unsigned* step_ctr, unsigned step_tag) {
  __shared__ int cnt[1024];
  const int t = threadIdx.x;
  const int per = (count + 1023) / 1024;
  const int a = lo + min(count, t * per);
  const int e = min(lo + count, a + per);
  int c = 0;
  for (int i = a; i < e; ++i) c += (status[i] == ST_RUNNING);
  cnt[t] = c;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = (t >= off) ? cnt[t - off] : 0;
    __syncthreads();
    cnt[t] += v;
    __syncthreads();
  }
  int pos = cnt[t] - c;
  for (int i = a; i < e; ++i)
    if (status[i] == ST_RUNNING) active[pos++] = i;
  if (t == 1023) {
    *n_active = cnt[1023];
    unsigned tag = step_tag;
    if (step_ctr != nullptr) {  // this kernel alone writes the counter, in stream order
      tag = *step_ctr + 1u;
      *step_ctr = tag;
    }
    // the host reads this word while later steps are still queued
    if (progress != nullptr)
      __hip_atomic_store(progress, ((unsigned long long)tag << 32) | (unsigned)cnt[1023],
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

void launch_compact(const RelaxDev& r, hipStream_t st, int lo, int count, unsigned step_tag) {
  hipLaunchKernelGGL(compact_kernel, dim3(1), dim3(1024), 0, st, r.status, r.active, r.n_active,
                     lo, count, r.progress, r.step_ctr, step_tag);
}

// ---------------------------------------------------------------------------
// adaptive descent (descent.py:92-104 status, :153-175 step)
// ---------------------------------------------------------------------------
__device__ int descent_status(double f, bool all_zero, double step, long long it,
                              const DescentCfg& c) {
  int code = it >= c.max_iters ? ST_BUDGET : ST_RUNNING;
  if (step <= c.step_min) code = ST_STEP_FLOOR;
  if (all_zero) code = ST_GRAD_TOL;
  if (f < c.energy_tol) code = ST_ENERGY_TOL;
  return code;
}

__global__ __launch_bounds__(kAdvBlock) void descent_advance_kernel(RelaxDev r, DescentCfg c) {
  __shared__ double red[16];
  const int slot = blockIdx.x;
  if (slot >= *r.n_active) return;
  const int lane = r.active[slot];
  const int nd = rv_nd(r, lane);
  const size_t off = rv_off(r, lane);
  double* s = r.s + off;
  double* g = r.g + off;
  double* se = r.s_eval + off;
  const double* ge = r.g_eval + off;
  DescentLane L = r.dl[lane];
  const double fe = r.phi_eval[lane];

  if (L.mode == MODE_INIT) {
    // init_one (descent.py:121-141): s = s0 already, f, g from the first evaluation
    for (int i = threadIdx.x; i < nd; i += blockDim.x) g[i] = ge[i];
    L.f = fe;
    L.step = c.step_init;
    L.it = 0;
    L.n_evals = 0;
    L.mode = MODE_EVAL;
  } else {
    const double f2 = fe;
    const double df = L.f - f2;
    const double df2 = L.step * L.norm;
    const bool reject = fabs(df - df2) / df2 > c.rel_tol;  // NaN accepts
    if (reject) {
      L.step = L.step * c.shrink;
    } else {
      L.step = L.step * c.grow;
      for (int i = threadIdx.x; i < nd; i += blockDim.x) {
        s[i] = se[i];
        g[i] = ge[i];
      }
      L.f = f2;
    }
    L.it += 1;
    L.n_evals += 2;
  }
  // norm and the all-zero test of the (new) gradient
  double ss = 0.0, nz = 0.0;
  for (int i = threadIdx.x; i < nd; i += blockDim.x) {
    const double v = g[i];
    ss += v * v;
    nz += (v != 0.0) ? 1.0 : 0.0;
  }
  ss = block_sum<kAdvBlock>(ss, red);
  nz = block_sum<kAdvBlock>(nz, red);
  const int code = descent_status(L.f, nz == 0.0, L.step, L.it, c);
  // trajectory row (minimize/_loop.py:221-235), iteration 0 written at init
  if (c.record_every > 0 && c.record_cap > 0) {
    const bool init_row = (L.it == 0 && L.rec_count == 0);
    const bool due = init_row || ((L.it % c.record_every) == 0 && L.rec_count < c.record_cap &&
                                  L.it > 0);
    if (due) {
      const size_t row = (size_t)lane * c.record_cap + L.rec_count;
      const size_t trow = rv_trow(r, lane, c.record_cap, L.rec_count);
      for (int i = threadIdx.x; i < nd; i += blockDim.x) r.tr_s[trow + i] = wrap01(s[i]);
      if (threadIdx.x == 0) {
        r.tr_e[row] = L.f;
        r.tr_step[row] = L.step;
        r.tr_it[row] = L.it;
      }
      L.rec_count += 1;
    }
  }
  if (code == ST_RUNNING) {
    const double norm = sqrt(ss);
    L.norm = norm;
    const double step = L.step;
    for (int i = threadIdx.x; i < nd; i += blockDim.x) {
      const double gn = norm > 0.0 ? g[i] / norm : 0.0 * g[i];
      se[i] = s[i] - gn * step;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    r.dl[lane] = L;
    r.status[lane] = code;
  }
}

void launch_descent_advance(const RelaxDev& r, const DescentCfg& c, int slots_ub,
                            hipStream_t st) {
  if (slots_ub <= 0) return;
  hipLaunchKernelGGL(descent_advance_kernel, dim3(slots_ub), dim3(kAdvBlock), 0, st, r, c);
}

// ---------------------------------------------------------------------------
// L-BFGS with zoom line search
// ---------------------------------------------------------------------------
constexpr double kSlopeRtol = 1e-4;
constexpr double kCurvRtol = 0.9;
constexpr double kApproxDecRtol = 1e-6;
constexpr double kIncrease = 2.0;
// The zoom phase stops once the bracket |high - low| is no wider than this: optax's
// zoom search threshold for 64-bit step sizes.
constexpr double kIntervalMin = 1e-10;

__device__ double ls_decrease_error(double a, double va, double sa, double v0, double s0) {
  double dec = va - (v0 + kSlopeRtol * a * s0);
  double approx = sa - (2.0 * kSlopeRtol - 1.0) * s0;
  const double delta = va - v0 - kApproxDecRtol * fabs(v0);
  approx = nanmax(approx, delta);
  dec = nanmin(approx, dec);
  return nanmax(dec, 0.0);
}

__device__ double ls_curvature_error(double sa, double s0) {
  return nanmax(fabs(sa) - kCurvRtol * fabs(s0), 0.0);
}

__device__ double cubicmin(double a, double fa, double fpa, double b, double fb, double c,
                           double fc) {
  const double C = fpa;
  const double db = b - a;
  const double dc = c - a;
  const double denom = (db * dc) * (db * dc) * (db - dc);
  const double d2a = fb - fa - C * db;
  const double d2b = fc - fa - C * dc;
  const double A = (dc * dc * d2a + (-(db * db)) * d2b) / denom;
  const double Bq = ((-(dc * dc * dc)) * d2a + db * db * db * d2b) / denom;
  const double radical = Bq * Bq - 3.0 * A * C;
  return a + (-Bq + sqrt(radical)) / (3.0 * A);
}

__device__ double quadmin(double a, double fa, double fpa, double b, double fb) {
  const double db = b - a;
  const double Bq = (fb - fa - fpa * db) / (db * db);
  return a - fpa / (2.0 * Bq);
}

// Propose the next trial step of a running line search.
__device__ void ls_propose(LsState& L) {
  if (!L.interval_found) {
    L.a = L.count == 0 ? 1.0 : kIncrease * L.prev_a;
    L.delta = 0.0;
    return;
  }
  const double delta = fabs(L.high_a - L.low_a);
  const double left = fmin(L.high_a, L.low_a);
  const double right = fmax(L.high_a, L.low_a);
  const double mc = cubicmin(L.low_a, L.low_v, L.low_s, L.high_a, L.high_v, L.cref_a, L.cref_v);
  const bool use_cubic = (mc > left + 0.2 * delta) && (mc < right - 0.2 * delta);
  const double mq = quadmin(L.low_a, L.low_v, L.low_s, L.high_a, L.high_v);
  const bool use_quad = !use_cubic && (mq > left + 0.1 * delta) && (mq < right - 0.1 * delta);
  L.a = use_cubic ? mc : (use_quad ? mq : (L.low_a + L.high_a) / 2.0);
  L.delta = delta;
}

// Consume the value and slope at the current trial step.  Returns flags:
// bit 0 = record as safe step, bit 1 = search finished, bit 2 = failed.
__device__ int ls_consume(LsState& L, double va, double sa, int max_steps) {
  const double a = L.a;
  const double dec = ls_decrease_error(a, va, sa, L.f0, L.slope0);
  const double err = nanmax(dec, ls_curvature_error(sa, L.slope0));
  bool safe, done, failed;
  if (!L.interval_found) {
    safe = dec <= 0.0;
    const bool set_high = (dec > 0.0) || (va >= L.prev_v && L.count > 0);
    const bool set_low = (sa >= 0.0) && !set_high;
    if (set_low) {
      L.high_a = L.prev_a; L.high_v = L.prev_v; L.high_s = L.prev_s;
      L.low_a = a; L.low_v = va; L.low_s = sa;
    } else {
      L.low_a = L.prev_a; L.low_v = L.prev_v; L.low_s = L.prev_s;
      L.high_a = a; L.high_v = va; L.high_s = sa;
    }
    L.interval_found = (set_high || set_low || err <= 0.0) ? 1 : 0;
    done = err <= 0.0;
    L.cref_a = L.low_a;
    L.cref_v = L.low_v;
    L.count += 1;
    failed = (L.count >= max_steps) && !done;
  } else {
    safe = (dec <= 0.0) && (va < L.safe_v);
    done = err <= 0.0;
    const bool hi_to_mid = (dec > 0.0) || (va >= L.low_v);
    const bool hi_to_low = (sa * (L.high_a - L.low_a) >= 0.0) && !hi_to_mid;
    const double ola = L.low_a, olv = L.low_v, ols = L.low_s;
    const double oha = L.high_a, ohv = L.high_v;
    if (hi_to_mid) {
      L.high_a = a; L.high_v = va; L.high_s = sa;
    } else if (hi_to_low) {
      L.high_a = ola; L.high_v = olv; L.high_s = ols;
    }
    if (!hi_to_mid) {
      L.low_a = a; L.low_v = va; L.low_s = sa;
    }
    if (hi_to_mid || hi_to_low) {
      L.cref_a = oha; L.cref_v = ohv;
    } else {
      L.cref_a = ola; L.cref_v = olv;
    }
    L.count += 1;
    failed = ((L.count >= max_steps) || (L.delta <= kIntervalMin)) && !done;
  }
  L.prev_a = a; L.prev_v = va; L.prev_s = sa;
  L.last_dec = dec;
  if (safe) {
    L.safe_a = a;
    L.safe_v = va;
  }
  const bool finished = done || failed;
  if (!finished) ls_propose(L);
  return (safe ? 1 : 0) | (finished ? 2 : 0) | (failed ? 4 : 0);
}

// Start an L-BFGS iteration: memory update, two-loop recursion, line-search
// init, first trial point x + 1 * u.  All threads call it.
template <int BLOCK>
__device__ void lbfgs_start_iteration(const RelaxDev& r, const LbfgsCfg& c, int lane,
                                      LbfgsLane& L, double* red) {
  const int nd = rv_nd(r, lane);
  const size_t off = rv_off(r, lane);
  const int m = c.memory;
  const double* x = r.x + off;
  const double* gx = r.g + off;
  double* px = r.px + off;
  double* pg = r.pg + off;
  double* u = r.u + off;
  double* S = r.S + rv_hist(r, lane, m);
  double* Y = r.Y + rv_hist(r, lane, m);
  double* rho = r.rho + (size_t)lane * m;
  const double rs = c.rescales[L.pass_idx];
  const int count = L.opt_count;
  double gamma;
  int slot_new = -1;
  double rho_new = 0.0;
  if (count > 0) {
    const int slot = (count - 1) % m;
    slot_new = slot;
    double* Ss = S + (size_t)slot * nd;
    double* Ys = Y + (size_t)slot * nd;
    double sy = 0.0, yy = 0.0;
    for (int i = threadIdx.x; i < nd; i += blockDim.x) {
      const double ds = x[i] - px[i];
      const double dy = gx[i] - pg[i];
      Ss[i] = ds;
      Ys[i] = dy;
      sy += dy * ds;
      yy += dy * dy;
    }
    sy = block_sum<BLOCK>(sy, red);
    yy = block_sum<BLOCK>(yy, red);
    rho_new = (sy == 0.0) ? 0.0 : 1.0 / sy;
    if (threadIdx.x == 0) rho[slot] = rho_new;
    gamma = (yy > 0.0) ? sy / yy : 1.0;
  } else {
    double gg = 0.0;
    for (int i = threadIdx.x; i < nd; i += blockDim.x) gg += gx[i] * gx[i];
    gg = block_sum<BLOCK>(gg, red);
    gamma = nanmin(1.0, 1.0 / sqrt(gg));
  }
  __syncthreads();
  // two-loop recursion over the written slots, newest first (q lives in u)
  const int nvalid = count < m ? count : m;
  const int midx = count % m;
  __shared__ double alphas[1024];
  // Each pair's update of q is applied in the pass that forms the next pair's
  // dot product (one pass over q per pair; every element sees the same
  // operations in the same order as with separate passes).  Elements are
  // owned by one thread throughout, so q needs no barrier between passes.
  for (int i = threadIdx.x; i < nd; i += blockDim.x) u[i] = gx[i];
  const double* pend = nullptr;  // the pair vector of the pending update
  double pc = 0.0;               // its coefficient
  for (int t = m - 1; t >= m - nvalid; --t) {
    const int sl = (midx + t) % m;
    const double* Si = S + (size_t)sl * nd;
    double dot = 0.0;
    for (int i = threadIdx.x; i < nd; i += blockDim.x) {
      double q = u[i];
      if (pend != nullptr) {
        q = q + pc * pend[i];
        u[i] = q;
      }
      dot += Si[i] * q;
    }
    dot = block_sum<BLOCK>(dot, red);
    const double a = (sl == slot_new ? rho_new : rho[sl]) * dot;
    if (threadIdx.x == 0) alphas[t] = a;
    pend = Y + (size_t)sl * nd;
    pc = -a;
  }
  for (int i = threadIdx.x; i < nd; i += blockDim.x) {
    double q = u[i];
    if (pend != nullptr) q = q + pc * pend[i];
    u[i] = gamma * q;
  }
  __syncthreads();
  pend = nullptr;
  for (int t = m - nvalid; t < m; ++t) {
    const int sl = (midx + t) % m;
    const double* Yi = Y + (size_t)sl * nd;
    double dot = 0.0;
    for (int i = threadIdx.x; i < nd; i += blockDim.x) {
      double q = u[i];
      if (pend != nullptr) {
        q = q + pc * pend[i];
        u[i] = q;
      }
      dot += Yi[i] * q;
    }
    dot = block_sum<BLOCK>(dot, red);
    const double bcoef = (sl == slot_new ? rho_new : rho[sl]) * dot;
    pend = S + (size_t)sl * nd;
    pc = alphas[t] - bcoef;
  }
  // direction u = -H g, remember (x, g) for the next memory update
  double slope = 0.0;
  for (int i = threadIdx.x; i < nd; i += blockDim.x) {
    double q = u[i];
    if (pend != nullptr) q = q + pc * pend[i];
    const double ui = -q;
    u[i] = ui;
    px[i] = x[i];
    pg[i] = gx[i];
    slope += ui * gx[i];
  }
  slope = block_sum<BLOCK>(slope, red);
  L.opt_count = count + 1;
  L.pairs += nvalid;
  LsState& Q = L.ls;
  Q.f0 = L.f;
  Q.slope0 = slope;
  Q.count = 0;
  Q.interval_found = 0;
  Q.prev_a = 0.0; Q.prev_v = L.f; Q.prev_s = slope;
  Q.low_a = 0.0; Q.low_v = L.f; Q.low_s = slope;
  Q.high_a = 0.0; Q.high_v = L.f; Q.high_s = slope;
  Q.cref_a = 0.0; Q.cref_v = L.f;
  Q.safe_a = 0.0; Q.safe_v = L.f;
  Q.last_dec = __longlong_as_double(0x7ff0000000000000LL);
  ls_propose(Q);
  const double a = Q.a;
  double* se = r.s_eval + off;
  for (int i = threadIdx.x; i < nd; i += blockDim.x) se[i] = (x[i] + a * u[i]) / rs;
}

// Gram form, first kernel: when a lane starts an iteration, store the new
// (s, y) pair (x - px, gx - pg, as lbfgs_start_iteration does) and flag the
// lane for the dot-product, recursion and direction kernels.
__device__ void lbfgs_gram_begin(const RelaxDev& r, const LbfgsCfg& c, int lane,
                                 const LbfgsLane& L) {
  const int count = L.opt_count;
  if (count <= 0) return;
  const int m = c.memory;
  const int slot = (count - 1) % m;
  const size_t off = rv_off(r, lane);
  const int nd = rv_nd(r, lane);
  const double* x = r.x + off;
  const double* gx = r.g + off;
  const double* px = r.px + off;
  const double* pg = r.pg + off;
  double* Ss = r.S + rv_hist(r, lane, m) + (size_t)slot * nd;
  double* Ys = r.Y + rv_hist(r, lane, m) + (size_t)slot * nd;
  for (int i = threadIdx.x; i < nd; i += blockDim.x) {
    Ss[i] = x[i] - px[i];
    Ys[i] = gx[i] - pg[i];
  }
}

// The first slope along a new direction arrives with the first trial
// evaluation (the Gram form's recursion kernel does not reduce over x).
__device__ __forceinline__ void ls_set_slope0(LsState& Q, double slope) {
  Q.slope0 = slope;
  Q.prev_s = slope;
  Q.low_s = slope;
  Q.high_s = slope;
  Q.pad0 = 0;
}

template <int BLOCK, bool GRAM>
__global__ __launch_bounds__(BLOCK) void lbfgs_advance_kernel(RelaxDev r, LbfgsCfg c) {
  __shared__ double red[16];
  const int slot = blockIdx.x;
  if (slot >= *r.n_active) return;
  const int lane = r.active[slot];
  const int nd = rv_nd(r, lane);
  const size_t off = rv_off(r, lane);
  double* s = r.s + off;
  double* x = r.x + off;
  double* gx = r.g + off;
  double* u = r.u + off;
  double* sg = r.safe_g + off;
  double* ge = r.g_eval + off;
  LbfgsLane L = r.ll[lane];
  const double fe = r.phi_eval[lane];
  int code = ST_RUNNING;
  bool started = false;
  const bool split = r.prec != nullptr && r.prec[lane] != 0;
  bool rescue = false;

  if (L.mode == MODE_REEVAL) {
    // the FP64 evaluation at the current point after a split-precision rescue:
    // the iteration restarts from it with an empty history
    const double rs = c.rescales[L.pass_idx];
    for (int i = threadIdx.x; i < nd; i += blockDim.x) gx[i] = ge[i] / rs;
    L.f = fe;
    L.n_evals += 1;
    L.pass_evals += 1;
    L.mode = MODE_EVAL;
    code = (fe < c.energy_tol) ? ST_ENERGY_TOL : ST_RUNNING;
    __syncthreads();
    if (code == ST_RUNNING) {
      if constexpr (GRAM) {
        lbfgs_gram_begin(r, c, lane, L);
        started = true;
      } else {
        lbfgs_start_iteration<BLOCK>(r, c, lane, L, red);
      }
    }
  } else if (L.mode == MODE_INIT) {
    // init_one (lbfgs.py:163-197)
    const double r0 = c.rescales[0];
    for (int i = threadIdx.x; i < nd; i += blockDim.x) {
      x[i] = s[i] * r0;
      gx[i] = ge[i] / r0;
    }
    L.f = fe;
    L.last_step = 0.0;
    L.it = 0;
    L.n_evals = 1;
    L.pass_idx = 0;
    L.pass_evals = 0;
    L.stag = 0;
    L.opt_count = 0;
    L.mode = MODE_EVAL;
    code = (fe < c.energy_tol) ? ST_ENERGY_TOL : ST_RUNNING;
    if (c.record_every > 0 && c.record_cap > 0 && L.rec_count == 0) {
      for (int i = threadIdx.x; i < nd; i += blockDim.x) r.tr_s[rv_trow(r, lane, c.record_cap, 0) + i] = wrap01(s[i]);
      if (threadIdx.x == 0) {
        r.tr_e[(size_t)lane * c.record_cap] = L.f;
        r.tr_step[(size_t)lane * c.record_cap] = 0.0;
        r.tr_it[(size_t)lane * c.record_cap] = 0;
      }
      L.rec_count = 1;
    }
    __syncthreads();
    if (code == ST_RUNNING) {
      if constexpr (GRAM) {
        lbfgs_gram_begin(r, c, lane, L);
        started = true;
      } else {
        lbfgs_start_iteration<BLOCK>(r, c, lane, L, red);
      }
    }
  } else {
    const double rs = c.rescales[L.pass_idx];
    // gradient w.r.t. x at the trial point, and its slope along u
    double sa = 0.0;
    double s0 = 0.0;
    const bool need_slope0 = GRAM && L.ls.pad0;
    for (int i = threadIdx.x; i < nd; i += blockDim.x) {
      const double gi = ge[i] / rs;
      ge[i] = gi;
      sa += gi * u[i];
      if (need_slope0) s0 += u[i] * gx[i];
    }
    sa = block_sum<BLOCK>(sa, red);
    if (need_slope0) ls_set_slope0(L.ls, block_sum<BLOCK>(s0, red));
    // every thread runs the same scalar update, so L stays uniform
    const int flags = ls_consume(L.ls, fe, sa, c.max_linesearch_steps);
    if (flags & 1) {
      for (int i = threadIdx.x; i < nd; i += blockDim.x) sg[i] = ge[i];
    }
    if (flags & 2) {
      const bool use_safe = (flags & 4) && (L.ls.safe_a > 0.0);
      const double alpha = use_safe ? L.ls.safe_a : L.ls.prev_a;
      const double value = use_safe ? L.ls.safe_v : fe;
      const double dec = use_safe ? 0.0 : L.ls.last_dec;
      const double* gnew = use_safe ? sg : ge;
      const bool fail = (dec > 0.0) || !isfinite(value);
      const long long evals = L.ls.count > 1 ? L.ls.count : 1;
      const double f_old = L.f;
      __syncthreads();
      if (!fail) {
        for (int i = threadIdx.x; i < nd; i += blockDim.x) {
          const double xi = x[i] + alpha * u[i];
          x[i] = xi;
          s[i] = xi / rs;
          gx[i] = gnew[i];
        }
        L.f = value;
      }
      L.stag = (L.f < f_old) ? 0 : L.stag + 1;
      L.pass_evals += evals;
      double gg = 0.0;
      for (int i = threadIdx.x; i < nd; i += blockDim.x) gg += gx[i] * gx[i];
      gg = block_sum<BLOCK>(gg, red);
      const double gnorm = sqrt(gg) * rs;
      int glob = (gnorm < c.grad_tol) ? ST_GRAD_TOL : ST_RUNNING;
      if (L.f < c.energy_tol) glob = ST_ENERGY_TOL;
      int local = (L.pass_evals >= c.max_evals_per_pass) ? ST_BUDGET : ST_RUNNING;
      if (L.stag >= c.stagnation_patience) local = ST_STAGNATED;
      if (fail) local = ST_LINESEARCH_FAIL;
      // split precision: a search that fails, or a lane that stagnates, on
      // split-precision energies turns the lane FP64 and re-evaluates its
      // current point instead of ending the pass
      rescue = split && glob == ST_RUNNING && (local == ST_LINESEARCH_FAIL || local == ST_STAGNATED);
      if (rescue) local = ST_RUNNING;
      const bool last = L.pass_idx >= c.n_rescales - 1;
      const bool pass_end = (local != ST_RUNNING) && (glob == ST_RUNNING);
      const bool advance = pass_end && !last;
      code = (glob != ST_RUNNING) ? glob : ST_RUNNING;
      if (pass_end && last) code = local;
      if (advance) {
        const double rn = c.rescales[L.pass_idx + 1];
        for (int i = threadIdx.x; i < nd; i += blockDim.x) {
          x[i] = s[i] * rn;
          gx[i] = gx[i] * rs / rn;
        }
        L.opt_count = 0;
        L.pass_idx += 1;
        L.pass_evals = 0;
        L.stag = 0;
      }
      L.last_step = alpha;
      L.it += 1;
      L.n_evals += evals;
      if (c.record_every > 0 && c.record_cap > 0 && (L.it % c.record_every) == 0 &&
          L.rec_count < c.record_cap) {
        const size_t row = (size_t)lane * c.record_cap + L.rec_count;
        const size_t trow = rv_trow(r, lane, c.record_cap, L.rec_count);
        for (int i = threadIdx.x; i < nd; i += blockDim.x) r.tr_s[trow + i] = wrap01(s[i]);
        if (threadIdx.x == 0) {
          r.tr_e[row] = L.f;
          r.tr_step[row] = L.last_step;
          r.tr_it[row] = L.it;
        }
        L.rec_count += 1;
      }
      __syncthreads();
      if (rescue && code == ST_RUNNING) {
        double* se = r.s_eval + off;
        for (int i = threadIdx.x; i < nd; i += blockDim.x) se[i] = x[i] / rs;
        L.mode = MODE_REEVAL;
        L.stag = 0;
        L.opt_count = 0;
      } else if (code == ST_RUNNING) {
        if constexpr (GRAM) {
          lbfgs_gram_begin(r, c, lane, L);
          started = true;
        } else {
          lbfgs_start_iteration<BLOCK>(r, c, lane, L, red);
        }
      }
    } else {
      const double a = L.ls.a;
      double* se = r.s_eval + off;
      for (int i = threadIdx.x; i < nd; i += blockDim.x) se[i] = (x[i] + a * u[i]) / rs;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    r.ll[lane] = L;
    r.status[lane] = code;
    if constexpr (GRAM) r.dir[lane] = make_int2(started ? 1 : 0, L.opt_count);
    if (code == ST_RUNNING && r.prec != nullptr &&
        (rescue || (c.mixed_switch > 0.0 && L.f <= c.mixed_switch)))
      r.prec[lane] = 0;
  }
}

// ---------------------------------------------------------------------------
// L-BFGS advance with register-resident vectors: one workgroup of NT threads
// per lane, element reg_elem(tid, e) of every vector owned by thread tid.
// The two-loop recursion keeps q in registers, keeps the next pairs (s, y) in
// flight while one is reduced (enough bytes in flight per CU to stream at HBM
// rate with one lane per CU), and pays one barrier per pair (double-buffered
// reduction slots).  The per-pair chain (dot, wave sum, LDS, barrier, sum of
// the wave partials) sets a lane's latency: 512 threads with 16-byte history
// loads (V = 2, even N d up to 2048 in a uniform batch) run it in about 60% of
// the time of 1024 threads, which serve larger vectors and uneven batches.
// Same arithmetic per element as lbfgs_advance_kernel; the dot products are
// summed in another order.
// ---------------------------------------------------------------------------
constexpr int kRegBlock = 1024;  // the largest register-resident workgroup
constexpr size_t kLdsPerCu = 160 * 1024;  // gfx950

// Element e of thread tid in a register-resident vector of NT threads.  V = 2:
// each thread owns adjacent element pairs (one 16-byte load per pair).
template <int NT, int V>
__device__ __forceinline__ int reg_elem(int tid, int e) {
  return V == 2 ? (e >> 1) * (2 * NT) + 2 * tid + (e & 1) : e * NT + tid;
}

struct Red2 {
  double* buf;  // [2][3 * W]
  int parity;
};

// Sum of the W wave partials p[stride * w] (w = lane % W of every row) with
// row shifts: lane W - 1 ends with the pairwise tree ((p0 + p1) + (p2 + p3)) + ...
// up to the order of each addition's operands, which does not change a
// floating-point sum.  Uniform result.
template <int W>
__device__ __forceinline__ double partials_sum(const double* p, int stride) {
  static_assert(W == 8 || W == 16, "the wave partials fill one 16-lane DPP row");
  double x = p[stride * (threadIdx.x % W)];
  x += dpp_d<0x111, 0xf, true>(x);  // row_shr:1
  x += dpp_d<0x112, 0xf, true>(x);  // row_shr:2
  x += dpp_d<0x114, 0xf, true>(x);  // row_shr:4
  if constexpr (W == 16) x += dpp_d<0x118, 0xf, true>(x);  // row_shr:8
  return readlane_d(x, W - 1);
}

template <int W>
__device__ __forceinline__ double reg_sum(double v, Red2& R) {
  v = wave_sum(v);
  double* b = R.buf + R.parity * 3 * W;
  R.parity ^= 1;
  if ((threadIdx.x % kWave) == 0) b[threadIdx.x / kWave] = v;
  __syncthreads();
  return partials_sum<W>(b, 1);
}

template <int W>
__device__ __forceinline__ void reg_sum2(double& v0, double& v1, Red2& R) {
  v0 = wave_sum(v0);
  v1 = wave_sum(v1);
  double* b = R.buf + R.parity * 3 * W;
  R.parity ^= 1;
  if ((threadIdx.x % kWave) == 0) {
    b[2 * (threadIdx.x / kWave)] = v0;
    b[2 * (threadIdx.x / kWave) + 1] = v1;
  }
  __syncthreads();
  v0 = partials_sum<W>(b, 2);
  v1 = partials_sum<W>(b + 1, 2);
}

// H32: a split-precision lane's (s, y) rows are FP32 (r.S32 / r.Y32, same
// element offsets as the FP64 rows), which halves the bytes each pair streams;
// the arithmetic stays FP64.
template <int NT, int EPT, int V, bool H32>
__device__ void lbfgs_start_reg(const RelaxDev& r, const LbfgsCfg& c, int lane, LbfgsLane& L,
                                Red2& R, double* alphas, double* rhos) {
  using HT = std::conditional_t<H32, float, double>;
  const int nd = rv_nd(r, lane);
  const int tid = threadIdx.x;
  const size_t off = rv_off(r, lane);
  const int m = c.memory;
  const double* x = r.x + off;
  const double* gx = r.g + off;
  double* px = r.px + off;
  double* pg = r.pg + off;
  double* u = r.u + off;
  double* S = r.S + rv_hist(r, lane, m);
  double* Y = r.Y + rv_hist(r, lane, m);
  float* S32 = r.S32 + rv_hist(r, lane, m);  // dereferenced only when H32
  float* Y32 = r.Y32 + rv_hist(r, lane, m);
  const HT* Sh;
  const HT* Yh;
  if constexpr (H32) {
    Sh = S32;
    Yh = Y32;
  } else {
    Sh = S;
    Yh = Y;
  }
  double* rho = r.rho + (size_t)lane * m;
  const double rs = c.rescales[L.pass_idx];
  const int count = L.opt_count;
  double q[EPT];
  double gamma;
  int slot_new = -1;
  double rho_new = 0.0;
  if (count > 0) {
    const int slot = (count - 1) % m;
    slot_new = slot;
    double* Ss = S + (size_t)slot * nd;
    double* Ys = Y + (size_t)slot * nd;
    double sy = 0.0, yy = 0.0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int i = reg_elem<NT, V>(tid, e);
      q[e] = 0.0;
      if (i < nd) {
        const double gi = gx[i];
        double ds = x[i] - px[i];
        double dy = gi - pg[i];
        if constexpr (H32) {
          // rho and gamma from the rounded pair the two-loop will read
          const float fs = (float)ds, fy = (float)dy;
          S32[(size_t)slot * nd + i] = fs;
          Y32[(size_t)slot * nd + i] = fy;
          ds = fs;
          dy = fy;
        } else {
          Ss[i] = ds;
          Ys[i] = dy;
        }
        sy += dy * ds;
        yy += dy * dy;
        q[e] = gi;
      }
    }
    reg_sum2<NT / kWave>(sy, yy, R);
    rho_new = (sy == 0.0) ? 0.0 : 1.0 / sy;
    if (tid == 0) rho[slot] = rho_new;
    gamma = (yy > 0.0) ? sy / yy : 1.0;
  } else {
    double gg = 0.0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int i = reg_elem<NT, V>(tid, e);
      q[e] = 0.0;
      if (i < nd) {
        q[e] = gx[i];
        gg += q[e] * q[e];
      }
    }
    gg = reg_sum<NT / kWave>(gg, R);
    gamma = nanmin(1.0, 1.0 / sqrt(gg));
  }
  const int nvalid = count < m ? count : m;
  const int midx = count % m;
  // rho of every valid pair by age, staged once in LDS (off the critical path)
  for (int t = m - nvalid + tid; t < m; t += NT) {
    const int sl = (midx + t) % m;
    rhos[t] = (sl == slot_new) ? rho_new : rho[sl];
  }
  __syncthreads();
  if constexpr (EPT <= 4) {
    // Pairs stream D ahead of the one being reduced.  Every load is
    // unconditional (element indices clamped into the lane's vector, values
    // past its end masked at use; pair indices clamped into the valid range)
    // and both loops run to a multiple of D, the extra pairs with zero
    // weight (which leaves q unchanged), so the loop body is straight-line
    // and each use waits for its own pair only, not for the pairs in flight.
    constexpr int D = NT == 512 ? (H32 ? 8 : 4) : (EPT == 1 ? 6 : (EPT == 2 ? 4 : 2));
    int ic[EPT];
    bool mk[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int i = reg_elem<NT, V>(tid, e);
      mk[e] = i < nd;
      // V = 2 keeps pairs whole: an even element and its successor (nd is even)
      ic[e] = mk[e] ? i : (V == 2 ? nd - 2 + (e & 1) : nd - 1);
    }
    HT bs[D][EPT], by[D][EPT];
    auto load_pair = [&](HT (&ds)[EPT], HT (&dy)[EPT], int t) {
      const int sl = (midx + t) % m;
      const HT* Si = Sh + (size_t)sl * nd;
      const HT* Yi = Yh + (size_t)sl * nd;
      if constexpr (V == 2) {
        using HT2 = std::conditional_t<H32, float2, double2>;
#pragma unroll
        for (int e = 0; e < EPT; e += 2) {
          const HT2 a = *reinterpret_cast<const HT2*>(Si + ic[e]);
          const HT2 b = *reinterpret_cast<const HT2*>(Yi + ic[e]);
          ds[e] = a.x;
          ds[e + 1] = a.y;
          dy[e] = b.x;
          dy[e + 1] = b.y;
        }
      } else {
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
          ds[e] = Si[ic[e]];
          dy[e] = Yi[ic[e]];
        }
      }
    };
    const int npad = ((nvalid + D - 1) / D) * D;
    const int last = nvalid - 1;
    // first loop: newest -> oldest (pair k has age m - 1 - k)
    if (nvalid > 0) {
#pragma unroll
      for (int p = 0; p < D; ++p) load_pair(bs[p], by[p], m - 1 - min(p, last));
    }
    for (int k0 = 0; k0 < npad; k0 += D) {
#pragma unroll
      for (int p = 0; p < D; ++p) {
        const int k = k0 + p;
        const int t = m - 1 - min(k, last);
        const double w = k < nvalid ? rhos[t] : 0.0;
        double sc[EPT], yc[EPT];
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
          sc[e] = bs[p][e];
          yc[e] = by[p][e];
        }
        load_pair(bs[p], by[p], m - 1 - min(k + D, last));
        double dot = 0.0;
#pragma unroll
        for (int e = 0; e < EPT; ++e) dot += sc[e] * q[e];  // q is zero past the vector's end
        dot = reg_sum<NT / kWave>(dot, R);
        const double a = w * dot;
        if (tid == 0 && k < nvalid) alphas[t] = a;
#pragma unroll
        for (int e = 0; e < EPT; ++e) q[e] = mk[e] ? q[e] + (-a) * yc[e] : 0.0;
      }
    }
#pragma unroll
    for (int e = 0; e < EPT; ++e) q[e] = gamma * q[e];
    // second loop: oldest -> newest (pair k has age m - nvalid + k)
    if (nvalid > 0) {
#pragma unroll
      for (int p = 0; p < D; ++p) load_pair(bs[p], by[p], m - nvalid + min(p, last));
    }
    for (int k0 = 0; k0 < npad; k0 += D) {
#pragma unroll
      for (int p = 0; p < D; ++p) {
        const int k = k0 + p;
        const int t = m - nvalid + min(k, last);
        double sc[EPT], yc[EPT];
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
          sc[e] = bs[p][e];
          yc[e] = by[p][e];
        }
        load_pair(bs[p], by[p], m - nvalid + min(k + D, last));
        double dot = 0.0;
#pragma unroll
        for (int e = 0; e < EPT; ++e) dot += yc[e] * q[e];
        dot = reg_sum<NT / kWave>(dot, R);
        const double coef = k < nvalid ? alphas[t] - rhos[t] * dot : 0.0;
#pragma unroll
        for (int e = 0; e < EPT; ++e) q[e] = mk[e] ? q[e] + coef * sc[e] : 0.0;
      }
    }
  } else {
    // One pair per reduction, one pair in flight (vectors of 8 elements per
    // thread leave no registers for more).
    HT bs[EPT], by[EPT];
    auto load_into = [&](HT (&ds)[EPT], HT (&dy)[EPT], int t) {
      const int sl = (midx + t) % m;
      const HT* Si = Sh + (size_t)sl * nd;
      const HT* Yi = Yh + (size_t)sl * nd;
  #pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int i = reg_elem<NT, V>(tid, e);
        ds[e] = i < nd ? Si[i] : (HT)0;
        dy[e] = i < nd ? Yi[i] : (HT)0;
      }
    };
    // first loop: newest -> oldest (k = 0 is age m - 1)
    if (nvalid > 0) load_into(bs, by, m - 1);
    for (int k = 0; k < nvalid; ++k) {
      const int t = m - 1 - k;
      double sc[EPT], yc[EPT];
  #pragma unroll
      for (int e = 0; e < EPT; ++e) {
        sc[e] = bs[e];
        yc[e] = by[e];
      }
      if (k + 1 < nvalid) load_into(bs, by, t - 1);
      double dot = 0.0;
  #pragma unroll
      for (int e = 0; e < EPT; ++e) dot += sc[e] * q[e];
      dot = reg_sum<NT / kWave>(dot, R);
      const double a = rhos[t] * dot;
      if (tid == 0) alphas[t] = a;
  #pragma unroll
      for (int e = 0; e < EPT; ++e) q[e] = q[e] + (-a) * yc[e];
    }
  #pragma unroll
    for (int e = 0; e < EPT; ++e) q[e] = gamma * q[e];
    // second loop: oldest -> newest (k = 0 is age m - nvalid)
    if (nvalid > 0) load_into(bs, by, m - nvalid);
    for (int k = 0; k < nvalid; ++k) {
      const int t = m - nvalid + k;
      double sc[EPT], yc[EPT];
  #pragma unroll
      for (int e = 0; e < EPT; ++e) {
        sc[e] = bs[e];
        yc[e] = by[e];
      }
      if (k + 1 < nvalid) load_into(bs, by, t + 1);
      double dot = 0.0;
  #pragma unroll
      for (int e = 0; e < EPT; ++e) dot += yc[e] * q[e];
      dot = reg_sum<NT / kWave>(dot, R);
      const double coef = alphas[t] - rhos[t] * dot;
  #pragma unroll
      for (int e = 0; e < EPT; ++e) q[e] = q[e] + coef * sc[e];
    }
  }
  double slope = 0.0;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int i = reg_elem<NT, V>(tid, e);
    if (i < nd) {
      const double ui = -q[e];
      q[e] = ui;
      const double gi = gx[i];
      u[i] = ui;
      px[i] = x[i];
      pg[i] = gi;
      slope += ui * gi;
    }
  }
  slope = reg_sum<NT / kWave>(slope, R);
  L.opt_count = count + 1;
  L.pairs += nvalid;
  LsState& Q = L.ls;
  Q.f0 = L.f;
  Q.slope0 = slope;
  Q.count = 0;
  Q.interval_found = 0;
  Q.prev_a = 0.0; Q.prev_v = L.f; Q.prev_s = slope;
  Q.low_a = 0.0; Q.low_v = L.f; Q.low_s = slope;
  Q.high_a = 0.0; Q.high_v = L.f; Q.high_s = slope;
  Q.cref_a = 0.0; Q.cref_v = L.f;
  Q.safe_a = 0.0; Q.safe_v = L.f;
  Q.last_dec = __longlong_as_double(0x7ff0000000000000LL);
  ls_propose(Q);
  const double a = Q.a;
  double* se = r.s_eval + off;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int i = reg_elem<NT, V>(tid, e);
    if (i < nd) se[i] = (x[i] + a * q[e]) / rs;
  }
}

template <int NT, int EPT, int V>
__global__ __launch_bounds__(NT) void lbfgs_advance_reg_kernel(RelaxDev r, LbfgsCfg c) {
  constexpr int W = NT / kWave;
  __shared__ double redbuf[6 * W];
  __shared__ double alphas[1024];
  __shared__ double rhos[1024];
  Red2 R{redbuf, 0};
  const int slot = blockIdx.x;
  if (slot >= *r.n_active) return;
  const int lane = r.active[slot];
  const int nd = rv_nd(r, lane);
  const int tid = threadIdx.x;
  const size_t off = rv_off(r, lane);
  double* s = r.s + off;
  double* x = r.x + off;
  double* gx = r.g + off;
  const double* u = r.u + off;
  double* sg = r.safe_g + off;
  const double* ge = r.g_eval + off;
  LbfgsLane L = r.ll[lane];
  const double fe = r.phi_eval[lane];
  int code = ST_RUNNING;
  bool start = false;
  const bool split = r.prec != nullptr && r.prec[lane] != 0;
  const bool h32 = split && r.S32 != nullptr;
  bool rescue = false;

  if (L.mode == MODE_REEVAL) {
    // the FP64 evaluation at the current point after a split-precision rescue:
    // the iteration restarts from it with an empty history
    const double rs = c.rescales[L.pass_idx];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int i = reg_elem<NT, V>(tid, e);
      if (i < nd) gx[i] = ge[i] / rs;
    }
    L.f = fe;
    L.n_evals += 1;
    L.pass_evals += 1;
    L.mode = MODE_EVAL;
    code = (fe < c.energy_tol) ? ST_ENERGY_TOL : ST_RUNNING;
    start = code == ST_RUNNING;
  } else if (L.mode == MODE_INIT) {
    const double r0 = c.rescales[0];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int i = reg_elem<NT, V>(tid, e);
      if (i < nd) {
        x[i] = s[i] * r0;
        gx[i] = ge[i] / r0;
      }
    }
    L.f = fe;
    L.last_step = 0.0;
    L.it = 0;
    L.n_evals = 1;
    L.pass_idx = 0;
    L.pass_evals = 0;
    L.stag = 0;
    L.opt_count = 0;
    L.mode = MODE_EVAL;
    code = (fe < c.energy_tol) ? ST_ENERGY_TOL : ST_RUNNING;
    if (c.record_every > 0 && c.record_cap > 0 && L.rec_count == 0) {
      const size_t row = (size_t)lane * c.record_cap;
      const size_t trow = rv_trow(r, lane, c.record_cap, 0);
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int i = reg_elem<NT, V>(tid, e);
        if (i < nd) r.tr_s[trow + i] = wrap01(s[i]);
      }
      if (tid == 0) {
        r.tr_e[row] = L.f;
        r.tr_step[row] = 0.0;
        r.tr_it[row] = 0;
      }
      L.rec_count = 1;
    }
    start = code == ST_RUNNING;
  } else {
    const double rs = c.rescales[L.pass_idx];
    double gt[EPT], uu[EPT];
    double sa = 0.0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int i = reg_elem<NT, V>(tid, e);
      gt[e] = 0.0;
      uu[e] = 0.0;
      if (i < nd) {
        gt[e] = ge[i] / rs;
        uu[e] = u[i];
        sa += gt[e] * uu[e];
      }
    }
    sa = reg_sum<NT / kWave>(sa, R);
    const int flags = ls_consume(L.ls, fe, sa, c.max_linesearch_steps);
    if (flags & 1) {
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int i = reg_elem<NT, V>(tid, e);
        if (i < nd) sg[i] = gt[e];
      }
    }
    if (flags & 2) {
      const bool use_safe = (flags & 4) && (L.ls.safe_a > 0.0);
      const double alpha = use_safe ? L.ls.safe_a : L.ls.prev_a;
      const double value = use_safe ? L.ls.safe_v : fe;
      const double dec = use_safe ? 0.0 : L.ls.last_dec;
      const bool fail = (dec > 0.0) || !isfinite(value);
      const long long evals = L.ls.count > 1 ? L.ls.count : 1;
      const double f_old = L.f;
      double gg = 0.0;
      double gv[EPT];
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int i = reg_elem<NT, V>(tid, e);
        gv[e] = 0.0;
        if (i < nd) {
          if (!fail) {
            const double xi = x[i] + alpha * uu[e];
            x[i] = xi;
            s[i] = xi / rs;
            gv[e] = use_safe ? sg[i] : gt[e];
            gx[i] = gv[e];
          } else {
            gv[e] = gx[i];
          }
          gg += gv[e] * gv[e];
        }
      }
      if (!fail) L.f = value;
      L.stag = (L.f < f_old) ? 0 : L.stag + 1;
      L.pass_evals += evals;
      gg = reg_sum<NT / kWave>(gg, R);
      const double gnorm = sqrt(gg) * rs;
      int glob = (gnorm < c.grad_tol) ? ST_GRAD_TOL : ST_RUNNING;
      if (L.f < c.energy_tol) glob = ST_ENERGY_TOL;
      int local = (L.pass_evals >= c.max_evals_per_pass) ? ST_BUDGET : ST_RUNNING;
      if (L.stag >= c.stagnation_patience) local = ST_STAGNATED;
      if (fail) local = ST_LINESEARCH_FAIL;
      // split precision: a failed search or a stagnating lane turns FP64 and
      // re-evaluates its current point instead of ending the pass
      rescue = split && glob == ST_RUNNING && (local == ST_LINESEARCH_FAIL || local == ST_STAGNATED);
      if (rescue) local = ST_RUNNING;
      const bool last = L.pass_idx >= c.n_rescales - 1;
      const bool pass_end = (local != ST_RUNNING) && (glob == ST_RUNNING);
      const bool advance = pass_end && !last;
      code = (glob != ST_RUNNING) ? glob : ST_RUNNING;
      if (pass_end && last) code = local;
      if (advance) {
        const double rn = c.rescales[L.pass_idx + 1];
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
          const int i = reg_elem<NT, V>(tid, e);
          if (i < nd) {
            x[i] = s[i] * rn;
            gx[i] = gv[e] * rs / rn;
          }
        }
        L.opt_count = 0;
        L.pass_idx += 1;
        L.pass_evals = 0;
        L.stag = 0;
      }
      L.last_step = alpha;
      L.it += 1;
      L.n_evals += evals;
      if (c.record_every > 0 && c.record_cap > 0 && (L.it % c.record_every) == 0 &&
          L.rec_count < c.record_cap) {
        const size_t row = (size_t)lane * c.record_cap + L.rec_count;
        const size_t trow = rv_trow(r, lane, c.record_cap, L.rec_count);
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
          const int i = reg_elem<NT, V>(tid, e);
          if (i < nd) r.tr_s[trow + i] = wrap01(s[i]);
        }
        if (tid == 0) {
          r.tr_e[row] = L.f;
          r.tr_step[row] = L.last_step;
          r.tr_it[row] = L.it;
        }
        L.rec_count += 1;
      }
      start = code == ST_RUNNING && !rescue;
      if (rescue && code == ST_RUNNING) {
        double* se = r.s_eval + off;
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
          const int i = reg_elem<NT, V>(tid, e);
          if (i < nd) se[i] = x[i] / rs;
        }
        L.mode = MODE_REEVAL;
        L.stag = 0;
        L.opt_count = 0;
      }
    } else {
      const double a = L.ls.a;
      double* se = r.s_eval + off;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int i = reg_elem<NT, V>(tid, e);
        if (i < nd) se[i] = (x[i] + a * uu[e]) / rs;
      }
    }
  }
  if (start) {
    if (h32)
      lbfgs_start_reg<NT, EPT, V, true>(r, c, lane, L, R, alphas, rhos);
    else
      lbfgs_start_reg<NT, EPT, V, false>(r, c, lane, L, R, alphas, rhos);
  }
  const bool to_fp64 = code == ST_RUNNING && r.prec != nullptr &&
                       (rescue || (c.mixed_switch > 0.0 && L.f <= c.mixed_switch));
  if (h32 && to_fp64 && !rescue) {
    // the lane's later iterations read FP64 rows: widen its valid FP32 pairs
    // (slots 0 .. min(count, m) - 1; a rescue empties the history instead)
    const int nv = min(L.opt_count - 1, c.memory);
    const size_t hb = rv_hist(r, lane, c.memory);
    for (int t = 0; t < nv; ++t) {
      const size_t row = hb + (size_t)t * nd;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int i = reg_elem<NT, V>(tid, e);
        if (i < nd) {
          r.S[row + i] = (double)r.S32[row + i];
          r.Y[row + i] = (double)r.Y32[row + i];
        }
      }
    }
  }
  if (tid == 0) {
    r.ll[lane] = L;
    r.status[lane] = code;
    if (to_fp64) r.prec[lane] = 0;
  }
}

int lbfgs_register_ept(int nd) {
  for (int e : {1, 2, 4, 6, 8})
    if (nd <= e * kRegBlock) return e;
  return 0;
}

void launch_lbfgs_advance(const RelaxDev& r, const LbfgsCfg& c, int slots_ub, hipStream_t st,
                          int variant) {
  if (slots_ub <= 0) return;
  if (variant == 3) {
    hipLaunchKernelGGL((lbfgs_advance_kernel<kRegBlock, true>), dim3(slots_ub), dim3(kRegBlock), 0,
                       st, r, c);
    return;
  }
  const int ept = lbfgs_register_ept(r.nd);
  if (variant == 1 || ept == 0) {
    hipLaunchKernelGGL((lbfgs_advance_kernel<kRegBlock, false>), dim3(slots_ub), dim3(kRegBlock), 0,
                       st, r, c);
    return;
  }
  // 512 threads and 16-byte history loads: uniform batches of even N d <= 2048
  // (every lane's vectors and history rows then start 16-byte aligned).  One
  // such workgroup per CU: without the LDS reservation the dispatcher packs
  // two or three lanes onto a CU while the other shards' kernels hold the
  // rest, and lanes that share a CU share its memory pipeline (measured 25%
  // slower per launch); more than half of the CU's LDS keeps a second one out.
  if (r.loff == nullptr && r.nd % 2 == 0 && r.nd <= 2048) {
    constexpr size_t kStatic = sizeof(double) * (6 * (512 / kWave) + 2 * 1024);
    const size_t pad = kLdsPerCu / 2 + 1024 - kStatic;
    if (r.nd <= 1024)
      hipLaunchKernelGGL((lbfgs_advance_reg_kernel<512, 2, 2>), dim3(slots_ub), dim3(512), pad, st, r, c);
    else
      hipLaunchKernelGGL((lbfgs_advance_reg_kernel<512, 4, 2>), dim3(slots_ub), dim3(512), pad, st, r, c);
    return;
  }
  switch (ept) {
    case 1:
      hipLaunchKernelGGL((lbfgs_advance_reg_kernel<kRegBlock, 1, 1>), dim3(slots_ub), dim3(kRegBlock),
                         0, st, r, c);
      break;
    case 2:
      hipLaunchKernelGGL((lbfgs_advance_reg_kernel<kRegBlock, 2, 1>), dim3(slots_ub), dim3(kRegBlock),
                         0, st, r, c);
      break;
    case 4:
      hipLaunchKernelGGL((lbfgs_advance_reg_kernel<kRegBlock, 4, 1>), dim3(slots_ub), dim3(kRegBlock),
                         0, st, r, c);
      break;
    case 6:
      hipLaunchKernelGGL((lbfgs_advance_reg_kernel<kRegBlock, 6, 1>), dim3(slots_ub), dim3(kRegBlock),
                         0, st, r, c);
      break;
    default:
      hipLaunchKernelGGL((lbfgs_advance_reg_kernel<kRegBlock, 8, 1>), dim3(slots_ub), dim3(kRegBlock),
                         0, st, r, c);
      break;
  }
}

// ---------------------------------------------------------------------------
// L-BFGS in Gram form.  The two-loop recursion only ever forms dot products of
// s_i, y_i and g, so it is restated on the tables S^T Y, Y^T Y, S^T g, Y^T g
// (the vector-free two-loop of Chen, Wang and Zhou, "Large-scale L-BFGS using
// MapReduce", NIPS 2014):
//   alpha_t = rho_t (s_t.g - sum_{j newer} alpha_j s_t.y_j)          (newest first)
//   w_t     = gamma (y_t.g - sum_i alpha_i y_t.y_i)
//   c_t     = alpha_t - rho_t (w_t + sum_{j older} c_j s_j.y_t)       (oldest first)
//   u       = -(gamma g - gamma sum_i alpha_i y_i + sum_t c_t s_t)
// The same direction as lbfgs_start_iteration up to rounding.  Per iteration:
//   dot kernel    one pass over the history (5 dot products per slot, and
//                 g.g), spread over ceil(m / kGramSG) workgroups per lane;
//   solve kernel  the recursion on LDS-resident S^T Y, one workgroup per lane;
//   comb kernel   one streaming pass forming u and the first trial point,
//                 kCombBlock elements per workgroup.
// Every reduction has a fixed shape per lane size, so results do not depend
// on which other lanes are active.
// ---------------------------------------------------------------------------
constexpr int kGramSG = 4;       // history slots per dot workgroup
constexpr int kGramBlock = 256;
constexpr int kSolveBlock = 256;
constexpr int kCombBlock = 256;
constexpr int kGramVals = kGramSG * 5 + 1;

struct GramView {
  double* sy;    // [m][m]  sy[a][b] = s_a . y_b (physical slots)
  double* yy;    // [m][m]  symmetric
  double* sg;    // [m]     s_a . g
  double* yg;    // [m]     y_a . g
  double* cs;    // [m]     coefficient of s_a in q
  double* cy;    // [m]     coefficient of y_a in q
  double* misc;  // [8]     0: g.g, 1: gamma
};

__device__ __forceinline__ GramView gram_view(const RelaxDev& r, int m, int lane) {
  double* b = r.gram + (size_t)lane * r.gram_stride;
  GramView v;
  v.sy = b;
  v.yy = b + (size_t)m * m;
  v.sg = b + 2 * (size_t)m * m;
  v.yg = v.sg + m;
  v.cs = v.yg + m;
  v.cy = v.cs + m;
  v.misc = v.cy + m;
  return v;
}


// grid: slots_ub x ceil(m / kGramSG); workgroup (lane, group) owns slots
// group*kGramSG .. +kGramSG-1 and reduces over all of x.
__global__ __launch_bounds__(kGramBlock) void lbfgs_gram_dot_kernel(RelaxDev r, LbfgsCfg c) {
  __shared__ double red[kGramBlock / kWave][kGramVals];
  const int m = c.memory;
  const int n_groups = (m + kGramSG - 1) / kGramSG;
  const int slot = blockIdx.x / n_groups;
  const int grp = blockIdx.x % n_groups;
  if (slot >= *r.n_active) return;
  const int lane = r.active[slot];
  const int2 di = r.dir[lane];
  if (di.x == 0) return;
  const int count = di.y;
  const int nvalid = count < m ? count : m;
  const int midx = count % m;
  const bool newpair = count > 0;
  const int sn = newpair ? (count - 1) % m : 0;
  int sl[kGramSG];
  bool val[kGramSG];
  bool any = grp == 0;
#pragma unroll
  for (int k = 0; k < kGramSG; ++k) {
    sl[k] = grp * kGramSG + k;
    const int age = (sl[k] - midx + m) % m;
    val[k] = sl[k] < m && age >= m - nvalid;
    any = any || val[k];
  }
  if (!any) return;  // uniform per workgroup, before any barrier
  const int nd = r.nd;
  const size_t stride = r.stride;
  const double* g = r.g + (size_t)lane * stride;
  const double* S = r.S + (size_t)lane * m * stride;
  const double* Y = r.Y + (size_t)lane * m * stride;
  const double* Sn = S + (size_t)sn * stride;
  const double* Yn = Y + (size_t)sn * stride;
  double acc[kGramSG][5];
#pragma unroll
  for (int k = 0; k < kGramSG; ++k)
#pragma unroll
    for (int q = 0; q < 5; ++q) acc[k][q] = 0.0;
  double gg = 0.0;
  // two elements per pass (i, i + kGramBlock): twice the loads in flight
  for (int i0 = threadIdx.x; i0 < nd; i0 += 2 * kGramBlock) {
    double gi[2], sni[2], yni[2], sv[kGramSG][2], yv[kGramSG][2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = i0 + h * kGramBlock;
      const bool ok = i < nd;
      gi[h] = ok ? g[i] : 0.0;
      sni[h] = (ok && newpair) ? Sn[i] : 0.0;
      yni[h] = (ok && newpair) ? Yn[i] : 0.0;
#pragma unroll
      for (int k = 0; k < kGramSG; ++k) {
        sv[k][h] = (ok && val[k]) ? S[(size_t)sl[k] * stride + i] : 0.0;
        yv[k][h] = (ok && val[k]) ? Y[(size_t)sl[k] * stride + i] : 0.0;
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      gg += gi[h] * gi[h];
#pragma unroll
      for (int k = 0; k < kGramSG; ++k) {
        acc[k][0] += sv[k][h] * gi[h];
        acc[k][1] += yv[k][h] * gi[h];
        acc[k][2] += sni[h] * yv[k][h];
        acc[k][3] += sv[k][h] * yni[h];
        acc[k][4] += yni[h] * yv[k][h];
      }
    }
  }
  const int wid = threadIdx.x / kWave;
  const int lid = threadIdx.x % kWave;
#pragma unroll
  for (int k = 0; k < kGramSG; ++k) {
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const double v = wave_sum(acc[k][q]);
      if (lid == 0) red[wid][k * 5 + q] = v;
    }
  }
  {
    const double v = wave_sum(gg);
    if (lid == 0) red[wid][kGramSG * 5] = v;
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t >= kGramVals) return;
  double tot = 0.0;
#pragma unroll
  for (int w = 0; w < kGramBlock / kWave; ++w) tot += red[w][t];
  GramView G = gram_view(r, m, lane);
  if (t == kGramSG * 5) {
    if (grp == 0) G.misc[0] = tot;
    return;
  }
  const int k = t / 5;
  const int q = t % 5;
  int slk = grp * kGramSG;
  bool vk = false;
#pragma unroll
  for (int kk = 0; kk < kGramSG; ++kk)
    if (kk == k) {
      slk = sl[kk];
      vk = val[kk];
    }
  if (!vk) return;
  if (q == 0) G.sg[slk] = tot;
  if (q == 1) G.yg[slk] = tot;
  if (!newpair) return;
  if (q == 2) G.sy[(size_t)sn * m + slk] = tot;  // s_new . y_slot
  if (q == 3) G.sy[(size_t)slk * m + sn] = tot;  // s_slot . y_new
  if (q == 4) {
    G.yy[(size_t)sn * m + slk] = tot;
    G.yy[(size_t)slk * m + sn] = tot;
  }
}

// grid: slots_ub; the sequential parts run on wave 0, lane l owning relative
// ages l and l + 64 (age 0 = oldest valid pair).  S^T Y sits in LDS in
// physical slot order with an odd row stride.
constexpr int kSolveBatch = 16;  // global loads in flight per thread while staging

__global__ __launch_bounds__(kSolveBlock) void lbfgs_gram_solve_kernel(RelaxDev r, LbfgsCfg c) {
  __shared__ double Lsy[kGramMaxMem * (kGramMaxMem + 1)];
  __shared__ double La[kGramMaxMem];
  __shared__ double Lw[kGramMaxMem];
  __shared__ double Lr[kGramMaxMem];
  const int slot = blockIdx.x;
  if (slot >= *r.n_active) return;
  const int lane = r.active[slot];
  const int2 di = r.dir[lane];
  if (di.x == 0) return;
  const int m = c.memory;
  const int count = di.y;
  const int nvalid = count < m ? count : m;
  const int first = (count % m + m - nvalid) % m;  // physical slot of relative age 0
  const int tid = threadIdx.x;
  GramView G = gram_view(r, m, lane);
  double* rho = r.rho + (size_t)lane * m;
  const int sn = count > 0 ? (count - 1) % m : -1;
  double gamma;
  double rho_new = 0.0;
  if (count > 0) {
    const double sy = G.sy[(size_t)sn * m + sn];
    const double yy = G.yy[(size_t)sn * m + sn];
    rho_new = (sy == 0.0) ? 0.0 : 1.0 / sy;
    gamma = (yy > 0.0) ? sy / yy : 1.0;
  } else {
    gamma = nanmin(1.0, 1.0 / sqrt(G.misc[0]));
  }
  // waves 2 and 3: relative age a = tid - 128 prefetches y_b.y_a for every b
  // (overlapping the staging and the first loop)
  const int aw = tid - 2 * kWave;
  const bool wok = aw >= 0 && aw < nvalid;
  const int saw = wok ? (first + aw) % m : 0;
  double yyv[kGramMaxMem];
#pragma unroll
  for (int b = 0; b < kGramMaxMem; ++b)
    yyv[b] = (wok && b < nvalid) ? G.yy[(size_t)((first + b) % m) * m + saw] : 0.0;
  const int ld = m | 1;
  if (nvalid > 0) {
    const int total = m * m;
    for (int base = tid; base < total; base += kSolveBlock * kSolveBatch) {
      double v[kSolveBatch];
#pragma unroll
      for (int q = 0; q < kSolveBatch; ++q) {
        const int idx = base + q * kSolveBlock;
        v[q] = idx < total ? G.sy[idx] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < kSolveBatch; ++q) {
        const int idx = base + q * kSolveBlock;
        if (idx < total) {
          const int row = idx / m;
          Lsy[row * ld + (idx - row * m)] = v[q];
        }
      }
    }
  }
  for (int a = tid; a < nvalid; a += kSolveBlock) {
    const int sa = (first + a) % m;
    Lr[a] = (sa == sn) ? rho_new : rho[sa];
  }
  __syncthreads();
  const int a0 = tid;
  const int a1 = tid + kWave;
  const bool v0ok = a0 < nvalid;
  const bool v1ok = a1 < nvalid;
  const int s0 = (first + a0) % m;
  const int s1 = (first + a1) % m;
  if (tid < kWave) {
    // first loop, newest first: alpha_j = rho_j (s_j.g - sum_{t newer} alpha_t s_j.y_t)
    double v0 = v0ok ? G.sg[s0] : 0.0;
    double v1 = v1ok ? G.sg[s1] : 0.0;
    const double r0 = v0ok ? Lr[a0] : 0.0;
    const double r1 = v1ok ? Lr[a1] : 0.0;
    const double* row0 = Lsy + s0 * ld;
    const double* row1 = Lsy + s1 * ld;
    int sj = (first + nvalid - 1) % m;
    for (int j = nvalid - 1; j >= 0; --j) {
      const double l0 = row0[sj];
      const double l1 = row1[sj];
      sj = (sj == 0) ? m - 1 : sj - 1;
      double alpha;
      if (j >= kWave) {
        alpha = readlane_d(r1 * v1, j - kWave);
      } else {
        alpha = readlane_d(r0 * v0, j);
      }
      if (a0 < j) v0 = v0 - alpha * l0;
      if (a1 < j) v1 = v1 - alpha * l1;
      if (a0 == j) La[a0] = alpha;
      if (a1 == j) La[a1] = alpha;
    }
  }
  __syncthreads();
  // w_t = gamma (y_t.g - sum_i alpha_i y_t.y_i), oldest i first
  if (wok) {
    double w = G.yg[saw];
#pragma unroll
    for (int b = 0; b < kGramMaxMem; ++b)
      if (b < nvalid) w = w - La[b] * yyv[b];
    Lw[aw] = gamma * w;
  }
  __syncthreads();
  if (tid < kWave) {
    // second loop, oldest first: c_j = alpha_j - rho_j (w_j + sum_{t older} c_t s_t.y_j)
    const double r0 = v0ok ? Lr[a0] : 0.0;
    const double r1 = v1ok ? Lr[a1] : 0.0;
    const double al0 = v0ok ? La[a0] : 0.0;
    const double al1 = v1ok ? La[a1] : 0.0;
    const double w0 = v0ok ? Lw[a0] : 0.0;
    const double w1 = v1ok ? Lw[a1] : 0.0;
    double acc0 = 0.0, acc1 = 0.0, c0 = 0.0, c1 = 0.0;
    int sj = first;
    for (int j = 0; j < nvalid; ++j) {
      const double l0 = Lsy[sj * ld + s0];
      const double l1 = Lsy[sj * ld + s1];
      sj = (sj == m - 1) ? 0 : sj + 1;
      double cj;
      if (j >= kWave) {
        cj = readlane_d(al1 - r1 * (w1 + acc1), j - kWave);
      } else {
        cj = readlane_d(al0 - r0 * (w0 + acc0), j);
      }
      if (a0 == j) c0 = cj;
      if (a1 == j) c1 = cj;
      if (a0 > j) acc0 = acc0 + cj * l0;
      if (a1 > j) acc1 = acc1 + cj * l1;
    }
    if (v0ok) {
      G.cs[s0] = c0;
      G.cy[s0] = -gamma * al0;
    }
    if (v1ok) {
      G.cs[s1] = c1;
      G.cy[s1] = -gamma * al1;
    }
  }
  if (tid == 0) {
    if (count > 0) rho[sn] = rho_new;
    G.misc[1] = gamma;
    LbfgsLane L = r.ll[lane];
    L.opt_count = count + 1;
    L.pairs += nvalid;
    LsState& Q = L.ls;
    const double nan = __longlong_as_double(0x7ff8000000000000LL);
    Q.f0 = L.f;
    Q.slope0 = nan;
    Q.count = 0;
    Q.interval_found = 0;
    Q.prev_a = 0.0; Q.prev_v = L.f; Q.prev_s = nan;
    Q.low_a = 0.0; Q.low_v = L.f; Q.low_s = nan;
    Q.high_a = 0.0; Q.high_v = L.f; Q.high_s = nan;
    Q.cref_a = 0.0; Q.cref_v = L.f;
    Q.safe_a = 0.0; Q.safe_v = L.f;
    Q.last_dec = __longlong_as_double(0x7ff0000000000000LL);
    Q.pad0 = 1;  // slope0 = u.gx is reduced with the first trial evaluation
    ls_propose(Q);
    r.ll[lane] = L;
  }
}

// grid: slots_ub x ceil(nd / (2 kCombBlock)); two elements per thread.
__global__ __launch_bounds__(kCombBlock) void lbfgs_gram_comb_kernel(RelaxDev r, LbfgsCfg c,
                                                                   int n_ctiles) {
  const int slot = blockIdx.x / n_ctiles;
  const int tile = blockIdx.x % n_ctiles;
  if (slot >= *r.n_active) return;
  const int lane = r.active[slot];
  const int2 di = r.dir[lane];
  if (di.x == 0) return;
  const int i0 = tile * 2 * kCombBlock + (int)threadIdx.x;
  if (i0 >= r.nd) return;
  const bool two = i0 + kCombBlock < r.nd;
  const int m = c.memory;
  const int count = di.y;
  const int nvalid = count < m ? count : m;
  const int first = (count % m + m - nvalid) % m;
  const GramView G = gram_view(r, m, lane);
  const double a = r.ll[lane].ls.a;
  const double rs = c.rescales[r.ll[lane].pass_idx];
  const size_t stride = r.stride;
  const size_t off = (size_t)lane * stride;
  const double* S = r.S + (size_t)lane * m * stride + i0;
  const double* Y = r.Y + (size_t)lane * m * stride + i0;
  const double gamma = G.misc[1];
  const double g0 = r.g[off + i0];
  const double g1 = two ? r.g[off + i0 + kCombBlock] : 0.0;
  double q0 = gamma * g0;
  double q1 = gamma * g1;
#pragma unroll 4
  for (int t = 0; t < nvalid; ++t) {
    const int sl = (first + t) % m;
    const double cy = G.cy[sl];
    const double cs = G.cs[sl];
    const double* Ys = Y + (size_t)sl * stride;
    const double* Ss = S + (size_t)sl * stride;
    const double y0 = Ys[0];
    const double s0 = Ss[0];
    const double y1 = two ? Ys[kCombBlock] : 0.0;
    const double s1 = two ? Ss[kCombBlock] : 0.0;
    q0 = q0 + cy * y0;
    q0 = q0 + cs * s0;
    q1 = q1 + cy * y1;
    q1 = q1 + cs * s1;
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (h == 1 && !two) break;
    const size_t i = off + i0 + h * kCombBlock;
    const double ui = -(h == 0 ? q0 : q1);
    const double xi = r.x[i];
    r.u[i] = ui;
    r.px[i] = xi;
    r.pg[i] = h == 0 ? g0 : g1;
    r.s_eval[i] = (xi + a * ui) / rs;
  }
}

void launch_lbfgs_gram(const RelaxDev& r, const LbfgsCfg& c, int slots_ub, int part,
                       hipStream_t st) {
  if (slots_ub <= 0) return;
  if (part == 1) {
    const int n_groups = (c.memory + kGramSG - 1) / kGramSG;
    hipLaunchKernelGGL(lbfgs_gram_dot_kernel, dim3(slots_ub * n_groups), dim3(kGramBlock), 0, st,
                       r, c);
  } else if (part == 2) {
    hipLaunchKernelGGL(lbfgs_gram_solve_kernel, dim3(slots_ub), dim3(kSolveBlock), 0, st, r, c);
  } else {
    const int n_ctiles = (r.nd + 2 * kCombBlock - 1) / (2 * kCombBlock);
    hipLaunchKernelGGL(lbfgs_gram_comb_kernel, dim3(slots_ub * n_ctiles), dim3(kCombBlock), 0, st,
                       r, c, n_ctiles);
  }
}

// ---------------------------------------------------------------------------
// finalize (descent.py:186-197, lbfgs.py:316-329): wrapped s and summaries
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kAdvBlock) void finalize_kernel(RelaxDev r, int method, LbfgsCfg lc,
                                                             RelaxOutDev o) {
  __shared__ double red[16];
  const int lane = blockIdx.x;
  const size_t off = rv_off(r, lane);
  const int nd = rv_nd(r, lane);
  const double* s = r.s + off;
  const double* g = r.g + off;
  double gg = 0.0;
  for (int i = threadIdx.x; i < nd; i += blockDim.x) {
    o.s[off + i] = wrap01(s[i]);
    gg += g[i] * g[i];
  }
  gg = block_sum<kAdvBlock>(gg, red);
  if (threadIdx.x == 0) {
    if (method == 0) {
      const DescentLane L = r.dl[lane];
      o.energy[lane] = L.f;
      o.grad_norm[lane] = sqrt(gg);
      o.n_evals[lane] = L.n_evals;
      o.n_iters[lane] = L.it;
    } else {
      const LbfgsLane L = r.ll[lane];
      o.energy[lane] = L.f;
      o.grad_norm[lane] = sqrt(gg) * lc.rescales[L.pass_idx];
      o.n_evals[lane] = L.n_evals;
      o.n_iters[lane] = L.it;
    }
    o.status[lane] = r.status[lane];
  }
}

void launch_finalize(const RelaxDev& r, int method, const LbfgsCfg& lc, const RelaxOutDev& o,
                     hipStream_t st) {
  hipLaunchKernelGGL(finalize_kernel, dim3(r.B), dim3(kAdvBlock), 0, st, r, method, lc, o);
}

}  // namespace cco
