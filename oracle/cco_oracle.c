/*
 * cco_oracle.c -- plain-C restatement of the stealthy hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as a second oracle at sizes where
 * NumPy is slow, and by bench.py's cpu_baseline leg as the timed CPU
 * baseline.  The product never links or loads it.
 *
 * Follows (paths in the sdawley1/xla-cco tree):
 *   energy + gradient : src/xla_cco/energy.py:45-50, 206-225, 301-324
 *                       (trig per (k, j): theta = 2 pi n.s_j, cos and sin)
 *                       phases = 1 forms exp(2 pi i n.s_j) instead as the
 *                       product over axes of z_a^{n_a}, z_a = exp(2 pi i s_a)
 *                       rounded once from 80-bit cosl/sinl and raised to
 *                       integer powers by repeated multiplication: the
 *                       accuracy class of the factorised phase tables (errors
 *                       of a few ulps per power step, against 2 pi |n.s| ulps
 *                       of the rounded argument in the trig form), which
 *                       matters only at the round-off floor (Phi ~ 1e-27)
 *   adaptive descent  : src/xla_cco/minimize/descent.py:92-197
 *   L-BFGS            : src/xla_cco/minimize/lbfgs.py:163-329 with the optax
 *                       0.2.8 lbfgs + zoom line search restated in
 *                       oracle/oracle.py (same order of operations)
 * Lanes (configurations) run in parallel with OpenMP.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define TWO_PI (2.0 * 3.14159265358979323846)
/* zoom phase ends once |high - low| <= this (optax's float64 threshold); a test
   rebuilds with another value to show the published outcomes reject it */
#ifndef LS_INTERVAL_MIN
#define LS_INTERVAL_MIN 1e-10
#endif

enum { ST_RUNNING = 0, ST_ENERGY_TOL = 1, ST_STEP_FLOOR = 2, ST_BUDGET = 3, ST_STAGNATED = 4,
       ST_LS_FAIL = 5, ST_GRAD_TOL = 6 };

typedef struct {
  const double* n; /* M x d (float64 integers) */
  const double* w; /* M */
  int M, d;
  double volume;
  int phases; /* 0: trig of the rounded argument, 1: per-axis power tables */
} kset_t;

/* Per-axis power tables of one configuration: tab[(j*d + a)*(P+1) + p] =
   z_{j,a}^p, z = exp(2 pi i s_{j,a}), p = 0..P (complex, re/im interleaved). */
static double* phase_tables(const double* s, int N, const kset_t* K, int* P_out) {
  int P = 0;
  for (size_t i = 0; i < (size_t)K->M * K->d; ++i) {
    const int e = (int)fabs(K->n[i]);
    if (e > P) P = e;
  }
  const size_t per = (size_t)P + 1;
  double* tab = (double*)malloc(sizeof(double) * 2 * per * (size_t)N * K->d);
  for (size_t ja = 0; ja < (size_t)N * K->d; ++ja) {
    const long double th = 6.28318530717958647692528676655900577L * (long double)s[ja];
    const double zr = (double)cosl(th), zi = (double)sinl(th);
    double* t = tab + 2 * per * ja;
    t[0] = 1.0;
    t[1] = 0.0;
    for (size_t p = 1; p < per; ++p) {
      t[2 * p] = t[2 * p - 2] * zr - t[2 * p - 1] * zi;
      t[2 * p + 1] = t[2 * p - 2] * zi + t[2 * p - 1] * zr;
    }
  }
  *P_out = P;
  return tab;
}

/* exp(2 pi i n.s_j) as the product of the axes' table entries (conjugate for n_a < 0). */
static void table_phase(const double* tab, int P, int d, int j, const double* nk, double* cr,
                        double* ci) {
  double pr = 1.0, pi_ = 0.0;
  for (int a = 0; a < d; ++a) {
    const int e = (int)nk[a];
    const double* t = tab + 2 * ((size_t)P + 1) * ((size_t)j * d + a) + 2 * (size_t)(e < 0 ? -e : e);
    const double er = t[0], ei = e < 0 ? -t[1] : t[1];
    const double tr = pr * er - pi_ * ei;
    pi_ = pr * ei + pi_ * er;
    pr = tr;
  }
  *cr = pr;
  *ci = pi_;
}

/* Phi, dPhi/ds (N x d), Re/Im rho (M, may be NULL). */
void oracle_energy_grad(const double* s, int N, const kset_t* K, double* phi, double* grad,
                        double* re_out, double* im_out) {
  const int d = K->d, M = K->M;
  double* re = (double*)malloc(sizeof(double) * (size_t)M);
  double* im = (double*)malloc(sizeof(double) * (size_t)M);
  int P = 0;
  double* tab = K->phases ? phase_tables(s, N, K, &P) : NULL;
  double value = 0.0;
  for (int k = 0; k < M; ++k) {
    double a = 0.0, b = 0.0;
    for (int j = 0; j < N; ++j) {
      if (tab) {
        double cr, ci;
        table_phase(tab, P, d, j, K->n + (size_t)k * d, &cr, &ci);
        a += cr;
        b += ci;
        continue;
      }
      double dot = 0.0;
      for (int c = 0; c < d; ++c) dot += K->n[(size_t)k * d + c] * s[(size_t)j * d + c];
      const double th = TWO_PI * dot;
      a += cos(th);
      b += sin(th);
    }
    re[k] = a;
    im[k] = b;
    value += K->w[k] * (a * a + b * b);
  }
  *phi = value / K->volume;
  if (grad) {
    const double pref = 2.0 * TWO_PI / K->volume;
    for (int j = 0; j < N; ++j) {
      double acc[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < M; ++k) {
        double cs, sn;
        if (tab) {
          table_phase(tab, P, d, j, K->n + (size_t)k * d, &cs, &sn);
        } else {
          double dot = 0.0;
          for (int c = 0; c < d; ++c) dot += s[(size_t)j * d + c] * K->n[(size_t)k * d + c];
          const double th = TWO_PI * dot;
          cs = cos(th);
          sn = sin(th);
        }
        const double coef = K->w[k] * (im[k] * cs - re[k] * sn);
        for (int c = 0; c < d; ++c) acc[c] += coef * K->n[(size_t)k * d + c];
      }
      for (int c = 0; c < d; ++c) grad[(size_t)j * d + c] = pref * acc[c];
    }
  }
  if (re_out) memcpy(re_out, re, sizeof(double) * (size_t)M);
  if (im_out) memcpy(im_out, im, sizeof(double) * (size_t)M);
  free(tab);
  free(re);
  free(im);
}

/* Batched energy+force over B lanes (OpenMP over lanes). */
void oracle_energy_grad_batch(const double* s, int B, int N, const double* n, const double* w,
                              int M, int d, double volume, double* phi, double* grad,
                              int threads, int phases) {
  kset_t K = {n, w, M, d, volume, phases};
  const size_t nd = (size_t)N * d;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int b = 0; b < B; ++b)
    oracle_energy_grad(s + b * nd, N, &K, phi + b, grad ? grad + b * nd : NULL, NULL, NULL);
}

static double vdot(const double* a, const double* b, size_t n) {
  double acc = 0.0;
  for (size_t i = 0; i < n; ++i) acc += a[i] * b[i];
  return acc;
}

static double nanmax2(double a, double b) { return (a != a || b != b) ? NAN : (a > b ? a : b); }
static double nanmin2(double a, double b) { return (a != a || b != b) ? NAN : (a < b ? a : b); }

static double wrap01(double v) {
  const double x = v - floor(v);
  const double below = nextafter(1.0, 0.0);
  return x < below ? x : below;
}

/* ------------------------------------------------------------------ */
/* adaptive descent                                                     */
/* ------------------------------------------------------------------ */
typedef struct {
  double step_init, step_min, shrink, grow, rel_tol, energy_tol;
  int64_t max_evals;
} descent_cfg;

static int descent_status(double f, const double* g, size_t nd, double step, int64_t it,
                          const descent_cfg* c) {
  int code = it >= c->max_evals / 2 ? ST_BUDGET : ST_RUNNING;
  if (step <= c->step_min) code = ST_STEP_FLOOR;
  int all_zero = 1;
  for (size_t i = 0; i < nd; ++i)
    if (g[i] != 0.0) all_zero = 0;
  if (all_zero) code = ST_GRAD_TOL;
  if (f < c->energy_tol) code = ST_ENERGY_TOL;
  return code;
}

static void descent_one(const kset_t* K, const double* s0, int N, const descent_cfg* c,
                        int64_t eval_cap, double* s_out, double* e_out, int64_t* evals_out,
                        int64_t* iters_out, int64_t* status_out) {
  const size_t nd = (size_t)N * K->d;
  double* s = (double*)malloc(sizeof(double) * nd);
  double* g = (double*)malloc(sizeof(double) * nd);
  double* xn = (double*)malloc(sizeof(double) * nd);
  double* g2 = (double*)malloc(sizeof(double) * nd);
  memcpy(s, s0, sizeof(double) * nd);
  double f;
  oracle_energy_grad(s, N, K, &f, g, NULL, NULL);
  double step = c->step_init;
  int64_t it = 0, n_evals = 0, calls = 1;
  int status = descent_status(f, g, nd, step, it, c);
  while (status == ST_RUNNING && calls < eval_cap) {
    const double norm = sqrt(vdot(g, g, nd));
    for (size_t i = 0; i < nd; ++i) {
      const double gn = norm > 0.0 ? g[i] / norm : 0.0 * g[i];
      xn[i] = s[i] - gn * step;
    }
    double f2;
    oracle_energy_grad(xn, N, K, &f2, g2, NULL, NULL);
    ++calls;
    const double df = f - f2;
    const double df2 = step * norm;
    const int reject = fabs(df - df2) / df2 > c->rel_tol;
    if (reject) {
      step = step * c->shrink;
    } else {
      step = step * c->grow;
      memcpy(s, xn, sizeof(double) * nd);
      memcpy(g, g2, sizeof(double) * nd);
      f = f2;
    }
    it += 1;
    n_evals += 2;
    status = descent_status(f, g, nd, step, it, c);
  }
  for (size_t i = 0; i < nd; ++i) s_out[i] = wrap01(s[i]);
  *e_out = f;
  *evals_out = n_evals;
  *iters_out = it;
  *status_out = status;
  free(s);
  free(g);
  free(xn);
  free(g2);
}

/* ------------------------------------------------------------------ */
/* L-BFGS with zoom line search                                         */
/* ------------------------------------------------------------------ */
typedef struct {
  double rescales[8];
  int n_rescales, memory;
  int64_t max_evals_per_pass;
  double energy_tol, grad_tol;
  int stagnation_patience, max_ls;
} lbfgs_cfg;

static double cubicmin(double a, double fa, double fpa, double b, double fb, double c, double fc) {
  const double C = fpa, db = b - a, dc = c - a;
  const double denom = (db * dc) * (db * dc) * (db - dc);
  const double d2a = fb - fa - C * db, d2b = fc - fa - C * dc;
  const double A = (dc * dc * d2a + (-(db * db)) * d2b) / denom;
  const double Bq = ((-(dc * dc * dc)) * d2a + db * db * db * d2b) / denom;
  const double radical = Bq * Bq - 3.0 * A * C;
  return a + (-Bq + sqrt(radical)) / (3.0 * A);
}

static double quadmin(double a, double fa, double fpa, double b, double fb) {
  const double db = b - a;
  const double Bq = (fb - fa - fpa * db) / (db * db);
  return a - fpa / (2.0 * Bq);
}

static double dec_err(double a, double va, double sa, double v0, double s0) {
  double dec = va - (v0 + 1e-4 * a * s0);
  double approx = sa - (2.0 * 1e-4 - 1.0) * s0;
  const double delta = va - v0 - 1e-6 * fabs(v0);
  approx = nanmax2(approx, delta);
  dec = nanmin2(approx, dec);
  return nanmax2(dec, 0.0);
}

static double curv_err(double sa, double s0) { return nanmax2(fabs(sa) - 0.9 * fabs(s0), 0.0); }

typedef struct {
  const kset_t* K;
  int N;
  double r;
  double* tmp_s;
  int64_t calls;
} phi_x_t;

/* Phi(x / r) and dPhi/dx (lbfgs.py:75-97). */
static void phi_x(phi_x_t* P, const double* x, double* f, double* gx) {
  const size_t nd = (size_t)P->N * P->K->d;
  for (size_t i = 0; i < nd; ++i) P->tmp_s[i] = x[i] / P->r;
  oracle_energy_grad(P->tmp_s, P->N, P->K, f, gx, NULL, NULL);
  for (size_t i = 0; i < nd; ++i) gx[i] = gx[i] / P->r;
  P->calls++;
}

/* One zoom line search; on return x_new/g_new/f_new hold the chosen point. */
static void zoom(phi_x_t* P, const double* x, const double* u, double f0, const double* g0,
                 int max_steps, double* x_new, double* g_new, double* f_new, double* alpha,
                 int* steps, double* dec_out, double* gtrial, double* gsafe, size_t nd) {
  const double slope0 = vdot(u, g0, nd);
  int count = 0, found = 0, done = 0, failed = 0;
  double pa = 0, pv = f0, ps = slope0;
  double la = 0, lv = f0, ls = slope0, ha = 0, hv = f0, hs = slope0;
  double ca = 0, cv = f0;
  double sa_ = 0, sv_ = f0;
  int have_safe = 0;
  double a = 0, va = f0, dec = INFINITY;
  memcpy(gsafe, g0, sizeof(double) * nd);
  while (!(done || failed)) {
    double delta = 0.0;
    if (!found) {
      a = count == 0 ? 1.0 : 2.0 * pa;
    } else {
      delta = fabs(ha - la);
      const double left = fmin(ha, la), right = fmax(ha, la);
      const double mc = cubicmin(la, lv, ls, ha, hv, ca, cv);
      const int use_c = (mc > left + 0.2 * delta) && (mc < right - 0.2 * delta);
      const double mq = quadmin(la, lv, ls, ha, hv);
      const int use_q = !use_c && (mq > left + 0.1 * delta) && (mq < right - 0.1 * delta);
      a = use_c ? mc : (use_q ? mq : (la + ha) / 2.0);
    }
    for (size_t i = 0; i < nd; ++i) x_new[i] = x[i] + a * u[i];
    phi_x(P, x_new, &va, gtrial);
    const double sa = vdot(gtrial, u, nd);
    dec = dec_err(a, va, sa, f0, slope0);
    const double err = nanmax2(dec, curv_err(sa, slope0));
    int safe;
    if (!found) {
      safe = dec <= 0.0;
      const int set_high = (dec > 0.0) || (va >= pv && count > 0);
      const int set_low = (sa >= 0.0) && !set_high;
      if (set_low) {
        ha = pa; hv = pv; hs = ps;
        la = a; lv = va; ls = sa;
      } else {
        la = pa; lv = pv; ls = ps;
        ha = a; hv = va; hs = sa;
      }
      found = set_high || set_low || err <= 0.0;
      done = err <= 0.0;
      ca = la; cv = lv;
      count++;
      failed = (count >= max_steps) && !done;
    } else {
      safe = (dec <= 0.0) && (va < sv_);
      done = err <= 0.0;
      const int hi_mid = (dec > 0.0) || (va >= lv);
      const int hi_low = (sa * (ha - la) >= 0.0) && !hi_mid;
      const double ola = la, olv = lv, ols = ls, oha = ha, ohv = hv;
      if (hi_mid) {
        ha = a; hv = va; hs = sa;
      } else if (hi_low) {
        ha = ola; hv = olv; hs = ols;
      }
      if (!hi_mid) {
        la = a; lv = va; ls = sa;
      }
      if (hi_mid || hi_low) {
        ca = oha; cv = ohv;
      } else {
        ca = ola; cv = olv;
      }
      count++;
      failed = ((count >= max_steps) || (delta <= LS_INTERVAL_MIN)) && !done;
    }
    (void)hs;
    pa = a; pv = va; ps = sa;
    if (safe) {
      sa_ = a;
      sv_ = va;
      have_safe = 1;
      memcpy(gsafe, gtrial, sizeof(double) * nd);
    }
  }
  if (failed && have_safe && sa_ > 0.0) {
    for (size_t i = 0; i < nd; ++i) x_new[i] = x[i] + sa_ * u[i];
    memcpy(g_new, gsafe, sizeof(double) * nd);
    *f_new = sv_;
    *alpha = sa_;
    *dec_out = 0.0;
  } else {
    memcpy(g_new, gtrial, sizeof(double) * nd);
    *f_new = va;
    *alpha = a;
    *dec_out = dec;
  }
  *steps = count;
}

static void lbfgs_one(const kset_t* K, const double* s0, int N, const lbfgs_cfg* c,
                      int64_t eval_cap, double* s_out, double* e_out, int64_t* evals_out,
                      int64_t* iters_out, int64_t* status_out) {
  const size_t nd = (size_t)N * K->d;
  const int m = c->memory;
  double* buf = (double*)calloc(nd * (12 + 2 * (size_t)m) + (size_t)m, sizeof(double));
  double *s = buf, *x = s + nd, *gx = x + nd, *px = gx + nd, *pg = px + nd, *u = pg + nd;
  double *xn = u + nd, *gn = xn + nd, *gt = gn + nd, *gs = gt + nd, *tmp = gs + nd, *gs0 = tmp + nd;
  double *S = gs0 + nd, *Y = S + nd * m, *rho = Y + nd * m;
  double* alphas = (double*)malloc(sizeof(double) * (size_t)m);
  phi_x_t P = {K, N, 1.0, tmp, 0};
  memcpy(s, s0, sizeof(double) * nd);
  double f;
  oracle_energy_grad(s, N, K, &f, gs0, NULL, NULL);
  P.calls = 1;
  double r = c->rescales[0];
  for (size_t i = 0; i < nd; ++i) {
    x[i] = s0[i] * r;
    gx[i] = gs0[i] / r;
  }
  int status = f < c->energy_tol ? ST_ENERGY_TOL : ST_RUNNING;
  int64_t n_evals = 1, it = 0, pass_evals = 0, stag = 0;
  int pass = 0, count = 0;
  while (status == ST_RUNNING && P.calls < eval_cap) {
    r = c->rescales[pass];
    P.r = r;
    double gamma;
    if (count > 0) {
      const int slot = (count - 1) % m;
      double* Ss = S + nd * slot;
      double* Ys = Y + nd * slot;
      for (size_t i = 0; i < nd; ++i) {
        Ss[i] = x[i] - px[i];
        Ys[i] = gx[i] - pg[i];
      }
      const double sy = vdot(Ys, Ss, nd);
      rho[slot] = sy == 0.0 ? 0.0 : 1.0 / sy;
      const double yy = vdot(Ys, Ys, nd);
      gamma = yy > 0.0 ? sy / yy : 1.0;
    } else {
      const double gn0 = sqrt(vdot(gx, gx, nd));
      gamma = (gn0 != gn0) ? NAN : (gn0 > 0.0 ? fmin(1.0, 1.0 / gn0) : 1.0);
    }
    const int nvalid = count < m ? count : m;
    const int midx = count % m;
    memcpy(u, gx, sizeof(double) * nd);
    for (int t = m - 1; t >= m - nvalid; --t) {
      const int sl = (midx + t) % m;
      const double a = rho[sl] * vdot(S + nd * sl, u, nd);
      alphas[t] = a;
      for (size_t i = 0; i < nd; ++i) u[i] = u[i] + (-a) * Y[nd * sl + i];
    }
    for (size_t i = 0; i < nd; ++i) u[i] = gamma * u[i];
    for (int t = m - nvalid; t < m; ++t) {
      const int sl = (midx + t) % m;
      const double b = rho[sl] * vdot(Y + nd * sl, u, nd);
      const double coef = alphas[t] - b;
      for (size_t i = 0; i < nd; ++i) u[i] = u[i] + coef * S[nd * sl + i];
    }
    for (size_t i = 0; i < nd; ++i) {
      u[i] = -u[i];
      px[i] = x[i];
      pg[i] = gx[i];
    }
    count += 1;
    double fnew, alpha, dec;
    int steps;
    zoom(&P, x, u, f, gx, c->max_ls, xn, gn, &fnew, &alpha, &steps, &dec, gt, gs, nd);
    const int fail = (dec > 0.0) || !isfinite(fnew);
    const int64_t evals = steps > 1 ? steps : 1;
    const double f_old = f;
    if (!fail) {
      for (size_t i = 0; i < nd; ++i) {
        x[i] = xn[i];
        s[i] = xn[i] / r;
        gx[i] = gn[i];
      }
      f = fnew;
    }
    stag = f < f_old ? 0 : stag + 1;
    pass_evals += evals;
    const double gnorm = sqrt(vdot(gx, gx, nd)) * r;
    int glob = gnorm < c->grad_tol ? ST_GRAD_TOL : ST_RUNNING;
    if (f < c->energy_tol) glob = ST_ENERGY_TOL;
    int local = pass_evals >= c->max_evals_per_pass ? ST_BUDGET : ST_RUNNING;
    if (stag >= c->stagnation_patience) local = ST_STAGNATED;
    if (fail) local = ST_LS_FAIL;
    const int last = pass >= c->n_rescales - 1;
    const int pass_end = local != ST_RUNNING && glob == ST_RUNNING;
    status = glob != ST_RUNNING ? glob : ST_RUNNING;
    if (pass_end && last) status = local;
    if (pass_end && !last) {
      const double rn = c->rescales[pass + 1];
      for (size_t i = 0; i < nd; ++i) {
        x[i] = s[i] * rn;
        gx[i] = gx[i] * r / rn;
      }
      count = 0;
      pass += 1;
      pass_evals = 0;
      stag = 0;
    }
    it += 1;
    n_evals += evals;
  }
  for (size_t i = 0; i < nd; ++i) s_out[i] = wrap01(s[i]);
  *e_out = f;
  *evals_out = n_evals;
  *iters_out = it;
  *status_out = status;
  free(alphas);
  free(buf);
}

/* Batched relax (method 0 descent, 1 L-BFGS), OpenMP over lanes.  eval_cap
 * bounds the energy evaluations per lane (a bounded timing sample); a lane
 * stopped by the cap reports status 0 (running). */
void oracle_relax(const double* s0, int B, int N, const double* n, const double* w, int M, int d,
                  double volume, int method, const double* dcfg, const int64_t* dint,
                  const double* lcfg, const int64_t* lint, int64_t eval_cap, double* s_out,
                  double* e_out, int64_t* evals_out, int64_t* iters_out, int64_t* status_out,
                  int threads, int phases) {
  kset_t K = {n, w, M, d, volume, phases};
  const size_t nd = (size_t)N * d;
  descent_cfg dc;
  lbfgs_cfg lc;
  memset(&dc, 0, sizeof(dc));
  memset(&lc, 0, sizeof(lc));
  if (method == 0) {
    dc.step_init = dcfg[0];
    dc.step_min = dcfg[1];
    dc.shrink = dcfg[2];
    dc.grow = dcfg[3];
    dc.rel_tol = dcfg[4];
    dc.energy_tol = dcfg[5];
    dc.max_evals = dint[0];
  } else {
    lc.n_rescales = (int)lint[0];
    for (int i = 0; i < lc.n_rescales && i < 8; ++i) lc.rescales[i] = lcfg[i];
    lc.energy_tol = lcfg[8];
    lc.grad_tol = lcfg[9];
    lc.memory = (int)lint[1];
    lc.max_evals_per_pass = lint[2];
    lc.stagnation_patience = (int)lint[3];
    lc.max_ls = (int)lint[4];
  }
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int b = 0; b < B; ++b) {
    if (method == 0)
      descent_one(&K, s0 + b * nd, N, &dc, eval_cap, s_out + b * nd, e_out + b, evals_out + b,
                  iters_out + b, status_out + b);
    else
      lbfgs_one(&K, s0 + b * nd, N, &lc, eval_cap, s_out + b * nd, e_out + b, evals_out + b,
                iters_out + b, status_out + b);
  }
}
