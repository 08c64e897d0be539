// Device-side batched dynamics: MD steps (velocity Verlet, Andersen resets,
// BAOAB Langevin) with the tracked density waves, the Dormand-Prince 5(4)
// and ROS2 gradient flows, sharing the energy kernels.
//
// Restated algorithms:
//   MD      : src/xla_cco/dynamics/integrators.py:83-176 (verlet_step,
//             andersen_reset, baoab_step), run.py:229-256 (force at
//             s = r inv(B), F = -g inv(B)^T), :259-291 (tracked rho),
//             :415-503 (_step, _record), observables.py:56-125 (densities,
//             COM correction, winding), thermal.py:100-106 (KE)
//   flow    : src/xla_cco/minimize/flow.py:212-251 (status, velocity),
//             :337-364 (scaled error, DOPRI5 proposal), :401-451 (step)
// One workgroup per active lane; threads own whole particles (d components)
// because the metric transforms mix the components of one particle.
// Contraction is off so the update arithmetic rounds like its NumPy
// statement.
#pragma clang fp contract(off)

#include "minimize_kernels.h"

namespace cco {

constexpr int kDynBlock = 1024;

__device__ __forceinline__ double nanmax_d(double a, double b) {
  if (a != a || b != b) return __longlong_as_double(0x7ff8000000000000LL);
  return a > b ? a : b;
}

// ---------------------------------------------------------------------------
// counter-based draws (Philox-4x32-10, Salmon et al. SC 2011): the same
// split / normal / uniform / integer definitions as the host module
// xla_cco/dynamics/rng.py, so host and device thermostat steps draw alike.
// ---------------------------------------------------------------------------
constexpr unsigned kTagSplit = 0x53504C54u, kTagNormal = 0x4E524D4Cu, kTagUnif = 0x554E4946u;

__device__ inline void philox(unsigned c[4], unsigned k0, unsigned k1) {
  for (int r = 0; r < 10; ++r) {
    const unsigned long long p0 = 0xD2511F53ull * c[0];
    const unsigned long long p1 = 0xCD9E8D57ull * c[2];
    const unsigned n0 = (unsigned)(p1 >> 32) ^ c[1] ^ k0;
    const unsigned n1 = (unsigned)p1;
    const unsigned n2 = (unsigned)(p0 >> 32) ^ c[3] ^ k1;
    const unsigned n3 = (unsigned)p0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// keys[j] = words 2j, 2j + 1 of philox((i, 0, 0, TAG_SPLIT), key), num <= 4
__device__ inline void rng_split(const unsigned key[2], int num, unsigned out[][2]) {
  unsigned w[8];
  for (int i = 0; i < (2 * num + 3) / 4; ++i) {
    unsigned c[4] = {(unsigned)i, 0u, 0u, kTagSplit};
    philox(c, key[0], key[1]);
    for (int q = 0; q < 4; ++q) w[4 * i + q] = c[q];
  }
  for (int j = 0; j < num; ++j) {
    out[j][0] = w[2 * j];
    out[j][1] = w[2 * j + 1];
  }
}

__device__ inline double u53(unsigned hi, unsigned lo) {
  return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * (1.0 / 9007199254740992.0);
}

// Normal draw number e of the stream of `key` (Box-Muller pair e / 2).
__device__ inline double rng_normal(const unsigned key[2], int e) {
  unsigned c[4] = {(unsigned)(e >> 1), 0u, 0u, kTagNormal};
  philox(c, key[0], key[1]);
  const double u1 = u53(c[0], c[1]);
  const double u2 = u53(c[2], c[3]);
  const double rho = sqrt(-2.0 * log1p(-u1));
  return (e & 1) ? rho * sin(2.0 * M_PI * u2) : rho * cos(2.0 * M_PI * u2);
}

__device__ inline void rng_unif_words(const unsigned key[2], unsigned c[4]) {
  c[0] = 0u;
  c[1] = 0u;
  c[2] = 0u;
  c[3] = kTagUnif;
  philox(c, key[0], key[1]);
}

// ---------------------------------------------------------------------------
// molecular dynamics (dynamics/integrators.py:83-176, run.py:383-537)
// ---------------------------------------------------------------------------
// verlet / andersen: v <- v + dt/2 f, r <- r + dt v.  baoab: v <- v + dt/2 f,
// r <- r + dt/2 v, v <- c v + sqrt((1 - c^2) T) xi, r <- r + dt/2 v, with
// (key', k_xi) = split(key) and xi the normal stream of k_xi in (j, a)
// order.  Then s_eval = r inv(B).
// grid: slots_ub x chunks (workgroups per lane; one for baoab, whose draw keys
// the lane's workgroup reads and then replaces).
__global__ __launch_bounds__(kDynBlock) void md_prepare_kernel(RelaxDev R, MdCfg c, int chunks) {
  const int slot = blockIdx.x / chunks;
  const int part = blockIdx.x % chunks;
  if (slot >= *R.n_active) return;
  const int lane = R.active[slot];
  const size_t off = (size_t)lane * R.stride;
  double* r = R.r + off;
  double* v = R.v + off;
  const double* f = R.f + off;
  double* se = R.s_eval + off;
  const int d = c.d;
  MdLane L = R.ml[lane];
  unsigned ks[2][2];
  double amp = 0.0;
  if (c.scheme == 2) {
    rng_split(L.key, 2, ks);
    amp = sqrt(c.ou_var * L.temperature);
  }
  for (int j = part * blockDim.x + threadIdx.x; j < c.n; j += chunks * blockDim.x) {
    double rn[3];
    for (int a = 0; a < d; ++a) {
      const int e = j * d + a;
      double vh = v[e] + c.half * f[e];
      if (c.scheme == 2) {
        double x = r[e] + c.half * vh;
        vh = c.ou_c * vh + amp * rng_normal(ks[1], e);
        x = x + c.half * vh;
        rn[a] = x;
      } else {
        rn[a] = r[e] + c.dt * vh;
      }
      v[e] = vh;
      r[e] = rn[a];
    }
    for (int a = 0; a < d; ++a) {
      double sa = 0.0;
      for (int b = 0; b < d; ++b) sa = sa + rn[b] * c.inv[b * d + a];
      se[j * d + a] = sa;
    }
  }
  if (c.scheme == 2) {
    __syncthreads();  // every wave has read the old key
    if (threadIdx.x == 0) {
      R.ml[lane].key[0] = ks[0][0];
      R.ml[lane].key[1] = ks[0][1];
    }
  }
  // the step md_finish is about to complete is L.step + 1: a cadence step
  // sums the unconstrained waves (the same test as md_finish's)
  if (c.trk_skip != nullptr && part == 0 && threadIdx.x == 0)
    c.trk_skip[lane] = (c.wind_every == 1 || (L.step + 1) % c.wind_every == 0) ? 0 : 1;
}

// The second half kick's elementwise part, F = -g B^-T and v += dt/2 F
// (grid: slots_ub x chunks); md_finish_kernel does the rest of the step.
__global__ __launch_bounds__(kDynBlock) void md_kick_kernel(RelaxDev R, MdCfg c, int chunks) {
  const int slot = blockIdx.x / chunks;
  const int part = blockIdx.x % chunks;
  if (slot >= *R.n_active) return;
  const int lane = R.active[slot];
  const size_t off = (size_t)lane * R.stride;
  double* v = R.v + off;
  double* f = R.f + off;
  const double* g = R.g_eval + off;
  const int d = c.d;
  for (int j = part * blockDim.x + threadIdx.x; j < c.n; j += chunks * blockDim.x) {
    for (int a = 0; a < d; ++a) {
      double acc = 0.0;
      for (int b = 0; b < d; ++b) acc = acc + g[j * d + b] * c.inv_t[b * d + a];
      const double fa = -acc;
      f[j * d + a] = fa;
      v[j * d + a] = v[j * d + a] + c.half * fa;
    }
  }
}

// Second half kick with the new forces; andersen: reset one particle's
// velocity (count: when the accumulator reaches 1; bernoulli: with
// probability p) from (key', k_pick, k_draw, k_coin) = split(key, 4); the
// kinetic energy, the drift guard (verlet), the running sums; then what
// the tracked waves need per lane (centre of mass, this step's snapshot and
// trace rows with their scalar fields) for md_waves_kernel.
__global__ __launch_bounds__(kDynBlock) void md_finish_kernel(RelaxDev R, MdCfg c, MdObs o) {
  __shared__ double red[16];
  const int slot = blockIdx.x;
  if (slot >= *R.n_active) return;
  const int lane = R.active[slot];
  const size_t off = (size_t)lane * R.stride;
  double* r = R.r + off;
  double* v = R.v + off;
  MdLane L = R.ml[lane];
  const int d = c.d;
  if (c.scheme == 1) {
    unsigned ks[4][2];
    rng_split(L.key, 4, ks);
    bool fire;
    if (c.andersen_mode == 0) {
      L.reset_acc = L.reset_acc + c.andersen_p;
      fire = L.reset_acc >= 1.0;
      if (fire) L.reset_acc = L.reset_acc - 1.0;
    } else {
      unsigned w[4];
      rng_unif_words(ks[3], w);
      fire = u53(w[0], w[1]) < c.andersen_p;
    }
    if (fire) {
      unsigned w[4];
      rng_unif_words(ks[1], w);
      int pick = (int)floor(u53(w[2], w[3]) * (double)c.n);
      if (pick > c.n - 1) pick = c.n - 1;
      __syncthreads();  // every half kick is written before the row is replaced
      if ((int)threadIdx.x < d) v[(size_t)pick * d + threadIdx.x] = sqrt(L.temperature) * rng_normal(ks[2], threadIdx.x);
    }
    L.key[0] = ks[0][0];
    L.key[1] = ks[0][1];
  }
  __syncthreads();
  double ke = 0.0;
  for (int e = threadIdx.x; e < c.n * d; e += blockDim.x) ke += v[e] * v[e];
  ke = 0.5 * block_sum<kDynBlock>(ke, red);
  const double phi = R.phi_eval[lane];
  int code = ST_RUNNING;
  const double drift = fabs(phi + ke - L.e0);
  if (c.scheme == 0 && isfinite(c.drift_tol)) {
    const double limit = c.drift_tol * 0.5 * (double)d * (double)(c.n - 1) * L.temperature;
    if (!(drift <= limit)) code = ST_DRIFT;
  }
  L.phi = phi;
  L.kinetic = ke;
  L.step += 1;
  L.time = L.time + c.dt;
  L.max_drift = nanmax_d(L.max_drift, drift);
  L.sum_kinetic = L.sum_kinetic + ke;
  L.sum_phi = L.sum_phi + phi;
  L.sum_phi2 = L.sum_phi2 + phi * phi;
  L.n_sum += 1;
  if (c.w > 0) {
    // per lane: the centre of mass (COM-corrected schemes), this step's
    // snapshot row (run.py:477-503: the state after this step) and trace
    // row with their scalar fields; md_waves_kernel does the per-wave part
    if (c.scheme != 0) {
      for (int a = 0; a < d; ++a) {
        double s = 0.0;
        for (int j = threadIdx.x; j < c.n; j += blockDim.x) s += r[(size_t)j * d + a];
        s = block_sum<kDynBlock>(s, red);
        if (threadIdx.x == 0) o.cm[(size_t)lane * 3 + a] = s / (double)c.n;
      }
    }
    int jrec = -1;
    for (int q = 0; q < c.n_rec; ++q)
      if (c.rec_steps[q] == L.step) jrec = q;
    int trow = -1;
    if (c.rho_every > 0 && L.step % c.rho_every == 0 && L.trace_count < c.rho_cap)
      trow = (int)L.trace_count;
    if (jrec >= 0) {
      const size_t row = (size_t)lane * c.n_rec + jrec;
      for (int e = threadIdx.x; e < c.n * d; e += blockDim.x) o.snap_r[row * (size_t)c.n * d + e] = r[e];
      if (threadIdx.x == 0) {
        o.snap_step[row] = L.step;
        o.snap_time[row] = L.time;
        o.snap_phi[row] = L.phi;
        o.snap_kin[row] = ke;
      }
      L.snap_count += 1;
    }
    if (trow >= 0) {
      if (threadIdx.x == 0) o.tr_time[(size_t)lane * c.rho_cap + trow] = L.time;
      L.trace_count += 1;
    }
    if (threadIdx.x == 0) {
      o.rows[2 * lane] = jrec;
      o.rows[2 * lane + 1] = trow;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    R.ml[lane] = L;
    R.status[lane] = code;
  }
}


// The tracked waves of one lane's step, one thread per wave (grid: slots x
// wave tiles), after md_finish_kernel: the constrained rho from the force
// pass, the others from the tracked-wave pass (or summed over the particles
// here when off the lattice) on cadence steps, COM rotation unless verlet,
// winding where S = |rho|^2 / N passes the phase floor at both ends, the S
// sums, and the wave columns of this step's snapshot and trace rows.
__global__ __launch_bounds__(kDynBlock) void md_waves_kernel(RelaxDev R, MdCfg c, MdObs o,
                                                             int n_wtiles) {
  const int slot = blockIdx.x / n_wtiles;
  if (slot >= *R.n_active) return;
  const int lane = R.active[slot];
  const int i = (blockIdx.x % n_wtiles) * kDynBlock + (int)threadIdx.x;
  if (i >= c.w) return;  // no barrier below
  const int d = c.d;
  const double* r = R.r + (size_t)lane * R.stride;
  const long long step = R.ml[lane].step;  // the step md_finish_kernel completed
  const bool corrected = c.scheme != 0;
  const bool wind = c.w > c.nc && (c.wind_every == 1 || step % c.wind_every == 0);
  double* rp = o.rho_prev + (size_t)lane * c.w * 2;
  double* wa = o.wind_abs + (size_t)lane * c.w;
  double* wn = o.wind_net + (size_t)lane * c.w;
  long long* nv = o.n_valid + (size_t)lane * c.w;
  double* ss = o.sum_s + (size_t)lane * c.w;
  const double* k = c.kvec + (size_t)i * d;
  const bool active = i < c.nc || wind;
  double nre = rp[2 * i], nim = rp[2 * i + 1];
  if (i < c.nc) {
    const int code = o.kbox[i];
    const double2 v = o.rbox[(size_t)lane * c.mbox + (code < 0 ? ~code : code)];
    nre = v.x;
    nim = code < 0 ? -v.y : v.y;
  } else if (wind && o.trk_rbox != nullptr) {
    const int code = o.trk_kbox[i - c.nc];
    const double2 v = o.trk_rbox[(size_t)lane * c.trk_mbox + (code < 0 ? ~code : code)];
    nre = v.x;
    nim = code < 0 ? -v.y : v.y;
  } else if (wind) {
    nre = 0.0;
    nim = 0.0;
    for (int j = 0; j < c.n; ++j) {
      double th = 0.0;
      for (int a = 0; a < d; ++a) th = th + k[a] * r[(size_t)j * d + a];
      double sn, cs;
      sincos(th, &sn, &cs);
      nre += cs;
      nim += sn;
    }
  }
  if (active && corrected) {
    double ph = 0.0;
    for (int a = 0; a < d; ++a) ph = ph + k[a] * o.cm[(size_t)lane * 3 + a];
    double sn, cs;
    sincos(ph, &sn, &cs);
    const double cre = nre * cs + nim * sn;
    const double cim = nim * cs - nre * sn;
    nre = cre;
    nim = cim;
  }
  if (active) {
    const double ore = rp[2 * i], oim = rp[2 * i + 1];
    const double delta = atan2(nim * ore - nre * oim, nre * ore + nim * oim);
    const double s_old = (ore * ore + oim * oim) / (double)c.n;
    const double s_new = (nre * nre + nim * nim) / (double)c.n;
    if (s_old > c.phase_floor && s_new > c.phase_floor) {
      wa[i] = wa[i] + fabs(delta);
      wn[i] = wn[i] + delta;
      nv[i] += 1;
    }
    rp[2 * i] = nre;
    rp[2 * i + 1] = nim;
  }
  ss[i] = ss[i] + (rp[2 * i] * rp[2 * i] + rp[2 * i + 1] * rp[2 * i + 1]) / (double)c.n;
  const int jrec = o.rows[2 * lane], trow = o.rows[2 * lane + 1];
  if (jrec >= 0) {
    const size_t row = (size_t)lane * c.n_rec + jrec;
    o.snap_wabs[row * c.w + i] = wa[i];
    o.snap_wnet[row * c.w + i] = wn[i];
    o.snap_rho[(row * c.w + i) * 2] = rp[2 * i];
    o.snap_rho[(row * c.w + i) * 2 + 1] = rp[2 * i + 1];
  }
  if (trow >= 0) {
    const size_t row = (size_t)lane * c.rho_cap + trow;
    o.tr_rho[(row * c.w + i) * 2] = rp[2 * i];
    o.tr_rho[(row * c.w + i) * 2 + 1] = rp[2 * i + 1];
  }
}

// Workgroups per lane of the elementwise MD kernels: one per kDynBlock particles.
int md_chunks(const MdCfg& c) { return (c.n + kDynBlock - 1) / kDynBlock; }

void launch_md_prepare(const RelaxDev& r, const MdCfg& c, int slots_ub, hipStream_t st) {
  if (slots_ub <= 0) return;
  const int chunks = c.scheme == 2 ? 1 : md_chunks(c);
  hipLaunchKernelGGL(md_prepare_kernel, dim3(slots_ub * chunks), dim3(kDynBlock), 0, st, r, c, chunks);
}

void launch_md_finish(const RelaxDev& r, const MdCfg& c, const MdObs& o, int slots_ub,
                      hipStream_t st) {
  if (slots_ub <= 0) return;
  const int chunks = md_chunks(c);
  hipLaunchKernelGGL(md_kick_kernel, dim3(slots_ub * chunks), dim3(kDynBlock), 0, st, r, c, chunks);
  hipLaunchKernelGGL(md_finish_kernel, dim3(slots_ub), dim3(kDynBlock), 0, st, r, c, o);
  if (c.w > 0) {
    const int n_wtiles = (c.w + kDynBlock - 1) / kDynBlock;
    hipLaunchKernelGGL(md_waves_kernel, dim3(slots_ub * n_wtiles), dim3(kDynBlock), 0, st, r, c, o,
                       n_wtiles);
  }
}

// ---------------------------------------------------------------------------
// Dormand-Prince 5(4) gradient flow
// ---------------------------------------------------------------------------
__constant__ double kA[6][5] = {
    {0, 0, 0, 0, 0},
    {1.0 / 5, 0, 0, 0, 0},
    {3.0 / 40, 9.0 / 40, 0, 0, 0},
    {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0},
    {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0},
    {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656},
};
__constant__ double kB5[7] = {35.0 / 384, 0.0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784,
                              11.0 / 84, 0.0};
__constant__ double kB4[7] = {5179.0 / 57600, 0.0, 7571.0 / 16695, 393.0 / 640,
                              -92097.0 / 339200, 187.0 / 2100, 1.0 / 40};

// Velocity of the chosen parametrization from the gradient g (flow.py:230-250):
// writes k = v into kout, returns (gnorm, all-zero flag) via references.
__device__ void flow_velocity(const double* g, double* kout, const FlowCfg& c, int n,
                              double* red, double& gnorm, bool& gzero) {
  const int d = c.d;
  double gg = 0.0, nz = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    for (int a = 0; a < d; ++a) {
      double acc = 0.0, cart = 0.0;
      for (int b = 0; b < d; ++b) {
        acc = acc + (-g[j * d + b]) * c.metric_inv[b * d + a];
        cart = cart + g[j * d + b] * c.inv_t[b * d + a];
      }
      kout[j * d + a] = acc;
      gg += cart * cart;
      nz += (g[j * d + a] != 0.0) ? 1.0 : 0.0;
    }
  }
  gg = block_sum<kDynBlock>(gg, red);
  nz = block_sum<kDynBlock>(nz, red);
  gnorm = sqrt(gg);
  gzero = nz == 0.0;
}

__device__ void flow_rate(double gnorm, bool arclength, double rate[2], double& scale) {
  if (!arclength) {
    rate[0] = 1.0;
    rate[1] = gnorm;
    scale = 1.0;
  } else {
    const double safe = gnorm > 0.0 ? gnorm : 1.0;
    const double inv = gnorm > 0.0 ? 1.0 / safe : 0.0;
    rate[0] = inv;
    rate[1] = 1.0;
    scale = inv;
  }
}

__device__ int flow_status(double f, bool gzero, double gnorm, double t, double dt, long long steps,
                           long long tries, const FlowCfg& c) {
  int code = dt < c.dt_min ? ST_STEP_FLOOR : ST_RUNNING;
  const bool over = (t >= c.t_max) || (steps >= c.max_steps);
  if (over || tries >= 4 * c.max_steps) code = ST_BUDGET;
  if (gzero || gnorm < c.grad_tol) code = ST_GRAD_TOL;
  if (f < c.energy_tol) code = ST_ENERGY_TOL;
  return code;
}

// s_eval = s + dt * sum_{j < i} A[i][j] k_j   (stage i = 1..5)
__device__ void flow_stage_point(const double* s, const double* k, int stride, int nd, int i,
                                 double dt, double* se) {
  for (int e = threadIdx.x; e < nd; e += blockDim.x) {
    double incr = kA[i][0] * k[e];
    for (int j = 1; j < i; ++j) incr = incr + kA[i][j] * k[(size_t)j * stride + e];
    se[e] = s[e] + dt * incr;
  }
}

__device__ void flow_record(const RelaxDev& R, const FlowCfg& c, int lane, FlowLane& L,
                            const double* s, int nd) {
  const size_t row = (size_t)lane * c.record_cap + L.rec_count;
  for (int e = threadIdx.x; e < nd; e += blockDim.x) R.tr_s[row * R.stride + e] = s[e];
  if (threadIdx.x == 0) {
    R.tr_e[row] = L.f;
    R.tr_step[row] = L.t;
    R.tr_len[row] = L.length;
    R.tr_it[row] = L.n_steps;
  }
  L.rec_count += 1;
}

__global__ __launch_bounds__(kDynBlock) void flow_advance_kernel(RelaxDev R, FlowCfg c) {
  __shared__ double red[16];
  const int slot = blockIdx.x;
  if (slot >= *R.n_active) return;
  const int lane = R.active[slot];
  const int nd = R.nd;
  const int n = nd / c.d;
  const size_t off = (size_t)lane * R.stride;
  double* s = R.s + off;
  double* se = R.s_eval + off;
  const double* ge = R.g_eval + off;
  double* k = R.k + (size_t)lane * 7 * R.stride;
  FlowLane L = R.fl[lane];
  const double fe = R.phi_eval[lane];
  int code = ST_RUNNING;
  bool start_attempt = false;

  if (L.mode == MODE_INIT) {
    double gn;
    bool gz;
    flow_velocity(ge, k, c, n, red, gn, gz);
    double rate[2], scale;
    flow_rate(gn, c.arclength, rate, scale);
    if (c.arclength)
      for (int e = threadIdx.x; e < nd; e += blockDim.x) k[e] = k[e] * scale;
    L.f = fe;
    L.gnorm = gn;
    L.gzero = gz;
    L.rate[0] = rate[0];
    L.rate[1] = rate[1];
    L.var = 0.0;
    L.t = 0.0;
    L.length = 0.0;
    L.dt = fmin(c.dt0 != nullptr ? c.dt0[lane] : c.dt_init, c.dt_max);
    L.n_steps = 0;
    L.n_tries = 0;
    L.n_evals = 1;
    L.mode = MODE_EVAL;
    code = flow_status(L.f, gz, gn, 0.0, L.dt, 0, 0, c);
    if (c.record_every > 0 && c.record_cap > 0 && L.rec_count == 0) flow_record(R, c, lane, L, s, nd);
    start_attempt = code == ST_RUNNING;
  } else if (L.stage >= 1 && L.stage <= 5) {
    // velocity at stage point i; next stage point (or the candidate after stage 5)
    const int i = L.stage;
    double* ki = k + (size_t)i * R.stride;
    double gn;
    bool gz;
    flow_velocity(ge, ki, c, n, red, gn, gz);
    double rate[2], scale;
    flow_rate(gn, c.arclength, rate, scale);
    if (c.arclength)
      for (int e = threadIdx.x; e < nd; e += blockDim.x) ki[e] = ki[e] * scale;
    L.srate[i][0] = rate[0];
    L.srate[i][1] = rate[1];
    __syncthreads();
    if (i < 5) {
      flow_stage_point(s, k, R.stride, nd, i + 1, L.dt_try, se);
    } else {
      for (int e = threadIdx.x; e < nd; e += blockDim.x) {
        double acc = kB5[0] * k[e];
        for (int j = 1; j < 6; ++j) acc = acc + kB5[j] * k[(size_t)j * R.stride + e];
        se[e] = s[e] + L.dt_try * acc;
      }
    }
    L.stage = i + 1;
  } else {
    // stage 6: the candidate s_new = s_eval was evaluated
    double* k6 = k + (size_t)6 * R.stride;
    double gn;
    bool gz;
    flow_velocity(ge, k6, c, n, red, gn, gz);
    double rate[2], scale;
    flow_rate(gn, c.arclength, rate, scale);
    if (c.arclength)
      for (int e = threadIdx.x; e < nd; e += blockDim.x) k6[e] = k6[e] * scale;
    __syncthreads();
    L.srate[6][0] = rate[0];
    L.srate[6][1] = rate[1];
    const double dt = L.dt_try;
    double tl[2], etl[2];
    for (int q = 0; q < 2; ++q) {
      double acc = kB5[0] * L.srate[0][q];
      for (int j = 1; j < 6; ++j) acc = acc + kB5[j] * L.srate[j][q];
      tl[q] = dt * acc;
      double ea = (kB5[0] - kB4[0]) * L.srate[0][q];
      for (int j = 1; j < 7; ++j) ea = ea + (kB5[j] - kB4[j]) * L.srate[j][q];
      etl[q] = dt * ea;
    }
    // RMS Cartesian step and error (flow.py:337-343)
    const int d = c.d;
    double mv = 0.0, me = 0.0;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      double st[3], er[3];
      for (int a = 0; a < d; ++a) {
        const int e = j * d + a;
        st[a] = se[e] - s[e];
        double ea = (kB5[0] - kB4[0]) * k[e];
        for (int q = 1; q < 7; ++q) ea = ea + (kB5[q] - kB4[q]) * k[(size_t)q * R.stride + e];
        er[a] = dt * ea;
      }
      for (int a = 0; a < d; ++a) {
        double cs = 0.0, ce = 0.0;
        for (int b = 0; b < d; ++b) {
          cs = cs + st[b] * c.basis[b * d + a];
          ce = ce + er[b] * c.basis[b * d + a];
        }
        mv += cs * cs;
        me += ce * ce;
      }
    }
    mv = block_sum<kDynBlock>(mv, red);
    me = block_sum<kDynBlock>(me, red);
    const double move = sqrt(mv / (double)nd);
    const double tol = c.atol + c.rtol * move;
    double err = sqrt(me / (double)nd) / tol;
    double etl_max = fabs(etl[0]) / (c.atol + c.rtol * fabs(tl[0]));
    etl_max = nanmax_d(etl_max, fabs(etl[1]) / (c.atol + c.rtol * fabs(tl[1])));
    err = nanmax_d(err, etl_max);
    const bool accept = err <= 1.0;
    double factor = 0.9 * pow(err, -1.0 / 5.0);
    factor = fmin(fmax(factor, 0.2), 5.0);
    if (!isfinite(factor) || factor != factor) factor = 0.2;
    const double dt_next = fmin(dt * factor, c.dt_max);
    if (accept) {
      for (int e = threadIdx.x; e < nd; e += blockDim.x) {
        s[e] = se[e];
        k[e] = k6[e];
      }
      L.f = fe;
      L.gnorm = gn;
      L.gzero = gz;
      L.rate[0] = rate[0];
      L.rate[1] = rate[1];
      L.t = L.t + tl[0];
      L.length = L.length + tl[1];
      L.var = L.var + dt;
      L.n_steps += 1;
    }
    L.n_tries += 1;
    L.n_evals += 6;
    L.dt = dt_next;
    code = flow_status(L.f, L.gzero, L.gnorm, L.t, dt_next, L.n_steps, L.n_tries, c);
    if (accept && c.record_every > 0 && (L.n_steps % c.record_every) == 0 &&
        L.rec_count < c.record_cap)
      flow_record(R, c, lane, L, s, nd);
    start_attempt = code == ST_RUNNING;
  }
  if (start_attempt) {
    __syncthreads();
    double dt = L.dt;
    if (!c.arclength) dt = fmin(dt, c.t_max - L.t);
    L.dt_try = dt;
    L.srate[0][0] = L.rate[0];
    L.srate[0][1] = L.rate[1];
    flow_stage_point(s, k, R.stride, nd, 1, dt, se);
    L.stage = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    R.fl[lane] = L;
    R.status[lane] = code;
    // a split-precision lane turns FP64 for good when a step starts below the switch
    if (start_attempt && R.prec != nullptr && c.mixed_switch > 0.0 && L.f <= c.mixed_switch)
      R.prec[lane] = 0;
  }
}

void launch_flow_advance(const RelaxDev& r, const FlowCfg& c, int slots_ub, hipStream_t st) {
  if (slots_ub <= 0) return;
  hipLaunchKernelGGL(flow_advance_kernel, dim3(slots_ub), dim3(kDynBlock), 0, st, r, c);
}

// ---------------------------------------------------------------------------
// ROS2 gradient flow (flow.py:367-398, step control :401-451).  A round
// attempts one step for every lane of a chunk of the active list; the host
// runs the Hessian, the LU factorisation and the solves (rocSOLVER) between
// these kernels.  Slot-indexed buffers: S / rhs [slot][nd], dts [slot].
// Per lane, k holds v(s) (row 0), k1 (1), v at s + h k1 (2), k2 (3) and
// v at the candidate (5).
// ---------------------------------------------------------------------------
constexpr double kRosGamma = 1.0 + 0.70710678118654752440;  // 1 + 1 / sqrt(2)

// S[slot] = s[lane]; dts[slot] = the attempted step min(dt, t_max - t).
__global__ __launch_bounds__(kDynBlock) void ros_gather_kernel(RelaxDev R, FlowCfg c, double* S,
                                                               double* dts) {
  const int slot = blockIdx.x;
  if (slot >= *R.n_active) return;
  const int lane = R.active[slot];
  const int nd = R.nd;
  const double* s = R.s + (size_t)lane * R.stride;
  for (int e = threadIdx.x; e < nd; e += blockDim.x) S[(size_t)slot * nd + e] = s[e];
  if (threadIdx.x == 0) {
    const FlowLane L = R.fl[lane];
    dts[slot] = fmin(L.dt, c.t_max - L.t);
  }
}

// In place on the symmetric Hessian H (row-major, (n d)^2 per slot): the
// column-major W = I - gamma h J with J = -H (I (x) inv(B B^T)), i.e.
// W[(i,a),(j,c)] = delta + gamma h sum_b H[(i,b),(j,c)] Minv[b][a]; entry
// ((i,a), col) of the column-major W sits where H[col][(i,a)] = H[(i,a)][col]
// sits, so each thread rewrites the d entries it read.
__global__ void ros_build_kernel(double* H, const double* dts, const int* n_active, FlowCfg c,
                                 int nd) {
  const int slot = blockIdx.z;
  if (slot >= *n_active) return;
  const int col = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int d = c.d;
  if (i * d >= nd) return;
  double* colp = H + (size_t)slot * nd * nd + (size_t)col * nd + (size_t)i * d;
  const double gh = kRosGamma * dts[slot];
  double h[3];
  for (int b = 0; b < d; ++b) h[b] = colp[b];
  for (int a = 0; a < d; ++a) {
    double acc = 0.0;
    for (int b = 0; b < d; ++b) acc = acc + h[b] * c.metric_inv[b * d + a];
    colp[a] = (col == i * d + a ? 1.0 : 0.0) + gh * acc;
  }
}

// mode 0: rhs = v(s) (k row 0).
// mode 1: k2's right-hand side: k row 2 = v(s + h k1) from g_eval, rhs = that - 2 k1.
// mode 2: the error's right-hand side, rhs = h (k1 + k2) / 2.
__global__ __launch_bounds__(kDynBlock) void ros_rhs_kernel(RelaxDev R, FlowCfg c, double* rhs,
                                                            const double* dts, int mode) {
  __shared__ double red[16];
  const int slot = blockIdx.x;
  if (slot >= *R.n_active) return;
  const int lane = R.active[slot];
  const int nd = R.nd;
  const double* k = R.k + (size_t)lane * 7 * R.stride;
  double* r = rhs + (size_t)slot * nd;
  if (mode == 0) {
    for (int e = threadIdx.x; e < nd; e += blockDim.x) r[e] = k[e];
  } else if (mode == 1) {
    double* k2row = R.k + ((size_t)lane * 7 + 2) * R.stride;
    double gn;
    bool gz;
    flow_velocity(R.g_eval + (size_t)lane * R.stride, k2row, c, nd / c.d, red, gn, gz);
    __syncthreads();
    for (int e = threadIdx.x; e < nd; e += blockDim.x)
      r[e] = k2row[e] - 2.0 * k[(size_t)R.stride + e];
  } else {
    const double h = dts[slot];
    for (int e = threadIdx.x; e < nd; e += blockDim.x)
      r[e] = 0.5 * h * (k[(size_t)R.stride + e] + k[(size_t)3 * R.stride + e]);
  }
}

// mode 0: k1 = rhs, s_eval = s + h k1.  mode 1: k2 = rhs, s_eval = s + h (3/2 k1 + 1/2 k2).
__global__ __launch_bounds__(kDynBlock) void ros_point_kernel(RelaxDev R, const double* rhs,
                                                              const double* dts, int mode) {
  const int slot = blockIdx.x;
  if (slot >= *R.n_active) return;
  const int lane = R.active[slot];
  const int nd = R.nd;
  const size_t off = (size_t)lane * R.stride;
  double* k = R.k + (size_t)lane * 7 * R.stride;
  const double* r = rhs + (size_t)slot * nd;
  const double h = dts[slot];
  for (int e = threadIdx.x; e < nd; e += blockDim.x) {
    if (mode == 0) {
      k[(size_t)R.stride + e] = r[e];
      R.s_eval[off + e] = R.s[off + e] + h * r[e];
    } else {
      k[(size_t)3 * R.stride + e] = r[e];
      R.s_eval[off + e] = R.s[off + e] + h * (1.5 * k[(size_t)R.stride + e] + 0.5 * r[e]);
    }
  }
}

// The candidate s_eval was evaluated; rhs holds W^-1 (h (k1 + k2) / 2).
__global__ __launch_bounds__(kDynBlock) void ros_finish_kernel(RelaxDev R, FlowCfg c,
                                                               const double* rhs,
                                                               const double* dts) {
  __shared__ double red[16];
  const int slot = blockIdx.x;
  if (slot >= *R.n_active) return;
  const int lane = R.active[slot];
  const int nd = R.nd;
  const int d = c.d;
  const int n = nd / d;
  const size_t off = (size_t)lane * R.stride;
  double* s = R.s + off;
  const double* se = R.s_eval + off;
  double* k = R.k + (size_t)lane * 7 * R.stride;
  double* k5 = k + (size_t)5 * R.stride;
  const double* errf = rhs + (size_t)slot * nd;
  FlowLane L = R.fl[lane];
  const double dt = dts[slot];
  double gn;
  bool gz;
  flow_velocity(R.g_eval + off, k5, c, n, red, gn, gz);
  double mv = 0.0, me = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    for (int a = 0; a < d; ++a) {
      double cs = 0.0, ce = 0.0;
      for (int b = 0; b < d; ++b) {
        cs = cs + (se[j * d + b] - s[j * d + b]) * c.basis[b * d + a];
        ce = ce + errf[j * d + b] * c.basis[b * d + a];
      }
      mv += cs * cs;
      me += ce * ce;
    }
  }
  mv = block_sum<kDynBlock>(mv, red);
  me = block_sum<kDynBlock>(me, red);
  const double move = sqrt(mv / (double)nd);
  const double err = sqrt(me / (double)nd) / (c.atol + c.rtol * move);
  const double chord = sqrt(mv);
  const bool accept = err <= 1.0;
  double factor = 0.9 * pow(err, -0.5);
  factor = fmin(fmax(factor, 0.2), 5.0);
  if (!isfinite(factor) || factor != factor) factor = 0.2;
  const double dt_next = fmin(dt * factor, c.dt_max);
  __syncthreads();
  if (accept) {
    for (int e = threadIdx.x; e < nd; e += blockDim.x) {
      s[e] = se[e];
      k[e] = k5[e];
    }
    L.f = R.phi_eval[lane];
    L.gnorm = gn;
    L.gzero = gz;
    L.t = L.t + dt;
    L.length = L.length + chord;
    L.var = L.var + dt;
    L.n_steps += 1;
  }
  L.n_tries += 1;
  L.n_evals += 2;
  L.dt = dt_next;
  const int code = flow_status(L.f, L.gzero, L.gnorm, L.t, dt_next, L.n_steps, L.n_tries, c);
  __syncthreads();
  if (accept && c.record_every > 0 && (L.n_steps % c.record_every) == 0 &&
      L.rec_count < c.record_cap)
    flow_record(R, c, lane, L, s, nd);
  __syncthreads();
  if (threadIdx.x == 0) {
    R.fl[lane] = L;
    R.status[lane] = code;
  }
}

void launch_ros_gather(const RelaxDev& r, const FlowCfg& c, int slots, double* S, double* dts,
                       hipStream_t st) {
  hipLaunchKernelGGL(ros_gather_kernel, dim3(slots), dim3(kDynBlock), 0, st, r, c, S, dts);
}

void launch_ros_build(double* H, const double* dts, const int* n_active, const FlowCfg& c,
                      int nd, int slots, hipStream_t st) {
  const int n = nd / c.d;
  hipLaunchKernelGGL(ros_build_kernel, dim3((n + 255) / 256, nd, slots), dim3(256), 0, st, H,
                     dts, n_active, c, nd);
}

void launch_ros_rhs(const RelaxDev& r, const FlowCfg& c, double* rhs, const double* dts, int mode,
                    int slots, hipStream_t st) {
  hipLaunchKernelGGL(ros_rhs_kernel, dim3(slots), dim3(kDynBlock), 0, st, r, c, rhs, dts, mode);
}

void launch_ros_point(const RelaxDev& r, const double* rhs, const double* dts, int mode, int slots,
                      hipStream_t st) {
  hipLaunchKernelGGL(ros_point_kernel, dim3(slots), dim3(kDynBlock), 0, st, r, rhs, dts, mode);
}

void launch_ros_finish(const RelaxDev& r, const FlowCfg& c, const double* rhs, const double* dts,
                       int slots, hipStream_t st) {
  hipLaunchKernelGGL(ros_finish_kernel, dim3(slots), dim3(kDynBlock), 0, st, r, c, rhs, dts);
}

}  // namespace cco
