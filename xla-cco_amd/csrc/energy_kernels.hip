// FP64 stealthy energy, collective density and forces for gfx950.
//
// Replaces the XLA fusions of src/xla_cco/energy.py:
//   _density_block (:45-50)        -> rho_kernel + reduce_kernel
//   value_grad_density (:206-225)  -> rho_kernel, reduce_kernel, force_kernel
//   _gradient scan (:301-324)      -> force_kernel
//
// The k-set is integer valued, so exp(2 pi i n.s_j) = z1^n1 z2^n2 z3^n3 with
// z_a = exp(2 pi i s_a) computed once per (j, a) by sincospi.  Phases along
// a column (fixed n2, n3) follow by complex recurrence e <- e z1, so each
// (k, j) pair costs one complex multiply and one complex add (rho) or one
// complex multiply and three FMAs (forces); no trig runs per pair.
//
// rho_kernel : one thread per unit (a run of <= kRowSeg rows of one column,
//              seeded by z1^(q kRowSeg) z2^n2 z3^n3 from LDS tables); loops
//              over a split of particles, accumulating the run in registers;
//              the phase tables of a chunk of particles sit in LDS.
// reduce_kernel : per lane, sums the particle splits in a fixed order,
//              forms Phi = sum_k w_k |rho_k|^2 / V, C = w rho on the box,
//              and writes Re/Im rho in k order (padding rows: N, 0).
// force_kernel : one thread per particle, one wave per 64 particles; the
//              columns are walked by (|n2|, |n3|) groups whose four sign
//              members share one z2^u z3^v product; C is read with uniform
//              (scalar) loads.  grad_j = (4 pi / V) sum_k w_k n_k Im(rho_k
//              conj e_kj).
#include "cco_internal.h"

namespace cco {

template <int D>
__global__ __launch_bounds__(256) void rho_kernel(Geo g, Batch b) {
  extern __shared__ double2 lds[];
  const int tiles_per_lane = g.n_utiles * b.n_splits;
  const int slot = blockIdx.x / tiles_per_lane;
  if (slot >= *b.n_active) return;
  const int t = blockIdx.x % tiles_per_lane;
  const int split = t % b.n_splits;
  const int utile = t / b.n_splits;
  const int lane = b.active[slot];
  const int u = utile * g.UT + (int)threadIdx.x;
  const bool uvalid = u < g.n_units;
  const int4 un = uvalid ? g.unit[u] : make_int4(0, 0, 0, 0);
  const int nrows = un.y & 0xff;
  const int q = un.y >> 8;
  const int j0 = split * b.split_len;
  const int j1 = min(b.N, j0 + b.split_len);
  const double* s = b.s + (size_t)lane * b.lane_stride;

  const int CH = g.rho_chunk;
  const int w2 = g.P2 + 1;
  const int w3 = g.P3 + 1;
  // LDS per chunk particle jj: z1, z1^(q kRowSeg) for q < Q, z2^p (p <= P2), z3^p (p <= P3)
  double2* Lz1 = lds;
  double2* LZR = lds + CH;
  double2* LT2 = LZR + CH * g.Q;
  double2* LT3 = LT2 + CH * w2;

  double2 acc[kRowSeg];
#pragma unroll
  for (int r = 0; r < kRowSeg; ++r) acc[r] = make_double2(0.0, 0.0);

  const int a2 = abs(un.z);
  const int a3 = abs(un.w);
  const double sg2 = un.z < 0 ? -1.0 : 1.0;
  const double sg3 = un.w < 0 ? -1.0 : 1.0;

  for (int jc = j0; jc < j1; jc += CH) {
    const int nj = min(CH, j1 - jc);
    for (int idx = threadIdx.x; idx < nj * D; idx += blockDim.x) {
      const int jj = idx % nj;
      const int ax = idx / nj;
      const double sv = s[(size_t)(jc + jj) * D + ax];
      double sn, cs;
      sincospi(2.0 * sv, &sn, &cs);
      const double2 z = make_double2(cs, sn);
      if (ax == 0) {
        Lz1[jj] = z;
        LZR[jj * g.Q] = make_double2(1.0, 0.0);
        for (int qq = 1; qq < g.Q; ++qq) {
          double s0, c0;
          sincospi(2.0 * (double)(qq * kRowSeg) * sv, &s0, &c0);
          LZR[jj * g.Q + qq] = make_double2(c0, s0);
        }
      } else {
        double2* T = (ax == 1) ? (LT2 + jj * w2) : (LT3 + jj * w3);
        const int P = (ax == 1) ? g.P2 : g.P3;
        double2 p = make_double2(1.0, 0.0);
        T[0] = p;
        for (int k = 1; k <= P; ++k) {
          p = cmul(p, z);
          T[k] = p;
        }
      }
    }
    __syncthreads();
    if (uvalid) {
      int jj = 0;
      // two particles per pass: two independent recurrence chains
      for (; jj + 1 < nj; jj += 2) {
        double2 e0 = LZR[jj * g.Q + q];
        double2 e1 = LZR[(jj + 1) * g.Q + q];
        if (D >= 2) {
          const double2 t0 = LT2[jj * w2 + a2];
          const double2 t1 = LT2[(jj + 1) * w2 + a2];
          e0 = cmul(e0, make_double2(t0.x, sg2 * t0.y));
          e1 = cmul(e1, make_double2(t1.x, sg2 * t1.y));
        }
        if (D >= 3) {
          const double2 t0 = LT3[jj * w3 + a3];
          const double2 t1 = LT3[(jj + 1) * w3 + a3];
          e0 = cmul(e0, make_double2(t0.x, sg3 * t0.y));
          e1 = cmul(e1, make_double2(t1.x, sg3 * t1.y));
        }
        const double2 z0 = Lz1[jj];
        const double2 z1 = Lz1[jj + 1];
#pragma unroll
        for (int r = 0; r < kRowSeg; ++r) {
          if (r < nrows) {
            acc[r].x += e0.x + e1.x;
            acc[r].y += e0.y + e1.y;
            e0 = cmul(e0, z0);
            e1 = cmul(e1, z1);
          }
        }
      }
      if (jj < nj) {
        double2 e0 = LZR[jj * g.Q + q];
        if (D >= 2) {
          const double2 t0 = LT2[jj * w2 + a2];
          e0 = cmul(e0, make_double2(t0.x, sg2 * t0.y));
        }
        if (D >= 3) {
          const double2 t0 = LT3[jj * w3 + a3];
          e0 = cmul(e0, make_double2(t0.x, sg3 * t0.y));
        }
        const double2 z0 = Lz1[jj];
#pragma unroll
        for (int r = 0; r < kRowSeg; ++r) {
          if (r < nrows) {
            acc[r].x += e0.x;
            acc[r].y += e0.y;
            e0 = cmul(e0, z0);
          }
        }
      }
    }
    __syncthreads();
  }
  if (uvalid) {
    double2* out = b.part + ((size_t)lane * b.n_splits + split) * g.Mbox + un.x;
#pragma unroll
    for (int r = 0; r < kRowSeg; ++r) {
      if (r < nrows) out[r] = acc[r];
    }
  }
}

__global__ __launch_bounds__(256) void reduce_kernel(Geo g, Batch b) {
  __shared__ double red[16];
  const int slot = blockIdx.x;
  if (slot >= *b.n_active) return;
  const int lane = b.active[slot];
  const double2* part = b.part + (size_t)lane * b.n_splits * g.Mbox;
  double2* cbox = b.cbox + (size_t)lane * g.Mbox;
  double acc = 0.0;
  for (int i = threadIdx.x; i < g.Mbox; i += blockDim.x) {
    double2 rho = part[i];
    for (int sp = 1; sp < b.n_splits; ++sp) {
      const double2 v = part[(size_t)sp * g.Mbox + i];
      rho.x += v.x;
      rho.y += v.y;
    }
    const double w = g.wbox[i];
    cbox[i] = make_double2(w * rho.x, w * rho.y);
    acc += w * (rho.x * rho.x + rho.y * rho.y);
  }
  if (b.re != nullptr) {
    for (int k = threadIdx.x; k < g.M; k += blockDim.x) {
      const int code = g.kbox[k];
      const int i = code < 0 ? ~code : code;
      double2 rho = part[i];
      for (int sp = 1; sp < b.n_splits; ++sp) {
        const double2 v = part[(size_t)sp * g.Mbox + i];
        rho.x += v.x;
        rho.y += v.y;
      }
      b.re[(size_t)lane * g.M_pad + k] = rho.x;
      b.im[(size_t)lane * g.M_pad + k] = code < 0 ? -rho.y : rho.y;
    }
    for (int k = g.M + threadIdx.x; k < g.M_pad; k += blockDim.x) {
      b.re[(size_t)lane * g.M_pad + k] = (double)b.N;
      b.im[(size_t)lane * g.M_pad + k] = 0.0;
    }
  }
  const double tot = block_sum<256>(acc, red);
  if (threadIdx.x == 0) b.phi[lane] = tot / g.volume;
}

template <int D>
__device__ __forceinline__ void force_member(const int4 m, const double2 B, const double2 z1,
                                             const double2* __restrict__ C, double& g1,
                                             double& g2, double& g3) {
  if (m.y == 0) return;
  double2 e = B;
  double us = 0.0;
  double rr = 0.0;
  const double2* Cm = C + m.x;
#pragma unroll 4
  for (int r = 0; r < m.y; ++r) {
    const double2 c = Cm[r];
    const double uu = c.y * e.x - c.x * e.y;
    g1 += rr * uu;
    us += uu;
    rr += 1.0;
    e = cmul(e, z1);
  }
  if (D >= 2) g2 += (double)m.z * us;
  if (D >= 3) g3 += (double)m.w * us;
}

template <int D>
__global__ __launch_bounds__(kForceTile) void force_kernel(Geo g, Batch b) {
  const int slot = blockIdx.x / b.n_ftiles;
  if (slot >= *b.n_active) return;
  const int tile = blockIdx.x % b.n_ftiles;
  const int lane = b.active[slot];
  const int j = tile * kForceTile + (int)threadIdx.x;
  const bool valid = j < b.N;
  const double* sp = b.s + (size_t)lane * b.lane_stride + (size_t)(valid ? j : 0) * D;
  double2 z[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) z[a] = make_double2(1.0, 0.0);
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double sn, cs;
    sincospi(2.0 * sp[a], &sn, &cs);
    z[a] = make_double2(cs, sn);
  }
  const double2* __restrict__ C = b.cbox + (size_t)lane * g.Mbox;
  double g1 = 0.0, g2 = 0.0, g3 = 0.0;
  double2 zu = make_double2(1.0, 0.0);
  for (int u = 0; u < g.n_u; ++u) {
    const int2 us = g.ustart[u];
    double2 zv = make_double2(1.0, 0.0);
    for (int v = 0; v < us.y; ++v) {
      const int4* mem = g.gmem + 4 * (size_t)(us.x + v);
      const double2 B1 = cmul(zu, zv);
      const double2 B2 = cmul(zu, cconj(zv));
      force_member<D>(mem[0], B1, z[0], C, g1, g2, g3);
      force_member<D>(mem[1], B2, z[0], C, g1, g2, g3);
      force_member<D>(mem[2], cconj(B1), z[0], C, g1, g2, g3);
      force_member<D>(mem[3], cconj(B2), z[0], C, g1, g2, g3);
      zv = cmul(zv, z[2]);
    }
    zu = cmul(zu, z[1]);
  }
  if (valid && b.grad != nullptr) {
    double* out = b.grad + (size_t)lane * b.lane_stride + (size_t)j * D;
    out[0] = g.pref * g1;
    if (D >= 2) out[1] = g.pref * g2;
    if (D >= 3) out[2] = g.pref * g3;
  }
}

size_t rho_lds_bytes(const Geo& g) {
  return sizeof(double2) * (size_t)g.rho_chunk * (1 + g.Q + (g.P2 + 1) + (g.P3 + 1));
}

// Launch one stage (0 = rho, 1 = reduce, 2 = force) of an evaluation.
template <int D>
static void launch_stage(int stage, const Geo& g, const Batch& b, int slots_ub, hipStream_t st) {
  if (stage == 0) {
    const int rho_tiles = g.n_utiles * b.n_splits;
    hipLaunchKernelGGL(rho_kernel<D>, dim3(slots_ub * rho_tiles), dim3(g.UT), rho_lds_bytes(g),
                       st, g, b);
  } else if (stage == 1) {
    hipLaunchKernelGGL(reduce_kernel, dim3(slots_ub), dim3(256), 0, st, g, b);
  } else {
    hipLaunchKernelGGL(force_kernel<D>, dim3(slots_ub * b.n_ftiles), dim3(kForceTile), 0, st, g,
                       b);
  }
}

void launch_energy_stage(int stage, const Geo& g, const Batch& b, int slots_ub, hipStream_t st) {
  if (slots_ub <= 0) return;
  if (g.d == 1) launch_stage<1>(stage, g, b, slots_ub, st);
  if (g.d == 2) launch_stage<2>(stage, g, b, slots_ub, st);
  if (g.d == 3) launch_stage<3>(stage, g, b, slots_ub, st);
}

void launch_energy(const Geo& g, const Batch& b, int slots_ub, bool with_grad,
                   hipStream_t stream) {
  launch_energy_stage(0, g, b, slots_ub, stream);
  launch_energy_stage(1, g, b, slots_ub, stream);
  if (with_grad) launch_energy_stage(2, g, b, slots_ub, stream);
}

__global__ __launch_bounds__(256) void copy_kernel(const double4* __restrict__ src,
                                                   double4* __restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

void launch_copy(const void* src, void* dst, size_t bytes, hipStream_t st) {
  const size_t n = bytes / sizeof(double4);
  hipLaunchKernelGGL(copy_kernel, dim3(2048), dim3(256), 0, st, (const double4*)src, (double4*)dst,
                     n);
}

}  // namespace cco
