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
// rho_kernel : one or two threads per unit (a run of <= kRowSeg rows of one
//              column, seeded by z1^(q kRowSeg) z2^n2 z3^n3 from LDS tables),
//              each taking its share of the particle pairs of a split and
//              accumulating the run in registers; the phase tables of a chunk
//              of particles sit in LDS.  Two threads per unit when all of a
//              lane's units fit one workgroup (more waves per LDS table).
// reduce_kernel : per (lane, tile of box entries), sums the particle splits
//              in a fixed order, forms C = w rho on the box and the tile's
//              part of Phi = sum_k w_k |rho_k|^2 / V (summed over tiles in
//              tile order by the lane's last workgroup), and writes Re/Im rho
//              in k order (padding rows: N, 0).
// force_kernel : one thread per (particle, k part), four waves per 64
//              particles, each wave walking a u = |n2| range of the
//              (|n2|, |n3|) groups whose four sign members share one
//              z2^u z3^v product; C is read with uniform (scalar) loads.  grad_j = (4 pi / V) sum_k w_k n_k Im(rho_k
//              conj e_kj).
#include "cco_internal.h"

namespace cco {

// grid: slots_ub x n_utiles x n_splits; block: H UT threads.  Thread
// (h, u) runs unit u over the particle pairs {2Hi + 2h, 2Hi + 2h + 1} of each
// LDS chunk; with H = 2, half 1 hands its sums to half 0 through LDS.
// Entries per table row (one power, every particle of the chunk): the chunk
// plus room for the pair loop's prefetch past the last pair (read, never
// used), rounded to 1 mod 16 entries so that rows 16 bytes x that length
// apart start 4 banks apart.
__host__ __device__ inline int rho_table_len(int chunk) { return ((chunk + 8 + 15) / 16) * 16 + 1; }

static_assert(kRowSeg == 8, "rho_kernel dispatches row counts 1..8");

// Rows 0 .. NR-1 of one unit for the particle pairs jb, jb + 1 (jb = jfirst,
// jfirst + 2H, ... < nj; an odd nj has zero table entries at index nj).  zr,
// t2, t3 point at the table rows of the unit's powers (z1^(8q), z2^|n2|,
// z3^|n3|); each pair's values are loaded one pair ahead.
template <int D, int NR, int H>
__device__ __forceinline__ void rho_unit_rows(double2 (&acc)[kRowSeg], const double2* zr,
                                              const double2* t2, const double2* t3,
                                              const double2* z1t, int jb, int nj, double sg2,
                                              double sg3) {
  if (jb >= nj) return;
  const double2* pa = zr + jb;
  const double2* pb = t2 + jb;
  const double2* pc = t3 + jb;
  const double2* py = z1t + jb;
  double2 a0 = pa[0], a1 = pa[1];
  double2 b0 = make_double2(1.0, 0.0), b1 = b0, c0 = b0, c1 = b0;
  if (D >= 2) {
    b0 = pb[0];
    b1 = pb[1];
  }
  if (D >= 3) {
    c0 = pc[0];
    c1 = pc[1];
  }
  double2 y0 = py[0], y1 = py[1];
  for (; jb < nj; jb += 2 * H) {
    pa += 2 * H;
    pb += 2 * H;
    pc += 2 * H;
    py += 2 * H;
    const double2 na0 = pa[0], na1 = pa[1];
    double2 nb0 = b0, nb1 = b1, nc0 = c0, nc1 = c1;
    if (D >= 2) {
      nb0 = pb[0];
      nb1 = pb[1];
    }
    if (D >= 3) {
      nc0 = pc[0];
      nc1 = pc[1];
    }
    const double2 ny0 = py[0], ny1 = py[1];
    double2 e0 = a0, e1 = a1;
    if (D >= 2) {
      e0 = cmul(e0, make_double2(b0.x, sg2 * b0.y));
      e1 = cmul(e1, make_double2(b1.x, sg2 * b1.y));
    }
    if (D >= 3) {
      e0 = cmul(e0, make_double2(c0.x, sg3 * c0.y));
      e1 = cmul(e1, make_double2(c1.x, sg3 * c1.y));
    }
    acc[0].x += e0.x + e1.x;
    acc[0].y += e0.y + e1.y;
    if constexpr (NR >= 2) {
      // row 1 by a complex multiply, rows 2.. by the three-term recurrence
      // e_{r+1} = 2 Re(z) e_r - e_{r-1} (z^{r+1} + z^{r-1} = 2 Re(z) z^r for
      // |z| = 1): two FMAs per particle and row instead of four operations
      double2 p0 = e0, p1 = e1;
      e0 = cmul(e0, y0);
      e1 = cmul(e1, y1);
      acc[1].x += e0.x + e1.x;
      acc[1].y += e0.y + e1.y;
      const double w0 = 2.0 * y0.x, w1 = 2.0 * y1.x;
#pragma unroll
      for (int r = 2; r < NR; ++r) {
        const double2 q0 = make_double2(fma(w0, e0.x, -p0.x), fma(w0, e0.y, -p0.y));
        const double2 q1 = make_double2(fma(w1, e1.x, -p1.x), fma(w1, e1.y, -p1.y));
        p0 = e0;
        p1 = e1;
        e0 = q0;
        e1 = q1;
        acc[r].x += e0.x + e1.x;
        acc[r].y += e0.y + e1.y;
      }
    }
    a0 = na0;
    a1 = na1;
    b0 = nb0;
    b1 = nb1;
    c0 = nc0;
    c1 = nc1;
    y0 = ny0;
    y1 = ny1;
  }
}

template <int D, int H>
__global__ __launch_bounds__(H * kRhoUnitsMax) void rho_kernel(Geo g0, Batch b) {
  extern __shared__ double2 lds[];
  const int tiles_per_lane = b.lay != nullptr ? b.tiles_rho : g0.n_utiles * b.n_splits;
  const int slot = blockIdx.x / tiles_per_lane;
  if (slot >= *b.n_active) return;
  const int t = blockIdx.x % tiles_per_lane;
  const int lane = b.active[slot];
  if (b.prec != nullptr && b.prec[lane]) return;  // split-precision lane
  if (b.skip != nullptr && b.skip[lane]) return;
  const Geo g = b.lay != nullptr ? b.geos[lane] : g0;
  if (g.d != D) return;  // uneven batches launch once per dimension
  const LaneIO io = lane_io(g, b, lane);
  const int split = t % io.n_splits;
  const int utile = t / io.n_splits;
  if (utile >= g.n_utiles) return;  // uniform per workgroup, before any barrier
  const int h = H == 1 ? 0 : (int)threadIdx.x / g.UT;
  const int ul = (int)threadIdx.x - h * g.UT;
  const int u = utile * g.UT + ul;
  const bool uvalid = u < g.n_units;
  const int4 un = uvalid ? g.unit[u] : make_int4(0, 0, 0, 0);
  const int nrows = un.y & 0xff;
  const int q = (un.y >> 8) & 0xff;
  const bool work = uvalid && !(un.y & kUnitEmpty);  // empty runs only write zero sums
  const int j0 = split * b.split_len;
  const int j1 = min(io.N, j0 + b.split_len);
  const double* s = io.s;

  // One row count per wave (its longest unit): shorter units also compute the
  // rows past their end, which are never stored, so a wave never loops over
  // the particles once per distinct length.
  int wave_rows = work ? nrows : 0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) wave_rows = max(wave_rows, __shfl_xor(wave_rows, off));
  wave_rows = __builtin_amdgcn_readfirstlane(wave_rows);

  const int CH = g.rho_chunk;
  const int w2 = g.P2 + 1;
  const int w3 = g.P3 + 1;
  // LDS tables, one row of TL entries (the chunk's particles jj) per power:
  // z1, z1^(q kRowSeg) for q < Q, z2^p (p <= P2), z3^p (p <= P3)
  const int TL = rho_table_len(CH);
  double2* Lz1 = lds;
  double2* LZR = lds + TL;
  double2* LT2 = LZR + TL * g.Q;
  double2* LT3 = LT2 + TL * w2;

  double2 acc[kRowSeg];
#pragma unroll
  for (int r = 0; r < kRowSeg; ++r) acc[r] = make_double2(0.0, 0.0);

  const int a2 = abs(un.z);
  const int a3 = abs(un.w);
  const double sg2 = un.z < 0 ? -1.0 : 1.0;
  const double sg3 = un.w < 0 ? -1.0 : 1.0;

  // empty runs sort last: a tile whose first run is empty holds no k-vector
  // and skips the tables (uniform per workgroup), writing zero sums
  const bool tile_empty = (g.unit[utile * g.UT].y & kUnitEmpty) != 0;
  for (int jc = j0; jc < j1 && !tile_empty; jc += CH) {
    const int nj = min(CH, j1 - jc);
    const int nje = (nj + 1) & ~1;  // an odd chunk gets a zero particle (CH is even)
    // Work items per particle: the power tables of axes 1 .. D-1 (serial
    // recurrences, listed first), then the Q row-run seeds z1^(q kRowSeg)
    // (item q = 0 also stores z1).
    const int n_items = D - 1 + g.Q;
    for (int idx = threadIdx.x; idx < nje * n_items; idx += blockDim.x) {
      const int jj = idx % nje;
      const int it = idx / nje;
      const bool real = jj < nj;
      if (it < D - 1) {
        const int ax = it + 1;
        double2* T = (ax == 1 ? LT2 : LT3) + jj;
        const int P = (ax == 1) ? g.P2 : g.P3;
        double2 z = make_double2(0.0, 0.0);
        double2 p = make_double2(0.0, 0.0);
        if (real) {
          double sn, cs;
          sincospi(2.0 * s[(size_t)(jc + jj) * D + ax], &sn, &cs);
          z = make_double2(cs, sn);
          p = make_double2(1.0, 0.0);
        }
        T[0] = p;
        for (int k = 1; k <= P; ++k) {
          p = cmul(p, z);
          T[k * TL] = p;
        }
      } else {
        const int qq = it - (D - 1);
        double2 zq = make_double2(0.0, 0.0);
        if (real) {
          const double sv = s[(size_t)(jc + jj) * D];
          if (qq == 0) {
            double sn, cs;
            sincospi(2.0 * sv, &sn, &cs);
            Lz1[jj] = make_double2(cs, sn);
            zq = make_double2(1.0, 0.0);
          } else {
            double sn, cs;
            sincospi(2.0 * (double)(qq * kRowSeg) * sv, &sn, &cs);
            zq = make_double2(cs, sn);
          }
        } else if (qq == 0) {
          Lz1[jj] = make_double2(0.0, 0.0);
        }
        LZR[qq * TL + jj] = zq;
      }
    }
    __syncthreads();
    if (work) {
      const double2* zr = LZR + q * TL;
      const double2* t2 = LT2 + a2 * TL;
      const double2* t3 = LT3 + a3 * TL;
      const int jb = 2 * h;
      switch (wave_rows) {
#define CCO_RHO_ROWS(NR)                                                                    \
  case NR:                                                                                  \
    rho_unit_rows<D, NR, H>(acc, zr, t2, t3, Lz1, jb, nj, sg2, sg3);              \
    break;
        CCO_RHO_ROWS(1)
        CCO_RHO_ROWS(2)
        CCO_RHO_ROWS(3)
        CCO_RHO_ROWS(4)
        CCO_RHO_ROWS(5)
        CCO_RHO_ROWS(6)
        CCO_RHO_ROWS(7)
        CCO_RHO_ROWS(8)
#undef CCO_RHO_ROWS
        default:
          break;
      }
    }
    __syncthreads();
  }
  // half 1 -> LDS -> half 0 (the tables are no longer read)
  double2* hand = lds;
  if (H == 2 && h == 1 && uvalid) {
#pragma unroll
    for (int r = 0; r < kRowSeg; ++r)
      if (r < nrows) hand[r * g.UT + ul] = acc[r];
  }
  if (H == 2) __syncthreads();
  if (h == 0 && uvalid) {
    double2* out = io.part + (size_t)split * g.Mbox + un.x;
#pragma unroll
    for (int r = 0; r < kRowSeg; ++r) {
      if (r < nrows) {
        if (H == 2) {
          const double2 o = hand[r * g.UT + ul];
          out[r] = make_double2(acc[r].x + o.x, acc[r].y + o.y);
        } else {
          out[r] = acc[r];
        }
      }
    }
  }
}

// grid: slots_ub x n_rtiles.  Workgroup t of a lane owns box entries and k
// rows [t kReduceTile, (t+1) kReduceTile) and writes its part of the energy;
// the force kernel (or phi_kernel) sums the parts in tile order.
__global__ __launch_bounds__(kReduceTile) void reduce_kernel(Geo g0, Batch b) {
  __shared__ double red[kReduceTile / kWave];
  const int tiles_per_lane = b.lay != nullptr ? b.tiles_reduce : g0.n_rtiles;
  const int slot = blockIdx.x / tiles_per_lane;
  if (slot >= *b.n_active) return;
  const int tile = blockIdx.x % tiles_per_lane;
  const int lane = b.active[slot];
  if (b.skip != nullptr && b.skip[lane]) return;  // uniform per workgroup, before any barrier
  if (b.mp_lane && b.prec[lane] != 0) return;      // evaluated whole by mp_lane_kernel
  const Geo g = b.lay != nullptr ? b.geos[lane] : g0;
  if (tile >= g.n_rtiles) return;  // uniform per workgroup, before any barrier
  const LaneIO io = lane_io(g, b, lane);
  const double2* part = io.part;
  double2* cbox = io.cbox;
  const int idx = tile * kReduceTile + (int)threadIdx.x;
  // split-precision lanes of the pair layout: its entries, weights and k map
  // (the column box's Mbox stays the stride)
  const bool pair = g.mp_geo != nullptr && b.prec != nullptr && b.prec[lane] != 0;
  const int nbox = pair ? g.mp_Mbox : g.Mbox;
  const double* wbox = pair ? g.mp_wbox : g.wbox;
  const int* kbox = pair ? g.mp_kbox : g.kbox;
  double acc = 0.0;
  if (idx < nbox) {
    double2 rho = part[idx];
    for (int sp = 1; sp < io.n_splits; ++sp) {
      const double2 v = part[(size_t)sp * g.Mbox + idx];
      rho.x += v.x;
      rho.y += v.y;
    }
    const double w = wbox[idx];
    cbox[idx] = make_double2(w * rho.x, w * rho.y);
    if (io.rbox != nullptr) {  // column-box order: coalesced, gathered by k later
      const int o = pair ? g.mp_p2b[idx] : idx;
      if (o >= 0) io.rbox[o] = rho;
    }
    acc = w * (rho.x * rho.x + rho.y * rho.y);
  }
  if (io.re != nullptr && idx < g.M_pad) {
    double re = (double)io.N;
    double im = 0.0;
    if (idx < g.M) {
      const int code = kbox[idx];
      const int i = code < 0 ? ~code : code;
      double2 rho = part[i];
      for (int sp = 1; sp < io.n_splits; ++sp) {
        const double2 v = part[(size_t)sp * g.Mbox + i];
        rho.x += v.x;
        rho.y += v.y;
      }
      re = rho.x;
      im = code < 0 ? -rho.y : rho.y;
    }
    io.re[idx] = re;
    io.im[idx] = im;
  }
  const double tot = block_sum<kReduceTile>(acc, red);
  if (threadIdx.x == 0) io.phi_part[tile] = tot;
}

// Phi of a lane: its reduce tiles' partials summed in tile order.
__device__ __forceinline__ double phi_from_parts(const Geo& g, const LaneIO& io) {
  double phi = 0.0;
  for (int t = 0; t < g.n_rtiles; ++t) phi += io.phi_part[t];
  return phi / g.volume;
}

// Energy-only evaluations: one thread per lane finishes Phi.
__global__ void phi_kernel(Geo g0, Batch b, int slots_ub) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= slots_ub || slot >= *b.n_active) return;
  const int lane = b.active[slot];
  const Geo g = b.lay != nullptr ? b.geos[lane] : g0;
  b.phi[lane] = phi_from_parts(g, lane_io(g, b, lane));
}

// One column (member) of a force group: with u_r = Im(conj(C_r) e_r) along the
// rows r = 0 .. L-1, g1 gains sum_r r u_r = L P - Q (P = sum_r u_r, Q = sum_r of
// the running P), and g2, g3 gain n2 P, n3 P.
template <int D>
__device__ __forceinline__ void force_member(const int4 m, const double2 B, const double2 z1,
                                             const double2* __restrict__ C, double& g1,
                                             double& g2, double& g3) {
  if (m.y == 0) return;
  // rows advance by the three-term recurrence e_{r+1} = 2 Re(z) e_r - e_{r-1}
  // (|z| = 1): two FMAs per row instead of a complex multiply.  The pair
  // (e_{r-1}, e_r) restarts from e_r conj(z) every 8 rows, so rounding errors
  // grow linearly along a long column (as for a chain of complex multiplies),
  // not quadratically as an unbroken recurrence near z = +-1 would let them.
  double2 e = B;
  const double w = 2.0 * z1.x;
  double P = 0.0;
  double Q = 0.0;
  const double2* Cm = C + m.x;
  for (int r0 = 0; r0 < m.y; r0 += 8) {
    double2 ep = cmul(e, cconj(z1));
    const int nr = min(8, m.y - r0);
#pragma unroll 4
    for (int r = 0; r < nr; ++r) {
      const double2 c = Cm[r0 + r];
      P = fma(c.y, e.x, P);  // P += Im(C conj e), folded into two FMAs
      P = fma(-c.x, e.y, P);
      Q += P;
      const double2 en = make_double2(fma(w, e.x, -ep.x), fma(w, e.y, -ep.y));
      ep = e;
      e = en;
    }
  }
  g1 += (double)m.y * P - Q;
  if (D >= 2) g2 += (double)m.z * P;
  if (D >= 3) g3 += (double)m.w * P;
}

// The force scan of one workgroup: particles [64 tile, +64) of `lane`, k part
// `part` walking the u range ur; the parts' partial gradients are summed in
// part order through LDS.
template <int D>
__device__ __forceinline__ void force_tile(const Geo& g, const LaneIO& io, const Batch& b,
                                           int lane, int tile, int part, int2 ur,
                                           double (*red)[3][kForceTile]) {
  const int nparts = (int)blockDim.x / kForceTile;
  const int jl = (int)threadIdx.x % kForceTile;
  const int j = tile * kForceTile + jl;
  const bool valid = j < io.N;
  const double* sp = io.s + (size_t)(valid ? j : 0) * D;
  double2 z[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) z[a] = make_double2(1.0, 0.0);
  double sv[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    sv[a] = sp[a];
    double sn, cs;
    sincospi(2.0 * sv[a], &sn, &cs);
    z[a] = make_double2(cs, sn);
  }
  const double2* __restrict__ C = io.cbox;
  double g1 = 0.0, g2 = 0.0, g3 = 0.0;
  double2 zu = make_double2(1.0, 0.0);
  if (D >= 2 && ur.x > 0) {
    double sn, cs;
    sincospi(2.0 * (double)ur.x * sv[1], &sn, &cs);
    zu = make_double2(cs, sn);
  }
  for (int u = ur.x; u < ur.y; ++u) {
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
  if (nparts > 1) {
    if (part > 0) {
      red[part - 1][0][jl] = g1;
      red[part - 1][1][jl] = g2;
      red[part - 1][2][jl] = g3;
    }
    __syncthreads();
    if (part > 0) return;
    for (int p = 1; p < nparts; ++p) {
      g1 += red[p - 1][0][jl];
      g2 += red[p - 1][1][jl];
      g3 += red[p - 1][2][jl];
    }
  }
  if (tile == 0 && threadIdx.x == 0) b.phi[lane] = phi_from_parts(g, io);
  if (valid && io.grad != nullptr) {
    double* out = io.grad + (size_t)j * D;
    out[0] = g.pref * g1;
    if (D >= 2) out[1] = g.pref * g2;
    if (D >= 3) out[2] = g.pref * g3;
  }
}

// grid: slots_ub x n_ftiles; block: kForceTile particles x k parts (g.fparts
// for a uniform batch, kForceParts for an uneven one).  UNEVEN reads each
// lane's geometry from b.geos; the uniform instance reads the argument (the
// group tables then stay on the scalar unit).
template <int D, bool UNEVEN>
__global__ __launch_bounds__(kForceTile * kForceParts) void force_kernel(Geo g0, Batch b) {
  __shared__ double red[kForceParts > 1 ? kForceParts - 1 : 1][3][kForceTile];
  const int tiles_per_lane = UNEVEN ? b.tiles_force : b.n_ftiles;
  const int slot = blockIdx.x / tiles_per_lane;
  if (slot >= *b.n_active) return;
  const int tile = blockIdx.x % tiles_per_lane;
  const int lane = b.active[slot];
  if (b.prec != nullptr && b.prec[lane]) return;  // split-precision lane
  // wave-uniform: keeps the k walk (and the loads of C) on the scalar unit
  const int part = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kForceTile);
  if constexpr (UNEVEN) {
    const Geo* gp = b.geos + lane;
    if (gp->d != D) return;  // uneven batches launch once per dimension
    const Geo g = *gp;
    const LaneIO io = lane_io<2>(g, b, lane);
    if (tile >= io.n_ftiles) return;  // uniform per workgroup, before any barrier
    force_tile<D>(g, io, b, lane, tile, part, gp->fpart[part], red);
  } else {
    const LaneIO io = lane_io<1>(g0, b, lane);
    force_tile<D>(g0, io, b, lane, tile, part, g0.fpart[part], red);
  }
}

size_t rho_lds_bytes(const Geo& g) {
  const size_t tables = (size_t)rho_table_len(g.rho_chunk) * (1 + g.Q + (g.P2 + 1) + (g.P3 + 1));
  const size_t hand = g.rho_halves == 2 ? (size_t)kRowSeg * g.UT : 0;  // half-1 sums
  return sizeof(double2) * (tables > hand ? tables : hand);
}

// Launch one stage (0 = rho, 1 = reduce, 2 = force) of an evaluation.
template <int D>
static void launch_stage(int stage, const Geo& g, const Batch& b, int slots_ub, hipStream_t st) {
  if (b.lay != nullptr) {  // uneven batch: per-lane geometry, launch-wide maxima
    if (stage == 0)
      hipLaunchKernelGGL((rho_kernel<D, 2>), dim3(slots_ub * b.tiles_rho), dim3(2 * kWave),
                         b.rho_lds, st, g, b);
    else if (stage == 1)
      hipLaunchKernelGGL(reduce_kernel, dim3(slots_ub * b.tiles_reduce), dim3(kReduceTile), 0, st,
                         g, b);
    else
      hipLaunchKernelGGL((force_kernel<D, true>), dim3(slots_ub * b.tiles_force),
                         dim3(kForceTile * kForceParts), 0, st, g, b);
    return;
  }
  if (stage == 0) {
    const int rho_tiles = g.n_utiles * b.n_splits;
    if (g.rho_halves == 2)
      hipLaunchKernelGGL((rho_kernel<D, 2>), dim3(slots_ub * rho_tiles), dim3(2 * g.UT),
                         rho_lds_bytes(g), st, g, b);
    else
      hipLaunchKernelGGL((rho_kernel<D, 1>), dim3(slots_ub * rho_tiles), dim3(g.UT),
                         rho_lds_bytes(g), st, g, b);
  } else if (stage == 1) {
    hipLaunchKernelGGL(reduce_kernel, dim3(slots_ub * g.n_rtiles), dim3(kReduceTile), 0, st, g, b);
  } else {
    hipLaunchKernelGGL((force_kernel<D, false>), dim3(slots_ub * b.n_ftiles),
                       dim3(kForceTile * g.fparts), 0, st, g, b);
  }
}

void launch_energy_stage(int stage, const Geo& g, const Batch& b, int slots_ub, hipStream_t st) {
  if (slots_ub <= 0) return;
  if (b.lay != nullptr) {  // uneven: reduce once, rho and forces once per dimension present
    if (stage == 1) {
      Batch bs = b;
      bs.mp_lane = mp_lane_used(g, b) ? 1 : 0;  // the split-precision lanes are evaluated whole
      launch_stage<1>(stage, g, bs, slots_ub, st);
      return;
    }
    if (b.dmask & 2) launch_stage<1>(stage, g, b, slots_ub, st);
    if (b.dmask & 4) launch_stage<2>(stage, g, b, slots_ub, st);
    if (b.dmask & 8) launch_stage<3>(stage, g, b, slots_ub, st);
    launch_energy_mixed(stage, g, b, slots_ub, st);
    return;
  }
  Batch bs = b;
  bs.mp_lane = stage == 1 && mp_lane_used(g, b) ? 1 : 0;  // the reduce skips those lanes
  if (g.d == 1) launch_stage<1>(stage, g, bs, slots_ub, st);
  if (g.d == 2) launch_stage<2>(stage, g, bs, slots_ub, st);
  if (g.d == 3) launch_stage<3>(stage, g, bs, slots_ub, st);
  if (stage != 1) launch_energy_mixed(stage, g, b, slots_ub, st);
}

void launch_rho_reduce(const Geo& g, const Batch& b, int slots_ub, hipStream_t st) {
  if (slots_ub <= 0) return;
  for (int stage = 0; stage < 2; ++stage) {
    if (g.d == 1) launch_stage<1>(stage, g, b, slots_ub, st);
    if (g.d == 2) launch_stage<2>(stage, g, b, slots_ub, st);
    if (g.d == 3) launch_stage<3>(stage, g, b, slots_ub, st);
  }
}

void launch_energy(const Geo& g, const Batch& b, int slots_ub, bool with_grad,
                   hipStream_t stream) {
  launch_energy_stage(0, g, b, slots_ub, stream);
  launch_energy_stage(1, g, b, slots_ub, stream);
  if (with_grad) {
    launch_energy_stage(2, g, b, slots_ub, stream);
  } else if (slots_ub > 0) {
    hipLaunchKernelGGL(phi_kernel, dim3((slots_ub + 255) / 256), dim3(256), 0, stream, g, b,
                       slots_ub);
  }
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
