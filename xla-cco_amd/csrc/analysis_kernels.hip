// Ensemble analysis on the GPU: isotropic structure-factor moments and
// minimum-image pair statistics.
//
//   sk_kernel / sk_sum_kernel : S(k) = |rho(k)|^2 / N on the C++
//       IsotropicStructureFactor k-set (primitive directions n_c, each with
//       its multiples m n_c, m = 1..L_c), summed over configurations as
//       sum S and sum S^2 per k (src/xla_cco/analysis/structure.py:111-173).
//       One thread per direction and 32 consecutive multiples; the phase of
//       multiple m follows by recurrence from exp(2 pi i m0 t), t = n_c . s_j.
//   pair_kernel : for every particle the minimum-image distances to all
//       others over the 3^d neighbouring images (analysis/pairs.py:42-209):
//       pair-distance histogram (np.histogram bin semantics, each unordered
//       pair once) with integer atomics, and the nearest-neighbour distance.
#include "cco_internal.h"

namespace cco {

constexpr int kSkSeg = 32;

// grid: (config slot) x (direction tile) x (segment); block = direction tile
__global__ __launch_bounds__(256) void sk_kernel(const double* __restrict__ s, int N, int d,
                                                 const int4* __restrict__ dirs,  // {n1, n2, n3, L}
                                                 const int* __restrict__ koff, int D, int Mk,
                                                 int n_tiles, int n_segs,
                                                 double* __restrict__ out) {
  const int cfg = blockIdx.x / (n_tiles * n_segs);
  const int rem = blockIdx.x % (n_tiles * n_segs);
  const int tile = rem / n_segs;
  const int seg = rem % n_segs;
  const int c = tile * blockDim.x + threadIdx.x;
  if (c >= D) return;
  const int4 dir = dirs[c];
  const int L = dir.w;
  const int m0 = seg * kSkSeg + 1;  // first multiple of this segment
  if (m0 > L) return;
  const int nrow = min(kSkSeg, L - m0 + 1);
  const double* sc = s + (size_t)cfg * N * d;
  double2 acc[kSkSeg];
#pragma unroll
  for (int r = 0; r < kSkSeg; ++r) acc[r] = make_double2(0.0, 0.0);
  const double n1 = dir.x, n2 = dir.y, n3 = dir.z;
  for (int j = 0; j < N; ++j) {
    double t = n1 * sc[(size_t)j * d];
    if (d >= 2) t = t + n2 * sc[(size_t)j * d + 1];
    if (d >= 3) t = t + n3 * sc[(size_t)j * d + 2];
    double sz, cz, se, ce;
    sincospi(2.0 * t, &sz, &cz);
    sincospi(2.0 * ((double)m0 * t), &se, &ce);
    const double2 z = make_double2(cz, sz);
    double2 e = make_double2(ce, se);
#pragma unroll
    for (int r = 0; r < kSkSeg; ++r) {
      if (r < nrow) {
        acc[r].x += e.x;
        acc[r].y += e.y;
        e = cmul(e, z);
      }
    }
  }
  const double invn = 1.0 / (double)N;
  double* o = out + (size_t)cfg * Mk + koff[c] + (m0 - 1);
#pragma unroll
  for (int r = 0; r < kSkSeg; ++r)
    if (r < nrow) o[r] = (acc[r].x * acc[r].x + acc[r].y * acc[r].y) * invn;
}

// tot[k] += sum over the chunk's configurations of S, tot2[k] of S^2 (fixed order)
__global__ void sk_sum_kernel(const double* __restrict__ sk, int n_cfg, int Mk,
                              double* __restrict__ tot, double* __restrict__ tot2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= Mk) return;
  double a = tot[k], b = tot2[k];
  for (int c = 0; c < n_cfg; ++c) {
    const double v = sk[(size_t)c * Mk + k];
    a += v;
    b += v * v;
  }
  tot[k] = a;
  tot2[k] = b;
}

// grid: (config) x (particle block of 256); block = 256
__global__ __launch_bounds__(256) void pair_kernel(const double* __restrict__ s, int N, int d,
                                                   const double* __restrict__ basis,  // d x d
                                                   const double* __restrict__ edges, int n_edges,
                                                   unsigned long long* __restrict__ counts,
                                                   double* __restrict__ nn, int n_blocks) {
  extern __shared__ unsigned int hist[];
  __shared__ double E[1024];
  const int cfg = blockIdx.x / n_blocks;
  const int blk = blockIdx.x % n_blocks;
  const int n_bins = n_edges - 1;
  for (int b = threadIdx.x; b < n_bins; b += blockDim.x) hist[b] = 0u;
  for (int b = threadIdx.x; b < n_edges && b < 1024; b += blockDim.x) E[b] = edges[b];
  __syncthreads();
  const int i = blk * blockDim.x + threadIdx.x;
  const double* sc = s + (size_t)cfg * N * d;
  if (i < N) {
    double si[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < d; ++a) si[a] = sc[(size_t)i * d + a];
    double B[9];
    for (int q = 0; q < d * d; ++q) B[q] = basis[q];
    int nshift = 1;
    for (int a = 0; a < d; ++a) nshift *= 3;
    double nn2 = __longlong_as_double(0x7ff0000000000000LL);
    for (int j = 0; j < N; ++j) {
      double ds[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < d; ++a) {
        const double v = sc[(size_t)j * d + a] - si[a];
        ds[a] = v - rint(v);
      }
      double best = __longlong_as_double(0x7ff0000000000000LL);
      for (int sh = 0; sh < nshift; ++sh) {
        int code = sh;
        double vec[3] = {0.0, 0.0, 0.0};
        bool zero_shift = true;
        double sv[3];
        for (int a = 0; a < d; ++a) {
          const int q = code % 3;
          code /= 3;
          sv[a] = (double)(q - 1);
          if (q != 1) zero_shift = false;
        }
        for (int a = 0; a < d; ++a) {
          double acc = 0.0;
          for (int b = 0; b < d; ++b) acc = acc + (ds[b] + sv[b]) * B[b * d + a];
          vec[a] = acc;
        }
        double r2 = 0.0;
        for (int a = 0; a < d; ++a) r2 = r2 + vec[a] * vec[a];
        if (zero_shift && j == i) continue;  // the particle itself
        best = fmin(best, r2);
      }
      nn2 = fmin(nn2, best);
      if (j > i) {
        const double dist = sqrt(best);
        if (isfinite(dist) && dist >= E[0] && dist <= E[n_bins]) {
          // np.histogram: bins [e_k, e_k+1), the last one closed
          int lo = 0, hi = n_bins;  // find the last edge <= dist
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (E[mid] <= dist) lo = mid; else hi = mid;
          }
          int bin = lo;
          if (dist == E[n_bins]) bin = n_bins - 1;
          atomicAdd(&hist[bin], 1u);
        }
      }
    }
    if (nn != nullptr) nn[(size_t)cfg * N + i] = sqrt(nn2);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < n_bins; b += blockDim.x)
    if (hist[b]) atomicAdd(&counts[b], (unsigned long long)hist[b]);
}

void launch_sk(const double* s, int n_cfg, int N, int d, const int4* dirs, const int* koff, int D,
               int Mk, int max_L, double* out, hipStream_t st) {
  const int n_tiles = (D + 255) / 256;
  const int n_segs = (max_L + kSkSeg - 1) / kSkSeg;
  hipLaunchKernelGGL(sk_kernel, dim3(n_cfg * n_tiles * n_segs), dim3(256), 0, st, s, N, d, dirs,
                     koff, D, Mk, n_tiles, n_segs, out);
}

void launch_sk_sum(const double* sk, int n_cfg, int Mk, double* tot, double* tot2,
                   hipStream_t st) {
  hipLaunchKernelGGL(sk_sum_kernel, dim3((Mk + 255) / 256), dim3(256), 0, st, sk, n_cfg, Mk, tot,
                     tot2);
}

void launch_pair(const double* s, int B, int N, int d, const double* basis, const double* edges,
                 int n_edges, unsigned long long* counts, double* nn, hipStream_t st) {
  const int n_blocks = (N + 255) / 256;
  hipLaunchKernelGGL(pair_kernel, dim3(B * n_blocks), dim3(256),
                     sizeof(unsigned int) * (size_t)(n_edges - 1), st, s, N, d, basis, edges,
                     n_edges, counts, nn, n_blocks);
}

}  // namespace cco
