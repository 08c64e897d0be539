// Closed-form Hessian of the stealthy energy (src/xla_cco/energy.py:272-299):
//
//   H[(j,a),(l,b)] = (2 (2 pi)^2 / V) sum_k w_k n_ka n_kb
//                    [cos th_kj cos th_kl + sin th_kj sin th_kl
//                     - delta_jl (Re rho_k cos th_kj + Im rho_k sin th_kj)],
//   th_kj = 2 pi n_k . s_j,
//
// flattened row-major to (N d) x (N d) per configuration.  The pair term of
// block (a, b) is one product per configuration, G = Xw X^T with
// X = [cos | sin] (N x 2M, column k of cos holds th_k over the particles) and
// Xw = X with column k scaled by w_k n_ka n_kb: an O(N^2 M) FP64 GEMM, run by
// rocBLAS (strided batched over the configurations); the trigonometric
// table, the density sums, the diagonal term and the scatter into the (N d)^2
// layout are the kernels below.  Blocks (a, b) and (b, a) share G (it is
// symmetric and w_ab = w_ba), so d (d + 1) / 2 products cover a Hessian.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "cco_internal.h"

namespace cco {

// X[c][k][j] = cos th_kj, X[c][M + k][j] = sin th_kj (column-major N x 2M per
// configuration c).  kn holds n_k1, n_k2, n_k3, w_k per k.
__global__ void hess_trig_kernel(const double* __restrict__ s, int N, int d, int M,
                                 const double* __restrict__ kn, double* __restrict__ X) {
  const int k = blockIdx.x;
  const int c = blockIdx.y;
  const double n1 = kn[4 * k], n2 = kn[4 * k + 1], n3 = kn[4 * k + 2];
  const double* sc = s + (size_t)c * N * d;
  double* xc = X + (size_t)c * N * 2 * M;
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    double t = n1 * sc[(size_t)j * d];
    if (d >= 2) t += n2 * sc[(size_t)j * d + 1];
    if (d >= 3) t += n3 * sc[(size_t)j * d + 2];
    double sn, cs;
    sincospi(2.0 * t, &sn, &cs);
    xc[(size_t)k * N + j] = cs;
    xc[(size_t)(M + k) * N + j] = sn;
  }
}

// rho[c][k] = sum_j cos th_kj (k < M), sum_j sin th_kj (M <= k < 2M), in a
// fixed tree order.
__global__ __launch_bounds__(256) void hess_rho_kernel(const double* __restrict__ X, int N,
                                                       int M, double* __restrict__ rho) {
  __shared__ double red[256 / kWave];
  const int k = blockIdx.x;
  const int c = blockIdx.y;
  const double* col = X + ((size_t)c * 2 * M + k) * N;
  double acc = 0.0;
  for (int j = threadIdx.x; j < N; j += 256) acc += col[j];
  acc = block_sum<256>(acc, red);
  if (threadIdx.x == 0) rho[(size_t)c * 2 * M + k] = acc;
}

// Xw = X with both halves of column k scaled by w_k n_ka n_kb.
__global__ void hess_scale_kernel(const double* __restrict__ X, int N, int M,
                                  const double* __restrict__ kn, int a, int b,
                                  double* __restrict__ Xw) {
  const int kk = blockIdx.x;  // 0 .. 2M - 1
  const int c = blockIdx.y;
  const int k = kk < M ? kk : kk - M;
  const double wab = kn[4 * k + 3] * kn[4 * k + a] * kn[4 * k + b];
  const size_t base = ((size_t)c * 2 * M + kk) * N;
  for (int j = threadIdx.x; j < N; j += blockDim.x) Xw[base + j] = wab * X[base + j];
}

// D[c][j] = sum_k w_k n_ka n_kb (Re rho_k cos th_kj + Im rho_k sin th_kj).
__global__ void hess_diag_kernel(const double* __restrict__ X, const double* __restrict__ rho,
                                 int N, int M, const double* __restrict__ kn, int a, int b,
                                 double* __restrict__ D) {
  const int c = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const double* xc = X + (size_t)c * 2 * M * N;
  const double* rc = rho + (size_t)c * 2 * M;
  double acc = 0.0;
  for (int k = 0; k < M; ++k) {
    const double wab = kn[4 * k + 3] * kn[4 * k + a] * kn[4 * k + b];
    acc += wab * (rc[k] * xc[(size_t)k * N + j] + rc[M + k] * xc[(size_t)(M + k) * N + j]);
  }
  D[(size_t)c * N + j] = acc;
}

// H[c][(j,a),(l,b)] = H[c][(j,b),(l,a)] = pref (G[c][j][l] - delta_jl D[c][j]).
__global__ void hess_scatter_kernel(const double* __restrict__ G, const double* __restrict__ D,
                                    int N, int d, int a, int b, double pref,
                                    double* __restrict__ H) {
  const int c = blockIdx.z;
  const int j = blockIdx.y;
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= N) return;
  const size_t nd = (size_t)N * d;
  double v = G[((size_t)c * N + l) * N + j];  // column-major G: entry (j, l)
  if (l == j) v -= D[(size_t)c * N + j];
  v *= pref;
  double* hc = H + (size_t)c * nd * nd;
  hc[((size_t)j * d + a) * nd + (size_t)l * d + b] = v;
  if (a != b) hc[((size_t)j * d + b) * nd + (size_t)l * d + a] = v;
}

void* blas_create() {
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
  return (void*)h;
}

void blas_destroy(void* blas) {
  if (blas != nullptr) (void)rocblas_destroy_handle((rocblas_handle)blas);
}

// LU factorisation of `batch` column-major n x n matrices stored n^2 apart
// (partial pivoting, rocSOLVER), and solves with one right-hand side each.
int lu_factor_batched(void* blas, hipStream_t st, double* A, int n, int* ipiv, int* info,
                      int batch) {
  rocblas_handle h = (rocblas_handle)blas;
  if (rocblas_set_stream(h, st) != rocblas_status_success) return 1;
  return rocsolver_dgetrf_strided_batched(h, n, n, A, n, (rocblas_stride)n * n, ipiv,
                                          (rocblas_stride)n, info, batch) == rocblas_status_success
             ? 0
             : 2;
}

int lu_solve_batched(void* blas, hipStream_t st, double* A, int n, const int* ipiv, double* B,
                     int batch) {
  rocblas_handle h = (rocblas_handle)blas;
  if (rocblas_set_stream(h, st) != rocblas_status_success) return 1;
  return rocsolver_dgetrs_strided_batched(h, rocblas_operation_none, n, 1, A, n,
                                          (rocblas_stride)n * n, ipiv, (rocblas_stride)n, B, n,
                                          (rocblas_stride)n, batch) == rocblas_status_success
             ? 0
             : 2;
}

int hessian_batch(hipStream_t st, void* blas, const double* s, int B, int N, int d, int M,
                  const double* kn, double volume, double* X, double* Xw, double* G,
                  double* rho, double* D, double* H) {
  rocblas_handle h = (rocblas_handle)blas;
  if (rocblas_set_stream(h, st) != rocblas_status_success) return 1;
  const int tb = 256;
  hipLaunchKernelGGL(hess_trig_kernel, dim3(M, B), dim3(tb), 0, st, s, N, d, M, kn, X);
  hipLaunchKernelGGL(hess_rho_kernel, dim3(2 * M, B), dim3(256), 0, st, X, N, M, rho);
  const double pref = 2.0 * (2.0 * M_PI) * (2.0 * M_PI) / volume;
  const double one = 1.0, zero = 0.0;
  for (int a = 0; a < d; ++a) {
    for (int b = a; b < d; ++b) {
      hipLaunchKernelGGL(hess_scale_kernel, dim3(2 * M, B), dim3(tb), 0, st, X, N, M, kn, a, b,
                         Xw);
      // G = Xw X^T (N x N, column-major), one product per configuration
      const rocblas_status r = rocblas_dgemm_strided_batched(
          h, rocblas_operation_none, rocblas_operation_transpose, N, N, 2 * M, &one, Xw, N,
          (rocblas_stride)N * 2 * M, X, N, (rocblas_stride)N * 2 * M, &zero, G, N,
          (rocblas_stride)N * N, B);
      if (r != rocblas_status_success) return 2;
      hipLaunchKernelGGL(hess_diag_kernel, dim3((N + tb - 1) / tb, B), dim3(tb), 0, st, X, rho,
                         N, M, kn, a, b, D);
      hipLaunchKernelGGL(hess_scatter_kernel, dim3((N + tb - 1) / tb, N, B), dim3(tb), 0, st, G,
                         D, N, d, a, b, pref, H);
    }
  }
  return hipGetLastError() == hipSuccess ? 0 : 3;
}

}  // namespace cco
