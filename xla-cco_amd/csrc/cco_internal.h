// Internal declarations shared by the gfx950 kernels and the C-ABI layer.
//
// Layout conventions (all device memory, all FP64 unless noted):
//   positions  s[lane][j][a]      lane stride = lane_stride doubles (>= N*d)
//   gradients  g[lane][j][a]      same strides as positions
//   rho box    part[lane][split][i], cbox[lane][i]   i < Mbox (complex)
// The k-set is stored as a "box" of columns: a column is one tail
// (n_2, ..., n_d) of the integer vector, and holds rows n_1 = 0 .. L-1.
// rho over the box is the complex contraction
//   rho[c][r] = sum_j z1_j^r * B_c(j),   B_c(j) = z2_j^{n2(c)} z3_j^{n3(c)},
//   z_a = exp(2 pi i s_a),
// built by phase recurrence along r, so no trig is evaluated per (k, j).
// The rho kernel's work unit is a run of at most kRowSeg consecutive rows of
// one column (a "unit"); units are sorted by length so the lanes of a wave
// run the same number of rows.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

namespace cco {

constexpr int kWave = 64;
constexpr int kRowSeg = 8;       // rows of one rho unit (accumulated in registers)
constexpr int kRhoChunkMax = 64; // particles staged in LDS per rho chunk (upper bound)
constexpr int kUnitEmpty = 1 << 30;  // rho unit flag (in .y): the run holds no k-vector
constexpr int kRhoUnitsMax = 256;  // rho units per workgroup (two threads per unit)
constexpr int kForceTile = 64;   // particles per force workgroup (one wave per k part)
constexpr int kForceParts = 4;   // k parts per force workgroup (waves sharing 64 particles)
constexpr int kReduceTile = 256; // box entries (and k rows) per reduce workgroup
constexpr int kMpChunk = 64;     // particles per LDS chunk of the split-precision kernels

// Geometry of one k-set, as the kernels read it.  Pointers are device memory.
struct Geo {
  int d;
  int M;        // real k count
  int M_pad;    // padded k count (multiple of pad_to)
  int Mbox;     // entries in the column box
  int Cn;       // number of columns
  int P1, P2, P3;  // max |n_a| per axis (P2 = P3 = 0 when absent)
  int n_units;  // rho units (column, row run) over the box
  int UT;       // rho unit tile width (threads per rho workgroup)
  int n_utiles; // rho workgroups per (lane, split)
  int Q;        // row-run seeds per particle: z1^(q kRowSeg), q < Q
  int rho_chunk;  // particles staged in LDS per rho chunk (16, 32 or 64)
  int rho_split;  // requested particles per rho split (0: the default rule)
  int n_rtiles;   // reduce workgroups per lane (kReduceTile box / k entries each)
  int rho_halves; // threads per rho unit (1 or 2)
  // split-precision (MFMA) path
  int mp_ok;      // the k-set fits the MFMA kernels (rows <= 48, LDS <= 64 KB)
  int RP;         // rows padded to a multiple of 16
  int mp_pg;      // force MFMA kernel: 64-particle groups per workgroup (1 or 2)
  int mp_cb;      // force MFMA kernel: 16-column tiles of C per LDS batch at mp_pg groups
  int mp_lane;    // split-precision lanes may run the one-workgroup-per-lane kernel (fits its LDS)
  double wmax, wmin;  // largest and smallest positive box weight (C scaling)
  const int4* gcol;   // per column (rho order): {box offset, L, n2, n3}
  int n_u;      // force traversal: u = 0 .. n_u-1 (u = |n2|)
  int fparts;   // force k parts in use (1 .. kForceParts)
  int2 fpart[kForceParts];  // per part: u range [x, y), balanced by rows
  double volume;
  double pref;  // 2 * (2 pi) / V
  const int4* unit;      // rho units: {box offset of the first row, nrows | q << 8, n2, n3}
  const int* kbox;       // k -> box index (M); ~i marks a conjugated row
  const double* wbox;    // summed weight of the k rows on each box entry (Mbox)
  const int2* ustart;    // per u: {first group, number of v values}
  const int4* gmem;      // per group: 4 members {box offset, L, n2, n3}
  // Split precision on the pair layout (d = 3): columns are (n1, n2) pairs and
  // rows r = n3 + mp_shift, the column phase carrying z3^-mp_shift.  In the
  // column-box Geo: the pair layout's Geo for the MFMA kernels (a host pointer,
  // null when not used) and, for the reduce of those lanes, its entry count,
  // weights, k map and map to this box (-1: no entry here).  Both layouts use
  // this box's Mbox as the stride of the partial sums and of C.
  const Geo* mp_geo;
  int mp_perm;   // this Geo is a pair layout
  int mp_shift;  // max |n3| of the k-set (pair layout)
  int mp_Mbox;
  const double* mp_wbox;
  const int* mp_kbox;
  const int* mp_p2b;
};

// Uneven batches: where one lane's data lives (offsets in elements).
struct LaneLayout {
  long long off;       // positions / gradients (doubles)
  long long part_off;  // rho partials (double2)
  long long cbox_off;  // C on the box (double2)
  long long rt_off;    // energy partials (doubles)
  int N;               // particles of the lane
  int pad;
};

// Per-launch batch description for the energy kernels.
struct Batch {
  const int* active;     // slot -> lane
  const int* n_active;   // device count (slots >= count exit)
  int N;                 // particles per lane
  int lane_stride;       // doubles between lanes in s and grad
  int n_splits;          // rho particle splits per lane
  int split_len;         // particles per split
  int n_ftiles;          // force tiles per lane
  const double* s;       // positions
  double* grad;          // gradient out (may be null for energy only)
  double* phi;           // energy out [lane]
  double* re;            // Re rho out [lane][M_pad] (may be null)
  double* im;            // Im rho out
  double2* part;         // rho partials [lane][split][Mbox]
  double2* cbox;         // w * rho per box entry [lane][Mbox]
  double* phi_part;      // energy partials [lane][n_rtiles]
  double2* rbox = nullptr;  // rho per box entry [lane][Mbox] (null: not written; uniform only)
  const int* prec = nullptr;  // [lane] 1 = split-precision MFMA path (null: every lane FP64)
  const int* skip = nullptr;  // [lane] nonzero: rho and reduce leave the lane alone (null: none)
  void* cpl = nullptr;        // split-precision C planes [lane][8][Cn padded to 16][RP + 8] fp16
  double* csig = nullptr;     // [lane] the C scale of those planes
  void* z1g = nullptr;        // split-precision z1 planes [lane][64-particle chunk][4][P1 + 1][64] fp16
  // uneven batches (null: every lane uses the launch's Geo and the strides above)
  const LaneLayout* lay = nullptr;  // [lane]
  const Geo* geos = nullptr;        // [lane] geometry (rho: UT = 64 units, two threads per unit)
  int tiles_rho = 0, tiles_reduce = 0, tiles_force = 0;  // workgroups per lane (max)
  int rho_lds = 0;                  // dynamic LDS of the rho kernel (bytes, max over lanes)
  int dmask = 0;                    // bit d set when some lane has dimension d
  int mp_lane = 0;  // set at launch: the split-precision lanes run mp_lane_kernel (reduce skips them)
  // uneven batches: LDS of mp_lane_kernel per (d, row tiles) class of the
  // lanes that may run split precision (0: no such lane)
  int lane_lds[4][4] = {};
};

// One lane's geometry and buffers inside a launch.
struct LaneIO {
  const double* s;
  double* grad;
  double2* part;
  double2* cbox;
  double* phi_part;
  double* re;
  double* im;
  double2* rbox;
  int N, n_splits, n_ftiles;
};

// MODE 0: decided by b.lay at run time; 1: uniform batch; 2: uneven batch.
template <int MODE = 0>
__device__ __forceinline__ LaneIO lane_io(const Geo& g, const Batch& b, int lane) {
  LaneIO io;
  if (MODE == 1 || (MODE == 0 && b.lay == nullptr)) {
    io.N = b.N;
    io.n_splits = b.n_splits;
    io.n_ftiles = b.n_ftiles;
    io.s = b.s + (size_t)lane * b.lane_stride;
    io.grad = b.grad != nullptr ? b.grad + (size_t)lane * b.lane_stride : nullptr;
    io.part = b.part + (size_t)lane * b.n_splits * g.Mbox;
    io.cbox = b.cbox + (size_t)lane * g.Mbox;
    io.phi_part = b.phi_part + (size_t)lane * g.n_rtiles;
    io.re = b.re != nullptr ? b.re + (size_t)lane * g.M_pad : nullptr;
    io.im = b.im != nullptr ? b.im + (size_t)lane * g.M_pad : nullptr;
    io.rbox = b.rbox != nullptr ? b.rbox + (size_t)lane * g.Mbox : nullptr;
  } else {
    const LaneLayout L = b.lay[lane];
    io.N = L.N;
    io.n_splits = (L.N + b.split_len - 1) / b.split_len;
    io.n_ftiles = (L.N + kForceTile - 1) / kForceTile;
    io.s = b.s + L.off;
    io.grad = b.grad != nullptr ? b.grad + L.off : nullptr;
    io.part = b.part + L.part_off;
    io.cbox = b.cbox + L.cbox_off;
    io.phi_part = b.phi_part + L.rt_off;
    io.re = nullptr;
    io.im = nullptr;
    io.rbox = nullptr;
  }
  return io;
}

void launch_energy(const Geo& g, const Batch& b, int slots_ub, bool with_grad,
                   hipStream_t stream);
// FP64 rho and reduce only (Re/Im rho in k order into b.re / b.im), uniform batches
void launch_rho_reduce(const Geo& g, const Batch& b, int slots_ub, hipStream_t stream);
void launch_energy_stage(int stage, const Geo& g, const Batch& b, int slots_ub, hipStream_t st);
size_t mp_cplane_bytes(const Geo& g);  // split-precision C planes of one lane
size_t mp_z1g_bytes(const Geo& g, int N);  // split-precision z1 planes of one lane
void launch_copy(const void* src, void* dst, size_t bytes, hipStream_t st);
void launch_sk(const double* s, int n_cfg, int N, int d, const int4* dirs, const int* koff, int D,
               int Mk, int max_L, double* out, hipStream_t st);
void launch_sk_sum(const double* sk, int n_cfg, int Mk, double* tot, double* tot2, hipStream_t st);
void launch_pair(const double* s, int B, int N, int d, const double* basis, const double* edges,
                 int n_edges, unsigned long long* counts, double* nn, hipStream_t st);
size_t rho_lds_bytes(const Geo& g);
size_t mp_lds_bytes(const Geo& g);
int mp_force_batch_for(const Geo& g);  // 16-column tiles per force batch (0: does not fit)
// The force MFMA kernel's particle groups per workgroup and C tiles per batch
// for a geometry (two groups when their LDS fits half a CU, else one).
void mp_force_plan(const Geo& g, int* pg, int* cb);
void launch_energy_mixed(int stage, const Geo& g, const Batch& b, int slots_ub, hipStream_t st);
// Density only, split precision, tiled kernels (z1 planes + rho_mp) for the
// batch's split-precision lanes; its reduce is the FP64 reduce_kernel.
void launch_rho_mixed(const Geo& g, const Batch& b, int slots_ub, hipStream_t st);
// 1 when a geometry's split-precision rho tables fit the rho kernel (density-only passes)
int mp_rho_fits(const Geo& g);
// 1 when a geometry's rho tiles fit mp_lane_kernel (the batch's N decides the rest)
int mp_lane_tiles_ok(const Geo& g);
// LDS bytes of mp_lane_kernel for one lane of this geometry and N (0: the lane does not fit it)
size_t mp_lane_fit(const Geo& g, int N);
// The split-precision lanes of this launch run mp_lane_kernel (one workgroup per lane).
bool mp_lane_used(const Geo& g, const Batch& b);

// Closed-form Hessians of B configurations (hessian.hip).  kn: n_k1, n_k2,
// n_k3, w_k per k (M rows); workspaces X, Xw: B x 2M x N, G: B x N x N,
// rho: B x 2M, D: B x N; H: B x (N d)^2 row-major.  blas is a rocBLAS handle
// (blas_create).  Returns 0 on success.
int hessian_batch(hipStream_t st, void* blas, const double* s, int B, int N, int d, int M,
                  const double* kn, double volume, double* X, double* Xw, double* G,
                  double* rho, double* D, double* H);
void* blas_create();
void blas_destroy(void* blas);
// Batched LU (hessian.hip): column-major n x n matrices n^2 apart.
int lu_factor_batched(void* blas, hipStream_t st, double* A, int n, int* ipiv, int* info,
                      int batch);
int lu_solve_batched(void* blas, hipStream_t st, double* A, int n, const int* ipiv, double* B,
                     int batch);

// ---------------- small helpers ----------------

struct Error {
  int code;
  std::string msg;
};

void set_last_error(const std::string& msg);

#define CCO_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::cco::set_last_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
      return CCO_ERR_DEVICE;                                                 \
    }                                                                        \
  } while (0)

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// One DPP move of a double (two 32-bit halves); lanes whose source is outside
// the row, or whose row is masked off, read 0.
template <int CTRL, int ROW_MASK, bool BOUND_ZERO>
__device__ __forceinline__ double dpp_d(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffffLL), CTRL, ROW_MASK, 0xf,
                                             BOUND_ZERO);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), CTRL, ROW_MASK, 0xf,
                                             BOUND_ZERO);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over a full wave (every lane active), returned in every lane.  A scan
// within each row of 16 lanes (row_shr 1, 2, 4, 8), then the row totals
// carried by row_bcast15 / row_bcast31 into lane 63; fixed order.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_d<0x111, 0xf, true>(v);   // row_shr:1
  v += dpp_d<0x112, 0xf, true>(v);   // row_shr:2
  v += dpp_d<0x114, 0xf, true>(v);   // row_shr:4
  v += dpp_d<0x118, 0xf, true>(v);   // row_shr:8
  v += dpp_d<0x142, 0xa, false>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_d<0x143, 0xc, false>(v);  // row_bcast:31 into rows 2 and 3
  return readlane_d(v, kWave - 1);
}

// Deterministic sum over a workgroup of up to 1024 threads (fixed tree order).

template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  constexpr int nw = BLOCK / kWave;
  v = wave_sum(v);
  const int wid = threadIdx.x / kWave;
  const int lid = threadIdx.x % kWave;
  __syncthreads();
  if (lid == 0) lds[wid] = v;
  __syncthreads();
  double t[nw];
#pragma unroll
  for (int i = 0; i < nw; ++i) t[i] = lds[i];
#pragma unroll
  for (int w = 1; w < nw; w *= 2)
#pragma unroll
    for (int i = 0; i + w < nw; i += 2 * w) t[i] += t[i + w];
  return t[0];
}

}  // namespace cco
