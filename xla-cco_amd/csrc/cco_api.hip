// C-ABI layer: contexts, k-set geometry upload, batched energy and relax
// sessions.  See include/cco.h for the contract of every entry point.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/cco.h"
#include "cco_internal.h"
#include "minimize_kernels.h"

namespace cco {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

static int fail(int code, const std::string& msg) {
  set_last_error(msg);
  return code;
}

}  // namespace cco

using namespace cco;

// Device scratch of cco_value_grad_density_dev_noalloc: a file-scope array that
// every GPU the library runs on holds from module load to process end.  Host
// code must not pass `scratch` itself to a kernel (on the host it names a
// stand-in, not device memory): scratch_ptr() gives the device's own address,
// looked up once per device by scratch_setup() when a context is created.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// With `device` current: record the device address of `scratch` on it.
int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(CCO_ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device] != nullptr) return CCO_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr)
    return fail(CCO_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
  return CCO_OK;
}

void* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}
}  // namespace

struct cco_ctx {
  int device;
  hipStream_t stream;
  void* blas = nullptr;  // rocBLAS handle, created on first use (Hessians)
};

// RAII device scope: make the context's device current for the call.
struct DevScope {
  explicit DevScope(int dev) { (void)hipSetDevice(dev); }
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

static int dev_reserve(DevBuf& b, size_t bytes) {
  if (bytes <= b.bytes && b.p != nullptr) return CCO_OK;
  if (b.p != nullptr) {
    CCO_HIP(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  if (bytes == 0) return CCO_OK;
  CCO_HIP(hipMalloc(&b.p, bytes));
  b.bytes = bytes;
  return CCO_OK;
}

static void dev_release(DevBuf& b) {
  if (b.p != nullptr) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

struct cco_energy {
  cco_ctx* ctx;
  Geo geo;
  Geo geo_mp;     // d = 3: the pair layout of the split-precision tiled kernels (mp_ok 0: unused)
  int nmax[3];
  DevBuf tables;
  DevBuf tables_mp;
  // batch workspace for cco_value_grad_density*
  DevBuf s, phi, grad, re, im, part, cbox, active, n_active, rphi, prec, cpl, csig;
  int ident_B = 0;
  int precision = 0;  // 1: split-precision MFMA evaluations (when the k-set allows)
  int prec_B = 0;     // lanes of prec filled with 1
  // the k-set as given (n_k1, n_k2, n_k3, w_k per row) and the Hessian workspaces
  std::vector<double> kn_host;
  DevBuf kn, hs, hX, hXw, hG, hrho, hD, hH;
  bool kn_ready = false;
};

namespace {

constexpr int kSplitLen = 128;  // particles per rho split up to N = 2048 and in uneven batches
// rho LDS tables: at most half the CU's 160 KB, so two workgroups share a CU
constexpr size_t kRhoLdsMax = 80 * 1024;

// Particles per rho split of a uniform batch: 128 up to N = 2048, then longer
// (a multiple of 128) so that a lane has at most 16 splits.  The partial sums
// (splits x box entries per lane) are written by the rho kernel and read back
// by the reduce kernel: 128-particle splits made that 0.8 GB per launch at 32
// lanes of N = 10000 (d = 3, M = 13540).
int split_len_for(int N) {
  constexpr int kMaxSplits = 16;
  if (N <= kMaxSplits * kSplitLen) return kSplitLen;
  return ((N + kMaxSplits * kSplitLen - 1) / (kMaxSplits * kSplitLen)) * kSplitLen;
}
int n_splits_for(int N) { return (N + split_len_for(N) - 1) / split_len_for(N); }
// The split length a uniform batch runs: the energy's request when it is longer
// than the default (never more splits than the buffers sized by n_splits_for).
int split_len_of(const Geo& g, int N) {
  const int def = split_len_for(N);
  return g.rho_split > def ? g.rho_split : def;
}
int n_splits_of(const Geo& g, int N) { return (N + split_len_of(g, N) - 1) / split_len_of(g, N); }
int n_splits_uneven(int N) { return (N + kSplitLen - 1) / kSplitLen; }
int n_ftiles_for(int N) { return (N + kForceTile - 1) / kForceTile; }

// Energy partials of the tiled reduce.
int reserve_reduce(DevBuf& rphi, int B, const Geo& g) {
  return dev_reserve(rphi, sizeof(double) * (size_t)B * g.n_rtiles);
}

struct ColKey {
  int n2, n3;
  bool operator<(const ColKey& o) const { return std::tie(n2, n3) < std::tie(o.n2, o.n3); }
};

struct BoxCol {
  int n2, n3, L, off;
};

// The pair layout of a d = 3 k-set for the split-precision tiled kernels:
// columns (n1, n2) with n1 >= 0, each holding the contiguous rows
// r = n3 + P3 in [a, a + L) of its k-vectors, the column phase carrying
// z3^-P3.  It is used when every column's rows are contiguous, its box is no
// larger than the column box (whose Mbox stays the stride of the partial sums
// and of C), it has fewer columns and its tiles fit the kernels.
int build_pair_geometry(cco_energy* E, const std::vector<int>& n1, const std::vector<int>& n2,
                        const std::vector<int>& n3, const std::vector<char>& flip, const double* w,
                        const std::vector<BoxCol>& cols, const std::map<ColKey, int>& col_index) {
  const Geo& g = E->geo;
  Geo& p = E->geo_mp;
  p = g;
  p.mp_ok = 0;
  const int M = g.M;
  const int shift = g.P3;
  std::map<ColKey, std::pair<int, int>> range;  // (n1, n2) -> rows [lo, hi]
  for (int k = 0; k < M; ++k) {
    const int r = n3[k] + shift;
    auto it = range.find({n1[k], n2[k]});
    if (it == range.end())
      range[{n1[k], n2[k]}] = {r, r};
    else {
      it->second.first = std::min(it->second.first, r);
      it->second.second = std::max(it->second.second, r);
    }
  }
  std::vector<int4> pc;  // {off - a, a << 16 | (a + L), n1, n2}
  std::map<ColKey, int> poff;
  int off = 0;
  for (auto& kv : range) {
    const int a = kv.second.first, L = kv.second.second - a + 1;
    pc.push_back(make_int4(off - a, (a << 16) | (a + L), kv.first.n2, kv.first.n3));
    poff[kv.first] = off - a;
    off += L;
  }
  if (off > g.Mbox) return CCO_OK;  // larger than the column box: not used
  std::vector<int> kmap(M);
  std::vector<double> wp(off, 0.0);
  std::vector<char> hit(off, 0);
  for (int k = 0; k < M; ++k) {
    const int i = poff[{n1[k], n2[k]}] + n3[k] + shift;
    kmap[k] = flip[k] ? ~i : i;
    wp[i] += w[k];
    hit[i] = 1;
  }
  for (int i = 0; i < off; ++i)
    if (!hit[i]) return CCO_OK;  // a gap inside a column's rows: not used
  // pair-layout entry -> column-box entry of the same (n1, n2, n3)
  std::vector<int> p2b(off, -1);
  for (auto& kv : range) {
    const int base = poff[kv.first];
    for (int r = kv.second.first; r <= kv.second.second; ++r) {
      auto c = col_index.find({kv.first.n3, r - shift});
      if (c != col_index.end() && kv.first.n2 < cols[c->second].L) p2b[base + r] = cols[c->second].off + kv.first.n2;
    }
  }
  p.mp_perm = 1;
  p.mp_shift = shift;
  p.Cn = (int)pc.size();
  p.P1 = 2 * shift;  // row powers 0 .. 2 P3
  p.P2 = g.P1;       // column tables: n1 >= 0, |n2|
  p.P3 = g.P2;
  p.RP = ((p.P1 + 1 + 15) / 16) * 16;
  p.Mbox = g.Mbox;   // the stride of the partial sums and of C
  p.mp_Mbox = off;
  p.mp_geo = nullptr;
  // fewer columns means fewer column-operand products on the VALU, which bound
  // the tiled kernels more than the MFMA row padding does
  p.mp_ok = (p.RP <= 48 && p.wmin > 0.0 && mp_force_batch_for(p) >= 1 && mp_lds_bytes(p) <= 64 * 1024 &&
             p.Cn < g.Cn) ? 1 : 0;
  if (!p.mp_ok) return CCO_OK;
  p.mp_lane = 0;
  mp_force_plan(p, &p.mp_pg, &p.mp_cb);
  pc.resize(((pc.size() + 127) / 128) * 128, make_int4(0, 0, 0, 0));
  auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_gc = 0;
  const size_t o_kb = align(sizeof(int4) * pc.size());
  const size_t o_wb = align(o_kb + sizeof(int) * kmap.size());
  const size_t o_pb = align(o_wb + sizeof(double) * wp.size());
  const size_t total = align(o_pb + sizeof(int) * p2b.size());
  std::vector<char> host(total, 0);
  std::memcpy(host.data() + o_gc, pc.data(), sizeof(int4) * pc.size());
  std::memcpy(host.data() + o_kb, kmap.data(), sizeof(int) * kmap.size());
  std::memcpy(host.data() + o_wb, wp.data(), sizeof(double) * wp.size());
  std::memcpy(host.data() + o_pb, p2b.data(), sizeof(int) * p2b.size());
  if (int rc = dev_reserve(E->tables_mp, total)) return rc;
  CCO_HIP(hipMemcpy(E->tables_mp.p, host.data(), total, hipMemcpyHostToDevice));
  char* base = (char*)E->tables_mp.p;
  p.gcol = (const int4*)(base + o_gc);
  p.kbox = (const int*)(base + o_kb);
  p.wbox = (const double*)(base + o_wb);
  p.mp_kbox = p.kbox;
  p.mp_wbox = p.wbox;
  p.mp_p2b = (const int*)(base + o_pb);
  return CCO_OK;
}

// Point the column-box Geo at the pair layout (on: when it was built) or not.
void use_pair_layout(cco_energy* E, bool on) {
  Geo& g = E->geo;
  const Geo& p = E->geo_mp;
  const bool use = on && p.mp_ok && p.mp_perm;
  g.mp_geo = use ? &E->geo_mp : nullptr;
  g.mp_Mbox = use ? p.mp_Mbox : 0;
  g.mp_wbox = use ? p.mp_wbox : nullptr;
  g.mp_kbox = use ? p.mp_kbox : nullptr;
  g.mp_p2b = use ? p.mp_p2b : nullptr;
}

// Build the column box and traversal tables of a k-set (host), upload them.
int build_geometry(cco_energy* E, const int64_t* n, const double* w, int M, int d, double volume,
                   int pad_to) {
  Geo& g = E->geo;
  g.d = d;
  g.M = M;
  g.M_pad = ((M + pad_to - 1) / pad_to) * pad_to;
  g.volume = volume;
  g.pref = (2.0 * (2.0 * M_PI)) / volume;
  // canonical rows: n1 >= 0 (a row with n1 < 0 is negated and flagged conj;
  // |rho|^2 and the gradient term of -n equal those of n)
  std::vector<int> n1(M), n2(M), n3(M);
  std::vector<char> flip(M, 0);
  int P1 = 0, P2 = 0, P3 = 0;
  for (int k = 0; k < M; ++k) {
    int a = (int)n[(size_t)k * d + 0];
    int b = d >= 2 ? (int)n[(size_t)k * d + 1] : 0;
    int c = d >= 3 ? (int)n[(size_t)k * d + 2] : 0;
    if (a < 0) {
      a = -a;
      b = -b;
      c = -c;
      flip[k] = 1;
    }
    n1[k] = a;
    n2[k] = b;
    n3[k] = c;
    P1 = std::max(P1, a);
    P2 = std::max(P2, std::abs(b));
    P3 = std::max(P3, std::abs(c));
  }
  g.P1 = P1;
  g.P2 = P2;
  g.P3 = P3;
  E->nmax[0] = P1;
  E->nmax[1] = P2;
  E->nmax[2] = P3;
  std::map<ColKey, int> colL;
  for (int k = 0; k < M; ++k) {
    auto it = colL.find({n2[k], n3[k]});
    if (it == colL.end())
      colL[{n2[k], n3[k]}] = n1[k] + 1;
    else
      it->second = std::max(it->second, n1[k] + 1);
  }
  using Col = BoxCol;
  std::vector<Col> cols;
  for (auto& kv : colL) cols.push_back({kv.first.n2, kv.first.n3, kv.second, 0});
  std::stable_sort(cols.begin(), cols.end(), [](const Col& a, const Col& b) {
    if (a.L != b.L) return a.L > b.L;
    return std::tie(a.n2, a.n3) < std::tie(b.n2, b.n3);
  });
  int off = 0;
  std::map<ColKey, int> col_index;
  for (size_t c = 0; c < cols.size(); ++c) {
    cols[c].off = off;
    off += cols[c].L;
    col_index[{cols[c].n2, cols[c].n3}] = (int)c;
  }
  g.Mbox = off;
  g.Cn = (int)cols.size();
  // box rows holding a k-vector (a k-set with an inner cutoff leaves runs
  // near n1 = 0 empty)
  std::vector<char> occ(g.Mbox, 0);
  for (int k = 0; k < M; ++k) occ[cols[col_index[{n2[k], n3[k]}]].off + n1[k]] = 1;
  // rho units: runs of <= kRowSeg rows of one column, longest first (stable);
  // a run without a k-vector is flagged empty and sorted last, so that whole
  // waves of them write zero sums without the particle loop
  std::vector<int4> unit_tab;
  g.Q = 1;
  for (size_t c = 0; c < cols.size(); ++c) {
    const int nq = (cols[c].L + kRowSeg - 1) / kRowSeg;
    g.Q = std::max(g.Q, nq);
    for (int q = 0; q < nq; ++q) {
      const int nr = std::min(kRowSeg, cols[c].L - q * kRowSeg);
      bool any = false;
      for (int r = 0; r < nr; ++r) any = any || occ[cols[c].off + q * kRowSeg + r];
      unit_tab.push_back(make_int4(cols[c].off + q * kRowSeg, nr | (q << 8) | (any ? 0 : kUnitEmpty),
                                   cols[c].n2, cols[c].n3));
    }
  }
  if (g.Q > 255) return fail(CCO_ERR_ARG, "k-set too long along axis 1 for the rho units");
  auto eff_rows = [](const int4& a) { return (a.y & kUnitEmpty) ? 0 : (a.y & 0xff); };
  std::stable_sort(unit_tab.begin(), unit_tab.end(),
                   [&](const int4& a, const int4& b) { return eff_rows(a) > eff_rows(b); });
  g.n_units = (int)unit_tab.size();
  g.n_utiles = (g.n_units + kRhoUnitsMax - 1) / kRhoUnitsMax;
  const int per_tile = (g.n_units + g.n_utiles - 1) / g.n_utiles;
  g.UT = ((per_tile + kWave - 1) / kWave) * kWave;
  g.rho_halves = g.n_utiles == 1 ? 2 : 1;
  std::vector<int> kbox(M);
  std::vector<double> wbox(g.Mbox, 0.0);
  for (int k = 0; k < M; ++k) {
    const Col& c = cols[col_index[{n2[k], n3[k]}]];
    const int i = c.off + n1[k];
    kbox[k] = flip[k] ? ~i : i;
    wbox[i] += w[k];
  }
  // force traversal: groups (u = |n2|, v = |n3|) with four sign members
  auto member = [&](int a2, int a3) -> int4 {
    auto it = col_index.find({a2, a3});
    if (it == col_index.end()) return make_int4(0, 0, 0, 0);
    const Col& c = cols[it->second];
    return make_int4(c.off, c.L, c.n2, c.n3);
  };
  std::vector<int2> ustart;
  std::vector<int4> gmem;
  int ngroups = 0;
  for (int u = 0; u <= P2; ++u) {
    int vmax = -1;
    for (int v = 0; v <= P3; ++v) {
      int4 m[4] = {member(u, v), v > 0 ? member(u, -v) : make_int4(0, 0, 0, 0),
                   u > 0 ? member(-u, -v) : make_int4(0, 0, 0, 0),
                   (u > 0 && v > 0) ? member(-u, v) : make_int4(0, 0, 0, 0)};
      if (m[0].y || m[1].y || m[2].y || m[3].y) vmax = v;
    }
    ustart.push_back(make_int2(ngroups, vmax + 1));
    for (int v = 0; v <= vmax; ++v) {
      gmem.push_back(member(u, v));
      gmem.push_back(v > 0 ? member(u, -v) : make_int4(0, 0, 0, 0));
      gmem.push_back(u > 0 ? member(-u, -v) : make_int4(0, 0, 0, 0));
      gmem.push_back((u > 0 && v > 0) ? member(-u, v) : make_int4(0, 0, 0, 0));
      ++ngroups;
    }
  }
  g.n_u = P2 + 1;
  g.n_rtiles = (std::max(g.Mbox, g.M_pad) + kReduceTile - 1) / kReduceTile;
  // force parts: contiguous u ranges with about equal row counts
  {
    std::vector<long> rows_u(g.n_u, 0);
    long total = 0;
    for (int u = 0; u < g.n_u; ++u) {
      for (int v = 0; v < ustart[u].y; ++v)
        for (int q = 0; q < 4; ++q) rows_u[u] += gmem[4 * (size_t)(ustart[u].x + v) + q].y;
      total += rows_u[u];
    }
    g.fparts = std::max(1, std::min(kForceParts, g.n_u));
    int u = 0;
    long acc = 0;
    for (int p = 0; p < kForceParts; ++p) {
      const int begin = u;
      if (p < g.fparts) {
        const long target = (total * (p + 1) + g.fparts - 1) / g.fparts;
        // leave at least one u for each later part
        while (u < g.n_u - (g.fparts - 1 - p) && (acc < target || u == begin)) acc += rows_u[u++];
        if (p == g.fparts - 1) u = g.n_u;
      }
      g.fpart[p] = make_int2(begin, u);
    }
  }
  g.rho_chunk = kRhoChunkMax;
  while (g.rho_chunk > 16 && rho_lds_bytes(g) > kRhoLdsMax) g.rho_chunk /= 2;
  if (rho_lds_bytes(g) > kRhoLdsMax)
    return fail(CCO_ERR_ARG, "k-set too wide for the rho kernel's LDS tables (max |n| too large)");
  // split-precision path: per-column table, padded rows, weight range
  // padded to whole 128-column tiles with empty columns (L = 0)
  std::vector<int4> gcol_tab(((cols.size() + 127) / 128) * 128, make_int4(0, 0, 0, 0));
  for (size_t c = 0; c < cols.size(); ++c)
    gcol_tab[c] = make_int4(cols[c].off, cols[c].L, cols[c].n2, cols[c].n3);
  g.RP = ((P1 + 1 + 15) / 16) * 16;
  g.wmax = 0.0;
  g.wmin = 0.0;
  for (double wv : wbox) {
    if (wv > g.wmax) g.wmax = wv;
    if (wv > 0.0 && (g.wmin == 0.0 || wv < g.wmin)) g.wmin = wv;
  }
  g.mp_ok = (g.RP <= 48 && g.wmin > 0.0 && mp_force_batch_for(g) >= 1 &&
              mp_lds_bytes(g) <= 64 * 1024) ? 1 : 0;
  g.mp_pg = g.mp_cb = 0;
  if (g.mp_ok) mp_force_plan(g, &g.mp_pg, &g.mp_cb);
  g.mp_lane = (g.mp_ok && mp_lane_tiles_ok(g)) ? 1 : 0;
  g.mp_geo = nullptr;
  g.mp_perm = 0;
  // one device allocation holding every table
  auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_un = 0;
  const size_t o_gc = align(o_un + sizeof(int4) * unit_tab.size());
  const size_t o_kb = align(o_gc + sizeof(int4) * gcol_tab.size());
  const size_t o_wb = align(o_kb + sizeof(int) * kbox.size());
  const size_t o_us = align(o_wb + sizeof(double) * wbox.size());
  const size_t o_gm = align(o_us + sizeof(int2) * ustart.size());
  const size_t total = align(o_gm + sizeof(int4) * std::max<size_t>(gmem.size(), 1));
  std::vector<char> host(total, 0);
  std::memcpy(host.data() + o_un, unit_tab.data(), sizeof(int4) * unit_tab.size());
  std::memcpy(host.data() + o_gc, gcol_tab.data(), sizeof(int4) * gcol_tab.size());
  std::memcpy(host.data() + o_kb, kbox.data(), sizeof(int) * kbox.size());
  std::memcpy(host.data() + o_wb, wbox.data(), sizeof(double) * wbox.size());
  std::memcpy(host.data() + o_us, ustart.data(), sizeof(int2) * ustart.size());
  if (!gmem.empty()) std::memcpy(host.data() + o_gm, gmem.data(), sizeof(int4) * gmem.size());
  int rc = dev_reserve(E->tables, total);
  if (rc) return rc;
  CCO_HIP(hipMemcpy(E->tables.p, host.data(), total, hipMemcpyHostToDevice));
  char* base = (char*)E->tables.p;
  g.unit = (const int4*)(base + o_un);
  g.gcol = (const int4*)(base + o_gc);
  g.kbox = (const int*)(base + o_kb);
  g.wbox = (const double*)(base + o_wb);
  g.ustart = (const int2*)(base + o_us);
  g.gmem = (const int4*)(base + o_gm);
  if (d == 3 && g.mp_ok) {
    if ((rc = build_pair_geometry(E, n1, n2, n3, flip, w, cols, col_index))) return rc;
    use_pair_layout(E, true);
  }
  return CCO_OK;
}

__global__ void fill_int_kernel(int* a, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void iota_kernel(int* a, int n, int* count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
  if (i == 0) *count = n;
}

// Scratch of one batched evaluation: rho partials, C on the box, energy
// partials of the tiled reduce and the identity lane list.
struct EnergyWs {
  double2* part;
  double2* cbox;
  double* rphi;
  int* active;
  int* n_active;
};

// Bytes of each EnergyWs region for B lanes of N particles (FP64 path).
struct EnergyWsSizes {
  size_t part, cbox, rphi, active, n_active;
};

EnergyWsSizes energy_ws_sizes(const Geo& g, int B, int N) {
  EnergyWsSizes z;
  z.part = sizeof(double2) * (size_t)B * n_splits_for(N) * g.Mbox;
  z.cbox = sizeof(double2) * (size_t)B * g.Mbox;
  z.rphi = sizeof(double) * (size_t)B * g.n_rtiles;
  z.active = sizeof(int) * (size_t)B;
  z.n_active = sizeof(int);
  return z;
}

Batch energy_batch_desc(const Geo& g, const EnergyWs& ws, const double* s_dev, int B, int N,
                        double* phi_dev, double* grad_dev, double* re_dev, double* im_dev) {
  (void)B;
  Batch b;
  b.active = ws.active;
  b.n_active = ws.n_active;
  b.N = N;
  b.lane_stride = N * g.d;
  b.n_splits = n_splits_of(g, N);
  b.split_len = split_len_of(g, N);
  b.n_ftiles = n_ftiles_for(N);
  b.s = s_dev;
  b.grad = grad_dev;
  b.phi = phi_dev;
  b.re = re_dev;
  b.im = im_dev;
  b.part = ws.part;
  b.cbox = ws.cbox;
  b.phi_part = ws.rphi;
  b.prec = nullptr;
  return b;
}

// The split-precision planes of B lanes share one buffer: the C planes of every
// lane, then the z1 planes of every lane.
// (sized for the layout the tiled kernels use: the pair layout when set)
size_t mp_buffer_bytes(const Geo& g0, int B, int N) {
  const Geo& g = g0.mp_geo != nullptr ? *g0.mp_geo : g0;
  return (mp_cplane_bytes(g) + mp_z1g_bytes(g, N)) * (size_t)B;
}
void mp_bind(Batch& b, const Geo& g0, void* base, int B) {
  const Geo& g = g0.mp_geo != nullptr ? *g0.mp_geo : g0;
  b.cpl = base;
  b.z1g = base != nullptr ? (char*)base + mp_cplane_bytes(g) * (size_t)B : nullptr;
}

int energy_batch(cco_energy* E, const double* s_dev, int B, int N, double* phi_dev,
                 double* grad_dev, double* re_dev, double* im_dev) {
  const Geo& g = E->geo;
  const EnergyWsSizes z = energy_ws_sizes(g, B, N);
  int rc;
  if ((rc = dev_reserve(E->part, z.part))) return rc;
  if ((rc = dev_reserve(E->cbox, z.cbox))) return rc;
  if ((rc = dev_reserve(E->active, z.active))) return rc;
  if ((rc = dev_reserve(E->n_active, z.n_active))) return rc;
  hipStream_t st = E->ctx->stream;
  if ((rc = reserve_reduce(E->rphi, B, g))) return rc;
  if (E->ident_B != B) {
    hipLaunchKernelGGL(iota_kernel, dim3((B + 255) / 256), dim3(256), 0, st, (int*)E->active.p, B,
                       (int*)E->n_active.p);
    CCO_HIP(hipGetLastError());
    E->ident_B = B;
  }
  EnergyWs ws{(double2*)E->part.p, (double2*)E->cbox.p, (double*)E->rphi.p, (int*)E->active.p,
              (int*)E->n_active.p};
  Batch b = energy_batch_desc(g, ws, s_dev, B, N, phi_dev, grad_dev, re_dev, im_dev);
  if (E->precision == 1 && g.mp_ok) {
    if ((rc = dev_reserve(E->prec, sizeof(int) * (size_t)B))) return rc;
    if ((rc = dev_reserve(E->cpl, mp_buffer_bytes(g, B, N)))) return rc;
    if ((rc = dev_reserve(E->csig, sizeof(double) * (size_t)B))) return rc;
    mp_bind(b, g, E->cpl.p, B);
    b.csig = (double*)E->csig.p;
    if (E->prec_B < B) {
      hipLaunchKernelGGL(fill_int_kernel, dim3((B + 255) / 256), dim3(256), 0, st, (int*)E->prec.p,
                         B, 1);
      E->prec_B = B;
    }
    b.prec = (const int*)E->prec.p;
  }
  launch_energy(g, b, B, grad_dev != nullptr, st);
  CCO_HIP(hipGetLastError());
  return CCO_OK;
}

}  // namespace

// ===========================================================================
// relax session
// ===========================================================================
// Stages of one evaluation step; profiling accumulates per stage id.
constexpr int kStages = CCO_PROFILE_STAGES;
enum StageId { STG_RHO = 0, STG_REDUCE = 1, STG_FORCE = 2, STG_ADVANCE = 3, STG_COMPACT = 4,
               STG_PREPARE = 5, STG_GRAM_DOT = 6, STG_GRAM_SOLVE = 7, STG_GRAM_COMB = 8,
               STG_TRACK = 9 };
constexpr int METHOD_FLOW = CCO_METHOD_FLOW;
constexpr int METHOD_MD = CCO_METHOD_VERLET;

// A run splits a session's lanes into shards [lo, lo + count), each with its
// own stream, active list, lane count and progress word, so that while one
// shard streams its L-BFGS history through HBM another evaluates energies on
// the vector units.  Every kernel indexes lanes, so a lane's arithmetic does
// not depend on its shard.
constexpr int kMaxShards = 4;
constexpr int kShardMinLanes = 32;  // k shards by default from k x this many lanes (k <= 3)

struct Shard {
  hipStream_t stream = nullptr;
  int lo = 0, count = 0;
  RelaxDev dev;  // the session's arrays with this shard's active list, count, progress
  int slots_ub = 0;
  int64_t steps = 0;  // evaluation steps queued since the last init
  unsigned long long* progress_host = nullptr;
};

struct cco_relax {
  cco_ctx* ctx;
  cco_energy* E;
  int B, N, d, method;
  DescentCfg dc;
  LbfgsCfg lc;
  FlowCfg fc;
  MdCfg mc;
  int md_stage = 0;  // MD sessions: 0 created, 1 configured, 2 state uploaded
  double* flow_dt0 = nullptr;  // device copy of per-lane first steps
  RelaxDev dev;
  std::vector<DevBuf> bufs;
  DevBuf part, cbox, rphi, out_s, out_scalars, pack, pack_idx, cpl, csig, hist32;
  int split_history = 0;    // 1: split-precision lanes keep FP32 L-BFGS history
  int slots_ub = 0;
  int64_t steps = 0;
  int64_t lane_evals = 0;
  bool ready = false;
  int* host_count = nullptr;  // pinned, kMaxShards
  unsigned long long* progress_host = nullptr;  // pinned, mapped, kMaxShards: compact_kernel
  unsigned long long* progress_dev = nullptr;   // device address of progress_host
  int* shard_counts = nullptr;                  // [kMaxShards] running lanes per shard
  std::vector<Shard> shards;
  hipStream_t extra_streams[kMaxShards] = {};   // streams of shards 1 ..
  hipEvent_t fork_ev = nullptr;
  int shard_req = 0;        // requested shards: 0 auto
  unsigned* shard_steps = nullptr;  // [kMaxShards] device step counters (compact_kernel)
  // HIP graphs of graph_steps evaluation steps per shard (0: plain launches),
  // instantiated per (shard, launch-bound bucket) and replayed
  int graph_steps = 0;
  std::map<std::pair<int, int>, hipGraphExec_t> graphs;
  int64_t graph_builds = 0, graph_launches = 0;
  int64_t launches = 0;     // evaluation steps queued over all shards since the last init
  // optional per-stage timing: events around the launches of every
  // profile_every-th step of each shard (0: off)
  int profile_every = 0;
  int advance_variant = 0;  // requested: 0 auto, 1 global-memory, 2 register-resident, 3 Gram
  int variant_run = 0;      // the variant latched by the last init (never 0)
  int precision = 0;        // requested: 1 = split-precision evaluations
  int* prec_store = nullptr;  // [B] per-lane precision flags
  int* prec_ok = nullptr;     // [B] uneven sessions: lanes that fit the split-precision lane kernel
  int lane_lds[4][4] = {};    // uneven sessions: that kernel's LDS per (d, row tiles) class
  // uneven batch (per-lane energy and N): layouts and geometry copies on device
  bool uneven = false;
  size_t vec_total = 0;     // doubles of one position array (B N d when uniform)
  DevBuf un_lay, un_geo;
  int tiles_rho = 0, tiles_reduce = 0, tiles_force = 0, rho_lds = 0, dmask = 0;
  // flow by ROS2: positions, Hessians / LU factors, right-hand sides, attempted
  // steps, pivots, factorisation info and the chunk's lane count
  bool ros2 = false;
  DevBuf ros_S, ros_H, ros_rhs, ros_dts, ros_ipiv, ros_info, ros_cnt;
  // MD observables (cco_md_configure): device buffers behind md_obs and the
  // waves / recorded steps behind mc.kvec, mc.rec_steps
  MdObs md_obs{};
  std::vector<DevBuf> md_bufs;
  double md_basis[9] = {};     // lattice vectors as rows (cco_md_create)
  // unconstrained tracked waves on the reciprocal lattice: their k-set as an
  // internal energy (unit weights; only its rho and reduce run) and that
  // pass's partial sums, box values and energy parts (in md_bufs)
  cco_energy* trk = nullptr;
  int trk_splits = 0, trk_split_len = 0;  // its particle splits (fewer than the energy's)
  double2 *trk_part = nullptr, *trk_cbox = nullptr, *trk_rbox = nullptr;
  unsigned char* trk_z1g = nullptr;  // split precision: that pass's z1 planes (null: FP64 pass)
  double2* md_rbox = nullptr;  // the force pass's rho per box entry (md_obs.rbox)
  double* trk_rphi = nullptr;
  std::vector<hipEvent_t> evs;
  double stage_ms[kStages] = {};
  int64_t stage_n[kStages] = {};
};

namespace {

template <typename T>
int alloc_into(cco_relax* R, T*& out, size_t count) {
  DevBuf b;
  int rc = dev_reserve(b, std::max<size_t>(count, 1) * sizeof(T));
  if (rc) return rc;
  CCO_HIP(hipMemsetAsync(b.p, 0, b.bytes, R->ctx->stream));
  R->bufs.push_back(b);
  out = (T*)b.p;
  return CCO_OK;
}

__global__ void gather_rows_kernel(const double* __restrict__ src, const int* __restrict__ active,
                                   const int* __restrict__ n_active, double* __restrict__ dst,
                                   int nd, int stride) {
  const int slot = blockIdx.x;
  if (slot >= *n_active) return;
  const int lane = active[slot];
  for (int i = threadIdx.x; i < nd; i += blockDim.x)
    dst[(size_t)slot * nd + i] = src[(size_t)lane * stride + i];
}

__global__ void scatter_eval_kernel(const double* __restrict__ phi_p,
                                    const double* __restrict__ grad_p,
                                    const int* __restrict__ active,
                                    const int* __restrict__ n_active, double* __restrict__ phi,
                                    double* __restrict__ grad, int nd, int stride) {
  const int slot = blockIdx.x;
  if (slot >= *n_active) return;
  const int lane = active[slot];
  if (threadIdx.x == 0) phi[lane] = phi_p[slot];
  for (int i = threadIdx.x; i < nd; i += blockDim.x)
    grad[(size_t)lane * stride + i] = grad_p[(size_t)slot * nd + i];
}

__global__ void relax_init_kernel(RelaxDev r, int method) {
  const int lane = blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= r.B) return;
  if (r.prec != nullptr) r.prec[lane] = r.prec_ok != nullptr ? r.prec_ok[lane] : 1;
  r.status[lane] = ST_RUNNING;
  r.active[lane] = lane;
  if (lane == 0) *r.n_active = r.B;
  if (method == CCO_METHOD_DESCENT) {
    DescentLane L = {};
    L.mode = MODE_INIT;
    r.dl[lane] = L;
  } else if (method == CCO_METHOD_LBFGS) {
    LbfgsLane L = {};
    L.mode = MODE_INIT;
    r.ll[lane] = L;
  } else {
    FlowLane L = {};
    L.mode = MODE_INIT;
    r.fl[lane] = L;
  }
}

Batch relax_batch(cco_relax* R, const RelaxDev& dv) {
  Batch b;
  b.active = dv.active;
  b.n_active = dv.n_active;
  b.N = R->N;
  b.lane_stride = R->dev.stride;
  b.n_splits = R->uneven ? n_splits_for(R->N) : n_splits_of(R->E->geo, R->N);
  b.split_len = R->uneven ? kSplitLen : split_len_of(R->E->geo, R->N);
  b.n_ftiles = n_ftiles_for(R->N);
  b.s = R->dev.s_eval;
  b.grad = R->dev.g_eval;
  b.phi = R->dev.phi_eval;
  b.re = nullptr;
  b.im = nullptr;
  b.part = (double2*)R->part.p;
  b.cbox = (double2*)R->cbox.p;
  b.phi_part = (double*)R->rphi.p;
  b.prec = R->dev.prec;
  mp_bind(b, R->E->geo, R->cpl.p, R->B);
  b.csig = (double*)R->csig.p;
  if (R->method == METHOD_MD && R->mc.w > 0) {  // the tracked constrained waves read re / im
    b.rbox = R->md_rbox;  // md_waves_kernel gathers the constrained waves from it
  }
  if (R->uneven) {
    b.lay = (const LaneLayout*)R->un_lay.p;
    b.geos = (const Geo*)R->un_geo.p;
    b.tiles_rho = R->tiles_rho;
    b.tiles_reduce = R->tiles_reduce;
    b.tiles_force = R->tiles_force;
    b.rho_lds = R->rho_lds;
    b.dmask = R->dmask;
    std::memcpy(b.lane_lds, R->lane_lds, sizeof(b.lane_lds));
  }
  return b;
}

// The tracked-wave pass of an MD step: FP64 rho + reduce of the unconstrained
// waves at the positions md_prepare wrote, for the lanes whose cadence flag
// md_prepare cleared; rho per box entry lands in trk_rbox.
Batch track_batch(cco_relax* R, const RelaxDev& dv) {
  Batch b;
  b.active = dv.active;
  b.n_active = dv.n_active;
  b.N = R->N;
  b.lane_stride = R->dev.stride;
  b.n_splits = R->trk_splits;
  b.split_len = R->trk_split_len;
  b.n_ftiles = n_ftiles_for(R->N);
  b.s = R->dev.s_eval;
  b.grad = nullptr;
  b.phi = nullptr;
  b.re = nullptr;
  b.im = nullptr;
  b.rbox = R->trk_rbox;
  b.part = R->trk_part;
  b.cbox = R->trk_cbox;
  b.phi_part = R->trk_rphi;
  b.skip = R->mc.trk_skip;
  if (R->trk_z1g != nullptr && R->dev.prec != nullptr) {  // split-precision lanes: MFMA rho
    b.prec = R->dev.prec;
    b.z1g = R->trk_z1g;
  }
  return b;
}

void relax_stage(cco_relax* R, const Shard& sh, int stage) {
  hipStream_t st = sh.stream;
  const RelaxDev& dv = sh.dev;
  const int ub = sh.slots_ub;
  if (stage <= STG_FORCE) {
    launch_energy_stage(stage, R->E->geo, relax_batch(R, dv), ub, st);
  } else if (stage == STG_ADVANCE) {
    if (R->method == CCO_METHOD_DESCENT)
      launch_descent_advance(dv, R->dc, ub, st);
    else if (R->method == CCO_METHOD_LBFGS)
      launch_lbfgs_advance(dv, R->lc, ub, st, R->variant_run);
    else if (R->method == METHOD_FLOW)
      launch_flow_advance(dv, R->fc, ub, st);
    else
      launch_md_finish(dv, R->mc, R->md_obs, ub, st);
  } else if (stage == STG_COMPACT) {
    launch_compact(dv, st, sh.lo, sh.count, (unsigned)(sh.steps + 1));
  } else if (stage == STG_TRACK) {
    const Batch tb = track_batch(R, dv);
    launch_rho_mixed(R->trk->geo, tb, ub, st);  // split-precision lanes (the FP64 rho skips them)
    launch_rho_reduce(R->trk->geo, tb, ub, st);
  } else if (stage >= STG_GRAM_DOT) {
    launch_lbfgs_gram(dv, R->lc, ub, stage - STG_GRAM_DOT + 1, st);
  } else {
    launch_md_prepare(dv, R->mc, ub, st);
  }
}

// The whole session as one shard on the context stream, for the
// external-energy path (no progress word: that path syncs every step).
Shard whole_view(const cco_relax* R) {
  Shard sh;
  sh.stream = R->ctx->stream;
  sh.lo = 0;
  sh.count = R->B;
  sh.dev = R->dev;
  sh.dev.progress = nullptr;
  sh.slots_ub = R->slots_ub;
  sh.steps = R->steps;
  return sh;
}

// Stage sequence of one evaluation step of the session's method.
int step_stages(const cco_relax* R, int* out) {
  int n = 0;
  if (R->method == METHOD_MD) out[n++] = STG_PREPARE;
  if (R->method == METHOD_MD && R->trk != nullptr) out[n++] = STG_TRACK;
  out[n++] = STG_RHO;
  out[n++] = STG_REDUCE;
  out[n++] = STG_FORCE;
  out[n++] = STG_ADVANCE;
  if (R->method == CCO_METHOD_LBFGS && R->variant_run == 3) {
    out[n++] = STG_GRAM_DOT;
    out[n++] = STG_GRAM_SOLVE;
    out[n++] = STG_GRAM_COMB;
  }
  out[n++] = STG_COMPACT;
  return n;
}

// The L-BFGS advance variant a session runs: the requested one, or for 0 the
// register-resident two-loop when N*d fits it, else the Gram form (its dot
// and direction passes spread a lane over many workgroups, where the
// global-memory two-loop keeps one workgroup per lane: at d = 3, N = 4000 it
// relaxes 2.6x faster at 64 lanes and 1.2x at 256), else (uneven batches,
// memory above its limit) the global-memory two-loop.  The choice depends on
// N*d and the memory only, never on the lane count, so a lane's bits do not
// depend on the batch it is in.
int resolve_variant(const cco_relax* R) {
  if (R->advance_variant != 0) return R->advance_variant;
  if (lbfgs_register_ept(R->dev.nd)) return 2;
  return (R->dev.gram != nullptr && !R->uneven) ? 3 : 1;
}

// Per-lane precision flags for the run being initialised: split precision
// when asked for, the k-set fits the MFMA kernels and the method uses it.
// Device buffers of the split-precision C planes, when the run uses them.
// FP32 history rows for the split-precision lanes of an L-BFGS run on the
// register-resident two-loop (the other advance variants keep FP64 rows).
int reserve_split_history(cco_relax* R) {
  R->dev.S32 = nullptr;
  R->dev.Y32 = nullptr;
  if (R->dev.prec == nullptr || R->method != CCO_METHOD_LBFGS || R->variant_run != 2 ||
      !R->split_history)
    return CCO_OK;
  const size_t n = R->vec_total * (size_t)R->lc.memory;
  const size_t half = (n + 63) / 64 * 64;  // Y rows start 256-byte aligned
  int rc = dev_reserve(R->hist32, sizeof(float) * 2 * half);
  if (rc != CCO_OK) return rc;
  R->dev.S32 = (float*)R->hist32.p;
  R->dev.Y32 = R->dev.S32 + half;
  return CCO_OK;
}

int reserve_mp_planes(cco_relax* R) {
  if (int rc = reserve_split_history(R)) return rc;
  if (R->dev.prec == nullptr || R->uneven) return CCO_OK;  // uneven: one workgroup per lane, no planes in HBM
  int rc = dev_reserve(R->cpl, mp_buffer_bytes(R->E->geo, R->B, R->N));
  if (rc == CCO_OK) rc = dev_reserve(R->csig, sizeof(double) * (size_t)R->B);
  return rc;
}

int* resolve_prec(const cco_relax* R) {
  if (R->precision != 1) return nullptr;
  if (R->uneven) return (R->method == CCO_METHOD_LBFGS && R->prec_ok != nullptr) ? R->prec_store : nullptr;
  if (R->E == nullptr || !R->E->geo.mp_ok) return nullptr;
  if (R->method != CCO_METHOD_LBFGS && R->method != METHOD_MD && R->method != METHOD_FLOW) return nullptr;
  if (R->method == METHOD_FLOW && R->ros2) return nullptr;  // ROS2 evaluates in FP64
  return R->prec_store;
}

int relax_advance(cco_relax* R) {
  const Shard sh = whole_view(R);
  relax_stage(R, sh, STG_ADVANCE);
  if (R->method == CCO_METHOD_LBFGS && R->variant_run == 3) {
    relax_stage(R, sh, STG_GRAM_DOT);
    relax_stage(R, sh, STG_GRAM_SOLVE);
    relax_stage(R, sh, STG_GRAM_COMB);
  }
  relax_stage(R, sh, STG_COMPACT);
  CCO_HIP(hipGetLastError());
  return CCO_OK;
}

// Pinned host words of a session: the lane counts read back at a sync, and
// the progress words the compact kernel stores to (coherent, mapped, so the
// host sees them while later steps are queued); the per-shard lane counts on
// the device.
int alloc_host_words(cco_relax* R) {
  if (hipHostMalloc((void**)&R->host_count, sizeof(int) * kMaxShards) != hipSuccess)
    return fail(CCO_ERR_DEVICE, "hipHostMalloc failed");
  if (hipHostMalloc((void**)&R->progress_host, sizeof(unsigned long long) * kMaxShards,
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
    return fail(CCO_ERR_DEVICE, "hipHostMalloc (mapped) failed");
  for (int k = 0; k < kMaxShards; ++k) R->progress_host[k] = ~0ull;
  void* dp = nullptr;
  if (hipHostGetDevicePointer(&dp, R->progress_host, 0) != hipSuccess)
    return fail(CCO_ERR_DEVICE, "hipHostGetDevicePointer failed");
  R->progress_dev = (unsigned long long*)dp;
  R->dev.progress = nullptr;
  R->dev.step_ctr = nullptr;
  if (int rc = alloc_into(R, R->shard_steps, kMaxShards)) return rc;
  return alloc_into(R, R->shard_counts, kMaxShards);
}

void drop_graphs(cco_relax* R) {
  for (auto& kv : R->graphs) (void)hipGraphExecDestroy(kv.second);
  R->graphs.clear();
}

// Split the lanes into shards for the run being initialised and queue each
// shard's first compaction (its active list and count) on the context stream.
// Built-in-energy descent and L-BFGS sessions use two shards from
// 2 kShardMinLanes lanes on (or the requested count); the others one.
int setup_shards(cco_relax* R) {
  int n = 1;
  drop_graphs(R);  // their launches capture this run's shard layout and kernels
  if (R->E != nullptr && (R->method == CCO_METHOD_DESCENT || R->method == CCO_METHOD_LBFGS))
    n = R->shard_req > 0 ? R->shard_req : std::min(3, R->B / kShardMinLanes);
  n = std::max(1, std::min(n, std::min(kMaxShards, R->B)));
  R->shards.assign(n, Shard());
  for (int k = 0; k < n; ++k) {
    Shard& sh = R->shards[k];
    if (k == 0) {
      sh.stream = R->ctx->stream;
    } else {
      if (R->extra_streams[k] == nullptr)
        CCO_HIP(hipStreamCreateWithFlags(&R->extra_streams[k], hipStreamNonBlocking));
      sh.stream = R->extra_streams[k];
    }
    sh.lo = (int)((int64_t)R->B * k / n);
    sh.count = (int)((int64_t)R->B * (k + 1) / n) - sh.lo;
    sh.dev = R->dev;
    sh.dev.active = R->dev.active + sh.lo;
    sh.dev.n_active = R->shard_counts + k;
    sh.dev.progress = R->progress_dev + k;
    sh.dev.step_ctr = nullptr;
    sh.progress_host = R->progress_host + k;
    sh.slots_ub = sh.count;
    sh.steps = 0;
    *sh.progress_host = ~0ull;  // no step of this run has finished
    launch_compact(sh.dev, R->ctx->stream, sh.lo, sh.count, 0u);  // tag 0 is never read
    sh.dev.step_ctr = R->shard_steps + k;  // the run's steps count from 1 on the device too
  }
  hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(kMaxShards), 0, R->ctx->stream, (int*)R->shard_steps,
                     kMaxShards, 0);
  if (R->fork_ev == nullptr) CCO_HIP(hipEventCreateWithFlags(&R->fork_ev, hipEventDisableTiming));
  CCO_HIP(hipGetLastError());
  R->steps = 0;
  R->launches = 0;
  return CCO_OK;
}

// Steps of a shard still queued after the last one its progress word reports
// finished.  The word (completed steps, running lanes) also lowers the
// shard's launch bound as lanes stop (the count only falls during a run),
// without draining the stream.  `first` is the shard's step count when the
// current call began: only words tagged after it are this call's.
int64_t shard_in_flight(Shard& sh, int64_t first) {
  const unsigned long long w = __atomic_load_n(sh.progress_host, __ATOMIC_ACQUIRE);
  const int64_t tag = (int64_t)(w >> 32);
  int64_t done = first;
  if (w != ~0ull && tag > first && tag <= sh.steps) {
    done = tag;
    sh.slots_ub = std::min(sh.slots_ub, (int)(w & 0xffffffffu));
  }
  return sh.steps - done;
}

// Steps are queued at most this many ahead of the device, per shard.
constexpr int64_t kRunAhead = 16;

int relax_sync_count(cco_relax* R, int* n_active) {
  CCO_HIP(hipMemcpyAsync(R->host_count, R->dev.n_active, sizeof(int), hipMemcpyDeviceToHost,
                         R->ctx->stream));
  CCO_HIP(hipStreamSynchronize(R->ctx->stream));
  R->slots_ub = *R->host_count;
  if (n_active) *n_active = R->slots_ub;
  return CCO_OK;
}

}  // namespace

// ===========================================================================
// extern "C"
// ===========================================================================
extern "C" {

int cco_version(void) { return 100; }

const char* cco_last_error(void) { return g_last_error.c_str(); }

int cco_device_count(int* count) {
  if (!count) return fail(CCO_ERR_ARG, "count is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return CCO_OK;
}

int cco_ctx_create(int device, cco_ctx** out) {
  if (!out) return fail(CCO_ERR_ARG, "out is NULL");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c == 0)
    return fail(CCO_ERR_DEVICE, "no GPU visible to the HIP runtime");
  if (device < 0 || device >= c) return fail(CCO_ERR_ARG, "device index out of range");
  CCO_HIP(hipSetDevice(device));
  if (const int rc = scratch_setup(device)) return rc;
  cco_ctx* ctx = new cco_ctx();
  ctx->device = device;
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return fail(CCO_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  *out = ctx;
  return CCO_OK;
}

int cco_ctx_destroy(cco_ctx* ctx) {
  if (!ctx) return CCO_OK;
  DevScope ds(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  blas_destroy(ctx->blas);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CCO_OK;
}

int cco_ctx_synchronize(cco_ctx* ctx) {
  if (!ctx) return fail(CCO_ERR_ARG, "ctx is NULL");
  DevScope ds(ctx->device);
  CCO_HIP(hipStreamSynchronize(ctx->stream));
  return CCO_OK;
}

int cco_dev_alloc(cco_ctx* ctx, size_t bytes, void** out) {
  if (!ctx || !out) return fail(CCO_ERR_ARG, "NULL argument");
  DevScope ds(ctx->device);
  CCO_HIP(hipMalloc(out, std::max<size_t>(bytes, 1)));
  return CCO_OK;
}

int cco_dev_free(cco_ctx* ctx, void* ptr) {
  if (!ctx) return fail(CCO_ERR_ARG, "ctx is NULL");
  DevScope ds(ctx->device);
  if (ptr) CCO_HIP(hipFree(ptr));
  return CCO_OK;
}

int cco_memcpy_h2d(cco_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return fail(CCO_ERR_ARG, "ctx is NULL");
  DevScope ds(ctx->device);
  CCO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  CCO_HIP(hipStreamSynchronize(ctx->stream));
  return CCO_OK;
}

int cco_memcpy_d2h(cco_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return fail(CCO_ERR_ARG, "ctx is NULL");
  DevScope ds(ctx->device);
  CCO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  CCO_HIP(hipStreamSynchronize(ctx->stream));
  return CCO_OK;
}

int cco_energy_create(cco_ctx* ctx, const int64_t* n, const double* w, int M, int d,
                      double volume, int pad_to, cco_energy** out) {
  if (!ctx || !out) return fail(CCO_ERR_ARG, "NULL argument");
  if (M <= 0) return fail(CCO_ERR_ARG, "empty k-set: no k-vectors in the shell");
  if (d < 1 || d > 3) return fail(CCO_ERR_ARG, "d must be 1, 2 or 3");
  if (!n || !w) return fail(CCO_ERR_ARG, "n and w must not be NULL");
  if (pad_to < 1) return fail(CCO_ERR_ARG, "pad_to must be >= 1");
  if (!(volume > 0.0) || !std::isfinite(volume))
    return fail(CCO_ERR_ARG, "volume must be positive and finite");
  DevScope ds(ctx->device);
  cco_energy* E = new cco_energy();
  E->ctx = ctx;
  E->kn_host.assign((size_t)4 * M, 0.0);
  for (int k = 0; k < M; ++k) {
    for (int a = 0; a < d; ++a) E->kn_host[4 * (size_t)k + a] = (double)n[(size_t)k * d + a];
    E->kn_host[4 * (size_t)k + 3] = w[k];
  }
  int rc = build_geometry(E, n, w, M, d, volume, pad_to);
  if (rc) {
    dev_release(E->tables);
    dev_release(E->tables_mp);
    delete E;
    return rc;
  }
  *out = E;
  return CCO_OK;
}

int cco_energy_destroy(cco_energy* E) {
  if (!E) return CCO_OK;
  DevScope ds(E->ctx->device);
  (void)hipStreamSynchronize(E->ctx->stream);
  for (DevBuf* b : {&E->tables, &E->tables_mp, &E->s, &E->phi, &E->grad, &E->re, &E->im, &E->part, &E->cbox,
                    &E->active, &E->n_active, &E->rphi, &E->prec, &E->cpl, &E->csig, &E->kn, &E->hs, &E->hX, &E->hXw,
                    &E->hG, &E->hrho, &E->hD, &E->hH})
    dev_release(*b);
  delete E;
  return CCO_OK;
}

namespace {

// Hessians of B configurations (device s, device h) on stream st with the
// energy's workspaces and its context's rocBLAS handle.
int hessian_on(cco_energy* E, hipStream_t st, const double* s_dev, int B, int N, double* h_dev) {
  cco_ctx* ctx = E->ctx;
  if (ctx->blas == nullptr && (ctx->blas = blas_create()) == nullptr)
    return fail(CCO_ERR_DEVICE, "rocblas_create_handle failed");
  const int M = E->geo.M, d = E->geo.d;
  int rc;
  if (!E->kn_ready) {
    if ((rc = dev_reserve(E->kn, sizeof(double) * E->kn_host.size()))) return rc;
    CCO_HIP(hipMemcpyAsync(E->kn.p, E->kn_host.data(), sizeof(double) * E->kn_host.size(),
                           hipMemcpyHostToDevice, st));
    E->kn_ready = true;
  }
  const size_t xs = (size_t)B * 2 * M * N;
  if ((rc = dev_reserve(E->hX, sizeof(double) * xs))) return rc;
  if ((rc = dev_reserve(E->hXw, sizeof(double) * xs))) return rc;
  if ((rc = dev_reserve(E->hG, sizeof(double) * B * (size_t)N * N))) return rc;
  if ((rc = dev_reserve(E->hrho, sizeof(double) * B * (size_t)2 * M))) return rc;
  if ((rc = dev_reserve(E->hD, sizeof(double) * B * (size_t)N))) return rc;
  if (hessian_batch(st, ctx->blas, s_dev, B, N, d, M, (const double*)E->kn.p, E->geo.volume,
                    (double*)E->hX.p, (double*)E->hXw.p, (double*)E->hG.p, (double*)E->hrho.p,
                    (double*)E->hD.p, h_dev) != 0)
    return fail(CCO_ERR_DEVICE, "Hessian launch failed (rocBLAS or kernel)");
  return CCO_OK;
}

}  // namespace

int cco_hessian_dev(cco_energy* E, const double* s_dev, int B, int N, double* h_dev) {
  if (!E || !s_dev || !h_dev) return fail(CCO_ERR_ARG, "NULL argument");
  if (B < 1 || N < 1) return fail(CCO_ERR_ARG, "need B >= 1 and N >= 1");
  DevScope ds(E->ctx->device);
  return hessian_on(E, E->ctx->stream, s_dev, B, N, h_dev);
}

int cco_hessian(cco_energy* E, const double* s_host, int B, int N, double* h_host) {
  if (!E || !s_host || !h_host) return fail(CCO_ERR_ARG, "NULL argument");
  if (B < 1 || N < 1) return fail(CCO_ERR_ARG, "need B >= 1 and N >= 1");
  DevScope ds(E->ctx->device);
  hipStream_t st = E->ctx->stream;
  const size_t nd = (size_t)N * E->geo.d;
  int rc;
  if ((rc = dev_reserve(E->hs, sizeof(double) * B * nd))) return rc;
  if ((rc = dev_reserve(E->hH, sizeof(double) * B * nd * nd))) return rc;
  CCO_HIP(hipMemcpyAsync(E->hs.p, s_host, sizeof(double) * B * nd, hipMemcpyHostToDevice, st));
  if ((rc = cco_hessian_dev(E, (const double*)E->hs.p, B, N, (double*)E->hH.p))) return rc;
  CCO_HIP(hipMemcpyAsync(h_host, E->hH.p, sizeof(double) * B * nd * nd, hipMemcpyDeviceToHost, st));
  CCO_HIP(hipStreamSynchronize(st));
  return CCO_OK;
}

int cco_energy_info(const cco_energy* E, int* M, int* M_pad, int* d, double* volume,
                    int* nmax3) {
  if (!E) return fail(CCO_ERR_ARG, "energy is NULL");
  if (M) *M = E->geo.M;
  if (M_pad) *M_pad = E->geo.M_pad;
  if (d) *d = E->geo.d;
  if (volume) *volume = E->geo.volume;
  if (nmax3) {
    nmax3[0] = E->nmax[0];
    nmax3[1] = E->nmax[1];
    nmax3[2] = E->nmax[2];
  }
  return CCO_OK;
}

int cco_value_grad_density_dev(cco_energy* E, const double* s_dev, int B, int N, double* phi_dev,
                               double* grad_dev, double* re_dev, double* im_dev) {
  if (!E || !s_dev || !phi_dev) return fail(CCO_ERR_ARG, "NULL argument");
  if (B < 1 || N < 1) return fail(CCO_ERR_ARG, "need B >= 1 and N >= 1");
  if ((re_dev == nullptr) != (im_dev == nullptr))
    return fail(CCO_ERR_ARG, "re and im must both be given or both be NULL");
  DevScope ds(E->ctx->device);
  return energy_batch(E, s_dev, B, N, phi_dev, grad_dev, re_dev, im_dev);
}

size_t cco_noalloc_scratch_bytes(void) { return kScratchBytes; }

int cco_value_grad_density_dev_noalloc(cco_energy* E, const double* s_dev, int B, int N,
                                       double* phi_dev, double* grad_dev, double* re_dev,
                                       double* im_dev) {
  if (!E || !s_dev || !phi_dev) return fail(CCO_ERR_ARG, "NULL argument");
  if (B < 1 || N < 1) return fail(CCO_ERR_ARG, "need B >= 1 and N >= 1");
  if ((re_dev == nullptr) != (im_dev == nullptr))
    return fail(CCO_ERR_ARG, "re and im must both be given or both be NULL");
  const Geo& g = E->geo;
  if (E->precision == 1 && g.mp_ok)
    return fail(CCO_ERR_STATE, "the allocation-free evaluation is FP64 only: set the energy's precision to fp64");
  const EnergyWsSizes z = energy_ws_sizes(g, B, N);
  auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
  const size_t need = up(z.part) + up(z.cbox) + up(z.rphi) + up(z.active) + up(z.n_active);
  if (need > kScratchBytes)
    return fail(CCO_ERR_ARG, "batch needs " + std::to_string(need) + " bytes of scratch, more than the " +
                                 std::to_string(kScratchBytes) + " of the allocation-free evaluation");
  DevScope ds(E->ctx->device);
  unsigned char* base = (unsigned char*)scratch_ptr(E->ctx->device);
  if (base == nullptr) return fail(CCO_ERR_STATE, "scratch not set up on this device");
  EnergyWs ws;
  size_t at = 0;
  ws.part = (double2*)(base + at);
  at += up(z.part);
  ws.cbox = (double2*)(base + at);
  at += up(z.cbox);
  ws.rphi = (double*)(base + at);
  at += up(z.rphi);
  ws.active = (int*)(base + at);
  at += up(z.active);
  ws.n_active = (int*)(base + at);
  hipStream_t st = E->ctx->stream;
  // the lane list and its count are rewritten on every call: the previous
  // call's contents (another B) must not be read
  hipLaunchKernelGGL(iota_kernel, dim3((B + 255) / 256), dim3(256), 0, st, ws.active, B, ws.n_active);
  Batch b = energy_batch_desc(g, ws, s_dev, B, N, phi_dev, grad_dev, re_dev, im_dev);
  launch_energy(g, b, B, grad_dev != nullptr, st);
  CCO_HIP(hipGetLastError());
  return CCO_OK;
}

int cco_value_grad_density(cco_energy* E, const double* s_host, int B, int N, double* phi_host,
                           double* grad_host, double* re_host, double* im_host) {
  if (!E || !s_host || !phi_host) return fail(CCO_ERR_ARG, "NULL argument");
  if (B < 1 || N < 1) return fail(CCO_ERR_ARG, "need B >= 1 and N >= 1");
  if ((re_host == nullptr) != (im_host == nullptr))
    return fail(CCO_ERR_ARG, "re and im must both be given or both be NULL");
  DevScope ds(E->ctx->device);
  const Geo& g = E->geo;
  const size_t nsd = (size_t)B * N * g.d;
  int rc;
  if ((rc = dev_reserve(E->s, nsd * sizeof(double)))) return rc;
  if ((rc = dev_reserve(E->phi, (size_t)B * sizeof(double)))) return rc;
  if (grad_host && (rc = dev_reserve(E->grad, nsd * sizeof(double)))) return rc;
  if (re_host) {
    if ((rc = dev_reserve(E->re, (size_t)B * g.M_pad * sizeof(double)))) return rc;
    if ((rc = dev_reserve(E->im, (size_t)B * g.M_pad * sizeof(double)))) return rc;
  }
  hipStream_t st = E->ctx->stream;
  CCO_HIP(hipMemcpyAsync(E->s.p, s_host, nsd * sizeof(double), hipMemcpyHostToDevice, st));
  rc = energy_batch(E, (const double*)E->s.p, B, N, (double*)E->phi.p,
                    grad_host ? (double*)E->grad.p : nullptr, re_host ? (double*)E->re.p : nullptr,
                    re_host ? (double*)E->im.p : nullptr);
  if (rc) return rc;
  CCO_HIP(hipMemcpyAsync(phi_host, E->phi.p, (size_t)B * sizeof(double), hipMemcpyDeviceToHost,
                         st));
  if (grad_host)
    CCO_HIP(hipMemcpyAsync(grad_host, E->grad.p, nsd * sizeof(double), hipMemcpyDeviceToHost, st));
  if (re_host) {
    CCO_HIP(hipMemcpyAsync(re_host, E->re.p, (size_t)B * g.M_pad * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    CCO_HIP(hipMemcpyAsync(im_host, E->im.p, (size_t)B * g.M_pad * sizeof(double),
                           hipMemcpyDeviceToHost, st));
  }
  CCO_HIP(hipStreamSynchronize(st));
  return CCO_OK;
}

namespace {
// Per-lane energies and sizes of an uneven batch.
struct UnevenSpec {
  std::vector<cco_energy*> E;
  std::vector<int> N;
};
}  // namespace

static int relax_create_impl(cco_ctx* ctx, cco_energy* E, int B, int N, int d, int method,
                             const cco_descent_settings* ds_, const cco_lbfgs_settings* ls_,
                             const UnevenSpec* un, cco_relax** out);

int cco_relax_create(cco_ctx* ctx, cco_energy* E, int B, int N, int d, int method,
                     const cco_descent_settings* ds_, const cco_lbfgs_settings* ls_,
                     cco_relax** out) {
  return relax_create_impl(ctx, E, B, N, d, method, ds_, ls_, nullptr, out);
}

int cco_relax_uneven_create(cco_ctx* ctx, int n_energies, cco_energy* const* energies, int B,
                            const int32_t* lane_energy, const int32_t* lane_n, int method,
                            const cco_descent_settings* ds_, const cco_lbfgs_settings* ls_,
                            cco_relax** out) {
  if (!ctx || !out || !energies || !lane_energy || !lane_n) return fail(CCO_ERR_ARG, "NULL argument");
  if (n_energies < 1 || B < 1) return fail(CCO_ERR_ARG, "need n_energies >= 1 and B >= 1");
  for (int e = 0; e < n_energies; ++e) {
    if (!energies[e]) return fail(CCO_ERR_ARG, "energy is NULL");
    if (energies[e]->ctx != ctx) return fail(CCO_ERR_ARG, "energies must live on the context");
  }
  UnevenSpec un;
  un.E.resize(B);
  un.N.resize(B);
  for (int i = 0; i < B; ++i) {
    if (lane_energy[i] < 0 || lane_energy[i] >= n_energies)
      return fail(CCO_ERR_ARG, "lane energy index out of range");
    if (lane_n[i] < 1) return fail(CCO_ERR_ARG, "every lane needs N >= 1");
    un.E[i] = energies[lane_energy[i]];
    un.N[i] = lane_n[i];
  }
  return relax_create_impl(ctx, un.E[0], B, un.N[0], un.E[0]->geo.d, method, ds_, ls_, &un, out);
}

static int relax_create_impl(cco_ctx* ctx, cco_energy* E, int B, int N, int d, int method,
                             const cco_descent_settings* ds_, const cco_lbfgs_settings* ls_,
                             const UnevenSpec* un, cco_relax** out) {
  if (!ctx || !out) return fail(CCO_ERR_ARG, "NULL argument");
  if (B < 1 || N < 1) return fail(CCO_ERR_ARG, "need B >= 1 and N >= 1");
  if (d < 1 || d > 3) return fail(CCO_ERR_ARG, "d must be 1, 2 or 3");
  if (E && E->geo.d != d) return fail(CCO_ERR_ARG, "dimension mismatch");
  if (method == CCO_METHOD_DESCENT && !ds_) return fail(CCO_ERR_ARG, "descent settings missing");
  if (method == CCO_METHOD_LBFGS && !ls_) return fail(CCO_ERR_ARG, "L-BFGS settings missing");
  if (method != CCO_METHOD_DESCENT && method != CCO_METHOD_LBFGS)
    return fail(CCO_ERR_ARG, "unknown method");
  if (method == CCO_METHOD_DESCENT && ds_->max_evals < 0)
    return fail(CCO_ERR_ARG, "max_evals must be >= 0");
  if (method == CCO_METHOD_LBFGS) {
    if (ls_->memory < 1 || ls_->memory > 1024) return fail(CCO_ERR_ARG, "memory must be in [1, 1024]");
    if (ls_->n_rescales < 1 || ls_->n_rescales > kMaxRescales)
      return fail(CCO_ERR_ARG, "need 1 to 8 rescales");
    if (ls_->max_linesearch_steps < 1)
      return fail(CCO_ERR_ARG, "max_linesearch_steps must be >= 1");
  }
  DevScope dscope(ctx->device);
  cco_relax* R = new cco_relax();
  R->ctx = ctx;
  R->E = E;
  R->B = B;
  R->N = N;
  R->d = d;
  R->method = method;
  std::memset(&R->dc, 0, sizeof(R->dc));
  std::memset(&R->lc, 0, sizeof(R->lc));
  int cap = 0;
  if (method == CCO_METHOD_DESCENT) {
    const cco_descent_settings& s = *ds_;
    R->dc.step_init = s.step_init;
    R->dc.step_min = s.step_min;
    R->dc.shrink = s.shrink;
    R->dc.grow = s.grow;
    R->dc.rel_tol = s.rel_tol;
    R->dc.energy_tol = s.energy_tol;
    R->dc.max_iters = s.max_evals / 2;
    R->dc.record_every = s.record_every;
    R->dc.record_cap = (s.record_every > 0) ? s.max_records : 0;
    cap = R->dc.record_cap;
  } else {
    const cco_lbfgs_settings& s = *ls_;
    for (int i = 0; i < s.n_rescales; ++i) R->lc.rescales[i] = s.rescales[i];
    R->lc.n_rescales = s.n_rescales;
    R->lc.memory = s.memory;
    R->lc.max_evals_per_pass = s.max_evals_per_pass;
    R->lc.energy_tol = s.energy_tol;
    R->lc.grad_tol = s.grad_tol;
    R->lc.stagnation_patience = s.stagnation_patience;
    R->lc.max_linesearch_steps = s.max_linesearch_steps;
    R->lc.record_every = s.record_every;
    R->lc.record_cap = (s.record_every > 0) ? s.max_records : 0;
    R->lc.mixed_switch = 1e-8;
    cap = R->lc.record_cap;
  }
  RelaxDev& v = R->dev;
  v.B = B;
  v.nd = N * d;
  v.stride = N * d;
  size_t vec = (size_t)B * v.stride;
  // uneven batch: per-lane layouts (vectors packed lane after lane)
  std::vector<long long> h_loff, h_lhoff;
  std::vector<int> h_lnd;
  std::vector<LaneLayout> h_lay;
  std::vector<Geo> h_geo;
  std::vector<int> h_ok;
  size_t part_total = 0, cbox_total = 0, rphi_total = 0;
  if (un) {
    R->uneven = true;
    const int m = method == CCO_METHOD_LBFGS ? ls_->memory : 0;
    h_loff.resize(B);
    h_lhoff.resize(B);
    h_lnd.resize(B);
    h_lay.resize(B);
    h_geo.resize(B);
    h_ok.assign(B, 0);
    long long vt = 0, ht = 0;
    int maxnd = 0;
    for (int i = 0; i < B; ++i) {
      Geo g = un->E[i]->geo;
      g.UT = kWave;  // uneven launches: 64 units per tile, two threads per unit
      g.n_utiles = (g.n_units + kWave - 1) / kWave;
      g.rho_halves = 2;
      const int nd = un->N[i] * g.d;
      const int ns = n_splits_uneven(un->N[i]);
      h_loff[i] = vt;
      h_lhoff[i] = ht;
      h_lnd[i] = nd;
      h_lay[i].off = vt;
      h_lay[i].part_off = (long long)part_total;
      h_lay[i].cbox_off = (long long)cbox_total;
      h_lay[i].rt_off = (long long)rphi_total;
      h_lay[i].N = un->N[i];
      h_lay[i].pad = 0;
      vt += nd;
      ht += (long long)m * nd;
      part_total += (size_t)ns * g.Mbox;
      cbox_total += (size_t)g.Mbox;
      rphi_total += (size_t)g.n_rtiles;
      R->tiles_rho = std::max(R->tiles_rho, g.n_utiles * ns);
      R->tiles_reduce = std::max(R->tiles_reduce, g.n_rtiles);
      R->tiles_force = std::max(R->tiles_force, n_ftiles_for(un->N[i]));
      R->rho_lds = std::max(R->rho_lds, (int)rho_lds_bytes(g));
      R->dmask |= 1 << g.d;
      maxnd = std::max(maxnd, nd);
      h_geo[i] = g;
      // split precision (when asked for at init) only for lanes that fit the
      // one-workgroup-per-lane kernel
      const size_t lds = method == CCO_METHOD_LBFGS ? mp_lane_fit(g, un->N[i]) : 0;
      h_ok[i] = lds > 0 ? 1 : 0;
      if (lds > 0) {
        int& cls = R->lane_lds[g.d][g.RP / 16];
        cls = std::max(cls, (int)lds);
      }
    }
    v.nd = maxnd;
    v.stride = maxnd;
    vec = (size_t)vt;
  }
  R->vec_total = vec;
  int rc = 0;
  auto A = [&](auto*& p, size_t cnt) {
    if (rc == 0) rc = alloc_into(R, p, cnt);
  };
  A(v.status, B);
  A(R->prec_store, B);
  A(v.active, B);
  A(v.n_active, 1);
  A(v.s_eval, vec);
  A(v.g_eval, vec);
  A(v.phi_eval, B);
  A(v.s, vec);
  A(v.g, vec);
  if (method == CCO_METHOD_DESCENT) {
    A(v.dl, B);
  } else {
    const size_t m = R->lc.memory;
    A(v.ll, B);
    A(v.x, vec);
    A(v.px, vec);
    A(v.pg, vec);
    A(v.u, vec);
    A(v.safe_g, vec);
    A(v.S, vec * m);
    A(v.Y, vec * m);
    A(v.rho, (size_t)B * m);
    if (m <= (size_t)kGramMaxMem) {
      v.gram_stride = gram_stride_for((int)m);
      A(v.dir, B);
      A(v.gram, (size_t)B * v.gram_stride);
    }
  }
  if (cap > 0) {
    A(v.tr_s, vec * cap);
    A(v.tr_e, (size_t)B * cap);
    A(v.tr_step, (size_t)B * cap);
    A(v.tr_it, (size_t)B * cap);
  }
  if (rc == 0 && E && !un) {
    rc = dev_reserve(R->part, sizeof(double2) * (size_t)B * n_splits_for(N) * E->geo.Mbox);
    if (rc == 0) rc = dev_reserve(R->cbox, sizeof(double2) * (size_t)B * E->geo.Mbox);
    if (rc == 0) rc = reserve_reduce(R->rphi, B, E->geo);
  }
  if (rc == 0 && un) {
    long long* d_loff = nullptr;
    long long* d_lhoff = nullptr;
    int* d_lnd = nullptr;
    A(d_loff, B);
    A(d_lhoff, B);
    A(d_lnd, B);
    if (rc == 0) rc = dev_reserve(R->part, sizeof(double2) * part_total);
    if (rc == 0) rc = dev_reserve(R->cbox, sizeof(double2) * cbox_total);
    if (rc == 0) rc = dev_reserve(R->rphi, sizeof(double) * rphi_total);
    if (rc == 0) rc = dev_reserve(R->un_lay, sizeof(LaneLayout) * (size_t)B);
    if (rc == 0) rc = dev_reserve(R->un_geo, sizeof(Geo) * (size_t)B);
    const bool any_ok = std::find(h_ok.begin(), h_ok.end(), 1) != h_ok.end();
    if (any_ok) A(R->prec_ok, B);
    if (rc == 0) {
      hipStream_t st = ctx->stream;
      if (hipMemcpyAsync(d_loff, h_loff.data(), sizeof(long long) * B, hipMemcpyHostToDevice, st) !=
              hipSuccess ||
          hipMemcpyAsync(d_lhoff, h_lhoff.data(), sizeof(long long) * B, hipMemcpyHostToDevice,
                         st) != hipSuccess ||
          hipMemcpyAsync(d_lnd, h_lnd.data(), sizeof(int) * B, hipMemcpyHostToDevice, st) !=
              hipSuccess ||
          hipMemcpyAsync(R->un_lay.p, h_lay.data(), sizeof(LaneLayout) * B, hipMemcpyHostToDevice,
                         st) != hipSuccess ||
          hipMemcpyAsync(R->un_geo.p, h_geo.data(), sizeof(Geo) * B, hipMemcpyHostToDevice, st) !=
              hipSuccess ||
          (R->prec_ok != nullptr &&
           hipMemcpyAsync(R->prec_ok, h_ok.data(), sizeof(int) * B, hipMemcpyHostToDevice, st) !=
               hipSuccess) ||
          hipStreamSynchronize(st) != hipSuccess)
        rc = fail(CCO_ERR_DEVICE, "uploading the uneven layout failed");
    }
    v.loff = d_loff;
    v.lhoff = d_lhoff;
    v.lnd = d_lnd;
    v.prec_ok = R->prec_ok;
  }
  if (rc == 0) rc = alloc_host_words(R);
  if (rc) {
    cco_relax_destroy(R);
    return rc;
  }
  *out = R;
  return CCO_OK;
}

int cco_relax_destroy(cco_relax* R) {
  if (!R) return CCO_OK;
  DevScope ds(R->ctx->device);
  (void)hipStreamSynchronize(R->ctx->stream);
  for (const Shard& sh : R->shards) (void)hipStreamSynchronize(sh.stream);
  drop_graphs(R);
  for (DevBuf& b : R->bufs) dev_release(b);
  for (DevBuf& b : R->md_bufs) dev_release(b);
  if (R->trk) cco_energy_destroy(R->trk);
  for (DevBuf* b : {&R->cpl, &R->csig, &R->hist32, &R->part, &R->cbox, &R->rphi, &R->out_s, &R->out_scalars, &R->pack,
                    &R->pack_idx, &R->un_lay, &R->un_geo, &R->ros_S, &R->ros_H, &R->ros_rhs,
                    &R->ros_dts, &R->ros_ipiv, &R->ros_info, &R->ros_cnt})
    dev_release(*b);
  if (R->host_count) (void)hipHostFree(R->host_count);
  if (R->progress_host) (void)hipHostFree(R->progress_host);
  for (hipStream_t s : R->extra_streams)
    if (s != nullptr) (void)hipStreamDestroy(s);
  if (R->fork_ev) (void)hipEventDestroy(R->fork_ev);
  for (hipEvent_t e : R->evs) (void)hipEventDestroy(e);
  delete R;
  return CCO_OK;
}

int cco_relax_init_dev(cco_relax* R, const double* s0_dev) {
  if (!R || !s0_dev) return fail(CCO_ERR_ARG, "NULL argument");
  if (R->method > CCO_METHOD_LBFGS) return fail(CCO_ERR_STATE, "not a relax session");
  DevScope ds(R->ctx->device);
  hipStream_t st = R->ctx->stream;
  const size_t bytes = sizeof(double) * R->vec_total;
  CCO_HIP(hipMemcpyAsync(R->dev.s, s0_dev, bytes, hipMemcpyDeviceToDevice, st));
  CCO_HIP(hipMemcpyAsync(R->dev.s_eval, s0_dev, bytes, hipMemcpyDeviceToDevice, st));
  R->variant_run = resolve_variant(R);
  R->dev.prec = resolve_prec(R);
  if (int rc_mp = reserve_mp_planes(R)) return rc_mp;
  hipLaunchKernelGGL(relax_init_kernel, dim3((R->B + 255) / 256), dim3(256), 0, st, R->dev,
                     R->method);
  CCO_HIP(hipGetLastError());
  R->slots_ub = R->B;
  if (int rc_sh = setup_shards(R)) return rc_sh;
  R->lane_evals = 0;
  R->ready = true;
  return CCO_OK;
}

int cco_relax_init(cco_relax* R, const double* s0_host) {
  if (!R || !s0_host) return fail(CCO_ERR_ARG, "NULL argument");
  if (R->method > CCO_METHOD_LBFGS) return fail(CCO_ERR_STATE, "not a relax session");
  DevScope ds(R->ctx->device);
  hipStream_t st = R->ctx->stream;
  const size_t bytes = sizeof(double) * R->vec_total;
  CCO_HIP(hipMemcpyAsync(R->dev.s, s0_host, bytes, hipMemcpyHostToDevice, st));
  CCO_HIP(hipMemcpyAsync(R->dev.s_eval, R->dev.s, bytes, hipMemcpyDeviceToDevice, st));
  R->variant_run = resolve_variant(R);
  R->dev.prec = resolve_prec(R);
  if (int rc_mp = reserve_mp_planes(R)) return rc_mp;
  hipLaunchKernelGGL(relax_init_kernel, dim3((R->B + 255) / 256), dim3(256), 0, st, R->dev,
                     R->method);
  CCO_HIP(hipGetLastError());
  CCO_HIP(hipStreamSynchronize(st));
  R->slots_ub = R->B;
  if (int rc_sh = setup_shards(R)) return rc_sh;
  R->lane_evals = 0;
  R->ready = true;
  return CCO_OK;
}

namespace {

// Launch bound of a graph: the shard's bound rounded up to a power of two (at
// most the shard's lanes), so a run instantiates a handful of graphs per shard
// while its lanes stop; surplus workgroups exit at once, as for any bound.
int graph_bucket(const Shard& sh) {
  int b = 1;
  while (b < sh.slots_ub) b <<= 1;
  return std::min(b, sh.count);
}

bool graph_eligible(const cco_relax* R) {
  return R->graph_steps > 0 && R->E != nullptr &&
         (R->method == CCO_METHOD_DESCENT || R->method == CCO_METHOD_LBFGS);
}

// The HIP graph of R->graph_steps evaluation steps of shard k at launch bound
// ub: the stage launches of each step, captured from the shard's stream.  The
// compaction kernel tags the progress word from the shard's device step
// counter, so every replay reports its own step numbers.
int step_graph(cco_relax* R, int k, int ub, hipGraphExec_t* out) {
  const auto key = std::make_pair(k, ub);
  const auto it = R->graphs.find(key);
  if (it != R->graphs.end()) {
    *out = it->second;
    return CCO_OK;
  }
  Shard sh = R->shards[k];
  sh.slots_ub = ub;
  int seq[kStages];
  const int ns = step_stages(R, seq);
  CCO_HIP(hipStreamBeginCapture(sh.stream, hipStreamCaptureModeRelaxed));
  for (int g = 0; g < R->graph_steps; ++g)
    for (int q = 0; q < ns; ++q) relax_stage(R, sh, seq[q]);
  hipGraph_t graph = nullptr;
  CCO_HIP(hipStreamEndCapture(sh.stream, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(CCO_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  R->graphs[key] = exec;
  R->graph_builds += 1;
  *out = exec;
  return CCO_OK;
}

}  // namespace

int cco_relax_step(cco_relax* R, int64_t n_evals, int* n_active) {
  if (!R) return fail(CCO_ERR_ARG, "relax is NULL");
  if (!R->ready) return fail(CCO_ERR_STATE, "session not initialised");
  if (!R->E) return fail(CCO_ERR_STATE, "external-energy session: use pending/submit");
  DevScope ds(R->ctx->device);
  const int nsh = (int)R->shards.size();
  if (nsh > 1) {  // the other shards' streams start after the work queued on the context stream
    CCO_HIP(hipEventRecord(R->fork_ev, R->ctx->stream));
    for (int k = 1; k < nsh; ++k) CCO_HIP(hipStreamWaitEvent(R->shards[k].stream, R->fork_ev, 0));
  }
  constexpr int kEv = kStages + 1;
  int64_t first[kMaxShards];
  bool live[kMaxShards], force[kMaxShards];
  for (int k = 0; k < nsh; ++k) {
    first[k] = R->shards[k].steps;
    live[k] = R->shards[k].slots_ub > 0;
    force[k] = false;
  }
  std::vector<int> seq_ids;  // per profiled step: its stage ids, then -1
  size_t n_prof = 0;
  // Round robin over the shards: queue a step on every shard that has fewer
  // than kRunAhead steps in flight; stop a shard when its lanes have all
  // stopped or it has queued n_evals steps.  When no shard can queue, wait;
  // a stream that went idle without its word catching up may queue one step.
  for (int64_t idle = 0;;) {
    bool pending = false, queued = false;
    for (int k = 0; k < nsh; ++k) {
      if (!live[k]) continue;
      Shard& sh = R->shards[k];
      if (sh.steps - first[k] >= n_evals) {
        live[k] = false;
        continue;
      }
      const int64_t in_flight = shard_in_flight(sh, first[k]);
      if (sh.slots_ub == 0) {
        live[k] = false;
        continue;
      }
      pending = true;
      if (in_flight >= kRunAhead && !force[k]) continue;
      force[k] = false;
      const bool prof = R->profile_every > 0 && sh.steps % R->profile_every == 0;
      if (!prof && graph_eligible(R) && n_evals - (sh.steps - first[k]) >= R->graph_steps) {
        hipGraphExec_t exec = nullptr;
        if (int rc = step_graph(R, k, graph_bucket(sh), &exec)) return rc;
        CCO_HIP(hipGraphLaunch(exec, sh.stream));
        sh.steps += R->graph_steps;
        R->launches += R->graph_steps;
        R->graph_launches += 1;
        queued = true;
        continue;
      }
      while (prof && R->evs.size() < kEv * (n_prof + 1)) {
        hipEvent_t ev;
        CCO_HIP(hipEventCreate(&ev));
        R->evs.push_back(ev);
      }
      int seq[kStages];
      const int ns = step_stages(R, seq);
      for (int q = 0; q < ns; ++q) {
        if (prof && q == 0) CCO_HIP(hipEventRecord(R->evs[kEv * n_prof], sh.stream));
        relax_stage(R, sh, seq[q]);
        if (prof) {
          CCO_HIP(hipEventRecord(R->evs[kEv * n_prof + q + 1], sh.stream));
          seq_ids.push_back(seq[q]);
        }
      }
      if (prof) {
        seq_ids.push_back(-1);
        ++n_prof;
      }
      CCO_HIP(hipGetLastError());
      sh.steps += 1;
      R->launches += 1;
      queued = true;
    }
    if (!pending) break;
    if (queued) {
      idle = 0;
      continue;
    }
    if ((++idle & 1023) == 0) {
      for (int k = 0; k < nsh; ++k) {
        if (!live[k]) continue;
        const hipError_t q = hipStreamQuery(R->shards[k].stream);
        if (q == hipSuccess) force[k] = true;
        else if (q != hipErrorNotReady) return fail(CCO_ERR_DEVICE, hipGetErrorString(q));
      }
    }
    std::this_thread::yield();
  }
  for (int k = 0; k < nsh; ++k) CCO_HIP(hipStreamSynchronize(R->shards[k].stream));
  CCO_HIP(hipMemcpy(R->host_count, R->shard_counts, sizeof(int) * nsh, hipMemcpyDeviceToHost));
  int total = 0;
  for (int k = 0; k < nsh; ++k) {
    Shard& sh = R->shards[k];
    sh.slots_ub = R->host_count[k];
    total += sh.slots_ub;
    R->steps = std::max(R->steps, sh.steps);
  }
  R->slots_ub = total;
  if (n_active) *n_active = total;
  size_t p = 0;
  for (size_t i = 0; i < n_prof; ++i) {
    for (int q = 0; p < seq_ids.size() && seq_ids[p] >= 0; ++q, ++p) {
      float ms = 0.0f;
      CCO_HIP(hipEventElapsedTime(&ms, R->evs[kEv * i + q], R->evs[kEv * i + q + 1]));
      R->stage_ms[seq_ids[p]] += ms;
      R->stage_n[seq_ids[p]] += 1;
    }
    ++p;  // step separator
  }
  return CCO_OK;
}

int cco_relax_set_precision(cco_relax* R, int mode, double switch_phi) {
  if (!R) return fail(CCO_ERR_ARG, "relax is NULL");
  if (mode != 0 && mode != 1) return fail(CCO_ERR_ARG, "precision mode must be 0 (FP64) or 1 (split)");
  if (mode == 1 && R->method != CCO_METHOD_LBFGS && R->method != METHOD_MD && R->method != METHOD_FLOW)
    return fail(CCO_ERR_ARG, "split precision is available for L-BFGS, gradient-flow and MD sessions");
  if (!(switch_phi >= 0.0)) return fail(CCO_ERR_ARG, "switch_phi must be >= 0");
  R->precision = mode;
  R->lc.mixed_switch = switch_phi;
  R->fc.mixed_switch = switch_phi;
  return CCO_OK;
}

int cco_relax_precision_state(cco_relax* R, int* active_mode, int* lanes_split) {
  if (!R || !active_mode || !lanes_split) return fail(CCO_ERR_ARG, "NULL argument");
  DevScope ds(R->ctx->device);
  *active_mode = R->dev.prec != nullptr ? 1 : 0;
  *lanes_split = 0;
  if (R->dev.prec == nullptr) return CCO_OK;
  std::vector<int> h(R->B);
  CCO_HIP(hipMemcpyAsync(h.data(), R->dev.prec, sizeof(int) * R->B, hipMemcpyDeviceToHost,
                         R->ctx->stream));
  CCO_HIP(hipStreamSynchronize(R->ctx->stream));
  for (int v : h) *lanes_split += v != 0;
  return CCO_OK;
}

int cco_energy_set_option(cco_energy* E, int key, int64_t value) {
  if (!E) return fail(CCO_ERR_ARG, "energy is NULL");
  if (key == CCO_EOPT_RHO_CHUNK) {
    if (value != 16 && value != 32 && value != 64)
      return fail(CCO_ERR_ARG, "rho chunk must be 16, 32 or 64 particles");
    Geo g = E->geo;
    g.rho_chunk = (int)value;
    if (rho_lds_bytes(g) > kRhoLdsMax)
      return fail(CCO_ERR_ARG, "rho chunk too long for this k-set's LDS tables");
    E->geo.rho_chunk = (int)value;
    return CCO_OK;
  }
  if (key == CCO_EOPT_RHO_SPLIT) {
    if (value < 0 || value > 65536 || value % kSplitLen != 0)
      return fail(CCO_ERR_ARG, "rho split must be 0 or a multiple of 128 up to 65536 particles");
    E->geo.rho_split = (int)value;
    return CCO_OK;
  }
  if (key == CCO_EOPT_MP_LAYOUT) {
    if (value != 0 && value != 1) return fail(CCO_ERR_ARG, "split-precision layout option must be 0 or 1");
    use_pair_layout(E, value == 1);
    return CCO_OK;
  }
  if (key == CCO_EOPT_MP_LANE) {
    if (value != 0 && value != 1) return fail(CCO_ERR_ARG, "split-precision lane kernel option must be 0 or 1");
    E->geo.mp_lane = (value == 1 && E->geo.mp_ok && mp_lane_tiles_ok(E->geo)) ? 1 : 0;
    return CCO_OK;
  }
  return fail(CCO_ERR_ARG, "unknown energy option key");
}

int cco_energy_get_option(const cco_energy* E, int key, int64_t* value) {
  if (!E || !value) return fail(CCO_ERR_ARG, "NULL argument");
  if (key == CCO_EOPT_RHO_CHUNK) {
    *value = E->geo.rho_chunk;
    return CCO_OK;
  }
  if (key == CCO_EOPT_MP_LANE) {
    *value = E->geo.mp_lane;
    return CCO_OK;
  }
  if (key == CCO_EOPT_RHO_SPLIT) {
    *value = E->geo.rho_split;
    return CCO_OK;
  }
  if (key == CCO_EOPT_MP_LAYOUT) {
    *value = E->geo.mp_geo != nullptr ? 1 : 0;
    return CCO_OK;
  }
  return fail(CCO_ERR_ARG, "unknown energy option key");
}

int cco_energy_set_precision(cco_energy* E, int mode, int* applied) {
  if (!E) return fail(CCO_ERR_ARG, "energy is NULL");
  if (mode != 0 && mode != 1) return fail(CCO_ERR_ARG, "precision mode must be 0 (FP64) or 1 (split)");
  E->precision = mode;
  if (applied) *applied = (mode == 1 && E->geo.mp_ok) ? 1 : 0;
  return CCO_OK;
}

int cco_relax_get_option(cco_relax* R, int key, int64_t* value) {
  if (!R || !value) return fail(CCO_ERR_ARG, "NULL argument");
  if (key == CCO_OPT_ADVANCE_VARIANT) {
    *value = R->variant_run;
    return CCO_OK;
  }
  if (key == CCO_OPT_SHARDS) {
    *value = (int64_t)R->shards.size();
    return CCO_OK;
  }
  if (key == CCO_OPT_GRAPH) {
    *value = R->graph_steps;
    return CCO_OK;
  }
  if (key == CCO_OPT_GRAPH_LAUNCHES) {
    *value = R->graph_launches;
    return CCO_OK;
  }
  if (key == CCO_OPT_SPLIT_HISTORY) {
    *value = R->split_history;
    return CCO_OK;
  }
  return fail(CCO_ERR_ARG, "unknown option key");
}

int cco_relax_set_option(cco_relax* R, int key, int64_t value) {
  if (!R) return fail(CCO_ERR_ARG, "relax is NULL");
  if (key == CCO_OPT_ADVANCE_VARIANT) {
    if (value < 0 || value > 3) return fail(CCO_ERR_ARG, "advance variant must be 0, 1, 2 or 3");
    if (value == 2 && lbfgs_register_ept(R->dev.nd) == 0)
      return fail(CCO_ERR_ARG, "register-resident advance needs N*d <= 8192");
    if (value == 3 && (R->dev.gram == nullptr || R->uneven))
      return fail(CCO_ERR_ARG,
                  "Gram-form advance needs a uniform L-BFGS session with memory <= 128");
    R->advance_variant = (int)value;
    return CCO_OK;
  }
  if (key == CCO_OPT_SHARDS) {
    if (value < 0 || value > kMaxShards) return fail(CCO_ERR_ARG, "shards must be 0 (auto) to 4");
    R->shard_req = (int)value;
    return CCO_OK;
  }
  if (key == CCO_OPT_GRAPH) {
    if (value < 0 || value > 64) return fail(CCO_ERR_ARG, "graph steps must be 0 (off) to 64");
    if (value != R->graph_steps) drop_graphs(R);
    R->graph_steps = (int)value;
    return CCO_OK;
  }
  if (key == CCO_OPT_SPLIT_HISTORY) {
    if (value != 0 && value != 1) return fail(CCO_ERR_ARG, "split history must be 0 or 1");
    R->split_history = (int)value;
    return CCO_OK;
  }
  return fail(CCO_ERR_ARG, "unknown option key");
}

int cco_relax_counters(cco_relax* R, int64_t* out4) {
  if (!R || !out4) return fail(CCO_ERR_ARG, "NULL argument");
  DevScope ds(R->ctx->device);
  const size_t B = R->B;
  int64_t ev = 0, it = 0, pairs = 0;
  if (R->method == METHOD_FLOW) {
    std::vector<FlowLane> L(B);
    CCO_HIP(hipMemcpy(L.data(), R->dev.fl, sizeof(FlowLane) * B, hipMemcpyDeviceToHost));
    for (auto& l : L) {
      ev += l.n_evals;
      it += l.n_tries;
    }
  } else if (R->method == METHOD_MD) {
    std::vector<MdLane> L(B);
    CCO_HIP(hipMemcpy(L.data(), R->dev.ml, sizeof(MdLane) * B, hipMemcpyDeviceToHost));
    for (auto& l : L) {
      ev += l.step + 1;
      it += l.step;
    }
  } else if (R->method == CCO_METHOD_DESCENT) {
    std::vector<DescentLane> L(B);
    CCO_HIP(hipMemcpy(L.data(), R->dev.dl, sizeof(DescentLane) * B, hipMemcpyDeviceToHost));
    for (auto& l : L) {
      ev += l.it + 1;
      it += l.it;
    }
  } else {
    std::vector<LbfgsLane> L(B);
    CCO_HIP(hipMemcpy(L.data(), R->dev.ll, sizeof(LbfgsLane) * B, hipMemcpyDeviceToHost));
    for (auto& l : L) {
      ev += l.n_evals;
      it += l.it;
      pairs += l.pairs;
    }
  }
  out4[0] = ev;
  out4[1] = it;
  out4[2] = pairs;
  out4[3] = R->launches;
  return CCO_OK;
}

int cco_relax_set_profiling(cco_relax* R, int every) {
  if (!R) return fail(CCO_ERR_ARG, "relax is NULL");
  if (every < 0) return fail(CCO_ERR_ARG, "profiling period must be >= 0");
  R->profile_every = every;
  for (int s = 0; s < kStages; ++s) {
    R->stage_ms[s] = 0.0;
    R->stage_n[s] = 0;
  }
  return CCO_OK;
}

int cco_relax_profile(cco_relax* R, double* ms_sum, int64_t* launches) {
  if (!R || !ms_sum || !launches) return fail(CCO_ERR_ARG, "NULL argument");
  for (int s = 0; s < kStages; ++s) {
    ms_sum[s] = R->stage_ms[s];
    launches[s] = R->stage_n[s];
  }
  return CCO_OK;
}

int cco_energy_profile(cco_energy* E, const double* s_dev, int B, int N, int reps,
                       double* avg_ms) {
  if (!E || !s_dev || !avg_ms || reps < 1 || B < 1 || N < 1)
    return fail(CCO_ERR_ARG, "bad argument");
  DevScope ds(E->ctx->device);
  const Geo& g = E->geo;
  int rc;
  const size_t nsd = (size_t)B * N * g.d;
  if ((rc = dev_reserve(E->phi, (size_t)B * sizeof(double)))) return rc;
  if ((rc = dev_reserve(E->grad, nsd * sizeof(double)))) return rc;
  // one untimed evaluation sets up the workspace and the lane list
  rc = energy_batch(E, s_dev, B, N, (double*)E->phi.p, (double*)E->grad.p, nullptr, nullptr);
  if (rc) return rc;
  Batch b;
  b.active = (const int*)E->active.p;
  b.n_active = (const int*)E->n_active.p;
  b.N = N;
  b.lane_stride = N * g.d;
  b.n_splits = n_splits_of(g, N);
  b.split_len = split_len_of(g, N);
  b.n_ftiles = n_ftiles_for(N);
  b.s = s_dev;
  b.grad = (double*)E->grad.p;
  b.phi = (double*)E->phi.p;
  b.re = nullptr;
  b.im = nullptr;
  b.part = (double2*)E->part.p;
  b.cbox = (double2*)E->cbox.p;
  b.phi_part = (double*)E->rphi.p;
  b.prec = (E->precision == 1 && g.mp_ok) ? (const int*)E->prec.p : nullptr;
  mp_bind(b, g, E->cpl.p, B);
  b.csig = (double*)E->csig.p;
  hipStream_t st = E->ctx->stream;
  std::vector<hipEvent_t> ev(4 * (size_t)reps);
  for (auto& e : ev) CCO_HIP(hipEventCreate(&e));
  for (int r = 0; r < reps; ++r) {
    CCO_HIP(hipEventRecord(ev[4 * r], st));
    for (int stage = 0; stage < 3; ++stage) {
      launch_energy_stage(stage, g, b, B, st);
      CCO_HIP(hipEventRecord(ev[4 * r + stage + 1], st));
    }
  }
  CCO_HIP(hipGetLastError());
  CCO_HIP(hipStreamSynchronize(st));
  double acc[4] = {0, 0, 0, 0};
  for (int r = 0; r < reps; ++r) {
    for (int stage = 0; stage < 3; ++stage) {
      float ms = 0.0f;
      CCO_HIP(hipEventElapsedTime(&ms, ev[4 * r + stage], ev[4 * r + stage + 1]));
      acc[stage] += ms;
    }
    float tot = 0.0f;
    CCO_HIP(hipEventElapsedTime(&tot, ev[4 * r], ev[4 * r + 3]));
    acc[3] += tot;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (int i = 0; i < 4; ++i) avg_ms[i] = acc[i] / reps;
  return CCO_OK;
}

int cco_copy_dev(cco_ctx* ctx, void* dst_dev, const void* src_dev, size_t bytes, double* ms) {
  if (!ctx || !dst_dev || !src_dev) return fail(CCO_ERR_ARG, "NULL argument");
  if (bytes % 32 != 0) return fail(CCO_ERR_ARG, "bytes must be a multiple of 32");
  DevScope ds(ctx->device);
  hipEvent_t a, b;
  CCO_HIP(hipEventCreate(&a));
  CCO_HIP(hipEventCreate(&b));
  CCO_HIP(hipEventRecord(a, ctx->stream));
  launch_copy(src_dev, dst_dev, bytes, ctx->stream);
  CCO_HIP(hipEventRecord(b, ctx->stream));
  CCO_HIP(hipGetLastError());
  CCO_HIP(hipStreamSynchronize(ctx->stream));
  float t = 0.0f;
  CCO_HIP(hipEventElapsedTime(&t, a, b));
  if (ms) *ms = t;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return CCO_OK;
}

namespace {

__global__ void set_int_kernel(int* p, int v) { *p = v; }

// Gradient flow by ROS2 (flow.py:367-398): after the first evaluation (the
// flow's init step), each round attempts one step for every running lane,
// in chunks of lanes whose (N d)^2 factors fit kRosBytes: Hessian, W in
// place, LU, k1, v(s + h k1), k2, v(candidate), the error solve and the step
// control, then one compaction and one sync.  Two evaluations per round.
constexpr size_t kRosBytes = (size_t)2 << 30;

int flow_run_ros2(cco_relax* R, int64_t max_evals) {
  if (!R->ready) return fail(CCO_ERR_STATE, "session not initialised");
  DevScope ds(R->ctx->device);
  int na = R->slots_ub;
  int rc;
  if (R->steps == 0 && (rc = cco_relax_step(R, 1, &na))) return rc;
  cco_ctx* ctx = R->ctx;
  hipStream_t st = ctx->stream;
  if (ctx->blas == nullptr && (ctx->blas = blas_create()) == nullptr)
    return fail(CCO_ERR_DEVICE, "rocblas_create_handle failed");
  const int nd = R->dev.nd;
  const size_t mat = (size_t)nd * nd;
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>(R->B, kRosBytes / (mat * sizeof(double))));
  if ((rc = dev_reserve(R->ros_S, sizeof(double) * chunk * (size_t)nd))) return rc;
  if ((rc = dev_reserve(R->ros_H, sizeof(double) * chunk * mat))) return rc;
  if ((rc = dev_reserve(R->ros_rhs, sizeof(double) * chunk * (size_t)nd))) return rc;
  if ((rc = dev_reserve(R->ros_dts, sizeof(double) * chunk))) return rc;
  if ((rc = dev_reserve(R->ros_ipiv, sizeof(int) * chunk * (size_t)nd))) return rc;
  if ((rc = dev_reserve(R->ros_info, sizeof(int) * chunk))) return rc;
  if ((rc = dev_reserve(R->ros_cnt, sizeof(int)))) return rc;
  double* S = (double*)R->ros_S.p;
  double* H = (double*)R->ros_H.p;
  double* rhs = (double*)R->ros_rhs.p;
  double* dts = (double*)R->ros_dts.p;
  int* ipiv = (int*)R->ros_ipiv.p;
  int* info = (int*)R->ros_info.p;
  int* cnt = (int*)R->ros_cnt.p;
  while (na > 0) {
    if (R->steps >= max_evals)
      return fail(CCO_ERR_BOUND, "flow did not finish within the round bound");
    for (int c0 = 0; c0 < na; c0 += chunk) {
      const int cn = std::min(chunk, na - c0);
      hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, st, cnt, cn);
      RelaxDev dv = R->dev;
      dv.active = R->dev.active + c0;
      dv.n_active = cnt;
      dv.progress = nullptr;
      const Batch eb = relax_batch(R, dv);
      launch_ros_gather(dv, R->fc, cn, S, dts, st);
      if ((rc = hessian_on(R->E, st, S, cn, R->N, H))) return rc;
      launch_ros_build(H, dts, cnt, R->fc, nd, cn, st);
      if (lu_factor_batched(ctx->blas, st, H, nd, ipiv, info, cn))
        return fail(CCO_ERR_DEVICE, "rocsolver getrf failed");
      for (int q = 0; q < 3; ++q) {
        launch_ros_rhs(dv, R->fc, rhs, dts, q, cn, st);
        if (lu_solve_batched(ctx->blas, st, H, nd, ipiv, rhs, cn))
          return fail(CCO_ERR_DEVICE, "rocsolver getrs failed");
        if (q == 2) break;
        launch_ros_point(dv, rhs, dts, q, cn, st);
        for (int stage = STG_RHO; stage <= STG_FORCE; ++stage)
          launch_energy_stage(stage, R->E->geo, eb, cn, st);
      }
      launch_ros_finish(dv, R->fc, rhs, dts, cn, st);
      CCO_HIP(hipGetLastError());
    }
    R->steps += 2;
    R->launches += 2;
    launch_compact(whole_view(R).dev, st, 0, R->B, 0u);
    if ((rc = relax_sync_count(R, &na))) return rc;
  }
  return CCO_OK;
}

}  // namespace

int cco_relax_run(cco_relax* R, int64_t max_evals, int sync_every) {
  if (!R) return fail(CCO_ERR_ARG, "relax is NULL");
  if (sync_every < 1) return fail(CCO_ERR_ARG, "sync_every must be >= 1");
  if (R->method == METHOD_FLOW && R->ros2) return flow_run_ros2(R, max_evals);
  int n_active = R->slots_ub;
  while (n_active > 0) {
    if (R->steps >= max_evals)
      return fail(CCO_ERR_BOUND, "relax did not finish within the round bound");
    // each call stops queueing steps as soon as the device reports that no
    // lane runs, so the chunk length only bounds the time between syncs
    const int64_t chunk = std::min<int64_t>(sync_every, max_evals - R->steps);
    int rc = cco_relax_step(R, chunk, &n_active);
    if (rc) return rc;
  }
  return CCO_OK;
}

int cco_relax_pending(cco_relax* R, int* n_active, int32_t* lanes_host, double* s_trial_host) {
  if (!R || !n_active) return fail(CCO_ERR_ARG, "NULL argument");
  if (!R->ready) return fail(CCO_ERR_STATE, "relax not initialised");
  if (R->uneven) return fail(CCO_ERR_STATE, "uneven batches use the built-in energies");
  DevScope ds(R->ctx->device);
  int rc = relax_sync_count(R, n_active);
  if (rc) return rc;
  const int na = *n_active;
  if (na == 0) return CCO_OK;
  hipStream_t st = R->ctx->stream;
  if (lanes_host)
    CCO_HIP(hipMemcpyAsync(lanes_host, R->dev.active, sizeof(int) * na, hipMemcpyDeviceToHost, st));
  if (s_trial_host) {
    const int nd = R->dev.nd;
    if ((rc = dev_reserve(R->pack, sizeof(double) * (size_t)R->B * nd))) return rc;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(na), dim3(256), 0, st, R->dev.s_eval,
                       R->dev.active, R->dev.n_active, (double*)R->pack.p, nd, R->dev.stride);
    CCO_HIP(hipGetLastError());
    CCO_HIP(hipMemcpyAsync(s_trial_host, R->pack.p, sizeof(double) * (size_t)na * nd,
                           hipMemcpyDeviceToHost, st));
  }
  CCO_HIP(hipStreamSynchronize(st));
  return CCO_OK;
}

int cco_relax_submit(cco_relax* R, const double* phi_host, const double* grad_host) {
  if (!R || !phi_host || !grad_host) return fail(CCO_ERR_ARG, "NULL argument");
  if (!R->ready) return fail(CCO_ERR_STATE, "relax not initialised");
  DevScope ds(R->ctx->device);
  const int na = R->slots_ub;
  if (na == 0) return fail(CCO_ERR_STATE, "no lane is awaiting an evaluation");
  hipStream_t st = R->ctx->stream;
  const int nd = R->dev.nd;
  int rc;
  if ((rc = dev_reserve(R->pack, sizeof(double) * (size_t)R->B * nd))) return rc;
  if ((rc = dev_reserve(R->pack_idx, sizeof(double) * (size_t)R->B))) return rc;
  CCO_HIP(hipMemcpyAsync(R->pack.p, grad_host, sizeof(double) * (size_t)na * nd,
                         hipMemcpyHostToDevice, st));
  CCO_HIP(hipMemcpyAsync(R->pack_idx.p, phi_host, sizeof(double) * na, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(scatter_eval_kernel, dim3(na), dim3(256), 0, st, (const double*)R->pack_idx.p,
                     (const double*)R->pack.p, R->dev.active, R->dev.n_active, R->dev.phi_eval,
                     R->dev.g_eval, nd, R->dev.stride);
  CCO_HIP(hipGetLastError());
  rc = relax_advance(R);
  if (rc) return rc;
  R->steps += 1;
  R->launches += 1;
  CCO_HIP(hipStreamSynchronize(st));
  return CCO_OK;
}

int cco_relax_result_dev(cco_relax* R, double* s_dev, double* energy_dev, double* grad_norm_dev,
                         int64_t* n_evals_dev, int64_t* n_iters_dev, int64_t* status_dev) {
  if (!R || !s_dev || !energy_dev || !grad_norm_dev || !n_evals_dev || !n_iters_dev || !status_dev)
    return fail(CCO_ERR_ARG, "NULL argument");
  if (!R->ready) return fail(CCO_ERR_STATE, "relax not initialised");
  if (R->method > CCO_METHOD_LBFGS) return fail(CCO_ERR_STATE, "not a relax session");
  DevScope ds(R->ctx->device);
  RelaxOutDev o;
  o.s = s_dev;
  o.energy = energy_dev;
  o.grad_norm = grad_norm_dev;
  o.n_evals = (long long*)n_evals_dev;
  o.n_iters = (long long*)n_iters_dev;
  o.status = (long long*)status_dev;
  launch_finalize(R->dev, R->method, R->lc, o, R->ctx->stream);
  CCO_HIP(hipGetLastError());
  return CCO_OK;
}

int cco_relax_result(cco_relax* R, double* s_host, double* energy_host, double* grad_norm_host,
                     int64_t* n_evals_host, int64_t* n_iters_host, int64_t* status_host) {
  if (!R || !s_host || !energy_host || !grad_norm_host || !n_evals_host || !n_iters_host ||
      !status_host)
    return fail(CCO_ERR_ARG, "NULL argument");
  DevScope ds(R->ctx->device);
  const size_t vec = R->vec_total;
  int rc;
  if ((rc = dev_reserve(R->out_s, sizeof(double) * vec))) return rc;
  if ((rc = dev_reserve(R->out_scalars, sizeof(double) * 6 * (size_t)R->B))) return rc;
  double* sc = (double*)R->out_scalars.p;
  const size_t B = R->B;
  rc = cco_relax_result_dev(R, (double*)R->out_s.p, sc, sc + B, (int64_t*)(sc + 2 * B),
                            (int64_t*)(sc + 3 * B), (int64_t*)(sc + 4 * B));
  if (rc) return rc;
  hipStream_t st = R->ctx->stream;
  CCO_HIP(hipMemcpyAsync(s_host, R->out_s.p, sizeof(double) * vec, hipMemcpyDeviceToHost, st));
  CCO_HIP(hipMemcpyAsync(energy_host, sc, sizeof(double) * B, hipMemcpyDeviceToHost, st));
  CCO_HIP(hipMemcpyAsync(grad_norm_host, sc + B, sizeof(double) * B, hipMemcpyDeviceToHost, st));
  CCO_HIP(hipMemcpyAsync(n_evals_host, sc + 2 * B, sizeof(int64_t) * B, hipMemcpyDeviceToHost, st));
  CCO_HIP(hipMemcpyAsync(n_iters_host, sc + 3 * B, sizeof(int64_t) * B, hipMemcpyDeviceToHost, st));
  CCO_HIP(hipMemcpyAsync(status_host, sc + 4 * B, sizeof(int64_t) * B, hipMemcpyDeviceToHost, st));
  CCO_HIP(hipStreamSynchronize(st));
  return CCO_OK;
}

int cco_relax_trajectory(cco_relax* R, int64_t* count_host, int64_t* iteration_host,
                         double* s_host, double* energy_host, double* step_host) {
  if (!R || !count_host) return fail(CCO_ERR_ARG, "NULL argument");
  if (R->method > CCO_METHOD_LBFGS) return fail(CCO_ERR_STATE, "not a relax session");
  DevScope ds(R->ctx->device);
  const int cap = R->method == CCO_METHOD_DESCENT ? R->dc.record_cap : R->lc.record_cap;
  hipStream_t st = R->ctx->stream;
  const size_t B = R->B;
  if (R->method == CCO_METHOD_DESCENT) {
    std::vector<DescentLane> L(B);
    CCO_HIP(hipMemcpyAsync(L.data(), R->dev.dl, sizeof(DescentLane) * B, hipMemcpyDeviceToHost, st));
    CCO_HIP(hipStreamSynchronize(st));
    for (size_t i = 0; i < B; ++i) count_host[i] = L[i].rec_count;
  } else {
    std::vector<LbfgsLane> L(B);
    CCO_HIP(hipMemcpyAsync(L.data(), R->dev.ll, sizeof(LbfgsLane) * B, hipMemcpyDeviceToHost, st));
    CCO_HIP(hipStreamSynchronize(st));
    for (size_t i = 0; i < B; ++i) count_host[i] = L[i].rec_count;
  }
  if (cap == 0) return CCO_OK;
  const size_t rows = B * cap;
  if (iteration_host)
    CCO_HIP(hipMemcpyAsync(iteration_host, R->dev.tr_it, sizeof(int64_t) * rows,
                           hipMemcpyDeviceToHost, st));
  if (s_host)
    CCO_HIP(hipMemcpyAsync(s_host, R->dev.tr_s, sizeof(double) * cap * R->vec_total,
                           hipMemcpyDeviceToHost, st));
  if (energy_host)
    CCO_HIP(hipMemcpyAsync(energy_host, R->dev.tr_e, sizeof(double) * rows, hipMemcpyDeviceToHost, st));
  if (step_host)
    CCO_HIP(hipMemcpyAsync(step_host, R->dev.tr_step, sizeof(double) * rows, hipMemcpyDeviceToHost,
                           st));
  CCO_HIP(hipStreamSynchronize(st));
  return CCO_OK;
}

int cco_relax_stats(cco_relax* R, int64_t* steps, int64_t* lane_evals) {
  if (!R) return fail(CCO_ERR_ARG, "relax is NULL");
  if (steps) *steps = R->steps;
  if (lane_evals) *lane_evals = R->lane_evals;
  return CCO_OK;
}


// ---------------------------------------------------------------------------
// gradient flow and velocity Verlet sessions
// ---------------------------------------------------------------------------
static bool invert3(const double* a, int d, double* out) {
  // Gauss-Jordan with partial pivoting on a d x d (d <= 3) row-major matrix
  double m[3][6];
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < 2 * d; ++j) m[i][j] = j < d ? a[i * d + j] : (j - d == i ? 1.0 : 0.0);
  for (int c = 0; c < d; ++c) {
    int p = c;
    for (int r = c + 1; r < d; ++r)
      if (std::fabs(m[r][c]) > std::fabs(m[p][c])) p = r;
    if (!(std::fabs(m[p][c]) > 0.0)) return false;
    if (p != c)
      for (int j = 0; j < 2 * d; ++j) std::swap(m[p][j], m[c][j]);
    const double piv = m[c][c];
    for (int j = 0; j < 2 * d; ++j) m[c][j] /= piv;
    for (int r = 0; r < d; ++r) {
      if (r == c) continue;
      const double f = m[r][c];
      for (int j = 0; j < 2 * d; ++j) m[r][j] -= f * m[c][j];
    }
  }
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) out[i * d + j] = m[i][d + j];
  return true;
}

static int new_session(cco_ctx* ctx, cco_energy* E, int B, int N, int d, int method,
                       cco_relax** out) {
  if (!ctx || !out) return fail(CCO_ERR_ARG, "NULL argument");
  if (B < 1 || N < 1) return fail(CCO_ERR_ARG, "need B >= 1 and N >= 1");
  if (d < 1 || d > 3) return fail(CCO_ERR_ARG, "d must be 1, 2 or 3");
  if (E && E->geo.d != d) return fail(CCO_ERR_ARG, "dimension mismatch");
  cco_relax* R = new cco_relax();
  R->ctx = ctx;
  R->E = E;
  R->B = B;
  R->N = N;
  R->d = d;
  R->method = method;
  std::memset(&R->dc, 0, sizeof(R->dc));
  std::memset(&R->lc, 0, sizeof(R->lc));
  std::memset(&R->fc, 0, sizeof(R->fc));
  std::memset(&R->mc, 0, sizeof(R->mc));
  R->dev.B = B;
  R->dev.nd = N * d;
  R->dev.stride = N * d;
  R->vec_total = (size_t)B * N * d;
  *out = R;
  return CCO_OK;
}

static int finish_session(cco_relax* R) {
  int rc = 0;
  if (R->E) {
    rc = dev_reserve(R->part, sizeof(double2) * (size_t)R->B * n_splits_for(R->N) * R->E->geo.Mbox);
    if (rc == 0) rc = dev_reserve(R->cbox, sizeof(double2) * (size_t)R->B * R->E->geo.Mbox);
    if (rc == 0) rc = reserve_reduce(R->rphi, R->B, R->E->geo);
  }
  if (rc == 0) rc = alloc_host_words(R);
  return rc;
}

int cco_flow_create(cco_ctx* ctx, cco_energy* E, int B, int N, int d,
                    const cco_flow_settings* fs, const double* basis, cco_relax** out) {
  if (!fs || !basis) return fail(CCO_ERR_ARG, "NULL argument");
  if (!(fs->atol > 0.0) || !(fs->rtol > 0.0) || !(fs->dt_init > 0.0) || fs->max_steps <= 0)
    return fail(CCO_ERR_ARG, "atol, rtol, dt_init and max_steps must be positive");
  if (fs->record_every < 0 || fs->max_records < 0)
    return fail(CCO_ERR_ARG, "record_every and max_records must be >= 0");
  if ((fs->record_every > 0) != (fs->max_records > 1))
    return fail(CCO_ERR_ARG, "recording needs record_every > 0 and max_records >= 2");
  if (fs->arclength && std::isfinite(fs->t_max))
    return fail(CCO_ERR_ARG, "t_max is supported only with parametrization='time'");
  double inv[9], bbt[9], minv[9];
  if (!invert3(basis, d, inv)) return fail(CCO_ERR_ARG, "basis must be invertible");
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      double acc = 0.0;
      for (int k = 0; k < d; ++k) acc += basis[i * d + k] * basis[j * d + k];
      bbt[i * d + j] = acc;
    }
  if (!invert3(bbt, d, minv)) return fail(CCO_ERR_ARG, "basis must be invertible");
  DevScope dscope(ctx ? ctx->device : 0);
  cco_relax* R = nullptr;
  int rc = new_session(ctx, E, B, N, d, METHOD_FLOW, &R);
  if (rc) return rc;
  FlowCfg& c = R->fc;
  c.rtol = fs->rtol;
  c.atol = fs->atol;
  c.dt_init = fs->dt_init;
  c.dt_min = fs->dt_min;
  c.dt_max = fs->dt_max;
  c.t_max = fs->t_max;
  c.energy_tol = fs->energy_tol;
  c.grad_tol = fs->grad_tol;
  c.max_steps = fs->max_steps;
  for (int i = 0; i < d * d; ++i) {
    c.metric_inv[i] = minv[i];
    c.basis[i] = basis[i];
  }
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) c.inv_t[i * d + j] = inv[j * d + i];
  c.d = d;
  c.arclength = fs->arclength ? 1 : 0;
  c.record_every = fs->record_every;
  c.record_cap = fs->record_every > 0 ? fs->max_records : 0;
  c.dt0 = nullptr;
  RelaxDev& v = R->dev;
  const size_t vec = (size_t)B * v.stride;
  auto A = [&](auto*& p, size_t cnt) {
    if (rc == 0) rc = alloc_into(R, p, cnt);
  };
  A(v.status, B);
  A(R->prec_store, B);
  A(v.active, B);
  A(v.n_active, 1);
  A(v.s_eval, vec);
  A(v.g_eval, vec);
  A(v.phi_eval, B);
  A(v.s, vec);
  A(v.k, vec * 7);
  A(v.fl, B);
  double* dt0 = nullptr;
  A(dt0, B);
  c.dt0 = nullptr;
  R->flow_dt0 = dt0;
  if (c.record_cap > 0) {
    A(v.tr_s, vec * c.record_cap);
    A(v.tr_e, (size_t)B * c.record_cap);
    A(v.tr_step, (size_t)B * c.record_cap);
    A(v.tr_len, (size_t)B * c.record_cap);
    A(v.tr_it, (size_t)B * c.record_cap);
  }
  if (rc == 0) rc = finish_session(R);
  if (rc) {
    cco_relax_destroy(R);
    return rc;
  }
  *out = R;
  return CCO_OK;
}

int cco_flow_set_method(cco_relax* R, int method) {
  if (!R) return fail(CCO_ERR_ARG, "NULL argument");
  if (R->method != METHOD_FLOW) return fail(CCO_ERR_STATE, "not a flow session");
  if (method != 0 && method != 1) return fail(CCO_ERR_ARG, "method must be 0 (dopri5) or 1 (rosenbrock)");
  if (method == 1 && R->fc.arclength)
    return fail(CCO_ERR_ARG, "method='rosenbrock' integrates the time parametrization");
  if (method == 1 && R->E == nullptr)
    return fail(CCO_ERR_ARG, "method='rosenbrock' needs the built-in stealthy energy (GPU Hessian)");
  R->ros2 = method == 1;
  return CCO_OK;
}

int cco_flow_init(cco_relax* R, const double* s0_host, const double* dt0_host) {
  if (!R || !s0_host) return fail(CCO_ERR_ARG, "NULL argument");
  if (R->method != METHOD_FLOW) return fail(CCO_ERR_STATE, "not a flow session");
  DevScope ds(R->ctx->device);
  hipStream_t st = R->ctx->stream;
  const size_t bytes = sizeof(double) * R->vec_total;
  CCO_HIP(hipMemcpyAsync(R->dev.s, s0_host, bytes, hipMemcpyHostToDevice, st));
  CCO_HIP(hipMemcpyAsync(R->dev.s_eval, R->dev.s, bytes, hipMemcpyDeviceToDevice, st));
  if (dt0_host) {
    CCO_HIP(hipMemcpyAsync(R->flow_dt0, dt0_host, sizeof(double) * R->B, hipMemcpyHostToDevice, st));
    R->fc.dt0 = R->flow_dt0;
  } else {
    R->fc.dt0 = nullptr;
  }
  R->variant_run = resolve_variant(R);
  R->dev.prec = resolve_prec(R);
  if (int rc_mp = reserve_mp_planes(R)) return rc_mp;
  hipLaunchKernelGGL(relax_init_kernel, dim3((R->B + 255) / 256), dim3(256), 0, st, R->dev,
                     R->method);
  CCO_HIP(hipGetLastError());
  CCO_HIP(hipStreamSynchronize(st));
  R->slots_ub = R->B;
  if (int rc_sh = setup_shards(R)) return rc_sh;
  R->ready = true;
  return CCO_OK;
}

int cco_flow_result(cco_relax* R, double* s, double* energy, double* t, double* length,
                    double* dt, int64_t* n_steps, int64_t* n_evals, int64_t* status) {
  if (!R || !s || !energy || !t || !length || !dt || !n_steps || !n_evals || !status)
    return fail(CCO_ERR_ARG, "NULL argument");
  if (R->method != METHOD_FLOW) return fail(CCO_ERR_STATE, "not a flow session");
  DevScope ds(R->ctx->device);
  const size_t B = R->B;
  std::vector<FlowLane> L(B);
  std::vector<int> stv(B);
  CCO_HIP(hipMemcpy(L.data(), R->dev.fl, sizeof(FlowLane) * B, hipMemcpyDeviceToHost));
  CCO_HIP(hipMemcpy(stv.data(), R->dev.status, sizeof(int) * B, hipMemcpyDeviceToHost));
  CCO_HIP(hipMemcpy(s, R->dev.s, sizeof(double) * B * R->dev.stride, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < B; ++i) {
    energy[i] = L[i].f;
    t[i] = L[i].t;
    length[i] = L[i].length;
    dt[i] = L[i].dt;
    n_steps[i] = L[i].n_steps;
    n_evals[i] = L[i].n_evals;
    status[i] = stv[i];
  }
  return CCO_OK;
}

int cco_flow_trajectory(cco_relax* R, int64_t* count, double* s, double* t, double* length,
                        double* energy) {
  if (!R || !count) return fail(CCO_ERR_ARG, "NULL argument");
  if (R->method != METHOD_FLOW) return fail(CCO_ERR_STATE, "not a flow session");
  DevScope ds(R->ctx->device);
  const size_t B = R->B;
  std::vector<FlowLane> L(B);
  CCO_HIP(hipMemcpy(L.data(), R->dev.fl, sizeof(FlowLane) * B, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < B; ++i) count[i] = L[i].rec_count;
  const size_t rows = B * (size_t)R->fc.record_cap;
  if (rows == 0) return CCO_OK;
  if (s) CCO_HIP(hipMemcpy(s, R->dev.tr_s, sizeof(double) * rows * R->dev.stride, hipMemcpyDeviceToHost));
  if (t) CCO_HIP(hipMemcpy(t, R->dev.tr_step, sizeof(double) * rows, hipMemcpyDeviceToHost));
  if (length) CCO_HIP(hipMemcpy(length, R->dev.tr_len, sizeof(double) * rows, hipMemcpyDeviceToHost));
  if (energy) CCO_HIP(hipMemcpy(energy, R->dev.tr_e, sizeof(double) * rows, hipMemcpyDeviceToHost));
  return CCO_OK;
}

int cco_md_create(cco_ctx* ctx, cco_energy* E, int B, int N, int d, const double* basis,
                  double dt, double drift_tol, cco_relax** out) {
  if (!E) return fail(CCO_ERR_ARG, "MD needs the built-in stealthy energy");
  if (!basis) return fail(CCO_ERR_ARG, "NULL basis");
  if (!(std::isfinite(dt) && dt > 0.0)) return fail(CCO_ERR_ARG, "dt must be finite and positive");
  if (!(drift_tol > 0.0)) return fail(CCO_ERR_ARG, "drift_tol must be positive (inf disables)");
  double inv[9];
  if (!invert3(basis, d, inv)) return fail(CCO_ERR_ARG, "basis must be invertible");
  DevScope dscope(ctx ? ctx->device : 0);
  cco_relax* R = nullptr;
  int rc = new_session(ctx, E, B, N, d, METHOD_MD, &R);
  if (rc) return rc;
  MdCfg& c = R->mc;
  c.dt = dt;
  c.half = 0.5 * dt;
  c.drift_tol = drift_tol;
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      c.inv[i * d + j] = inv[i * d + j];
      c.inv_t[i * d + j] = inv[j * d + i];
    }
  c.d = d;
  c.n = N;
  for (int i = 0; i < d * d; ++i) R->md_basis[i] = basis[i];
  RelaxDev& v = R->dev;
  const size_t vec = (size_t)B * v.stride;
  auto A = [&](auto*& p, size_t cnt) {
    if (rc == 0) rc = alloc_into(R, p, cnt);
  };
  A(v.status, B);
  A(R->prec_store, B);
  A(v.active, B);
  A(v.n_active, 1);
  A(v.s_eval, vec);
  A(v.g_eval, vec);
  A(v.phi_eval, B);
  A(v.r, vec);
  A(v.v, vec);
  A(v.f, vec);
  A(v.ml, B);
  if (rc == 0) rc = finish_session(R);
  if (rc) {
    cco_relax_destroy(R);
    return rc;
  }
  *out = R;
  return CCO_OK;
}

int cco_md_run(cco_relax* R, int64_t n_steps, int sync_every, int* n_active) {
  if (!R) return fail(CCO_ERR_ARG, "NULL argument");
  if (R->method != METHOD_MD) return fail(CCO_ERR_STATE, "not an MD session");
  if (n_steps < 0 || sync_every < 1) return fail(CCO_ERR_ARG, "need n_steps >= 0, sync_every >= 1");
  if (R->md_stage < 2) return fail(CCO_ERR_STATE, "cco_md_put must come before cco_md_run");
  int na = R->slots_ub;
  int64_t done = 0;
  while (done < n_steps && na > 0) {
    const int64_t chunk = std::min<int64_t>(sync_every, n_steps - done);
    int rc = cco_relax_step(R, chunk, &na);
    if (rc) return rc;
    done += chunk;
  }
  if (n_active) *n_active = na;
  return CCO_OK;
}

extern "C++" {
namespace {

template <typename T>
int md_alloc(cco_relax* R, T*& out, size_t count) {
  DevBuf b;
  int rc = dev_reserve(b, std::max<size_t>(count, 1) * sizeof(T));
  if (rc) return rc;
  CCO_HIP(hipMemsetAsync(b.p, 0, b.bytes, R->ctx->stream));
  R->md_bufs.push_back(b);
  out = (T*)b.p;
  return CCO_OK;
}

template <typename T>
int md_copy(T* dst, const T* src, size_t count, hipMemcpyKind kind) {
  if (count == 0 || dst == nullptr || src == nullptr) return CCO_OK;
  CCO_HIP(hipMemcpy(dst, src, sizeof(T) * count, kind));
  return CCO_OK;
}

}  // namespace
}  // extern "C++"

extern "C++" {
namespace {

// The unconstrained waves k_i (i = nc .. w-1) as a k-set of their own when
// every one lies on the reciprocal lattice: n = B k / (2 pi) integral to
// 1e-9, so exp(i k.r) = exp(2 pi i n.s) and the factorised rho kernel sums
// them (O(w N) complex multiplies instead of w N sincos).  Waves off the
// lattice, or a set too wide for the rho kernel's LDS tables, stay with
// md_finish's direct sums.
int md_tracked_pass(cco_relax* R, int n_waves, int n_constrained, const double* kvec_host) {
  const int d = R->d, mt = n_waves - n_constrained;
  if (mt <= 0) return CCO_OK;
  std::vector<int64_t> nt((size_t)mt * d);
  std::vector<double> ones(mt, 1.0);
  for (int i = 0; i < mt; ++i) {
    const double* k = kvec_host + (size_t)(n_constrained + i) * d;
    bool nonzero = false;
    for (int a = 0; a < d; ++a) {
      double x = 0.0;
      for (int b = 0; b < d; ++b) x += R->md_basis[a * d + b] * k[b];
      x /= 2.0 * M_PI;
      const double r = std::nearbyint(x);
      if (!(std::fabs(x - r) <= 1e-9 * std::max(1.0, std::fabs(r))) || std::fabs(r) > 1e6)
        return CCO_OK;  // off the lattice: direct sums
      nt[(size_t)i * d + a] = (int64_t)r;
      nonzero = nonzero || r != 0.0;
    }
    if (!nonzero) return CCO_OK;
  }
  cco_energy* T = nullptr;
  const int rc_t = cco_energy_create(R->ctx, nt.data(), ones.data(), mt, d, 1.0, 16, &T);
  if (rc_t == CCO_ERR_ARG) {
    g_last_error.clear();  // too wide for the LDS tables: direct sums
    return CCO_OK;
  }
  if (rc_t != CCO_OK) return rc_t;
  R->trk = T;
  const Geo& g = T->geo;
  const size_t B = R->B;
  // a large wave set has many unit tiles per lane: about 4096 workgroups over
  // all lanes need fewer particle splits than the energy's, and fewer splits
  // mean fewer partial sums for the reduce to read
  const int want = (4096 + R->B * g.n_utiles - 1) / (R->B * g.n_utiles);
  const int ns = std::max(1, std::min(n_splits_for(R->N), want));
  R->trk_split_len = ((R->N + ns - 1) / ns + kSplitLen - 1) / kSplitLen * kSplitLen;
  R->trk_splits = (R->N + R->trk_split_len - 1) / R->trk_split_len;
  MdObs& o = R->md_obs;
  int rc = CCO_OK;
  auto go = [&](auto*& p, size_t cnt) {
    if (rc == 0) rc = md_alloc(R, p, cnt);
  };
  go(R->trk_part, B * R->trk_splits * g.Mbox);
  go(R->trk_cbox, B * g.Mbox);
  go(R->trk_rphi, B * g.n_rtiles);
  go(R->trk_rbox, B * g.Mbox);
  go(R->mc.trk_skip, B);
  // a split-precision session sums the tracked waves on the MFMA kernels too
  // when their k-set fits them (row tables, LDS): their z1 planes
  const Geo& gm = g.mp_geo != nullptr ? *g.mp_geo : g;
  if (R->precision == 1 && R->E->geo.mp_ok && (gm.mp_ok || mp_rho_fits(gm)))
    go(R->trk_z1g, B * mp_z1g_bytes(gm, R->N));
  if (rc) return rc;
  R->mc.trk_mbox = g.Mbox;
  o.trk_rbox = R->trk_rbox;
  o.trk_kbox = g.kbox;
  return CCO_OK;
}

}  // namespace
}  // extern "C++"

int cco_md_configure(cco_relax* R, const cco_md_settings* ms, int n_waves, int n_constrained,
                     const double* kvec_host, const int64_t* record_steps_host) {
  if (!R || !ms) return fail(CCO_ERR_ARG, "NULL argument");
  if (R->method != METHOD_MD) return fail(CCO_ERR_STATE, "not an MD session");
  if (ms->scheme < 0 || ms->scheme > 2) return fail(CCO_ERR_ARG, "scheme must be 0, 1 or 2");
  if (ms->andersen_mode < 0 || ms->andersen_mode > 1)
    return fail(CCO_ERR_ARG, "andersen_mode must be 0 (count) or 1 (bernoulli)");
  if (!(ms->andersen_p >= 0.0 && ms->andersen_p <= 1.0) || !(ms->gamma >= 0.0) ||
      !std::isfinite(ms->gamma))
    return fail(CCO_ERR_ARG, "need 0 <= andersen_p <= 1 and a finite gamma >= 0");
  if (ms->wind_every < 1 || ms->record_rho_every < 0 || ms->max_rho_records < 0 ||
      ms->n_record_steps < 0 || ms->n_record_steps > 256 || !(ms->phase_floor >= 0.0))
    return fail(CCO_ERR_ARG, "bad wind_every, trace, record or phase_floor setting");
  if (n_waves < 0 || n_constrained < 0 || n_constrained > n_waves ||
      n_constrained > R->E->geo.M || (n_waves > 0 && !kvec_host) ||
      (ms->n_record_steps > 0 && !record_steps_host))
    return fail(CCO_ERR_ARG, "bad waves (need n_constrained <= n_waves, <= M) or NULL arrays");
  for (int i = 1; i < ms->n_record_steps; ++i)
    if (record_steps_host[i] <= record_steps_host[i - 1])
      return fail(CCO_ERR_ARG, "record_steps must be strictly increasing");
  DevScope ds(R->ctx->device);
  R->md_stage = 0;  // the observable buffers are replaced: the state must be put again
  R->ready = false;
  for (DevBuf& b : R->md_bufs) dev_release(b);
  R->md_bufs.clear();
  if (R->trk) cco_energy_destroy(R->trk);
  R->trk = nullptr;
  R->trk_part = R->trk_cbox = nullptr;
  R->trk_rphi = nullptr;
  R->trk_z1g = nullptr;
  MdCfg& c = R->mc;
  c.trk_skip = nullptr;
  c.trk_mbox = 0;
  R->md_rbox = nullptr;
  R->trk_rbox = nullptr;
  c.scheme = ms->scheme;
  c.andersen_mode = ms->andersen_mode;
  c.andersen_p = ms->andersen_p;
  c.ou_c = std::exp(-ms->gamma * c.dt);
  c.ou_var = -std::expm1(-2.0 * ms->gamma * c.dt);  // 1 - c^2 without cancellation
  c.phase_floor = ms->phase_floor;
  c.w = n_waves;
  c.nc = n_constrained;
  c.wind_every = ms->wind_every;
  c.mbox = R->E->geo.Mbox;
  c.n_rec = ms->n_record_steps;
  c.rho_every = ms->record_rho_every;
  c.rho_cap = ms->record_rho_every > 0 ? ms->max_rho_records : 0;
  const size_t B = R->B, w = n_waves, nr = c.n_rec, q = c.rho_cap;
  const size_t nd = (size_t)R->N * R->d;
  MdObs& o = R->md_obs;
  o = MdObs{};
  int rc = CCO_OK;
  double* kvec = nullptr;
  long long* rec = nullptr;
  auto go = [&](auto*& p, size_t cnt) {
    if (rc == 0) rc = md_alloc(R, p, cnt);
  };
  go(kvec, w * R->d);
  go(rec, nr);
  go(o.rho_prev, B * w * 2);
  go(o.wind_abs, B * w);
  go(o.wind_net, B * w);
  go(o.n_valid, B * w);
  go(o.sum_s, B * w);
  go(R->md_rbox, B * c.mbox);
  go(o.snap_step, B * nr);
  go(o.snap_time, B * nr);
  go(o.snap_phi, B * nr);
  go(o.snap_kin, B * nr);
  go(o.snap_wabs, B * nr * w);
  go(o.snap_wnet, B * nr * w);
  go(o.snap_rho, B * nr * w * 2);
  go(o.snap_r, B * nr * nd);
  go(o.tr_rho, B * q * w * 2);
  go(o.tr_time, B * q);
  go(o.cm, B * 3);
  go(o.rows, B * 2);
  if (rc) return rc;
  o.rbox = R->md_rbox;
  o.kbox = R->E->geo.kbox;
  if (w > 0) CCO_HIP(hipMemcpy(kvec, kvec_host, sizeof(double) * w * R->d, hipMemcpyHostToDevice));
  if (nr > 0)
    CCO_HIP(hipMemcpy(rec, record_steps_host, sizeof(long long) * nr, hipMemcpyHostToDevice));
  c.kvec = kvec;
  c.rec_steps = rec;
  if ((rc = md_tracked_pass(R, n_waves, n_constrained, kvec_host))) return rc;
  CCO_HIP(hipStreamSynchronize(R->ctx->stream));
  R->md_stage = 1;
  return CCO_OK;
}

int cco_md_put(cco_relax* R, const cco_md_state_io* s) {
  if (!R || !s || !s->r || !s->v || !s->f || !s->lane_f64 || !s->lane_i64 || !s->key)
    return fail(CCO_ERR_ARG, "NULL argument");
  if (R->method != METHOD_MD) return fail(CCO_ERR_STATE, "not an MD session");
  if (R->md_stage < 1) return fail(CCO_ERR_STATE, "cco_md_configure must come before cco_md_put");
  DevScope ds(R->ctx->device);
  hipStream_t st = R->ctx->stream;
  CCO_HIP(hipStreamSynchronize(st));
  const size_t B = R->B, nd = (size_t)R->dev.stride;
  const MdCfg& c = R->mc;
  const size_t w = c.w, nr = c.n_rec, q = c.rho_cap;
  // the counters index device rows: reject what would address outside them
  for (size_t b = 0; b < B; ++b) {
    const int64_t* y = s->lane_i64 + b * 5;
    if (y[0] < 0 || y[1] < 0 || y[3] < 0 || y[3] > (int64_t)nr || y[4] < 0 || y[4] > (int64_t)q)
      return fail(CCO_ERR_ARG, "lane " + std::to_string(b) +
                                   ": need step, n_sum >= 0, 0 <= snapshot count <= " +
                                   std::to_string(nr) + ", 0 <= trace count <= " + std::to_string(q));
    if (y[2] != (int64_t)(int32_t)y[2])
      return fail(CCO_ERR_ARG, "lane " + std::to_string(b) + ": status " + std::to_string(y[2]) +
                                   " does not fit 32 bits");
  }
  R->md_stage = 1;  // until every copy below has landed
  int rc;
  if ((rc = md_copy(R->dev.r, s->r, B * nd, hipMemcpyHostToDevice))) return rc;
  if ((rc = md_copy(R->dev.v, s->v, B * nd, hipMemcpyHostToDevice))) return rc;
  if ((rc = md_copy(R->dev.f, s->f, B * nd, hipMemcpyHostToDevice))) return rc;
  std::vector<MdLane> L(B);
  std::vector<int> status(B);
  for (size_t b = 0; b < B; ++b) {
    const double* x = s->lane_f64 + b * 10;
    const int64_t* y = s->lane_i64 + b * 5;
    MdLane& l = L[b];
    l = MdLane{};
    l.phi = x[0];
    l.temperature = x[1];
    l.time = x[2];
    l.reset_acc = x[3];
    l.e0 = x[4];
    l.max_drift = x[5];
    l.sum_kinetic = x[6];
    l.sum_phi = x[7];
    l.sum_phi2 = x[8];
    l.kinetic = x[9];
    l.step = y[0];
    l.n_sum = y[1];
    status[b] = (int)y[2];
    l.snap_count = y[3];
    l.trace_count = y[4];
    l.key[0] = s->key[2 * b];
    l.key[1] = s->key[2 * b + 1];
    l.mode = MODE_EVAL;
  }
  if ((rc = md_copy(R->dev.ml, L.data(), B, hipMemcpyHostToDevice))) return rc;
  if ((rc = md_copy(R->dev.status, status.data(), B, hipMemcpyHostToDevice))) return rc;
  const MdObs& o = R->md_obs;
  if (w > 0) {
    if ((rc = md_copy(o.rho_prev, s->rho_prev, B * w * 2, hipMemcpyHostToDevice))) return rc;
    if ((rc = md_copy(o.wind_abs, s->wind_abs, B * w, hipMemcpyHostToDevice))) return rc;
    if ((rc = md_copy(o.wind_net, s->wind_net, B * w, hipMemcpyHostToDevice))) return rc;
    if ((rc = md_copy(o.n_valid, (const long long*)s->n_valid, B * w, hipMemcpyHostToDevice)))
      return rc;
    if ((rc = md_copy(o.sum_s, s->sum_s, B * w, hipMemcpyHostToDevice))) return rc;
    if ((rc = md_copy(o.snap_step, (const long long*)s->snap_step, B * nr, hipMemcpyHostToDevice)))
      return rc;
    if ((rc = md_copy(o.snap_time, s->snap_time, B * nr, hipMemcpyHostToDevice))) return rc;
    if ((rc = md_copy(o.snap_phi, s->snap_phi, B * nr, hipMemcpyHostToDevice))) return rc;
    if ((rc = md_copy(o.snap_kin, s->snap_kinetic, B * nr, hipMemcpyHostToDevice))) return rc;
    if ((rc = md_copy(o.snap_wabs, s->snap_wind_abs, B * nr * w, hipMemcpyHostToDevice))) return rc;
    if ((rc = md_copy(o.snap_wnet, s->snap_wind_net, B * nr * w, hipMemcpyHostToDevice))) return rc;
    if ((rc = md_copy(o.snap_rho, s->snap_rho, B * nr * w * 2, hipMemcpyHostToDevice))) return rc;
    if ((rc = md_copy(o.snap_r, s->snap_r, B * nr * nd, hipMemcpyHostToDevice))) return rc;
    if ((rc = md_copy(o.tr_rho, s->trace_rho, B * q * w * 2, hipMemcpyHostToDevice))) return rc;
    if ((rc = md_copy(o.tr_time, s->trace_time, B * q, hipMemcpyHostToDevice))) return rc;
  }
  R->variant_run = resolve_variant(R);
  R->dev.prec = resolve_prec(R);
  if (int rc_mp = reserve_mp_planes(R)) return rc_mp;
  if (R->dev.prec != nullptr)
    hipLaunchKernelGGL(fill_int_kernel, dim3((R->B + 255) / 256), dim3(256), 0, st, R->dev.prec,
                       R->B, 1);
  R->slots_ub = R->B;
  if ((rc = setup_shards(R))) return rc;  // active lists from the uploaded status
  CCO_HIP(hipStreamSynchronize(st));
  R->ready = true;
  R->md_stage = 2;
  return CCO_OK;
}

int cco_md_get(cco_relax* R, cco_md_state_io* s) {
  if (!R || !s || !s->r || !s->v || !s->f || !s->lane_f64 || !s->lane_i64 || !s->key)
    return fail(CCO_ERR_ARG, "NULL argument");
  if (R->method != METHOD_MD) return fail(CCO_ERR_STATE, "not an MD session");
  if (R->md_stage < 2) return fail(CCO_ERR_STATE, "cco_md_put must come before cco_md_get");
  DevScope ds(R->ctx->device);
  CCO_HIP(hipStreamSynchronize(R->ctx->stream));
  const size_t B = R->B, nd = (size_t)R->dev.stride;
  const MdCfg& c = R->mc;
  const size_t w = c.w, nr = c.n_rec, q = c.rho_cap;
  int rc;
  if ((rc = md_copy(s->r, R->dev.r, B * nd, hipMemcpyDeviceToHost))) return rc;
  if ((rc = md_copy(s->v, R->dev.v, B * nd, hipMemcpyDeviceToHost))) return rc;
  if ((rc = md_copy(s->f, R->dev.f, B * nd, hipMemcpyDeviceToHost))) return rc;
  std::vector<MdLane> L(B);
  std::vector<int> status(B);
  if ((rc = md_copy(L.data(), R->dev.ml, B, hipMemcpyDeviceToHost))) return rc;
  if ((rc = md_copy(status.data(), R->dev.status, B, hipMemcpyDeviceToHost))) return rc;
  for (size_t b = 0; b < B; ++b) {
    double* x = s->lane_f64 + b * 10;
    int64_t* y = s->lane_i64 + b * 5;
    const MdLane& l = L[b];
    x[0] = l.phi;
    x[1] = l.temperature;
    x[2] = l.time;
    x[3] = l.reset_acc;
    x[4] = l.e0;
    x[5] = l.max_drift;
    x[6] = l.sum_kinetic;
    x[7] = l.sum_phi;
    x[8] = l.sum_phi2;
    x[9] = l.kinetic;
    y[0] = l.step;
    y[1] = l.n_sum;
    y[2] = status[b];
    y[3] = l.snap_count;
    y[4] = l.trace_count;
    s->key[2 * b] = l.key[0];
    s->key[2 * b + 1] = l.key[1];
  }
  const MdObs& o = R->md_obs;
  if (w > 0) {
    if ((rc = md_copy(s->rho_prev, o.rho_prev, B * w * 2, hipMemcpyDeviceToHost))) return rc;
    if ((rc = md_copy(s->wind_abs, o.wind_abs, B * w, hipMemcpyDeviceToHost))) return rc;
    if ((rc = md_copy(s->wind_net, o.wind_net, B * w, hipMemcpyDeviceToHost))) return rc;
    if ((rc = md_copy((long long*)s->n_valid, o.n_valid, B * w, hipMemcpyDeviceToHost))) return rc;
    if ((rc = md_copy(s->sum_s, o.sum_s, B * w, hipMemcpyDeviceToHost))) return rc;
    if ((rc = md_copy((long long*)s->snap_step, o.snap_step, B * nr, hipMemcpyDeviceToHost)))
      return rc;
    if ((rc = md_copy(s->snap_time, o.snap_time, B * nr, hipMemcpyDeviceToHost))) return rc;
    if ((rc = md_copy(s->snap_phi, o.snap_phi, B * nr, hipMemcpyDeviceToHost))) return rc;
    if ((rc = md_copy(s->snap_kinetic, o.snap_kin, B * nr, hipMemcpyDeviceToHost))) return rc;
    if ((rc = md_copy(s->snap_wind_abs, o.snap_wabs, B * nr * w, hipMemcpyDeviceToHost))) return rc;
    if ((rc = md_copy(s->snap_wind_net, o.snap_wnet, B * nr * w, hipMemcpyDeviceToHost))) return rc;
    if ((rc = md_copy(s->snap_rho, o.snap_rho, B * nr * w * 2, hipMemcpyDeviceToHost))) return rc;
    if ((rc = md_copy(s->snap_r, o.snap_r, B * nr * nd, hipMemcpyDeviceToHost))) return rc;
    if ((rc = md_copy(s->trace_rho, o.tr_rho, B * q * w * 2, hipMemcpyDeviceToHost))) return rc;
    if ((rc = md_copy(s->trace_time, o.tr_time, B * q, hipMemcpyDeviceToHost))) return rc;
  }
  return CCO_OK;
}

// ---------------------------------------------------------------------------
// ensemble analysis
// ---------------------------------------------------------------------------
int cco_sk_moments(cco_ctx* ctx, const double* s_host, int B, int N, int d, const int64_t* dir_n,
                   const int32_t* dir_L, int D, double* tot_host, double* tot2_host) {
  if (!ctx || !s_host || !dir_n || !dir_L || !tot_host || !tot2_host)
    return fail(CCO_ERR_ARG, "NULL argument");
  if (B < 1 || N < 1 || D < 1 || d < 1 || d > 3) return fail(CCO_ERR_ARG, "bad sizes");
  DevScope ds(ctx->device);
  std::vector<int> koff(D);
  int64_t Mk = 0;
  int max_L = 0;
  for (int c = 0; c < D; ++c) {
    if (dir_L[c] < 1) return fail(CCO_ERR_ARG, "every direction needs at least one multiple");
    koff[c] = (int)Mk;
    Mk += dir_L[c];
    max_L = std::max(max_L, (int)dir_L[c]);
  }
  if (Mk > (int64_t)1 << 30) return fail(CCO_ERR_ARG, "k-set too large");
  // directions sorted by their number of multiples (uniform work per tile)
  std::vector<int> order(D);
  for (int c = 0; c < D; ++c) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return dir_L[a] > dir_L[b]; });
  std::vector<int4> dirs(D);
  std::vector<int> koff_sorted(D);
  for (int q = 0; q < D; ++q) {
    const int c = order[q];
    dirs[q] = make_int4((int)dir_n[(size_t)c * d], d >= 2 ? (int)dir_n[(size_t)c * d + 1] : 0,
                        d >= 3 ? (int)dir_n[(size_t)c * d + 2] : 0, dir_L[c]);
    koff_sorted[q] = koff[c];
  }
  const int chunk = std::max(1, std::min(B, (int)std::max<int64_t>(1, ((int64_t)1 << 27) / Mk)));
  DevBuf d_s, d_dirs, d_koff, d_out, d_tot, d_tot2;
  int rc = 0;
  hipStream_t st = ctx->stream;
  auto cleanup = [&]() {
    for (DevBuf* b : {&d_s, &d_dirs, &d_koff, &d_out, &d_tot, &d_tot2}) dev_release(*b);
  };
  if ((rc = dev_reserve(d_s, sizeof(double) * (size_t)B * N * d)) ||
      (rc = dev_reserve(d_dirs, sizeof(int4) * D)) || (rc = dev_reserve(d_koff, sizeof(int) * D)) ||
      (rc = dev_reserve(d_out, sizeof(double) * (size_t)chunk * Mk)) ||
      (rc = dev_reserve(d_tot, sizeof(double) * Mk)) || (rc = dev_reserve(d_tot2, sizeof(double) * Mk))) {
    cleanup();
    return rc;
  }
  hipError_t e = hipMemcpyAsync(d_s.p, s_host, sizeof(double) * (size_t)B * N * d,
                                hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(d_dirs.p, dirs.data(), sizeof(int4) * D, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(d_koff.p, koff_sorted.data(), sizeof(int) * D, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(d_tot.p, 0, sizeof(double) * Mk, st);
  if (e == hipSuccess) e = hipMemsetAsync(d_tot2.p, 0, sizeof(double) * Mk, st);
  for (int c0 = 0; e == hipSuccess && c0 < B; c0 += chunk) {
    const int nc = std::min(chunk, B - c0);
    launch_sk((const double*)d_s.p + (size_t)c0 * N * d, nc, N, d, (const int4*)d_dirs.p,
              (const int*)d_koff.p, D, (int)Mk, max_L, (double*)d_out.p, st);
    launch_sk_sum((const double*)d_out.p, nc, (int)Mk, (double*)d_tot.p, (double*)d_tot2.p, st);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(tot_host, d_tot.p, sizeof(double) * Mk, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(tot2_host, d_tot2.p, sizeof(double) * Mk, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  cleanup();
  if (e != hipSuccess) return fail(CCO_ERR_DEVICE, std::string("cco_sk_moments: ") + hipGetErrorString(e));
  return CCO_OK;
}

int cco_pair_stats(cco_ctx* ctx, const double* s_host, int B, int N, int d, const double* basis,
                   const double* edges, int n_edges, int64_t* counts_host, double* nn_host) {
  if (!ctx || !s_host || !basis || !edges || !counts_host) return fail(CCO_ERR_ARG, "NULL argument");
  if (B < 1 || N < 1 || d < 1 || d > 3) return fail(CCO_ERR_ARG, "bad sizes");
  if (n_edges < 2 || n_edges > 1024) return fail(CCO_ERR_ARG, "need 2 to 1024 bin edges");
  DevScope ds(ctx->device);
  DevBuf d_s, d_b, d_e, d_c, d_nn;
  int rc = 0;
  hipStream_t st = ctx->stream;
  auto cleanup = [&]() {
    for (DevBuf* b : {&d_s, &d_b, &d_e, &d_c, &d_nn}) dev_release(*b);
  };
  if ((rc = dev_reserve(d_s, sizeof(double) * (size_t)B * N * d)) ||
      (rc = dev_reserve(d_b, sizeof(double) * d * d)) || (rc = dev_reserve(d_e, sizeof(double) * n_edges)) ||
      (rc = dev_reserve(d_c, sizeof(unsigned long long) * (n_edges - 1))) ||
      (nn_host && (rc = dev_reserve(d_nn, sizeof(double) * (size_t)B * N)))) {
    cleanup();
    return rc;
  }
  hipError_t e = hipMemcpyAsync(d_s.p, s_host, sizeof(double) * (size_t)B * N * d, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(d_b.p, basis, sizeof(double) * d * d, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(d_e.p, edges, sizeof(double) * n_edges, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(d_c.p, 0, sizeof(unsigned long long) * (n_edges - 1), st);
  if (e == hipSuccess) {
    launch_pair((const double*)d_s.p, B, N, d, (const double*)d_b.p, (const double*)d_e.p, n_edges,
                (unsigned long long*)d_c.p, nn_host ? (double*)d_nn.p : nullptr, st);
    e = hipGetLastError();
  }
  std::vector<unsigned long long> cnt(n_edges - 1);
  if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), d_c.p, sizeof(unsigned long long) * (n_edges - 1), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && nn_host) e = hipMemcpyAsync(nn_host, d_nn.p, sizeof(double) * (size_t)B * N, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  cleanup();
  if (e != hipSuccess) return fail(CCO_ERR_DEVICE, std::string("cco_pair_stats: ") + hipGetErrorString(e));
  for (int b = 0; b < n_edges - 1; ++b) counts_host[b] = (int64_t)cnt[b];
  return CCO_OK;
}

}  // extern "C"
