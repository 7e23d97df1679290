// sovlp.hip -- kernels and C ABI of libsovlp.so (gfx950).
//
// Kernels:
//   spmv_vector_kernel   Backend surface SpMV, TPR lanes per row
//   elementwise kernels  Backend surface steps (backend.py:204-291)
//   reduce_kernel        Backend surface dot / nrm2 / nrm_inf, fixed order
//   spmv_tiled_kernel    the engines' CSR-Adaptive tile pass as a bare SpMV
//   pdlp_run_kernel      PDLP phases of each LP, one workgroup per LP
//   pdlp_phase_kernel    one PDLP phase of one LP over W workgroups
//   pdlp_decide_kernel   reduces the W partial sums and takes the transition
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sovlp.h"
#include "pdlp_kernels.hpp"
#include "certify_kernels.hpp"
#include "bnr_kernels.hpp"

using namespace sov;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(2, std::string(#expr) + ": " + hipGetErrorString(e_));                \
  } while (0)

inline hipStream_t S(sovlp_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kBlock = 256;
constexpr int kTnnz64 = 2048;   // FP64 engines: 2 vectors x 2048 x 8 B = 32 KiB LDS
constexpr int kTnnz32 = 4096;   // FP32 engines: 2 x 4096 x 4 B (or 4096 doubles) = 32 KiB

// ---------------------------------------------------------------- Backend
template <class T, int TPR>
__global__ void spmv_vector_kernel(int rows, const int64_t* __restrict__ rp,
                                   const int32_t* __restrict__ ri, const T* __restrict__ rx,
                                   const T* __restrict__ x, T* __restrict__ y) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long row = g / TPR;
  const int ln = (int)(g % TPR);
  T s = T(0);
  if (row < rows) {
    const int64_t e = rp[row + 1];
    for (int64_t p = rp[row] + ln; p < e; p += TPR) s += rx[p] * x[ri[p]];
  }
#pragma unroll
  for (int o = TPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (row < rows && ln == 0) y[row] = s;
}

template <class T>
int launch_spmv_vector(int rows, const int64_t* rp, const int32_t* ri, const T* rx, const T* x,
                       T* y, int tpr, hipStream_t st) {
  if (rows < 0) return fail(1, "rows < 0");
  if (rows == 0) return 0;
  const long long threads = (long long)rows * tpr;
  const int blocks = (int)((threads + 255) / 256);
  switch (tpr) {
    case 1: spmv_vector_kernel<T, 1><<<blocks, 256, 0, st>>>(rows, rp, ri, rx, x, y); break;
    case 2: spmv_vector_kernel<T, 2><<<blocks, 256, 0, st>>>(rows, rp, ri, rx, x, y); break;
    case 4: spmv_vector_kernel<T, 4><<<blocks, 256, 0, st>>>(rows, rp, ri, rx, x, y); break;
    case 8: spmv_vector_kernel<T, 8><<<blocks, 256, 0, st>>>(rows, rp, ri, rx, x, y); break;
    case 16: spmv_vector_kernel<T, 16><<<blocks, 256, 0, st>>>(rows, rp, ri, rx, x, y); break;
    case 32: spmv_vector_kernel<T, 32><<<blocks, 256, 0, st>>>(rows, rp, ri, rx, x, y); break;
    case 64: spmv_vector_kernel<T, 64><<<blocks, 256, 0, st>>>(rows, rp, ri, rx, x, y); break;
    default: return fail(1, "lanes_per_row must be a power of two in [1, 64]");
  }
  HIPCHK(hipGetLastError());
  return 0;
}

__global__ void primal_step_kernel(double* __restrict__ xnew, const double* __restrict__ x,
                                   const double* __restrict__ c, const double* __restrict__ aty,
                                   const double* __restrict__ lo, const double* __restrict__ hi,
                                   double tau, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  xnew[i] = clampT(x[i] - tau * (c[i] + aty[i]), lo[i], hi[i]);
}

__global__ void dual_step_kernel(double* __restrict__ ynew, const double* __restrict__ y,
                                 const double* __restrict__ ax, const double* __restrict__ rl,
                                 const double* __restrict__ ru, double sigma, int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double w = y[i] + sigma * ax[i];
  ynew[i] = w - sigma * clampT(w / sigma, rl[i], ru[i]);
}

__global__ void pdhg_primal_kernel(double* __restrict__ x, double* __restrict__ dx,
                                   const double* __restrict__ c, const double* __restrict__ aty,
                                   const double* __restrict__ lo, const double* __restrict__ hi,
                                   double tau, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = x[i];
  const double v = clampT(xi - tau * (c[i] + aty[i]), lo[i], hi[i]);
  dx[i] = v - xi;
  x[i] = v;
}

__global__ void pdhg_dual_kernel(double* __restrict__ y, double* __restrict__ dy,
                                 double* __restrict__ ax, const double* __restrict__ adx,
                                 const double* __restrict__ rl, const double* __restrict__ ru,
                                 double sigma, int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double a = adx[i];
  const double axn = ax[i] + a;
  ax[i] = axn;
  const double yo = y[i];
  const double w = yo + sigma * (axn + a);
  const double yv = w - sigma * clampT(w / sigma, rl[i], ru[i]);
  dy[i] = yv - yo;
  y[i] = yv;
}

__global__ void pdhg_accum_kernel(double* __restrict__ acc, const double* __restrict__ inc,
                                  double* __restrict__ sum, const double* __restrict__ v,
                                  double eta, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  acc[i] += inc[i];
  sum[i] += eta * v[i];
}

// one 1024-thread workgroup; OP 0 = dot, 1 = 2-norm, 2 = inf-norm
template <class T, int OP>
__global__ void __launch_bounds__(1024) reduce_kernel(const T* __restrict__ a,
                                                      const T* __restrict__ b, long long n,
                                                      double* __restrict__ out) {
  __shared__ double red[(1024 / 64 + 2) * kMaxRed];
  double v[1] = {OP == 2 ? 0.0 : 0.0};
  for (long long i = threadIdx.x; i < n; i += 1024) {
    const double ai = (double)a[i];
    if (OP == 0) v[0] += ai * (double)b[i];
    else if (OP == 1) v[0] += ai * ai;
    else v[0] = nanmax(v[0], fabs(ai));
  }
  block_reduce<1024, 1>(v, OP == 2 ? 1u : 0u, red);
  if (threadIdx.x == 0) *out = OP == 1 ? sqrt(v[0]) : v[0];
}

template <class T, int OP>
int launch_reduce(const T* a, const T* b, int64_t n, double* out, hipStream_t st) {
  if (n < 0) return fail(1, "n < 0");
  reduce_kernel<T, OP><<<1, 1024, 0, st>>>(a, b, (long long)n, out);
  HIPCHK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- tiled SpMV
template <class T, int BLOCK, int TNNZ>
__global__ void __launch_bounds__(BLOCK)
    spmv_tiled_kernel(const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                      const T* __restrict__ val, const int32_t* __restrict__ tiles,
                      const int32_t* __restrict__ part, const T* __restrict__ x,
                      T* __restrict__ y) {
  __shared__ __align__(16) T lds[TNNZ];
  const int t0 = part[blockIdx.x], t1 = part[blockIdx.x + 1];
  auto prod = [&](int q, int j, T (&v)[1]) { v[0] = val[q] * x[j]; };
  auto epi = [&](int r, const T (&s)[1]) { y[r] = s[0]; };
  tile_pass<BLOCK, TNNZ, 1, false, T>(ptr, idx, tiles, t0, t1, prod, epi, lds);
}

// ------------------------------------------------------- column-panel SpMV
// y = A x with A split into P column panels, each its own CSR (pointers
// ptrs[p * (rows + 1) + r] index the concatenated idx / val).  Workgroup b
// works on panel b % P, so with workgroups dispatched round-robin over the
// XCDs every XCD gathers from one slice of x, which its L2 holds; each panel
// writes its row sums to partial[p * rows + r].
template <class T, int BLOCK, int TNNZ>
__global__ void __launch_bounds__(BLOCK)
    spmv_panel_kernel(int panels, int rows, const int32_t* __restrict__ ptrs,
                      const int32_t* __restrict__ idx, const T* __restrict__ val,
                      const int32_t* __restrict__ tiles, const int32_t* __restrict__ tile_off,
                      const int32_t* __restrict__ parts, int groups, const T* __restrict__ x,
                      T* __restrict__ partial) {
  __shared__ __align__(16) T lds[TNNZ];
  const int p = (int)blockIdx.x % panels, g = (int)blockIdx.x / panels;
  const int32_t* pp = parts + (size_t)p * (groups + 1);
  const int t0 = pp[g], t1 = pp[g + 1];
  T* out = partial + (size_t)p * rows;
  auto prod = [&](int q, int j, T (&v)[1]) { v[0] = val[q] * x[j]; };
  auto epi = [&](int r, const T (&s)[1]) { out[r] = s[0]; };
  tile_pass<BLOCK, TNNZ, 1, false, T>(ptrs + (size_t)p * (rows + 1), idx, tiles + tile_off[p],
                                      t0, t1, prod, epi, lds);
}

// y[r] = sum of the panels' partial row sums, in panel order
template <class T>
__global__ void panel_sum_kernel(int panels, int rows, const T* __restrict__ partial,
                                 T* __restrict__ y) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  T s = partial[r];
  for (int p = 1; p < panels; ++p) s += partial[(size_t)p * rows + r];
  y[r] = s;
}

// ---------------------------------------------------------------- workspace
struct Carver {
  char* base;
  size_t off;
  size_t cap;
  template <class U>
  U* take(size_t count) {
    off = (off + 255) & ~size_t(255);
    U* p = reinterpret_cast<U*>(base ? base + off : nullptr);
    off += sizeof(U) * (count ? count : 1);
    return p;
  }
};

template <class T>
void carve_lp(Carver& cv, LpDev<T>& L, int m, int n, long long nnz, long long hist_cap) {
  L.sav = cv.take<T>(nnz);
  L.stv = cv.take<T>(nnz);
  L.rowf = cv.take<double>(m);
  L.colf = cv.take<double>(n);
  L.tmpm = cv.take<double>(m);
  L.tmpn = cv.take<double>(n);
  T** cols[] = {&L.c, &L.lo, &L.hi, &L.x, &L.xn, &L.dx, &L.aty, &L.sumx, &L.xr, &L.atyf, &L.atya};
  for (T** p : cols) *p = cv.take<T>(n);
  T** rows[] = {&L.rl, &L.ru, &L.y, &L.yn, &L.dy, &L.ax, &L.adx, &L.sumy, &L.yr, &L.axf, &L.axa};
  for (T** p : rows) *p = cv.take<T>(m);
  L.xo = cv.take<double>(n);
  L.yo = cv.take<double>(m);
  L.dout = cv.take<double>(n);
  L.hist = cv.take<double>(5 * (size_t)(hist_cap > 0 ? hist_cap : 1));
  L.hist_cap = hist_cap;
  L.rhist = cv.take<double>(5 * kMaxRounds);
  L.st = cv.take<LpState>(1);
  if (sizeof(T) == 4) {      // FP32 iterates: room for iterative refinement
    L.wrl = cv.take<double>(m);
    L.wru = cv.take<double>(m);
    L.wlo = cv.take<double>(n);
    L.whi = cv.take<double>(n);
    L.xacc = cv.take<double>(n);
    L.axacc = cv.take<double>(m);
    L.yacc = cv.take<double>(m);
    L.qc = cv.take<double>(m);
    L.dacc = cv.take<double>(n);
    L.cc = cv.take<double>(n);
    L.q = reinterpret_cast<T*>(cv.take<float>(m));
    L.xbest = cv.take<double>(n);
    L.ybest = cv.take<double>(m);
  }
}

int clock_hz(double* hz) {
  int dev = 0, khz = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  *hz = (double)khz * 1000.0;
  return 0;
}

void host_partition(int ntiles, const int32_t* tiles, const int64_t* ptr_at_tile, int parts,
                    int32_t* part) {
  // ptr_at_tile[t] = pointer value at row tiles[t]; cost = nonzeros + rows
  std::vector<double> pre(ntiles + 1, 0.0);
  for (int t = 0; t < ntiles; ++t)
    pre[t + 1] = pre[t] + (double)(ptr_at_tile[t + 1] - ptr_at_tile[t]) +
                 (double)(tiles[t + 1] - tiles[t]);
  const double total = pre[ntiles];
  part[0] = 0;
  int t = 0;
  for (int p = 1; p < parts; ++p) {
    const double target = total * (double)p / (double)parts;
    while (t < ntiles && pre[t] < target) ++t;
    part[p] = t;
  }
  part[parts] = ntiles;
}

}  // namespace

extern "C" {

const char* sovlp_last_error(void) { return g_err.c_str(); }

int sovlp_version(void) { return 100; }

int sovlp_clock_rate(int device, double* hz) {
  int khz = 0;
  HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
  *hz = (double)khz * 1000.0;
  return 0;
}

int sovlp_tile_geometry(int precision, int* block, int* tile_rows, int* tile_nnz) {
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  *block = kBlock;
  *tile_rows = kRowsPerLane * kBlock;
  *tile_nnz = precision == 64 ? kTnnz64 : kTnnz32;
  return 0;
}

int sovlp_spmv_csr_f64(int rows, const int64_t* rp, const int32_t* ri, const double* rx,
                       const double* x, double* y, int lanes_per_row, sovlp_stream_t stream) {
  return launch_spmv_vector<double>(rows, rp, ri, rx, x, y, lanes_per_row, S(stream));
}
int sovlp_spmv_csr_f32(int rows, const int64_t* rp, const int32_t* ri, const float* rx,
                       const float* x, float* y, int lanes_per_row, sovlp_stream_t stream) {
  return launch_spmv_vector<float>(rows, rp, ri, rx, x, y, lanes_per_row, S(stream));
}

// ---------------------------------------------------------------- certificate
int sovlp_residual_workspace_bytes(int rows, int64_t* bytes) {
  if (!bytes) return fail(1, "null argument");
  if (rows < 0) return fail(1, "negative row count");
  *bytes = 4 * ((int64_t)rows + 4 + kCertMaxGroups);
  return 0;
}

// The lines of more than kLongLine nonzeros, in index order, into list[0..*total):
// chunks of the lines over at most kCertMaxGroups workgroups, counted, scanned
// (counts[kCertMaxGroups] holds the counts, then the offsets) and written.
static_assert(kCertMaxGroups <= 4 * kCertBlock, "long_scan_kernel covers 4 counts per thread");
static int list_long_lines(int lines, const int64_t* ptr, int* counts, int* total, int* list,
                           hipStream_t st) {
  const int groups = std::max(1, std::min(kCertMaxGroups, (lines + kCertBlock - 1) / kCertBlock));
  const int chunk = (int)(((int64_t)lines + groups - 1) / groups);
  long_count_kernel<<<groups, kCertBlock, 0, st>>>(lines, chunk, ptr, counts);
  long_scan_kernel<<<1, kCertBlock, 0, st>>>(groups, counts, total);
  long_write_kernel<<<groups, kCertBlock, 0, st>>>(lines, chunk, ptr, counts, list);
  HIPCHK(hipGetLastError());
  return 0;
}

int sovlp_residual_f64(int rows, const int64_t* rp, const int32_t* ri, const double* rx,
                       const double* x, const double* b, double* out, void* workspace,
                       int64_t workspace_bytes, sovlp_stream_t stream) {
  if (rows < 0) return fail(1, "negative row count");
  if (rows == 0) return 0;
  if (!rp || !ri || !rx || !x || !b || !out || !workspace) return fail(1, "null argument");
  int64_t need = 0;
  sovlp_residual_workspace_bytes(rows, &need);
  if (workspace_bytes < need) return fail(1, "workspace too small (sovlp_residual_workspace_bytes)");
  if (reinterpret_cast<uintptr_t>(workspace) % 4) return fail(1, "workspace not 4-byte aligned");
  // the count of long rows, their list, then the listing's per-chunk counts
  int* cnt = static_cast<int*>(workspace);
  hipStream_t st = S(stream);
  const int groups = std::min(kCertMaxGroups * 4, (rows + kCertBlock - 1) / kCertBlock);
  if (int rc = list_long_lines(rows, rp, cnt + 4 + rows, cnt, cnt + 4, st)) return rc;
  residual_kernel<<<groups, kCertBlock, 0, st>>>(rows, rp, ri, rx, x, b, out);
  HIPCHK(hipGetLastError());
  long_lines_kernel<true><<<kCertMaxGroups, kCertBlock, 0, st>>>(cnt, cnt + 4, rp, ri, rx, x, b,
                                                                   1.0, out);
  HIPCHK(hipGetLastError());
  return 0;
}

static int cert_groups(int len) {
  return std::max(1, std::min(kCertMaxGroups, (len + kCertBlock - 1) / kCertBlock));
}

int sovlp_certify_workspace_bytes(int m, int n, int64_t* bytes) {
  if (!bytes) return fail(1, "null argument");
  if (m < 0 || n < 0) return fail(1, "negative size");
  *bytes = 8 * (2 * (int64_t)m + n + 15 * (int64_t)kCertMaxGroups + 8) +
           4 * ((int64_t)m + n + 4 + 2 * (int64_t)kCertMaxGroups);
  return 0;
}

int sovlp_certify_f64(int m, int n, const int64_t* rp, const int32_t* ri, const double* rx,
                      const int64_t* cp, const int32_t* ci, const double* cx, const double* c,
                      const double* rl, const double* ru, const double* lo, const double* hi,
                      double offset, int maximise, const double* x, const double* y,
                      double dual_tol, void* workspace, int64_t workspace_bytes, double* d_out,
                      double* act_out, sovlp_certificate* out, sovlp_stream_t stream) {
  if (!out || !workspace) return fail(1, "null argument");
  if (m < 0 || n < 0) return fail(1, "negative size");
  if ((m > 0 && (!rp || !ri || !rx || !rl || !ru || !y)) ||
      (n > 0 && (!cp || !ci || !cx || !c || !lo || !hi || !x)))
    return fail(1, "null array");
  int64_t need = 0;
  sovlp_certify_workspace_bytes(m, n, &need);
  if (workspace_bytes < need) return fail(1, "workspace too small (sovlp_certify_workspace_bytes)");
  if (reinterpret_cast<uintptr_t>(workspace) % 8) return fail(1, "workspace not 8-byte aligned");
  double* ws = static_cast<double*>(workspace);
  CertPtrs P{};
  P.m = m; P.n = n; P.maximise = maximise ? 1 : 0; P.dual_tol = dual_tol;
  P.rp = rp; P.cp = cp; P.ri = ri; P.ci = ci; P.rx = rx; P.cx = cx;
  P.c = c; P.rl = rl; P.ru = ru; P.lo = lo; P.hi = hi; P.x = x; P.y = y;
  P.yy = ws;
  P.d = d_out ? d_out : ws + m;
  P.act = act_out ? act_out : ws + m + n;
  P.p1 = ws + 2 * (int64_t)m + n;
  P.p2 = P.p1 + 5 * kCertMaxGroups;
  P.p3 = P.p2 + 6 * kCertMaxGroups;
  P.res = P.p3 + 4 * kCertMaxGroups;
  int* cnt = reinterpret_cast<int*>(P.res + 8);     // two counters, then the lists
  int* list_r = cnt + 4;
  int* list_c = list_r + m;
  int* counts_r = list_c + n;                       // per-chunk counts, then offsets
  int* counts_c = counts_r + kCertMaxGroups;
  const int g1 = cert_groups(m), g2 = cert_groups(n);
  hipStream_t st = S(stream);
  // long rows' activities, then the rows (yy), long columns' d, the columns
  zero_ints_kernel<<<1, 64, 0, st>>>(cnt, 2);
  HIPCHK(hipGetLastError());
  if (m > 0) {
    if (int rc = list_long_lines(m, rp, counts_r, cnt, list_r, st)) return rc;
    long_lines_kernel<true><<<kCertMaxGroups, kCertBlock, 0, st>>>(cnt, list_r, rp, ri, rx, x,
                                                                     nullptr, -1.0, P.act);
    HIPCHK(hipGetLastError());
  }
  cert_rows_kernel<<<g1, kCertBlock, 0, st>>>(P);
  HIPCHK(hipGetLastError());
  if (n > 0) {
    if (int rc = list_long_lines(n, cp, counts_c, cnt + 1, list_c, st)) return rc;
    long_lines_kernel<false><<<kCertMaxGroups, kCertBlock, 0, st>>>(
        cnt + 1, list_c, cp, ci, cx, P.yy, c, maximise ? -1.0 : 1.0, P.d);
    HIPCHK(hipGetLastError());
  }
  cert_cols_kernel<<<g2, kCertBlock, 0, st>>>(P);
  HIPCHK(hipGetLastError());
  cert_colterms_kernel<<<g2, kCertBlock, 0, st>>>(P, g2);
  HIPCHK(hipGetLastError());
  cert_final_kernel<<<1, kCertBlock, 0, st>>>(P, g1, g2, offset);
  HIPCHK(hipGetLastError());
  double r[6];
  HIPCHK(hipMemcpyAsync(r, P.res, sizeof(r), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  out->bound = r[0];
  out->perturbation = r[1];
  out->objective = r[2];
  out->col_violation = r[3];
  out->row_violation = r[4];
  out->vacuous = r[5] > 0.0 ? 1 : 0;
  out->reserved = 0;
  return 0;
}

// ------------------------------------------------- batched node relaxation
static void bnr_shape(int K, int* KP, int* RB) {
  int kp = 1;
  while (kp < K) kp <<= 1;
  *KP = kp;
  *RB = kBnrBlock / kp;
}

static int64_t bnr_groups(int len, int RB) { return std::max<int64_t>(1, (len + RB - 1) / RB); }

int sovlp_bnr_workspace_bytes(int m, int n, int K, int64_t* bytes) {
  if (!bytes) return fail(1, "null argument");
  if (m < 0 || n < 1 || K < 1 || K > 256) return fail(1, "need m >= 0, n >= 1, 1 <= K <= 256");
  int KP, RB;
  bnr_shape(K, &KP, &RB);
  const int64_t gm = bnr_groups(m, RB), gn = bnr_groups(n, RB);
  const int64_t parts = 3 * (int64_t)K * (gn + gm + (gm + gn));
  *bytes = 8 * (4 * (int64_t)n * K + 4 * (int64_t)m * K + parts) + 4 * (int64_t)K + 256;
  return 0;
}

int sovlp_bnr_solve_batch(const sovlp_bnr_batch* b, void* workspace, int64_t workspace_bytes,
                          sovlp_stream_t stream) {
  if (!b || !workspace) return fail(1, "null argument");
  const int m = b->m, n = b->n, K = b->K;
  int64_t need = 0;
  if (int rc = sovlp_bnr_workspace_bytes(m, n, K, &need)) return rc;
  if (workspace_bytes < need) return fail(1, "workspace too small (sovlp_bnr_workspace_bytes)");
  if (reinterpret_cast<uintptr_t>(workspace) % 8) return fail(1, "workspace not 8-byte aligned");
  if (b->iters < 0) return fail(1, "negative iteration count");
  // index and value arrays may be null when A has no nonzeros (never read then)
  if (!b->cp || !b->c || !b->L || !b->U || !b->X || !b->eta || !b->omega || !b->bounds ||
      !b->empty || (m > 0 && (!b->rp || !b->rl || !b->ru || !b->Y)))
    return fail(1, "null array");
  int KP, RB;
  bnr_shape(K, &KP, &RB);
  const int gm = (int)bnr_groups(m, RB), gn = (int)bnr_groups(n, RB);
  double* w = static_cast<double*>(workspace);
  BnrPtrs P{};
  P.m = m; P.n = n; P.K = K; P.KP = KP; P.RB = RB;
  P.rp = b->rp; P.cp = b->cp; P.ri = b->ri; P.ci = b->ci; P.rx = b->rx; P.cx = b->cx;
  P.c = b->c; P.rl = b->rl; P.ru = b->ru; P.L = b->L; P.U = b->U;
  P.X = b->X; P.Y = b->Y; P.eta = b->eta; P.omega = b->omega;
  P.growth = b->step_growth; P.damp = b->step_damp;
  if (b->dual_sign != 1.0 && b->dual_sign != -1.0) return fail(1, "dual_sign must be 1 or -1");
  P.ysign = b->dual_sign;
  const int64_t nk = (int64_t)n * K, mk = (int64_t)m * K;
  P.Xn = w; P.dX = w + nk; P.AtY = w + 2 * nk; P.AtdY = w + 3 * nk;
  P.Yn = w + 4 * nk; P.dY = P.Yn + mk; P.AX = P.Yn + 2 * mk; P.AdX = P.Yn + 3 * mk;
  double* part_p = P.Yn + 4 * mk;                       // primal pass: gn groups
  double* part_d = part_p + 3 * (int64_t)K * gn;        // dual pass: gm groups
  double* part_b = part_d + 3 * (int64_t)K * gm;        // bound terms: gm + gn groups
  P.accept = reinterpret_cast<int*>(part_b + 3 * (int64_t)K * (gm + gn));
  P.empty = b->empty;
  P.bounds = b->bounds;
  hipStream_t st = S(stream);
  const int gk = (K + 63) / 64;
  bnr_zero_kernel<<<gk, 64, 0, st>>>(P.empty, K);
  HIPCHK(hipGetLastError());
  bnr_start_kernel<<<gn, kBnrBlock, 0, st>>>(P);
  HIPCHK(hipGetLastError());
  if (m > 0) bnr_spmm_kernel<<<gm, kBnrBlock, 0, st>>>(m, K, KP, RB, P.rp, P.ri, P.rx, P.X, P.AX);
  bnr_spmm_kernel<<<gn, kBnrBlock, 0, st>>>(n, K, KP, RB, P.cp, P.ci, P.cx, P.Y, P.AtY);
  HIPCHK(hipGetLastError());
  if (m == 0) HIPCHK(hipMemsetAsync(P.AtdY, 0, sizeof(double) * nk, st));   // A^T dY = 0
  BnrPtrs Pp = P, Pd = P, Pb = P;
  Pp.part = part_p;
  Pd.part = part_d;
  Pb.part = part_b;
  for (int it = 0; it < b->iters; ++it) {
    // from the second iteration on, the previous iteration's update runs in
    // the primal launch (its row part in gm extra blocks)
    const int pending = it > 0 ? 1 : 0;
    bnr_primal_kernel<<<gn + (pending && m > 0 ? gm : 0), kBnrBlock, 0, st>>>(Pp, pending, gn);
    if (m > 0) {
      bnr_spmm_kernel<<<gm, kBnrBlock, 0, st>>>(m, K, KP, RB, P.rp, P.ri, P.rx, P.dX, P.AdX);
      bnr_dual_kernel<<<gm, kBnrBlock, 0, st>>>(Pd);
      bnr_spmm_kernel<<<gn, kBnrBlock, 0, st>>>(n, K, KP, RB, P.cp, P.ci, P.cx, P.dY, P.AtdY);
    }
    const double t = (double)(it + 2);
    bnr_decide_kernel<<<K, kBnrLanes, 0, st>>>(P, part_p, gn, part_d, m > 0 ? gm : 0,
                                         1.0 - std::pow(t, -b->step_growth),
                                         1.0 + std::pow(t, -b->step_damp));
    HIPCHK(hipGetLastError());
  }
  if (b->iters > 0) bnr_update_kernel<<<gn + (m > 0 ? gm : 0), kBnrBlock, 0, st>>>(P, gn);
  // the bound from the final Y: A^T Y into AtdY, then the terms
  if (m > 0) {
    bnr_spmm_kernel<<<gn, kBnrBlock, 0, st>>>(n, K, KP, RB, P.cp, P.ci, P.cx, P.Y, P.AtdY);
  } else {
    HIPCHK(hipMemsetAsync(P.AtdY, 0, sizeof(double) * nk, st));
  }
  bnr_bound_terms_kernel<<<(m > 0 ? gm : 0) + gn, kBnrBlock, 0, st>>>(Pb, m > 0 ? gm : 0);
  bnr_bound_final_kernel<<<K, kBnrLanes, 0, st>>>(Pb, m > 0 ? gm : 0, gn);
  HIPCHK(hipGetLastError());
  return 0;
}

#define ELEMWISE_LAUNCH(kernel, len, ...)                                     \
  do {                                                                        \
    if ((len) < 0) return fail(1, "negative length");                         \
    if ((len) == 0) return 0;                                                 \
    kernel<<<((len) + 255) / 256, 256, 0, S(stream)>>>(__VA_ARGS__);          \
    HIPCHK(hipGetLastError());                                                \
    return 0;                                                                 \
  } while (0)

int sovlp_primal_step_f64(double* xnew, const double* x, const double* c, const double* aty,
                          const double* lo, const double* hi, double tau, int n,
                          sovlp_stream_t stream) {
  ELEMWISE_LAUNCH(primal_step_kernel, n, xnew, x, c, aty, lo, hi, tau, n);
}
int sovlp_dual_step_f64(double* ynew, const double* y, const double* ax, const double* rl,
                        const double* ru, double sigma, int m, sovlp_stream_t stream) {
  ELEMWISE_LAUNCH(dual_step_kernel, m, ynew, y, ax, rl, ru, sigma, m);
}
int sovlp_pdhg_primal_f64(double* x, double* dx, const double* c, const double* aty,
                          const double* lo, const double* hi, double tau, int n,
                          sovlp_stream_t stream) {
  ELEMWISE_LAUNCH(pdhg_primal_kernel, n, x, dx, c, aty, lo, hi, tau, n);
}
int sovlp_pdhg_dual_f64(double* y, double* dy, double* ax, const double* adx, const double* rl,
                        const double* ru, double sigma, int m, sovlp_stream_t stream) {
  ELEMWISE_LAUNCH(pdhg_dual_kernel, m, y, dy, ax, adx, rl, ru, sigma, m);
}
int sovlp_pdhg_accum_f64(double* acc, const double* inc, double* sum, const double* v,
                         double eta, int n, sovlp_stream_t stream) {
  ELEMWISE_LAUNCH(pdhg_accum_kernel, n, acc, inc, sum, v, eta, n);
}

int sovlp_dot_f64(const double* a, const double* b, int64_t n, double* out, sovlp_stream_t s) {
  return launch_reduce<double, 0>(a, b, n, out, S(s));
}
int sovlp_nrm2_f64(const double* a, int64_t n, double* out, sovlp_stream_t s) {
  return launch_reduce<double, 1>(a, a, n, out, S(s));
}
int sovlp_nrminf_f64(const double* a, int64_t n, double* out, sovlp_stream_t s) {
  return launch_reduce<double, 2>(a, a, n, out, S(s));
}
int sovlp_dot_f32(const float* a, const float* b, int64_t n, double* out, sovlp_stream_t s) {
  return launch_reduce<float, 0>(a, b, n, out, S(s));
}
int sovlp_nrm2_f32(const float* a, int64_t n, double* out, sovlp_stream_t s) {
  return launch_reduce<float, 1>(a, a, n, out, S(s));
}

// Serial additions one lane makes to reduce the rows of a tile with `rows`
// rows, the longest of length `maxlen`, on `block` lanes (tile_pass_pf: up
// to 64 lanes per row when the rows fit the lanes, else one lane for up to
// kRowsPerLane rows).
static int64_t tile_serial_work(int64_t rows, int64_t maxlen, int block) {
  if (rows <= block) {
    int tpr = 64;
    while (tpr > 1 && (int64_t)tpr * rows > block) tpr >>= 1;
    return (maxlen + tpr - 1) / tpr;
  }
  return ((rows + block - 1) / block) * maxlen;
}

int sovlp_plan_tiles(int rows, const int64_t* ptr, int tile_nnz, int tile_rows, int block,
                     int32_t* tiles, int32_t* ntiles_out) {
  if (rows < 0 || tile_nnz < 1 || tile_rows < 1 || block < 0)
    return fail(1, "bad tile plan arguments");
  if (ptr[rows] >= (int64_t)INT32_MAX) return fail(4, "more than 2^31-1 nonzeros");
  // a tile also ends before a row that would make one lane's serial share of
  // the row reductions exceed kMaxSerial additions (a long row among many
  // short ones would otherwise be summed by a single lane)
  constexpr int64_t kMaxSerial = 64;
  int t = 0, r = 0;
  tiles[0] = 0;
  while (r < rows) {
    const int64_t len = ptr[r + 1] - ptr[r];
    if (len < 0) return fail(1, "row pointers must be non-decreasing");
    if (len > tile_nnz) {
      tiles[++t] = ++r;
      continue;
    }
    int64_t acc = 0, maxlen = 0;
    const int start = r;
    while (r < rows && r - start < tile_rows) {
      const int64_t l = ptr[r + 1] - ptr[r];
      if (l > tile_nnz || acc + l > tile_nnz) break;
      const int64_t ml = l > maxlen ? l : maxlen;
      if (block > 0 && r > start && tile_serial_work(r - start + 1, ml, block) > kMaxSerial) break;
      acc += l;
      maxlen = ml;
      ++r;
    }
    tiles[++t] = r;
  }
  *ntiles_out = t;
  return 0;
}

int sovlp_partition_tiles(int ntiles, const int32_t* tiles, const int64_t* ptr, int parts,
                          int32_t* part_out) {
  if (ntiles < 0 || parts < 1) return fail(1, "bad partition arguments");
  std::vector<int64_t> pt(ntiles + 1);
  for (int t = 0; t <= ntiles; ++t) pt[t] = ptr[tiles[t]];
  host_partition(ntiles, tiles, pt.data(), parts, part_out);
  return 0;
}

}  // extern "C"

template <class T, int TNNZ>
static int spmv_tiled(int rows, const int32_t* ptr32, const int32_t* idx, const T* val,
                      const int32_t* tiles, int ntiles, const int32_t* part, int grid,
                      const T* x, T* y, hipStream_t st) {
  if (rows < 0 || ntiles < 0 || grid < 1) return fail(1, "bad tiled SpMV arguments");
  if (rows == 0 || ntiles == 0) return 0;
  spmv_tiled_kernel<T, kBlock, TNNZ><<<grid, kBlock, 0, st>>>(ptr32, idx, val, tiles, part, x, y);
  HIPCHK(hipGetLastError());
  return 0;
}

template <class T, int TNNZ>
static int spmv_panels(int panels, int rows, const int32_t* ptrs, const int32_t* idx,
                       const T* val, const int32_t* tiles, const int32_t* tile_off,
                       const int32_t* parts, int groups, const T* x, T* partial, T* y,
                       hipStream_t st) {
  if (panels < 1 || rows < 0 || groups < 1) return fail(1, "bad panel SpMV arguments");
  if ((long long)panels * groups > (1LL << 30)) return fail(1, "too many workgroups");
  if (rows == 0) return 0;
  spmv_panel_kernel<T, kBlock, TNNZ><<<panels * groups, kBlock, 0, st>>>(
      panels, rows, ptrs, idx, val, tiles, tile_off, parts, groups, x, partial);
  HIPCHK(hipGetLastError());
  panel_sum_kernel<T><<<(rows + 255) / 256, 256, 0, st>>>(panels, rows, partial, y);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" {

int sovlp_spmv_panels_f64(int panels, int rows, const int32_t* ptrs, const int32_t* idx,
                          const double* val, const int32_t* tiles, const int32_t* tile_off,
                          const int32_t* parts, int groups, const double* x, double* partial,
                          double* y, sovlp_stream_t stream) {
  return spmv_panels<double, kTnnz64>(panels, rows, ptrs, idx, val, tiles, tile_off, parts,
                                      groups, x, partial, y, S(stream));
}
int sovlp_spmv_panels_f32(int panels, int rows, const int32_t* ptrs, const int32_t* idx,
                          const float* val, const int32_t* tiles, const int32_t* tile_off,
                          const int32_t* parts, int groups, const float* x, float* partial,
                          float* y, sovlp_stream_t stream) {
  return spmv_panels<float, kTnnz32>(panels, rows, ptrs, idx, val, tiles, tile_off, parts,
                                     groups, x, partial, y, S(stream));
}

int sovlp_spmv_tiled_f64(int rows, const int32_t* ptr32, const int32_t* idx, const double* val,
                         const int32_t* tiles, int ntiles, const int32_t* part, int grid,
                         const double* x, double* y, sovlp_stream_t stream) {
  return spmv_tiled<double, kTnnz64>(rows, ptr32, idx, val, tiles, ntiles, part, grid, x, y,
                                     S(stream));
}
int sovlp_spmv_tiled_f32(int rows, const int32_t* ptr32, const int32_t* idx, const float* val,
                         const int32_t* tiles, int ntiles, const int32_t* part, int grid,
                         const float* x, float* y, sovlp_stream_t stream) {
  return spmv_tiled<float, kTnnz32>(rows, ptr32, idx, val, tiles, ntiles, part, grid, x, y,
                                    S(stream));
}

int sovlp_selftest(void) {
  // [[1, 0, 2], [0, 3, 0]] . [1, 2, 3] = [7, 6]
  const int64_t rp[3] = {0, 2, 3};
  const int32_t rp32[3] = {0, 2, 3}, ri[3] = {0, 2, 1}, tiles[2] = {0, 2}, part[2] = {0, 1};
  const double rx[3] = {1.0, 2.0, 3.0}, xv[3] = {1.0, 2.0, 3.0};
  const float rxf[3] = {1.0f, 2.0f, 3.0f}, xf[3] = {1.0f, 2.0f, 3.0f};
  char* d = nullptr;
  HIPCHK(hipMalloc(&d, 4096));
  int64_t* d_rp = (int64_t*)d;
  int32_t* d_rp32 = (int32_t*)(d + 64);
  int32_t* d_ri = (int32_t*)(d + 128);
  int32_t* d_tiles = (int32_t*)(d + 192);
  int32_t* d_part = (int32_t*)(d + 256);
  double* d_rx = (double*)(d + 320);
  double* d_x = (double*)(d + 384);
  double* d_y = (double*)(d + 448);
  float* d_rxf = (float*)(d + 512);
  float* d_xf = (float*)(d + 576);
  float* d_yf = (float*)(d + 640);
  int rc = 0;
  std::string why;
  auto run = [&]() -> int {
    HIPCHK(hipMemcpy(d_rp, rp, sizeof rp, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_rp32, rp32, sizeof rp32, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_ri, ri, sizeof ri, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_tiles, tiles, sizeof tiles, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_part, part, sizeof part, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_rx, rx, sizeof rx, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_x, xv, sizeof xv, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_rxf, rxf, sizeof rxf, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_xf, xf, sizeof xf, hipMemcpyHostToDevice));
    for (int variant = 0; variant < 4; ++variant) {
      double got[2] = {12345.0, 12345.0};
      float gotf[2] = {12345.0f, 12345.0f};
      HIPCHK(hipMemcpy(d_y, got, sizeof got, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(d_yf, gotf, sizeof gotf, hipMemcpyHostToDevice));
      int r = 0;
      if (variant == 0) r = sovlp_spmv_csr_f64(2, d_rp, d_ri, d_rx, d_x, d_y, 64, nullptr);
      if (variant == 1) r = sovlp_spmv_tiled_f64(2, d_rp32, d_ri, d_rx, d_tiles, 1, d_part, 1, d_x, d_y, nullptr);
      if (variant == 2) r = sovlp_spmv_csr_f32(2, d_rp, d_ri, d_rxf, d_xf, d_yf, 4, nullptr);
      if (variant == 3) r = sovlp_spmv_tiled_f32(2, d_rp32, d_ri, d_rxf, d_tiles, 1, d_part, 1, d_xf, d_yf, nullptr);
      if (r) return r;
      HIPCHK(hipDeviceSynchronize());
      HIPCHK(hipMemcpy(got, d_y, sizeof got, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(gotf, d_yf, sizeof gotf, hipMemcpyDeviceToHost));
      const bool ok = variant < 2 ? (got[0] == 7.0 && got[1] == 6.0)
                                  : (gotf[0] == 7.0f && gotf[1] == 6.0f);
      if (!ok) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "selftest variant %d returned [%g, %g], expected [7, 6]",
                      variant, variant < 2 ? got[0] : gotf[0], variant < 2 ? got[1] : gotf[1]);
        why = buf;
        return 5;
      }
    }
    return 0;
  };
  rc = run();
  (void)hipFree(d);
  if (rc == 5) return fail(5, why);
  return rc;
}

}  // extern "C"

// ====================================================================== engines
struct sovlp_pdlp {
  int precision, W, m, n;
  void* lpdev;               // LpDev<T> (device)
  LpState* st;
  StepCtl* ctl;
  unsigned* done;
  double *xo, *yo, *dout, *hist, *rhist;
  long long* pcount;         // [kPhaseSlots]
  int phases_per_replay;
  long long hist_cap;
  double hz;
  bool started;
  hipGraph_t graph;
  hipGraphExec_t exec;
  hipStream_t cap;
  long long ticks_set;
  LpState* hstate;           // [2] pinned host copies of the state
  hipEvent_t ev[2];
};

struct sovlp_batch {
  int count, precision, block;
  void* lps;                 // LpDev<T>[count] (device)
  StepCtl* ctl;
  unsigned* done;
  long long* pcount;         // [count][kPhaseSlots]
  std::vector<double*> xo, yo, dout, rhist;
  std::vector<LpState*> st;
  hipGraph_t graph;
  hipGraphExec_t exec;
  hipStream_t cap;
  double hz;
  long long ticks_set;
  unsigned* hdone;           // [2] pinned host copies of the done counter
  hipEvent_t ev[2];
  hipStream_t side[2];       // capture branches of the second and third groups
  hipEvent_t fork, join[2];
};

namespace {

constexpr int kTnnzSmall = 512;   // 64-thread workgroups: 2 x 512 x 8 B = 8 KiB LDS
constexpr int kPairsPerGraph = 64;     // W > 1: phase + decide launches per graph

// Workgroup sizes of the batched engine, and their tile nonzero caps (16 KiB
// of LDS for 128 threads, 32 KiB for 256)
bool batch_block_ok(int block) { return block == 64 || block == 128 || block == 256; }
int batch_tnnz(int precision, int block) {
  if (block == 64) return kTnnzSmall;
  if (block == 128) return precision == 64 ? kTnnz64 / 2 : kTnnz32 / 2;
  return precision == 64 ? kTnnz64 : kTnnz32;
}

// One launch of the run kernel over `count` LP descriptors with `block` threads each
template <class T>
void launch_run(int block, const LpDev<T>* lps, int count, const StepCtl* ctl, unsigned* done,
                int phases, hipStream_t st) {
  constexpr int T256 = sizeof(T) == 8 ? kTnnz64 : kTnnz32;
  if (block == 64)
    pdlp_run_kernel<T, 64, kTnnzSmall><<<count, 64, 0, st>>>(lps, ctl, done, phases);
  else if (block == 128)
    pdlp_run_kernel<T, 128, T256 / 2><<<count, 128, 0, st>>>(lps, ctl, done, phases);
  else
    pdlp_run_kernel<T, 256, T256><<<count, 256, 0, st>>>(lps, ctl, done, phases);
}
// run kernel: phases per LP per launch. Every launch ends when its slowest LP
// has run its phases, so shorter launches leave CUs idle at each end (4096
// uneven LPs: 1024 phases 8470 LPs/s, 16384 11170, 65536 12012, 262144 12101);
// time and iteration limits are enforced on the device, not by this count.
constexpr int kPhasesPerLaunch = 65536;

template <class T>
void fill_lp(LpDev<T>& L, const sovlp_lp& lp) {
  L.m = lp.m; L.n = lp.n; L.ntA = lp.ntiles_a; L.ntT = lp.ntiles_t;
  L.flip = lp.maximise ? 1 : 0;
  L.offset0 = lp.obj_offset;
  L.ap = lp.rp; L.ai = lp.ri; L.av = lp.rx;
  L.tp = lp.cp; L.ti = lp.ci; L.tv = lp.cx;
  L.c0 = lp.c; L.lo0 = lp.lo; L.hi0 = lp.hi; L.rl0 = lp.rl; L.ru0 = lp.ru; L.v0 = lp.v0;
  L.tilesA = lp.tiles_a;
  L.tilesT = lp.tiles_t;
  if (L.xacc) {              // the rounds' bounds live in the workspace
    L.crl = L.wrl; L.cru = L.wru; L.clo = L.wlo; L.chi = L.whi;
  } else {                   // no refinement: the original bounds
    L.crl = lp.rl; L.cru = lp.ru; L.clo = lp.lo; L.chi = lp.hi;
  }
}

// workgroups sharing one LP's phases: one per ~4 tiles, at most what is resident
template <class T>
int pick_width(const sovlp_lp* lp, int team, int* W) {
  int dev = 0, cus = 0, nb = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &nb, pdlp_phase_kernel<T, kBlock, sizeof(T) == 8 ? kTnnz64 : kTnnz32>, kBlock, 0));
  const long long tiles = (long long)lp->ntiles_a + lp->ntiles_t;
  const long long resident = (long long)std::max(nb, 1) * cus;
  long long w = std::max<long long>(1, tiles / 4);
  if (team == 1) w = 1;
  if (team == 0 && lp->nnz < 200000) w = 1;
  *W = (int)std::min(w, resident);
  return 0;
}

template <class T>
size_t single_ws_T(const sovlp_lp* lp, long long hist_cap, int W) {
  Carver cv{nullptr, 0, 0};
  cv.take<LpDev<T>>(1);
  cv.take<StepCtl>(1);
  cv.take<unsigned>(64);
  cv.take<double>((size_t)W * kMaxRed);
  cv.take<int32_t>(W + 1);
  cv.take<int32_t>(W + 1);
  cv.take<long long>(kPhaseSlots);
  LpDev<T> L{};
  carve_lp<T>(cv, L, lp->m, lp->n, lp->nnz, hist_cap);
  return cv.off + 256;
}

template <class T>
int single_create_T(const sovlp_lp* lp, const sovlp_params* prm, void* ws, int64_t ws_bytes,
                    int64_t hist_cap, int team, hipStream_t st, sovlp_pdlp* h) {
  int W = 1;
  int rc = pick_width<T>(lp, team, &W);
  if (rc) return rc;
  if ((size_t)ws_bytes < single_ws_T<T>(lp, hist_cap, W)) return fail(1, "workspace too small");
  Carver cv{reinterpret_cast<char*>(ws), 0, (size_t)ws_bytes};
  LpDev<T>* d_lp = cv.take<LpDev<T>>(1);
  StepCtl* ctl = cv.take<StepCtl>(1);
  unsigned* done = cv.take<unsigned>(64);
  double* partials = cv.take<double>((size_t)W * kMaxRed);
  int32_t* partA = cv.take<int32_t>(W + 1);
  int32_t* partT = cv.take<int32_t>(W + 1);
  long long* pcount = cv.take<long long>(kPhaseSlots);
  LpDev<T> L{};
  carve_lp<T>(cv, L, lp->m, lp->n, lp->nnz, hist_cap);
  fill_lp<T>(L, *lp);
  L.partials = partials;
  L.pcount = pcount;
  if (W > 1) {
    // equal-cost split of both tile lists (needs the pointer values at tile starts)
    std::vector<int32_t> ta(lp->ntiles_a + 1), tt(lp->ntiles_t + 1), rp(lp->m + 1), cp(lp->n + 1);
    HIPCHK(hipMemcpyAsync(ta.data(), lp->tiles_a, sizeof(int32_t) * ta.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(tt.data(), lp->tiles_t, sizeof(int32_t) * tt.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(rp.data(), lp->rp, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(cp.data(), lp->cp, sizeof(int32_t) * cp.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<int64_t> pa(ta.size()), pt(tt.size());
    for (size_t t = 0; t < ta.size(); ++t) pa[t] = rp[ta[t]];
    for (size_t t = 0; t < tt.size(); ++t) pt[t] = cp[tt[t]];
    std::vector<int32_t> ha(W + 1), ht(W + 1);
    host_partition(lp->ntiles_a, ta.data(), pa.data(), W, ha.data());
    host_partition(lp->ntiles_t, tt.data(), pt.data(), W, ht.data());
    HIPCHK(hipMemcpyAsync(partA, ha.data(), sizeof(int32_t) * (W + 1), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(partT, ht.data(), sizeof(int32_t) * (W + 1), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    L.partA = partA;
    L.partT = partT;
  }
  LpState s0{};
  s0.phase = PH_INIT;
  s0.status = ST_RUNNING;
  StepCtl c0{};
  std::memcpy(&c0.P, prm, sizeof(Params));
  c0.ticks = (long long)9.0e18;
  HIPCHK(hipMemcpyAsync(L.st, &s0, sizeof s0, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ctl, &c0, sizeof c0, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(done, 0, sizeof(unsigned), st));
  HIPCHK(hipMemsetAsync(pcount, 0, sizeof(long long) * kPhaseSlots, st));
  HIPCHK(hipMemcpyAsync(d_lp, &L, sizeof L, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));

  h->precision = sizeof(T) == 8 ? 64 : 32;
  h->W = W;
  h->m = lp->m;
  h->n = lp->n;
  h->lpdev = d_lp;
  h->st = L.st;
  h->ctl = ctl;
  h->done = done;
  h->xo = L.xo; h->yo = L.yo; h->dout = L.dout; h->hist = L.hist; h->rhist = L.rhist;
  h->pcount = pcount;
  h->hist_cap = hist_cap;
  h->ticks_set = c0.ticks;
  // record the launches of one replay into a graph on a private capture
  // stream: W == 1, one run launch (kPhasesPerLaunch phases); W > 1,
  // kPairsPerGraph phase + decide pairs (one phase each)
  HIPCHK(hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
  HIPCHK(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
  constexpr int TN = sizeof(T) == 8 ? kTnnz64 : kTnnz32;
  if (W == 1) {
    pdlp_run_kernel<T, kBlock, TN><<<1, kBlock, 0, h->cap>>>(d_lp, ctl, done, kPhasesPerLaunch);
    h->phases_per_replay = kPhasesPerLaunch;
  } else {
    for (int i = 0; i < kPairsPerGraph; ++i) {
      pdlp_phase_kernel<T, kBlock, TN><<<W, kBlock, 0, h->cap>>>(d_lp);
      pdlp_decide_kernel<T, kBlock><<<1, kBlock, 0, h->cap>>>(d_lp, W, ctl, done);
    }
    h->phases_per_replay = kPairsPerGraph;
  }
  HIPCHK(hipStreamEndCapture(h->cap, &h->graph));
  HIPCHK(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
  HIPCHK(hipHostMalloc((void**)&h->hstate, 2 * sizeof(LpState), hipHostMallocDefault));
  for (hipEvent_t& e : h->ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return clock_hz(&h->hz);
}

template <class T>
size_t batch_ws_T(int count, const sovlp_lp* lps, long long hist_cap) {
  Carver cv{nullptr, 0, 0};
  cv.take<LpDev<T>>(count);
  cv.take<StepCtl>(1);
  cv.take<unsigned>(64);
  cv.take<long long>((size_t)count * kPhaseSlots);
  for (int i = 0; i < count; ++i) {
    LpDev<T> L{};
    carve_lp<T>(cv, L, lps[i].m, lps[i].n, lps[i].nnz, hist_cap);
  }
  return cv.off + 256;
}

template <class T>
int batch_create_T(int count, const sovlp_lp* lps, const sovlp_params* prm, const int32_t* blocks,
                   void* ws, int64_t ws_bytes, int64_t hist_cap, hipStream_t st, sovlp_batch* h) {
  if ((size_t)ws_bytes < batch_ws_T<T>(count, lps, hist_cap)) return fail(1, "workspace too small");
  // descriptors grouped by workgroup size (one launch per group), largest LP
  // first within a group, so that the longest workgroups of a launch start
  // first and do not extend its tail; pos[i] is LP i's place in the device
  // array (every output stays indexed by i)
  const int kBlocks[3] = {64, 128, 256};
  std::vector<int> order, pos(count), gcount(3, 0);
  for (int g = 0; g < 3; ++g) {
    std::vector<int> idx;
    for (int i = 0; i < count; ++i)
      if (blocks[i] == kBlocks[g]) idx.push_back(i);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return lps[a].nnz > lps[b].nnz; });
    for (int i : idx) {
      pos[i] = (int)order.size();
      order.push_back(i);
      ++gcount[g];
    }
  }
  if ((int)order.size() != count) return fail(1, "every LP's block must be 64, 128 or 256");
  Carver cv{reinterpret_cast<char*>(ws), 0, (size_t)ws_bytes};
  LpDev<T>* d_lps = cv.take<LpDev<T>>(count);
  StepCtl* ctl = cv.take<StepCtl>(1);
  unsigned* done = cv.take<unsigned>(64);
  long long* pcount = cv.take<long long>((size_t)count * kPhaseSlots);
  std::vector<LpDev<T>> host(count);
  h->xo.resize(count);
  h->yo.resize(count);
  h->dout.resize(count);
  h->rhist.resize(count);
  h->st.resize(count);
  for (int i = 0; i < count; ++i) {
    if (lps[i].nnz >= (int64_t)INT32_MAX) return fail(4, "an LP has more than 2^31-1 nonzeros");
    LpDev<T>& L = host[pos[i]];
    L = LpDev<T>{};
    carve_lp<T>(cv, L, lps[i].m, lps[i].n, lps[i].nnz, hist_cap);
    fill_lp<T>(L, lps[i]);
    L.pcount = pcount + (size_t)i * kPhaseSlots;
    h->xo[i] = L.xo; h->yo[i] = L.yo; h->dout[i] = L.dout; h->st[i] = L.st;
    h->rhist[i] = L.rhist;
  }
  StepCtl c0{};
  std::memcpy(&c0.P, prm, sizeof(Params));
  c0.ticks = (long long)9.0e18;
  HIPCHK(hipMemcpyAsync(d_lps, host.data(), sizeof(LpDev<T>) * count, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ctl, &c0, sizeof c0, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  h->count = count;
  h->precision = sizeof(T) == 8 ? 64 : 32;
  h->block = blocks[0];
  h->lps = d_lps;
  h->ctl = ctl;
  h->done = done;
  h->pcount = pcount;
  h->ticks_set = c0.ticks;
  HIPCHK(hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
  // one launch per group; the groups are independent branches of the graph
  // (forked from and joined to the capture stream), so they can run side by side
  int first = -1, nside = 0;
  for (int g = 0; g < 3; ++g)
    if (gcount[g] > 0 && first < 0) first = g;
  HIPCHK(hipEventCreateWithFlags(&h->fork, hipEventDisableTiming));
  for (int k = 0; k < 2; ++k) {
    HIPCHK(hipStreamCreateWithFlags(&h->side[k], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&h->join[k], hipEventDisableTiming));
  }
  HIPCHK(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
  HIPCHK(hipEventRecord(h->fork, h->cap));
  for (int g = 0, off = 0; g < 3; off += gcount[g], ++g) {
    if (gcount[g] == 0) continue;
    if (g == first) {
      launch_run<T>(kBlocks[g], d_lps + off, gcount[g], ctl, done, kPhasesPerLaunch, h->cap);
    } else {
      hipStream_t sd = h->side[nside];
      HIPCHK(hipStreamWaitEvent(sd, h->fork, 0));
      launch_run<T>(kBlocks[g], d_lps + off, gcount[g], ctl, done, kPhasesPerLaunch, sd);
      HIPCHK(hipEventRecord(h->join[nside], sd));
      ++nside;
    }
  }
  for (int k = 0; k < nside; ++k) HIPCHK(hipStreamWaitEvent(h->cap, h->join[k], 0));
  HIPCHK(hipStreamEndCapture(h->cap, &h->graph));
  HIPCHK(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
  HIPCHK(hipHostMalloc((void**)&h->hdone, 2 * sizeof(unsigned), hipHostMallocDefault));
  for (hipEvent_t& e : h->ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return clock_hz(&h->hz);
}

int set_ticks(StepCtl* ctl, long long* cached, double seconds, double hz, hipStream_t st) {
  const double secs = seconds > 0.0 ? seconds : 0.0;
  long long ticks = (long long)std::min(secs * hz, 9.0e18);
  if (ticks != *cached) {
    *cached = ticks;
    HIPCHK(hipMemcpyAsync(&ctl->ticks, cached, sizeof(long long), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));   // the source is host memory owned by the handle
  }
  return 0;
}

void fill_info(const LpState& t, sovlp_info* f) {
  std::memset(f, 0, sizeof *f);
  const bool opt = t.status == ST_OPTIMAL;
  const double nan = std::nan("");
  f->status = t.status;
  f->stage = t.phase == PH_DONE ? 3
           : t.phase < PH_KKT0 ? 0
           : (t.phase == PH_FINAL || t.phase == PH_FINAL_D) ? 2 : 1;
  f->iterations = t.k;
  f->restarts = t.restarts;
  f->rejected_steps = t.rejected;
  f->history_len = t.hist_len;
  f->primal_res = opt ? t.fin[0] : nan;
  f->dual_res = opt ? t.fin[1] : nan;
  f->gap = opt ? t.fin[2] : nan;
  f->primal_obj = opt ? t.fin[3] : nan;
  f->dual_obj = opt ? t.fin[4] : nan;
  f->step_size = t.eta;
  f->primal_weight = t.omega;
  f->objective = t.objective;
  f->scale_obj = t.sobj0;
  f->device_ticks = t.ticks;
  f->iterations = t.k_done + t.k;
  f->refine_rows = t.ref_rows;
  f->refine_round = t.round;
  f->refine_scale = t.ref_scale;
  f->start_tick = t.tick0;
}

int check_params(const sovlp_params* p, int precision) {
  if (p->check_every < 1 || p->stepsize_every < 1) return fail(1, "check/stepsize period < 1");
  if (p->refine_rounds < 0 || p->refine_rounds > kMaxRounds)
    return fail(1, "refine_rounds must be in [0, 16]");
  if (p->refine_rounds > 0 && precision != 32)
    return fail(4, "iterative refinement is provided for precision 32 only");
  if (p->refine_rounds > 0 && !(p->refine_eps > 0.0 && p->refine_tol >= 0.0))
    return fail(1, "refine_eps must be > 0 and refine_tol >= 0");
  if (p->refine_stall < 0 || p->refine_stall_open < 0)
    return fail(1, "refine_stall and refine_stall_open must be >= 0");
  return 0;
}

}  // namespace

extern "C" {

int sovlp_pdlp_workspace_bytes(const sovlp_lp* lp, int precision, int64_t history_cap,
                               int64_t* bytes) {
  if (!lp || !bytes) return fail(1, "null argument");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  int W = 1;
  int rc = precision == 64 ? pick_width<double>(lp, 2, &W) : pick_width<float>(lp, 2, &W);
  if (rc) return rc;
  *bytes = (int64_t)(precision == 64 ? single_ws_T<double>(lp, history_cap, W)
                                     : single_ws_T<float>(lp, history_cap, W));
  return 0;
}

int sovlp_pdlp_create(const sovlp_lp* lp, const sovlp_params* params, int precision,
                      void* workspace, int64_t workspace_bytes, int64_t history_cap, int team,
                      sovlp_stream_t stream, sovlp_pdlp** out) {
  static_assert(sizeof(sovlp_params) == sizeof(Params), "params layout");
  if (!lp || !params || !workspace || !out) return fail(1, "null argument");
  if (lp->m < 0 || lp->n < 0 || lp->nnz < 0) return fail(1, "negative dimension");
  if (lp->nnz >= (int64_t)INT32_MAX) return fail(4, "more than 2^31-1 nonzeros");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  if (int rc = check_params(params, precision)) return rc;
  if (team < 0 || team > 2) return fail(1, "team must be 0 (auto), 1 (one workgroup) or 2 (many)");
  sovlp_pdlp* h = new sovlp_pdlp{};
  int rc = precision == 64
               ? single_create_T<double>(lp, params, workspace, workspace_bytes, history_cap, team,
                                         S(stream), h)
               : single_create_T<float>(lp, params, workspace, workspace_bytes, history_cap, team,
                                        S(stream), h);
  if (rc) {
    sovlp_pdlp_destroy(h);
    return rc;
  }
  *out = h;
  return 0;
}

int sovlp_pdlp_advance(sovlp_pdlp* h, int64_t iteration_budget, double seconds_left,
                       sovlp_stream_t stream) {
  if (!h) return fail(1, "null handle");
  hipStream_t st = S(stream);
  if (!h->started) {
    int rc = set_ticks(h->ctl, &h->ticks_set, seconds_left, h->hz, st);
    if (rc) return rc;
    h->started = true;
  }
  HIPCHK(hipMemcpyAsync(&h->hstate[0], h->st, sizeof(LpState), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const long long k0 = h->hstate[0].k;
  const long long budget = iteration_budget < 1 ? 1 : iteration_budget;
  // every replay runs phases_per_replay phases; an iteration takes at most 61
  // phases (60 rejected trials), so this bound is never the one that stops a
  // healthy solve
  const long long max_replays = (budget * 61 + 200) / h->phases_per_replay + 16;
  // one replay stays queued while the host reads the state the previous one
  // left (pinned copy + event), so the device does not idle on the host
  long long replays = 0;
  int cur = 0;
  if (h->hstate[0].phase == PH_DONE) return 0;
  HIPCHK(hipGraphLaunch(h->exec, st));
  HIPCHK(hipMemcpyAsync(&h->hstate[0], h->st, sizeof(LpState), hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(h->ev[0], st));
  ++replays;
  for (;;) {
    const int nxt = cur ^ 1;
    const bool more = replays < max_replays;
    if (more) {
      HIPCHK(hipGraphLaunch(h->exec, st));
      HIPCHK(hipMemcpyAsync(&h->hstate[nxt], h->st, sizeof(LpState), hipMemcpyDeviceToHost, st));
      HIPCHK(hipEventRecord(h->ev[nxt], st));
      ++replays;
    }
    HIPCHK(hipEventSynchronize(h->ev[cur]));
    const LpState& s = h->hstate[cur];
    if (s.phase == PH_DONE || s.k - k0 >= budget || !more) break;
    cur = nxt;
  }
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int sovlp_pdlp_info(sovlp_pdlp* h, sovlp_info* info, sovlp_stream_t stream) {
  if (!h || !info) return fail(1, "null argument");
  LpState s;
  HIPCHK(hipMemcpyAsync(&s, h->st, sizeof s, hipMemcpyDeviceToHost, S(stream)));
  HIPCHK(hipStreamSynchronize(S(stream)));
  fill_info(s, info);
  return 0;
}

int sovlp_pdlp_outputs(sovlp_pdlp* h, double** x, double** y, double** d, double** history) {
  if (!h) return fail(1, "null handle");
  if (x) *x = h->xo;
  if (y) *y = h->yo;
  if (d) *d = h->dout;
  if (history) *history = h->hist;
  return 0;
}

int sovlp_pdlp_refinement(sovlp_pdlp* h, double** rows) {
  if (!h || !rows) return fail(1, "null argument");
  *rows = h->rhist;
  return 0;
}

int sovlp_pdlp_width(sovlp_pdlp* h, int* workgroups) {
  if (!h || !workgroups) return fail(1, "null argument");
  *workgroups = h->W;
  return 0;
}

int sovlp_phase_slots(int* slots) {
  if (!slots) return fail(1, "null argument");
  *slots = kPhaseSlots;
  return 0;
}

const char* sovlp_phase_name(int slot) {
  static const char* names[kPhaseSlots] = {
      "INIT", "RUIZ_MAX", "RUIZ_APPLY", "PC_SUM", "PC_APPLY", "SCALE", "AX0", "POW_A",
      "POW_T", "KKT0", "P_TRIAL", "A_TRIAL", "B_ADAPT", "A_FUSED", "B_FUSED", "CHECK",
      "RESTART", "COPY_AVG", "FINAL", "FINAL_D", "REF_X", "REF_KKT", "REF_SHIFT", "DONE",
      "B_ADAPT_NEXT", "B_FUSED_NEXT", "", ""};
  return slot >= 0 && slot < kPhaseSlots ? names[slot] : "";
}

int sovlp_pdlp_phase_counts(sovlp_pdlp* h, int64_t* counts, sovlp_stream_t stream) {
  if (!h || !counts) return fail(1, "null argument");
  static_assert(sizeof(long long) == sizeof(int64_t), "counter width");
  HIPCHK(hipMemcpyAsync(counts, h->pcount, sizeof(long long) * kPhaseSlots,
                        hipMemcpyDeviceToHost, S(stream)));
  HIPCHK(hipStreamSynchronize(S(stream)));
  return 0;
}

int sovlp_pdlp_destroy(sovlp_pdlp* h) {
  if (!h) return 0;
  for (hipEvent_t e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->hstate) (void)hipHostFree(h->hstate);
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  if (h->cap) (void)hipStreamDestroy(h->cap);
  delete h;
  return 0;
}

int sovlp_batch_geometry(int precision, int block, int* tile_rows, int* tile_nnz) {
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  if (!batch_block_ok(block)) return fail(1, "block must be 64, 128 or 256");
  *tile_rows = kRowsPerLane * block;
  *tile_nnz = batch_tnnz(precision, block);
  return 0;
}

int sovlp_batch_workspace_bytes(int count, const sovlp_lp* lps, int precision,
                                int64_t history_cap, int64_t* bytes) {
  if (count < 1 || !lps || !bytes) return fail(1, "bad batch arguments");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  *bytes = (int64_t)(precision == 64 ? batch_ws_T<double>(count, lps, history_cap)
                                     : batch_ws_T<float>(count, lps, history_cap));
  return 0;
}

int sovlp_batch_create_mixed(int count, const sovlp_lp* lps, const sovlp_params* params,
                             int precision, const int32_t* blocks, void* workspace,
                             int64_t workspace_bytes, int64_t history_cap, sovlp_stream_t stream,
                             sovlp_batch** out) {
  if (count < 1 || !lps || !params || !blocks || !workspace || !out)
    return fail(1, "bad batch arguments");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  for (int i = 0; i < count; ++i)
    if (!batch_block_ok(blocks[i])) return fail(1, "block must be 64, 128 or 256");
  if (int rc = check_params(params, precision)) return rc;
  sovlp_batch* h = new sovlp_batch{};
  int rc = precision == 64
               ? batch_create_T<double>(count, lps, params, blocks, workspace, workspace_bytes,
                                        history_cap, S(stream), h)
               : batch_create_T<float>(count, lps, params, blocks, workspace, workspace_bytes,
                                       history_cap, S(stream), h);
  if (rc) {
    sovlp_batch_destroy(h);
    return rc;
  }
  *out = h;
  return 0;
}

int sovlp_batch_create(int count, const sovlp_lp* lps, const sovlp_params* params,
                       int precision, int block, void* workspace, int64_t workspace_bytes,
                       int64_t history_cap, sovlp_stream_t stream, sovlp_batch** out) {
  if (count < 1) return fail(1, "bad batch arguments");
  if (!batch_block_ok(block)) return fail(1, "block must be 64, 128 or 256");
  std::vector<int32_t> blocks(count, block);
  return sovlp_batch_create_mixed(count, lps, params, precision, blocks.data(), workspace,
                                  workspace_bytes, history_cap, stream, out);
}

int sovlp_batch_run(sovlp_batch* h, double seconds_per_lp, int64_t max_graphs,
                    sovlp_stream_t stream, int64_t* graphs_run) {
  if (!h) return fail(1, "null handle");
  hipStream_t st = S(stream);
  int rc = set_ticks(h->ctl, &h->ticks_set, seconds_per_lp, h->hz, st);
  if (rc) return rc;
  if (h->precision == 64)
    state_reset_kernel<double><<<(h->count + 255) / 256, 256, 0, st>>>(
        (const LpDev<double>*)h->lps, h->count, h->done);
  else
    state_reset_kernel<float><<<(h->count + 255) / 256, 256, 0, st>>>(
        (const LpDev<float>*)h->lps, h->count, h->done);
  HIPCHK(hipGetLastError());
  // one replay stays queued while the host reads the done counter the
  // previous one left (pinned copy + event)
  long long g = 0;
  int cur = 0;
  HIPCHK(hipGraphLaunch(h->exec, st));
  HIPCHK(hipMemcpyAsync(&h->hdone[0], h->done, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(h->ev[0], st));
  ++g;
  for (;;) {
    const int nxt = cur ^ 1;
    const bool more = max_graphs <= 0 || g < max_graphs;
    if (more) {
      HIPCHK(hipGraphLaunch(h->exec, st));
      HIPCHK(hipMemcpyAsync(&h->hdone[nxt], h->done, sizeof(unsigned), hipMemcpyDeviceToHost, st));
      HIPCHK(hipEventRecord(h->ev[nxt], st));
      ++g;
    }
    HIPCHK(hipEventSynchronize(h->ev[cur]));
    if (h->hdone[cur] >= (unsigned)h->count || !more) break;
    cur = nxt;
  }
  HIPCHK(hipStreamSynchronize(st));
  if (graphs_run) *graphs_run = g;
  return 0;
}

int sovlp_batch_infos(sovlp_batch* h, sovlp_info* infos, sovlp_stream_t stream) {
  if (!h || !infos) return fail(1, "null argument");
  std::vector<LpState> s(h->count);
  for (int i = 0; i < h->count; ++i)
    HIPCHK(hipMemcpyAsync(&s[i], h->st[i], sizeof(LpState), hipMemcpyDeviceToHost, S(stream)));
  HIPCHK(hipStreamSynchronize(S(stream)));
  for (int i = 0; i < h->count; ++i) fill_info(s[i], &infos[i]);
  return 0;
}

int sovlp_batch_outputs(sovlp_batch* h, int index, double** x, double** y, double** d) {
  if (!h || index < 0 || index >= h->count) return fail(1, "bad batch index");
  if (x) *x = h->xo[index];
  if (y) *y = h->yo[index];
  if (d) *d = h->dout[index];
  return 0;
}

int sovlp_batch_refinement(sovlp_batch* h, int index, double** rows) {
  if (!h || !rows || index < 0 || index >= h->count) return fail(1, "bad batch index");
  *rows = h->rhist[index];
  return 0;
}

int sovlp_batch_phase_counts(sovlp_batch* h, int64_t* counts, sovlp_stream_t stream) {
  if (!h || !counts) return fail(1, "null argument");
  HIPCHK(hipMemcpyAsync(counts, h->pcount, sizeof(long long) * kPhaseSlots * (size_t)h->count,
                        hipMemcpyDeviceToHost, S(stream)));
  HIPCHK(hipStreamSynchronize(S(stream)));
  return 0;
}

int sovlp_batch_destroy(sovlp_batch* h) {
  if (!h) return 0;
  for (int k = 0; k < 2; ++k) {
    if (h->side[k]) (void)hipStreamDestroy(h->side[k]);
    if (h->join[k]) (void)hipEventDestroy(h->join[k]);
  }
  if (h->fork) (void)hipEventDestroy(h->fork);
  for (hipEvent_t e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->hdone) (void)hipHostFree(h->hdone);
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  if (h->cap) (void)hipStreamDestroy(h->cap);
  delete h;
  return 0;
}

}  // extern "C"

// ====================================================================== server
// A slot table for continuous batching: `slots` LPs share every launch of one
// recorded graph (one workgroup per slot, up to `phases` phases each).  An
// LP is admitted into a free slot between launches (its descriptor and its
// fresh state are copied in stream order), runs until its state reaches
// DONE, and is released by the host after the result is read.  Empty and
// finished slots point at a DONE state in the server's own workspace, so
// their workgroups return at once and never touch a released LP's memory.
struct sovlp_server {
  int slots, precision, block, phases;
  int team;                  // workgroups per slot and phase (1: run kernel)
  double* partials;          // [slots][team][kMaxRed] (team > 1)
  int32_t* parts;            // [slots][2][team + 1] tile partitions of A and A^T (team > 1)
  int32_t* hparts;           // pinned staging of the partitions
  void* lps;                 // LpDev<T>[slots] (device)
  LpState* states;           // [slots] (device)
  long long* pcount;         // [slots][kPhaseSlots] (device)
  StepCtl* ctl;
  unsigned* done;
  char* hstage;              // pinned LpDev<T>[slots]: descriptors being copied in
  LpState* hinit;            // pinned [slots]: fresh states being copied in
  LpState* hsnap;            // pinned [slots]: states after the last step
  std::vector<char> busy, seen_done;
  std::vector<double*> xo, yo, dout, rhist;
  hipGraph_t graph;
  hipGraphExec_t exec;
  hipStream_t cap;
  double hz;
};

namespace {

template <class T>
size_t server_ws_T(int slots, int team) {
  Carver cv{nullptr, 0, 0};
  cv.take<LpDev<T>>(slots);
  cv.take<LpState>(slots);
  cv.take<long long>((size_t)slots * kPhaseSlots);
  cv.take<StepCtl>(1);
  cv.take<unsigned>(64);
  if (team > 1) {
    cv.take<double>((size_t)slots * team * kMaxRed);
    cv.take<int32_t>((size_t)slots * 2 * (team + 1));
  }
  return cv.off + 256;
}

template <class T>
size_t server_lp_ws_T(const sovlp_lp* lp, long long hist_cap) {
  Carver cv{nullptr, 0, 0};
  LpDev<T> L{};
  carve_lp<T>(cv, L, lp->m, lp->n, lp->nnz, hist_cap);
  return cv.off + 256;
}

template <class T>
int server_create_T(int slots, int team, const sovlp_params* prm, int block, int phases, void* ws,
                    int64_t ws_bytes, hipStream_t st, sovlp_server* h) {
  if ((size_t)ws_bytes < server_ws_T<T>(slots, team)) return fail(1, "workspace too small");
  Carver cv{reinterpret_cast<char*>(ws), 0, (size_t)ws_bytes};
  LpDev<T>* d_lps = cv.take<LpDev<T>>(slots);
  LpState* states = cv.take<LpState>(slots);
  long long* pcount = cv.take<long long>((size_t)slots * kPhaseSlots);
  StepCtl* ctl = cv.take<StepCtl>(1);
  unsigned* done = cv.take<unsigned>(64);
  h->team = team;
  if (team > 1) {
    h->partials = cv.take<double>((size_t)slots * team * kMaxRed);
    h->parts = cv.take<int32_t>((size_t)slots * 2 * (team + 1));
    HIPCHK(hipHostMalloc((void**)&h->hparts, sizeof(int32_t) * (size_t)slots * 2 * (team + 1),
                         hipHostMallocDefault));
  }
  h->slots = slots;
  h->precision = sizeof(T) == 8 ? 64 : 32;
  h->block = block;
  h->phases = phases;
  h->lps = d_lps;
  h->states = states;
  h->pcount = pcount;
  h->ctl = ctl;
  h->done = done;
  h->busy.assign(slots, 0);
  h->seen_done.assign(slots, 0);
  h->xo.assign(slots, nullptr);
  h->yo.assign(slots, nullptr);
  h->dout.assign(slots, nullptr);
  h->rhist.assign(slots, nullptr);
  HIPCHK(hipHostMalloc((void**)&h->hstage, sizeof(LpDev<T>) * (size_t)slots, hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void**)&h->hinit, sizeof(LpState) * (size_t)slots, hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void**)&h->hsnap, sizeof(LpState) * (size_t)slots, hipHostMallocDefault));
  LpDev<T>* stage = reinterpret_cast<LpDev<T>*>(h->hstage);
  for (int i = 0; i < slots; ++i) {
    stage[i] = LpDev<T>{};
    stage[i].st = states + i;
    stage[i].pcount = pcount + (size_t)i * kPhaseSlots;
    h->hinit[i] = LpState{};
    h->hinit[i].phase = PH_DONE;
    h->hinit[i].status = ST_RUNNING;
  }
  StepCtl c0{};
  std::memcpy(&c0.P, prm, sizeof(Params));
  const double secs = prm->time_limit > 0.0 ? prm->time_limit : 0.0;
  double hz = 0.0;
  if (int rc = clock_hz(&hz)) return rc;
  h->hz = hz;
  c0.ticks = (long long)std::min(secs * hz, 9.0e18);
  HIPCHK(hipMemcpyAsync(d_lps, stage, sizeof(LpDev<T>) * (size_t)slots, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(states, h->hinit, sizeof(LpState) * (size_t)slots, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(pcount, 0, sizeof(long long) * kPhaseSlots * (size_t)slots, st));
  HIPCHK(hipMemcpyAsync(ctl, &c0, sizeof c0, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(done, 0, sizeof(unsigned), st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
  HIPCHK(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
  if (team > 1) {
    // `phases` pairs: one phase of every slot over `team` workgroups, then one
    // workgroup per slot reduces the partial sums and takes the transition
    constexpr int TN = sizeof(T) == 8 ? kTnnz64 : kTnnz32;
    for (int i = 0; i < phases; ++i) {
      pdlp_slot_phase_kernel<T, kBlock, TN><<<slots * team, kBlock, 0, h->cap>>>(d_lps, team);
      pdlp_slot_decide_kernel<T, kBlock><<<slots, kBlock, 0, h->cap>>>(d_lps, team, ctl, done);
    }
  } else {
    launch_run<T>(block, d_lps, slots, ctl, done, phases, h->cap);
  }
  HIPCHK(hipStreamEndCapture(h->cap, &h->graph));
  HIPCHK(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
  return 0;
}

template <class T>
int server_admit_T(sovlp_server* h, int slot, const sovlp_lp* lp, const double* x0,
                   const double* y0, double omega0, void* ws, int64_t ws_bytes,
                   int64_t hist_cap, hipStream_t st) {
  if ((size_t)ws_bytes < server_lp_ws_T<T>(lp, hist_cap)) return fail(1, "LP workspace too small");
  Carver cv{reinterpret_cast<char*>(ws), 0, (size_t)ws_bytes};
  LpDev<T> L{};
  carve_lp<T>(cv, L, lp->m, lp->n, lp->nnz, hist_cap);
  fill_lp<T>(L, *lp);
  L.st = h->states + slot;   // the state lives in the server's table, not the LP's workspace
  L.pcount = h->pcount + (size_t)slot * kPhaseSlots;
  if (h->team > 1) {
    // equal-cost split of both tile lists over the team (host plan from the
    // device copies of the tile starts and row pointers)
    const int W = h->team;
    std::vector<int32_t> ta(lp->ntiles_a + 1), tt(lp->ntiles_t + 1), rp(lp->m + 1), cp(lp->n + 1);
    HIPCHK(hipMemcpyAsync(ta.data(), lp->tiles_a, sizeof(int32_t) * ta.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(tt.data(), lp->tiles_t, sizeof(int32_t) * tt.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(rp.data(), lp->rp, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(cp.data(), lp->cp, sizeof(int32_t) * cp.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<int64_t> pa(ta.size()), pt(tt.size());
    for (size_t t = 0; t < ta.size(); ++t) pa[t] = rp[ta[t]];
    for (size_t t = 0; t < tt.size(); ++t) pt[t] = cp[tt[t]];
    int32_t* hp = h->hparts + (size_t)slot * 2 * (W + 1);
    host_partition(lp->ntiles_a, ta.data(), pa.data(), W, hp);
    host_partition(lp->ntiles_t, tt.data(), pt.data(), W, hp + W + 1);
    int32_t* dp = h->parts + (size_t)slot * 2 * (W + 1);
    HIPCHK(hipMemcpyAsync(dp, hp, sizeof(int32_t) * 2 * (W + 1), hipMemcpyHostToDevice, st));
    L.partA = dp;
    L.partT = dp + W + 1;
    L.partials = h->partials + (size_t)slot * W * kMaxRed;
  }
  L.wx0 = x0;
  L.wy0 = y0;
  L.womega = omega0 > 0.0 ? omega0 : 0.0;
  LpDev<T>* stage = reinterpret_cast<LpDev<T>*>(h->hstage);
  stage[slot] = L;
  h->hinit[slot] = LpState{};
  h->hinit[slot].phase = PH_INIT;
  h->hinit[slot].status = ST_RUNNING;
  LpDev<T>* d_lps = reinterpret_cast<LpDev<T>*>(h->lps);
  HIPCHK(hipMemcpyAsync(d_lps + slot, &stage[slot], sizeof(LpDev<T>), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(h->states + slot, &h->hinit[slot], sizeof(LpState), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(h->pcount + (size_t)slot * kPhaseSlots, 0, sizeof(long long) * kPhaseSlots, st));
  h->xo[slot] = L.xo;
  h->yo[slot] = L.yo;
  h->dout[slot] = L.dout;
  h->rhist[slot] = L.rhist;
  return 0;
}

}  // namespace

extern "C" {

int sovlp_server_workspace_bytes(int slots, int precision, int64_t* bytes) {
  return sovlp_server_team_workspace_bytes(slots, 1, precision, bytes);
}

int sovlp_server_team_workspace_bytes(int slots, int team, int precision, int64_t* bytes) {
  if (slots < 1 || team < 1 || !bytes) return fail(1, "slots and team must be >= 1");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  *bytes = (int64_t)(precision == 64 ? server_ws_T<double>(slots, team)
                                     : server_ws_T<float>(slots, team));
  return 0;
}

int sovlp_server_lp_bytes(const sovlp_lp* lp, int precision, int64_t history_cap,
                          int64_t* bytes) {
  if (!lp || !bytes) return fail(1, "null argument");
  if (lp->m < 0 || lp->n < 0 || lp->nnz < 0) return fail(1, "negative dimension");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  *bytes = (int64_t)(precision == 64 ? server_lp_ws_T<double>(lp, history_cap)
                                     : server_lp_ws_T<float>(lp, history_cap));
  return 0;
}

static int server_create(int slots, int team, const sovlp_params* params, int precision,
                         int block, int phases_per_launch, void* workspace,
                         int64_t workspace_bytes, sovlp_stream_t stream, sovlp_server** out) {
  if (slots < 1 || team < 1 || !params || !workspace || !out)
    return fail(1, "bad server arguments");
  if (precision != 64 && precision != 32) return fail(1, "precision must be 64 or 32");
  if (!batch_block_ok(block)) return fail(1, "block must be 64, 128 or 256");
  if (phases_per_launch < 1 || phases_per_launch > 4096)
    return fail(1, "phases_per_launch must be in [1, 4096]");
  if ((long long)slots * team > (1LL << 24)) return fail(1, "slots x team too large");
  if (int rc = check_params(params, precision)) return rc;
  sovlp_server* h = new sovlp_server{};
  int rc = precision == 64
               ? server_create_T<double>(slots, team, params, block, phases_per_launch, workspace,
                                         workspace_bytes, S(stream), h)
               : server_create_T<float>(slots, team, params, block, phases_per_launch, workspace,
                                        workspace_bytes, S(stream), h);
  if (rc) {
    sovlp_server_destroy(h);
    return rc;
  }
  *out = h;
  return 0;
}

int sovlp_server_create(int slots, const sovlp_params* params, int precision, int block,
                        int phases_per_launch, void* workspace, int64_t workspace_bytes,
                        sovlp_stream_t stream, sovlp_server** out) {
  return server_create(slots, 1, params, precision, block, phases_per_launch, workspace,
                       workspace_bytes, stream, out);
}

int sovlp_server_create_team(int slots, int team, const sovlp_params* params, int precision,
                             int pairs_per_launch, void* workspace, int64_t workspace_bytes,
                             sovlp_stream_t stream, sovlp_server** out) {
  return server_create(slots, team, params, precision, kBlock, pairs_per_launch, workspace,
                       workspace_bytes, stream, out);
}

int sovlp_server_admit(sovlp_server* h, int slot, const sovlp_lp* lp, const double* x0,
                       const double* y0, double omega0, void* workspace,
                       int64_t workspace_bytes, int64_t history_cap, sovlp_stream_t stream) {
  if (!h || !lp || !workspace) return fail(1, "null argument");
  if (slot < 0 || slot >= h->slots) return fail(1, "slot out of range");
  if (h->busy[slot]) return fail(1, "slot is occupied (release it first)");
  if (lp->m < 0 || lp->n < 0 || lp->nnz < 0) return fail(1, "negative dimension");
  if (lp->nnz >= (int64_t)INT32_MAX) return fail(4, "more than 2^31-1 nonzeros");
  int rc = h->precision == 64
               ? server_admit_T<double>(h, slot, lp, x0, y0, omega0, workspace, workspace_bytes,
                                        history_cap, S(stream))
               : server_admit_T<float>(h, slot, lp, x0, y0, omega0, workspace, workspace_bytes,
                                       history_cap, S(stream));
  if (rc) return rc;
  h->busy[slot] = 1;
  h->seen_done[slot] = 0;
  return 0;
}

int sovlp_server_step(sovlp_server* h, int launches, sovlp_stream_t stream, int32_t* finished,
                      int32_t* running) {
  if (!h) return fail(1, "null handle");
  hipStream_t st = S(stream);
  for (int i = 0; i < launches; ++i) HIPCHK(hipGraphLaunch(h->exec, st));
  HIPCHK(hipMemcpyAsync(h->hsnap, h->states, sizeof(LpState) * (size_t)h->slots,
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  int run = 0;
  for (int i = 0; i < h->slots; ++i) {
    const bool fin = h->busy[i] && h->hsnap[i].phase == PH_DONE;
    if (fin) h->seen_done[i] = 1;
    if (finished) finished[i] = fin ? 1 : 0;
    run += (h->busy[i] && !fin) ? 1 : 0;
  }
  if (running) *running = run;
  return 0;
}

int sovlp_server_info(sovlp_server* h, int slot, sovlp_info* info) {
  if (!h || !info) return fail(1, "null argument");
  if (slot < 0 || slot >= h->slots || !h->busy[slot]) return fail(1, "slot holds no LP");
  fill_info(h->hsnap[slot], info);
  return 0;
}

int sovlp_server_outputs(sovlp_server* h, int slot, double** x, double** y, double** d) {
  if (!h) return fail(1, "null handle");
  if (slot < 0 || slot >= h->slots || !h->busy[slot]) return fail(1, "slot holds no LP");
  if (x) *x = h->xo[slot];
  if (y) *y = h->yo[slot];
  if (d) *d = h->dout[slot];
  return 0;
}

int sovlp_server_refinement(sovlp_server* h, int slot, double** rows) {
  if (!h || !rows) return fail(1, "null argument");
  if (slot < 0 || slot >= h->slots || !h->busy[slot]) return fail(1, "slot holds no LP");
  *rows = h->rhist[slot];
  return 0;
}

int sovlp_server_phase_counts(sovlp_server* h, int slot, int64_t* counts,
                              sovlp_stream_t stream) {
  if (!h || !counts) return fail(1, "null argument");
  if (slot < 0 || slot >= h->slots) return fail(1, "slot out of range");
  HIPCHK(hipMemcpyAsync(counts, h->pcount + (size_t)slot * kPhaseSlots,
                        sizeof(long long) * kPhaseSlots, hipMemcpyDeviceToHost, S(stream)));
  HIPCHK(hipStreamSynchronize(S(stream)));
  return 0;
}

int sovlp_server_release(sovlp_server* h, int slot) {
  if (!h) return fail(1, "null handle");
  if (slot < 0 || slot >= h->slots || !h->busy[slot]) return fail(1, "slot holds no LP");
  if (!h->seen_done[slot]) return fail(1, "the slot's LP has not finished (step until it has)");
  h->busy[slot] = 0;
  h->seen_done[slot] = 0;
  h->xo[slot] = h->yo[slot] = h->dout[slot] = h->rhist[slot] = nullptr;
  return 0;
}

int sovlp_server_clock_hz(sovlp_server* h, double* hz) {
  if (!h || !hz) return fail(1, "null argument");
  *hz = h->hz;
  return 0;
}

int sovlp_server_destroy(sovlp_server* h) {
  if (!h) return 0;
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  if (h->cap) (void)hipStreamDestroy(h->cap);
  if (h->hstage) (void)hipHostFree(h->hstage);
  if (h->hinit) (void)hipHostFree(h->hinit);
  if (h->hsnap) (void)hipHostFree(h->hsnap);
  if (h->hparts) (void)hipHostFree(h->hparts);
  delete h;
  return 0;
}

}  // extern "C"
