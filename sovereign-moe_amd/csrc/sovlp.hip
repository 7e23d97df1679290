// sovlp.hip -- Backend surface, tiled SpMV, certificate and batched node
// relaxation kernels, and their C ABI (gfx950).  The PDLP engines are in
// engine.hpp (compiled per precision by engine_inst.hip) and engine_api.hip.
//
// Kernels:
//   spmv_vector_kernel   Backend surface SpMV, TPR lanes per row
//   elementwise kernels  Backend surface steps (backend.py:204-291)
//   reduce_kernel        Backend surface dot / nrm2 / nrm_inf, fixed order
//   spmv_tiled_kernel    the engines' CSR-Adaptive tile pass as a bare SpMV
#include "common.hpp"

#include <mutex>
#include "certify_kernels.hpp"
#include "bnr_kernels.hpp"

namespace sovh {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
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

}  // namespace sovh

using namespace sovh;

namespace {

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

// out = A X with X dense (cols, k) and out (rows, k), batch index fastest:
// thread g takes row g / k and column g % k and walks the row's nonzeros in
// CSR order from 0, multiplying and adding separately (the reference's
// _csr_spmm, core/sparse.py:72-83), so results equal its CPU ones bit for bit.
template <class T>
__global__ void __launch_bounds__(256)
    spmm_csr_kernel(int rows, int k, const int64_t* __restrict__ rp,
                    const int32_t* __restrict__ ri, const T* __restrict__ rx,
                    const T* __restrict__ X, T* __restrict__ out) {
#pragma clang fp contract(off)
  const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
  if (g >= (long long)rows * k) return;
  const long long row = g / k;
  const int t = (int)(g % k);
  T acc = T(0);
  const int64_t p1 = rp[row + 1];
  int64_t p = rp[row];
  for (; p + 4 <= p1; p += 4) {          // four nonzeros' loads, then the sums in order
    int jj[4];
    T a[4], v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      jj[u] = ri[p + u];
      a[u] = rx[p + u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = X[(long long)jj[u] * k + t];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += a[u] * v[u];
  }
  for (; p < p1; ++p) acc += rx[p] * X[(long long)ri[p] * k + t];
  out[g] = acc;
}

template <class T>
int launch_spmm(int rows, const int64_t* rp, const int32_t* ri, const T* rx, const T* X, T* out,
                int k, hipStream_t st) {
  if (rows < 0 || k < 0) return fail(1, "rows and k must be >= 0");
  if (rows == 0 || k == 0) return 0;
  const long long total = (long long)rows * k;
  if (total > (1LL << 40)) return fail(1, "rows x k too large");
  spmm_csr_kernel<T><<<(unsigned)((total + 255) / 256), 256, 0, st>>>(rows, k, rp, ri, rx, X,
                                                                      out);
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


}  // namespace

extern "C" {

const char* sovlp_last_error(void) { return sovh::g_err.c_str(); }

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

int sovlp_spmm_csr_f64(int rows, const int64_t* rp, const int32_t* ri, const double* rx,
                       const double* X, double* out, int k, sovlp_stream_t stream) {
  return launch_spmm<double>(rows, rp, ri, rx, X, out, k, S(stream));
}
int sovlp_spmm_csr_f32(int rows, const int64_t* rp, const int32_t* ri, const float* rx,
                       const float* X, float* out, int k, sovlp_stream_t stream) {
  return launch_spmm<float>(rows, rp, ri, rx, X, out, k, S(stream));
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

// The residual's launches over a scratch of sovlp_residual_workspace_bytes:
// the long rows' count, their list, then the listing's per-chunk counts.
// Every word the kernels read is written first, so the scratch needs no
// clearing.
static int residual_launch(int rows, const int64_t* rp, const int32_t* ri, const double* rx,
                           const double* x, const double* b, double* out, int* cnt,
                           hipStream_t st) {
  const int groups = std::min(kCertMaxGroups * 4, (rows + kCertBlock - 1) / kCertBlock);
  if (int rc = list_long_lines(rows, rp, cnt + 4 + rows, cnt, cnt + 4, st)) return rc;
  residual_kernel<<<groups, kCertBlock, 0, st>>>(rows, rp, ri, rx, x, b, out);
  HIPCHK(hipGetLastError());
  long_lines_kernel<true><<<kCertMaxGroups, kCertBlock, 0, st>>>(cnt, cnt + 4, rp, ri, rx, x, b,
                                                                   1.0, out);
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
  return residual_launch(rows, rp, ri, rx, x, b, out, static_cast<int*>(workspace), S(stream));
}

}  // extern "C"

// Scratch of sovlp_residual_f64_noalloc: the residual's workspace for up to
// kNoAllocRows rows (4 MiB), one copy on each GPU for the life of the process.
// Its name is a host-side stand-in: host code uses only the device address
// that hipGetSymbolAddress returns for the current device.
constexpr int kNoAllocRows = 1 << 20;
static __device__ __align__(256) int32_t scratch[4 + kNoAllocRows + kCertMaxGroups];

namespace {
std::mutex g_scratch_mu;
std::vector<int*> g_scratch;     // per device: the address of its copy of `scratch`

// The current device's copy of `scratch`, looked up once per device.
int scratch_on_current_device(int** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0) return fail(2, "no current HIP device");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1, nullptr);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(2, "hipGetSymbolAddress returned no address for the scratch");
    g_scratch[dev] = static_cast<int*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}
}  // namespace

extern "C" {

int sovlp_residual_noalloc_max_rows(void) { return kNoAllocRows; }

int sovlp_residual_f64_noalloc(int rows, const int64_t* rp, const int32_t* ri, const double* rx,
                               const double* x, const double* b, double* out,
                               sovlp_stream_t stream) {
  if (rows < 0) return fail(1, "negative row count");
  if (rows > kNoAllocRows)
    return fail(1, "sovlp_residual_f64_noalloc takes at most 1048576 rows "
                   "(use sovlp_residual_f64 with a workspace)");
  if (rows == 0) return 0;
  if (!rp || !ri || !rx || !x || !b || !out) return fail(1, "null argument");
  int* cnt = nullptr;
  if (int rc = scratch_on_current_device(&cnt)) return rc;
  return residual_launch(rows, rp, ri, rx, x, b, out, cnt, S(stream));
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

}  // extern "C"

namespace {

// The launches of one node-relaxation call (start, `iters` iterations of five
// launches, the bound), enqueued on `st`.
int bnr_enqueue(const BnrPtrs& P, double* part_p, double* part_d, double* part_b, int gm,
                int gn, int iters, double growth, double damp, hipStream_t st) {
  const int m = P.m, n = P.n, K = P.K;
  const int64_t nk = (int64_t)n * K;
  const int gk = (K + 63) / 64;
  bnr_zero_kernel<<<gk, 64, 0, st>>>(P.empty, K);
  HIPCHK(hipGetLastError());
  bnr_start_kernel<<<gn, kBnrBlock, 0, st>>>(P);
  HIPCHK(hipGetLastError());
  if (m > 0) bnr_spmm_kernel<<<gm, kBnrBlock, 0, st>>>(m, K, P.KP, P.RB, P.rp, P.ri, P.rx, P.X, P.AX);
  bnr_spmm_kernel<<<gn, kBnrBlock, 0, st>>>(n, K, P.KP, P.RB, P.cp, P.ci, P.cx, P.Y, P.AtY);
  HIPCHK(hipGetLastError());
  if (m == 0) HIPCHK(hipMemsetAsync(P.AtdY, 0, sizeof(double) * nk, st));   // A^T dY = 0
  BnrPtrs Pp = P, Pd = P, Pb = P;
  Pp.part = part_p;
  Pd.part = part_d;
  Pb.part = part_b;
  for (int it = 0; it < iters; ++it) {
    // from the second iteration on, the previous iteration's update runs in
    // the primal launch (its row part in gm extra blocks)
    const int pending = it > 0 ? 1 : 0;
    bnr_primal_kernel<<<gn + (pending && m > 0 ? gm : 0), kBnrBlock, 0, st>>>(Pp, pending, gn);
    if (m > 0) {
      bnr_spmm_kernel<<<gm, kBnrBlock, 0, st>>>(m, K, P.KP, P.RB, P.rp, P.ri, P.rx, P.dX, P.AdX);
      bnr_dual_kernel<<<gm, kBnrBlock, 0, st>>>(Pd);
      bnr_spmm_kernel<<<gn, kBnrBlock, 0, st>>>(n, K, P.KP, P.RB, P.cp, P.ci, P.cx, P.dY, P.AtdY);
    }
    const double t = (double)(it + 2);
    bnr_decide_kernel<<<K, kBnrLanes, 0, st>>>(P, part_p, gn, part_d, m > 0 ? gm : 0,
                                               1.0 - std::pow(t, -growth),
                                               1.0 + std::pow(t, -damp));
    HIPCHK(hipGetLastError());
  }
  if (iters > 0) bnr_update_kernel<<<gn + (m > 0 ? gm : 0), kBnrBlock, 0, st>>>(P, gn);
  // the bound from the final Y: A^T Y into AtdY, then the terms
  if (m > 0) {
    bnr_spmm_kernel<<<gn, kBnrBlock, 0, st>>>(n, K, P.KP, P.RB, P.cp, P.ci, P.cx, P.Y, P.AtdY);
  } else {
    HIPCHK(hipMemsetAsync(P.AtdY, 0, sizeof(double) * nk, st));
  }
  bnr_bound_terms_kernel<<<(m > 0 ? gm : 0) + gn, kBnrBlock, 0, st>>>(Pb, m > 0 ? gm : 0);
  bnr_bound_final_kernel<<<K, kBnrLanes, 0, st>>>(Pb, m > 0 ? gm : 0, gn);
  HIPCHK(hipGetLastError());
  return 0;
}

// A call's launches depend only on the call's arguments, so a repeated call
// (a tree visiting batch after batch through the same buffers) replays a
// HIP graph of them: captured on the second call with the same arguments,
// replayed from the third.  At most kBnrGraphs argument sets per process
// are kept; beyond that, and for LPs without rows (whose zeroing is a
// memset), calls launch directly.
constexpr int kBnrGraphs = 16;
struct BnrGraph {
  std::vector<uint64_t> key;
  int seen = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};
std::mutex g_bnr_mu;
std::vector<BnrGraph> g_bnr_graphs;
std::vector<hipStream_t> g_bnr_capture;   // per device: the capture stream

std::vector<uint64_t> bnr_key(const sovlp_bnr_batch* b, const void* ws, int dev) {
  auto u = [](const void* p) { return (uint64_t)reinterpret_cast<uintptr_t>(p); };
  auto f = [](double v) {
    uint64_t r;
    std::memcpy(&r, &v, sizeof r);
    return r;
  };
  return {(uint64_t)dev, (uint64_t)b->m, (uint64_t)b->n, (uint64_t)b->K, (uint64_t)b->iters,
          u(b->rp), u(b->cp), u(b->ri), u(b->ci), u(b->rx), u(b->cx), u(b->c), u(b->rl),
          u(b->ru), u(b->L), u(b->U), u(b->X), u(b->Y), u(b->eta), u(b->omega),
          f(b->step_growth), f(b->step_damp), u(b->bounds), u(b->empty), f(b->dual_sign),
          u(ws)};
}

}  // namespace

extern "C" {

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
  if (m == 0)
    return bnr_enqueue(P, part_p, part_d, part_b, gm, gn, b->iters, b->step_growth,
                       b->step_damp, st);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  const std::vector<uint64_t> key = bnr_key(b, workspace, dev);
  std::lock_guard<std::mutex> lock(g_bnr_mu);
  BnrGraph* g = nullptr;
  for (BnrGraph& e : g_bnr_graphs)
    if (e.key == key) g = &e;
  if (!g && (int)g_bnr_graphs.size() < kBnrGraphs) {
    g_bnr_graphs.push_back(BnrGraph{key});
    g = &g_bnr_graphs.back();
  }
  if (!g || ++g->seen < 2)
    return bnr_enqueue(P, part_p, part_d, part_b, gm, gn, b->iters, b->step_growth,
                       b->step_damp, st);
  if (!g->exec) {
    if ((int)g_bnr_capture.size() <= dev) g_bnr_capture.resize(dev + 1, nullptr);
    if (!g_bnr_capture[dev])
      HIPCHK(hipStreamCreateWithFlags(&g_bnr_capture[dev], hipStreamNonBlocking));
    hipStream_t cs = g_bnr_capture[dev];
    HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    const int rc = bnr_enqueue(P, part_p, part_d, part_b, gm, gn, b->iters, b->step_growth,
                               b->step_damp, cs);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (ec != hipSuccess) return fail(2, std::string("bnr graph capture: ") + hipGetErrorString(ec));
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return fail(2, std::string("bnr graph instantiate: ") + hipGetErrorString(ei));
    }
    g->graph = graph;
    g->exec = exec;
  }
  HIPCHK(hipGraphLaunch(g->exec, st));
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

