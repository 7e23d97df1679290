// certify_kernels.hpp -- the certificate of a point (x, y) on the original LP.
//
//  * compensated residual b - A x: one lane per row walks the row in CSR order
//    with Dekker's two_product (split 2^27 + 1) and Knuth's two_sum, the
//    operations and order of the reference's _residual_csr
//    (numerics/refine.py:55-84), so each row's result is bit for bit the
//    reference's.  Contraction is off: a fused multiply-add would change the
//    rounding of the products these transformations take apart.
//  * certified_bound (mip/safebound.py:252-307 with safe_dual_bound strict,
//    102-143): the Neumaier-Shcherbina bound of min c'x from any y,
//      sum_i (yy_i > 0 ? yy_i rl_i : yy_i ru_i) + sum_j (d_j > 0 ? d_j lo_j : d_j hi_j),
//    d = c - A^T yy (each column summed in CSC order, as _csc_spmv_t in
//    core/sparse.py:61-69), less 4 eps sum |terms|.  The two sums are
//    double-double (two_sum accumulators, fixed-order trees), where the
//    reference runs a sequential Neumaier sum; both are exact to a few ulps.
//  * the feasibility figures of bench/verify.py:208-224: column-bound and
//    row violations (row activities compensated), and c'x compensated.
// A line (row or column) of up to kLongLine nonzeros is walked by one lane;
// a longer one by a workgroup that stages its products in LDS and has one
// lane accumulate them in order: the same operations either way.
// Launch sequence (one stream): long rows, rows, long columns, columns,
// column terms, final.  Every sum is a fixed tree over a grid fixed by the
// sizes, so results repeat bit for bit.
#pragma once
#include "sov_device.hpp"

namespace sov {

constexpr double kDekkerSplit = 134217729.0;
constexpr int kCertBlock = 256;
constexpr int kCertMaxGroups = 1024;
// Lines (rows of the residual, columns of A^T y) longer than this are left by
// the lane-per-line kernels to long_lines_kernel: one workgroup per line
// computes the products (and their exact errors) into LDS, and one lane then
// accumulates them in the line's order, the same operations as the lane walk
constexpr int kLongLine = 128;
constexpr int kLineStage = 2048;

// Dekker's two_product as refine.py:59-65 writes it: err = a xv - fl(a xv) exactly
__device__ __forceinline__ double dekker_err(double a, double xv, double prod) {
#pragma clang fp contract(off)
  const double ca = kDekkerSplit * a;
  const double ah = ca - (ca - a);
  const double al = a - ah;
  const double cb = kDekkerSplit * xv;
  const double bh = cb - (cb - xv);
  const double bl = xv - bh;
  return ((ah * bh - prod) + ah * bl + al * bh) + al * bl;
}

// refine.py:55-84 for row i: b_i - (s + comp) as the reference rounds it
__device__ __forceinline__ double compensated_row(const int64_t* __restrict__ rp,
                                                  const int32_t* __restrict__ ri,
                                                  const double* __restrict__ rx,
                                                  const double* __restrict__ x, double bi,
                                                  long long i) {
#pragma clang fp contract(off)
  double s = 0.0, comp = 0.0;
  const long long p1 = rp[i + 1];
  long long p = rp[i];
  // the loads of 8 nonzeros are issued before their arithmetic (they do not
  // depend on it); the arithmetic stays in row order
  for (; p + 8 <= p1; p += 8) {
    double a[8], xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = rx[p + u];
    int jj[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) jj[u] = ri[p + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = x[jj[u]];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double prod = a[u] * xv[u];
      const double err = dekker_err(a[u], xv[u], prod);
      const double t = s + prod;
      const double bb = t - s;
      const double e = (s - (t - bb)) + (prod - bb);
      s = t;
      comp += e + err;
    }
  }
  for (; p < p1; ++p) {
    const double a = rx[p], xv = x[ri[p]];
    const double prod = a * xv;
    const double err = dekker_err(a, xv, prod);
    const double t = s + prod;
    const double bb = t - s;
    const double e = (s - (t - bb)) + (prod - bb);
    s = t;
    comp += e + err;
  }
  const double t = bi - s;
  const double bb = t - bi;
  const double e = (bi - (t - bb)) + (-s - bb);
  return t + (e - comp);
}

__device__ __forceinline__ bool long_line(const int64_t* ptr, long long i) {
  return ptr[i + 1] - ptr[i] > kLongLine;
}

// Short rows: one lane each (long ones are long_lines_kernel's).  b null: b = 0.
// sign -1 stores -(b - A x).
__global__ void __launch_bounds__(kCertBlock)
    residual_kernel(int rows, const int64_t* __restrict__ rp, const int32_t* __restrict__ ri,
                    const double* __restrict__ rx, const double* __restrict__ x,
                    const double* __restrict__ b, double* __restrict__ out) {
  const long long stride = (long long)gridDim.x * kCertBlock;
  for (long long i = (long long)blockIdx.x * kCertBlock + threadIdx.x; i < rows; i += stride)
    if (!long_line(rp, i)) out[i] = compensated_row(rp, ri, rx, x, b[i], i);
}

// The lines longer than kLongLine, listed in index order (each line's result
// depends on the line alone) so that the long-line kernel can deal them out
// one per workgroup.  Three launches over G workgroups, workgroup g taking
// lines [g chunk, (g + 1) chunk): the long lines of each chunk counted, the
// counts scanned by one workgroup (offsets in place, the total into *total),
// each chunk's lines written from its offset.  No atomics: same-address
// atomics serialise beyond the per-XCD L2, and a wave-aggregated atomic list
// of 331 long columns took 55 us.
__device__ __forceinline__ long long chunk_end(int lines, int chunk, long long b0) {
  const long long e = b0 + chunk;
  return e < lines ? e : (long long)lines;
}

__global__ void __launch_bounds__(kCertBlock)
    long_count_kernel(int lines, int chunk, const int64_t* __restrict__ ptr, int* counts) {
  __shared__ int wsum[kCertBlock / kWave];
  const long long b0 = (long long)blockIdx.x * chunk, b1 = chunk_end(lines, chunk, b0);
  int c = 0;
  for (long long i = b0 + threadIdx.x; i < b1; i += kCertBlock) c += long_line(ptr, i) ? 1 : 0;
  for (int o = kWave / 2; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kCertBlock / kWave; ++w) t += wsum[w];
    counts[blockIdx.x] = t;
  }
}

// groups <= 4 * kCertBlock; one workgroup
__global__ void __launch_bounds__(kCertBlock)
    long_scan_kernel(int groups, int* counts, int* total) {
  __shared__ int sc[kCertBlock];
  const int t = (int)threadIdx.x;
  int v[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int g = 4 * t + k;
    v[k] = g < groups ? counts[g] : 0;
    s += v[k];
  }
  sc[t] = s;
  __syncthreads();
  for (int o = 1; o < kCertBlock; o <<= 1) {
    const int add = t >= o ? sc[t - o] : 0;
    __syncthreads();
    sc[t] += add;
    __syncthreads();
  }
  int base = sc[t] - s;                    // exclusive prefix of thread t's four counts
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int g = 4 * t + k;
    if (g < groups) counts[g] = base;
    base += v[k];
  }
  if (t == kCertBlock - 1) *total = sc[t];
}

__global__ void __launch_bounds__(kCertBlock)
    long_write_kernel(int lines, int chunk, const int64_t* __restrict__ ptr,
                      const int* __restrict__ offsets, int* list) {
  __shared__ int wcnt[kCertBlock / kWave];
  const long long b0 = (long long)blockIdx.x * chunk, b1 = chunk_end(lines, chunk, b0);
  const int lane = (int)threadIdx.x & (kWave - 1), w = (int)threadIdx.x / kWave;
  int base = offsets[blockIdx.x];
  for (long long i0 = b0; i0 < b1; i0 += kCertBlock) {
    const long long i = i0 + threadIdx.x;
    const bool lng = i < b1 && long_line(ptr, i);
    const unsigned long long mask = __ballot(lng);
    if (lane == 0) wcnt[w] = (int)__popcll(mask);
    __syncthreads();
    int before = 0, all = 0;
    for (int u = 0; u < kCertBlock / kWave; ++u) {
      before += u < w ? wcnt[u] : 0;
      all += wcnt[u];
    }
    const int pos = base + before + (int)__popcll(mask & ((1ull << lane) - 1ull));
    if (lng && pos < lines) list[pos] = (int)i;
    base += all;
    __syncthreads();                     // wcnt is rewritten by the next round
  }
}

__global__ void zero_ints_kernel(int* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// Stage nonzeros [c0, c0 + len) of a line: products (and their exact errors)
// into LDS, by the threads `first`.. of the workgroup
template <bool COMP>
__device__ __forceinline__ void stage_line(const int32_t* __restrict__ idx,
                                           const double* __restrict__ val,
                                           const double* __restrict__ x, long long c0, int len,
                                           int first, double* sp, double* se) {
#pragma clang fp contract(off)
  constexpr int N = kCertBlock - kWave;
  constexpr int PER = (kLineStage + N - 1) / N;
  const int t = (int)threadIdx.x - first;
  int jj[PER];
  double a[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int k = t + u * N;
    const int kc = k < len ? k : len - 1;         // in the line; discarded past len
    jj[u] = idx[c0 + kc];
    a[u] = val[c0 + kc];
  }
  double xv[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) xv[u] = x[jj[u]];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int k = t + u * N;
    if (k < len) {
      const double prod = a[u] * xv[u];
      sp[k] = prod;
      if constexpr (COMP) se[k] = dekker_err(a[u], xv[u], prod);
    }
  }
}

// The listed long lines, workgroup g taking entries g, g + gridDim.x, ...
// Waves 1..3 stage chunk c + 1 while thread 0 sums chunk c in line order
// (two LDS buffers, one barrier per chunk).
//   COMP true : out[i] = sign * (b_i - A_i x), compensated (refine.py:55-84);
//               b null means b = 0
//   COMP false: out[i] = sign b_i - sum_p A_ip x_p, the plain sequential sum
//               of _csc_spmv_t (sparse.py:61-69) subtracted from sg c_i
template <bool COMP>
__global__ void __launch_bounds__(kCertBlock)
    long_lines_kernel(const int* __restrict__ count, const int* __restrict__ list,
                      const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                      const double* __restrict__ val, const double* __restrict__ x,
                      const double* __restrict__ b, double sign, double* __restrict__ out) {
  __shared__ double sp[2][kLineStage];
  __shared__ double se[COMP ? 2 : 1][COMP ? kLineStage : 1];
  const int nl = *count;
  const bool stager = threadIdx.x >= kWave;
  for (int q = blockIdx.x; q < nl; q += gridDim.x) {
    const long long line = list[q];
    const long long p0 = ptr[line], p1 = ptr[line + 1];
    const int chunks = (int)((p1 - p0 + kLineStage - 1) / kLineStage);
    auto chunk_len = [&](int c) {
      const long long r = p1 - (p0 + (long long)c * kLineStage);
      return (int)(r < kLineStage ? r : kLineStage);
    };
    if (stager) stage_line<COMP>(idx, val, x, p0, chunk_len(0), kWave, sp[0], se[0]);
    __syncthreads();
    double s = 0.0, comp = 0.0;        // thread 0's accumulators
    for (int c = 0; c < chunks; ++c) {
      const int cur = c & 1;
      if (stager) {
        if (c + 1 < chunks)
          stage_line<COMP>(idx, val, x, p0 + (long long)(c + 1) * kLineStage,
                           chunk_len(c + 1), kWave, sp[cur ^ 1], se[COMP ? cur ^ 1 : 0]);
      } else if (threadIdx.x == 0) {
#pragma clang fp contract(off)
        const int len = chunk_len(c);
        const double* bp = sp[cur];
        const double* be = se[COMP ? cur : 0];
        int k = 0;
        for (; k + 8 <= len; k += 8) {
          double v[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) v[u] = bp[k + u];
          if constexpr (COMP) {
            double ev[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) ev[u] = be[k + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const double t = s + v[u];
              const double bb = t - s;
              const double e = (s - (t - bb)) + (v[u] - bb);
              s = t;
              comp += e + ev[u];
            }
          } else {
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
          }
        }
        for (; k < len; ++k) {
          if constexpr (COMP) {
            const double prod = bp[k];
            const double t = s + prod;
            const double bb = t - s;
            const double e = (s - (t - bb)) + (prod - bb);
            s = t;
            comp += e + be[k];
          } else {
            s += bp[k];
          }
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma clang fp contract(off)
      if constexpr (COMP) {
        const double bi = b ? b[line] : 0.0;
        const double t = bi - s;
        const double bb = t - bi;
        const double e = (bi - (t - bb)) + (-s - bb);
        out[line] = sign * (t + (e - comp));
      } else {
        out[line] = sign * b[line] - s;
      }
    }
  }
}

// ---------------------------------------------------------------- double-double
struct DD {
  double h, l;
};

__device__ __forceinline__ DD dd_add(DD a, double b) {   // a + b, b's error kept in l
#pragma clang fp contract(off)
  const double s = a.h + b;
  const double bb = s - a.h;
  const double e = (a.h - (s - bb)) + (b - bb);
  return DD{s, a.l + e};
}

__device__ __forceinline__ DD dd_merge(DD a, DD b) {
#pragma clang fp contract(off)
  const double s = a.h + b.h;
  const double bb = s - a.h;
  double e = (a.h - (s - bb)) + (b.h - bb);
  e += a.l + b.l;
  const double h = s + e;
  return DD{h, e - (h - s)};
}

// value of term coef * (bound it points at); -inf at an infinite bound
// (safebound.py:93-99)
__device__ __forceinline__ double cert_term(double coef, double lo, double hi) {
  const bool pos = coef > 0.0;
  const double chosen = pos ? lo : hi;
  const bool bad = (pos ? lo <= -kInf : hi >= kInf) && coef != 0.0;
  return bad ? -__builtin_inf() : coef * chosen;
}

// Workgroup combine, fixed order: lanes by xor butterfly, then waves in order.
// Every thread returns the workgroup's value.
template <class V, class Op>
__device__ __forceinline__ V cert_block_reduce(V v, Op op, V* lds) {
  constexpr int NW = kCertBlock / kWave;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    V u;
    if constexpr (sizeof(V) == sizeof(DD)) {
      u.h = __shfl_xor(v.h, o);
      u.l = __shfl_xor(v.l, o);
    } else {
      u = __shfl_xor(v, o);
    }
    v = op(v, u);
  }
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if ((threadIdx.x % kWave) == 0) lds[w] = v;
  __syncthreads();
  V r = lds[0];
  for (int q = 1; q < NW; ++q) r = op(r, lds[q]);
  return r;
}

struct CertPtrs {
  int m, n, maximise;
  double dual_tol;
  const int64_t *rp, *cp;
  const int32_t *ri, *ci;
  const double *rx, *cx, *c, *rl, *ru, *lo, *hi, *x, *y;
  double *yy, *d, *act;
  double *p1, *p2, *p3;     // per-group partials of the three passes
  double* res;              // the result (8 doubles)
};

// rows: activity (compensated), violation, yy, row term
// p1[g]: {sum_h, sum_l, mag, viol, nonfinite}
__global__ void __launch_bounds__(kCertBlock) cert_rows_kernel(CertPtrs P) {
  __shared__ DD ldd[kCertBlock / kWave];
  __shared__ double lmx[kCertBlock / kWave];
  const double sg = P.maximise ? -1.0 : 1.0;
  DD sum{0.0, 0.0};
  double mag = 0.0, viol = 0.0, nonfinite = 0.0;
  const long long stride = (long long)gridDim.x * kCertBlock;
  for (long long i = (long long)blockIdx.x * kCertBlock + threadIdx.x; i < P.m; i += stride) {
    double a;
    if (long_line(P.rp, i)) {
      a = P.act[i];                    // long_lines_kernel's
    } else {
      a = -compensated_row(P.rp, P.ri, P.rx, P.x, 0.0, i);
      P.act[i] = a;
    }
    const double rl = P.rl[i], ru = P.ru[i];
    viol = fmax(viol, fmax(rl - a, 0.0));
    viol = fmax(viol, fmax(a - ru, 0.0));
    double v = sg * P.y[i];
    if (v > 0.0 && rl <= -kInf) v = 0.0;
    if (v < 0.0 && ru >= kInf) v = 0.0;
    P.yy[i] = v;
    const double t = cert_term(v, rl, ru);
    if (!isfinite(t)) {
      nonfinite = 1.0;
    } else {
      sum = dd_add(sum, t);
      mag += fabs(t);
    }
  }
  sum = cert_block_reduce(sum, [](DD a, DD b) { return dd_merge(a, b); }, ldd);
  mag = cert_block_reduce(mag, [](double a, double b) { return a + b; }, lmx);
  viol = cert_block_reduce(viol, [](double a, double b) { return fmax(a, b); }, lmx);
  nonfinite = cert_block_reduce(nonfinite, [](double a, double b) { return fmax(a, b); }, lmx);
  if (threadIdx.x == 0) {
    double* o = P.p1 + 5 * blockIdx.x;
    o[0] = sum.h; o[1] = sum.l; o[2] = mag; o[3] = viol; o[4] = nonfinite;
  }
}

__device__ __forceinline__ bool points_at_infinity(double d, double lo, double hi) {
  return (d > 0.0 && lo <= -kInf) || (d < 0.0 && hi >= kInf);
}

// columns: d = sg c - A^T yy (CSC order), column violation, |c| max, the
// largest |d| pointing at an infinite bound, c'x (compensated)
// p2[g]: {cmax, pert, anybad, viol, obj_h, obj_l}
__global__ void __launch_bounds__(kCertBlock) cert_cols_kernel(CertPtrs P) {
  __shared__ DD ldd[kCertBlock / kWave];
  __shared__ double lmx[kCertBlock / kWave];
  const double sg = P.maximise ? -1.0 : 1.0;
  double cmax = 0.0, pert = 0.0, anybad = 0.0, viol = 0.0;
  DD obj{0.0, 0.0};
  const long long stride = (long long)gridDim.x * kCertBlock;
  for (long long j = (long long)blockIdx.x * kCertBlock + threadIdx.x; j < P.n; j += stride) {
    const double cj = P.c[j], xj = P.x[j], lo = P.lo[j], hi = P.hi[j];
    double dj;
    if (long_line(P.cp, j)) {
      dj = P.d[j];                     // long_lines_kernel's
    } else {
#pragma clang fp contract(off)
      double acc = 0.0;
      const long long q1 = P.cp[j + 1];
      long long q = P.cp[j];
      for (; q + 4 <= q1; q += 4) {     // 4 nonzeros' loads, then the sums in order
        double a[4], yv[4];
        int ii[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          a[u] = P.cx[q + u];
          ii[u] = P.ci[q + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) yv[u] = P.yy[ii[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc += a[u] * yv[u];
      }
      for (; q < q1; ++q) acc += P.cx[q] * P.yy[P.ci[q]];
      dj = sg * cj - acc;
      P.d[j] = dj;
    }
    cmax = fmax(cmax, fabs(cj));
    if (points_at_infinity(dj, lo, hi)) {
      anybad = 1.0;
      pert = fmax(pert, fabs(dj));
    }
    viol = fmax(viol, fmax(lo - xj, 0.0));
    viol = fmax(viol, fmax(xj - hi, 0.0));
    const double p = cj * xj;
    const double err = fma(cj, xj, -p);          // exact error of the product
    obj = dd_add(obj, p);
    obj.l += err;
  }
  obj = cert_block_reduce(obj, [](DD a, DD b) { return dd_merge(a, b); }, ldd);
  cmax = cert_block_reduce(cmax, [](double a, double b) { return fmax(a, b); }, lmx);
  pert = cert_block_reduce(pert, [](double a, double b) { return fmax(a, b); }, lmx);
  anybad = cert_block_reduce(anybad, [](double a, double b) { return fmax(a, b); }, lmx);
  viol = cert_block_reduce(viol, [](double a, double b) { return fmax(a, b); }, lmx);
  if (threadIdx.x == 0) {
    double* o = P.p2 + 6 * blockIdx.x;
    o[0] = cmax; o[1] = pert; o[2] = anybad; o[3] = viol; o[4] = obj.h; o[5] = obj.l;
  }
}

// column terms, after the absorb decision of safebound.py:299-306 (each group
// takes the maxima of every column group's partials: order-free)
// p3[g]: {sum_h, sum_l, mag, nonfinite}
__global__ void __launch_bounds__(kCertBlock) cert_colterms_kernel(CertPtrs P, int groups2) {
  __shared__ DD ldd[kCertBlock / kWave];
  __shared__ double lmx[kCertBlock / kWave];
  double cmax = 0.0, pert = 0.0, anybad = 0.0;
  for (int g = threadIdx.x; g < groups2; g += kCertBlock) {
    cmax = fmax(cmax, P.p2[6 * g]);
    pert = fmax(pert, P.p2[6 * g + 1]);
    anybad = fmax(anybad, P.p2[6 * g + 2]);
  }
  cmax = cert_block_reduce(cmax, [](double a, double b) { return fmax(a, b); }, lmx);
  pert = cert_block_reduce(pert, [](double a, double b) { return fmax(a, b); }, lmx);
  anybad = cert_block_reduce(anybad, [](double a, double b) { return fmax(a, b); }, lmx);
  const bool absorb = anybad > 0.0 && pert <= P.dual_tol * (1.0 + cmax);
  DD sum{0.0, 0.0};
  double mag = 0.0, nonfinite = 0.0;
  const long long stride = (long long)gridDim.x * kCertBlock;
  for (long long j = (long long)blockIdx.x * kCertBlock + threadIdx.x; j < P.n; j += stride) {
    const double lo = P.lo[j], hi = P.hi[j];
    double dj = P.d[j];
    if (absorb && points_at_infinity(dj, lo, hi)) dj = 0.0;
    const double t = cert_term(dj, lo, hi);
    if (!isfinite(t)) {
      nonfinite = 1.0;
    } else {
      sum = dd_add(sum, t);
      mag += fabs(t);
    }
  }
  sum = cert_block_reduce(sum, [](DD a, DD b) { return dd_merge(a, b); }, ldd);
  mag = cert_block_reduce(mag, [](double a, double b) { return a + b; }, lmx);
  nonfinite = cert_block_reduce(nonfinite, [](double a, double b) { return fmax(a, b); }, lmx);
  if (threadIdx.x == 0) {
    double* o = P.p3 + 4 * blockIdx.x;
    o[0] = sum.h; o[1] = sum.l; o[2] = mag; o[3] = nonfinite;
  }
}

// one workgroup: the partials in group order, then the bound.
// res: {bound, pert, objective, col_viol, row_viol, vacuous}
__global__ void __launch_bounds__(kCertBlock)
    cert_final_kernel(CertPtrs P, int groups1, int groups2, double offset) {
  __shared__ DD ldd[kCertBlock / kWave];
  __shared__ double lmx[kCertBlock / kWave];
  DD rs{0.0, 0.0}, cs{0.0, 0.0}, obj{0.0, 0.0};
  double mag = 0.0, rowv = 0.0, colv = 0.0, nonfinite = 0.0, cmax = 0.0, pert = 0.0,
         anybad = 0.0;
  for (int g = threadIdx.x; g < groups1; g += kCertBlock) {
    const double* o = P.p1 + 5 * g;
    rs = dd_merge(rs, DD{o[0], o[1]});
    mag += o[2];
    rowv = fmax(rowv, o[3]);
    nonfinite = fmax(nonfinite, o[4]);
  }
  for (int g = threadIdx.x; g < groups2; g += kCertBlock) {
    const double* o = P.p2 + 6 * g;
    cmax = fmax(cmax, o[0]);
    pert = fmax(pert, o[1]);
    anybad = fmax(anybad, o[2]);
    colv = fmax(colv, o[3]);
    obj = dd_merge(obj, DD{o[4], o[5]});
    const double* q = P.p3 + 4 * g;
    cs = dd_merge(cs, DD{q[0], q[1]});
    mag += q[2];
    nonfinite = fmax(nonfinite, q[3]);
  }
  auto ddop = [](DD a, DD b) { return dd_merge(a, b); };
  auto sum = [](double a, double b) { return a + b; };
  auto mx = [](double a, double b) { return fmax(a, b); };
  rs = cert_block_reduce(rs, ddop, ldd);
  cs = cert_block_reduce(cs, ddop, ldd);
  obj = cert_block_reduce(obj, ddop, ldd);
  mag = cert_block_reduce(mag, sum, lmx);
  rowv = cert_block_reduce(rowv, mx, lmx);
  colv = cert_block_reduce(colv, mx, lmx);
  nonfinite = cert_block_reduce(nonfinite, mx, lmx);
  pert = cert_block_reduce(pert, mx, lmx);
  anybad = cert_block_reduce(anybad, mx, lmx);
  if (threadIdx.x == 0) {
    const double sg = P.maximise ? -1.0 : 1.0;
    double b;
    if (nonfinite > 0.0) {
      b = -__builtin_inf();
    } else {
      const DD t = dd_merge(rs, cs);
      b = (t.h + t.l) - 4.0 * 2.220446049250313e-16 * mag;
    }
    b = b + sg * offset;
    P.res[0] = P.maximise ? -b : b;
    P.res[1] = anybad > 0.0 ? pert : 0.0;
    P.res[2] = (obj.h + obj.l) + offset;
    P.res[3] = colv;
    P.res[4] = rowv;
    P.res[5] = nonfinite;
  }
}

}  // namespace sov
