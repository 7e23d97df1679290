// bnr_kernels.hpp -- batched node relaxation: K LPs that share A, c and the row
// bounds and differ only in their column bounds (the nodes of a
// branch-and-bound frontier), advanced together by a fixed number of PDHG
// iterations with a per-node adaptive step (reference mip/bnr.py:176-301),
// then bounded by the batched safe dual bound (mip/safebound.py:194-235).
//
// Layout: every (length, K) block is row-major with the node index fastest,
// so the K values of one row or column are one contiguous, coalesced run.  A
// workgroup covers RB = 256 / KP rows (KP = K rounded up to a power of two,
// at most 256) times KP lanes.  The shared-A product walks each row's
// nonzeros in CSR order for every node, the reference's _csr_spmm order
// (core/sparse.py:72-83).  Per-node sums over rows are a fixed tree (lanes of
// one node within a workgroup, then the workgroups in order), so results
// repeat bit for bit.
#pragma once
#include "sov_device.hpp"

namespace sov {

constexpr int kBnrBlock = 256;
constexpr int kBnrLanes = 256;    // lanes of the per-node cross-group sums (64: slower)
constexpr double kQuotientMax = 0x1p1000;   // core/tolerances.py:23

struct BnrPtrs {
  int m, n, K, KP, RB;
  const int64_t *rp, *cp;
  const int32_t *ri, *ci;
  const double *rx, *cx, *c, *rl, *ru, *L, *U;
  double *X, *Y, *Xn, *Yn, *dX, *dY, *AX, *AtY, *AdX, *AtdY;
  double *eta, *omega;
  double* part;          // per-group per-node partial sums: [group][3][K]
  int* accept;           // [K]
  int* empty;            // [K]
  double* bounds;        // [K]
  double growth, damp;
  double ysign;          // the bound's dual: ysign * Y
};

__device__ __forceinline__ bool bnr_lane(const BnrPtrs& P, int len, long long* row, int* k) {
  const int t = threadIdx.x;
  *k = t & (P.KP - 1);
  *row = (long long)blockIdx.x * P.RB + t / P.KP;
  return *row < len && *k < P.K;
}

// out = A V over `rows` rows (CSR), V and out node-fastest
__global__ void __launch_bounds__(kBnrBlock)
    bnr_spmm_kernel(int rows, int K, int KP, int RB, const int64_t* __restrict__ rp,
                    const int32_t* __restrict__ ri, const double* __restrict__ rx,
                    const double* __restrict__ V, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int t = threadIdx.x;
  const int k = t & (KP - 1);
  const long long row = (long long)blockIdx.x * RB + t / KP;
  if (row >= rows || k >= K) return;
  double acc = 0.0;
  const long long p1 = rp[row + 1];
  long long p = rp[row];
  for (; p + 4 <= p1; p += 4) {        // 4 nonzeros' loads, then the sums in order
    int jj[4];
    double a[4], v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      jj[u] = ri[p + u];
      a[u] = rx[p + u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = V[(long long)jj[u] * K + k];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += a[u] * v[u];
  }
  for (; p < p1; ++p) acc += rx[p] * V[(long long)ri[p] * K + k];
  out[row * K + k] = acc;
}

// Workgroup sum of each node's lane values (rows of the group), into
// part[group * 3 * K + slot * K + k]: lanes of one node sit KP apart.
__device__ __forceinline__ void bnr_group_sums(const BnrPtrs& P, int nslots, const double* v,
                                               double* lds) {
  const int t = threadIdx.x;
  for (int s = 0; s < nslots; ++s) {
    __syncthreads();
    lds[t] = v[s];
    __syncthreads();
    if (t < P.KP) {
      double a = 0.0;
      for (int r = 0; r < P.RB; ++r) a += lds[r * P.KP + t];
      if (t < P.K) P.part[((long long)blockIdx.x * 3 + s) * P.K + t] = a;
    }
  }
}

// X clamped into each node's box; a node whose box is empty is marked
// (bnr.py:190, 205)
__global__ void __launch_bounds__(kBnrBlock) bnr_start_kernel(BnrPtrs P) {
  long long j;
  int k;
  if (!bnr_lane(P, P.n, &j, &k)) return;
  const long long q = j * P.K + k;
  const double lo = P.L[q], hi = P.U[q];
  if (lo > hi + 1e-12) atomicMax(P.empty + k, 1);
  P.X[q] = fmin(fmax(P.X[q], lo), hi);
}

// Xn = clip(X - tau (c + AtY), L, U), dX = Xn - X; sum dX^2 per node.  With
// `pending`, the previous iteration's accepted nodes first take their trial
// point (bnr_update_kernel, fused): X, AtY in the groups_n column blocks, Y, AX
// in the blocks after them (launched only then; they write no group sums).
__global__ void __launch_bounds__(kBnrBlock)
    bnr_primal_kernel(BnrPtrs P, int pending, int groups_n) {
  __shared__ double lds[kBnrBlock];
  if ((int)blockIdx.x >= groups_n) {
    const int t = threadIdx.x;
    const int kr = t & (P.KP - 1);
    const long long r = (long long)((int)blockIdx.x - groups_n) * P.RB + t / P.KP;
    if (kr >= P.K || r >= P.m || !P.accept[kr]) return;
    const long long q = r * P.K + kr;
    P.Y[q] = P.Yn[q];
    P.AX[q] = P.AX[q] + P.AdX[q];
    return;
  }
  long long j;
  int k;
  double v[1] = {0.0};
  if (bnr_lane(P, P.n, &j, &k)) {
#pragma clang fp contract(off)
    const long long q = j * P.K + k;
    const double tau = P.eta[k] / P.omega[k];
    double x, aty;
    if (pending && P.accept[k]) {
      x = P.Xn[q];
      aty = P.AtY[q] + P.AtdY[q];
      P.X[q] = x;
      P.AtY[q] = aty;
    } else {
      x = P.X[q];
      aty = P.AtY[q];
    }
    double xn = x - tau * (P.c[j] + aty);
    xn = fmin(fmax(xn, P.L[q]), P.U[q]);
    const double d = xn - x;
    P.Xn[q] = xn;
    P.dX[q] = d;
    v[0] = d * d;
  }
  bnr_group_sums(P, 1, v, lds);
}

// W = Y + sigma (AX + 2 AdX), Yn = W - sigma clip(W / sigma, rl, ru), dY = Yn - Y;
// sums dY^2 and dY . AdX per node
__global__ void __launch_bounds__(kBnrBlock) bnr_dual_kernel(BnrPtrs P) {
  __shared__ double lds[kBnrBlock];
  long long i;
  int k;
  double v[2] = {0.0, 0.0};
  if (bnr_lane(P, P.m, &i, &k)) {
#pragma clang fp contract(off)
    const long long q = i * P.K + k;
    const double sigma = P.eta[k] * P.omega[k];
    const double y = P.Y[q], adx = P.AdX[q];
    const double w = y + sigma * (P.AX[q] + 2.0 * adx);
    const double yn = w - sigma * fmin(fmax(w / sigma, P.rl[i]), P.ru[i]);
    const double d = yn - y;
    P.Yn[q] = yn;
    P.dY[q] = d;
    v[0] = d * d;
    v[1] = d * adx;
  }
  bnr_group_sums(P, 2, v, lds);
}

// Sum over groups g of part[(g * 3 + slot) * K + k] by one workgroup per
// node: lane t adds groups t, t + 256, ... in order, then lane 0 adds the 256
// lane sums in order (the oracle sums in this same order).
__device__ __forceinline__ double bnr_groups_sum(const double* part, int groups, int slot, int K,
                                                 int k, double* lds) {
#pragma clang fp contract(off)
  double a = 0.0;
  for (int g = threadIdx.x; g < groups; g += kBnrLanes)
    a += part[((long long)g * 3 + slot) * K + k];
  __syncthreads();
  lds[threadIdx.x] = a;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int t = 0; t < kBnrLanes; ++t) s += lds[t];
  return s;                            // lane 0's value is the sum
}

// One workgroup per node: the step-size test and update (bnr.py:265-288).
// gp: groups of the primal pass, gd: of the dual pass; gfac = 1 - t^-growth and
// dfac = 1 + t^-damp (t = attempts + 1), computed on the host with its libm.
__global__ void __launch_bounds__(kBnrLanes)
    bnr_decide_kernel(BnrPtrs P, const double* part_p, int gp, const double* part_d, int gd,
                      double gfac, double dfac) {
#pragma clang fp contract(off)
  // the three sums in the order of bnr_groups_sum, with their loads issued
  // together and their 256-term lane chains run side by side on lanes 0-2
  __shared__ double lds[3 * kBnrLanes];
  __shared__ double tot[3];
  const int k = blockIdx.x;
  const int t = threadIdx.x;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int g = t; g < gp; g += kBnrLanes) a0 += part_p[(long long)g * 3 * P.K + k];
  for (int g = t; g < gd; g += kBnrLanes) {
    a1 += part_d[(long long)g * 3 * P.K + k];
    a2 += part_d[((long long)g * 3 + 1) * P.K + k];
  }
  lds[t] = a0;
  lds[kBnrLanes + t] = a1;
  lds[2 * kBnrLanes + t] = a2;
  __syncthreads();
  if (t < 3) {
    double sum = 0.0;
    for (int i = 0; i < kBnrLanes; ++i) sum += lds[t * kBnrLanes + i];
    tot[t] = sum;
  }
  __syncthreads();
  if (t != 0) return;
  const double dxx = tot[0], dyy = tot[1], inter = tot[2];
  const double omega = P.omega[k], eta = P.eta[k];
  const double interaction = fabs(inter);
  const double movement = 0.5 * (omega * dxx + dyy / omega);
  const double floor_ = fmax(movement, 1.0) / kQuotientMax;
  const double limit = interaction > 0.0 ? movement / fmax(interaction, floor_)
                                         : __builtin_inf();
  P.accept[k] = (eta <= limit) || (movement == 0.0);
  double e = fmin(gfac * limit, dfac * eta);
  e = isfinite(e) ? e : 1.0;
  P.eta[k] = fmax(e, 1e-14);
}

// accepted nodes take the trial point: X, AtY over columns; Y, AX over rows
__global__ void __launch_bounds__(kBnrBlock) bnr_update_kernel(BnrPtrs P, int groups_n) {
#pragma clang fp contract(off)
  const bool cols = (int)blockIdx.x < groups_n;
  const int g = cols ? (int)blockIdx.x : (int)blockIdx.x - groups_n;
  const int t = threadIdx.x;
  const int k = t & (P.KP - 1);
  const long long r = (long long)g * P.RB + t / P.KP;
  if (k >= P.K || !P.accept[k]) return;
  if (cols) {
    if (r >= P.n) return;
    const long long q = r * P.K + k;
    P.X[q] = P.Xn[q];
    P.AtY[q] = P.AtY[q] + P.AtdY[q];
  } else {
    if (r >= P.m) return;
    const long long q = r * P.K + k;
    P.Y[q] = P.Yn[q];
    P.AX[q] = P.AX[q] + P.AdX[q];
  }
}

// safe_dual_bound_batch terms (safebound.py:214-232): rows from y = ysign Y,
// columns from d = c - A^T y (A^T Y in AtdY); sums and vacuous flags per node.  Row groups first,
// then column groups.
__global__ void __launch_bounds__(kBnrBlock) bnr_bound_terms_kernel(BnrPtrs P, int groups_m) {
#pragma clang fp contract(off)
  __shared__ double lds[kBnrBlock];
  const bool rows = (int)blockIdx.x < groups_m;
  const int g = rows ? (int)blockIdx.x : (int)blockIdx.x - groups_m;
  const int t = threadIdx.x;
  const int k = t & (P.KP - 1);
  const long long r = (long long)g * P.RB + t / P.KP;
  double v[2] = {0.0, 0.0};            // contribution, vacuous
  if (rows) {
    if (r < P.m && k < P.K) {
      const long long q = r * P.K + k;
      const double y = P.ysign * P.Y[q], rl = P.rl[r], ru = P.ru[r];
      const bool pos = y > 0.0;
      const bool bad = (pos ? rl <= -kInf : ru >= kInf) && y != 0.0;
      if (bad) v[1] = 1.0;
      else v[0] = y * (pos ? rl : ru);
    }
  } else {
    if (r < P.n && k < P.K) {
      const long long q = r * P.K + k;
      const double d = P.c[r] - P.ysign * P.AtdY[q];
      const double lo = P.L[q], hi = P.U[q];
      const bool pos = d > 0.0;
      const bool bad = (pos ? lo <= -kInf : hi >= kInf) && d != 0.0;
      if (bad) v[1] = 1.0;
      else v[0] = d * (pos ? lo : hi);
    }
  }
  // one partial slot per group: [group][0..1][K]; the column groups follow
  // the row groups in `part`
  const int t2 = threadIdx.x;
  for (int s = 0; s < 2; ++s) {
    __syncthreads();
    lds[t2] = v[s];
    __syncthreads();
    if (t2 < P.KP) {
      double a = 0.0;
      for (int rr = 0; rr < P.RB; ++rr) a = s == 0 ? a + lds[rr * P.KP + t2]
                                                   : fmax(a, lds[rr * P.KP + t2]);
      if (t2 < P.K) P.part[((long long)blockIdx.x * 3 + s) * P.K + t2] = a;
    }
  }
}

// per node: row sum plus column sum (each over its groups, in the order of
// bnr_groups_sum), -inf when vacuous (safebound.py:231-232)
__global__ void __launch_bounds__(kBnrLanes)
    bnr_bound_final_kernel(BnrPtrs P, int groups_m, int groups_n) {
#pragma clang fp contract(off)
  __shared__ double lds[kBnrLanes];
  const int k = blockIdx.x;
  const double rs = bnr_groups_sum(P.part, groups_m, 0, P.K, k, lds);
  const double cs = bnr_groups_sum(P.part + (long long)groups_m * 3 * P.K, groups_n, 0, P.K, k,
                                   lds);
  double bad = 0.0;
  for (int g = threadIdx.x; g < groups_m + groups_n; g += kBnrLanes)
    bad = fmax(bad, P.part[((long long)g * 3 + 1) * P.K + k]);
  __syncthreads();
  lds[threadIdx.x] = bad;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int t = 0; t < kBnrLanes; ++t) bad = fmax(bad, lds[t]);
    P.bounds[k] = bad > 0.0 ? -__builtin_inf() : rs + cs;
  }
}

__global__ void bnr_zero_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace sov
