// sov_device.hpp -- device primitives for the gfx950 PDLP engine.
//
//  * deterministic wave64 / workgroup reductions (fixed butterfly order, so a
//    result never depends on timing);
//  * the CSR-Adaptive tile pass: a sparse matrix is cut on the host into
//    tiles of consecutive rows holding at most TNNZ nonzeros and at most
//    BLOCK rows, or a single row longer than TNNZ.  A workgroup streams a
//    tile's products (value * gathered vector entry) into LDS with fully
//    coalesced loads, then reduces each row with 1..64 lanes chosen from the
//    tile's row count, and hands every row sum to a fused epilogue.  Long
//    rows are reduced by the whole workgroup.  The summation order is a
//    function of the tile plan only, so products are bitwise repeatable.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sov {

// A pointer into LDS (address space 3): loads and stores through it are LDS
// instructions, and it takes one 32-bit register.
#define SOV_LDS __attribute__((address_space(3)))

constexpr double kInf = 1.0e30;   // infinite-bound sentinel (tolerances.py:16)
constexpr int kWave = 64;
constexpr int kMaxRed = 24;       // values per team reduction

enum : int {               // Status codes of src/sovopt/core/problem.py:58-69
  ST_OPTIMAL = 0,
  ST_ITERATION_LIMIT = 4,
  ST_TIME_LIMIT = 5,
  ST_NUMERICAL = 8,
  ST_RUNNING = -1,
};

// A value every lane of the wave holds equally, moved to a scalar register so
// that branches and loop bounds depending on it are scalar branches: barriers
// inside them then never sit in an EXEC-masked region.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ double nanmax(double a, double b) {
  return (a != a) ? a : ((b != b) ? b : (a > b ? a : b));
}

// The 32 bits of lane (l ^ O), O a power of two below 64, by the cheapest
// exchange for that distance: DPP within quads (1: quad_perm [1,0,3,2], 2:
// [2,3,0,1]) and rows of 16 (8: row_ror:8), ds_swizzle's bitmask mode within
// halves of 32 (4, 16: xor mask in bits 14:10, and mask 0x1f), ds_bpermute
// across the halves (32).  Every lane of the wave must be active.
template <int O>
__device__ __forceinline__ int xor_lane32(int x) {
  if constexpr (O == 1) return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);
  else if constexpr (O == 2) return __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);
  else if constexpr (O == 8) return __builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false);
  else if constexpr (O == 4) return __builtin_amdgcn_ds_swizzle(x, (4 << 10) | 0x1F);
  else if constexpr (O == 16) return __builtin_amdgcn_ds_swizzle(x, (16 << 10) | 0x1F);
  else return __shfl_xor(x, O);
}
template <int O>
__device__ __forceinline__ double xor_lane(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)xor_lane32<O>((int)(unsigned)b);
  const unsigned hi = (unsigned)xor_lane32<O>((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Butterfly over the wave, partners at distance 32, 16, ..., 1 in that order
// (each lane adds its partner's value to its own): every lane ends with the
// same total, formed in the same order whichever exchange moves the bits.
__device__ __forceinline__ double wave_sum(double v) {
  v += xor_lane<32>(v);
  v += xor_lane<16>(v);
  v += xor_lane<8>(v);
  v += xor_lane<4>(v);
  v += xor_lane<2>(v);
  v += xor_lane<1>(v);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
  v = nanmax(v, xor_lane<32>(v));
  v = nanmax(v, xor_lane<16>(v));
  v = nanmax(v, xor_lane<8>(v));
  v = nanmax(v, xor_lane<4>(v));
  v = nanmax(v, xor_lane<2>(v));
  v = nanmax(v, xor_lane<1>(v));
  return v;
}

template <class T>
__device__ __forceinline__ T tmax(T a, T b) {
  return a > b ? a : b;
}

// ------------------------------------------------------------------------
// workgroup reduction of K doubles; bit k of maxmask selects max, else sum.
// `red` is LDS with room for (BLOCK/64 + 2) * kMaxRed doubles.  Every thread
// receives the totals.  Leaves the workgroup's totals in red[NW*kMaxRed + k].
// ------------------------------------------------------------------------
template <int BLOCK, int K>
__device__ __forceinline__ void block_reduce_to_lds(const double (&v)[K], uint32_t maxmask,
                                                    double* red) {
  constexpr int NW = BLOCK / kWave;
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = ((maxmask >> k) & 1u) ? wave_max(v[k]) : wave_sum(v[k]);
    if (l == 0) red[w * kMaxRed + k] = s;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    const bool mx = (maxmask >> k) & 1u;
    double s = red[k];
    for (int q = 1; q < NW; ++q) {
      double o = red[q * kMaxRed + k];
      s = mx ? nanmax(s, o) : s + o;
    }
    red[NW * kMaxRed + k] = s;
  }
  __syncthreads();
}

template <int BLOCK, int K>
__device__ __forceinline__ void block_reduce(double (&v)[K], uint32_t maxmask, double* red) {
  constexpr int NW = BLOCK / kWave;
  block_reduce_to_lds<BLOCK, K>(v, maxmask, red);
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = red[NW * kMaxRed + k];
  __syncthreads();
}

// workgroup sum of NV values of type T (tile pass, long rows)
template <int BLOCK, int NV, bool MAXOP, class T>
__device__ __forceinline__ void block_combine(T (&s)[NV], T* scratch) {
  constexpr int NW = BLOCK / kWave;
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    T v = s[k];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
      T u = __shfl_xor(v, o);
      v = MAXOP ? tmax(v, u) : v + u;
    }
    if (l == 0) scratch[k * NW + w] = v;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    T v = scratch[k * NW];
    for (int q = 1; q < NW; ++q) v = MAXOP ? tmax(v, scratch[k * NW + q]) : v + scratch[k * NW + q];
    s[k] = v;
  }
  __syncthreads();
}

// Rows per lane of the short-row tiles: a tile holds at most kRowsPerLane *
// BLOCK rows (the host plans tiles with that row cap).
constexpr int kRowsPerLane = 4;

// Fold the staged products lds[u * TNNZ + q], q = a, a + step, ... < b, into
// s[u] (sum, or max if MAXOP) in that order.  Four LDS reads are issued before
// the first of them is used, so a row costs about len / 4 LDS round trips
// instead of len; the operations and their order are those of a loop over
// one product at a time, so the result is the same bit for bit.  The last
// 1-3 products read in-range slots (clamped to b - 1) and use only their own.
template <int TNNZ, int NV, bool MAXOP, class T>
__device__ __forceinline__ void staged_row_fold(const T* lds, int a, int b, int step,
                                                T (&s)[NV]) {
  int q = a;
  for (; q + 3 * step < b; q += 4 * step) {
    T v[4][NV];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int u = 0; u < NV; ++u) v[r][u] = lds[u * TNNZ + q + r * step];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int u = 0; u < NV; ++u) s[u] = MAXOP ? tmax(s[u], v[r][u]) : s[u] + v[r][u];
    }
  }
  if (q < b) {
    T v[3][NV];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int qr = q + r * step < b ? q + r * step : b - 1;
#pragma unroll
      for (int u = 0; u < NV; ++u) v[r][u] = lds[u * TNNZ + qr];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      if (q + r * step < b) {
#pragma unroll
        for (int u = 0; u < NV; ++u) s[u] = MAXOP ? tmax(s[u], v[r][u]) : s[u] + v[r][u];
      }
    }
  }
}

// Nonzeros per lane whose loads are issued together: the column indices of a
// chunk, then its products (values and gathered entries), then the LDS
// stores; the next chunk's indices are loaded with this chunk's products.  A
// plain loop over the nonzeros would wait for each index before its gather,
// two dependent memory round trips per nonzero.  (8 per lane spills registers
// under the run kernel's 128-VGPR cap and is slower.)
constexpr int kChunk = 4;

// How a tile pass reads nonzero q: fetch(q) issues the loads that depend on no
// other load, product(q, e, v) forms the NV products from what fetch loaded.
//   IndexFetch  the column index; prod(q, j, v) loads the value and gathers
//   LdsGather   the column index and the value together; the entry of the
//               gathered vector comes from LDS (one memory latency per chunk
//               of nonzeros, the gather an LDS read)
template <class Prod>
struct IndexFetch {
  const int32_t* __restrict__ idx;
  Prod& prod;
  using E = int;
  static constexpr int kCh = kChunk;
  __device__ __forceinline__ E fetch(int q) const { return idx[q]; }
  template <class V>
  __device__ __forceinline__ void product(int q, E j, V& v) const { prod(q, j, v); }
};

template <class T>
struct LdsGather {
  const int32_t* __restrict__ idx;
  const T* __restrict__ val;
  const SOV_LDS T* __restrict__ g;
  struct E {
    int j;
    T a;
  };
  static constexpr int kCh = 8;
  __device__ __forceinline__ E fetch(int q) const { return E{idx[q], val[q]}; }
  __device__ __forceinline__ void product(int, const E& e, T (&v)[1]) const {
    v[0] = e.a * g[e.j];
  }
};

// Stage the products of nonzeros [p0, p0 + nz) of a tile (nz <= TNNZ) into
// LDS slot q = nonzero - p0.  Lanes past the end load the tile's last nonzero
// (always valid) and discard it.
template <int BLOCK, int TNNZ, int NV, class T, class Ld>
__device__ __forceinline__ void stage_products(const Ld& ld, int p0, int nz, T* lds) {
  constexpr int PER = (TNNZ + BLOCK - 1) / BLOCK;
  constexpr int CH = PER < Ld::kCh ? PER : Ld::kCh;
  using E = typename Ld::E;
  if (nz <= 0) return;
  E jj[CH];
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    const int q = (int)threadIdx.x + k * BLOCK;
    jj[k] = ld.fetch(p0 + (q < nz ? q : nz - 1));
  }
#pragma unroll
  for (int c = 0; c < PER; c += CH) {
    if (c * BLOCK >= nz) break;          // nz is uniform: a scalar branch
    // the next chunk's column indices load while this chunk's products do:
    // one memory latency per chunk instead of two
    E jn[CH];
    const bool next = c + CH < PER && (c + CH) * BLOCK < nz;
    if (next) {
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int q = (int)threadIdx.x + (c + CH + k) * BLOCK;
        jn[k] = ld.fetch(p0 + (q < nz ? q : nz - 1));
      }
    }
    T v[CH][NV];
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int q = (int)threadIdx.x + (c + k) * BLOCK;
      ld.product(p0 + (q < nz ? q : nz - 1), jj[k], v[k]);
    }
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int q = (int)threadIdx.x + (c + k) * BLOCK;
      if (q < nz) {
#pragma unroll
        for (int u = 0; u < NV; ++u) lds[u * TNNZ + q] = v[k][u];
      }
    }
    if (next) {
#pragma unroll
      for (int k = 0; k < CH; ++k) jj[k] = jn[k];
    }
  }
}

// Where a tile pass finds its tiles: tile t covers rows [row(t), row(t+1)) and
// nonzeros [nz(t), nz(t+1)), nz(t) = ptr[row(t)].  GlobalBounds reads the
// plan and the row pointers in memory (two dependent loads at the start of a
// pass and one per tile after it); LdsBounds reads both from a copy in LDS.
struct GlobalBounds {
  const int32_t* __restrict__ ptr;
  const int32_t* __restrict__ tiles;
  __device__ __forceinline__ int row(int t) const { return uniform(tiles[t]); }
  __device__ __forceinline__ int nz(int, int r) const { return uniform(ptr[r]); }
};
struct LdsBounds {
  const SOV_LDS int32_t* rows;
  const SOV_LDS int32_t* nzs;
  __device__ __forceinline__ int row(int t) const { return uniform(rows[t]); }
  __device__ __forceinline__ int nz(int t, int) const { return uniform(nzs[t]); }
};

// ------------------------------------------------------------------------
// CSR-Adaptive tile pass.
//   ptr   : row pointers (LP-local int32 offsets), rows+1 entries
//   idx   : column index of each nonzero
//   tiles : tile t covers rows [tiles[t], tiles[t+1])
//   prod(q, j, v[NV]) writes the NV products of nonzero q (column j)
//   pre(row) -> R     loads what the row's epilogue needs; it is called
//                     before the products are staged, so those loads are in
//                     flight together with the product loads
//   epi(row, s[NV], r) receives the NV row reductions (sum, or max if MAXOP)
//                     and pre's result
// The bounds of tile t+1 are fetched while tile t is processed.  A workgroup
// runs the tiles one after another, so the number of dependent memory
// latencies per tile, not bandwidth, sets the time of a small pass.
// ------------------------------------------------------------------------
template <int BLOCK, int TNNZ, int NV, bool MAXOP, class T, class Bd, class Ld, class Pre,
          class Epi>
__device__ __forceinline__ void tile_pass_ld(const int32_t* __restrict__ ptr, const Bd& bd,
                                             int t0, int t1, const Ld& ld, Pre& pre, Epi& epi,
                                             T* lds) {
  using E = typename Ld::E;
  t0 = uniform(t0);
  t1 = uniform(t1);
  if (t0 >= t1) return;
  int r0 = bd.row(t0), r1 = bd.row(t0 + 1);
  int p0 = bd.nz(t0, r0), p1 = bd.nz(t0 + 1, r1);
  for (int t = t0; t < t1; ++t) {
    const bool more = t + 1 < t1;
    const int nr1 = more ? bd.row(t + 2) : r1;   // next tile's end row
    const int nz = p1 - p0, rows = r1 - r0;
    int np1 = p1;
    if (nz > TNNZ) {
      // one long row: the whole workgroup strides over it
      decltype(pre(0)) pr{};
      if (threadIdx.x == 0) pr = pre(r0);
      T s[NV];
#pragma unroll
      for (int k = 0; k < NV; ++k) s[k] = T(0);
      // lane order q = p0 + tid, + BLOCK, ...; kChunk of them in flight at
      // once, and the next chunk's column indices load with this chunk's
      // products (one memory latency per chunk instead of two)
      E jj[kChunk];
#pragma unroll
      for (int c = 0; c < kChunk; ++c) {
        const int q = p0 + (int)threadIdx.x + c * BLOCK;
        jj[c] = ld.fetch(q < p1 ? q : p1 - 1);
      }
      for (int base = p0; base < p1; base += kChunk * BLOCK) {
        const int nb = base + kChunk * BLOCK;
        const bool next = nb < p1;           // uniform: a scalar branch
        E jn[kChunk];
        if (next) {
#pragma unroll
          for (int c = 0; c < kChunk; ++c) {
            const int q = nb + (int)threadIdx.x + c * BLOCK;
            jn[c] = ld.fetch(q < p1 ? q : p1 - 1);
          }
        }
        T v[kChunk][NV];
#pragma unroll
        for (int c = 0; c < kChunk; ++c) {
          const int q = base + (int)threadIdx.x + c * BLOCK;
          ld.product(q < p1 ? q : p1 - 1, jj[c], v[c]);
        }
#pragma unroll
        for (int c = 0; c < kChunk; ++c) {
          const int q = base + (int)threadIdx.x + c * BLOCK;
          if (q < p1) {
#pragma unroll
            for (int k = 0; k < NV; ++k) s[k] = MAXOP ? tmax(s[k], v[c][k]) : s[k] + v[c][k];
          }
        }
        if (next) {
#pragma unroll
          for (int c = 0; c < kChunk; ++c) jj[c] = jn[c];
        }
      }
      if (more) np1 = bd.nz(t + 2, nr1);
      block_combine<BLOCK, NV, MAXOP, T>(s, lds);
      if (threadIdx.x == 0) epi(r0, s, pr);
      __syncthreads();
    } else if (rows > BLOCK) {
      // many short rows: one lane per row, up to kRowsPerLane rows per lane,
      // every row's extent and epilogue operands loaded before the products
      int a[kRowsPerLane], b[kRowsPerLane];
      decltype(pre(0)) pr[kRowsPerLane];
#pragma unroll
      for (int k = 0; k < kRowsPerLane; ++k) {
        // rows past the tile load its last row (valid) and are not used
        const int row = (int)threadIdx.x + k * BLOCK;
        const int rc = r0 + (row < rows ? row : rows - 1);
        a[k] = ptr[rc] - p0;
        b[k] = ptr[rc + 1] - p0;
        pr[k] = pre(rc);
      }
      stage_products<BLOCK, TNNZ, NV, T>(ld, p0, nz, lds);
      __syncthreads();
      if (more) np1 = bd.nz(t + 2, nr1);
#pragma unroll
      for (int k = 0; k < kRowsPerLane; ++k) {
        const int row = (int)threadIdx.x + k * BLOCK;
        if (row < rows) {
          T s[NV];
#pragma unroll
          for (int u = 0; u < NV; ++u) s[u] = T(0);
          staged_row_fold<TNNZ, NV, MAXOP, T>(lds, a[k], b[k], 1, s);
          epi(r0 + row, s, pr[k]);
        }
      }
      // rows past kRowsPerLane * BLOCK (a plan with a larger row cap): no prefetch
      for (int row = (int)threadIdx.x + kRowsPerLane * BLOCK; row < rows; row += BLOCK) {
        const int ra = ptr[r0 + row] - p0, rb = ptr[r0 + row + 1] - p0;
        const auto pq = pre(r0 + row);
        T s[NV];
#pragma unroll
        for (int u = 0; u < NV; ++u) s[u] = T(0);
        staged_row_fold<TNNZ, NV, MAXOP, T>(lds, ra, rb, 1, s);
        epi(r0 + row, s, pq);
      }
      __syncthreads();
    } else {
      int tpr = BLOCK / (rows > 0 ? rows : 1);
      tpr = tpr >= 64 ? 64 : (tpr >= 32 ? 32 : (tpr >= 16 ? 16 : (tpr >= 8 ? 8 : (tpr >= 4 ? 4 : (tpr >= 2 ? 2 : 1)))));
      const int grp = (int)threadIdx.x / tpr, ln = (int)threadIdx.x & (tpr - 1);
      const bool active = grp < rows;
      // row extent and epilogue operands, issued before the product loads and
      // without a branch (idle groups load the tile's last row and ignore it)
      const int rc = r0 + (active ? grp : rows - 1);
      const int a = ptr[rc] - p0, b = ptr[rc + 1] - p0;
      const auto pr = pre(rc);
      stage_products<BLOCK, TNNZ, NV, T>(ld, p0, nz, lds);
      __syncthreads();
      if (more) np1 = bd.nz(t + 2, nr1);
      T s[NV];
#pragma unroll
      for (int k = 0; k < NV; ++k) s[k] = T(0);
      if (active) staged_row_fold<TNNZ, NV, MAXOP, T>(lds, a + ln, b, tpr, s);
      for (int o = tpr >> 1; o > 0; o >>= 1) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
          T u = __shfl_xor(s[k], o);
          s[k] = MAXOP ? tmax(s[k], u) : s[k] + u;
        }
      }
      if (active && ln == 0) epi(r0 + grp, s, pr);
      __syncthreads();
    }
    r0 = r1;
    p0 = p1;
    r1 = nr1;
    p1 = np1;
  }
}

// ------------------------------------------------------------------------
// The tile pass over a matrix held in LDS (the run kernel keeps an LP's
// scaled matrix there when it fits): val / idx / ptr / tiles are LDS copies
// (16-bit indices and pointers), g the gathered vector (LDS too).  Each row
// sum is formed in exactly the order, and from exactly the rounded products,
// of tile_pass_pf's short-row branches, so the results are bit for bit those
// of the staged pass: a product is rounded to T before it is added
// (contraction off), as it is when staged through LDS.  There is nothing to
// stage, so a tile needs no workgroup barrier; the only global loads are
// pre()'s, issued before the products.  Every tile must be a short-row tile
// (at most TNNZ nonzeros), which holds for any matrix of at most TNNZ
// nonzeros.
// ------------------------------------------------------------------------
// Sum of the products val[q] * g[idx[q]] for q = a, a + step, ... < b, in
// that order, each product rounded before it is added; four products at a
// time, so that their LDS reads overlap.
template <class T>
__device__ __forceinline__ T lds_row_sum(const SOV_LDS T* __restrict__ val,
                                         const SOV_LDS uint16_t* __restrict__ idx,
                                         const SOV_LDS T* __restrict__ g, int a, int b,
                                         int step, T s) {
#pragma clang fp contract(off)
  int q = a;
  for (; q + 3 * step < b; q += 4 * step) {
    const int j0 = idx[q], j1 = idx[q + step], j2 = idx[q + 2 * step], j3 = idx[q + 3 * step];
    const T v0 = val[q], v1 = val[q + step], v2 = val[q + 2 * step], v3 = val[q + 3 * step];
    const T p0 = v0 * g[j0], p1 = v1 * g[j1], p2 = v2 * g[j2], p3 = v3 * g[j3];
    s = s + p0;
    s = s + p1;
    s = s + p2;
    s = s + p3;
  }
  if (q < b) {
    // the last 1-3 products: their reads issued together from in-range
    // slots (clamped to q), each product added only if it is the row's
    int jr[3];
    T vr[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int qr = q + r * step < b ? q + r * step : q;
      jr[r] = idx[qr];
      vr[r] = val[qr];
    }
    T pr[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) pr[r] = vr[r] * g[jr[r]];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      if (q + r * step < b) s = s + pr[r];
    }
  }
  return s;
}

template <int BLOCK, class T, class Pre, class Epi>
__device__ __forceinline__ void tile_pass_res(const SOV_LDS T* __restrict__ val,
                                              const SOV_LDS uint16_t* __restrict__ idx,
                                              const SOV_LDS uint16_t* __restrict__ ptr,
                                              const SOV_LDS uint16_t* __restrict__ tiles,
                                              const SOV_LDS T* __restrict__ g, int t0, int t1,
                                              Pre& pre, Epi& epi) {
#pragma clang fp contract(off)
  for (int t = uniform(t0); t < uniform(t1); ++t) {
    const int r0 = uniform(tiles[t]), r1 = uniform(tiles[t + 1]);
    const int rows = r1 - r0;
    if (rows > BLOCK) {
      // one lane per row, up to kRowsPerLane rows per lane (tile_pass_pf's
      // many-short-rows branch): the rows' operands first, then the sums
      int a[kRowsPerLane], b[kRowsPerLane];
      decltype(pre(0)) pr[kRowsPerLane];
#pragma unroll
      for (int k = 0; k < kRowsPerLane; ++k) {
        const int row = (int)threadIdx.x + k * BLOCK;
        const int rc = r0 + (row < rows ? row : rows - 1);
        a[k] = ptr[rc];
        b[k] = ptr[rc + 1];
        pr[k] = pre(rc);
      }
#pragma unroll
      for (int k = 0; k < kRowsPerLane; ++k) {
        const int row = (int)threadIdx.x + k * BLOCK;
        if (row < rows) {
          T s[1] = {lds_row_sum<T>(val, idx, g, a[k], b[k], 1, T(0))};
          epi(r0 + row, s, pr[k]);
        }
      }
      for (int row = (int)threadIdx.x + kRowsPerLane * BLOCK; row < rows; row += BLOCK) {
        const auto pq = pre(r0 + row);
        T s[1] = {lds_row_sum<T>(val, idx, g, ptr[r0 + row], ptr[r0 + row + 1], 1, T(0))};
        epi(r0 + row, s, pq);
      }
    } else {
      int tpr = BLOCK / (rows > 0 ? rows : 1);
      tpr = tpr >= 64 ? 64 : (tpr >= 32 ? 32 : (tpr >= 16 ? 16 : (tpr >= 8 ? 8 : (tpr >= 4 ? 4 : (tpr >= 2 ? 2 : 1)))));
      const int grp = (int)threadIdx.x / tpr, ln = (int)threadIdx.x & (tpr - 1);
      const bool active = grp < rows;
      const int rc = r0 + (active ? grp : rows - 1);
      const int a = ptr[rc], b = ptr[rc + 1];
      const auto pr = pre(rc);
      T s[1] = {T(0)};
      if (active) s[0] = lds_row_sum<T>(val, idx, g, a + ln, b, tpr, T(0));
      for (int o = tpr >> 1; o > 0; o >>= 1) s[0] = s[0] + __shfl_xor(s[0], o);
      if (active && ln == 0) epi(r0 + grp, s, pr);
    }
  }
}

// The tile pass with the generic loader: column index, then prod(q, j, v).
template <int BLOCK, int TNNZ, int NV, bool MAXOP, class T, class Prod, class Pre, class Epi>
__device__ __forceinline__ void tile_pass_pf(const int32_t* __restrict__ ptr,
                                             const int32_t* __restrict__ idx,
                                             const int32_t* __restrict__ tiles, int t0, int t1,
                                             Prod& prod, Pre& pre, Epi& epi, T* lds) {
  const IndexFetch<Prod> ld{idx, prod};
  const GlobalBounds bd{ptr, tiles};
  tile_pass_ld<BLOCK, TNNZ, NV, MAXOP, T>(ptr, bd, t0, t1, ld, pre, epi, lds);
}

// The tile pass without epilogue prefetch: epi(row, s[NV]).
template <int BLOCK, int TNNZ, int NV, bool MAXOP, class T, class Prod, class Epi>
__device__ __forceinline__ void tile_pass(const int32_t* __restrict__ ptr,
                                          const int32_t* __restrict__ idx,
                                          const int32_t* __restrict__ tiles, int t0, int t1,
                                          Prod& prod, Epi& epi, T* lds) {
  auto pre = [](int) { return 0; };
  auto epi3 = [&](int i, const T (&s)[NV], int) { epi(i, s); };
  tile_pass_pf<BLOCK, TNNZ, NV, MAXOP, T>(ptr, idx, tiles, t0, t1, prod, pre, epi3, lds);
}

// Visit every nonzero of a tile range with its row: fn(q, row, col).
template <int BLOCK, int TNNZ, class Fn>
__device__ __forceinline__ void tile_foreach(const int32_t* __restrict__ ptr,
                                             const int32_t* __restrict__ idx,
                                             const int32_t* __restrict__ tiles, int t0, int t1,
                                             Fn& fn) {
  for (int t = uniform(t0); t < uniform(t1); ++t) {
    const int r0 = uniform(tiles[t]), r1 = uniform(tiles[t + 1]);
    const int p0 = uniform(ptr[r0]), p1 = uniform(ptr[r1]);
    for (int q = p0 + (int)threadIdx.x; q < p1; q += BLOCK) {
      // binary search the row of q among [r0, r1)
      int lo = r0, hi = r1 - 1;
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (ptr[mid] <= q) lo = mid; else hi = mid - 1;
      }
      fn(q, lo, idx[q]);
    }
  }
}

}  // namespace sov
