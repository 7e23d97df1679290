// common.hpp -- host-side helpers shared by the translation units of libsovlp.so.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sovlp.h"
#include "pdlp_kernels.hpp"

namespace sovh {

using namespace sov;

// Records the message of a failure for sovlp_last_error and returns `code`.
int fail(int code, const std::string& msg);

#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::sovh::fail(2, std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

inline hipStream_t S(sovlp_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kBlock = 256;
constexpr int kTnnz64 = 2048;   // FP64 engines: 2 vectors x 2048 x 8 B = 32 KiB LDS
constexpr int kTnnz32 = 4096;   // FP32 engines: 2 x 4096 x 4 B (or 4096 doubles) = 32 KiB

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

int clock_hz(double* hz);
// ptr_at_tile[t] = pointer value at row tiles[t]; cost = nonzeros + rows
void host_partition(int ntiles, const int32_t* tiles, const int64_t* ptr_at_tile, int parts,
                    int32_t* part);

}  // namespace sovh
