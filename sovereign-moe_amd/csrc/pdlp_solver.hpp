// pdlp_solver.hpp -- restarted-average PDHG (PDLP) on the device, as a state machine.
//
// The reference loop (src/sovopt/lp/pdlp.py:238-494; line numbers below are
// to that file unless stated) is cut into PHASES.  A phase is the work between
// two reductions: one tile pass over A or A^T with a fused epilogue, or one
// elementwise sweep.  Every kernel launch executes exactly one phase of each LP
// it serves, so no launch contains a loop over iterations, and the iteration
// state lives in HBM (LpState) between launches:
//
//   batched LPs   one workgroup per LP per launch; the workgroup reduces its
//                 partial sums in LDS and applies the transition itself.
//   one large LP  W workgroups share a phase and write partial sums; a
//                 one-workgroup "decide" launch reduces them in a fixed order
//                 and applies the transition.  Kernel boundaries take the
//                 place of grid barriers.
//
// Phases and the reference lines they restate:
//   INIT, RUIZ_MAX, RUIZ_APPLY   Ruiz equilibration (numerics/scaling.py:149-170)
//   PC_SUM, PC_APPLY             Pock-Chambolle + power-of-two rounding
//                                (scaling.py:274-309, 119-131, 349-352) and the
//                                objective scale (scaling.py:422-444)
//   SCALE                        scaled copy, bounds (scaling.py:446-456), x0 =
//                                clip(0, lo, hi), y0 = 0 (274-275), norms for
//                                omega0 (292-296) and for _converged (523-528)
//   AX0, POW_A, POW_T            ax = A x0 (279) and 20 power iterations for
//                                ||A||_2 (140-159, 289-290)
//   KKT0                         initial merit (304)
//   P_TRIAL, A_TRIAL             adaptive step trial (337-381): primal step
//                                (backend.py:204-215), A dx, dual step on
//                                ax + 2 adx (backend.py:219-231), three dots
//   B_ADAPT                      accepted trial: A^T dy, x,y <- x',y', ax += adx,
//                                aty += ady, sums with the updated eta (383-389)
//   A_FUSED, B_FUSED             fused iteration (390-399; backend.py:241-291);
//                                B phases also form the NEXT iteration's primal
//                                projection in their epilogue
//   CHECK                        scaled KKT of current and average points
//                                (191-225), unscaled KKT of both (497-520),
//                                termination (523-530), restart test (405-441)
//   RESTART                      restart to the candidate and primal weight
//                                update (443-463)
//   COPY_AVG                     converged on the average point (424-425)
//   FINAL, FINAL_D               unscale and report (465-494)
#pragma once
#include "sov_device.hpp"

namespace sov {

struct Params {              // same layout as sovlp_params in include/sovlp.h
  double eps_abs, eps_rel, feas_cap;
  long long max_iter;
  double time_limit;
  int check_every, stepsize_every;
  double beta_sufficient, beta_necessary, beta_artificial, primal_weight_smoothing;
  int refine_rounds;         // 0: no refinement (FP64 engines always 0)
  int refine_stall;          // a round whose KKT ratio has not halved for this many
                             // iterations stops (0: 1024); it stops at 8x this in any case
  double refine_eps;         // relative KKT tolerance of each refinement round
  double refine_tol;         // primal residual target: refine_tol (1 + ||b||_inf)
  int refine_stall_open;     // refine_stall while no evaluated X has met _converged
                             // (0: 4096)
  int flags;                 // kFlagNoLdsMatrix: the run kernel keeps no matrix in LDS
};

constexpr int kFlagNoLdsMatrix = 1;
constexpr int kKnownFlags = kFlagNoLdsMatrix;

constexpr int kMaxRounds = 16;   // refinement rounds recorded per LP
// Trust region of a refinement round, in its zoomed units: every finite
// shifted bound is clipped to [-kTrustRegion, kTrustRegion], so a round can
// move X by at most kTrustRegion / Delta (~ 4096 x the last residual).  From
// an (X, Y) that is only accurate to ~1e-4, an unclipped round starts with a
// box ~1e5 wide in zoomed units and its first steps overshoot across it.
constexpr double kTrustRegion = 4096.0;

enum Phase : int {
  PH_INIT = 0, PH_RUIZ_MAX, PH_RUIZ_APPLY, PH_PC_SUM, PH_PC_APPLY, PH_SCALE, PH_AX0, PH_POW_A,
  PH_POW_T, PH_KKT0, PH_P_TRIAL, PH_A_TRIAL, PH_B_ADAPT, PH_A_FUSED, PH_B_FUSED, PH_CHECK,
  PH_RESTART, PH_COPY_AVG, PH_FINAL, PH_FINAL_D, PH_REF_X, PH_REF_KKT, PH_REF_SHIFT, PH_DONE
};

enum NextMode : int { NEXT_NONE = 0, NEXT_FUSED = 1, NEXT_TRIAL = 2 };

// Per-LP phase counters: slot ph counts executions of phase ph; the two extra
// slots count the B phases that also formed the next primal step (they move
// more bytes).  The host turns the counts into algorithmic bytes.
constexpr int kSlotBAdaptNext = PH_DONE + 1;
constexpr int kSlotBFusedNext = PH_DONE + 2;
constexpr int kPhaseSlots = 28;

struct LpState {             // everything the loop carries between phases
  double eta, omega, sum_eta, mu_restart, mu_last, mu_cand, anorm, sobj, ptol, dtol;
  double dev, nv, nrm, est, dxdx;
  double fin[5];             // unscaled (pres, dres, gap, pobj, dobj) at OPTIMAL
  double objective;
  long long k, epoch, attempts, restarts, rejected, hist_len;
  long long tick0;           // device clock when the solve started (phase INIT)
  long long ticks;           // device clock ticks used (set when the solve ends)
  long long ticks_left;      // time budget in ticks
  // iterative refinement (FP32 iterates): X accumulates in FP64 (xacc); each
  // round solves min c'x^ over bounds shifted by X and scaled by ref_scale
  double ref_scale;          // Delta of the current round (power of two)
  double rhs0;               // ||finite row bounds||_inf of the original problem
  double cinf0;              // ||c||_inf of the original problem
  double gap_last;           // duality gap of the last X, Y on the original problem
  double sobj0;              // objective scale of the original problem
  double ptol0, dtol0;       // _converged tolerances of the original problem
  long long k_done;          // iterations of finished rounds
  long long kmax;            // iteration limit of the current round
  long long k_round0;        // iterations of round 0
  long long k_best;          // iteration of the round's last KKT-ratio halving
  double kkt_best;           // best KKT ratio of the round (stall detection)
  int phase, status, sweep, pit, trial, primal_ready, time_flag, use_avg, next_mode;
  int round;                 // refinement round (0: the plain solve)
  int conv_last;             // the last X evaluated satisfied _converged
  int absorbed;              // xacc already holds the current round's x
  int ref_rows;              // rows written to rhist
  // best (X, Y) of the refinement so far, by the _converged ratio
  // max(pres / ptol, dres / dtol, gap / gtol) (<= 1: converged)
  double best_q;
  double best_fin[5];        // (pres, dres, gap, pobj, dobj) of the best
  int nobetter;              // consecutive evaluations that did not improve on it
  int save_best;             // the next REF_SHIFT copies X, Y to xbest, ybest
  int use_best;              // report xbest, ybest rather than X, Y
  int pending;               // W workgroups: `phase` has run and its transition
                             // (from the partial sums) is still to be taken
};

template <class T>
struct LpDev {
  int m, n, ntA, ntT, flip, pad;
  double offset0;            // objective offset of the minimisation form
  // original problem (minimisation form), FP64
  const int32_t *ap, *ai;
  const double* av;          // CSR of A
  const int32_t *tp, *ti;
  const double* tv;          // CSC of A == CSR of A^T
  const double *c0, *lo0, *hi0, *rl0, *ru0, *v0;
  const int32_t *tilesA, *tilesT, *partA, *partT;
  // scaled problem
  T *sav, *stv;
  double *rowf, *colf;
  T *c, *lo, *hi, *rl, *ru;
  // iterates (columns)
  T *x, *xn, *dx, *aty, *sumx, *xr, *atyf, *atya;
  // iterates (rows)
  T *y, *yn, *dy, *ax, *adx, *sumy, *yr, *axf, *axa;
  double *tmpm, *tmpn;
  // outputs
  double *xo, *yo, *dout;
  double* hist;
  long long hist_cap;
  // bounds of the problem the rounds solve (FP64, unscaled): the original
  // arrays when there is no refinement, the shifted ones otherwise
  const double *crl, *cru, *clo, *chi;
  double *wrl, *wru, *wlo, *whi;   // writable copies (refinement only)
  double *xacc, *axacc;      // X and A X (refinement only; null otherwise)
  // primal-dual refinement (FP32 engines): Y, the reduced cost c + A^T Y,
  // and the current round's objective and row-slack cost, unscaled FP64;
  // q is the slack cost in scaled units (zero in round 0)
  double *yacc, *dacc, *cc, *qc;
  T* q;
  double *xbest, *ybest;     // the refinement's best X, Y (FP32 engines)
  double* rhist;             // [kMaxRounds][5] per round: k, pres, dres, gap, pobj
  LpState* st;
  LpState* st2;              // W workgroups: the other state of the step kernels' ping-pong
  long long* pcount;         // [kPhaseSlots] phase counters
  long long* pticks;         // [kPhaseSlots] device clock ticks spent in each phase (run kernel)
  double* partials;          // [2][W][kMaxRed] per-workgroup partial sums (W-way phases;
                             // the step kernels alternate between the halves)
  // warm start (served resolves; null / 0 = the reference's cold start
  // x0 = clip(0), y0 = 0, omega0 from the norms): x and y in the units and
  // signs of the report (xo, yo), and a primal weight to start from
  const double *wx0, *wy0;
  double womega;
};

// ---------------------------------------------------------------------------
// An LP's scaled matrix in both orientations (values, 16-bit column indices,
// row pointers and tile plans) and its two step vectors dx, dy, held in the
// run kernel's LDS when they fit.  The phases of the PDHG loop (P_TRIAL,
// A_TRIAL, B_ADAPT, A_FUSED, B_FUSED) then read the matrix and gather the
// step vectors from LDS; each of them that writes dx or dy writes both
// copies.  Any other phase may use the same LDS for staging, so the copy is
// reloaded from global memory before the next loop phase after it.  The
// products and sums are those of the staged tile pass bit for bit.
// ---------------------------------------------------------------------------
template <class T>
struct ResidentLds {
  SOV_LDS T *valA, *valT, *dx, *dy;
  SOV_LDS uint16_t *idxA, *idxT, *ptrA, *ptrT, *tilesA, *tilesT;
  // only dx, dy in LDS: the tile bounds of both orientations (first row and
  // first nonzero of each tile, and the ends), so the streamed passes find
  // their tiles without dependent loads from memory
  SOV_LDS int32_t *rowsA, *nzsA, *rowsT, *nzsT;
};

// LDS bytes of an LP's resident copy
template <class T>
__host__ __device__ inline long long resident_bytes(int m, int n, long long nnz, int ntA,
                                                    int ntT) {
  return (2 * nnz + m + n) * (long long)sizeof(T) +
         2LL * (2 * nnz + (m + 1) + (n + 1) + (ntA + 1) + (ntT + 1));
}

// Whether an LP's copy fits `cap` bytes of LDS and 16-bit indices, with every
// tile a short-row tile (at most tnnz nonzeros).
template <class T>
__host__ __device__ inline bool resident_fits(int m, int n, long long nnz, int ntA, int ntT,
                                              long long cap, int tnnz) {
  return m < 65536 && n < 65536 && nnz < 65536 && nnz <= tnnz &&
         resident_bytes<T>(m, n, nnz, ntA, ntT) <= cap;
}

// `base`: the start of the LDS region (16-byte aligned)
template <class T>
__device__ inline ResidentLds<T> carve_resident(void* base, int m, int n, int nnz, int ntA) {
  ResidentLds<T> r;
  SOV_LDS T* t = (SOV_LDS T*)base;
  r.valA = t;
  r.valT = t + nnz;
  r.dx = t + 2 * nnz;
  r.dy = t + 2 * nnz + n;
  SOV_LDS uint16_t* u = (SOV_LDS uint16_t*)(t + 2 * nnz + n + m);
  r.idxA = u;
  r.idxT = u + nnz;
  r.ptrA = u + 2 * nnz;
  r.ptrT = r.ptrA + (m + 1);
  r.tilesA = r.ptrT + (n + 1);
  r.tilesT = r.tilesA + (ntA + 1);
  return r;
}

// An LP whose matrix does not fit may still keep its two step vectors in LDS,
// above the loop phases' staging area (TNNZ values of T): the products of the
// staged tile pass then gather dx or dy from LDS (the matrix streams from
// memory as before).  ResidentLds then holds only dx and dy.
template <class T>
__host__ __device__ inline long long vec_offset(int tnnz) {
  return ((long long)tnnz * (long long)sizeof(T) + 15) / 16 * 16;
}
template <class T>
__host__ __device__ inline bool vec_fits(int m, int n, int ntA, int ntT, long long cap,
                                         int tnnz) {
  return vec_offset<T>(tnnz) + (long long)(m + n) * (long long)sizeof(T) +
             8LL * (ntA + ntT + 2) <=
         cap;
}
template <class T>
__device__ inline ResidentLds<T> carve_vec(void* base, int m, int n, int ntA, int ntT,
                                           int tnnz) {
  ResidentLds<T> r{};
  r.dx = (SOV_LDS T*)((char*)base + vec_offset<T>(tnnz));
  r.dy = r.dx + n;
  r.rowsA = (SOV_LDS int32_t*)(r.dy + m);
  r.nzsA = r.rowsA + (ntA + 1);
  r.rowsT = r.nzsA + (ntA + 1);
  r.nzsT = r.rowsT + (ntT + 1);
  return r;
}

// Every thread of the workgroup: copy the current dx and dy into LDS.
template <class T, int BLOCK>
__device__ inline void load_vec(const LpDev<T>& l, const ResidentLds<T>& r) {
  for (int j = (int)threadIdx.x; j < l.n; j += BLOCK) r.dx[j] = l.dx[j];
  for (int i = (int)threadIdx.x; i < l.m; i += BLOCK) r.dy[i] = l.dy[i];
  for (int t = (int)threadIdx.x; t <= l.ntA; t += BLOCK) {
    const int row = l.tilesA[t];
    r.rowsA[t] = row;
    r.nzsA[t] = l.ap[row];
  }
  for (int t = (int)threadIdx.x; t <= l.ntT; t += BLOCK) {
    const int row = l.tilesT[t];
    r.rowsT[t] = row;
    r.nzsT[t] = l.tp[row];
  }
}

// Every thread of the workgroup: fill the resident copy from global memory
// (the scaled matrix, its plans, and the current dx and dy).  The caller
// places a workgroup barrier before the copy is read.
template <class T, int BLOCK>
__device__ inline void load_resident(const LpDev<T>& l, int nnz, const ResidentLds<T>& r) {
  for (int q = (int)threadIdx.x; q < nnz; q += BLOCK) {
    r.valA[q] = l.sav[q];
    r.valT[q] = l.stv[q];
    r.idxA[q] = (uint16_t)l.ai[q];
    r.idxT[q] = (uint16_t)l.ti[q];
  }
  for (int i = (int)threadIdx.x; i <= l.m; i += BLOCK) r.ptrA[i] = (uint16_t)l.ap[i];
  for (int j = (int)threadIdx.x; j <= l.n; j += BLOCK) r.ptrT[j] = (uint16_t)l.tp[j];
  for (int t = (int)threadIdx.x; t <= l.ntA; t += BLOCK) r.tilesA[t] = (uint16_t)l.tilesA[t];
  for (int t = (int)threadIdx.x; t <= l.ntT; t += BLOCK) r.tilesT[t] = (uint16_t)l.tilesT[t];
  for (int j = (int)threadIdx.x; j < l.n; j += BLOCK) r.dx[j] = l.dx[j];
  for (int i = (int)threadIdx.x; i < l.m; i += BLOCK) r.dy[i] = l.dy[i];
}

__device__ __forceinline__ double pow2_round(double s) {
  if (!(s > 0.0 && s <= 1.79769313486231570e308)) return 1.0;   // non-finite or <= 0
  return exp2(rint(log2(s)));
}

template <class T>
__device__ __forceinline__ T clampT(T v, T l, T u) {
  return v < l ? l : (v > u ? u : v);
}

__device__ __forceinline__ double bsel_row(double y, double rl, double ru) {
  return y > 0.0 ? (ru < kInf ? ru : 0.0) : (rl > -kInf ? rl : 0.0);
}
__device__ __forceinline__ double bsel_col(double d, double lo, double hi) {
  return d > 0.0 ? (lo > -kInf ? lo : 0.0) : (hi < kInf ? hi : 0.0);
}
__device__ __forceinline__ double dres_of(double d, double lo, double hi) {
  return (lo > -kInf ? 0.0 : fmax(d, 0.0)) + (hi < kInf ? 0.0 : fmax(-d, 0.0));
}

// Reduction slots of the CHECK phase
enum { PC = 0, DC, CXC, SUC, BNC, PA, DA, CXA, SUA, BNA, PUC, DUC, CXUC, SUUC, BNUC,
       PUA, DUA, CXUA, SUUA, BNUA, TFLAG, NCHK };

// Which reduced values of a phase are maxima (the rest are sums)
__host__ __device__ constexpr uint32_t phase_maxmask(int ph) {
  return ph == PH_RUIZ_MAX ? 1u : ph == PH_PC_APPLY ? 3u : ph == PH_SCALE ? ((1u << 3) | (1u << 4))
       : ph == PH_KKT0 ? (1u | 2u | 32u)
       : ph == PH_REF_KKT ? (1u | 2u | 32u)
       : ph == PH_REF_SHIFT ? (1u | 8u)
       : ph == PH_CHECK ? ((1u << PC) | (1u << DC) | (1u << PA) | (1u << DA) | (1u << PUC) |
                           (1u << DUC) | (1u << PUA) | (1u << DUA) | (1u << TFLAG))
       : 0u;
}

// ---------------------------------------------------------------------------
// Phase work.  `rank`/`size` split the LP's rows, columns and tiles among the
// workgroups that share the phase (size == 1 for batched LPs).  Each call adds
// this workgroup's contribution to part[]; the caller reduces.
// ---------------------------------------------------------------------------
template <class T, int BLOCK, int TNNZ>
struct PhaseWork {
  const LpDev<T>& l;
  const LpState& s;
  T* lds;
  double* ldsd;
  int rank, size;
  // the LP's resident copy in LDS (loop phases) when has_res, or only its dx
  // and dy when has_vec; held by value so that its pointers stay in registers
  ResidentLds<T> res;
  bool has_res, has_vec, lds_dxdy;

  __device__ PhaseWork(const LpDev<T>& l_, const LpState& s_, T* lds_, int rank_, int size_,
                       const ResidentLds<T>& res_ = ResidentLds<T>{}, bool has_res_ = false,
                       bool has_vec_ = false)
      : l(l_), s(s_), lds(lds_), ldsd(reinterpret_cast<double*>(lds_)), rank(rank_),
        size(size_), res(res_), has_res(has_res_), has_vec(has_vec_),
        lds_dxdy(has_res_ || has_vec_) {}

  __device__ __forceinline__ void erange(long long len, long long* a, long long* b) const {
    *a = len * rank / size;
    *b = len * (rank + 1) / size;
  }
  __device__ __forceinline__ void trange(int ntiles, const int32_t* part, int* a, int* b) const {
    if (size == 1) {
      *a = 0;
      *b = ntiles;
    } else {
      *a = part[rank];
      *b = part[rank + 1];
    }
  }

  // the row-slack cost of a refinement round (scaled; zero without refinement)
  __device__ __forceinline__ T qs(long long i) const {
    if constexpr (sizeof(T) == 4) return l.q[i];
    else return T(0);
  }
  // the same, unscaled FP64, and the current round's objective
  __device__ __forceinline__ double qu(long long i) const {
    if constexpr (sizeof(T) == 4) return l.qc[i];
    else return 0.0;
  }
  __device__ __forceinline__ double cu(long long j) const {
    if constexpr (sizeof(T) == 4) return l.cc[j];
    else return l.c0[j];
  }

  // dual prox of a row: y' = w - sigma clamp((w - q) / sigma, rl, ru)
  // (backend.py:219-231 with the refinement round's slack cost q).  FP32: when
  // the clamp is inactive the result is q itself, not w - sigma ((w - q) /
  // sigma) with its rounding residue: that residue, multiplied by a shifted
  // row bound of a refinement round, would swamp the round's duality gap.
  // FP64 keeps the reference's expression.
  __device__ __forceinline__ T dual_prox(T w, T q, T sg, T rl, T ru) const {
    const T t = (w - q) / sg;
    if constexpr (sizeof(T) == 4) {
      return t < rl ? w - sg * rl : (t > ru ? w - sg * ru : q);
    } else {
      return w - sg * clampT<T>(t, rl, ru);
    }
  }

  // next iteration's primal projection, in place (fused) or out of place (trial)
  __device__ __forceinline__ void next_primal(int j, T xj, T aty, int nm, T tt, double& dxdx) {
    if (nm == NEXT_NONE) return;
    const T v = clampT<T>(xj - tt * (l.c[j] + aty), l.lo[j], l.hi[j]);
    const T d = v - xj;
    l.dx[j] = d;
    if (lds_dxdy) res.dx[j] = d;
    if (nm == NEXT_FUSED) {
      l.x[j] = v;
    } else {
      l.xn[j] = v;
      dxdx += (double)d * (double)d;
    }
  }

  // rank 0's thread 0 reports whether the time budget is spent
  __device__ __forceinline__ double clock_flag() const {
    if (rank == 0 && threadIdx.x == 0)
      return ((long long)wall_clock64() - s.tick0) > s.ticks_left ? 1.0 : 0.0;
    return 0.0;
  }

  __device__ __forceinline__ void run(int ph, double (&part)[kMaxRed]) {
    long long i0, i1, j0, j1;
    erange(l.m, &i0, &i1);
    erange(l.n, &j0, &j1);
    int a0, a1, b0, b1;
    trange(l.ntA, l.partA, &a0, &a1);
    trange(l.ntT, l.partT, &b0, &b1);
    const int tid = threadIdx.x;

    switch (ph) {
      case PH_INIT: {
        for (long long i = i0 + tid; i < i1; i += BLOCK) l.rowf[i] = 1.0;
        for (long long j = j0 + tid; j < j1; j += BLOCK) l.colf[j] = 1.0;
      } break;

      case PH_RUIZ_MAX: {   // row/col max of |a_ij| r_i c_j; deviation from 1
        auto prodA = [&](int q, int j, double (&v)[1]) { v[0] = fabs(l.av[q]) * l.colf[j]; };
        auto epiA = [&](int i, const double (&sm)[1]) {
          double r = l.rowf[i] * sm[0];
          if (r <= 0.0) r = 1.0;
          part[0] = nanmax(part[0], fabs(1.0 - r));
          l.tmpm[i] = r;
        };
        tile_pass<BLOCK, TNNZ, 1, true, double>(l.ap, l.ai, l.tilesA, a0, a1, prodA, epiA, ldsd);
        auto prodT = [&](int q, int i, double (&v)[1]) { v[0] = fabs(l.tv[q]) * l.rowf[i]; };
        auto epiT = [&](int j, const double (&sm)[1]) {
          double c = l.colf[j] * sm[0];
          if (c <= 0.0) c = 1.0;
          part[0] = nanmax(part[0], fabs(1.0 - c));
          l.tmpn[j] = c;
        };
        tile_pass<BLOCK, TNNZ, 1, true, double>(l.tp, l.ti, l.tilesT, b0, b1, prodT, epiT, ldsd);
      } break;

      case PH_RUIZ_APPLY: {
        for (long long i = i0 + tid; i < i1; i += BLOCK) l.rowf[i] *= 1.0 / sqrt(l.tmpm[i]);
        for (long long j = j0 + tid; j < j1; j += BLOCK) l.colf[j] *= 1.0 / sqrt(l.tmpn[j]);
      } break;

      case PH_PC_SUM: {     // Pock-Chambolle sums (alpha = 1) on the Ruiz-scaled matrix
        auto prodA = [&](int q, int j, double (&v)[1]) { v[0] = fabs(l.av[q]) * l.colf[j]; };
        auto epiA = [&](int i, const double (&sm)[1]) {
          const double r = l.rowf[i] * sm[0];
          l.tmpm[i] = r <= 0.0 ? 1.0 : r;
        };
        tile_pass<BLOCK, TNNZ, 1, false, double>(l.ap, l.ai, l.tilesA, a0, a1, prodA, epiA, ldsd);
        auto prodT = [&](int q, int i, double (&v)[1]) { v[0] = fabs(l.tv[q]) * l.rowf[i]; };
        auto epiT = [&](int j, const double (&sm)[1]) {
          const double c = l.colf[j] * sm[0];
          l.tmpn[j] = c <= 0.0 ? 1.0 : c;
        };
        tile_pass<BLOCK, TNNZ, 1, false, double>(l.tp, l.ti, l.tilesT, b0, b1, prodT, epiT, ldsd);
      } break;

      case PH_PC_APPLY: {   // power-of-two factors; extremes of |c_j col_j| over nonzeros
        for (long long i = i0 + tid; i < i1; i += BLOCK)
          l.rowf[i] = pow2_round(l.rowf[i] * (1.0 / sqrt(l.tmpm[i])));
        for (long long j = j0 + tid; j < j1; j += BLOCK) {
          const double cf = pow2_round(l.colf[j] * (1.0 / sqrt(l.tmpn[j])));
          l.colf[j] = cf;
          const double v = l.c0[j] * cf;
          if (v != 0.0) {
            part[0] = nanmax(part[0], fabs(v));
            part[1] = nanmax(part[1], -fabs(v));
          }
        }
      } break;

      case PH_SCALE: {      // scaled copy, bounds, x0, y0; ||c'||^2 ||b'||^2 ||v0||^2 max|c| rhs
        auto wA = [&](int q, int i, int j) { l.sav[q] = (T)(l.av[q] * (l.rowf[i] * l.colf[j])); };
        tile_foreach<BLOCK, TNNZ>(l.ap, l.ai, l.tilesA, a0, a1, wA);
        auto wT = [&](int q, int j, int i) { l.stv[q] = (T)(l.tv[q] * (l.rowf[i] * l.colf[j])); };
        tile_foreach<BLOCK, TNNZ>(l.tp, l.ti, l.tilesT, b0, b1, wT);
        const double so = s.sobj;
        for (long long j = j0 + tid; j < j1; j += BLOCK) {
          const double cf = l.colf[j];
          const T cj = (T)((l.c0[j] * cf) * so);
          const double lo0 = l.lo0[j], hi0 = l.hi0[j];
          const T lo = (T)(lo0 <= -kInf ? -kInf : (lo0 >= kInf ? kInf : lo0 / cf));
          const T hi = (T)(hi0 <= -kInf ? -kInf : (hi0 >= kInf ? kInf : hi0 / cf));
          if (l.xacc) {     // round 0 solves the original problem; X = 0
            l.wlo[j] = lo0;
            l.whi[j] = hi0;
            l.xacc[j] = 0.0;
            l.cc[j] = l.c0[j];
          }
          l.c[j] = cj;
          l.lo[j] = lo;
          l.hi[j] = hi;
          // warm start: the reported x back in scaled units (x' = x / col)
          const T x0 = clampT<T>(l.wx0 ? (T)(l.wx0[j] / cf) : T(0), lo, hi);
          l.x[j] = x0;
          l.xr[j] = x0;
          l.sumx[j] = T(0);
          l.aty[j] = T(0);
          part[0] += (double)cj * (double)cj;
          part[2] += l.v0[j] * l.v0[j];
          part[3] = nanmax(part[3], fabs(l.c0[j]));
        }
        for (long long i = i0 + tid; i < i1; i += BLOCK) {
          const double rf = l.rowf[i];
          const double rl0 = l.rl0[i], ru0 = l.ru0[i];
          const T rl = (T)(rl0 <= -kInf ? -kInf : (rl0 >= kInf ? kInf : rl0 * rf));
          const T ru = (T)(ru0 <= -kInf ? -kInf : (ru0 >= kInf ? kInf : ru0 * rf));
          if (l.xacc) {
            l.wrl[i] = rl0;
            l.wru[i] = ru0;
            l.yacc[i] = 0.0;
            l.qc[i] = 0.0;
            l.q[i] = T(0);
          }
          l.rl[i] = rl;
          l.ru[i] = ru;
          // warm start: the reported (textbook) y back to the internal sign
          // and scaled units, inverting FINAL: y' = -y s / row (y' = y s / row
          // for a maximisation)
          T y0 = T(0);
          if (l.wy0) {
            const double yw = l.flip ? l.wy0[i] : -l.wy0[i];
            y0 = (T)((yw * so) / rf);
          }
          l.y[i] = y0;
          l.yr[i] = y0;
          l.sumy[i] = T(0);
          const double b = (rl > (T)(-kInf) ? (double)rl : 0.0) + (ru < (T)kInf ? (double)ru : 0.0);
          part[1] += b * b;
          if (rl0 > -kInf) part[4] = nanmax(part[4], fabs(rl0));
          if (ru0 < kInf) part[4] = nanmax(part[4], fabs(ru0));
        }
        // a warm start's primal weight reaches the transition as a sum of one term
        if (rank == 0 && tid == 0 && l.womega > 0.0) part[5] = l.womega;
      } break;

      case PH_AX0: {        // ax = A x0 (aty = A^T y0 when warm); v = v0 / ||v0|| (in xn)
        auto prod = [&](int q, int j, T (&v)[1]) { v[0] = l.sav[q] * l.x[j]; };
        auto epi = [&](int i, const T (&sm)[1]) { l.ax[i] = sm[0]; };
        tile_pass<BLOCK, TNNZ, 1, false, T>(l.ap, l.ai, l.tilesA, a0, a1, prod, epi, lds);
        if (l.wy0 && s.round == 0) {
          auto prodT = [&](int q, int i, T (&v)[1]) { v[0] = l.stv[q] * l.y[i]; };
          auto epiT = [&](int j, const T (&sm)[1]) { l.aty[j] = sm[0]; };
          tile_pass<BLOCK, TNNZ, 1, false, T>(l.tp, l.ti, l.tilesT, b0, b1, prodT, epiT, lds);
        }
        const double nv = s.nv;
        if (nv != 0.0)
          for (long long j = j0 + tid; j < j1; j += BLOCK) l.xn[j] = (T)(l.v0[j] / nv);
      } break;

      case PH_POW_A: {      // yn = A v, v = xn (first) or dx / ||dx||
        const bool first = s.pit == 0;
        const double nrm = s.nrm;
        auto prod = [&](int q, int j, T (&v)[1]) {
          v[0] = l.sav[q] * (first ? l.xn[j] : (T)((double)l.dx[j] / nrm));
        };
        auto epi = [&](int i, const T (&sm)[1]) { l.yn[i] = sm[0]; };
        tile_pass<BLOCK, TNNZ, 1, false, T>(l.ap, l.ai, l.tilesA, a0, a1, prod, epi, lds);
      } break;

      case PH_POW_T: {      // dx = A^T yn; ||dx||^2
        auto prod = [&](int q, int i, T (&v)[1]) { v[0] = l.stv[q] * l.yn[i]; };
        auto epi = [&](int j, const T (&sm)[1]) {
          l.dx[j] = sm[0];
          part[0] += (double)sm[0] * (double)sm[0];
        };
        tile_pass<BLOCK, TNNZ, 1, false, T>(l.tp, l.ti, l.tilesT, b0, b1, prod, epi, lds);
      } break;

      case PH_KKT0: {       // scaled KKT of (x0, y0, A x0, 0): pres dres c.x supp bnd time
        for (long long i = i0 + tid; i < i1; i += BLOCK) {
          const double a = l.ax[i], rl = l.rl[i], ru = l.ru[i], q = qs(i);
          const double z = (double)l.y[i] - q;
          part[0] = nanmax(part[0], fabs(fmax(rl - a, 0.0) + fmax(a - ru, 0.0)));
          part[3] += bsel_row(z, rl, ru) * z;
          part[2] += q * a;
        }
        for (long long j = j0 + tid; j < j1; j += BLOCK) {
          const double c = l.c[j], d = c + (double)l.aty[j], lo = l.lo[j], hi = l.hi[j];
          part[1] = nanmax(part[1], fabs(dres_of(d, lo, hi)));
          part[2] += c * (double)l.x[j];
          part[4] += bsel_col(d, lo, hi) * d;
        }
        part[5] = clock_flag();
      } break;

      case PH_P_TRIAL: {    // x' = clip(x - tau (c + aty)); dx = x' - x
        const T tt = (T)(s.eta / s.omega);
        for (long long j = j0 + tid; j < j1; j += BLOCK) {
          const T xj = l.x[j];
          const T v = clampT<T>(xj - tt * (l.c[j] + l.aty[j]), l.lo[j], l.hi[j]);
          l.xn[j] = v;
          const T d = v - xj;
          l.dx[j] = d;
          if (lds_dxdy) res.dx[j] = d;
          part[0] += (double)d * (double)d;
        }
      } break;

      case PH_A_TRIAL: {    // adx = A dx; y' = prox(y + sigma (ax + 2 adx)); dy.adx, dy.dy
        const T sg = (T)(s.eta * s.omega);
        struct R { T y, ax, rl, ru, q; };
        auto prod = [&](int q, int j, T (&v)[1]) { v[0] = l.sav[q] * l.dx[j]; };
        auto pre = [&](int i) { return R{l.y[i], l.ax[i], l.rl[i], l.ru[i], qs(i)}; };
        auto epi = [&](int i, const T (&sm)[1], const R& r) {
          const T a = sm[0];
          l.adx[i] = a;
          const T w = r.y + sg * (r.ax + T(2) * a);
          const T yv = dual_prox(w, r.q, sg, r.rl, r.ru);
          l.yn[i] = yv;
          const T d = yv - r.y;
          l.dy[i] = d;
          if (lds_dxdy) res.dy[i] = d;
          part[0] += (double)d * (double)a;
          part[1] += (double)d * (double)d;
        };
        if (has_res) {
          tile_pass_res<BLOCK, T>(res.valA, res.idxA, res.ptrA, res.tilesA, res.dx, a0, a1,
                                  pre, epi);
        } else if (has_vec) {
          const LdsGather<T> ld{l.ai, l.sav, res.dx};
          const LdsBounds bd{res.rowsA, res.nzsA};
          tile_pass_ld<BLOCK, TNNZ, 1, false, T>(l.ap, bd, a0, a1, ld, pre, epi, lds);
        } else {
          tile_pass_pf<BLOCK, TNNZ, 1, false, T>(l.ap, l.ai, l.tilesA, a0, a1, prod, pre, epi, lds);
        }
      } break;

      case PH_B_ADAPT: {    // accepted trial
        const T et = (T)s.eta, tt = (T)(s.eta / s.omega);
        const int nm = s.next_mode;
        double dxdx = 0.0;
        struct R { T xn, aty, sumx; };
        auto prod = [&](int q, int i, T (&v)[1]) { v[0] = l.stv[q] * l.dy[i]; };
        auto pre = [&](int j) { return R{l.xn[j], l.aty[j], l.sumx[j]}; };
        auto epi = [&](int j, const T (&sm)[1], const R& r) {
          const T xj = r.xn;
          l.x[j] = xj;
          const T aty = r.aty + sm[0];
          l.aty[j] = aty;
          l.sumx[j] = r.sumx + et * xj;
          next_primal(j, xj, aty, nm, tt, dxdx);
        };
        if (has_res) {
          // the row sweep's operands are loaded before the column pass, so that
          // their latency overlaps it (the sweep and the pass touch disjoint
          // arrays: the results are the same in either order)
          constexpr int kSw = 2;
          T yn_[kSw], ax_[kSw], adx_[kSw], sy_[kSw];
          const bool early = i1 - i0 <= (long long)kSw * BLOCK;
          if (early) {
#pragma unroll
            for (int k = 0; k < kSw; ++k) {
              const long long i = i0 + tid + (long long)k * BLOCK;
              if (i < i1) {
                yn_[k] = l.yn[i];
                ax_[k] = l.ax[i];
                adx_[k] = l.adx[i];
                sy_[k] = l.sumy[i];
              }
            }
          }
          tile_pass_res<BLOCK, T>(res.valT, res.idxT, res.ptrT, res.tilesT, res.dy, b0, b1,
                                  pre, epi);
          if (early) {
#pragma unroll
            for (int k = 0; k < kSw; ++k) {
              const long long i = i0 + tid + (long long)k * BLOCK;
              if (i < i1) {
                l.y[i] = yn_[k];
                l.ax[i] = ax_[k] + adx_[k];
                l.sumy[i] = sy_[k] + et * yn_[k];
              }
            }
            part[0] += dxdx;
            break;
          }
        } else if (has_vec) {
          const LdsGather<T> ld{l.ti, l.stv, res.dy};
          const LdsBounds bd{res.rowsT, res.nzsT};
          tile_pass_ld<BLOCK, TNNZ, 1, false, T>(l.tp, bd, b0, b1, ld, pre, epi, lds);
        } else {
          tile_pass_pf<BLOCK, TNNZ, 1, false, T>(l.tp, l.ti, l.tilesT, b0, b1, prod, pre, epi, lds);
        }
        for (long long i = i0 + tid; i < i1; i += BLOCK) {
          const T yv = l.yn[i];
          l.y[i] = yv;
          l.ax[i] += l.adx[i];
          l.sumy[i] += et * yv;
        }
        part[0] += dxdx;
      } break;

      case PH_A_FUSED: {    // fused dual step (backend.py:260-279) and sum_y += eta y
        const T sg = (T)(s.eta * s.omega), et = (T)s.eta;
        struct R { T ax, y, rl, ru, sumy, q; };
        auto prod = [&](int q, int j, T (&v)[1]) { v[0] = l.sav[q] * l.dx[j]; };
        auto pre = [&](int i) { return R{l.ax[i], l.y[i], l.rl[i], l.ru[i], l.sumy[i], qs(i)}; };
        auto epi = [&](int i, const T (&sm)[1], const R& r) {
          const T a = sm[0];
          const T axn = r.ax + a;
          l.ax[i] = axn;
          const T w = r.y + sg * (axn + a);
          const T yv = dual_prox(w, r.q, sg, r.rl, r.ru);
          const T d = yv - r.y;
          l.dy[i] = d;
          if (lds_dxdy) res.dy[i] = d;
          l.y[i] = yv;
          l.sumy[i] = r.sumy + et * yv;
        };
        if (has_res) {
          tile_pass_res<BLOCK, T>(res.valA, res.idxA, res.ptrA, res.tilesA, res.dx, a0, a1,
                                  pre, epi);
        } else if (has_vec) {
          const LdsGather<T> ld{l.ai, l.sav, res.dx};
          const LdsBounds bd{res.rowsA, res.nzsA};
          tile_pass_ld<BLOCK, TNNZ, 1, false, T>(l.ap, bd, a0, a1, ld, pre, epi, lds);
        } else {
          tile_pass_pf<BLOCK, TNNZ, 1, false, T>(l.ap, l.ai, l.tilesA, a0, a1, prod, pre, epi, lds);
        }
      } break;

      case PH_B_FUSED: {    // aty += A^T dy; sum_x += eta x (backend.py:282-291); next primal
        const T et = (T)s.eta, tt = (T)(s.eta / s.omega);
        const int nm = s.next_mode;
        double dxdx = 0.0;
        struct R { T aty, x, sumx, c, lo, hi; };
        auto prod = [&](int q, int i, T (&v)[1]) { v[0] = l.stv[q] * l.dy[i]; };
        auto pre = [&](int j) {
          R r{l.aty[j], l.x[j], l.sumx[j], T(0), T(0), T(0)};
          if (nm != NEXT_NONE) {
            r.c = l.c[j];
            r.lo = l.lo[j];
            r.hi = l.hi[j];
          }
          return r;
        };
        auto epi = [&](int j, const T (&sm)[1], const R& r) {
          const T aty = r.aty + sm[0];
          l.aty[j] = aty;
          const T xj = r.x;
          l.sumx[j] = r.sumx + et * xj;
          if (nm == NEXT_NONE) return;
          const T v = clampT<T>(xj - tt * (r.c + aty), r.lo, r.hi);
          const T d = v - xj;
          l.dx[j] = d;
          if (lds_dxdy) res.dx[j] = d;
          if (nm == NEXT_FUSED) {
            l.x[j] = v;
          } else {
            l.xn[j] = v;
            dxdx += (double)d * (double)d;
          }
        };
        if (has_res) {
          tile_pass_res<BLOCK, T>(res.valT, res.idxT, res.ptrT, res.tilesT, res.dy, b0, b1,
                                  pre, epi);
        } else if (has_vec) {
          const LdsGather<T> ld{l.ti, l.stv, res.dy};
          const LdsBounds bd{res.rowsT, res.nzsT};
          tile_pass_ld<BLOCK, TNNZ, 1, false, T>(l.tp, bd, b0, b1, ld, pre, epi, lds);
        } else {
          tile_pass_pf<BLOCK, TNNZ, 1, false, T>(l.tp, l.ti, l.tilesT, b0, b1, prod, pre, epi, lds);
        }
        part[0] += dxdx;
      } break;

      case PH_CHECK: check(part, a0, a1, b0, b1); break;

      case PH_RESTART: {    // x, y <- candidate; ax, aty <- fresh products; distances
        const bool ua = s.use_avg;
        const double se = s.sum_eta;
        for (long long j = j0 + tid; j < j1; j += BLOCK) {
          const T xv = ua ? (T)((double)l.sumx[j] / se) : l.x[j];
          l.x[j] = xv;
          l.aty[j] = ua ? l.atya[j] : l.atyf[j];
          const double d = (double)xv - (double)l.xr[j];
          part[0] += d * d;
          l.xr[j] = xv;
          l.sumx[j] = T(0);
        }
        for (long long i = i0 + tid; i < i1; i += BLOCK) {
          const T yv = ua ? (T)((double)l.sumy[i] / se) : l.y[i];
          l.y[i] = yv;
          l.ax[i] = ua ? l.axa[i] : l.axf[i];
          const double d = (double)yv - (double)l.yr[i];
          part[1] += d * d;
          l.yr[i] = yv;
          l.sumy[i] = T(0);
        }
      } break;

      case PH_COPY_AVG: {
        const double se = s.sum_eta;
        for (long long j = j0 + tid; j < j1; j += BLOCK) l.x[j] = (T)((double)l.sumx[j] / se);
        for (long long i = i0 + tid; i < i1; i += BLOCK) l.y[i] = (T)((double)l.sumy[i] / se);
      } break;

      case PH_FINAL: {      // x = clip(x col), y = -(y row / s); objective partial
        const double rs = s.ref_scale;
        for (long long j = j0 + tid; j < j1; j += BLOCK) {
          double v = (double)l.x[j] * l.colf[j];
          if (l.xacc) {     // X plus the current round's correction, if not yet in X
            v = s.use_best ? l.xbest[j]
                           : l.xacc[j] + (s.absorbed ? 0.0 : fmin(fmax(v, l.clo[j]), l.chi[j]) / rs);
          }
          v = fmin(fmax(v, l.lo0[j]), l.hi0[j]);
          l.xo[j] = v;
          part[0] += l.c0[j] * v;
        }
        for (long long i = i0 + tid; i < i1; i += BLOCK) {
          double v = ((double)l.y[i] * l.rowf[i]) / s.sobj;
          if (l.xacc) v = s.use_best ? l.ybest[i] : l.yacc[i] + (s.absorbed ? 0.0 : v / rs);
          l.yo[i] = l.flip ? v : -v;
        }
      } break;

      case PH_FINAL_D: {    // reduced costs d = c - A^T y in original units
        auto prod = [&](int q, int i, double (&p)[1]) { p[0] = l.tv[q] * l.yo[i]; };
        auto epi = [&](int j, const double (&sm)[1]) {
          const double c = l.flip ? -l.c0[j] : l.c0[j];
          l.dout[j] = c - sm[0];
        };
        tile_pass<BLOCK, TNNZ, 1, false, double>(l.tp, l.ti, l.tilesT, b0, b1, prod, epi, ldsd);
      } break;

      case PH_REF_X: {      // X += clip(x col, lo_r, hi_r) / Delta; Y += (y row / s) / Delta
        const double rs = s.ref_scale, so = s.sobj;
        for (long long j = j0 + tid; j < j1; j += BLOCK) {
          const double v = fmin(fmax((double)l.x[j] * l.colf[j], l.clo[j]), l.chi[j]);
          l.xacc[j] += v / rs;
        }
        for (long long i = i0 + tid; i < i1; i += BLOCK)
          l.yacc[i] += (((double)l.y[i] * l.rowf[i]) / so) / rs;
      } break;

      case PH_REF_KKT: {    // FP64 KKT of (X, Y) on the original problem; A X and c + A^T Y kept
        struct RR { double rl, ru, y; };
        auto prodA = [&](int q, int j, double (&p)[1]) { p[0] = l.av[q] * l.xacc[j]; };
        auto preA = [&](int i) { return RR{l.rl0[i], l.ru0[i], l.yacc[i]}; };
        auto epiA = [&](int i, const double (&sm)[1], const RR& r) {
          const double a = sm[0];
          l.axacc[i] = a;
          part[0] = nanmax(part[0], fabs(fmax(r.rl - a, 0.0) + fmax(a - r.ru, 0.0)));
          part[3] += bsel_row(r.y, r.rl, r.ru) * r.y;
        };
        tile_pass_pf<BLOCK, TNNZ, 1, false, double>(l.ap, l.ai, l.tilesA, a0, a1, prodA, preA,
                                                    epiA, ldsd);
        struct RC { double c, lo, hi, x; };
        auto prodT = [&](int q, int i, double (&p)[1]) { p[0] = l.tv[q] * l.yacc[i]; };
        auto preT = [&](int j) { return RC{l.c0[j], l.lo0[j], l.hi0[j], l.xacc[j]}; };
        auto epiT = [&](int j, const double (&sm)[1], const RC& r) {
          const double du = r.c + sm[0];
          l.dacc[j] = du;
          part[0] = nanmax(part[0], fabs(fmax(r.lo - r.x, 0.0) + fmax(r.x - r.hi, 0.0)));
          part[1] = nanmax(part[1], fabs(dres_of(du, r.lo, r.hi)));
          part[2] += r.c * r.x;
          part[4] += bsel_col(du, r.lo, r.hi) * du;
        };
        tile_pass_pf<BLOCK, TNNZ, 1, false, double>(l.tp, l.ti, l.tilesT, b0, b1, prodT, preT,
                                                    epiT, ldsd);
        part[5] = clock_flag();
      } break;

      case PH_REF_SHIFT: {  // next round: bounds shifted by X (and A X), scaled by Delta
        const double rs = s.ref_scale;
        const bool save = s.save_best;   // X, Y just evaluated are the best so far
        for (long long j = j0 + tid; j < j1; j += BLOCK) {
          const double X = l.xacc[j], lo0 = l.lo0[j], hi0 = l.hi0[j], cf = l.colf[j];
          if (save) l.xbest[j] = X;
          const double lo = lo0 > -kInf ? fmax((lo0 - X) * rs, -kTrustRegion) : -kInf;
          const double hi = hi0 < kInf ? fmin((hi0 - X) * rs, kTrustRegion) : kInf;
          l.wlo[j] = lo;
          l.whi[j] = hi;
          const T los = (T)(lo > -kInf ? lo / cf : -kInf), his = (T)(hi < kInf ? hi / cf : kInf);
          l.lo[j] = los;
          l.hi[j] = his;
          const T x0 = clampT<T>(T(0), los, his);
          l.x[j] = x0;
          l.xr[j] = x0;
          l.sumx[j] = T(0);
          // objective of the round: Delta (c + A^T Y), scaled as the original
          const double cj = l.dacc[j] * rs;
          l.cc[j] = cj;
          const T cs = (T)((cj * cf) * s.sobj);
          l.c[j] = cs;
          l.aty[j] = T(0);      // the round's dual starts at Y (y^ = 0)
          part[2] += (double)cs * (double)cs;
          part[3] = nanmax(part[3], fabs(cj));
        }
        for (long long i = i0 + tid; i < i1; i += BLOCK) {
          const double a = l.axacc[i], rl0 = l.rl0[i], ru0 = l.ru0[i], rf = l.rowf[i];
          const double rl = rl0 > -kInf ? fmax((rl0 - a) * rs, -kTrustRegion) : -kInf;
          const double ru = ru0 < kInf ? fmin((ru0 - a) * rs, kTrustRegion) : kInf;
          l.wrl[i] = rl;
          l.wru[i] = ru;
          l.rl[i] = (T)(rl > -kInf ? rl * rf : -kInf);
          l.ru[i] = (T)(ru < kInf ? ru * rf : kInf);
          if (rl > -kInf) part[0] = nanmax(part[0], fabs(rl));
          if (ru < kInf) part[0] = nanmax(part[0], fabs(ru));
          if (save) l.ybest[i] = l.yacc[i];
          // row-slack cost -Delta Y (unscaled and scaled)
          const double qv = -(l.yacc[i] * rs);
          l.qc[i] = qv;
          l.q[i] = (T)((qv * s.sobj) / rf);
          l.y[i] = T(0);
          l.yr[i] = T(0);
          l.sumy[i] = T(0);
        }
      } break;

      default: break;
    }
  }

  // Two-vector passes give A x, A avg_x, A^T y, A^T avg_y; the epilogues
  // accumulate the scaled KKT of the current point (with the maintained ax,
  // aty, as _KKT.__call__ is given them) and of the average, and the unscaled
  // KKT of both.  FP64 iterates: the scaled products are exact power-of-two
  // multiples of the original ones (A' = R A C).  FP32 iterates: the unscaled
  // KKT is recomputed in FP64 from the original matrix, as _unscaled_kkt does.
  __device__ __forceinline__ void check(double (&v)[kMaxRed], int a0, int a1, int b0, int b1) {
    const double se = s.sum_eta;
    const double inv_se = se > 0.0 ? se : 1.0;
    const double so = s.sobj;
    constexpr bool kDerived = sizeof(T) == 8;

    auto unscaled_row = [&](int i, double au, double aua, double y, double ya) {
      const double rf = l.rowf[i], rl0 = l.crl[i], ru0 = l.cru[i], q = qu(i);
      const double yu = (y * rf) / so - q, yua = (ya * rf) / so - q;
      v[PUC] = nanmax(v[PUC], fabs(fmax(rl0 - au, 0.0) + fmax(au - ru0, 0.0)));
      v[SUUC] += bsel_row(yu, rl0, ru0) * yu;
      v[CXUC] += q * au;
      v[PUA] = nanmax(v[PUA], fabs(fmax(rl0 - aua, 0.0) + fmax(aua - ru0, 0.0)));
      v[SUUA] += bsel_row(yua, rl0, ru0) * yua;
      v[CXUA] += q * aua;
    };
    auto unscaled_col = [&](int j, double atu, double atua, double x, double xa) {
      const double cf = l.colf[j], lo0 = l.clo[j], hi0 = l.chi[j], c0 = cu(j);
      const double xu = x * cf, xua = xa * cf;
      const double du = c0 + atu, dua = c0 + atua;
      v[PUC] = nanmax(v[PUC], fabs(fmax(lo0 - xu, 0.0) + fmax(xu - hi0, 0.0)));
      v[DUC] = nanmax(v[DUC], fabs(dres_of(du, lo0, hi0)));
      v[CXUC] += c0 * xu;
      v[BNUC] += bsel_col(du, lo0, hi0) * du;
      v[PUA] = nanmax(v[PUA], fabs(fmax(lo0 - xua, 0.0) + fmax(xua - hi0, 0.0)));
      v[DUA] = nanmax(v[DUA], fabs(dres_of(dua, lo0, hi0)));
      v[CXUA] += c0 * xua;
      v[BNUA] += bsel_col(dua, lo0, hi0) * dua;
    };

    auto prodA = [&](int q, int j, T (&p)[2]) {
      const T a = l.sav[q];
      p[0] = a * l.x[j];
      p[1] = a * (T)((double)l.sumx[j] / inv_se);
    };
    auto epiA = [&](int i, const T (&sm)[2]) {
      l.axf[i] = sm[0];
      l.axa[i] = sm[1];
      const double rl = l.rl[i], ru = l.ru[i], y = l.y[i], q = qs(i);
      const double ya = (T)((double)l.sumy[i] / inv_se);
      const double am = l.ax[i], ag = sm[1];
      const double z = y - q, za = ya - q;
      v[PC] = nanmax(v[PC], fabs(fmax(rl - am, 0.0) + fmax(am - ru, 0.0)));
      v[SUC] += bsel_row(z, rl, ru) * z;
      v[CXC] += q * am;
      v[PA] = nanmax(v[PA], fabs(fmax(rl - ag, 0.0) + fmax(ag - ru, 0.0)));
      v[SUA] += bsel_row(za, rl, ru) * za;
      v[CXA] += q * ag;
      if (kDerived) {
        const double rf = l.rowf[i];
        unscaled_row(i, (double)sm[0] / rf, (double)sm[1] / rf, y, ya);
      }
    };
    tile_pass<BLOCK, TNNZ, 2, false, T>(l.ap, l.ai, l.tilesA, a0, a1, prodA, epiA, lds);

    auto prodT = [&](int q, int i, T (&p)[2]) {
      const T a = l.stv[q];
      p[0] = a * l.y[i];
      p[1] = a * (T)((double)l.sumy[i] / inv_se);
    };
    auto epiT = [&](int j, const T (&sm)[2]) {
      l.atyf[j] = sm[0];
      l.atya[j] = sm[1];
      const double c = l.c[j], lo = l.lo[j], hi = l.hi[j], x = l.x[j];
      const double xa = (T)((double)l.sumx[j] / inv_se);
      const double d = c + (double)l.aty[j], da = c + (double)sm[1];
      v[DC] = nanmax(v[DC], fabs(dres_of(d, lo, hi)));
      v[CXC] += c * x;
      v[BNC] += bsel_col(d, lo, hi) * d;
      v[DA] = nanmax(v[DA], fabs(dres_of(da, lo, hi)));
      v[CXA] += c * xa;
      v[BNA] += bsel_col(da, lo, hi) * da;
      if (kDerived) {
        const double kf = l.colf[j] * so;
        unscaled_col(j, (double)sm[0] / kf, (double)sm[1] / kf, x, xa);
      }
    };
    tile_pass<BLOCK, TNNZ, 2, false, T>(l.tp, l.ti, l.tilesT, b0, b1, prodT, epiT, lds);

    if constexpr (!kDerived) {
      // FP64 products of the original matrix, one vector per pass so that the
      // LDS holds TNNZ doubles: the current point's sums wait in tmpm / tmpn
      // (the same lane handles a row in both passes: same tile plan)
      auto prodA0 = [&](int q, int j, double (&p)[1]) {
        p[0] = l.av[q] * ((double)l.x[j] * l.colf[j]);
      };
      auto epiA0 = [&](int i, const double (&sm)[1]) { l.tmpm[i] = sm[0]; };
      tile_pass<BLOCK, TNNZ, 1, false, double>(l.ap, l.ai, l.tilesA, a0, a1, prodA0, epiA0, ldsd);
      auto prodA1 = [&](int q, int j, double (&p)[1]) {
        p[0] = l.av[q] * ((double)(T)((double)l.sumx[j] / inv_se) * l.colf[j]);
      };
      auto epiA1 = [&](int i, const double (&sm)[1]) {
        const double y = l.y[i], ya = (T)((double)l.sumy[i] / inv_se);
        unscaled_row(i, l.tmpm[i], sm[0], y, ya);
      };
      tile_pass<BLOCK, TNNZ, 1, false, double>(l.ap, l.ai, l.tilesA, a0, a1, prodA1, epiA1, ldsd);
      auto prodT0 = [&](int q, int i, double (&p)[1]) {
        p[0] = l.tv[q] * (((double)l.y[i] * l.rowf[i]) / so);
      };
      auto epiT0 = [&](int j, const double (&sm)[1]) { l.tmpn[j] = sm[0]; };
      tile_pass<BLOCK, TNNZ, 1, false, double>(l.tp, l.ti, l.tilesT, b0, b1, prodT0, epiT0, ldsd);
      auto prodT1 = [&](int q, int i, double (&p)[1]) {
        p[0] = l.tv[q] * (((double)(T)((double)l.sumy[i] / inv_se) * l.rowf[i]) / so);
      };
      auto epiT1 = [&](int j, const double (&sm)[1]) {
        const double x = l.x[j], xa = (T)((double)l.sumx[j] / inv_se);
        unscaled_col(j, l.tmpn[j], sm[0], x, xa);
      };
      tile_pass<BLOCK, TNNZ, 1, false, double>(l.tp, l.ti, l.tilesT, b0, b1, prodT1, epiT1, ldsd);
    }
    v[TFLAG] = clock_flag();
  }
};

// pow out of line: the step-size rule is taken inline in the run kernel, and
// pow's code there would cost the phases registers
__device__ __noinline__ double pow_call(double x, double y) { return pow(x, y); }

// ---------------------------------------------------------------------------
// Transition: applies the reduced values of phase `ph` to the state and picks
// the next phase.  Scalar logic only; runs on one thread.
// ---------------------------------------------------------------------------
struct Transition {
  const Params& P;
  double offset0;            // minimisation-form offset
  int flip;                  // the caller's problem maximises
  bool commit = true;        // write the history rows (false: only the state moves)

  __device__ __forceinline__ double merit(const LpState& s, double p, double d, double cx,
                                          double supp, double bnd, double w) const {
    const double soff = offset0 * s.sobj;
    const double pobj = cx + soff, dobj = (-supp + bnd) + soff;
    const double g = fabs(pobj - dobj), a = w * p, b = d / w;
    return sqrt(a * a + b * b + g * g);
  }

  __device__ __forceinline__ bool refining() const { return P.refine_rounds > 0; }

  // k mod a period (period >= 1): a mask when the period is a power of two
  // (the defaults, 64 and 8) and k >= 0, else the division, which is a long
  // instruction sequence on the one thread that takes the transition; the
  // same value either way
  __device__ __forceinline__ static long long kmod(long long k, int period) {
    if (k >= 0 && (period & (period - 1)) == 0) return k & (long long)(period - 1);
    return k % period;
  }

  __device__ __forceinline__ int next_mode(const LpState& s, long long k1) const {
    if (kmod(k1, P.check_every) == 0 || k1 >= s.kmax) return NEXT_NONE;
    return kmod(k1, P.stepsize_every) == 0 ? NEXT_TRIAL : NEXT_FUSED;
  }

  // top of the loop body (pdlp.py:313-317, 335): limits, then the step flavour
  // (the iteration limit counts the iterations of every refinement round)
  __device__ __forceinline__ int loop_top(LpState& s) const {
    if (s.k >= s.kmax) {
      s.status = ST_ITERATION_LIMIT;
      return PH_FINAL;
    }
    if (kmod(s.k, P.check_every) == 0 && s.time_flag) {
      s.status = ST_TIME_LIMIT;
      return PH_FINAL;
    }
    if (kmod(s.k, P.stepsize_every) == 0) {
      s.trial = 0;
      ++s.attempts;
      if (s.primal_ready) {
        s.primal_ready = 0;
        return PH_A_TRIAL;       // the previous B phase already formed x' and ||dx||^2
      }
      return PH_P_TRIAL;
    }
    s.primal_ready = 0;
    s.next_mode = next_mode(s, s.k + 1);
    return PH_A_FUSED;
  }

  // after the trial's A phase: the adaptive step-size rule (accept, or
  // shrink and try again), t = (dy.adx, ||dy||^2)
  __device__ __forceinline__ int trial_step(LpState& s, const double* t) const {
    const double interaction = fabs(t[0]);
    const double movement = 0.5 * (s.omega * s.dxdx + t[1] / s.omega);
    bool accepted = false;
    if (movement == 0.0) {
      accepted = true;
    } else {
      const double lim = interaction > 0.0 ? movement / interaction : __builtin_inf();
      const double tt = (double)(s.attempts + 1);
      const double nxt =
          fmax(fmin((1.0 - pow_call(tt, -0.3)) * lim, (1.0 + pow_call(tt, -0.6)) * s.eta), 1e-14);
      if (s.eta <= lim) accepted = true;
      else ++s.rejected;
      s.eta = nxt;
    }
    if (accepted) {
      s.next_mode = next_mode(s, s.k + 1);
      return PH_B_ADAPT;
    }
    ++s.trial;
    if (s.trial >= 60) {
      s.status = ST_NUMERICAL;
      return PH_FINAL;
    }
    ++s.attempts;
    return PH_P_TRIAL;
  }

  // after a B phase: the iteration is complete (pdlp.py:401-403)
  __device__ __forceinline__ int end_iteration(LpState& s, double dxdx) const {
    s.sum_eta += s.eta;
    ++s.k;
    ++s.epoch;
    s.primal_ready = s.next_mode != NEXT_NONE;
    if (s.next_mode == NEXT_TRIAL) s.dxdx = dxdx;
    if (kmod(s.k, P.check_every) == 0) return PH_CHECK;
    return loop_top(s);
  }

  // The transition of phase ph, by one thread.  The end of an iteration (after
  // a B phase), the most frequent transition, is taken inline on the caller's
  // state (in LDS: plain LDS accesses, no call); every other one out of line.
  __device__ __forceinline__ int next(LpState& s, int ph, const double* t, double* hist,
                                      long long hist_cap, double* rhist) const {
    if (ph == PH_B_FUSED || ph == PH_B_ADAPT) return end_iteration(s, t[0]);
    return apply(s, ph, t, hist, hist_cap, rhist);
  }

  // Called by one thread with `s` and `t` in LDS (out of line: the scalar logic
  // then costs the phase kernels neither registers nor scratch).
  __device__ __noinline__ int apply(LpState& s, int ph, const double* t, double* hist,
                                    long long hist_cap, double* rhist) const {
    switch (ph) {
      case PH_INIT: s.sweep = 0; return PH_RUIZ_MAX;
      case PH_RUIZ_MAX: s.dev = t[0]; return PH_RUIZ_APPLY;
      case PH_RUIZ_APPLY:
        if (s.dev < 1e-2 || s.sweep >= 9) return PH_PC_SUM;
        ++s.sweep;
        return PH_RUIZ_MAX;
      case PH_PC_SUM: return PH_PC_APPLY;
      case PH_PC_APPLY: {
        s.sobj = 1.0;
        if (t[0] > 0.0) {
          const double g = exp2(rint(log2(sqrt(t[0] * (-t[1])))));
          if (g > 0.0 && g <= 1.79769313486231570e308) s.sobj = 1.0 / g;
        }
        s.sobj0 = s.sobj;
        return PH_SCALE;
      }
      case PH_SCALE: {
        s.ptol0 = fmin(P.eps_abs + P.eps_rel * t[4], P.feas_cap);
        s.dtol0 = fmin(P.eps_abs + P.eps_rel * t[3], P.feas_cap);
        s.rhs0 = t[4];
        s.cinf0 = t[3];
        s.ref_scale = 1.0;
        s.round = 0;
        s.absorbed = 0;
        s.conv_last = 0;
        s.best_q = __builtin_inf();
        s.nobetter = 0;
        s.save_best = 0;
        s.use_best = 0;
        s.k_done = 0;
        s.kmax = P.max_iter;
        if (refining()) {
          // a round is solved to a relative tolerance that FP32 iterates can
          // reach; the absolute feasibility cap is met by the rounds together
          s.ptol = P.refine_eps + P.refine_eps * t[4];
          s.dtol = P.refine_eps + P.refine_eps * t[3];
        } else {
          s.ptol = s.ptol0;
          s.dtol = s.dtol0;
        }
        const double cn = sqrt(t[0]), bn = sqrt(t[1]);
        s.nv = sqrt(t[2]);
        const double w = bn > 0.0 ? exp(log(fmax(cn, 1e-12)) - log(fmax(bn, 1e-12))) : 1.0;
        s.omega = fmin(fmax(w, 1e-4), 1e4);
        // warm start: the primal weight the earlier solve ended with, in the
        // range the primal-weight update keeps it in (pdlp.py:459-461)
        if (t[5] > 0.0) s.omega = fmin(fmax(t[5], 1e-6), 1e6);
        return PH_AX0;
      }
      case PH_AX0:
        if (s.round > 0) return PH_KKT0;   // refinement round: ||A|| and eta carry over
        if (s.nv != 0.0) {
          s.pit = 0;
          s.est = 1.0;
          return PH_POW_A;
        }
        s.anorm = 1.0;
        s.eta = 1.0;
        return PH_KKT0;
      case PH_POW_A: return PH_POW_T;
      case PH_POW_T: {
        const double nrm = sqrt(t[0]);
        if (!(nrm > 0.0)) {
          s.anorm = 1.0;
          s.eta = 1.0;
          return PH_KKT0;
        }
        s.nrm = nrm;
        s.est = sqrt(nrm);
        ++s.pit;
        if (s.pit < 20) return PH_POW_A;
        s.anorm = fmax(s.est, 1e-12);
        s.eta = 1.0 / s.anorm;
        return PH_KKT0;
      }
      case PH_KKT0:
        s.mu_restart = merit(s, t[0], t[1], t[2], t[3], t[4], s.omega);
        s.mu_last = __builtin_inf();
        s.sum_eta = 0.0;
        s.k = s.epoch = s.attempts = 0;
        if (s.round == 0) s.restarts = s.rejected = s.hist_len = 0;
        s.kkt_best = __builtin_inf();
        s.k_best = 0;
        s.status = ST_RUNNING;
        s.time_flag = t[5] > 0.0;
        s.primal_ready = 0;
        return loop_top(s);
      case PH_P_TRIAL: s.dxdx = t[0]; return PH_A_TRIAL;
      case PH_A_TRIAL: return trial_step(s, t);
      case PH_B_ADAPT: return end_iteration(s, t[0]);
      case PH_A_FUSED: return PH_B_FUSED;
      case PH_B_FUSED: return end_iteration(s, t[0]);
      case PH_CHECK: {
        s.time_flag = t[TFLAG] > 0.0;
        const bool have_avg = s.sum_eta > 0.0;
        const double mu_c = merit(s, t[PC], t[DC], t[CXC], t[SUC], t[BNC], s.omega);
        const double mu_a = have_avg ? merit(s, t[PA], t[DA], t[CXA], t[SUA], t[BNA], s.omega)
                                     : __builtin_inf();
        const bool use_avg = mu_a < mu_c;
        const double mu_cand = use_avg ? mu_a : mu_c;
        const double pu = use_avg ? t[PUA] : t[PUC], du = use_avg ? t[DUA] : t[DUC];
        const double pou = (use_avg ? t[CXUA] : t[CXUC]) + offset0;
        const double dou = (-(use_avg ? t[SUUA] : t[SUUC]) + (use_avg ? t[BNUA] : t[BNUC])) + offset0;
        const double gu = fabs(pou - dou);
        s.use_avg = use_avg;
        s.primal_ready = 0;
        const double ea = refining() ? P.refine_eps : P.eps_abs;
        const double er = refining() ? P.refine_eps : P.eps_rel;
        // a correction round's gap is Delta^2 times the original gap at X + x^ /
        // Delta, Y + y^ / Delta, while its objective is the (tiny) change in
        // objective: its gap is measured against Delta^2 x the last original
        // gap, i.e. the round must shrink the original gap by the tolerance
        double gscale = fabs(pou) + fabs(dou);
        if (refining() && s.round > 0)
          gscale = fmax(gscale, s.ref_scale * s.ref_scale * s.gap_last);
        const double gtol = ea + er * gscale;
        if (refining()) {
          // a round stalls when its KKT ratio has not halved for
          // refine_stall iterations (FP32 iterates can stall short of the
          // round's tolerance), or after 8x that.  The round's current
          // candidate is then evaluated like a converged round's (REF_KKT):
          // the refinement goes on while evaluations improve on the best X
          // and stops, reporting the best, after two that do not (a round
          // that cuts the gap by less than 10x makes the next one zoom out)
          const double ratio = fmax(fmax(pu / s.ptol, du / s.dtol), gu / gtol);
          if (ratio < 0.5 * s.kkt_best) {
            s.kkt_best = ratio;
            s.k_best = s.k;
          }
          // rounds that only polish an X that already met _converged (for the
          // objective target) get refine_stall; rounds before that refine_stall_open
          const long long stall_c = P.refine_stall > 0 ? P.refine_stall : 1024;
          const long long stall = s.best_q <= 1.0 ? stall_c
                                : (P.refine_stall_open > 0 ? P.refine_stall_open : 4096);
          if (ratio > 1.0 && (s.k - s.k_best >= stall || s.k >= 8 * stall))
            return use_avg ? PH_COPY_AVG : PH_REF_X;
        }
        if (pu <= s.ptol && du <= s.dtol && gu <= gtol) {
          if (refining()) return use_avg ? PH_COPY_AVG : PH_REF_X;   // round converged
          s.status = ST_OPTIMAL;
          s.fin[0] = pu; s.fin[1] = du; s.fin[2] = gu; s.fin[3] = pou; s.fin[4] = dou;
          return use_avg ? PH_COPY_AVG : PH_FINAL;
        }
        if (commit && s.hist_len < hist_cap) {
          double* h = hist + 5 * s.hist_len;
          h[0] = (double)(s.k_done + s.k); h[1] = pu; h[2] = du; h[3] = gu; h[4] = pou;
        }
        ++s.hist_len;
        const bool restart = mu_cand <= P.beta_sufficient * s.mu_restart ||
                             (mu_cand <= P.beta_necessary * s.mu_restart && mu_cand > s.mu_last) ||
                             (double)s.epoch >= P.beta_artificial * (double)(s.k > 1 ? s.k : 1);
        s.mu_last = mu_cand;
        s.mu_cand = mu_cand;
        if (restart) return PH_RESTART;
        return loop_top(s);
      }
      case PH_RESTART: {
        const double ddx = sqrt(t[0]), ddy = sqrt(t[1]);
        if (ddx > 1e-12 && ddy > 1e-12) {
          const double th = P.primal_weight_smoothing;
          const double w = exp(th * log(ddy / ddx) + (1.0 - th) * log(s.omega));
          s.omega = fmin(fmax(w, 1e-6), 1e6);
        }
        s.mu_restart = s.mu_cand;
        s.sum_eta = 0.0;
        s.epoch = 0;
        ++s.restarts;
        return loop_top(s);
      }
      case PH_COPY_AVG: return s.status == ST_RUNNING && refining() ? PH_REF_X : PH_FINAL;
      case PH_REF_X:
        s.absorbed = 1;
        return PH_REF_KKT;
      case PH_REF_KKT: {
        // X against the original problem: the reference's _converged test
        // (pdlp.py:523-530) plus the refinement target on the primal residual
        s.time_flag = t[5] > 0.0;
        const double pres = t[0], dres = t[1];
        const double pou = t[2] + offset0, dou = (-t[3] + t[4]) + offset0;
        const double gap = fabs(pou - dou);
        if (s.ref_rows < kMaxRounds) {
          if (commit) {
            double* h = rhist + 5 * s.ref_rows;
            h[0] = (double)(s.k_done + s.k); h[1] = pres; h[2] = dres; h[3] = gap; h[4] = pou;
          }
          ++s.ref_rows;
        }
        if (s.round == 0) s.k_round0 = s.k;
        s.k_done += s.k;
        s.k = 0;
        const double gtol = P.eps_abs + P.eps_rel * (fabs(pou) + fabs(dou));
        const bool conv = pres <= s.ptol0 && dres <= s.dtol0 && gap <= gtol;
        // the _converged ratio of X, Y (<= 1: converged); the refinement keeps
        // the best X, Y by it and stops after two evaluations that do not
        // improve on the best (FP32 rounds can move X the wrong way)
        const double q = fmax(fmax(pres / fmax(s.ptol0, 1e-300), dres / fmax(s.dtol0, 1e-300)),
                              gap / fmax(gtol, 1e-300));
        const bool improved = q < s.best_q;
        if (improved) {
          s.best_q = q;
          s.best_fin[0] = pres; s.best_fin[1] = dres; s.best_fin[2] = gap;
          s.best_fin[3] = pou; s.best_fin[4] = dou;
          s.nobetter = 0;
        } else {
          ++s.nobetter;
        }
        s.save_best = 0;
        const double gap_before = s.gap_last;
        s.conv_last = conv;
        s.gap_last = gap;
        if (conv) {
          s.fin[0] = pres; s.fin[1] = dres; s.fin[2] = gap; s.fin[3] = pou; s.fin[4] = dou;
        }
        // refinement target: primal residual and duality gap of X, Y both at
        // refine_tol relative, so the objective is that close to the optimum
        const bool target = pres <= P.refine_tol * (1.0 + s.rhs0) &&
                            gap <= P.refine_tol * (1.0 + fabs(pou) + fabs(dou));
        if (conv && target) {
          s.status = ST_OPTIMAL;
          return PH_FINAL;
        }
        int stop = ST_RUNNING;
        if (s.k_done >= P.max_iter) stop = ST_ITERATION_LIMIT;
        else if (s.time_flag) stop = ST_TIME_LIMIT;
        else if (s.round + 1 >= P.refine_rounds || s.nobetter >= 2)
          stop = ST_NUMERICAL;       // the rounds could not reach the tolerance
        if (stop != ST_RUNNING) {
          // report the best X, Y (the current ones if they are the best);
          // OPTIMAL whenever the best met _converged
          s.use_best = !improved;
          if (s.best_q <= 1.0) {
            stop = ST_OPTIMAL;
            for (int k = 0; k < 5; ++k) s.fin[k] = s.best_fin[k];
          }
          s.status = stop;
          return PH_FINAL;
        }
        s.save_best = improved;
        // Delta: the power of two nearest below 1 / error, at most 2^20 times
        // the last one (and never below 1).  The error is the larger of the
        // residuals and the gap over the cost scale, which estimates how far X
        // is from optimal: a zoom set by the primal residual alone would trap
        // a feasible, suboptimal X inside the trust region.
        const double err = fmax(fmax(pres, dres), gap / (1.0 + s.cinf0));
        double d = exp2(floor(log2(1.0 / fmax(err, 1e-300))));
        d = fmin(d, s.ref_scale * 1048576.0);
        // a round that cut the gap by less than 10x was held back by its trust
        // region: zoom out 16x so the next one can move X further
        if (s.round > 0 && gap > 0.1 * gap_before) d = s.ref_scale / 16.0;
        d = fmax(d, 1.0);
        // the round works in x^ = Delta (x - X), y^ = Delta (y - Y) with
        // objective Delta (c + A^T Y) and bounds scaled by Delta: the same
        // PDHG in zoomed coordinates, so step size, primal weight and the
        // objective scale carry over unchanged
        s.ref_scale = d;
        ++s.round;
        return PH_REF_SHIFT;
      }
      case PH_REF_SHIFT: {
        // tolerances of the new round: relative to its own right-hand side,
        // capped at the original problem's scale
        s.ptol = fmin(P.refine_eps + P.refine_eps * t[0], P.refine_eps * (1.0 + s.rhs0));
        s.dtol = fmin(P.refine_eps + P.refine_eps * t[3], P.refine_eps * (1.0 + s.cinf0));
        s.absorbed = 0;
        s.kmax = P.max_iter - s.k_done;
        return PH_AX0;
      }
      case PH_FINAL: {
        // objective of the caller's problem (pdlp.py:473): prob.c = -c0 when maximising
        const double cx = flip ? -t[0] : t[0];
        s.objective = cx + (flip ? -offset0 : offset0);
        return PH_FINAL_D;
      }
      case PH_FINAL_D: return PH_DONE;
      default: return PH_DONE;
    }
  }
};

}  // namespace sov
