// crossover.cpp -- host code of libsovlp.so: a bounded primal revised simplex
// over [A | -I] and the basis identification that hands a PDLP point to it.
//
// Reference: src/sovopt/lp/crossover.py:80-222 (basis_from_point, crossover),
// the simplex it finishes with (src/sovopt/lp/simplex.py:1181, solve_simplex),
// and the basis conventions of src/sovopt/lp/basis.py:75-79, 203-246 (status
// codes, the all-logical basis B = -I, statuses from costs).
//
// The variables are z = [x ; s] with A x - s = 0: structurals j < n bounded by
// the column bounds, one logical per row bounded by the row bounds.  A basis
// is m of them; its matrix is factorised densely (LU with partial pivoting,
// column-major) and updated in product form between refactorisations.  Dense
// factors bound the size (kMaxRows rows); the crossover is a CPU step after
// the device solve, run once per LP.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/sovlp.h"

namespace sovh {
int fail(int code, const std::string& msg);
}

namespace {

using sovh::fail;

constexpr double kInf = 1.0e30;
constexpr uint8_t kBasic = 0, kLower = 1, kUpper = 2, kFree = 3, kFixed = 4;
// internal only: a nonbasic variable held at a value strictly between its
// bounds (the crossover's point); the simplex moves every one to a bound or
// into the basis before it reports
constexpr uint8_t kSuper = 5;
constexpr int kMaxRows = 8192;
constexpr double kSingular = 1e-11;   // smallest pivot the factorisation takes

enum { ST_OPTIMAL = 0, ST_INFEASIBLE = 1, ST_UNBOUNDED = 2, ST_ITERATION_LIMIT = 4,
       ST_TIME_LIMIT = 5, ST_NUMERICAL = 8 };

// ------------------------------------------------------------------ the LP
struct Lp {
  int m = 0, n = 0, N = 0;
  const int64_t* cp = nullptr;
  const int32_t* ci = nullptr;
  const double* cx = nullptr;
  std::vector<double> cost, lo, up;   // N entries: structurals, then logicals

  bool has_lo(int k) const { return lo[k] > -kInf; }
  bool has_up(int k) const { return up[k] < kInf; }
  // w . a_k for the column a_k of [A | -I]
  double dot(const double* w, int k) const {
    if (k >= n) return -w[k - n];
    double s = 0.0;
    for (int64_t p = cp[k]; p < cp[k + 1]; ++p) s += w[ci[p]] * cx[p];
    return s;
  }
  // out += f * a_k
  void axpy(double f, int k, double* out) const {
    if (k >= n) {
      out[k - n] -= f;
      return;
    }
    for (int64_t p = cp[k]; p < cp[k + 1]; ++p) out[ci[p]] += f * cx[p];
  }
};

int load(const sovlp_host_lp* h, Lp& L) {
  if (!h) return fail(1, "null LP");
  if (h->m < 0 || h->n < 0 || h->nnz < 0) return fail(1, "negative LP dimension");
  if (h->m > kMaxRows)
    return fail(4, "the simplex factorises its basis densely: at most 8192 rows");
  if (!h->cp || (h->nnz > 0 && (!h->ci || !h->cx)) || !h->c || !h->col_lb || !h->col_ub ||
      (h->m > 0 && (!h->row_lb || !h->row_ub)))
    return fail(1, "null LP array");
  if (h->cp[0] != 0 || h->cp[h->n] != h->nnz) return fail(1, "column pointers do not span nnz");
  for (int j = 0; j < h->n; ++j)
    if (h->cp[j + 1] < h->cp[j]) return fail(1, "column pointers decrease");
  for (int64_t p = 0; p < h->nnz; ++p)
    if (h->ci[p] < 0 || h->ci[p] >= h->m) return fail(1, "row index out of range");
  L.m = h->m;
  L.n = h->n;
  L.N = h->m + h->n;
  L.cp = h->cp;
  L.ci = h->ci;
  L.cx = h->cx;
  L.cost.assign(L.N, 0.0);
  L.lo.resize(L.N);
  L.up.resize(L.N);
  for (int j = 0; j < L.n; ++j) {
    L.cost[j] = h->c[j];
    L.lo[j] = h->col_lb[j];
    L.up[j] = h->col_ub[j];
  }
  for (int i = 0; i < L.m; ++i) {
    L.lo[L.n + i] = h->row_lb[i];
    L.up[L.n + i] = h->row_ub[i];
  }
  for (int k = 0; k < L.N; ++k) {
    if (std::isnan(L.lo[k]) || std::isnan(L.up[k]) || L.lo[k] > L.up[k])
      return fail(1, "a bound is NaN or a lower bound exceeds its upper bound");
    if (k < L.n && !std::isfinite(L.cost[k])) return fail(1, "non-finite cost");
  }
  return 0;
}

// -------------------------------------------------------- basis factors
// B = B0 E1 ... Ek: B0 = P^T L U (dense, column-major), each E replaces one
// column by alpha = B_prev^-1 a_q (product form of the inverse).
struct Eta {
  int r;
  double piv;
  std::vector<int> idx;
  std::vector<double> val;
};

struct Factor {
  int m = 0;
  std::vector<double> lu;   // m x m, column-major
  std::vector<int> perm;    // row i of P B0 is row perm[i] of B0
  std::vector<Eta> etas;

  // Factorises the columns head[0..m) of [A | -I].  A column without a pivot
  // of at least kSingular is replaced by the logical of a row that has none
  // yet (its position in `head` changes; `repairs` counts them) and the
  // factorisation restarts.  Returns the positions replaced.
  std::vector<int> factorize(const Lp& L, std::vector<int>& head) {
    std::vector<int> replaced;
    m = L.m;
    for (;;) {
      lu.assign((size_t)m * m, 0.0);
      for (int c = 0; c < m; ++c) {
        double* col = &lu[(size_t)c * m];
        const int k = head[c];
        if (k >= L.n) {
          col[k - L.n] = -1.0;
        } else {
          for (int64_t p = L.cp[k]; p < L.cp[k + 1]; ++p) col[L.ci[p]] += L.cx[p];
        }
      }
      perm.resize(m);
      std::iota(perm.begin(), perm.end(), 0);
      int bad = -1;
      for (int c = 0; c < m && bad < 0; ++c) {
        double* col = &lu[(size_t)c * m];
        int p = c;
        double best = std::fabs(col[c]);
        for (int i = c + 1; i < m; ++i)
          if (std::fabs(col[i]) > best) {
            best = std::fabs(col[i]);
            p = i;
          }
        if (best < kSingular) {
          bad = c;
          break;
        }
        if (p != c) {
          for (int j = 0; j < m; ++j) std::swap(lu[(size_t)j * m + c], lu[(size_t)j * m + p]);
          std::swap(perm[c], perm[p]);
        }
        const double inv = 1.0 / col[c];
        for (int i = c + 1; i < m; ++i) col[i] *= inv;
        for (int j = c + 1; j < m; ++j) {
          double* cj = &lu[(size_t)j * m];
          const double f = cj[c];
          if (f == 0.0) continue;
          for (int i = c + 1; i < m; ++i) cj[i] -= f * col[i];
        }
      }
      if (bad < 0) break;
      // rows perm[bad..m) have no pivot yet: the logical of the first one
      // is independent of the columns before `bad`
      head[bad] = L.n + perm[bad];
      replaced.push_back(bad);
    }
    etas.clear();
    return replaced;
  }

  // v <- B^-1 v
  void ftran(std::vector<double>& v, std::vector<double>& tmp) const {
    tmp.resize(m);
    for (int i = 0; i < m; ++i) tmp[i] = v[perm[i]];
    for (int c = 0; c < m; ++c) {           // L (unit lower)
      const double t = tmp[c];
      if (t == 0.0) continue;
      const double* col = &lu[(size_t)c * m];
      for (int i = c + 1; i < m; ++i) tmp[i] -= t * col[i];
    }
    for (int c = m - 1; c >= 0; --c) {      // U
      const double* col = &lu[(size_t)c * m];
      const double t = tmp[c] / col[c];
      tmp[c] = t;
      if (t == 0.0) continue;
      for (int i = 0; i < c; ++i) tmp[i] -= t * col[i];
    }
    v.swap(tmp);
    for (const Eta& e : etas) {
      const double t = v[e.r] / e.piv;
      v[e.r] = t;
      if (t == 0.0) continue;
      for (size_t q = 0; q < e.idx.size(); ++q) v[e.idx[q]] -= e.val[q] * t;
    }
  }

  // w <- B^-T w
  void btran(std::vector<double>& w, std::vector<double>& tmp) const {
    for (size_t k = etas.size(); k-- > 0;) {
      const Eta& e = etas[k];
      double s = w[e.r];
      for (size_t q = 0; q < e.idx.size(); ++q) s -= e.val[q] * w[e.idx[q]];
      w[e.r] = s / e.piv;
    }
    tmp.resize(m);
    for (int c = 0; c < m; ++c) {           // U^T (lower)
      const double* col = &lu[(size_t)c * m];
      double s = w[c];
      for (int i = 0; i < c; ++i) s -= col[i] * tmp[i];
      tmp[c] = s / col[c];
    }
    for (int c = m - 1; c >= 0; --c) {      // L^T (unit upper)
      const double* col = &lu[(size_t)c * m];
      double s = tmp[c];
      for (int i = c + 1; i < m; ++i) s -= col[i] * tmp[i];
      tmp[c] = s;
    }
    for (int i = 0; i < m; ++i) w[perm[i]] = tmp[i];
  }

  void update(int r, const std::vector<double>& alpha) {
    Eta e;
    e.r = r;
    e.piv = alpha[r];
    for (int i = 0; i < m; ++i)
      if (i != r && std::fabs(alpha[i]) > 1e-14) {
        e.idx.push_back(i);
        e.val.push_back(alpha[i]);
      }
    etas.push_back(std::move(e));
  }
};

// ------------------------------------------------------------- the simplex
struct Simplex {
  const Lp& L;
  sovlp_simplex_params P;
  std::vector<int> head, pos;
  std::vector<uint8_t> st;
  std::vector<double> x;       // every variable's value
  std::vector<double> pert;    // cost perturbation (phase 2, degenerate stalls)
  Factor F;
  std::vector<double> w1, w2, tmp;
  int64_t iters = 0, pivots = 0, phase1 = 0;
  int repairs = 0;
  bool perturbed = false;

  explicit Simplex(const Lp& l, const sovlp_simplex_params& p) : L(l), P(p) {}

  double ftol(int k) const {
    const double b = std::min(std::fabs(L.lo[k]), std::fabs(L.up[k]));
    return P.feas_tol * (1.0 + (b < kInf ? b : 0.0));
  }
  double nonbasic_value(int k) const {
    switch (st[k]) {
      case kUpper: return L.up[k];
      case kFree: return 0.0;
      case kSuper: return x[k];
      default: return L.lo[k];
    }
  }
  // a nonbasic status that names a finite bound
  uint8_t sane_status(int k, uint8_t s) const {
    const bool lo = L.has_lo(k), up = L.has_up(k);
    if (!lo && !up) return kFree;
    if (lo && up && L.up[k] - L.lo[k] <= 0.0) return kFixed;
    if (s == kUpper) return up ? kUpper : kLower;
    if (s == kLower || s == kFixed) return lo ? kLower : kUpper;
    return lo ? kLower : kUpper;     // kFree of a bounded variable
  }

  void set_logical() {
    head.resize(L.m);
    pos.assign(L.N, -1);
    st.assign(L.N, kLower);
    for (int i = 0; i < L.m; ++i) {
      head[i] = L.n + i;
      pos[L.n + i] = i;
      st[L.n + i] = kBasic;
    }
  }
  void status_from_costs() {   // basis.py:226-246
    for (int j = 0; j < L.n; ++j) {
      if (st[j] == kBasic) continue;
      const double lo = L.lo[j], up = L.up[j];
      if (lo <= -kInf && up >= kInf) st[j] = kFree;
      else if (up - lo <= 0.0) st[j] = kFixed;
      else if (L.cost[j] >= 0.0) st[j] = lo > -kInf ? kLower : kUpper;
      else st[j] = up < kInf ? kUpper : kLower;
    }
  }
  void refactor() {
    std::vector<int> rep = F.factorize(L, head);
    if (!rep.empty()) {
      // a column replaced by a logical rests at its nearest finite bound
      std::vector<uint8_t> was(st);
      std::vector<char> inb(L.N, 0);
      for (int k : head) inb[k] = 1;
      for (int k = 0; k < L.N; ++k) {
        if (inb[k]) {
          st[k] = kBasic;
        } else if (was[k] == kBasic) {
          const double v = x.empty() ? 0.0 : x[k];
          const bool nearer_lo = !L.has_up(k) || (L.has_lo(k) && std::fabs(v - L.lo[k]) <= std::fabs(L.up[k] - v));
          st[k] = sane_status(k, nearer_lo ? kLower : kUpper);
        }
      }
      repairs += (int)rep.size();
    }
    pos.assign(L.N, -1);
    for (int r = 0; r < L.m; ++r) pos[head[r]] = r;
  }
  // values of the nonbasics from their statuses; x_B = -B^-1 N x_N
  void refresh() {
    x.resize(L.N);
    w1.assign(L.m, 0.0);
    for (int k = 0; k < L.N; ++k) {
      if (st[k] == kBasic) continue;
      x[k] = nonbasic_value(k);
      if (x[k] != 0.0) L.axpy(-x[k], k, w1.data());
    }
    F.ftran(w1, tmp);
    for (int r = 0; r < L.m; ++r) x[head[r]] = w1[r];
  }
  double infeasibility() const {
    double s = 0.0;
    for (int r = 0; r < L.m; ++r) {
      const int k = head[r];
      if (x[k] < L.lo[k] - ftol(k)) s += L.lo[k] - x[k];
      else if (x[k] > L.up[k] + ftol(k)) s += x[k] - L.up[k];
    }
    return s;
  }
  double phase_cost(int k, bool ph1) const {
    if (!ph1) return L.cost[k] + (pert.empty() ? 0.0 : pert[k]);
    if (st[k] != kBasic) return 0.0;
    if (x[k] < L.lo[k] - ftol(k)) return -1.0;
    if (x[k] > L.up[k] + ftol(k)) return 1.0;
    return 0.0;
  }
  // duals of the phase's costs into w1
  void duals(bool ph1) {
    w1.assign(L.m, 0.0);
    for (int r = 0; r < L.m; ++r) w1[r] = phase_cost(head[r], ph1);
    F.btran(w1, tmp);
  }

  void perturb(std::mt19937_64& rng) {
    std::uniform_real_distribution<double> u(0.5, 1.0);
    pert.assign(L.N, 0.0);
    for (int k = 0; k < L.n; ++k) {
      if (st[k] == kBasic || st[k] == kFixed || st[k] == kSuper) continue;
      const double mag = P.perturb_scale * (1.0 + std::fabs(L.cost[k])) * u(rng);
      pert[k] = st[k] == kUpper ? -mag : mag;   // keeps the nonbasic where it rests
    }
    perturbed = true;
  }

  // Harris two-pass ratio test for a column alpha = B^-1 a_q entering in
  // direction dir (x_B moves by -dir t alpha): the blocking position (-1:
  // none), the step, and whether the blocking variable leaves at its upper
  // bound.  A basic variable outside its bounds blocks only when it would
  // reach the violated bound (phase 1's breakpoint).
  void ratio(const std::vector<double>& alpha, double dir, int& leave, double& t,
             bool& leave_up) const {
    double tmax = kInf;
    for (int r = 0; r < L.m; ++r) {
      const double a = alpha[r];
      if (std::fabs(a) < P.pivot_tol) continue;
      const int k = head[r];
      const double rate = -dir * a;
      const double tol = ftol(k);
      double lim = kInf;
      const bool below = x[k] < L.lo[k] - tol, above = x[k] > L.up[k] + tol;
      if (rate < 0.0) {
        if (above) lim = (x[k] - L.up[k] + tol) / -rate;
        else if (!below && L.has_lo(k)) lim = (x[k] - L.lo[k] + tol) / -rate;
      } else {
        if (below) lim = (L.lo[k] - x[k] + tol) / rate;
        else if (!above && L.has_up(k)) lim = (L.up[k] - x[k] + tol) / rate;
      }
      tmax = std::min(tmax, lim);
    }
    leave = -1;
    t = kInf;
    leave_up = false;
    if (tmax >= kInf) return;
    double big = 0.0;
    for (int r = 0; r < L.m; ++r) {
      const double a = alpha[r];
      if (std::fabs(a) < P.pivot_tol) continue;
      const int k = head[r];
      const double rate = -dir * a;
      const double tol = ftol(k);
      const bool below = x[k] < L.lo[k] - tol, above = x[k] > L.up[k] + tol;
      double lim = kInf;
      bool to_up = false;
      if (rate < 0.0) {
        if (above) { lim = (x[k] - L.up[k]) / -rate; to_up = true; }
        else if (!below && L.has_lo(k)) lim = (x[k] - L.lo[k]) / -rate;
      } else {
        if (below) lim = (L.lo[k] - x[k]) / rate;
        else if (!above && L.has_up(k)) { lim = (L.up[k] - x[k]) / rate; to_up = true; }
      }
      if (lim <= tmax && std::fabs(a) > big) {
        big = std::fabs(a);
        leave = r;
        t = std::max(lim, 0.0);
        leave_up = to_up;
      }
    }
  }

  // distance from x_q to its bound in direction dir (kInf: none)
  double own_room(int q, double dir) const {
    if (dir > 0.0) return L.has_up(q) ? std::max(L.up[q] - x[q], 0.0) : kInf;
    return L.has_lo(q) ? std::max(x[q] - L.lo[q], 0.0) : kInf;
  }

  // Moves x_q by dir t (the basics with it); then either x_q rests at the
  // bound it reached (`to_bound`) or it enters the basis at position `leave`.
  void step(int q, double dir, double t, const std::vector<double>& alpha, bool to_bound,
            int leave, bool leave_up) {
    x[q] += dir * t;
    for (int r = 0; r < L.m; ++r)
      if (alpha[r] != 0.0) x[head[r]] -= dir * t * alpha[r];
    ++iters;
    if (to_bound) {
      st[q] = sane_status(q, dir > 0.0 ? kUpper : kLower);
      x[q] = st[q] == kUpper ? L.up[q] : L.lo[q];
      return;
    }
    const int out = head[leave];
    const double bound = leave_up ? L.up[out] : L.lo[out];
    // the leaving variable rests exactly at its bound; when the ratio test
    // let it cross by up to its tolerance, the basic values take the
    // difference through its column, so they stay those of the nonbasics
    const double delta = bound - x[out];
    x[out] = bound;
    st[out] = sane_status(out, leave_up ? kUpper : kLower);
    head[leave] = q;
    pos[q] = leave;
    pos[out] = -1;
    st[q] = kBasic;
    F.update(leave, alpha);
    ++pivots;
    if ((int)F.etas.size() >= P.refactor_freq) {
      refactor();
      refresh();
    } else if (std::fabs(delta) > 1e-12 * (1.0 + std::fabs(bound))) {
      w2.assign(L.m, 0.0);
      L.axpy(1.0, out, w2.data());
      F.ftran(w2, tmp);
      for (int r = 0; r < L.m; ++r)
        if (w2[r] != 0.0) x[head[r]] -= delta * w2[r];
    }
  }

  // One phase of the primal simplex.  Phase 2 returns ST_OPTIMAL when no
  // column improves the cost; phase 1 returns ST_OPTIMAL once the basics are
  // within their bounds and ST_INFEASIBLE when no column reduces a positive
  // infeasibility.
  int loop(bool ph1, std::chrono::steady_clock::time_point t0) {
    std::mt19937_64 rng(12345);
    int64_t degen = 0;
    std::vector<double> alpha;
    for (;;) {
      if (iters >= P.max_iter) return ST_ITERATION_LIMIT;
      if ((iters & 63) == 0) {
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (el > P.time_limit) return ST_TIME_LIMIT;
      }
      if (ph1 && infeasibility() == 0.0) return ST_OPTIMAL;
      duals(ph1);
      // Dantzig pricing over the nonbasic columns
      int q = -1;
      double dq = 0.0, best = 0.0;
      for (int k = 0; k < L.N; ++k) {
        const uint8_t s = st[k];
        if (s == kBasic || s == kFixed) continue;
        const double d = (ph1 ? 0.0 : L.cost[k] + (pert.empty() ? 0.0 : pert[k])) - L.dot(w1.data(), k);
        const double tol = P.opt_tol * (ph1 ? 1.0 : 1.0 + std::fabs(L.cost[k]));
        bool ok = false;
        if (s == kLower) ok = d < -tol;
        else if (s == kUpper) ok = d > tol;
        else ok = std::fabs(d) > tol;        // free or held between its bounds
        if (ok && std::fabs(d) > best) {
          best = std::fabs(d);
          q = k;
          dq = d;
        }
      }
      if (q < 0) return ph1 ? ST_INFEASIBLE : ST_OPTIMAL;
      const double dir = dq < 0.0 ? 1.0 : -1.0;
      alpha.assign(L.m, 0.0);
      L.axpy(1.0, q, alpha.data());
      F.ftran(alpha, tmp);
      int leave;
      double t;
      bool leave_up;
      ratio(alpha, dir, leave, t, leave_up);
      const double room = own_room(q, dir);
      const bool to_bound = room < kInf && room <= t;
      if (!to_bound && leave < 0) return ph1 ? ST_NUMERICAL : ST_UNBOUNDED;
      if (to_bound) t = room;
      if (ph1) ++phase1;
      degen = t <= 1e-12 ? degen + 1 : 0;
      step(q, dir, t, alpha, to_bound, leave, leave_up);
      if (!ph1 && !perturbed && P.perturb_after > 0 && degen >= P.perturb_after) perturb(rng);
    }
  }

  bool has_super() const {
    for (int k = 0; k < L.N; ++k)
      if (st[k] == kSuper) return true;
    return false;
  }

  // Every variable still held between its bounds goes to its nearer bound,
  // or into the basis when a basic variable blocks it first (a free one
  // moves up, or down when nothing blocks it upward).  The reduced costs of
  // these variables are within tolerance of zero once phase 2 has ended, so
  // the pushes leave the objective where it is.
  int push(std::chrono::steady_clock::time_point t0) {
    std::vector<double> alpha;
    for (int k = 0; k < L.N; ++k) {
      if (st[k] != kSuper) continue;
      if (iters >= P.max_iter) return ST_ITERATION_LIMIT;
      if ((iters & 63) == 0) {
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (el > P.time_limit) return ST_TIME_LIMIT;
      }
      const double dl = L.has_lo(k) ? x[k] - L.lo[k] : kInf;
      const double du = L.has_up(k) ? L.up[k] - x[k] : kInf;
      double dir = du < dl ? 1.0 : -1.0;
      if (dl >= kInf && du >= kInf) dir = 1.0;
      alpha.assign(L.m, 0.0);
      L.axpy(1.0, k, alpha.data());
      F.ftran(alpha, tmp);
      int leave;
      double t;
      bool leave_up;
      ratio(alpha, dir, leave, t, leave_up);
      double room = own_room(k, dir);
      if (room >= kInf && leave < 0) {        // free and unblocked upward
        dir = -dir;
        ratio(alpha, dir, leave, t, leave_up);
        room = own_room(k, dir);
      }
      if (room >= kInf && leave < 0) {        // moves no basic variable at all
        st[k] = kFree;
        x[k] = 0.0;
        continue;
      }
      const bool to_bound = room < kInf && room <= t;
      step(k, dir, to_bound ? room : t, alpha, to_bound, leave, leave_up);
    }
    refactor();
    refresh();
    return ST_OPTIMAL;
  }

  int solve(std::chrono::steady_clock::time_point t0) {
    for (int attempt = 0; attempt < 6; ++attempt) {
      int rc = ST_OPTIMAL;
      if (infeasibility() > 0.0) {
        // phase 1 ends feasible (OPTIMAL) or with no column that reduces a
        // positive infeasibility (INFEASIBLE); fresh factors that show the
        // basics drifted out of their bounds send it round again
        rc = loop(true, t0);
        if (rc != ST_OPTIMAL) return rc;
        refactor();
        refresh();
        if (infeasibility() > 0.0) continue;
      }
      rc = loop(false, t0);
      if (rc != ST_OPTIMAL) return rc;
      if (has_super()) {          // to a basis, then the phases again
        rc = push(t0);
        if (rc != ST_OPTIMAL) return rc;
        continue;
      }
      if (perturbed) {            // the true costs: finish from this basis
        pert.clear();
        perturbed = false;
        rc = loop(false, t0);
        if (rc != ST_OPTIMAL) return rc;
      }
      refactor();
      refresh();
      if (infeasibility() == 0.0) return ST_OPTIMAL;   // else: drifted, go again
    }
    return ST_NUMERICAL;
  }
};

int check_params(const sovlp_simplex_params* p) {
  if (!p) return fail(1, "null simplex parameters");
  if (!(p->feas_tol > 0.0) || !(p->opt_tol > 0.0) || !(p->pivot_tol > 0.0))
    return fail(1, "simplex tolerances must be positive");
  if (p->refactor_freq < 1) return fail(1, "refactor_freq must be >= 1");
  return 0;
}

// The crossover's basis identification (crossover.py:80-188): writes the
// n + m statuses, z = [x ; A x] when asked, and returns the number of
// structurals placed in the basis.
int identify(const Lp& L, const double* xp, const double* yp, double pivot_tol,
             double tol_fixed, uint8_t* status, std::vector<double>* zout) {
  const int m = L.m, n = L.n, N = L.N;
  std::vector<double> z(N, 0.0);
  for (int j = 0; j < n; ++j) {
    z[j] = xp[j];
    for (int64_t p = L.cp[j]; p < L.cp[j + 1]; ++p) z[n + L.ci[p]] += L.cx[p] * xp[j];
  }
  // reduced costs over [A | -I] with zero cost on the logicals: c - A^T y
  // and +y (tie-break only)
  std::vector<double> dd(N, 0.0);
  if (yp) {
    for (int j = 0; j < n; ++j) dd[j] = L.cost[j] - L.dot(yp, j);
    for (int i = 0; i < m; ++i) dd[n + i] = yp[i];
  }
  std::vector<double> score(N);
  std::vector<char> fixed(N), free_(N);
  for (int k = 0; k < N; ++k) {
    const bool hl = L.has_lo(k), hu = L.has_up(k);
    fixed[k] = hl && hu && (L.up[k] - L.lo[k] <= tol_fixed * std::max(1.0, std::fabs(L.lo[k])));
    free_[k] = !hl && !hu;
    const double span = hl && hu ? L.up[k] - L.lo[k] : 0.0;
    const double scale = std::max(1.0, span > 0.0 ? span : std::fabs(z[k]));
    const double dl = hl ? z[k] - L.lo[k] : INFINITY;
    const double du = hu ? L.up[k] - z[k] : INFINITY;
    double interior = std::min(dl, du) / scale;
    if (!std::isfinite(interior)) interior = INFINITY;
    score[k] = interior - 1e-9 * std::fabs(dd[k]);
    if (fixed[k]) score[k] = -INFINITY;
    if (free_[k]) score[k] = INFINITY;
  }
  // candidates: free structurals first, then the structurals measurably off
  // their bounds by score, best first (ties in index order)
  std::vector<int> order;
  for (int j = 0; j < n; ++j)
    if (free_[j]) order.push_back(j);
  std::vector<int> live;
  for (int j = 0; j < n; ++j)
    if (score[j] > 1e-9 && std::isfinite(score[j])) live.push_back(j);
  std::stable_sort(live.begin(), live.end(), [&](int a, int b) { return score[a] > score[b]; });
  order.insert(order.end(), live.begin(), live.end());

  sovlp_simplex_params P{};
  P.feas_tol = 1e-9;
  P.opt_tol = 1e-9;
  P.pivot_tol = 1e-7;
  P.refactor_freq = 100;
  Simplex S(L, P);
  S.set_logical();
  S.refactor();
  std::vector<char> taken(m, 0);
  std::vector<double> alpha;
  const double ptol = std::max(pivot_tol, 1e-7);
  int got = 0;
  for (int j : order) {
    if (got == m) break;
    alpha.assign(m, 0.0);
    L.axpy(1.0, j, alpha.data());
    S.F.ftran(alpha, S.tmp);
    int r = -1;
    double best = 0.0;
    for (int i = 0; i < m; ++i)
      if (!taken[i] && std::fabs(alpha[i]) > best) {
        best = std::fabs(alpha[i]);
        r = i;
      }
    if (r < 0 || best <= ptol) continue;     // dependent on what is in
    S.F.update(r, alpha);
    S.head[r] = j;
    taken[r] = 1;
    ++got;
    if ((int)S.F.etas.size() >= 100) S.refactor();
  }
  for (int k = 0; k < N; ++k) {
    const bool hl = L.has_lo(k), hu = L.has_up(k);
    const bool at_lower = (std::fabs(z[k] - L.lo[k]) <= std::fabs(L.up[k] - z[k])) || !hu;
    status[k] = (at_lower && hl) ? kLower : (hu ? kUpper : kFree);
    if (fixed[k]) status[k] = kFixed;
  }
  for (int r = 0; r < m; ++r) status[S.head[r]] = kBasic;
  if (zout) zout->swap(z);
  return got;
}

// A simplex started from `status` (exactly m BASIC; statuses sanitised).
void start_from(Simplex& S, const uint8_t* status) {
  const Lp& L = S.L;
  S.head.clear();
  S.st.assign(L.N, kLower);
  for (int k = 0; k < L.N; ++k) {
    if (status[k] == kBasic) {
      S.head.push_back(k);
      S.st[k] = kBasic;
    } else {
      S.st[k] = S.sane_status(k, status[k]);
    }
  }
}

// x, duals and reduced costs of the true costs, statuses and counters.
void report(Simplex& S, int rc, int64_t superbasics, std::chrono::steady_clock::time_point t0,
            double* x, double* y, double* d, uint8_t* status, sovlp_simplex_result* res) {
  const Lp& L = S.L;
  S.pert.clear();
  for (int k = 0; k < L.N; ++k)            // only after a limit: none held between bounds
    if (S.st[k] == kSuper) {
      const bool lo = !L.has_up(k) || (L.has_lo(k) && S.x[k] - L.lo[k] <= L.up[k] - S.x[k]);
      S.st[k] = S.sane_status(k, lo ? kLower : kUpper);
    }
  S.duals(false);
  double obj = 0.0, dinf = 0.0;
  for (int j = 0; j < L.n; ++j) {
    x[j] = S.x[j];
    obj += L.cost[j] * S.x[j];
    d[j] = L.cost[j] - L.dot(S.w1.data(), j);
    if (S.st[j] == kLower) dinf = std::max(dinf, -d[j]);
    else if (S.st[j] == kUpper) dinf = std::max(dinf, d[j]);
    else if (S.st[j] == kFree) dinf = std::max(dinf, std::fabs(d[j]));
  }
  for (int i = 0; i < L.m; ++i) y[i] = S.w1[i];
  for (int k = 0; k < L.N; ++k) status[k] = S.st[k];
  res->status = rc;
  res->repairs = S.repairs;
  res->iterations = S.iters;
  res->pivots = S.pivots;
  res->phase1_iterations = S.phase1;
  res->superbasics = superbasics;
  res->objective = obj;
  res->primal_infeasibility = S.infeasibility();
  res->dual_infeasibility = dinf;
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

int sovlp_simplex_solve(const sovlp_host_lp* lp, const sovlp_simplex_params* params,
                        const uint8_t* warm_status, double* x, double* y, double* d,
                        uint8_t* status, sovlp_simplex_result* res) {
  const auto t0 = std::chrono::steady_clock::now();
  Lp L;
  if (int rc = load(lp, L)) return rc;
  if (int rc = check_params(params)) return rc;
  if (!x || !y || !d || !status || !res) return fail(1, "null output");
  Simplex S(L, *params);
  bool warm = false;
  if (warm_status) {
    int nb = 0;
    for (int k = 0; k < L.N; ++k) {
      if (warm_status[k] > kFixed) return fail(1, "basis status out of range");
      nb += warm_status[k] == kBasic;
    }
    if (nb == L.m) {
      warm = true;
      start_from(S, warm_status);
    }
  }
  if (!warm) {
    S.set_logical();
    S.status_from_costs();
  }
  S.refactor();
  S.refresh();
  report(S, S.solve(t0), 0, t0, x, y, d, status, res);
  return 0;
}

int sovlp_basis_from_point(const sovlp_host_lp* lp, const double* xp, const double* yp,
                           double pivot_tol, double tol_fixed, uint8_t* status, int32_t* placed) {
  Lp L;
  if (int rc = load(lp, L)) return rc;
  if (!xp || !status) return fail(1, "null argument");
  const int got = identify(L, xp, yp, pivot_tol, tol_fixed, status, nullptr);
  if (placed) *placed = got;
  return 0;
}

int sovlp_crossover(const sovlp_host_lp* lp, const sovlp_simplex_params* params,
                    const double* x0, const double* y0, double* x, double* y, double* d,
                    uint8_t* status, sovlp_simplex_result* res) {
  const auto t0 = std::chrono::steady_clock::now();
  Lp L;
  if (int rc = load(lp, L)) return rc;
  if (int rc = check_params(params)) return rc;
  if (!x0 || !x || !y || !d || !status || !res) return fail(1, "null argument");
  std::vector<uint8_t> st0(L.N);
  std::vector<double> z;
  identify(L, x0, y0, params->pivot_tol, 1e-12, st0.data(), &z);
  Simplex S(L, *params);
  start_from(S, st0.data());
  // nonbasic variables strictly between their bounds stay at the point's
  // value; the simplex pushes them to bounds itself
  S.x.assign(L.N, 0.0);
  int64_t held = 0;
  for (int k = 0; k < L.N; ++k) {
    const uint8_t s = S.st[k];
    if (s == kBasic || s == kFixed) continue;
    const double v = std::min(std::max(z[k], L.lo[k]), L.up[k]);
    const double tol = S.ftol(k);
    const bool off_lo = !L.has_lo(k) || v > L.lo[k] + tol;
    const bool off_up = !L.has_up(k) || v < L.up[k] - tol;
    if (off_lo && off_up && (s != kFree || v != 0.0)) {
      S.st[k] = kSuper;
      S.x[k] = v;
      ++held;
    }
  }
  S.refactor();
  S.refresh();
  report(S, S.solve(t0), held, t0, x, y, d, status, res);
  return 0;
}

int sovlp_simplex_ranging(const sovlp_host_lp* lp, const uint8_t* status, double* value,
                          double* cost_lo, double* cost_hi, double* rhs, double* rhs_lo,
                          double* rhs_hi, double* activity, double* dual, uint8_t* binding) {
  Lp L;
  if (int rc = load(lp, L)) return rc;
  if (!status || !value || !cost_lo || !cost_hi || !rhs || !rhs_lo || !rhs_hi || !activity ||
      !dual || !binding)
    return fail(1, "null argument");
  const int m = L.m, n = L.n, N = L.N;
  int nb = 0;
  for (int k = 0; k < N; ++k) {
    if (status[k] > kFixed) return fail(1, "basis status out of range");
    nb += status[k] == kBasic;
  }
  if (nb != m) return fail(1, "the status vector must hold exactly m basic entries");
  sovlp_simplex_params P{};
  P.feas_tol = 1e-9;
  P.opt_tol = 1e-9;
  P.pivot_tol = 1e-7;
  P.refactor_freq = 1 << 30;
  Simplex S(L, P);
  start_from(S, status);
  S.refactor();
  if (S.repairs) return fail(1, "the basis is singular");
  S.refresh();                            // z: nonbasics from statuses, x_B = -B^-1 N x_N
  S.duals(false);                         // y = B^-T c_B
  const std::vector<double> y(S.w1);
  std::vector<double> d(N);
  for (int k = 0; k < N; ++k) d[k] = L.cost[k] - L.dot(y.data(), k);
  const std::vector<uint8_t>& st = S.st;
  // cost ranging (sensitivity.py:154-185, 236-255): a nonbasic cost moves its
  // own reduced cost one for one; a basic one moves every nonbasic reduced
  // cost by -delta alpha_r[k], alpha_r the tableau row of its position
  std::vector<double> rho, tmp;
  for (int j = 0; j < n; ++j) {
    value[j] = S.x[j];
    cost_lo[j] = -INFINITY;
    cost_hi[j] = INFINITY;
    if (st[j] == kLower) cost_lo[j] = L.cost[j] - d[j];
    else if (st[j] == kUpper) cost_hi[j] = L.cost[j] - d[j];
  }
  for (int r = 0; r < m; ++r) {
    const int j = S.head[r];
    if (j >= n) continue;                 // a logical carries no cost
    rho.assign(m, 0.0);
    rho[r] = 1.0;
    S.F.btran(rho, tmp);
    double lo = -INFINITY, hi = INFINITY;
    for (int k = 0; k < N; ++k) {
      const uint8_t s = st[k];
      if (s == kBasic || s == kFixed) continue;
      const double a = L.dot(rho.data(), k);
      if (std::fabs(a) < 1e-12) continue;
      const double ratio = d[k] / a;
      const bool keep_nonneg = s == kLower || (s == kFree && d[k] >= 0.0);
      if (keep_nonneg == (a > 0.0)) hi = std::min(hi, ratio);
      else lo = std::max(lo, ratio);
    }
    cost_lo[j] = L.cost[j] + lo;
    cost_hi[j] = L.cost[j] + hi;
  }
  // right-hand-side ranging (sensitivity.py:187-212, 257-279): a binding row's
  // bound moves its logical, and with it x_B along beta = B^-1 e_i
  std::vector<double> beta;
  for (int i = 0; i < m; ++i) {
    const int k = n + i;
    const double lo_i = L.lo[k], hi_i = L.up[k];
    activity[i] = S.x[k];
    dual[i] = y[i];
    if (st[k] == kBasic) {
      binding[i] = 0;
      rhs[i] = lo_i > -kInf ? lo_i : hi_i;
      rhs_lo[i] = -INFINITY;
      rhs_hi[i] = S.x[k];
      if (hi_i < kInf && lo_i <= -kInf) {
        rhs_lo[i] = S.x[k];
        rhs_hi[i] = INFINITY;
      }
      continue;
    }
    binding[i] = 1;
    const double b = (st[k] == kLower || st[k] == kFixed) ? lo_i : hi_i;
    rhs[i] = b;
    beta.assign(m, 0.0);
    beta[i] = 1.0;
    S.F.ftran(beta, tmp);
    double lo = -INFINITY, hi = INFINITY;
    for (int r = 0; r < m; ++r) {
      const double a = beta[r];
      if (std::fabs(a) < 1e-12) continue;
      const int q = S.head[r];
      const double zb = S.x[q], l = L.lo[q], u = L.up[q];
      if (a > 0.0) {
        if (l > -kInf) lo = std::max(lo, (l - zb) / a);
        if (u < kInf) hi = std::min(hi, (u - zb) / a);
      } else {
        if (u < kInf) lo = std::max(lo, (u - zb) / a);
        if (l > -kInf) hi = std::min(hi, (l - zb) / a);
      }
    }
    rhs_lo[i] = b + lo;
    rhs_hi[i] = b + hi;
  }
  return 0;
}

int sovlp_basis_repairs(const sovlp_host_lp* lp, const uint8_t* status, int32_t* repairs) {
  Lp L;
  if (int rc = load(lp, L)) return rc;
  if (!status || !repairs) return fail(1, "null argument");
  std::vector<int> head;
  for (int k = 0; k < L.N; ++k)
    if (status[k] == kBasic) head.push_back(k);
  if ((int)head.size() != L.m) return fail(1, "the status vector must hold exactly m basic entries");
  Factor F;
  *repairs = (int)F.factorize(L, head).size();
  return 0;
}

}  // extern "C"
