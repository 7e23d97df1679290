/* certify_oracle.c -- CPU restatement of the reference's certificate path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ (through oracle/oracle.py) as the
 * checker of the device certificate; never linked into the sovlp library.
 *
 * Restates, in the reference's own operation order:
 *   oracle_residual   numerics/refine.py:48-84 (_residual_csr) and 103-110
 *                     (compensated_residual): out = b - A x per row, Dekker's
 *                     two_product with the 2^27+1 split and Knuth's two_sum,
 *                     the row's nonzeros in CSR order.
 *   oracle_certify    mip/safebound.py:252-307 (certified_bound) over
 *                     safe_dual_bound(strict=True) (:102-143), _term (:93-99)
 *                     and _neumaier_sum (:175-186); A^T y walks the CSC as
 *                     core/sparse.py:61-69 (_csc_spmv_t) does; plus the
 *                     feasibility lines of bench/verify.py:208-224 (column
 *                     bounds, compensated row activities) and the objective.
 * Two sums differ from NumPy's: |terms| is summed in index order here where
 * NumPy sums pairwise (it sizes the 4 eps allowance only), and c'x is a
 * compensated sum where NumPy's dot is BLAS order.  Built with
 * -ffp-contract=off so no product is fused into an addition.
 */
#include <math.h>
#include <stdint.h>

#define SOV_INF 1.0e30            /* core/tolerances.py:16 */
#define SPLIT 134217729.0          /* 2^27 + 1 (refine.py:46) */

typedef struct {
  double bound;          /* in the model's sense, offset included; -inf/+inf when vacuous */
  double perturbation;
  double objective;      /* c'x + offset */
  double col_violation;
  double row_violation;
  int32_t vacuous;
  int32_t pad;
} oracle_certificate;

/* refine.py:55-84 for one row: returns b - (s + comp) as the reference rounds it */
static double residual_row(const int64_t* rp, const int32_t* ri, const double* rx,
                           const double* x, double bi, int64_t i) {
  double s = 0.0, comp = 0.0;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
    const double a = rx[p], xv = x[ri[p]];
    const double prod = a * xv;
    const double ca = SPLIT * a;
    const double ah = ca - (ca - a);
    const double al = a - ah;
    const double cb = SPLIT * xv;
    const double bh = cb - (cb - xv);
    const double bl = xv - bh;
    const double err = ((ah * bh - prod) + ah * bl + al * bh) + al * bl;
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

void oracle_residual(int m, const int64_t* rp, const int32_t* ri, const double* rx,
                     const double* x, const double* b, double* out) {
  for (int64_t i = 0; i < m; ++i) out[i] = residual_row(rp, ri, rx, x, b[i], i);
}

/* safebound.py:175-186 */
static void neumaier(const double* a, int64_t n, double* sum, double* comp) {
  double s = 0.0, c = 0.0;
  for (int64_t k = 0; k < n; ++k) {
    const double v = a[k];
    const double t = s + v;
    if (fabs(s) >= fabs(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
  }
  *sum = s;
  *comp = c;
}

/* safebound.py:93-99: coef times the bound it points at; -inf at an infinite one */
static double term(double coef, double lo, double hi) {
  const int pos = coef > 0.0;
  const double chosen = pos ? lo : hi;
  const int bad = (pos ? lo <= -SOV_INF : hi >= SOV_INF) && coef != 0.0;
  return bad ? -INFINITY : coef * chosen;
}

/* `rterm` (m) and `cterm` (n) are scratch; `d_out` (n), `act_out` (m) receive
 * the reduced costs c - A^T yy (minimisation sense) and the row activities. */
void oracle_certify(int m, int n, const int64_t* rp, const int32_t* ri, const double* rx,
                    const int64_t* cp, const int32_t* ci, const double* cx, const double* c,
                    const double* rl, const double* ru, const double* lo, const double* hi,
                    double offset, int maximise, const double* x, const double* y,
                    double dual_tol, double* yy, double* d_out, double* act_out,
                    double* rterm, double* cterm, oracle_certificate* out) {
  const double sg = maximise ? -1.0 : 1.0;       /* safebound.py:283-291 */
  const double off = sg * offset;
  /* bench/verify.py:208-224: column bounds, compensated row activities */
  double colv = 0.0, rowv = 0.0;
  for (int64_t j = 0; j < n; ++j) {
    colv = fmax(colv, fmax(lo[j] - x[j], 0.0));
    colv = fmax(colv, fmax(x[j] - hi[j], 0.0));
  }
  for (int64_t i = 0; i < m; ++i) {
    const double act = -residual_row(rp, ri, rx, x, 0.0, i);
    act_out[i] = act;
    rowv = fmax(rowv, fmax(rl[i] - act, 0.0));
    rowv = fmax(rowv, fmax(act - ru[i], 0.0));
  }
  out->col_violation = colv;
  out->row_violation = rowv;
  /* objective, compensated (verify.py:227) */
  {
    double s = 0.0, cc = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      const double p = c[j] * x[j];
      const double err = fma(c[j], x[j], -p);
      const double t = s + p;
      const double bb = t - s;
      cc += ((s - (t - bb)) + (p - bb)) + err;
      s = t;
    }
    out->objective = (s + cc) + offset;
  }
  /* safebound.py:293-295 then 116-117: a dual pointing at an infinite row bound is 0 */
  for (int64_t i = 0; i < m; ++i) {
    double v = sg * y[i];
    if (v > 0.0 && rl[i] <= -SOV_INF) v = 0.0;
    if (v < 0.0 && ru[i] >= SOV_INF) v = 0.0;
    yy[i] = v;
  }
  /* d = g - A^T yy (safebound.py:296-297), A^T yy in CSC order (sparse.py:61-69) */
  double cmax = 0.0, pert = 0.0;
  int anybad = 0;
  for (int64_t j = 0; j < n; ++j) {
    double acc = 0.0;
    for (int64_t p = cp[j]; p < cp[j + 1]; ++p) acc += cx[p] * yy[ci[p]];
    const double d = sg * c[j] - acc;
    d_out[j] = d;
    cmax = fmax(cmax, fabs(c[j]));
    const int bad = (d > 0.0 && lo[j] <= -SOV_INF) || (d < 0.0 && hi[j] >= SOV_INF);
    if (bad) {
      anybad = 1;
      pert = fmax(pert, fabs(d));
    }
  }
  /* safebound.py:299-306: absorb small ones, else the bound is vacuous */
  const int absorb = anybad && pert <= dual_tol * (1.0 + cmax);
  if (!anybad) pert = 0.0;
  out->perturbation = pert;
  /* safe_dual_bound, strict (safebound.py:125-143) */
  int finite = 1;
  double mag = 0.0;
  for (int64_t i = 0; i < m; ++i) {
    rterm[i] = term(yy[i], rl[i], ru[i]);
    if (!isfinite(rterm[i])) finite = 0;
    mag += fabs(rterm[i]);
  }
  for (int64_t j = 0; j < n; ++j) {
    double d = d_out[j];
    if (absorb && ((d > 0.0 && lo[j] <= -SOV_INF) || (d < 0.0 && hi[j] >= SOV_INF))) d = 0.0;
    cterm[j] = term(d, lo[j], hi[j]);
    if (!isfinite(cterm[j])) finite = 0;
    mag += fabs(cterm[j]);
  }
  double b;
  if (!finite) {
    b = -INFINITY;
  } else {
    double t1, c1, t2, c2;
    neumaier(rterm, m, &t1, &c1);
    neumaier(cterm, n, &t2, &c2);
    const double value = (t1 + t2) + (c1 + c2);
    b = value - 4.0 * 2.220446049250313e-16 * mag;
  }
  b = b + off;
  out->vacuous = !finite;
  out->bound = maximise ? -b : b;
  out->pad = 0;
}
