/*
 * pdlp_oracle.c -- CPU restatement of the reference PDLP path, FP64.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker and the CPU baseline; no
 * product code links, loads or calls it. Only tests/, the smoke entry point and
 * the cpu_baseline leg of bench.py use it (through oracle/oracle.py).
 *
 * It restates, in plain C with OpenMP row-parallel sparse products:
 *   - scale_problem(prob, "pdlp")       src/sovopt/numerics/scaling.py:374-478
 *       _ruiz (10 sweeps, tol 1e-2)     scaling.py:149-170
 *       _pock_chambolle (alpha = 1)     scaling.py:274-295
 *       _pdlp                           scaling.py:298-309
 *       _pow2_round                     scaling.py:119-131
 *       objective scale                 scaling.py:422-444
 *       INF-preserving bound scaling    scaling.py:446-456
 *   - solve_pdlp                        src/sovopt/lp/pdlp.py:238-494
 *       _estimate_norm                  pdlp.py:140-159
 *       _KKT.__call__ / merit           pdlp.py:191-225
 *       adaptive step (Malitsky-Pock)   pdlp.py:337-389
 *       fused step                      pdlp.py:390-399 (backend.py:441-479)
 *       check / restart / weight        pdlp.py:405-463
 *       _unscaled_kkt, _converged       pdlp.py:497-530
 *       unscale and report              pdlp.py:465-494
 *
 * Every sparse product walks one row (CSR) or one column (CSC) sequentially in
 * stored order, exactly as the reference's _csr_spmv / bincount paths do, so the
 * products are bitwise equal to the reference's whatever the thread count.
 * Vector reductions are sequential left-to-right sums (the reference uses BLAS
 * ddot and NumPy pairwise sums; the difference is rounding-level).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_INF 1.0e30

typedef struct {
    double eps_abs, eps_rel, feas_cap;
    long long max_iter;
    double time_limit;
    int check_every, stepsize_every;
    double beta_sufficient, beta_necessary, beta_artificial;
    double primal_weight_smoothing;
} oracle_params;

typedef struct {
    int status;               /* Status codes of problem.py:58-69 */
    long long iterations, restarts, rejected_steps;
    double primal_res, dual_res, gap, primal_obj, dual_obj;
    double step_size, primal_weight, objective;
    double scale_obj;
    long long history_len;
} oracle_info;

enum { ST_OPTIMAL = 0, ST_ITERATION_LIMIT = 4, ST_TIME_LIMIT = 5, ST_NUMERICAL = 8 };

typedef struct {
    int rows;
    const int64_t *p;
    const int32_t *i;
    const double *x;
} csr_t;

static double now_s(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* out = M v  (M walked one row at a time, stored order) */
static void spmv(csr_t M, const double *v, double *out) {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < M.rows; ++r) {
        double acc = 0.0;
        for (int64_t q = M.p[r]; q < M.p[r + 1]; ++q) acc += M.x[q] * v[M.i[q]];
        out[r] = acc;
    }
}

static double dot(const double *a, const double *b, long n) {
    double s = 0.0;
    for (long k = 0; k < n; ++k) s += a[k] * b[k];
    return s;
}

static double nrm2(const double *a, long n) { return sqrt(dot(a, a, n)); }

static double clampd(double v, double l, double u) { return v < l ? l : (v > u ? u : v); }

static double pow2_round(double s) {
    if (!(isfinite(s) && s > 0.0)) return 1.0;
    return exp2(rint(log2(s)));
}

/* ------------------------------------------------------------------------ */
/* scaling                                                                   */
/* ------------------------------------------------------------------------ */

/* W <- diag(dr) W diag(dc) on both storage orders: every value multiplied by
 * the rounded product dr_i * dc_j, as SparseMatrix.scale does. */
static void scale_both(int m, int n, const int64_t *cp, const int32_t *ci, double *cx,
                       const int64_t *rp, const int32_t *ri, double *rx,
                       const double *dr, const double *dc) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int64_t q = cp[j]; q < cp[j + 1]; ++q) cx[q] *= dr[ci[q]] * dc[j];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m; ++i)
        for (int64_t q = rp[i]; q < rp[i + 1]; ++q) rx[q] *= dr[i] * dc[ri[q]];
}

static void absmax_by_slice(int s, const int64_t *p, const double *x, double *out) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < s; ++k) {
        double hi = 0.0;
        for (int64_t q = p[k]; q < p[k + 1]; ++q) {
            double v = fabs(x[q]);
            if (v > hi) hi = v;
        }
        out[k] = hi;
    }
}

static void abssum_by_slice(int s, const int64_t *p, const double *x, double *out) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < s; ++k) {
        double acc = 0.0;
        for (int64_t q = p[k]; q < p[k + 1]; ++q) acc += fabs(x[q]);
        out[k] = acc;
    }
}

/* Computes (row, col, obj) factors of scale_problem(prob, "pdlp"). */
int oracle_scale(int m, int n, const int64_t *cp, const int32_t *ci, const double *cx,
                 const int64_t *rp, const int32_t *ri, const double *rx, const double *c,
                 double *row_out, double *col_out, double *obj_out, int *ruiz_sweeps_out) {
    int64_t nnz = cp[n];
    double *wc = malloc(sizeof(double) * (nnz ? nnz : 1));
    double *wr = malloc(sizeof(double) * (nnz ? nnz : 1));
    double *rmax = malloc(sizeof(double) * (m ? m : 1));
    double *cmax = malloc(sizeof(double) * (n ? n : 1));
    double *r = malloc(sizeof(double) * (m ? m : 1));
    double *cc = malloc(sizeof(double) * (n ? n : 1));
    if (!wc || !wr || !rmax || !cmax || !r || !cc) return -1;
    memcpy(wc, cx, sizeof(double) * nnz);
    memcpy(wr, rx, sizeof(double) * nnz);
    for (int i = 0; i < m; ++i) r[i] = 1.0;
    for (int j = 0; j < n; ++j) cc[j] = 1.0;

    /* _ruiz(A, iters=10, tol=1e-2) */
    int sweeps = 0;
    for (int it = 0; it < 10; ++it) {
        absmax_by_slice(m, rp, wr, rmax);
        absmax_by_slice(n, cp, wc, cmax);
        double dev = 0.0;
        for (int i = 0; i < m; ++i) {
            if (rmax[i] <= 0.0) rmax[i] = 1.0;
            double d = fabs(1.0 - rmax[i]);
            if (d > dev) dev = d;
        }
        double devc = 0.0;
        for (int j = 0; j < n; ++j) {
            if (cmax[j] <= 0.0) cmax[j] = 1.0;
            double d = fabs(1.0 - cmax[j]);
            if (d > devc) devc = d;
        }
        if (devc > dev) dev = devc;
        for (int i = 0; i < m; ++i) rmax[i] = 1.0 / sqrt(rmax[i]);
        for (int j = 0; j < n; ++j) cmax[j] = 1.0 / sqrt(cmax[j]);
        scale_both(m, n, cp, ci, wc, rp, ri, wr, rmax, cmax);
        for (int i = 0; i < m; ++i) r[i] *= rmax[i];
        for (int j = 0; j < n; ++j) cc[j] *= cmax[j];
        ++sweeps;
        if (dev < 1e-2) break;
    }
    /* _pdlp: W = A.copy(); W.scale(r, c); Pock-Chambolle alpha = 1 */
    memcpy(wc, cx, sizeof(double) * nnz);
    memcpy(wr, rx, sizeof(double) * nnz);
    scale_both(m, n, cp, ci, wc, rp, ri, wr, r, cc);
    abssum_by_slice(m, rp, wr, rmax);
    abssum_by_slice(n, cp, wc, cmax);
    for (int i = 0; i < m; ++i) {
        double s = rmax[i] <= 0.0 ? 1.0 : rmax[i];
        row_out[i] = pow2_round(r[i] * (1.0 / sqrt(s)));
    }
    for (int j = 0; j < n; ++j) {
        double s = cmax[j] <= 0.0 ? 1.0 : cmax[j];
        col_out[j] = pow2_round(cc[j] * (1.0 / sqrt(s)));
    }
    /* objective scale on c' = c * col */
    double cmx = 0.0, cmn = INFINITY;
    int any = 0;
    for (int j = 0; j < n; ++j) {
        double v = c[j] * col_out[j];
        if (v != 0.0) {
            double a = fabs(v);
            if (a > cmx) cmx = a;
            if (a < cmn) cmn = a;
            any = 1;
        }
    }
    double obj = 1.0;
    if (any) {
        double g = exp2(rint(log2(sqrt(cmx * cmn))));
        if (g > 0.0 && isfinite(g)) obj = 1.0 / g;
    }
    *obj_out = obj;
    if (ruiz_sweeps_out) *ruiz_sweeps_out = sweeps;
    free(wc); free(wr); free(rmax); free(cmax); free(r); free(cc);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* KKT                                                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    int m, n;
    const double *c, *rl, *ru, *lo, *hi;
    double offset;
} lpvec_t;

/* _KKT.__call__ (pdlp.py:191-220) on already computed ax, aty. */
static void kkt_eval(const lpvec_t *L, const double *x, const double *y, const double *ax,
                     const double *aty, double out[5]) {
    double pres = 0.0, dres = 0.0, cx = 0.0, supp = 0.0, bnd = 0.0;
    for (int i = 0; i < L->m; ++i) {
        double a = ax[i];
        double pv = fmax(L->rl[i] - a, 0.0) + fmax(a - L->ru[i], 0.0);
        if (fabs(pv) > pres) pres = fabs(pv);
        double yi = y[i];
        double bsel = yi > 0.0 ? (L->ru[i] < ORACLE_INF ? L->ru[i] : 0.0)
                               : (L->rl[i] > -ORACLE_INF ? L->rl[i] : 0.0);
        supp += bsel * yi;
    }
    for (int j = 0; j < L->n; ++j) {
        double d = L->c[j] + aty[j];
        int lof = L->lo[j] > -ORACLE_INF, hif = L->hi[j] < ORACLE_INF;
        double dr = (lof ? 0.0 : fmax(d, 0.0)) + (hif ? 0.0 : fmax(-d, 0.0));
        if (fabs(dr) > dres) dres = fabs(dr);
        cx += L->c[j] * x[j];
        double bsel = d > 0.0 ? (lof ? L->lo[j] : 0.0) : (hif ? L->hi[j] : 0.0);
        bnd += bsel * d;
    }
    double pobj = cx + L->offset;
    double dobj = (-supp + bnd) + L->offset;
    out[0] = pres; out[1] = dres; out[2] = fabs(pobj - dobj); out[3] = pobj; out[4] = dobj;
}

static double merit_of(const double k[5], double omega) {
    double a = omega * k[0], b = k[1] / omega, g = k[2];
    return sqrt(a * a + b * b + g * g);
}

/* ------------------------------------------------------------------------ */
/* solve                                                                     */
/* ------------------------------------------------------------------------ */

int oracle_pdlp_solve(int m, int n,
                      const int64_t *cp, const int32_t *ci, const double *cx,
                      const int64_t *rp, const int32_t *ri, const double *rx,
                      const double *c_in, const double *rl_in, const double *ru_in,
                      const double *lo_in, const double *hi_in, double offset_in, int maximise,
                      const double *v0, const oracle_params *P,
                      double *x_out, double *y_out, double *d_out, oracle_info *info,
                      double *hist, long long hist_cap, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    double t0 = now_s();
    int64_t nnz = cp[n];
    size_t M = (size_t)(m ? m : 1), N = (size_t)(n ? n : 1), Z = (size_t)(nnz ? nnz : 1);

    /* work problem: objective flipped for maximisation (pdlp.py:244-250) */
    double *wc = malloc(sizeof(double) * N);
    for (int j = 0; j < n; ++j) wc[j] = maximise ? -c_in[j] : c_in[j];
    double woff = maximise ? -offset_in : offset_in;

    double *row = malloc(sizeof(double) * M), *col = malloc(sizeof(double) * N), sobj;
    if (oracle_scale(m, n, cp, ci, cx, rp, ri, rx, wc, row, col, &sobj, NULL)) return -1;
    info->scale_obj = sobj;

    /* scaled copy: A' = R A C, c' = s C c, rl' = R rl, lo' = lo / C */
    double *sx_c = malloc(sizeof(double) * Z), *sx_r = malloc(sizeof(double) * Z);
    memcpy(sx_c, cx, sizeof(double) * nnz);
    memcpy(sx_r, rx, sizeof(double) * nnz);
    scale_both(m, n, cp, ci, sx_c, rp, ri, sx_r, row, col);
    double *c = malloc(sizeof(double) * N), *lo = malloc(sizeof(double) * N),
           *hi = malloc(sizeof(double) * N);
    double *rl = malloc(sizeof(double) * M), *ru = malloc(sizeof(double) * M);
    for (int j = 0; j < n; ++j) {
        c[j] = (wc[j] * col[j]) * sobj;
        lo[j] = lo_in[j] <= -ORACLE_INF ? -ORACLE_INF : (lo_in[j] >= ORACLE_INF ? ORACLE_INF : lo_in[j] / col[j]);
        hi[j] = hi_in[j] <= -ORACLE_INF ? -ORACLE_INF : (hi_in[j] >= ORACLE_INF ? ORACLE_INF : hi_in[j] / col[j]);
    }
    for (int i = 0; i < m; ++i) {
        rl[i] = rl_in[i] <= -ORACLE_INF ? -ORACLE_INF : (rl_in[i] >= ORACLE_INF ? ORACLE_INF : rl_in[i] * row[i]);
        ru[i] = ru_in[i] <= -ORACLE_INF ? -ORACLE_INF : (ru_in[i] >= ORACLE_INF ? ORACLE_INF : ru_in[i] * row[i]);
    }
    double soff = woff * sobj;
    lpvec_t L = {m, n, c, rl, ru, lo, hi, soff};
    csr_t A = {m, rp, ri, sx_r}, AT = {n, cp, ci, sx_c};
    csr_t A0 = {m, rp, ri, rx}, AT0 = {n, cp, ci, cx};

    /* the original-problem constants of _converged (pdlp.py:523-530) */
    double rhs = 0.0, cn_inf = 0.0;
    for (int i = 0; i < m; ++i) {
        double a = rl_in[i] > -ORACLE_INF ? fabs(rl_in[i]) : 0.0;
        double b = ru_in[i] < ORACLE_INF ? fabs(ru_in[i]) : 0.0;
        if (a > rhs) rhs = a;
        if (b > rhs) rhs = b;
    }
    for (int j = 0; j < n; ++j) if (fabs(c_in[j]) > cn_inf) cn_inf = fabs(c_in[j]);
    double ptol = fmin(P->eps_abs + P->eps_rel * rhs, P->feas_cap);
    double dtol = fmin(P->eps_abs + P->eps_rel * cn_inf, P->feas_cap);

#define NEWV(nm, len) double *nm = calloc((len), sizeof(double))
    NEWV(x, N); NEWV(y, M); NEWV(ax, M); NEWV(aty, N);
    NEWV(xn, N); NEWV(yn, M); NEWV(adx, M); NEWV(ady, N); NEWV(dx, N); NEWV(dy, M);
    NEWV(sum_x, N); NEWV(sum_y, M); NEWV(xr, N); NEWV(yr, M);
    NEWV(avg_x, N); NEWV(avg_y, M); NEWV(tax, M); NEWV(taty, N);
    NEWV(ux, N); NEWV(uy, M);

    for (int j = 0; j < n; ++j) x[j] = fmin(fmax(0.0, lo[j]), hi[j]);
    spmv(A, x, ax);
    spmv(AT, y, aty);

    /* _estimate_norm */
    double anorm = 1.0;
    {
        double *v = malloc(sizeof(double) * N), *av = malloc(sizeof(double) * M),
               *atv = malloc(sizeof(double) * N);
        memcpy(v, v0, sizeof(double) * n);
        double nv = nrm2(v, n);
        if (nv != 0.0) {
            for (int j = 0; j < n; ++j) v[j] /= nv;
            double est = 1.0;
            int degenerate = 0;
            for (int it = 0; it < 20; ++it) {
                spmv(A, v, av);
                spmv(AT, av, atv);
                double nrm = nrm2(atv, n);
                if (nrm <= 0.0) { degenerate = 1; break; }
                for (int j = 0; j < n; ++j) v[j] = atv[j] / nrm;
                est = sqrt(nrm);
            }
            anorm = degenerate ? 1.0 : fmax(est, 1e-12);
        }
        free(v); free(av); free(atv);
    }
    double eta = 1.0 / anorm;

    double cn = nrm2(c, n), bn;
    {
        double s = 0.0;
        for (int i = 0; i < m; ++i) {
            double b = (rl[i] > -ORACLE_INF ? rl[i] : 0.0) + (ru[i] < ORACLE_INF ? ru[i] : 0.0);
            s += b * b;
        }
        bn = sqrt(s);
    }
    double omega = bn > 0.0 ? exp(log(fmax(cn, 1e-12)) - log(fmax(bn, 1e-12))) : 1.0;
    omega = fmin(fmax(omega, 1e-4), 1e4);

    double sum_eta = 0.0;
    memcpy(xr, x, sizeof(double) * n);
    memcpy(yr, y, sizeof(double) * m);
    double kv[5];
    kkt_eval(&L, x, y, ax, aty, kv);
    double mu_restart = merit_of(kv, omega), mu_last = INFINITY;
    long long epoch = 0, k = 0, attempts = 0, restarts = 0, rejected = 0, hlen = 0;
    int status = ST_ITERATION_LIMIT;
    double fin_p = NAN, fin_d = NAN, fin_g = NAN, fin_po = NAN, fin_do = NAN;
    const int CE = P->check_every, SE = P->stepsize_every;

    while (k < P->max_iter) {
        if (k % CE == 0 && (now_s() - t0) > P->time_limit) { status = ST_TIME_LIMIT; break; }
        if (k % SE == 0) {
            int accepted = 0;
            for (int trial = 0; trial < 60; ++trial) {
                ++attempts;
                double tau = eta / omega, sigma = eta * omega;
                for (int j = 0; j < n; ++j) {
                    double v = x[j] - tau * (c[j] + aty[j]);
                    xn[j] = v < lo[j] ? lo[j] : (v > hi[j] ? hi[j] : v);
                    dx[j] = xn[j] - x[j];
                }
                spmv(A, dx, adx);
                for (int i = 0; i < m; ++i) {
                    double w = y[i] + sigma * (ax[i] + 2.0 * adx[i]);
                    double pr = clampd(w / sigma, rl[i], ru[i]);
                    yn[i] = w - sigma * pr;
                    dy[i] = yn[i] - y[i];
                }
                double interaction = fabs(dot(dy, adx, m));
                double movement = 0.5 * (omega * dot(dx, dx, n) + dot(dy, dy, m) / omega);
                if (movement == 0.0) { accepted = 1; break; }
                double lim = interaction > 0.0 ? movement / interaction : INFINITY;
                double t = (double)(attempts + 1);
                double nxt = fmax(fmin((1.0 - pow(t, -0.3)) * lim, (1.0 + pow(t, -0.6)) * eta), 1e-14);
                if (eta <= lim) { accepted = 1; eta = nxt; break; }
                ++rejected;
                eta = nxt;
            }
            if (!accepted) { status = ST_NUMERICAL; break; }
            spmv(AT, dy, ady);
            double *t1 = x; x = xn; xn = t1;
            double *t2 = y; y = yn; yn = t2;
            for (int i = 0; i < m; ++i) ax[i] += adx[i];
            for (int j = 0; j < n; ++j) aty[j] += ady[j];
            for (int j = 0; j < n; ++j) sum_x[j] += eta * x[j];
            for (int i = 0; i < m; ++i) sum_y[i] += eta * y[i];
        } else {
            double tau = eta / omega, sigma = eta * omega;
            for (int j = 0; j < n; ++j) {
                double xi = x[j];
                double v = xi - tau * (c[j] + aty[j]);
                v = v < lo[j] ? lo[j] : (v > hi[j] ? hi[j] : v);
                dx[j] = v - xi;
                x[j] = v;
            }
            spmv(A, dx, adx);
            for (int i = 0; i < m; ++i) {
                double a = adx[i];
                double axn = ax[i] + a;
                ax[i] = axn;
                double yo = y[i];
                double w = yo + sigma * (axn + a);
                double pr = clampd(w / sigma, rl[i], ru[i]);
                double ynew = w - sigma * pr;
                dy[i] = ynew - yo;
                y[i] = ynew;
            }
            spmv(AT, dy, ady);
            for (int j = 0; j < n; ++j) { aty[j] += ady[j]; sum_x[j] += eta * x[j]; }
            for (int i = 0; i < m; ++i) sum_y[i] += eta * y[i];
        }
        sum_eta += eta;
        ++k;
        ++epoch;

        if (k % CE == 0) {
            double kc[5], ka[5], mu_a;
            kkt_eval(&L, x, y, ax, aty, kc);
            double mu_c = merit_of(kc, omega);
            const double *cxp, *cyp;
            if (sum_eta > 0.0) {
                for (int j = 0; j < n; ++j) avg_x[j] = sum_x[j] / sum_eta;
                for (int i = 0; i < m; ++i) avg_y[i] = sum_y[i] / sum_eta;
                spmv(A, avg_x, tax);
                spmv(AT, avg_y, taty);
                kkt_eval(&L, avg_x, avg_y, tax, taty, ka);
                mu_a = merit_of(ka, omega);
            } else {
                mu_a = INFINITY;
            }
            int use_avg = mu_a < mu_c;
            double mu_cand = use_avg ? mu_a : mu_c;
            cxp = use_avg ? avg_x : x;
            cyp = use_avg ? avg_y : y;

            /* _unscaled_kkt on the candidate (pdlp.py:497-520) */
            double pu = 0.0, du = 0.0, pou, dou, gu;
            {
                for (int j = 0; j < n; ++j) ux[j] = cxp[j] * col[j];
                for (int i = 0; i < m; ++i) uy[i] = (cyp[i] * row[i]) / (sobj / 1.0);
                spmv(A0, ux, tax);
                spmv(AT0, uy, taty);
                double pvm = 0.0, bvm = 0.0, supp = 0.0, bnd = 0.0, cxs = 0.0;
                for (int i = 0; i < m; ++i) {
                    double a = tax[i];
                    double pv = fmax(rl_in[i] - a, 0.0) + fmax(a - ru_in[i], 0.0);
                    if (fabs(pv) > pvm) pvm = fabs(pv);
                    double yi = uy[i];
                    double bsel = yi > 0.0 ? (ru_in[i] < ORACLE_INF ? ru_in[i] : 0.0)
                                           : (rl_in[i] > -ORACLE_INF ? rl_in[i] : 0.0);
                    supp += bsel * yi;
                }
                for (int j = 0; j < n; ++j) {
                    double xj = ux[j];
                    double bv = fmax(lo_in[j] - xj, 0.0) + fmax(xj - hi_in[j], 0.0);
                    if (fabs(bv) > bvm) bvm = fabs(bv);
                    double d = wc[j] + taty[j];
                    int lof = lo_in[j] > -ORACLE_INF, hif = hi_in[j] < ORACLE_INF;
                    double dr = (lof ? 0.0 : fmax(d, 0.0)) + (hif ? 0.0 : fmax(-d, 0.0));
                    if (fabs(dr) > du) du = fabs(dr);
                    cxs += wc[j] * xj;
                    double bsel = d > 0.0 ? (lof ? lo_in[j] : 0.0) : (hif ? hi_in[j] : 0.0);
                    bnd += bsel * d;
                }
                pu = fmax(pvm, bvm);
                pou = cxs + woff;
                dou = (-supp + bnd) + woff;
                gu = fabs(pou - dou);
            }
            if (pu <= ptol && du <= dtol && gu <= P->eps_abs + P->eps_rel * (fabs(pou) + fabs(dou))) {
                if (use_avg) {
                    memcpy(x, avg_x, sizeof(double) * n);
                    memcpy(y, avg_y, sizeof(double) * m);
                }
                status = ST_OPTIMAL;
                fin_p = pu; fin_d = du; fin_g = gu; fin_po = pou; fin_do = dou;
                break;
            }
            if (hist && hlen < hist_cap) {
                double *h = hist + 5 * hlen;
                h[0] = (double)k; h[1] = pu; h[2] = du; h[3] = gu; h[4] = pou;
            }
            ++hlen;
            double kk = (double)(k > 1 ? k : 1);
            int do_restart = (mu_cand <= P->beta_sufficient * mu_restart) ||
                             (mu_cand <= P->beta_necessary * mu_restart && mu_cand > mu_last) ||
                             ((double)epoch >= P->beta_artificial * kk);
            mu_last = mu_cand;
            if (do_restart) {
                if (use_avg) {
                    memcpy(x, avg_x, sizeof(double) * n);
                    memcpy(y, avg_y, sizeof(double) * m);
                }
                spmv(A, x, ax);
                spmv(AT, y, aty);
                double sx = 0.0, sy = 0.0;
                for (int j = 0; j < n; ++j) { double d = x[j] - xr[j]; sx += d * d; }
                for (int i = 0; i < m; ++i) { double d = y[i] - yr[i]; sy += d * d; }
                double ddx = sqrt(sx), ddy = sqrt(sy);
                if (ddx > 1e-12 && ddy > 1e-12) {
                    double th = P->primal_weight_smoothing;
                    omega = exp(th * log(ddy / ddx) + (1.0 - th) * log(omega));
                    omega = fmin(fmax(omega, 1e-6), 1e6);
                }
                memcpy(xr, x, sizeof(double) * n);
                memcpy(yr, y, sizeof(double) * m);
                mu_restart = mu_cand;
                memset(sum_x, 0, sizeof(double) * n);
                memset(sum_y, 0, sizeof(double) * m);
                sum_eta = 0.0;
                epoch = 0;
                ++restarts;
            }
        }
    }

    /* unscale and report (pdlp.py:465-494) */
    double obj = 0.0;
    for (int j = 0; j < n; ++j) {
        double v = x[j] * col[j];
        v = fmin(fmax(v, lo_in[j]), hi_in[j]);
        x_out[j] = v;
        obj += c_in[j] * v;
    }
    obj += offset_in;
    for (int i = 0; i < m; ++i) {
        double v = -((y[i] * row[i]) / (sobj / 1.0));
        y_out[i] = maximise ? -v : v;
    }
    spmv(AT0, y_out, taty);
    for (int j = 0; j < n; ++j) d_out[j] = c_in[j] - taty[j];

    info->status = status;
    info->iterations = k;
    info->restarts = restarts;
    info->rejected_steps = rejected;
    info->primal_res = fin_p; info->dual_res = fin_d; info->gap = fin_g;
    info->primal_obj = fin_po; info->dual_obj = fin_do;
    info->step_size = eta;
    info->primal_weight = omega;
    info->objective = obj;
    info->history_len = hlen;

    free(wc); free(row); free(col); free(sx_c); free(sx_r); free(c); free(lo); free(hi);
    free(rl); free(ru);
    free(x); free(y); free(ax); free(aty); free(xn); free(yn); free(adx); free(ady);
    free(dx); free(dy); free(sum_x); free(sum_y); free(xr); free(yr); free(avg_x);
    free(avg_y); free(tax); free(taty); free(ux); free(uy);
    return 0;
}

/* Single sparse product in the reference's order, for kernel parity tests. */
void oracle_spmv(int rows, const int64_t *p, const int32_t *i, const double *x,
                 const double *v, double *out) {
    csr_t M = {rows, p, i, x};
    spmv(M, v, out);
}
