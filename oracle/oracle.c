/* CPU oracle for ILU(k)-GMRES: a plain-C restatement of the reference
 * algorithm.  TEST INFRASTRUCTURE ONLY -- see oracle.h for what may use it
 * and what pins each routine. */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc_free(void* p) { free(p); }

/* ------------------------------------------------------------------ */
/* Symbolic ILU(k), src/iluk.cpp:5-54.                                 */
/* Row i is held as a sorted singly linked list over dense arrays      */
/* (nxt, lev, mark); the reference uses std::map<int,int> per row.     */
/* ------------------------------------------------------------------ */
typedef struct { int* p; long cap, len; } ivec;

static int ivec_push(ivec* v, int x) {
    if (v->len == v->cap) {
        long nc = v->cap ? 2 * v->cap : 1024;
        int* np = (int*)realloc(v->p, (size_t)nc * sizeof(int));
        if (!np) return -1;
        v->p = np;
        v->cap = nc;
    }
    v->p[v->len++] = x;
    return 0;
}

int orc_iluk_symbolic(int n, const int* arp, const int* aci, const double* aval,
                      int lof, int** frp_out, int** fci_out, int** flev_out,
                      double** fval_out, int* nzmap) {
    const int head = n;                 /* list head sentinel            */
    const int tail = n + 1;             /* list end marker (> any col)   */
    int* nxt = (int*)malloc((size_t)(n + 2) * sizeof(int));
    int* lev = (int*)malloc((size_t)(n + 1) * sizeof(int));
    int* mark = (int*)malloc((size_t)(n + 1) * sizeof(int));
    int* frp = (int*)malloc((size_t)(n + 1) * sizeof(int));
    int* dpos = (int*)malloc((size_t)(n + 1) * sizeof(int)); /* pos of (k,k) or -1 */
    ivec ci = {0, 0, 0}, lv = {0, 0, 0};
    int dup = 0;
    if (!nxt || !lev || !mark || !frp || !dpos) goto fail;
    for (int i = 0; i <= n; ++i) mark[i] = -1;

    frp[0] = 0;
    for (int i = 0; i < n; ++i) {
        /* :11-13 -- row i starts as A's columns at level 0 (a repeated
         * column collapses, as std::map assignment does) */
        nxt[head] = tail;
        for (int p = arp[i]; p < arp[i + 1]; ++p) {
            int c = aci[p];
            if (mark[c] == i) { dup = 1; continue; }
            int q = head;
            while (nxt[q] != tail && nxt[q] < c) q = nxt[q];
            nxt[c] = nxt[q];
            nxt[q] = c;
            mark[c] = i;
            lev[c] = 0;
        }
        /* :14-28 -- walk pivots k < i in ascending order, including
         * entries inserted during the walk (they are > k) */
        for (int k = nxt[head]; k != tail && k < i; k = nxt[k]) {
            int lk = lev[k];
            if (lk >= lof) continue;                  /* :19-21 */
            if (dpos[k] < 0) continue;                /* find(k) == end() */
            int prev = k;
            for (long q = dpos[k]; q < frp[k + 1]; ++q) {   /* :22 */
                int j = ci.p[q];
                int L = lk + lv.p[q] + 1;                    /* :23 */
                if (L > lof) continue;                       /* :24 */
                if (mark[j] == i) continue;                  /* try_emplace keeps the first level */
                while (nxt[prev] != tail && nxt[prev] < j) prev = nxt[prev];
                nxt[j] = nxt[prev];
                nxt[prev] = j;
                mark[j] = i;
                lev[j] = L;
                prev = j;
            }
        }
        /* :29, :34-47 -- emit the row */
        dpos[i] = -1;
        for (int c = nxt[head]; c != tail; c = nxt[c]) {
            if (c == i) dpos[i] = (int)ci.len;
            if (ivec_push(&ci, c) || ivec_push(&lv, lev[c])) goto fail;
        }
        if (ci.len > 0x7fffffffL) goto fail;
        frp[i + 1] = (int)ci.len;
    }

    {
        const int nnzf = frp[n];
        double* fval = (double*)calloc((size_t)(nnzf > 0 ? nnzf : 1), sizeof(double));
        if (!fval) goto fail;
        /* :36-44 -- level-0 slots take A's values in A's entry order */
        int innz = 0;
        for (int i = 0; i < n; ++i) {
            for (int q = frp[i]; q < frp[i + 1]; ++q) {
                if (lv.p[q] == 0) {
                    if (innz < arp[n]) {
                        fval[q] = aval ? aval[innz] : 0.0;
                        if (nzmap) nzmap[innz] = q;
                    }
                    ++innz;
                }
            }
        }
        if (dup || innz != arp[n]) { free(fval); goto fail; }   /* :49-52 */
        free(nxt); free(lev); free(mark); free(dpos);
        *frp_out = frp;
        *fci_out = ci.p ? ci.p : (int*)malloc(sizeof(int));
        *flev_out = lv.p ? lv.p : (int*)malloc(sizeof(int));
        *fval_out = fval;
        return nnzf;
    }
fail:
    free(nxt); free(lev); free(mark); free(frp); free(dpos); free(ci.p); free(lv.p);
    return -1;
}

/* ------------------------------------------------------------------ */
/* Diagonal pointer, src/ilu_solver.cpp:276-289.                       */
/* ------------------------------------------------------------------ */
void orc_diag_ptr(int n, const int* rp, const int* ci, int* diag) {
    for (int i = 0; i < n; ++i) {
        diag[i] = rp[i + 1];
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            if (ci[p] >= i) { diag[i] = p; break; }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Numeric ILU, left_looking_col<Skip,...,S=true>                       */
/* (src/ilu_solver.cpp:553-589): scatter row into a dense work row,     */
/* eliminate with pivots k ascending, gather back on the pattern.      */
/* ------------------------------------------------------------------ */
int orc_ilu_factorize(int n, const int* rp, const int* ci, const int* diag,
                      double* val) {
    double* w = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    int zero_pivots = 0;
    if (!w) return -1;
    for (int i = 0; i < n; ++i) {
        for (int p = rp[i]; p < rp[i + 1]; ++p) w[ci[p]] = val[p];        /* :565-568 */
        for (int p = rp[i]; p < diag[i]; ++p) {                            /* :571 */
            int k = ci[p];
            double piv = val[diag[k]];
            if (piv == 0.0) ++zero_pivots;
            w[k] = w[k] / piv;                                             /* :576 */
            double lik = w[k];
            for (int q = diag[k] + 1; q < rp[k + 1]; ++q)                  /* :577-579 */
                w[ci[q]] = fma(-val[q], lik, w[ci[q]]);
        }
        for (int p = rp[i]; p < rp[i + 1]; ++p) val[p] = w[ci[p]];        /* :581-583 */
    }
    free(w);
    return zero_pivots;
}

int orc_ilu_factorize_comp(int n, const int* rp, const int* ci, const int* diag,
                           double* val, const float* omega) {
    double* w = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    char* in = (char*)calloc((size_t)(n > 0 ? n : 1), 1);
    int zero_pivots = 0;
    if (!w || !in) { free(w); free(in); return -1; }
    for (int i = 0; i < n; ++i) {
        double dropped = 0.0;
        for (int p = rp[i]; p < rp[i + 1]; ++p) { w[ci[p]] = val[p]; in[ci[p]] = 1; }
        for (int p = rp[i]; p < diag[i]; ++p) {
            int k = ci[p];
            double piv = val[diag[k]];
            if (piv == 0.0) ++zero_pivots;
            w[k] = w[k] / piv;
            double lik = w[k];
            for (int q = diag[k] + 1; q < rp[k + 1]; ++q) {
                if (in[ci[q]]) w[ci[q]] = fma(-val[q], lik, w[ci[q]]);
                else dropped = fma(-val[q], lik, dropped);
            }
        }
        if (omega) w[i] = fma((double)omega[i], dropped, w[i]);
        for (int p = rp[i]; p < rp[i + 1]; ++p) { val[p] = w[ci[p]]; in[ci[p]] = 0; }
    }
    free(w);
    free(in);
    return zero_pivots;
}

/* ------------------------------------------------------------------ */
/* The factorization of the reference's USE_MKL_ILU build, which calls   */
/* MKL's dcsrilu0 (src/ilu_solver.cpp:248-259,631-639; MKL is closed and */
/* absent, so this restates its documented contract, not its code): ILU  */
/* on the given pattern -- the same IKJ elimination as                   */
/* orc_ilu_factorize -- with each row's final pivot checked before later */
/* rows use it.  policy 1 (ipar[30] = 0): a pivot that is zero or below  */
/* tiny (dpar[30]) in magnitude stops the factorization with -102 (zero) */
/* or -103 (small); policy 2 (ipar[30] != 0): it is replaced by repl      */
/* (dpar[31]) and the factorization goes on; policy 0: no check.  Also   */
/* dcsrilu0's input checks: -101 a row without its diagonal, -106 a row  */
/* whose columns are not ascending.  Returns 0 or that code.             */
/* ------------------------------------------------------------------ */
int orc_ilu0_guarded(int n, const int* rp, const int* ci, const int* diag, double* val,
                     int policy, double tiny, double repl) {
    for (int i = 0; i < n; ++i) {
        if (diag[i] >= rp[i + 1] || ci[diag[i]] != i) return -101;
        for (int p = rp[i]; p + 1 < rp[i + 1]; ++p)
            if (ci[p] >= ci[p + 1]) return -106;
    }
    double* w = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    if (!w) return -1;
    int status = 0;
    for (int i = 0; i < n && status == 0; ++i) {
        for (int p = rp[i]; p < rp[i + 1]; ++p) w[ci[p]] = val[p];
        for (int p = rp[i]; p < diag[i]; ++p) {
            int k = ci[p];
            w[k] = w[k] / val[diag[k]];
            double lik = w[k];
            for (int q = diag[k] + 1; q < rp[k + 1]; ++q) w[ci[q]] = fma(-val[q], lik, w[ci[q]]);
        }
        double d = w[i];
        if (policy != 0 && (d == 0.0 || fabs(d) < tiny)) {
            if (policy == 1) status = d == 0.0 ? -102 : -103;
            else w[i] = repl;
        }
        for (int p = rp[i]; p < rp[i + 1]; ++p) val[p] = w[ci[p]];
    }
    free(w);
    return status;
}

/* ------------------------------------------------------------------ */
/* Triangular sweeps, src/ilu_solver.cpp:775-796 (one-thread path).    */
/* ------------------------------------------------------------------ */
void orc_substitute(int n, const int* rp, const int* ci, const int* diag,
                    const double* val, const double* b, double* x) {
    if (b != x) memcpy(x, b, (size_t)n * sizeof(double));                 /* :775-777 */
    for (int i = 0; i < n; ++i) {                                          /* :783-788 */
        double xi = x[i];
        for (int p = rp[i]; p < diag[i]; ++p) xi = fma(-x[ci[p]], val[p], xi);
        x[i] = xi;
    }
    for (int i = n - 1; i >= 0; --i) {                                     /* :789-795 */
        double xi = x[i];
        for (int p = diag[i] + 1; p < rp[i + 1]; ++p) xi = fma(-x[ci[p]], val[p], xi);
        x[i] = xi / val[diag[i]];
    }
}

/* ------------------------------------------------------------------ */
/* The same sweeps with each row's sum split over g lanes (g a power of  */
/* two <= 64): lane t accumulates entries t, t+g, t+2g, ... of the row    */
/* in order with fma, lane 0 starting from the right-hand side; then the  */
/* lanes are combined pairwise, lane t += lane t+off for off = g/2, g/4,  */
/* ..., 1.  This is the order of the GPU sweeps with g lanes per row; g = */
/* 1 is orc_substitute.                                                   */
/* ------------------------------------------------------------------ */
static double lane_sum(int g, double init, int p0, int p1, const int* ci, const double* val, const double* x) {
    double part[64];
    for (int t = 0; t < g; ++t) part[t] = 0.0;
    part[0] = init;
    for (int p = p0; p < p1; ++p) {
        const int t = (p - p0) % g;
        part[t] = fma(-x[ci[p]], val[p], part[t]);
    }
    for (int off = g / 2; off > 0; off /= 2)
        for (int t = 0; t < off; ++t) part[t] += part[t + off];
    return part[0];
}

void orc_substitute_lanes(int n, const int* rp, const int* ci, const int* diag, const double* val,
                          const double* b, double* x, int g) {
    if (b != x) memcpy(x, b, (size_t)n * sizeof(double));
    for (int i = 0; i < n; ++i) x[i] = lane_sum(g, x[i], rp[i], diag[i], ci, val, x);
    for (int i = n - 1; i >= 0; --i) x[i] = lane_sum(g, x[i], diag[i] + 1, rp[i + 1], ci, val, x) / val[diag[i]];
}

/* ------------------------------------------------------------------ */
/* One triangle of orc_substitute_lanes, as the reference's USE_MKL_SV  */
/* build applies the preconditioner (two mkl_sparse_d_trsv calls,        */
/* src/ilu_solver.cpp:753-773): upper 0, y = alpha L^-1 x with L the     */
/* unit lower triangle; upper 1, y = alpha U^-1 x with U the upper       */
/* triangle and its diagonal.  Row sums in the g-lane order; alpha is    */
/* applied to the finished solution.  x may alias y.                     */
/* ------------------------------------------------------------------ */
void orc_trsv_lanes(int n, const int* rp, const int* ci, const int* diag, const double* val,
                    int upper, double alpha, const double* x, double* y, int g) {
    if (x != y) memcpy(y, x, (size_t)n * sizeof(double));
    if (!upper) {
        for (int i = 0; i < n; ++i) y[i] = lane_sum(g, y[i], rp[i], diag[i], ci, val, y);
    } else {
        for (int i = n - 1; i >= 0; --i) y[i] = lane_sum(g, y[i], diag[i] + 1, rp[i + 1], ci, val, y) / val[diag[i]];
    }
    if (alpha != 1.0)
        for (int i = 0; i < n; ++i) y[i] *= alpha;
}

/* ------------------------------------------------------------------ */
/* y = alpha A x + beta y (the mkl_sparse_d_mv calls of src/gmres.cpp). */
/* ------------------------------------------------------------------ */
void orc_spmv(int n, const int* rp, const int* ci, const double* val,
              double alpha, const double* x, double beta, double* y) {
    for (int i = 0; i < n; ++i) {
        double t = 0.0;
        for (int p = rp[i]; p < rp[i + 1]; ++p) t = fma(val[p], x[ci[p]], t);
        y[i] = (beta == 0.0) ? alpha * t : beta * y[i] + alpha * t;
    }
}

static double dot(int n, const double* a, const double* b) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = fma(a[i], b[i], s);
    return s;
}

static double nrm2(int n, const double* a) { return sqrt(dot(n, a, a)); }

static void axpy(int n, double a, const double* x, double* y) {
    for (int i = 0; i < n; ++i) y[i] = fma(a, x[i], y[i]);
}

static void scal(int n, double a, double* x) {
    for (int i = 0; i < n; ++i) x[i] *= a;
}

/* ------------------------------------------------------------------ */
/* GMRES(m), src/gmres.cpp:138-288, step by step.                       */
/* ------------------------------------------------------------------ */
int orc_gmres(int n, const int* arp, const int* aci, const double* aval,
              const int* frp, const int* fci, const int* fdiag, const double* fval,
              const double* rhs, double* sol, double tol, int m, int maxits,
              int* its_out, double* hist, int hist_cap, int* nhist) {
    const double epsmac = 1e-16;                                           /* :145 */
    const int im = m, im1 = m + 1;
    const double eps1 = tol;                                               /* :147 */
    double* vv = (double*)malloc((size_t)im1 * n * sizeof(double));
    double* z = (double*)malloc((size_t)im * n * sizeof(double));
    double* hh = (double*)calloc((size_t)im1 * (im + 3), sizeof(double));
    int ret = 1, its = 0, nh = 0;
    if (!vv || !z || !hh) { free(vv); free(z); free(hh); return -1; }
    double* c = hh + (size_t)im1 * im;                                     /* :158-160 */
    double* s = c + im1;
    double* rs = s + im1;
    int ptih = 0;

    while (its < maxits) {                                                 /* :179 */
        memcpy(vv, rhs, (size_t)n * sizeof(double));                      /* :181 */
        orc_spmv(n, arp, aci, aval, -1.0, sol, 1.0, vv);                   /* :183 */
        double beta = nrm2(n, vv);                                         /* :185 */
        if (beta <= tol) break;                                            /* :186-188 */
        /* Not in the reference: with a NaN residual its inner loop never
         * runs, `its` never advances and the outer loop never ends (after
         * writing rs[-1]).  Report failure instead.  Bit test, so that
         * -ffast-math builds keep it. */
        {
            unsigned long long bits;
            memcpy(&bits, &beta, sizeof bits);
            if ((bits & 0x7FF0000000000000ull) == 0x7FF0000000000000ull && (bits & 0x000FFFFFFFFFFFFFull)) {
                ret = 0;
                goto done;
            }
        }
        double t = 1.0 / beta;                                             /* :189 */
        scal(n, t, vv);                                                    /* :191 */
        rs[0] = beta;                                                      /* :196 */
        int i = -1;                                                        /* :198 */
        while ((i < im - 1) && (beta > eps1) && (its++ < maxits)) {        /* :199 */
            ++i;
            int i1 = i + 1;
            double* vi = vv + (size_t)i * n;
            double* vi1 = vv + (size_t)i1 * n;
            double* zi = z + (size_t)i * n;
            if (frp) orc_substitute(n, frp, fci, fdiag, fval, vi, zi);     /* :208 */
            else memcpy(zi, vi, (size_t)n * sizeof(double));
            orc_spmv(n, arp, aci, aval, 1.0, zi, 0.0, vi1);                /* :212 */
            ptih = i * im1;                                                /* :218 */
            for (int j = 0; j <= i; ++j) {                                 /* :219-224 */
                t = dot(n, vv + (size_t)j * n, vi1);
                hh[ptih + j] = t;
                axpy(n, -t, vv + (size_t)j * n, vi1);
            }
            t = nrm2(n, vi1);                                              /* :226 */
            hh[ptih + i1] = t;
            if (t <= epsmac) break;                                        /* :228-230 */
            t = 1.0 / t;
            scal(n, t, vi1);                                               /* :233 */
            for (int k = 1; k <= i; ++k) {                                 /* :238-243 */
                int k1 = k - 1;
                t = hh[ptih + k1];
                hh[ptih + k1] = c[k1] * t + s[k1] * hh[ptih + k];
                hh[ptih + k] = -s[k1] * t + c[k1] * hh[ptih + k];
            }
            double a = hh[ptih + i], b = hh[ptih + i1];
            double gam = sqrt(a * a + b * b);                              /* :244 */
            if (gam == 0.0) gam = epsmac;                                  /* :246-248 */
            c[i] = hh[ptih + i] / gam;                                     /* :250-253 */
            s[i] = hh[ptih + i1] / gam;
            rs[i1] = -s[i] * rs[i];
            rs[i] = c[i] * rs[i];
            hh[ptih + i] = c[i] * hh[ptih + i] + s[i] * hh[ptih + i1];     /* :255 */
            beta = fabs(rs[i1]);                                           /* :256 */
            if (hist && nh < hist_cap) hist[nh] = beta;
            ++nh;
            if (isnan(beta)) { ret = 0; goto done; }                       /* :259-262 */
        }
        rs[i] = rs[i] / hh[ptih + i];                                      /* :265 */
        for (int ii = i - 1; ii >= 0; --ii) {                              /* :266-272 */
            t = rs[ii];
            for (int j = ii + 1; j <= i; ++j) t -= hh[j * im1 + ii] * rs[j];
            rs[ii] = t / hh[ii * im1 + ii];
        }
        for (int j = 0; j <= i; ++j) axpy(n, rs[j], z + (size_t)j * n, sol); /* :274-276 */
        if (beta <= eps1) break;                                           /* :278-280 */
        else if (its > maxits) { ret = 0; goto done; }                     /* :281-283 */
    }
done:
    if (its_out) *its_out = its;
    if (nhist) *nhist = nh < hist_cap ? nh : hist_cap;
    free(vv); free(z); free(hh);
    return ret;
}
