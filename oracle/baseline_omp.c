/* CPU baseline with OpenMP threads (test infrastructure: timed by the
 * benchmark's cpu_baseline leg only; never linked by the product).
 *
 * The reference runs the factorization and the sweeps on an OpenMP team
 * (src/ilu_solver.cpp:591-614 factorization runner, :797-827 sweep
 * runners, scheduled by the subtree partition of src/subtree.hpp) and the
 * GMRES vector work through threaded MKL (src/gmres.cpp:172-276).  This
 * restatement keeps each row's arithmetic exactly as in oracle.c (IKJ
 * updates in ascending pivot order, fma, serial row sums) and replaces the
 * subtree/queue scheduler by level sets: the rows of one level of the L (or
 * U) dependency graph run in parallel, levels in order.  SpMV, dot
 * products, axpy and scaling are OpenMP loops; dot products reduce in the
 * threads' order (not bit-reproducible across thread counts, like the
 * reference's MKL reductions).
 *
 * baseline_ilu_gmres() times its phases separately: schedule (once per
 * pattern), factorization, and the GMRES(m) iteration of src/gmres.cpp:
 * 138-288 up to `maxits` Arnoldi steps. */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int nlev;
    int* ptr;  /* nlev + 1 */
    int* rows; /* n, rows by level */
} Levels;

static void levels_free(Levels* l) {
    free(l->ptr);
    free(l->rows);
}

/* level(i) = 1 + max level of the rows row i depends on (forward: its L
 * part; backward: its U part). */
static int levels_build(int n, const int* rp, const int* ci, const int* dg, int upper, Levels* out) {
    int* lev = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    if (!lev) return -1;
    int nlev = 0;
    for (int k = 0; k < n; ++k) {
        const int i = upper ? n - 1 - k : k;
        int l = 0;
        const int p0 = upper ? dg[i] + 1 : rp[i], p1 = upper ? rp[i + 1] : dg[i];
        for (int p = p0; p < p1; ++p)
            if (lev[ci[p]] + 1 > l) l = lev[ci[p]] + 1;
        lev[i] = l;
        if (l + 1 > nlev) nlev = l + 1;
    }
    out->nlev = nlev;
    out->ptr = (int*)calloc((size_t)nlev + 1, sizeof(int));
    out->rows = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    if (!out->ptr || !out->rows) { free(lev); return -1; }
    for (int i = 0; i < n; ++i) ++out->ptr[lev[i] + 1];
    for (int l = 0; l < nlev; ++l) out->ptr[l + 1] += out->ptr[l];
    int* fill = (int*)malloc((size_t)(nlev > 0 ? nlev : 1) * sizeof(int));
    memcpy(fill, out->ptr, (size_t)nlev * sizeof(int));
    for (int i = 0; i < n; ++i) out->rows[fill[lev[i]]++] = i;
    free(fill);
    free(lev);
    return 0;
}

/* Row IKJ ILU (oracle.c orc_ilu_factorize), rows of one L level in
 * parallel; each thread keeps a dense work row. */
static int factorize(int n, const int* rp, const int* ci, const int* dg, double* val, const Levels* L) {
    int zero = 0;
#pragma omp parallel reduction(+ : zero)
    {
        double* w = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
        for (int l = 0; l < L->nlev; ++l) {
#pragma omp for schedule(static)
            for (int k = L->ptr[l]; k < L->ptr[l + 1]; ++k) {
                const int i = L->rows[k];
                for (int p = rp[i]; p < rp[i + 1]; ++p) w[ci[p]] = val[p];
                for (int p = rp[i]; p < dg[i]; ++p) {
                    const int kk = ci[p];
                    const double piv = val[dg[kk]];
                    if (piv == 0.0) ++zero;
                    w[kk] = w[kk] / piv;
                    const double lik = w[kk];
                    for (int q = dg[kk] + 1; q < rp[kk + 1]; ++q) w[ci[q]] = fma(-val[q], lik, w[ci[q]]);
                }
                for (int p = rp[i]; p < rp[i + 1]; ++p) val[p] = w[ci[p]];
            }
        }
        free(w);
    }
    return zero;
}

/* x = U^-1 L^-1 b (oracle.c orc_substitute), level by level. */
static void substitute(int n, const int* rp, const int* ci, const int* dg, const double* val, const double* b,
                       double* x, const Levels* L, const Levels* U) {
    if (b != x) memcpy(x, b, (size_t)n * sizeof(double));
#pragma omp parallel
    {
        for (int l = 0; l < L->nlev; ++l) {
#pragma omp for schedule(static)
            for (int k = L->ptr[l]; k < L->ptr[l + 1]; ++k) {
                const int i = L->rows[k];
                double xi = x[i];
                for (int p = rp[i]; p < dg[i]; ++p) xi = fma(-x[ci[p]], val[p], xi);
                x[i] = xi;
            }
        }
        for (int l = 0; l < U->nlev; ++l) {
#pragma omp for schedule(static)
            for (int k = U->ptr[l]; k < U->ptr[l + 1]; ++k) {
                const int i = U->rows[k];
                double xi = x[i];
                for (int p = dg[i] + 1; p < rp[i + 1]; ++p) xi = fma(-x[ci[p]], val[p], xi);
                x[i] = xi / val[dg[i]];
            }
        }
    }
}

static void spmv(int n, const int* rp, const int* ci, const double* val, double alpha, const double* x, double beta,
                 double* y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double t = 0.0;
        for (int p = rp[i]; p < rp[i + 1]; ++p) t = fma(val[p], x[ci[p]], t);
        y[i] = (beta == 0.0) ? alpha * t : beta * y[i] + alpha * t;
    }
}

static double dot(int n, const double* a, const double* b) {
    double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void axpy(int n, double a, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) y[i] = fma(a, x[i], y[i]);
}

static void scal(int n, double a, double* x) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) x[i] *= a;
}

int baseline_threads(void) { return omp_get_max_threads(); }

void baseline_set_threads(int k) { omp_set_num_threads(k); }

/* times[0] schedule, [1] factorization, [2] GMRES (seconds); fval is
 * factorized in place.  Returns 1 converged, 0 not (or step cap reached),
 * -1 allocation failure. */
int baseline_ilu_gmres(int n, const int* arp, const int* aci, const double* aval, const int* frp, const int* fci,
                       const int* fdg, double* fval, const double* rhs, double* sol, double tol, int m, int maxits,
                       int* its_out, double* times) {
    Levels L = {0, NULL, NULL}, U = {0, NULL, NULL};
    double t0 = omp_get_wtime();
    if (levels_build(n, frp, fci, fdg, 0, &L) || levels_build(n, frp, fci, fdg, 1, &U)) return -1;
    double t1 = omp_get_wtime();
    factorize(n, frp, fci, fdg, fval, &L);
    double t2 = omp_get_wtime();

    /* GMRES(m), src/gmres.cpp:138-288 (as oracle.c orc_gmres) */
    const double epsmac = 1e-16;
    const int im = m, im1 = m + 1;
    double* vv = (double*)malloc((size_t)im1 * n * sizeof(double));
    double* z = (double*)malloc((size_t)im * n * sizeof(double));
    double* hh = (double*)calloc((size_t)im1 * (im + 3), sizeof(double));
    int ret = 1, its = 0;
    if (!vv || !z || !hh) { free(vv); free(z); free(hh); levels_free(&L); levels_free(&U); return -1; }
    double* c = hh + (size_t)im1 * im;
    double* s = c + im1;
    double* rs = s + im1;
    while (its < maxits) {
        memcpy(vv, rhs, (size_t)n * sizeof(double));
        spmv(n, arp, aci, aval, -1.0, sol, 1.0, vv);
        double beta = sqrt(dot(n, vv, vv));
        if (beta <= tol) break;
        if (!(beta == beta)) { ret = 0; break; }
        scal(n, 1.0 / beta, vv);
        rs[0] = beta;
        int i = -1;
        while ((i < im - 1) && (beta > tol) && (its++ < maxits)) {
            ++i;
            const int i1 = i + 1;
            double* vi = vv + (size_t)i * n;
            double* vi1 = vv + (size_t)i1 * n;
            double* zi = z + (size_t)i * n;
            substitute(n, frp, fci, fdg, fval, vi, zi, &L, &U);
            spmv(n, arp, aci, aval, 1.0, zi, 0.0, vi1);
            const int ptih = i * im1;
            for (int j = 0; j <= i; ++j) {
                const double t = dot(n, vv + (size_t)j * n, vi1);
                hh[ptih + j] = t;
                axpy(n, -t, vv + (size_t)j * n, vi1);
            }
            double t = sqrt(dot(n, vi1, vi1));
            hh[ptih + i1] = t;
            if (t <= epsmac) break;
            scal(n, 1.0 / t, vi1);
            for (int k = 1; k <= i; ++k) {
                const int k1 = k - 1;
                t = hh[ptih + k1];
                hh[ptih + k1] = c[k1] * t + s[k1] * hh[ptih + k];
                hh[ptih + k] = -s[k1] * t + c[k1] * hh[ptih + k];
            }
            const double a = hh[ptih + i], b = hh[ptih + i1];
            double gam = sqrt(a * a + b * b);
            if (gam == 0.0) gam = epsmac;
            c[i] = hh[ptih + i] / gam;
            s[i] = hh[ptih + i1] / gam;
            rs[i1] = -s[i] * rs[i];
            rs[i] = c[i] * rs[i];
            hh[ptih + i] = c[i] * hh[ptih + i] + s[i] * hh[ptih + i1];
            beta = fabs(rs[i1]);
            if (!(beta == beta)) { ret = 0; goto done; }
        }
        if (i < 0) break;
        rs[i] = rs[i] / hh[i * im1 + i];
        for (int ii = i - 1; ii >= 0; --ii) {
            double t = rs[ii];
            for (int j = ii + 1; j <= i; ++j) t -= hh[j * im1 + ii] * rs[j];
            rs[ii] = t / hh[ii * im1 + ii];
        }
        for (int j = 0; j <= i; ++j) axpy(n, rs[j], z + (size_t)j * n, sol);
        if (beta <= tol) break;
        else if (its >= maxits) { ret = 0; break; }
    }
done:
    if (its_out) *its_out = its;
    double t3 = omp_get_wtime();
    if (times) { times[0] = t1 - t0; times[1] = t2 - t1; times[2] = t3 - t2; }
    free(vv); free(z); free(hh);
    levels_free(&L);
    levels_free(&U);
    return ret;
}
