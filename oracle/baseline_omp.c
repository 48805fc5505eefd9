/* CPU baseline with OpenMP threads (test infrastructure: timed by the
 * benchmark's cpu_baseline leg only; never linked by the product).
 *
 * The reference runs the factorization and the sweeps on an OpenMP team
 * (src/ilu_solver.cpp:591-614 factorization runner, :797-827 sweep
 * runners) and the GMRES vector work through threaded MKL
 * (src/gmres.cpp:172-276).  This restatement keeps each row's arithmetic
 * exactly as in oracle.c (IKJ updates in ascending pivot order, fma, serial
 * row sums), so every schedule gives the serial oracle's factors and
 * sweeps bit for bit, and offers two schedules:
 *
 *   sched 1, the reference's (src/subtree.hpp:598-688, set up as in
 *   src/ilu_solver.cpp:295-358, run as in :375-480): an elimination forest
 *   of the dependency graph; its heaviest subtrees (weight = entries of the
 *   row in that direction + a base) are split until none outweighs the
 *   average per thread by more than 1/relax; whole subtrees go to threads
 *   largest first onto the least loaded, and run with no waits (a subtree
 *   holds all its rows' dependencies); the rows left over form a queue in
 *   topological order.  When the queue is large, the first levels of the
 *   graph (up to a fraction `strip` of the rows) are "stripped": they run
 *   level by level in parallel before the subtrees, which are then formed
 *   without them.  The factorization (all threads) and sweeps with at least
 *   40 entries per row run their queue in parallel, each row busy-waiting on
 *   its dependencies' done flags; sparser sweeps (at most 8 threads) run
 *   their queue on one thread after the parallel part.  The thresholds are
 *   the reference's (queue ratios 0.1 / 0.02, strip 0.2 then 0.5, relax 0.7
 *   when stripping).  Its packed-task form (8 queued rows per task, each
 *   first run up to its first unmet dependency) is a timing refinement of
 *   the same busy-wait queue and is not restated.
 *
 *   sched 0: level sets, the rows of one level of the L (or U) dependency
 *   graph in parallel, levels in order.
 *
 * SpMV, dot products, axpy and scaling are OpenMP loops; dot products
 * reduce in the threads' order (not bit-reproducible across thread counts,
 * like the reference's MKL reductions).
 *
 * baseline_ilu_gmres() times its phases separately: schedule (once per
 * pattern), factorization, and the GMRES(m) iteration of src/gmres.cpp:
 * 138-288 up to `maxits` Arnoldi steps. */
#include <math.h>
#include <omp.h>
#include <stdatomic.h>
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

/* ------------------------------------------------------------------ */
/* Subtree / queue schedule                                            */
/* ------------------------------------------------------------------ */

typedef struct {
    int nthreads;
    int nstrip; /* stripped levels */
    int* strip_ptr;
    int* strip_rows;
    int* part_ptr; /* nthreads + 1 */
    int* part_rows;
    int nqueue;
    int* queue;
} Subpart;

static void subpart_free(Subpart* s) {
    free(s->strip_ptr);
    free(s->strip_rows);
    free(s->part_ptr);
    free(s->part_rows);
    free(s->queue);
    memset(s, 0, sizeof(*s));
}

/* Dependencies of row i in one direction: L, the columns before the
 * diagonal; U, the columns after it.  Positions t put the rows in
 * topological order: t = i (L) or n - 1 - i (U). */
#define DEP_BEGIN(i) (upper ? dg[i] + 1 : rp[i])
#define DEP_END(i) (upper ? rp[(i) + 1] : dg[i])
#define ROW_OF(t) (upper ? n - 1 - (t) : (t))
#define POS_OF(i) (upper ? n - 1 - (i) : (i))

static int cmp_int(const void* a, const void* b) {
    const int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/* Builds the schedule; returns 0, or -1 on allocation failure.  base: the
 * weight added to a row's entry count; relax <= 1; strip <= 0: none. */
static int subpart_build(int n, const int* rp, const int* ci, const int* dg, int upper, int nthreads, int base,
                         double relax, double strip, Subpart* out) {
    memset(out, 0, sizeof(*out));
    out->nthreads = nthreads;
    const size_t N = (size_t)(n > 0 ? n : 1);
    char* stripped = (char*)calloc(N, 1);
    int* parent = (int*)malloc((N + 1) * sizeof(int));
    int* anc = (int*)malloc((N + 1) * sizeof(int));
    int* first = (int*)malloc((N + 1) * sizeof(int));
    int* next = (int*)malloc((N + 1) * sizeof(int));
    long* w = (long*)calloc(N + 1, sizeof(long));
    int* owner = (int*)malloc(N * sizeof(int));
    int* cand = (int*)malloc((N + 1) * sizeof(int));
    out->part_ptr = (int*)calloc((size_t)nthreads + 1, sizeof(int));
    out->part_rows = (int*)malloc(N * sizeof(int));
    out->queue = (int*)malloc(N * sizeof(int));
    out->strip_ptr = (int*)calloc(2, sizeof(int));
    out->strip_rows = (int*)malloc(N * sizeof(int));
    int rc = 0;
    if (!stripped || !parent || !anc || !first || !next || !w || !owner || !cand || !out->part_ptr ||
        !out->part_rows || !out->queue || !out->strip_ptr || !out->strip_rows) {
        rc = -1;
        goto end;
    }
    /* stripping: whole levels from the first, while fewer than strip * n
     * rows are stripped (the level that crosses the bound is included) */
    if (strip > 0) {
        int* lev = owner; /* scratch */
        int nlev = 0;
        for (int t = 0; t < n; ++t) {
            const int i = ROW_OF(t);
            int l = 0;
            for (int p = DEP_BEGIN(i); p < DEP_END(i); ++p) {
                const int d = lev[POS_OF(ci[p])] + 1;
                if (d > l) l = d;
            }
            lev[t] = l;
            if (l + 1 > nlev) nlev = l + 1;
        }
        int* cnt = (int*)calloc((size_t)nlev + 1, sizeof(int));
        if (!cnt) { rc = -1; goto end; }
        for (int t = 0; t < n; ++t) ++cnt[lev[t] + 1];
        int nl = 0;
        long acc = 0;
        while (nl < nlev && !((double)acc > strip * n)) {
            acc += cnt[nl + 1];
            ++nl;
        }
        free(out->strip_ptr);
        out->strip_ptr = (int*)calloc((size_t)nl + 1, sizeof(int));
        if (!out->strip_ptr) { free(cnt); rc = -1; goto end; }
        for (int l = 0; l < nl; ++l) out->strip_ptr[l + 1] = out->strip_ptr[l] + cnt[l + 1];
        int* fill = cand; /* scratch */
        for (int l = 0; l < nl; ++l) fill[l] = out->strip_ptr[l];
        for (int t = 0; t < n; ++t)
            if (lev[t] < nl) {
                stripped[t] = 1;
                out->strip_rows[fill[lev[t]]++] = ROW_OF(t);
            }
        out->nstrip = nl;
        free(cnt);
    }
    /* elimination forest over the rows not stripped: every dependency of t
     * ends up below t (union-find with path compression; n = the root) */
    for (int t = 0; t <= n; ++t) parent[t] = anc[t] = n;
    for (int t = 0; t < n; ++t) {
        if (stripped[t]) continue;
        const int i = ROW_OF(t);
        for (int p = DEP_BEGIN(i); p < DEP_END(i); ++p) {
            int r = POS_OF(ci[p]);
            if (stripped[r]) continue;
            while (anc[r] != n) {
                const int up = anc[r];
                anc[r] = t;
                r = up;
            }
            if (r != t) {
                anc[r] = t;
                parent[r] = t;
            }
        }
    }
    for (int t = 0; t <= n; ++t) first[t] = next[t] = -1;
    for (int t = n - 1; t >= 0; --t) {
        if (stripped[t]) continue;
        next[t] = first[parent[t]];
        first[parent[t]] = t;
    }
    for (int t = 0; t < n; ++t)
        if (!stripped[t]) {
            const int i = ROW_OF(t);
            w[t] += DEP_END(i) - DEP_BEGIN(i) + base;
            w[parent[t]] += w[t];
        }
    /* split the heaviest subtree (below its single-child chain) into its
     * children while it outweighs the average by more than 1/relax; a
     * subtree that is a bare chain is kept whole */
    int nsub = 0, ncand = 0;
    int* subs = anc; /* reuse: final subtree roots */
    long sum = w[n];
    cand[ncand++] = n;
    if (nthreads > 1) {
        while (ncand > 0) {
            int best = 0;
            for (int k = 1; k < ncand; ++k)
                if (w[cand[k]] > w[cand[best]]) best = k;
            if (!((double)w[cand[best]] * nthreads * relax > (double)sum)) break;
            int t = cand[best];
            while (first[t] >= 0 && next[first[t]] < 0) t = first[t];
            if (first[t] < 0) {  /* a chain: keep it whole */
                subs[nsub++] = cand[best];
                cand[best] = cand[--ncand];
                continue;
            }
            sum -= w[cand[best]];
            cand[best] = cand[--ncand];
            for (int c = first[t]; c >= 0; c = next[c]) {
                cand[ncand++] = c;
                sum += w[c];
            }
        }
    }
    for (int k = 0; k < ncand; ++k) subs[nsub++] = cand[k];
    if (nthreads <= 1 || nsub < nthreads || (nsub == 1 && subs[0] == n)) nsub = 0; /* everything queued */
    /* largest subtrees first, each to the least loaded thread */
    for (int a = 1; a < nsub; ++a) { /* insertion sort by weight, descending */
        const int x = subs[a];
        int b = a - 1;
        while (b >= 0 && w[subs[b]] < w[x]) { subs[b + 1] = subs[b]; --b; }
        subs[b + 1] = x;
    }
    long* load = (long*)calloc((size_t)nthreads, sizeof(long));
    if (!load) { rc = -1; goto end; }
    for (int t = 0; t < n; ++t) owner[t] = -1;
    for (int k = 0; k < nsub; ++k) {
        int th = 0;
        for (int q = 1; q < nthreads; ++q)
            if (load[q] < load[th]) th = q;
        load[th] += w[subs[k]];
        /* mark the subtree (iterative walk over first-child / next-sibling) */
        const int root = subs[k];
        if (root == n) continue;
        owner[root] = th;
        int v = first[root];
        while (v >= 0) {
            owner[v] = th;
            if (first[v] >= 0) { v = first[v]; continue; }
            while (v >= 0 && v != root && next[v] < 0) v = parent[v];
            if (v < 0 || v == root) break;
            v = next[v];
        }
    }
    free(load);
    /* each thread's rows and the queue, in topological (position) order */
    for (int t = 0; t < n; ++t)
        if (!stripped[t] && owner[t] >= 0) ++out->part_ptr[owner[t] + 1];
    for (int q = 0; q < nthreads; ++q) out->part_ptr[q + 1] += out->part_ptr[q];
    for (int q = 0; q < nthreads; ++q) cand[q] = out->part_ptr[q];
    int nq = 0;
    for (int t = 0; t < n; ++t) {
        if (stripped[t]) continue;
        if (owner[t] >= 0) out->part_rows[cand[owner[t]]++] = ROW_OF(t);
        else out->queue[nq++] = ROW_OF(t);
    }
    out->nqueue = nq;
end:
    free(stripped); free(parent); free(anc); free(first); free(next); free(w); free(owner); free(cand);
    if (rc) subpart_free(out);
    return rc;
}

/* The reference's choice of schedule for one direction (src/ilu_solver.cpp
 * :315-345): plain subtrees (relax 1); with a sequential sweep queue, the
 * factorization strips 20% of the rows (relax 1) when more than 10% are
 * queued, and a sweep strips 20%, then 50% (relax 0.7), while more than 2%
 * are queued. */
static int subpart_choose(int n, const int* rp, const int* ci, const int* dg, int upper, int nthreads,
                          int is_fact, int par_subs_queue, Subpart* out) {
    if (subpart_build(n, rp, ci, dg, upper, nthreads, 1, 1.0, -1.0, out)) return -1;
    if (par_subs_queue) return 0;
    if (is_fact) {
        if (out->nqueue <= 0.1 * n) return 0;
        subpart_free(out);
        return subpart_build(n, rp, ci, dg, upper, nthreads, 1, 1.0, 0.2, out);
    }
    if (out->nqueue <= 0.02 * n) return 0;
    subpart_free(out);
    if (subpart_build(n, rp, ci, dg, upper, nthreads, 1, 0.7, 0.2, out)) return -1;
    if (out->nqueue <= 0.02 * n) return 0;
    subpart_free(out);
    return subpart_build(n, rp, ci, dg, upper, nthreads, 1, 0.7, 0.5, out);
}

/* Row i of the IKJ ILU (oracle.c orc_ilu_factorize) with a dense work
 * row w; returns the zero pivots met. */
static inline int factor_row(int i, const int* rp, const int* ci, const int* dg, double* val, double* w) {
    int zero = 0;
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
    return zero;
}

/* Row IKJ ILU, rows of one L level in parallel; each thread keeps a dense
 * work row. */
static int factorize(int n, const int* rp, const int* ci, const int* dg, double* val, const Levels* L) {
    int zero = 0;
#pragma omp parallel reduction(+ : zero)
    {
        double* w = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
        for (int l = 0; l < L->nlev; ++l) {
#pragma omp for schedule(static)
            for (int k = L->ptr[l]; k < L->ptr[l + 1]; ++k) zero += factor_row(L->rows[k], rp, ci, dg, val, w);
        }
        free(w);
    }
    return zero;
}

static inline void sweep_row(int i, int upper, const int* rp, const int* ci, const int* dg, const double* val,
                             double* x) {
    double xi = x[i];
    if (!upper) {
        for (int p = rp[i]; p < dg[i]; ++p) xi = fma(-x[ci[p]], val[p], xi);
        x[i] = xi;
    } else {
        for (int p = dg[i] + 1; p < rp[i + 1]; ++p) xi = fma(-x[ci[p]], val[p], xi);
        x[i] = xi / val[dg[i]];
    }
}

/* One unit of the subtree schedule: the stripped levels in parallel, each
 * thread's subtrees without waits, then the queue -- in parallel with a
 * busy-wait on every dependency's done flag (par_queue), or on one thread
 * after the parallel part.  row(i, w) does the row's work; fact selects the
 * factorization's dependencies (its L part) over the sweep direction's. */
typedef struct {
    int n;
    const int *rp, *ci, *dg;
    double* val;
    double* x;
    int upper, fact;
} Work;

static inline int run_row(const Work* W, int i, double* w) {
    if (W->fact) return factor_row(i, W->rp, W->ci, W->dg, W->val, w);
    sweep_row(i, W->upper, W->rp, W->ci, W->dg, W->val, W->x);
    return 0;
}

static int run_subpart(const Work* W, const Subpart* S, int par_queue, _Atomic unsigned char* done) {
    const int n = W->n;
    const int *rp = W->rp, *ci = W->ci, *dg = W->dg;
    const int upper = W->fact ? 0 : W->upper;
    _Atomic int head = 0;
    int zero = 0;
#pragma omp parallel reduction(+ : zero)
    {
        double* w = W->fact ? (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)) : NULL;
        const int tid = omp_get_thread_num();
        if (par_queue) {
#pragma omp for schedule(static, 16)
            for (int j = 0; j < n; ++j) atomic_store_explicit(&done[j], 0, memory_order_relaxed);
        }
        for (int l = 0; l < S->nstrip; ++l) {
#pragma omp for schedule(static)
            for (int k = S->strip_ptr[l]; k < S->strip_ptr[l + 1]; ++k) {
                const int i = S->strip_rows[k];
                zero += run_row(W, i, w);
                if (par_queue) atomic_store_explicit(&done[i], 1, memory_order_release);
            }
        }
        if (tid < S->nthreads) {
            for (int k = S->part_ptr[tid]; k < S->part_ptr[tid + 1]; ++k) {
                const int i = S->part_rows[k];
                zero += run_row(W, i, w);
                if (par_queue) atomic_store_explicit(&done[i], 1, memory_order_release);
            }
            while (par_queue) {
                const int k = atomic_fetch_add_explicit(&head, 1, memory_order_relaxed);
                if (k >= S->nqueue) break;
                const int i = S->queue[k];
                const int b = upper ? dg[i] + 1 : rp[i], e = upper ? rp[i + 1] : dg[i];
                for (int p = b; p < e; ++p)
                    while (!atomic_load_explicit(&done[ci[p]], memory_order_acquire)) {
                    }
                zero += run_row(W, i, w);
                atomic_store_explicit(&done[i], 1, memory_order_release);
            }
        }
        free(w);
    }
    if (!par_queue) {
        double* w = W->fact ? (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)) : NULL;
        for (int k = 0; k < S->nqueue; ++k) zero += run_row(W, S->queue[k], w);
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

/* The three schedules of sched 1 and their shape. */
typedef struct {
    Subpart fact, l, u;
    int par_subs_queue;
    _Atomic unsigned char* done;
} SubSched;

static void subsched_free(SubSched* s) {
    subpart_free(&s->fact);
    subpart_free(&s->l);
    subpart_free(&s->u);
    free((void*)s->done);
}

/* The reference's set-up (src/ilu_solver.cpp:295-345): the factorization
 * on every thread, the sweeps on at most 8, their queues in parallel only
 * with at least 40 entries per row. */
static int subsched_build(int n, const int* rp, const int* ci, const int* dg, SubSched* s) {
    memset(s, 0, sizeof(*s));
    const int nt = omp_get_max_threads();
    const int st = nt < 8 ? nt : 8;
    const long nnz = n > 0 ? rp[n] : 0;
    s->par_subs_queue = !(st == 1 || (n > 0 && nnz / n < 40));
    s->done = (_Atomic unsigned char*)calloc((size_t)(n > 0 ? n : 1), 1);
    if (!s->done || subpart_choose(n, rp, ci, dg, 0, nt, 1, s->par_subs_queue, &s->fact) ||
        subpart_choose(n, rp, ci, dg, 0, st, 0, s->par_subs_queue, &s->l) ||
        subpart_choose(n, rp, ci, dg, 1, st, 0, s->par_subs_queue, &s->u)) {
        subsched_free(s);
        return -1;
    }
    return 0;
}

static void substitute_sub(int n, const int* rp, const int* ci, const int* dg, const double* val, const double* b,
                           double* x, const SubSched* S) {
    if (b != x) memcpy(x, b, (size_t)n * sizeof(double));
    Work W = {n, rp, ci, dg, (double*)val, x, 0, 0};
    run_subpart(&W, &S->l, S->par_subs_queue, S->done);
    W.upper = 1;
    run_subpart(&W, &S->u, S->par_subs_queue, S->done);
}

/* Shape of the subtree schedule of the pattern (for tests and reports):
 * info[0..2] = rows queued (factorization, L sweep, U sweep), info[3..5] =
 * stripped levels, info[6] = sweeps queue in parallel.  With check != 0,
 * also verifies that every row appears exactly once in each schedule and
 * that no row of a thread's subtrees depends on another thread's rows or on
 * a queued row (returns 1 then).  Returns 0 / 1, or -1 on failure. */
int baseline_subtree_info(int n, const int* rp, const int* ci, const int* dg, int* info, int check) {
    SubSched S;
    if (subsched_build(n, rp, ci, dg, &S)) return -1;
    const Subpart* parts[3] = {&S.fact, &S.l, &S.u};
    int ok = 1;
    for (int d = 0; d < 3; ++d) {
        info[d] = parts[d]->nqueue;
        info[3 + d] = parts[d]->nstrip;
    }
    info[6] = S.par_subs_queue;
    if (check) {
        int* where = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
        if (!where) { subsched_free(&S); return -1; }
        for (int d = 0; d < 3 && ok; ++d) {
            const Subpart* P = parts[d];
            const int upper = d == 2;
            for (int i = 0; i < n; ++i) where[i] = -3;
            int seen = 0;
            for (int k = 0; k < P->strip_ptr[P->nstrip]; ++k) { where[P->strip_rows[k]] = -2; ++seen; }
            for (int q = 0; q < P->nthreads; ++q)
                for (int k = P->part_ptr[q]; k < P->part_ptr[q + 1]; ++k) { where[P->part_rows[k]] = q; ++seen; }
            for (int k = 0; k < P->nqueue; ++k) { where[P->queue[k]] = -1; ++seen; }
            if (seen != n) ok = 0;
            for (int i = 0; i < n && ok; ++i) {
                if (where[i] == -3) ok = 0;
                if (where[i] < 0) continue;
                for (int p = upper ? dg[i] + 1 : rp[i]; p < (upper ? rp[i + 1] : dg[i]); ++p) {
                    const int j = where[ci[p]];
                    if (j != -2 && j != where[i]) ok = 0;
                }
            }
        }
        free(where);
    }
    subsched_free(&S);
    return check ? ok : 0;
}

/* times[0] schedule, [1] factorization, [2] GMRES (seconds); fval is
 * factorized in place.  sched 0: level sets; 1: the reference's subtree /
 * queue schedule.  Returns 1 converged, 0 not (or step cap reached), -1
 * allocation failure. */
int baseline_ilu_gmres(int n, const int* arp, const int* aci, const double* aval, const int* frp, const int* fci,
                       const int* fdg, double* fval, const double* rhs, double* sol, double tol, int m, int maxits,
                       int* its_out, double* times, int sched) {
    Levels L = {0, NULL, NULL}, U = {0, NULL, NULL};
    SubSched SS;
    memset(&SS, 0, sizeof(SS));
    double t0 = omp_get_wtime();
    if (sched == 1) {
        if (subsched_build(n, frp, fci, fdg, &SS)) return -1;
    } else if (levels_build(n, frp, fci, fdg, 0, &L) || levels_build(n, frp, fci, fdg, 1, &U)) {
        return -1;
    }
    double t1 = omp_get_wtime();
    if (sched == 1) {
        Work W = {n, frp, fci, fdg, fval, NULL, 0, 1};
        run_subpart(&W, &SS.fact, 1, SS.done);
    } else {
        factorize(n, frp, fci, fdg, fval, &L);
    }
    double t2 = omp_get_wtime();

    /* GMRES(m), src/gmres.cpp:138-288 (as oracle.c orc_gmres) */
    const double epsmac = 1e-16;
    const int im = m, im1 = m + 1;
    double* vv = (double*)malloc((size_t)im1 * n * sizeof(double));
    double* z = (double*)malloc((size_t)im * n * sizeof(double));
    double* hh = (double*)calloc((size_t)im1 * (im + 3), sizeof(double));
    int ret = 1, its = 0;
    if (!vv || !z || !hh) {
        free(vv); free(z); free(hh); levels_free(&L); levels_free(&U); subsched_free(&SS);
        return -1;
    }
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
            if (sched == 1) substitute_sub(n, frp, fci, fdg, fval, vi, zi, &SS);
            else substitute(n, frp, fci, fdg, fval, vi, zi, &L, &U);
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
    subsched_free(&SS);
    return ret;
}
