/* CPU oracle: a plain-C restatement of the reference algorithm
 * (lljbash/par_ilu0_cpp) for ILU(k)-preconditioned restarted GMRES.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/ as the checker, by the package's
 * smoke() as the checker, and by bench.py's cpu_baseline leg as the timed
 * CPU baseline.
 *
 * Every routine states the reference file:line it follows.  Floating-point
 * updates of the form `t -= a * b` are written as fma(-a, b, t), which is
 * what the reference's -O3 -march=native -ffast-math build emits for them;
 * everything else is evaluated in program order (this file is compiled with
 * -ffp-contract=off).
 *
 * Pinning (what checks each function against something other than itself):
 *  orc_iluk_symbolic  - hand-derived known-answer test for the first-found
 *                       rule, closed-form nnz counts of stencil patterns,
 *                       the reference's own src/iluk.cpp compiled into
 *                       oracle/_ref (bit-exact pattern and level-0 map),
 *                       and a brute-force O(n^2) Python restatement.
 *  orc_ilu_factorize  - (L U)_ij = A_ij on pattern(F); tridiagonal closed
 *                       form; exact dense LU (numpy) when lof >= bandwidth.
 *  orc_substitute     - dense triangular solves (numpy) and L(Ux) = b.
 *  orc_ilu0_guarded   - policy 0 equals orc_ilu_factorize; hand-worked
 *                       3x3/4x4 closed forms for the stop codes (-102,
 *                       -103, first failing row wins), the replaced pivot
 *                       and its effect on later rows, and -101/-106.
 *  orc_trsv_lanes     - L^-1 then U^-1 equals orc_substitute_lanes bit for
 *                       bit; L y = x and U y = x (scipy triangular solve);
 *                       alpha scaling exact.
 *  orc_spmv           - exact math.fsum dot products.
 *  orc_gmres          - invariants: monotone residual estimate within a
 *                       cycle, convergence in <= n steps with M = I and
 *                       m >= n, final true residual <= tol, the
 *                       max_iterations-at-restart-boundary quirk, and a
 *                       scipy.sparse.linalg.gmres solution cross-check.
 */
#ifndef PAR_ILU_ORACLE_H_
#define PAR_ILU_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Symbolic ILU(lof) with first-found levels (src/iluk.cpp:5-54).
 * Outputs are malloc'ed by the oracle (free with orc_free):
 *   *frp[n+1], *fci[nnzF], *flev[nnzF], *fval[nnzF]; nzmap[nnz(A)] optional.
 * Returns nnz(F), or -1 when the level-0 slot count disagrees with nnz(A)
 * (a repeated column in a row of A), where the reference calls exit(-1). */
int orc_iluk_symbolic(int n, const int* arp, const int* aci, const double* aval,
                      int lof, int** frp, int** fci, int** flev, double** fval,
                      int* nzmap);
void orc_free(void* p);

/* First entry of each row with column >= i (src/ilu_solver.cpp:276-289);
 * a row with no such entry gets rp[i+1]. */
void orc_diag_ptr(int n, const int* rp, const int* ci, int* diag);

/* In-place row IKJ ILU on the pattern (src/ilu_solver.cpp:553-589, the
 * threads==1 path :652-655).  Returns the number of exactly-zero pivots
 * divided by (the reference does not check; the factors then hold inf). */
int orc_ilu_factorize(int n, const int* rp, const int* ci, const int* diag,
                      double* val);

/* The same factorization with per-row diagonal compensation (modified ILU
 * with relaxation omega[i]): updates falling outside row i's pattern are
 * summed and omega[i] times the sum is added to a_ii when row i is done.
 * No reference counterpart (the learned-compensation extension). */
int orc_ilu_factorize_comp(int n, const int* rp, const int* ci, const int* diag,
                           double* val, const float* omega);

/* x = U^-1 L^-1 b, one-thread path (src/ilu_solver.cpp:775-796):
 * L ascending rows / ascending entries, U descending rows / ascending
 * entries, divide by the diagonal.  b may alias x. */
void orc_substitute(int n, const int* rp, const int* ci, const int* diag,
                    const double* val, const double* b, double* x);
/* orc_substitute with each row's sum split over g lanes in the GPU sweeps'
 * order (lane-strided fma chains, then a pairwise tree); g = 1 equals
 * orc_substitute. */
void orc_substitute_lanes(int n, const int* rp, const int* ci, const int* diag, const double* val,
                          const double* b, double* x, int g);

/* The factorization of the reference's USE_MKL_ILU build (dcsrilu0,
 * src/ilu_solver.cpp:248-259,631-639), restated from MKL's documented
 * contract: orc_ilu_factorize's elimination with each row's final pivot
 * checked -- policy 1 stops with -102 (zero) or -103 (|u_ii| < tiny),
 * policy 2 replaces it by repl, policy 0 does not check; -101 a row
 * without its diagonal, -106 columns not ascending.  Returns 0 or a code. */
int orc_ilu0_guarded(int n, const int* rp, const int* ci, const int* diag, double* val,
                     int policy, double tiny, double repl);

/* One triangle of orc_substitute_lanes, as the USE_MKL_SV build's
 * mkl_sparse_d_trsv calls (src/ilu_solver.cpp:753-773): upper 0,
 * y = alpha L^-1 x (unit lower); upper 1, y = alpha U^-1 x.  x may alias y. */
void orc_trsv_lanes(int n, const int* rp, const int* ci, const int* diag, const double* val,
                    int upper, double alpha, const double* x, double* y, int g);

/* y = alpha A x + beta y, row sums accumulated in entry order. */
void orc_spmv(int n, const int* rp, const int* ci, const double* val,
              double alpha, const double* x, double beta, double* y);

/* Restarted right-preconditioned GMRES(m), src/gmres.cpp:138-288.
 * Preconditioner: the ILU factors (frp, fci, fdiag, fval) through
 * orc_substitute, or the identity when frp is NULL.
 * hist (optional, capacity hist_cap) receives |rs[i+1]| after each step.
 * Returns 1 (converged / success) or 0; *its gets the step count,
 * *nhist the number of history entries written. */
int orc_gmres(int n, const int* arp, const int* aci, const double* aval,
              const int* frp, const int* fci, const int* fdiag, const double* fval,
              const double* rhs, double* sol, double tol, int m, int maxits,
              int* its, double* hist, int hist_cap, int* nhist);

#ifdef __cplusplus
}
#endif

#endif /* PAR_ILU_ORACLE_H_ */
