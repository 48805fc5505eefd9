/* A C host of the drop-in libraries, written only against include/ and
 * linked to libpar-ilu0-gmres_c.so / libpar-ilu0-gmres.so.  It runs the
 * flow of the reference's C driver (test/test.c:6-62): read a 1-based
 * MatrixMarket file, copy it into the ILU solver's matrix (ILU(0)),
 * b = A * ones, set up, factorize, GMRES with the default parameters from
 * x0 = 0, error against the exact solution.
 *
 * With a second argument "mkl" it also takes the path of the reference's
 * USE_MKL build through the Ex ABI: dcsrilu0's pivot check (ipar[30] = 0)
 * before the factorization, and y = U^-1 (L^-1 b) as two single triangular
 * solves after it.
 *
 * Output lines: "setup <rc>", "factorize <rc>", ["factor_status <code>",
 * "trsv <y[0]> <y[n/2]> <y[n-1]> <max |y|>",] "iterations <n>",
 * "converged <0|1>", "max_error <e>", "timers <init> <fgmr> <mv> <precon>".
 * Exit status: 0 converged; 2 GMRES failed before max_iterations; 3 not
 * converged; 4 setup or factorization failed; 5 a triangular solve failed;
 * 1 usage. */
#include <math.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gmres_c.h"
#include "par_ilu_ex.h"

int main(int argc, char** argv) {
    if (argc < 2 || argc > 3 || (argc == 3 && strcmp(argv[2], "mkl") != 0)) {
        fprintf(stderr, "usage: %s matrix.mtx [mkl]\n", argv[0]);
        return 1;
    }
    const bool mkl = argc == 3;
    CsrMatrix a = CSR_MATRIX_DEFAULT;
    ReadCsrMatrixMM1(&a, argv[1]);
    const int n = a.size;

    IluSolverHdl ilu = IluSolverCreate(0);
    CopyCsrMatrix(IluSolverGetMatrix(ilu), &a);

    double* ones = (double*)malloc(sizeof(double) * (size_t)n);
    double* b = (double*)malloc(sizeof(double) * (size_t)n);
    double* x = (double*)calloc((size_t)n, sizeof(double));
    for (int i = 0; i < n; ++i) ones[i] = 1.0;
    CsrMatVec(&a, ones, b);

    int status = 0;
    GmresHdl gmres = NULL;
    const int rs = IluSolverSetup(ilu);
    printf("setup %d\n", rs);
    if (mkl && rs == 0) IluSolverSetPivotPolicyEx(ilu, 1, 1e-16, 1e-10);
    const int rf = rs == 0 ? IluSolverFactorize(ilu) : -1;
    printf("factorize %d\n", rf);
    if (mkl) printf("factor_status %d\n", IluSolverFactorStatusEx(ilu));
    if (rs != 0 || rf != 0) {
        status = 4;
        goto done;
    }
    if (mkl) {
        double* t = (double*)malloc(sizeof(double) * (size_t)n);
        double* y = (double*)malloc(sizeof(double) * (size_t)n);
        const int r1 = IluSolverTrsvEx(ilu, 0, 1.0, b, t);
        const int r2 = IluSolverTrsvEx(ilu, 1, 1.0, t, y);
        double big = 0.0;
        for (int i = 0; i < n; ++i) big = fmax(big, fabs(y[i]));
        printf("trsv %.17e %.17e %.17e %.17e\n", y[0], y[n / 2], y[n - 1], big);
        free(t);
        free(y);
        if (r1 != 0 || r2 != 0) {
            status = 5;
            goto done;
        }
    }

    gmres = GmresCreate();
    GmresSetPreconditioner(gmres, ilu);
    const GmresParameters* prm = GmresGetParameters(gmres);
    int iterations = 0;
    const bool converged = GmresSolve(gmres, &a, b, x, &iterations);
    printf("iterations %d\nconverged %d\n", iterations, converged ? 1 : 0);
    if (!converged) {
        status = iterations < prm->max_iterations ? 2 : 3;
    } else {
        double worst = 0.0;
        for (int i = 0; i < n; ++i) worst = fmax(worst, fabs(x[i] - ones[i]));
        printf("max_error %.6e\n", worst);
    }
    const GmresStat* st = GmresGetStat(gmres);
    printf("timers %.6e %.6e %.6e %.6e\n", st->total_init_time, st->total_fgmr_time, st->total_mv_time,
           st->total_precon_time);

done:
    if (gmres) GmresDestroy(gmres);
    free(x);
    free(b);
    free(ones);
    DestroyCsrMatrix(&a);
    IluSolverDestroy(ilu);
    return status;
}
