/* Extension C ABI of the MI355X build (no counterpart in the reference).
 *
 * These LLJBASH_*Ex entry points add what the reference interface cannot
 * express: device-resident data, FP32 factor storage, sweep launch shapes
 * and diagnostics.  The reference's own structs and functions are left
 * unchanged.  "Device" pointers are HIP device allocations on the current
 * device; work is ordered on the solver's stream (the null stream), so
 * data produced by earlier null-stream work is ready.  Return values: 0 (or
 * a count) on success, -1 on failure, with a message on stderr.
 */
#ifndef LLJBASH_PAR_ILU_EX_H_
#define LLJBASH_PAR_ILU_EX_H_

#include "gmres_c.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 64 (default) or 32: precision of the stored L/U factors used by the
 * sweeps; vectors, Arnoldi and the residual stay FP64. */
#define IluSolverSetFactorPrecisionEx LLJBASH_DECL(IluSolverSetFactorPrecisionEx)
int IluSolverSetFactorPrecisionEx(IluSolverHdl hdl, int bits);

/* Lanes per row of the forward / backward sweep (0 = heuristic, else a
 * power of two <= 64). */
#define IluSolverSetSweepLanesEx LLJBASH_DECL(IluSolverSetSweepLanesEx)
int IluSolverSetSweepLanesEx(IluSolverHdl hdl, int lanes_l, int lanes_u);

/* Per-row diagonal compensation for the following factorizations: device
 * array of n floats (kept by pointer; NULL disables).  Row i adds
 * omega[i] * (sum of the fill updates dropped from row i) to its pivot. */
#define IluSolverSetDiagonalCompensationEx LLJBASH_DECL(IluSolverSetDiagonalCompensationEx)
int IluSolverSetDiagonalCompensationEx(IluSolverHdl hdl, const float* d_omega);

/* Pivot check of the following factorizations, as MKL's dcsrilu0
 * documents it (replaces the reference's USE_MKL_ILU factorization,
 * src/ilu_solver.cpp:248-259,631-639, which calls dcsrilu0 with
 * ipar[30] = 0): policy 0 none (default); 1 a pivot that is zero or below
 * `tiny` in magnitude makes Factorize fail (dpar[30] = tiny, 1e-16 in
 * dcsrilu0); 2 such a pivot is replaced by `replacement` (ipar[30] != 0,
 * dpar[31] = replacement, 1e-10 in dcsrilu0) and the factorization goes on.
 * Factors after a policy-1 failure are unspecified. */
#define IluSolverSetPivotPolicyEx LLJBASH_DECL(IluSolverSetPivotPolicyEx)
int IluSolverSetPivotPolicyEx(IluSolverHdl hdl, int policy, double tiny, double replacement);

/* dcsrilu0's ierr for the last factorization under policy 1: 0, -102 (the
 * first failing pivot, by row, is zero) or -103 (it is below tiny); -1 on
 * a bad handle.  (-101 and -106, a missing diagonal and unsorted columns,
 * are refused earlier by IluSolverSetup with -2 and -1.) */
#define IluSolverFactorStatusEx LLJBASH_DECL(IluSolverFactorStatusEx)
int IluSolverFactorStatusEx(IluSolverHdl hdl);

/* One triangular solve with the factors, replacing the reference's
 * mkl_sparse_d_trsv calls (USE_MKL_SV, src/ilu_solver.cpp:753-773):
 * upper 0, y = alpha L^-1 x with L the unit lower triangle; upper 1,
 * y = alpha U^-1 x with U the upper triangle and its diagonal.  Device
 * vectors of n doubles; x and y are the same vector or do not overlap. */
#define IluSolverTrsvDeviceEx LLJBASH_DECL(IluSolverTrsvDeviceEx)
int IluSolverTrsvDeviceEx(IluSolverHdl hdl, int upper, double alpha, const double* d_x, double* d_y);

/* The same with host vectors (copied to and from the device). */
#define IluSolverTrsvEx LLJBASH_DECL(IluSolverTrsvEx)
int IluSolverTrsvEx(IluSolverHdl hdl, int upper, double alpha, const double* x, double* y);

/* Wall-clock limit in seconds (0..3600, default 10) of every bounded
 * dependency wait in the analysis, factorization and sweep kernels; the
 * symbolic pass waits up to twice as long.  A wait past it makes the call
 * fail (-1) after every kernel has drained; the solver then refuses sweeps
 * until Setup / Factorize complete again. */
#define IluSolverSetWatchdogEx LLJBASH_DECL(IluSolverSetWatchdogEx)
int IluSolverSetWatchdogEx(IluSolverHdl hdl, double seconds);

/* Autotune the sweeps after Factorize: level-scheduled rows with 1..64
 * lanes per row, the single-workgroup level walk with 4..64 lanes (where
 * levels average at most 256 rows) and the block-local sweeps (contiguous blocks of
 * 4096/8192/16384 rows, and, once the grid shape is declared, tiles spanning
 * the x extent (up to 128 cells) and several planes; CSR factors, or packed
 * factors with level barriers or LDS value-flags) with 2..32 lanes are each
 * timed over `reps` applies; the fastest shape per direction is kept.  table (optional, cap rows x 8
 * doubles): mode, block rows, tile x, tile y, tile z, lanes, forward ms,
 * backward ms.  Returns the rows timed. */
#define IluSolverAutotuneEx LLJBASH_DECL(IluSolverAutotuneEx)
int IluSolverAutotuneEx(IluSolverHdl hdl, int reps, double* table, int cap);

/* Explicit sweep algorithm per direction (0 level rows, 1 block-local over
 * contiguous blocks of blk_rows rows, 2 block-local over packed factors,
 * software-pipelined, with a workgroup barrier per local level, 3 the same
 * with the block's x as LDS value-flags and no barriers, 4 one workgroup
 * walking the level schedule with a barrier per level, for deep schedules
 * with few rows per level, 5 block-local with one solver, one record-stream
 * loader and one importer wave per block, 1-16 lanes per row; blk_rows is
 * ignored by modes 0 and 4) and the current configuration (cfg[8]: mode_l,
 * lanes_l, mode_u, lanes_u, blk_rows, tile x, tile y, tile z). */
#define IluSolverSetSweepModeEx LLJBASH_DECL(IluSolverSetSweepModeEx)
int IluSolverSetSweepModeEx(IluSolverHdl hdl, int mode_l, int mode_u, int blk_rows);
/* Declares the natural-order grid of the matrix (row x + nx (y + ny z));
 * nx * ny * nz must equal n.  Enables tiled block-local sweeps. */
#define IluSolverSetGridShapeEx LLJBASH_DECL(IluSolverSetGridShapeEx)
int IluSolverSetGridShapeEx(IluSolverHdl hdl, int nx, int ny, int nz);
/* Block-local sweeps (mode 0, 2 or 3 per direction) over tx x ty x tz tiles
 * of the declared grid (at most 16384 rows per tile). */
#define IluSolverSetSweepTilesEx LLJBASH_DECL(IluSolverSetSweepTilesEx)
int IluSolverSetSweepTilesEx(IluSolverHdl hdl, int mode_l, int mode_u, int tx, int ty, int tz);
#define IluSolverGetSweepConfigEx LLJBASH_DECL(IluSolverGetSweepConfigEx)
int IluSolverGetSweepConfigEx(IluSolverHdl hdl, int* cfg);

/* Number of levels of the L and U dependency graphs (after setup). */
/* IluSolver::DescribeSchedule as a NUL-terminated JSON string in buf
 * (truncated to cap - 1 characters); returns the full length, -1 on error. */
#define IluSolverScheduleJsonEx LLJBASH_DECL(IluSolverScheduleJsonEx)
int IluSolverScheduleJsonEx(IluSolverHdl hdl, char* buf, int cap);

#define IluSolverGetLevelsEx LLJBASH_DECL(IluSolverGetLevelsEx)
int IluSolverGetLevelsEx(IluSolverHdl hdl, int* nlev_l, int* nlev_u);

/* Symbolic ILU(lof) of a device CSR `d_a` (row_ptr/col_idx device arrays,
 * size = n) into the solver, followed by the same analysis as Setup.
 * Returns nnz(F), -1 on a repeated column, -2 on a missing diagonal. */
#define IluSolverSymbolicDeviceEx LLJBASH_DECL(IluSolverSymbolicDeviceEx)
int IluSolverSymbolicDeviceEx(IluSolverHdl hdl, const CsrMatrix* d_a, int lof);

/* Factorize on the device.  values_are_a != 0: d_values are A's values
 * (nnz(A), mapped through the SymbolicDeviceEx map, fill = 0); otherwise
 * d_values are F's values (nnz(F)).  Returns 0, or -1 (also on a zero
 * pivot). */
#define IluSolverFactorizeDeviceEx LLJBASH_DECL(IluSolverFactorizeDeviceEx)
int IluSolverFactorizeDeviceEx(IluSolverHdl hdl, const double* d_values, int values_are_a);

/* x = U^-1 L^-1 b on device vectors (b may alias x). */
#define IluSolverSubstituteDeviceEx LLJBASH_DECL(IluSolverSubstituteDeviceEx)
int IluSolverSubstituteDeviceEx(IluSolverHdl hdl, const double* d_b, double* d_x);

/* The apply of IluSolverSubstituteDeviceEx without any device allocation,
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops).  No counterpart in the reference.  The forward sweep's
 * intermediate vector, the ticket counters and the level walk's progress
 * words live in one static device array per GPU (32 MiB, held until the
 * process ends), so n must not exceed 4194304 rows (-1 otherwise).  The
 * sweep structures of the current sweep configuration must already exist:
 * run IluSolverSubstituteDeviceEx once after Setup/Factorize or after a
 * change of sweep shape (-1 otherwise).  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (several host threads or
 * streams) race.  Results are bit-identical to IluSolverSubstituteDeviceEx. */
#define IluSolverSubstituteNoAllocEx LLJBASH_DECL(IluSolverSubstituteNoAllocEx)
int IluSolverSubstituteNoAllocEx(IluSolverHdl hdl, const double* d_b, double* d_x);

/* Copies of the device factors (nnz(F) doubles) and fill levels
 * (nnz(F) bytes, after SymbolicDeviceEx) to host memory. */
#define IluSolverCopyFactorsEx LLJBASH_DECL(IluSolverCopyFactorsEx)
int IluSolverCopyFactorsEx(IluSolverHdl hdl, double* out);
#define IluSolverCopyFillLevelsEx LLJBASH_DECL(IluSolverCopyFillLevelsEx)
int IluSolverCopyFillLevelsEx(IluSolverHdl hdl, signed char* out);
#define IluSolverNonzerosEx LLJBASH_DECL(IluSolverNonzerosEx)
int IluSolverNonzerosEx(IluSolverHdl hdl);

/* Symbolic ILU(lof) on host CSR like IlukSymbolic, also returning the level
 * of every entry of F in `levels` (caller-provided, capacity `cap`). */
#define IlukSymbolicLevelsEx LLJBASH_DECL(IlukSymbolicLevelsEx)
int IlukSymbolicLevelsEx(const CsrMatrix* a, int lof, CsrMatrix* f, int* nzmap, signed char* levels, long cap);

/* GMRES on device data (d_mat arrays on the device). */
#define GmresSolveDeviceEx LLJBASH_DECL(GmresSolveDeviceEx)
bool GmresSolveDeviceEx(GmresHdl hdl, const CsrMatrix* d_mat, const double* d_rhs, double* d_sol, int* iter);

/* ||b - A x||_2 of the last solve's result (FP64, on the device). */
#define GmresLastTrueResidualEx LLJBASH_DECL(GmresLastTrueResidualEx)
double GmresLastTrueResidualEx(GmresHdl hdl);

/* Residual estimates |rs[i+1]| after each Arnoldi step of the last solve. */
#define GmresResidualHistoryEx LLJBASH_DECL(GmresResidualHistoryEx)
int GmresResidualHistoryEx(GmresHdl hdl, double* out, int cap);

/* ---- batched solver: many independent systems x several RHS ----
 * The systems are passed as one block-diagonal device CSR matrix d_a
 * (system s owns rows sys_ptr[s] .. sys_ptr[s+1]-1, sys_ptr on the host);
 * right-hand sides and solutions are RHS-interleaved device arrays
 * X[row * nrhs + r], nrhs in {1, 2, 4, 8}.  Every (system, RHS) pair runs
 * the reference's GMRES(m) with tolerance rel_tol * ||b||_2 of its own RHS;
 * one restart cycle of all pairs is one HIP graph launch. */
#define BatchGmresHdl LLJBASH_DECL(BatchGmresHdl)
typedef void* BatchGmresHdl;
#define BatchGmresCreateEx LLJBASH_DECL(BatchGmresCreateEx)
BatchGmresHdl BatchGmresCreateEx(void);
#define BatchGmresDestroyEx LLJBASH_DECL(BatchGmresDestroyEx)
void BatchGmresDestroyEx(BatchGmresHdl hdl);
/* Symbolic ILU(lof) of the block-diagonal matrix plus analysis; returns nnz(F). */
#define BatchGmresSetupEx LLJBASH_DECL(BatchGmresSetupEx)
int BatchGmresSetupEx(BatchGmresHdl hdl, const CsrMatrix* d_a, int lof, int nsys, const int* sys_ptr, int nrhs);
#define BatchGmresSetFactorPrecisionEx LLJBASH_DECL(BatchGmresSetFactorPrecisionEx)
int BatchGmresSetFactorPrecisionEx(BatchGmresHdl hdl, int bits);
/* Factorize every system from A's device values (one launch). */
#define BatchGmresFactorizeEx LLJBASH_DECL(BatchGmresFactorizeEx)
int BatchGmresFactorizeEx(BatchGmresHdl hdl, const double* d_avalues);
/* Solve all pairs; d_sol holds the initial guesses.  Host outputs (length
 * nsys * nrhs, pair index s * nrhs + r): step counts, the reference's
 * success flag, FP64 true relative residuals.  Returns the number of
 * restart cycles run, or -1. */
#define BatchGmresSolveEx LLJBASH_DECL(BatchGmresSolveEx)
int BatchGmresSolveEx(BatchGmresHdl hdl, const CsrMatrix* d_a, const double* d_rhs, double* d_sol, double rel_tol,
                      int m, int maxits, int* its, int* converged, double* true_rel_res);

/* ---- learned diagonal compensation (3-layer message-passing network,
 * hidden width 32, BF16 MFMA with FP32 accumulation) ----
 * The network reads a device matrix and predicts omega (n floats, device)
 * for IluSolverSetDiagonalCompensationEx.  Parameters are a flat FP32
 * vector of GnnParameterCountEx() values. */
#define GnnHdl LLJBASH_DECL(GnnHdl)
typedef void* GnnHdl;
#define GnnCreateEx LLJBASH_DECL(GnnCreateEx)
GnnHdl GnnCreateEx(void);
#define GnnDestroyEx LLJBASH_DECL(GnnDestroyEx)
void GnnDestroyEx(GnnHdl hdl);
#define GnnParameterCountEx LLJBASH_DECL(GnnParameterCountEx)
int GnnParameterCountEx(void);
#define GnnSetParametersEx LLJBASH_DECL(GnnSetParametersEx)
int GnnSetParametersEx(GnnHdl hdl, const float* host_params);
#define GnnGetParametersEx LLJBASH_DECL(GnnGetParametersEx)
int GnnGetParametersEx(GnnHdl hdl, float* host_params);
/* omega = network(A); d_a holds device arrays (values included). */
#define GnnPredictEx LLJBASH_DECL(GnnPredictEx)
int GnnPredictEx(GnnHdl hdl, const CsrMatrix* d_a, float* d_omega);
/* Gradients of the mean squared error between omega and d_target (n
 * floats).  d_at: device CSR pattern of A^T; d_tmap[k]: index in A of the
 * k-th entry of A^T.  host_grads (optional) receives the gradient. */
#define GnnGradientsEx LLJBASH_DECL(GnnGradientsEx)
int GnnGradientsEx(GnnHdl hdl, const CsrMatrix* d_a, const CsrMatrix* d_at, const int* d_tmap, const float* d_target,
                   float* host_grads, double* loss);
/* One Adam step (lr) on that loss. */
#define GnnTrainStepEx LLJBASH_DECL(GnnTrainStepEx)
int GnnTrainStepEx(GnnHdl hdl, const CsrMatrix* d_a, const CsrMatrix* d_at, const int* d_tmap, const float* d_target,
                   float lr, double* loss);
/* Learned solve with fallback: omega = network(A), compensated factorization
 * of the solver's pattern (set up by IluSolverSymbolicDeviceEx) from A's
 * values, GMRES from d_sol.  If that solve fails or needs more than
 * `budget` steps, the plain factorization is solved again from the same
 * initial guess and *fallback is set.  Returns the converged flag. */
#define GmresSolveLearnedDeviceEx LLJBASH_DECL(GmresSolveLearnedDeviceEx)
bool GmresSolveLearnedDeviceEx(GmresHdl gmres, IluSolverHdl ilu, GnnHdl gnn, const CsrMatrix* d_a,
                               const double* d_rhs, double* d_sol, int budget, int* iter, int* fallback);

/* 1 if a HIP device is usable from this process, else 0. */
#define DeviceAvailableEx LLJBASH_DECL(DeviceAvailableEx)
int DeviceAvailableEx(void);

#ifdef __cplusplus
}
#endif

#endif /* LLJBASH_PAR_ILU_EX_H_ */
