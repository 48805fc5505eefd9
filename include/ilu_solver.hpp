// ILU preconditioner on the GPU.
//
// Drop-in for include/ilu_solver.hpp of lljbash/par_ilu0_cpp (lines 8-30).
// The user fills the factor matrix F through GetMatrix() (CopyCsrMatrix for
// ILU(0), IlukSymbolic for ILU(k)), then calls SetupMatrix() once per
// sparsity pattern and Factorize() once per value set.  Factorize() leaves
// the factors in F.value in place, as the reference does
// (src/ilu_solver.cpp:553-589): L strictly lower with an implied unit
// diagonal, U upper including the diagonal.
//
// The work happens on the current HIP device.  Host-pointer methods copy in
// and out; the *Device methods work on device-resident data and are the
// extension surface used by the GMRES solver and by batched callers.
#pragma once

#include <string>
#include "csr_matrix.h"

namespace lljbash {

class IluSolver {
public:
    IluSolver() = default;
    ~IluSolver();
    IluSolver(const IluSolver&) = delete;
    IluSolver& operator=(const IluSolver&) = delete;

    const CsrMatrix& GetILUMatrix() const { return aMatrix_; }
    CsrMatrix* GetMatrix() { return &aMatrix_; }

    // Kept for interface compatibility: the reference sets its OpenMP team
    // size here (src/ilu_solver.cpp:228-237).  The GPU path ignores it.
    void SetThreads(int threads) { threads_ = threads; }

    // Uploads the pattern, locates diagonals and builds the level schedules
    // of the L and U dependency graphs.  Throws std::runtime_error("missing
    // diag") if a row lacks its diagonal entry, std::runtime_error for other
    // failures.  reference: src/ilu_solver.cpp:225-373
    void SetupMatrix();

    // Numeric ILU on F's pattern; uploads F.value, factorizes on the GPU and
    // copies the factors back into F.value.  Returns false if a zero pivot
    // was met (the factors then hold inf/nan, as the reference's would).
    // reference: src/ilu_solver.cpp:616-671
    bool Factorize();

    // x = U^-1 L^-1 b with host vectors; b may alias x.
    // reference: src/ilu_solver.cpp:736-831
    void Substitute(const double* b, double* x);

    // ---- extension surface (no counterpart in the reference) ----------
    // Factorize from values already on the device (length nnz(F)); factors
    // stay on the device, F.value on the host is not touched.
    bool FactorizeDevice(const double* d_values);
    // Apply the preconditioner to device vectors; b may alias x.
    void SubstituteDevice(const double* d_b, double* d_x);
    // The same apply without allocating device memory: the intermediate
    // vector and the sweeps' counters live in a static per-device scratch
    // (n <= 4194304).  The sweep structures of the current configuration
    // must exist (one SubstituteDevice call builds them); otherwise it
    // throws.  Calls on one device share the scratch and must not overlap.
    void SubstituteDeviceNoAlloc(const double* d_b, double* d_x);
    // 64 (default) or 32: precision in which the factors are stored for the
    // triangular sweeps.  Vectors and accumulation stay FP64.
    void SetFactorPrecision(int bits);
    int FactorPrecision() const;
    // Lanes cooperating on one row in the L and U sweeps (1..64, power of
    // two); 0 picks a value from the average row length.
    void SetSweepLanes(int lanes_l, int lanes_u);
    // Wall-clock limit (seconds, 0..3600; default 10) of every bounded
    // dependency wait in the analysis, factorization and sweep kernels; the
    // symbolic pass waits up to twice as long.  A wait past the limit gives
    // up, the kernel drains and the call throws; a solver whose analysis or
    // factorization timed out refuses further sweeps until it is redone.
    void SetWatchdog(double seconds);
    // Pivot check of the following factorizations, as MKL's dcsrilu0
    // documents it (the reference's USE_MKL_ILU build sets ipar[30] = 0):
    // policy 0 none (the default build's behaviour); 1 a pivot that is zero
    // or below `tiny` in magnitude fails Factorize (ipar[30] = 0, dpar[30]
    // = tiny); 2 such a pivot is replaced by `replacement` and the
    // factorization goes on (ipar[30] != 0, dpar[31] = replacement).
    // reference: src/ilu_solver.cpp:248-259,631-639
    void SetPivotPolicy(int policy, double tiny, double replacement);
    // dcsrilu0-style code of the last factorization under policy 1: 0, -102
    // (the first failing pivot, by row, is zero) or -103 (it is below tiny).
    int FactorStatus() const;
    // One triangular solve with the factors, as mkl_sparse_d_trsv
    // (non-transposed; the reference's USE_MKL_SV build applies the
    // preconditioner as two of them): upper false, y = alpha L^-1 x with L
    // the unit lower triangle; upper true, y = alpha U^-1 x with U the upper
    // triangle and its diagonal.  x and y are the same vector or do not
    // overlap.  reference: src/ilu_solver.cpp:753-773
    void TrsvDevice(bool upper, double alpha, const double* d_x, double* d_y);
    // The same with host vectors.
    void Trsv(bool upper, double alpha, const double* x, double* y);
    // Per-row diagonal compensation used by the next factorizations:
    // d_omega (device, n floats, caller-owned; null turns it off).  Row i
    // adds omega[i] times the sum of its dropped fill updates to its pivot.
    void SetDiagonalCompensation(const float* d_omega);
    // Times the sweeps in every launch shape -- level-scheduled rows with
    // 1..64 lanes per row, and block-local sweeps over contiguous blocks of
    // 4096, 8192 or 16384 rows (CSR or packed factors) and, once the grid
    // shape is declared, over tiles spanning the x extent (up to 128 cells)
    // and several planes (packed factors) -- over `reps` applies each, and
    // keeps the fastest shape for each direction.
    // Results depend on the lanes per row only through the rounding of row
    // sums.  table (optional, cap rows of 8 doubles): mode, block rows, tile
    // x, y, z, lanes, forward ms, backward ms.  Returns the shapes timed.
    int AutotuneSweeps(int reps, double* table, int cap);
    // Sweep algorithm per direction: 0 level-scheduled rows, 1 block-local
    // over contiguous blocks of blk_rows rows (64..16384), 2 block-local over
    // a packed, schedule-ordered copy of the factors (refreshed at every
    // Factorize) with its loads issued several passes ahead and a workgroup
    // barrier per local level, 3 the same with the block's x held in LDS as
    // value-flags instead of barriers, 4 one workgroup walking the level
    // schedule with a barrier per level (no hand-offs between workgroups;
    // for deep schedules with few rows per level), 5 block-local with the
    // work of a block split over a solver wave (LDS only), a wave streaming
    // the block's records into an LDS ring and a wave importing its
    // out-of-block values (1-16 lanes per row).
    void SetSweepMode(int mode_l, int mode_u, int blk_rows);
    // Declares that row x + nx (y + ny z) is grid cell (x, y, z); enables
    // tiled partitions of the block-local sweeps.
    void SetGridShape(int nx, int ny, int nz);
    // Block-local sweeps (mode 0, 2, 3 or 5 per direction) over tx x ty x tz
    // tiles of the declared grid (at most 16384 rows per tile).  Blocks run
    // in a topological order of their dependency graph; the results are
    // those of the same lanes per row with any other partition.
    void SetSweepTiles(int mode_l, int mode_u, int tx, int ty, int tz);
    // cfg[8]: mode_l, lanes_l, mode_u, lanes_u, blk_rows, tile x, y, z.
    void GetSweepConfig(int* cfg) const;
    // Device pointers of the current pattern and factors (valid after
    // SetupMatrix / Factorize): row_ptr, col_idx, diag_ptr, factor values.
    const int* DeviceRowPtr() const;
    const int* DeviceColIdx() const;
    const int* DeviceDiagPtr() const;
    const double* DeviceFactors() const;
    int NumLevelsL() const;
    int NumLevelsU() const;
    // Make the solver use the pattern of a device-resident F; the pattern
    // arrays are copied.  Same checks and schedules as SetupMatrix().
    void SetupMatrixDevice(int n, int nnz, const int* d_row_ptr, const int* d_col_idx);
    // Symbolic ILU(lof) of a device-resident A straight into this solver
    // (pattern, schedules and the A-to-F value map stay on the device);
    // returns nnz(F), or -1 for a row of A with a repeated column.
    int SymbolicDevice(int n, const int* d_row_ptr, const int* d_col_idx, int lof);
    // Factorize from A's device values through the map kept by
    // SymbolicDevice (fill entries start at zero).
    bool FactorizeDeviceFromA(const double* d_a_values);
    // Copy the level of every entry of F (from the last SymbolicDevice) to
    // the host; returns nnz(F).
    int CopyFillLevels(signed char* out) const;
    // The schedules and sweep configuration as one JSON object: levels per
    // direction, block rows and blocks, sweep mode and lanes per direction,
    // packed entries (with padding) and passes.  Printed to stdout after
    // SetupMatrix and after a sweep-shape change when the environment
    // variable PAR_ILU_SHOW_ALGORITHM is set (the reference prints its
    // schedule under the SHOW_ALGORITHM macro, src/ilu_solver.cpp:196-223).
    std::string DescribeSchedule() const;
    int Size() const;
    int Nonzeros() const;

    struct Ext;
    // Library-internal device state (opaque to users).
    Ext* DeviceState() const { return ext_; }
private:
    int              threads_ = 0;
    CsrMatrix        aMatrix_;
    Ext* ext_ = nullptr;
    struct ThreadLocalExt;
    ThreadLocalExt* extt_ = nullptr;  // unused; keeps the reference's layout
};

} // namespace lljbash
