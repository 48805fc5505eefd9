// Extension C ABI (include/par_ilu_ex.h); no counterpart in the reference.
#include "par_ilu_ex.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <exception>

#include "gmres.hpp"
#include "iluk.h"

using lljbash::IluSolver;
using lljbash::PreconditionedGmres;

namespace {

IluSolver* S(IluSolverHdl h) { return static_cast<IluSolver*>(h); }
PreconditionedGmres* G(GmresHdl h) { return static_cast<PreconditionedGmres*>(h); }

template <typename F>
int guarded(const char* what, F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s: %s\n", what, e.what());
        return std::strcmp(e.what(), "missing diag") == 0 ? -2 : -1;
    }
}

}  // namespace

extern "C" {

int IluSolverSetFactorPrecisionEx(IluSolverHdl hdl, int bits) {
    return guarded("SetFactorPrecision", [&] { S(hdl)->SetFactorPrecision(bits); return 0; });
}

int IluSolverSetSweepLanesEx(IluSolverHdl hdl, int lanes_l, int lanes_u) {
    return guarded("SetSweepLanes", [&] { S(hdl)->SetSweepLanes(lanes_l, lanes_u); return 0; });
}

int IluSolverSetDiagonalCompensationEx(IluSolverHdl hdl, const float* d_omega) {
    return guarded("SetDiagonalCompensation", [&] { S(hdl)->SetDiagonalCompensation(d_omega); return 0; });
}

int IluSolverSetPivotPolicyEx(IluSolverHdl hdl, int policy, double tiny, double replacement) {
    return guarded("SetPivotPolicy", [&] { S(hdl)->SetPivotPolicy(policy, tiny, replacement); return 0; });
}

int IluSolverFactorStatusEx(IluSolverHdl hdl) {
    return guarded("FactorStatus", [&] { return S(hdl)->FactorStatus(); });
}

int IluSolverTrsvDeviceEx(IluSolverHdl hdl, int upper, double alpha, const double* d_x, double* d_y) {
    return guarded("TrsvDevice", [&] { S(hdl)->TrsvDevice(upper != 0, alpha, d_x, d_y); return 0; });
}

int IluSolverTrsvEx(IluSolverHdl hdl, int upper, double alpha, const double* x, double* y) {
    return guarded("Trsv", [&] { S(hdl)->Trsv(upper != 0, alpha, x, y); return 0; });
}

int IluSolverSetWatchdogEx(IluSolverHdl hdl, double seconds) {
    return guarded("SetWatchdog", [&] { S(hdl)->SetWatchdog(seconds); return 0; });
}

int IluSolverAutotuneEx(IluSolverHdl hdl, int reps, double* table, int cap) {
    return guarded("Autotune", [&] { return S(hdl)->AutotuneSweeps(reps, table, cap); });
}

int IluSolverSetSweepModeEx(IluSolverHdl hdl, int mode_l, int mode_u, int blk_rows) {
    return guarded("SetSweepMode", [&] { S(hdl)->SetSweepMode(mode_l, mode_u, blk_rows); return 0; });
}

int IluSolverSetGridShapeEx(IluSolverHdl hdl, int nx, int ny, int nz) {
    return guarded("SetGridShape", [&] { S(hdl)->SetGridShape(nx, ny, nz); return 0; });
}

int IluSolverSetSweepTilesEx(IluSolverHdl hdl, int mode_l, int mode_u, int tx, int ty, int tz) {
    return guarded("SetSweepTiles", [&] { S(hdl)->SetSweepTiles(mode_l, mode_u, tx, ty, tz); return 0; });
}

int IluSolverGetSweepConfigEx(IluSolverHdl hdl, int* cfg) {
    S(hdl)->GetSweepConfig(cfg);
    return 0;
}

int IluSolverScheduleJsonEx(IluSolverHdl hdl, char* buf, int cap) {
    return guarded("ScheduleJson", [&] {
        const std::string js = S(hdl)->DescribeSchedule();
        if (buf && cap > 0) {
            const size_t k = std::min(js.size(), static_cast<size_t>(cap - 1));
            std::memcpy(buf, js.data(), k);
            buf[k] = '\0';
        }
        return static_cast<int>(js.size());
    });
}

int IluSolverGetLevelsEx(IluSolverHdl hdl, int* nlev_l, int* nlev_u) {
    if (nlev_l) *nlev_l = S(hdl)->NumLevelsL();
    if (nlev_u) *nlev_u = S(hdl)->NumLevelsU();
    return 0;
}

int IluSolverSymbolicDeviceEx(IluSolverHdl hdl, const CsrMatrix* d_a, int lof) {
    return guarded("SymbolicDevice", [&] { return S(hdl)->SymbolicDevice(d_a->size, d_a->row_ptr, d_a->col_idx, lof); });
}

int IluSolverFactorizeDeviceEx(IluSolverHdl hdl, const double* d_values, int values_are_a) {
    return guarded("FactorizeDevice", [&] {
        const bool ok = values_are_a ? S(hdl)->FactorizeDeviceFromA(d_values) : S(hdl)->FactorizeDevice(d_values);
        return ok ? 0 : -1;
    });
}

int IluSolverSubstituteDeviceEx(IluSolverHdl hdl, const double* d_b, double* d_x) {
    return guarded("SubstituteDevice", [&] { S(hdl)->SubstituteDevice(d_b, d_x); return 0; });
}

int IluSolverSubstituteNoAllocEx(IluSolverHdl hdl, const double* d_b, double* d_x) {
    return guarded("SubstituteNoAlloc", [&] { S(hdl)->SubstituteDeviceNoAlloc(d_b, d_x); return 0; });
}

int IluSolverCopyFactorsEx(IluSolverHdl hdl, double* out) {
    return guarded("CopyFactors", [&] {
        const int nnz = S(hdl)->Nonzeros();
        if (nnz <= 0 || !S(hdl)->DeviceFactors()) return -1;
        if (hipMemcpy(out, S(hdl)->DeviceFactors(), sizeof(double) * nnz, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        return nnz;
    });
}

int IluSolverCopyFillLevelsEx(IluSolverHdl hdl, signed char* out) {
    return guarded("CopyFillLevels", [&] { return S(hdl)->CopyFillLevels(out); });
}

int IluSolverNonzerosEx(IluSolverHdl hdl) { return S(hdl)->Nonzeros(); }

int IlukSymbolicLevelsEx(const CsrMatrix* a, int lof, CsrMatrix* f, int* nzmap, signed char* levels, long cap) {
    return guarded("IlukSymbolicLevels", [&] {
        const int n = a->size;
        if (n <= 0) return -1;
        const int nnz_a = a->row_ptr[n];
        int *d_rp = nullptr, *d_ci = nullptr;
        if (hipMalloc(&d_rp, sizeof(int) * (n + 1)) != hipSuccess) return -1;
        if (hipMalloc(&d_ci, sizeof(int) * (nnz_a > 0 ? nnz_a : 1)) != hipSuccess) { (void)hipFree(d_rp); return -1; }
        int rc = -1;
        try {
            (void)hipMemcpy(d_rp, a->row_ptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice);
            (void)hipMemcpy(d_ci, a->col_idx, sizeof(int) * nnz_a, hipMemcpyHostToDevice);
            IluSolver tmp;
            rc = tmp.SymbolicDevice(n, d_rp, d_ci, lof);
            if (rc >= 0 && levels) {
                if (cap < rc) rc = -1;
                else tmp.CopyFillLevels(levels);
            }
        } catch (...) {
            (void)hipFree(d_rp);
            (void)hipFree(d_ci);
            throw;
        }
        (void)hipFree(d_rp);
        (void)hipFree(d_ci);
        if (rc < 0) return rc;
        return IlukSymbolic(a, lof, f, nzmap);
    });
}

bool GmresSolveDeviceEx(GmresHdl hdl, const CsrMatrix* d_mat, const double* d_rhs, double* d_sol, int* iter) {
    std::pair<bool, int> r{false, 0};
    try {
        r = G(hdl)->SolveDevice(d_mat, d_rhs, d_sol);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "GmresSolveDevice: %s\n", e.what());
    }
    if (iter) *iter = r.second;
    return r.first;
}

double GmresLastTrueResidualEx(GmresHdl hdl) { return G(hdl)->LastTrueResidual(); }

int GmresResidualHistoryEx(GmresHdl hdl, double* out, int cap) {
    return guarded("ResidualHistory", [&] { return G(hdl)->ResidualHistory(out, cap); });
}

int DeviceAvailableEx(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return 0;
    return count > 0 ? 1 : 0;
}

}  // extern "C"
