// Restarted, right-preconditioned GMRES(m) on the GPU.
//
// Drop-in for include/gmres.hpp of lljbash/par_ilu0_cpp (lines 9-47) with
// the iteration of src/gmres.cpp:138-288: modified Gram-Schmidt Arnoldi,
// Givens rotations, flexible form (Z = M^-1 V kept), absolute tolerance on
// ||b - A x||_2, `its` counting Arnoldi steps over all cycles.  Solve()
// takes host arrays; `sol` is the initial guess on entry.  Returns
// {converged, its}, with the reference's quirks kept: a NaN residual
// estimate returns false; exhausting max_iterations exactly at a restart
// boundary returns true (src/gmres.cpp:179,199,278-287).
#pragma once

#include <utility>
#include "gmres_param.h"
#include "ilu_solver.hpp"

namespace lljbash {

class PreconditionedGmres {
public:
    PreconditionedGmres() = default;
    ~PreconditionedGmres();
    PreconditionedGmres(const PreconditionedGmres&) = delete;
    PreconditionedGmres& operator=(const PreconditionedGmres&) = delete;

    GmresParameters& param() { return param_; }
    GmresStat& stat() { return stat_; }

    // M^-1 is this solver's Substitute; the pointer is kept, so the
    // IluSolver must outlive this object (same contract as the reference).
    void SetIluPreconditioner(IluSolver* ilu_solver) { ilu_ = ilu_solver; }

    std::pair<bool, int> Solve(const CsrMatrix* mat, double* rhs, double* sol);

    // ---- extension surface (no counterpart in the reference) ----------
    // Same iteration on device-resident data: d_mat holds device arrays.
    std::pair<bool, int> SolveDevice(const CsrMatrix* d_mat, const double* d_rhs, double* d_sol);
    // ||b - A x||_2 of the last solve's final iterate, in FP64 on the device.
    double LastTrueResidual() const;
    // Per-step residual estimates |rs[i+1]| of the last solve (host copy).
    int ResidualHistory(double* out, int cap) const;

    struct Ext;
private:
    GmresParameters param_;
    GmresStat stat_;
    IluSolver* ilu_ = nullptr;
    Ext* ext_ = nullptr;
};

} // namespace lljbash
