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

#include <functional>
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
    // The solver's device sweeps are used directly; precon_ is set as in
    // the reference, so code compiled against the reference's header (whose
    // inline version sets only precon_) still solves, through that
    // callback on host vectors.
    void SetIluPreconditioner(IluSolver* ilu_solver) {
        precon_ = [ilu_solver](const double* rhs, double* sol) { ilu_solver->Substitute(rhs, sol); };
        intermediate_.ilu = ilu_solver;
    }

    std::pair<bool, int> Solve(const CsrMatrix* mat, double* rhs, double* sol);

    // ---- extension surface (no counterpart in the reference) ----------
    // Any right preconditioner z = M^-1 v on host vectors of length n
    // (applied once per Arnoldi step, with a device-host round trip).
    void SetPreconditioner(std::function<void(const double*, double*)> precon) {
        precon_ = std::move(precon);
        intermediate_.ilu = nullptr;
    }
    // Same iteration on device-resident data: d_mat holds device arrays.
    std::pair<bool, int> SolveDevice(const CsrMatrix* d_mat, const double* d_rhs, double* d_sol);
    // ||b - A x||_2 of the last solve's final iterate, in FP64 on the device.
    double LastTrueResidual() const;
    // Per-step residual estimates |rs[i+1]| of the last solve (host copy).
    int ResidualHistory(double* out, int cap) const;

    struct Ext;
private:
    // The reference's members, in its order and with its sizes
    // (include/gmres.hpp:24-46 there), so that objects laid out by code
    // compiled against either header agree.  The reference's work-buffer
    // block {bool, double* vv, z, hh} holds the device solver here.
    GmresParameters param_;
    std::function<void(const double*, double*)> precon_;
    GmresStat stat_;
    struct {
        bool setup = false;
        IluSolver* ilu = nullptr;  // the IluSolver whose device sweeps are M^-1 (null: use precon_)
        Ext* ext = nullptr;        // device state of the iteration
        void* reserved = nullptr;
    } intermediate_;
};

} // namespace lljbash
