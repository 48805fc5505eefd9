// C ABI of the GMRES solver: include/gmres_c.h
// (reference include_c/gmres_c.h:10-29, src_c/gmres_c.cpp:6-33).
#include "gmres_c.h"

#include <cstdio>
#include <exception>

#include "gmres.hpp"

using lljbash::IluSolver;
using lljbash::PreconditionedGmres;

extern "C" {

GmresHdl GmresCreate(void) { return static_cast<GmresHdl>(new PreconditionedGmres); }

void GmresDestroy(GmresHdl hdl) { delete static_cast<PreconditionedGmres*>(hdl); }

GmresParameters* GmresGetParameters(GmresHdl hdl) { return &static_cast<PreconditionedGmres*>(hdl)->param(); }

const GmresStat* GmresGetStat(GmresHdl hdl) { return &static_cast<PreconditionedGmres*>(hdl)->stat(); }

void GmresSetPreconditioner(GmresHdl hdl, IluSolverHdl ilu) {
    static_cast<PreconditionedGmres*>(hdl)->SetIluPreconditioner(static_cast<IluSolver*>(ilu));
}

bool GmresSolve(GmresHdl hdl, const CsrMatrix* mat, double* rhs, double* sol, int* iter) {
    std::pair<bool, int> r{false, 0};
    try {
        r = static_cast<PreconditionedGmres*>(hdl)->Solve(mat, rhs, sol);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "GmresSolve: %s\n", e.what());
        r = {false, 0};
    }
    if (iter) *iter = r.second;
    return r.first;
}

}  // extern "C"
