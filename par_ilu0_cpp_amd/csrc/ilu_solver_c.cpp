// C ABI of the ILU preconditioner: include/ilu_solver_c.h
// (reference include_c/ilu_solver_c.h:15-31, src_c/ilu_solver_c.cpp:6-50).
#include "ilu_solver_c.h"

#include <cstring>
#include <exception>
#include <string>

#include "ilu_solver.hpp"

using lljbash::IluSolver;

namespace {

int code_of(const std::exception& e) { return std::strcmp(e.what(), "missing diag") == 0 ? -2 : -1; }

}  // namespace

extern "C" {

IluSolverHdl IluSolverCreate(int nthreads) {
    auto* s = new IluSolver;
    s->SetThreads(nthreads);
    return static_cast<IluSolverHdl>(s);
}

void IluSolverDestroy(IluSolverHdl hdl) { delete static_cast<IluSolver*>(hdl); }

CsrMatrix* IluSolverGetMatrix(IluSolverHdl hdl) { return static_cast<IluSolver*>(hdl)->GetMatrix(); }

int IluSolverSetup(IluSolverHdl hdl) {
    try {
        static_cast<IluSolver*>(hdl)->SetupMatrix();
    } catch (const std::exception& e) {
        return code_of(e);
    }
    return 0;
}

int IluSolverFactorize(IluSolverHdl hdl) {
    try {
        return static_cast<IluSolver*>(hdl)->Factorize() ? 0 : -1;
    } catch (const std::exception& e) {
        return code_of(e);
    }
}

}  // extern "C"
