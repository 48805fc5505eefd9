/* C ABI of the ILU preconditioner (opaque handle).
 *
 * Drop-in for include_c/ilu_solver_c.h of lljbash/par_ilu0_cpp (lines 15-31);
 * behaviour of src_c/ilu_solver_c.cpp:6-50.  Setup/Factorize return 0 on
 * success, -2 when a row has no diagonal entry, -1 for any other failure
 * (including a zero pivot met by Factorize).
 */
#ifndef LLJBASH_ILU_SOLVER_C_H_
#define LLJBASH_ILU_SOLVER_C_H_

#ifndef __cplusplus
#include <stdbool.h>
#endif

#include "csr_matrix.h"

#ifdef __cplusplus
extern "C" {
#endif

#define IluSolverHdl LLJBASH_DECL(IluSolverHdl)
typedef void* IluSolverHdl;

/* nthreads is accepted and ignored (GPU). */
#define IluSolverCreate LLJBASH_DECL(IluSolverCreate)
IluSolverHdl IluSolverCreate(int nthreads);

#define IluSolverDestroy LLJBASH_DECL(IluSolverDestroy)
void IluSolverDestroy(IluSolverHdl hdl);

/* The solver-owned F; fill it with CopyCsrMatrix or IlukSymbolic. */
#define IluSolverGetMatrix LLJBASH_DECL(IluSolverGetMatrix)
CsrMatrix* IluSolverGetMatrix(IluSolverHdl hdl);

#define IluSolverSetup LLJBASH_DECL(IluSolverSetup)
int IluSolverSetup(IluSolverHdl hdl);

#define IluSolverFactorize LLJBASH_DECL(IluSolverFactorize)
int IluSolverFactorize(IluSolverHdl hdl);

#ifdef __cplusplus
}
#endif

#endif /* LLJBASH_ILU_SOLVER_C_H_ */
