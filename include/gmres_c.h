/* C ABI of the GMRES solver (opaque handle).
 *
 * Drop-in for include_c/gmres_c.h of lljbash/par_ilu0_cpp (lines 10-29);
 * behaviour of src_c/gmres_c.cpp:6-33.  GmresSolve returns the converged
 * flag and, if iter is not null, the Arnoldi step count.  Unlike the
 * reference (which throws std::bad_function_call through the C ABI), a
 * solve without a preconditioner returns false with *iter = 0.
 */
#ifndef LLJBASH_GMRES_C_H_
#define LLJBASH_GMRES_C_H_

#include "gmres_param.h"
#include "ilu_solver_c.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GmresHdl LLJBASH_DECL(GmresHdl)
typedef void* GmresHdl;

#define GmresCreate LLJBASH_DECL(GmresCreate)
GmresHdl GmresCreate(void);

#define GmresDestroy LLJBASH_DECL(GmresDestroy)
void GmresDestroy(GmresHdl hdl);

#define GmresGetParameters LLJBASH_DECL(GmresGetParameters)
GmresParameters* GmresGetParameters(GmresHdl hdl);

#define GmresGetStat LLJBASH_DECL(GmresGetStat)
const GmresStat* GmresGetStat(GmresHdl hdl);

#define GmresSetPreconditioner LLJBASH_DECL(GmresSetPreconditioner)
void GmresSetPreconditioner(GmresHdl hdl, IluSolverHdl ilu);

#define GmresSolve LLJBASH_DECL(GmresSolve)
bool GmresSolve(GmresHdl hdl, const CsrMatrix* mat, double* rhs, double* sol, int* iter);

#ifdef __cplusplus
}
#endif

#endif /* LLJBASH_GMRES_C_H_ */
