/* GMRES parameters and accumulated timers (plain C structs).
 *
 * Drop-in for include/gmres_param.h of lljbash/par_ilu0_cpp (lines 10-23).
 * tolerance is ABSOLUTE on ||b - A x||_2, as in src/gmres.cpp:186,278; for a
 * relative target pass tol * ||b||_2 with x0 = 0.
 * Timers are in seconds, measured with HIP events on the solver stream and
 * accumulated over calls (never reset by the library).
 */
#ifndef LLJBASH_GMRES_PARAM_H_
#define LLJBASH_GMRES_PARAM_H_

#include "lljbash_decl.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GmresParameters LLJBASH_DECL(GmresParameters)
struct GmresParameters {
    double tolerance LLJBASH_DEFAULT_VALUE(1e-8);
    int krylov_subspace_dimension LLJBASH_DEFAULT_VALUE(60);
    int max_iterations LLJBASH_DEFAULT_VALUE(200);
};
LLJBASH_STRUCT_TYPEDEF(GmresParameters);

#define GmresStat LLJBASH_DECL(GmresStat)
struct GmresStat {
    double total_init_time LLJBASH_DEFAULT_VALUE(0.0);   /* uploads of A/rhs/x0 */
    double total_fgmr_time LLJBASH_DEFAULT_VALUE(0.0);   /* whole device solve   */
    double total_mv_time LLJBASH_DEFAULT_VALUE(0.0);     /* SpMV kernels         */
    double total_precon_time LLJBASH_DEFAULT_VALUE(0.0); /* L and U sweeps       */
};
LLJBASH_STRUCT_TYPEDEF(GmresStat);

#ifdef __cplusplus
}
#endif

#endif /* LLJBASH_GMRES_PARAM_H_ */
