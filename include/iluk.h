/* Symbolic ILU(k): level-of-fill pattern of the incomplete factor.
 *
 * Drop-in for include/iluk.h of lljbash/par_ilu0_cpp (line 15) with the
 * semantics of src/iluk.cpp:5-54, computed on the GPU:
 *  - levels are "first found": pivots k < i are visited in ascending column
 *    order and a fill entry keeps the level of the first pivot that created
 *    it, even if a later pivot would give a lower level;
 *  - only pivots of level < lof are expanded, and only if row k holds its
 *    diagonal; candidates above lof are dropped;
 *  - level-0 slots of row i receive a->value[row_ptr[i] + t] in order, where
 *    t counts the level-0 slots of the row; nzmap[row_ptr[i] + t] is that
 *    slot's index in f (nzmap may be null);
 *  - fill slots are set to 0 (the reference leaves them uninitialised).
 * Returns nnz(F).  Where the reference calls exit(-1) (a row of A with a
 * repeated column), this returns -1 and leaves f untouched; a device error
 * also returns -1.
 */
#ifndef LLJBASH_ILUK_H_
#define LLJBASH_ILUK_H_

#include "csr_matrix.h"
#include "lljbash_decl.h"

#ifdef __cplusplus
extern "C" {
#endif

#define IlukSymbolic LLJBASH_DECL(IlukSymbolic)
int IlukSymbolic(const CsrMatrix* a, int lof, CsrMatrix* f, int* nzmap);

#ifdef __cplusplus
}
#endif

#endif /* LLJBASH_ILUK_H_ */
