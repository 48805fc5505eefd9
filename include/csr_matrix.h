/* Compressed-sparse-row container and host utilities.
 *
 * Drop-in for include/csr_matrix.h of lljbash/par_ilu0_cpp (lines 10-49).
 * Layout is unchanged: 0-based int32 row_ptr[size+1], int32 col_idx[max_nnz],
 * double value[max_nnz], all host memory owned by the library (64-byte
 * aligned) and released by DestroyCsrMatrix.  A matrix starts life as
 * CSR_MATRIX_DEFAULT (size -1, null arrays).
 */
#ifndef LLJBASH_CSR_MATRIX_H_
#define LLJBASH_CSR_MATRIX_H_

#include "lljbash_decl.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CsrMatrix LLJBASH_DECL(CsrMatrix)
struct CsrMatrix {
    int size LLJBASH_DEFAULT_VALUE(-1);
    int max_nnz LLJBASH_DEFAULT_VALUE(0);
    int* row_ptr LLJBASH_DEFAULT_VALUE(nullptr);
    int* col_idx LLJBASH_DEFAULT_VALUE(nullptr);
    double* value LLJBASH_DEFAULT_VALUE(nullptr);
};
LLJBASH_STRUCT_TYPEDEF(CsrMatrix);

/* reference: include/csr_matrix.h:19-20, src/csr_matrix.cpp:25 */
#define CSR_MATRIX_DEFAULT LLJBASH_DECL(CSR_MATRIX_DEFAULT)
extern const CsrMatrix CSR_MATRIX_DEFAULT;

/* (Re)allocate the three arrays for `size` rows and `max_nnz` entries; old
 * arrays are freed.  size < 0 or max_nnz == 0 leaves the corresponding
 * arrays null.  reference: include/csr_matrix.h:22-23, src/csr_matrix.cpp:27-33 */
#define SetupCsrMatrix LLJBASH_DECL(SetupCsrMatrix)
void SetupCsrMatrix(CsrMatrix* mat, int size, int max_nnz);

/* Deep copy (pattern and values), dst resized to nnz(src).
 * reference: include/csr_matrix.h:25-26, src/csr_matrix.cpp:35-42 */
#define CopyCsrMatrix LLJBASH_DECL(CopyCsrMatrix)
void CopyCsrMatrix(CsrMatrix* dst, const CsrMatrix* src);

/* Copy nnz(src) values only.  reference: include/csr_matrix.h:28-29 */
#define CopyCsrMatrixValues LLJBASH_DECL(CopyCsrMatrixValues)
void CopyCsrMatrixValues(CsrMatrix* dst, const CsrMatrix* src);

/* Free and reset to the default state.  reference: include/csr_matrix.h:31-32 */
#define DestroyCsrMatrix LLJBASH_DECL(DestroyCsrMatrix)
void DestroyCsrMatrix(CsrMatrix* mat);

/* reference: include/csr_matrix.h:34-37 */
#define GetCsrNonzeros LLJBASH_DECL(GetCsrNonzeros)
static inline int GetCsrNonzeros(const CsrMatrix* mat) {
    return mat->row_ptr[mat->size];
}

/* Read a 1-based "coordinate real general" MatrixMarket file, then row-
 * permute by a maximum bipartite matching on |a| > 1e-12 so that every row
 * has a nonzero diagonal.  Rows are column-sorted.  Storage is sized nnz+n
 * like the reference.  On unreadable or inconsistent input it prints a
 * message and calls exit(-1), as the reference does.
 * reference: include/csr_matrix.h:39-40, src/csr_matrix.cpp:53-159 */
#define ReadCsrMatrixMM1 LLJBASH_DECL(ReadCsrMatrixMM1)
void ReadCsrMatrixMM1(CsrMatrix* mat, const char* filename);

/* y = A x on the host (serial, ascending entries).
 * reference: include/csr_matrix.h:42-43, src/csr_matrix.cpp:163-171 */
#define CsrMatVec LLJBASH_DECL(CsrMatVec)
void CsrMatVec(const CsrMatrix* mat, const double* vec, double* sol);

/* Fill- and depth-reducing symmetric orderings of the pattern of A + A^T:
 * p[k] = original index of the k-th row of the permuted matrix, and, if ip
 * is not null, ip[p[k]] = k; apply as B[ip[i]][ip[j]] = A[i][j].  The
 * reference delegates to SuiteSparse (src/csr_matrix.cpp:173-225); this
 * build has its own approximate minimum degree (CsrAmdOrder) and nested
 * dissection by recursive level-structure bisection (CsrNdOrder), so the
 * permutations are not SuiteSparse's.  Return values keep the reference's
 * conventions: CsrAmdOrder 0 on success and -1 on failure, CsrNdOrder 1 on
 * success and 0 on failure.
 * reference: include/csr_matrix.h:45-49 */
#define CsrAmdOrder LLJBASH_DECL(CsrAmdOrder)
int CsrAmdOrder(const CsrMatrix* mat, int* p, int* ip);

#define CsrNdOrder LLJBASH_DECL(CsrNdOrder)
int CsrNdOrder(const CsrMatrix* mat, int* p, int* ip);

#ifdef __cplusplus
}
#endif

#endif /* LLJBASH_CSR_MATRIX_H_ */
