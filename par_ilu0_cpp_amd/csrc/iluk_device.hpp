// Device entry of the symbolic ILU(k) pass (private to the library).
#pragma once

#include "common.hpp"

namespace pilu {

// Pattern of F on the device: row_ptr[n+1], col_idx/levels/values[nnz],
// nzmap[nnz(A)] (index in F of each entry of A).
struct SymbolicResult {
    int n = 0;
    int nnz = 0;
    DevBuf<int> rp, ci;
    DevBuf<signed char> lev;
    DevBuf<double> val;
    DevBuf<int> nzmap;
};

// Runs the pass on device arrays of A (values optional).  Returns nnz(F),
// or -1 for a row of A with a repeated column; throws std::runtime_error on
// device errors, capacity limits or a dependency wait longer than
// watchdog_s seconds.
int symbolic_device(int n, const int* d_arp, const int* d_aci, const double* d_aval, int lof, SymbolicResult& out,
                    hipStream_t s, double watchdog_s = 20.0);

}  // namespace pilu
