// Host CSR container utilities and the MatrixMarket reader.
// Behaviour follows src/csr_matrix.cpp of lljbash/par_ilu0_cpp; the
// implementation is this project's own (no MKL/SuiteSparse).
#include "csr_matrix.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

namespace {

// 64-byte aligned host allocation, the alignment of the reference's
// mkl_malloc(..., 64) (src/csr_matrix.cpp:12-23).
template <typename T>
void realloc_aligned(T*& arr, long count) {
    std::free(arr);
    arr = nullptr;
    if (count > 0) {
        void* p = nullptr;
        if (posix_memalign(&p, 64, static_cast<size_t>(count) * sizeof(T)) != 0) {
            std::fprintf(stderr, "SetupCsrMatrix: out of host memory\n");
            std::exit(-1);
        }
        arr = static_cast<T*>(p);
    }
}

constexpr double kMinDiag = 1e-12;  // src/csr_matrix.cpp:53

// Augmenting-path search of the reference's recursive `match`
// (src/csr_matrix.cpp:55-66), run with an explicit stack so deep chains do
// not overflow the call stack.  Candidates of a row are tried in the row's
// (value-descending) order; a column is visited once per top-level search.
bool augment(int root, const std::vector<std::vector<std::pair<int, double>>>& rows,
             std::vector<int>& col_owner, std::vector<char>& visited) {
    struct Frame { int row; size_t next; int col; };
    std::vector<Frame> stack;
    stack.push_back({root, 0, -1});
    while (!stack.empty()) {
        Frame& f = stack.back();
        const auto& cand = rows[f.row];
        bool pushed = false;
        while (f.next < cand.size()) {
            const int j = cand[f.next].first;
            const double a = cand[f.next].second;
            ++f.next;
            if (std::abs(a) > kMinDiag && !visited[j]) {
                visited[j] = 1;
                f.col = j;
                if (col_owner[j] == -1) {
                    for (auto it = stack.rbegin(); it != stack.rend(); ++it) col_owner[it->col] = it->row;
                    return true;
                }
                stack.push_back({col_owner[j], 0, -1});
                pushed = true;
                break;
            }
        }
        if (!pushed) stack.pop_back();
    }
    return false;
}

}  // namespace

extern "C" {

const CsrMatrix CSR_MATRIX_DEFAULT = CsrMatrix{};

void SetupCsrMatrix(CsrMatrix* mat, int size, int max_nnz) {
    mat->size = size;
    mat->max_nnz = max_nnz;
    realloc_aligned(mat->row_ptr, static_cast<long>(size) + 1);
    realloc_aligned(mat->col_idx, max_nnz);
    realloc_aligned(mat->value, max_nnz);
}

void CopyCsrMatrix(CsrMatrix* dst, const CsrMatrix* src) {
    const int n = src->size;
    const int nnz = GetCsrNonzeros(src);
    SetupCsrMatrix(dst, n, nnz);
    std::memcpy(dst->row_ptr, src->row_ptr, sizeof(int) * (static_cast<size_t>(n) + 1));
    if (nnz > 0) {
        std::memcpy(dst->col_idx, src->col_idx, sizeof(int) * static_cast<size_t>(nnz));
        std::memcpy(dst->value, src->value, sizeof(double) * static_cast<size_t>(nnz));
    }
}

void CopyCsrMatrixValues(CsrMatrix* dst, const CsrMatrix* src) {
    const int nnz = GetCsrNonzeros(src);
    if (nnz > 0) std::memcpy(dst->value, src->value, sizeof(double) * static_cast<size_t>(nnz));
}

void DestroyCsrMatrix(CsrMatrix* mat) { SetupCsrMatrix(mat, -1, 0); }

// src/csr_matrix.cpp:81-159: read entries, sort each row by value
// (descending), complete a row matching on |a| > 1e-12 starting from the
// rows whose own diagonal qualifies, then emit row i := old row p[i] with
// columns ascending; every emitted row must hold a qualifying diagonal.
void ReadCsrMatrixMM1(CsrMatrix* mat, const char* filename) {
    std::FILE* fin = std::fopen(filename, "r");
    if (!fin) {
        std::fprintf(stderr, "cannot open file %s\n", filename);
        std::exit(-1);
    }
    char buf[256];
    bool header_pending = true;
    int n = -1, nnz = 0, read = 0;
    std::vector<std::vector<std::pair<int, double>>> rows;
    std::vector<int> owner;
    while ((header_pending || read < nnz) && std::fgets(buf, sizeof(buf), fin)) {
        if (buf[0] == '%') continue;
        if (header_pending) {
            int m = -1;
            if (std::sscanf(buf, "%d%d%d", &n, &m, &nnz) != 3) continue;
            if (m > n) n = m;
            SetupCsrMatrix(mat, n, nnz + n);
            rows.assign(static_cast<size_t>(n), {});
            owner.assign(static_cast<size_t>(n), -1);
            header_pending = false;
        } else {
            int i = 0, j = 0;
            double a = 0;
            if (std::sscanf(buf, "%d%d%lf", &i, &j, &a) != 3) continue;
            if (i < 1 || i > n || j < 1 || j > n) {
                std::fprintf(stderr, "entry (%d, %d) out of range\n", i, j);
                std::fclose(fin);
                std::exit(-1);
            }
            rows[static_cast<size_t>(i - 1)].emplace_back(j - 1, a);
            if (i == j && std::abs(a) > kMinDiag) owner[static_cast<size_t>(i - 1)] = i - 1;
            ++read;
        }
    }
    std::fclose(fin);
    if (read != nnz) {
        std::puts("nnz number error");
        std::exit(-1);
    }
    for (auto& r : rows)
        std::stable_sort(r.begin(), r.end(), [](const std::pair<int, double>& x, const std::pair<int, double>& y) {
            return x.second > y.second;
        });
    const std::vector<int> had_diag(owner);
    std::vector<char> visited(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        std::fill(visited.begin(), visited.end(), 0);
        if (had_diag[static_cast<size_t>(i)] == -1 && !augment(i, rows, owner, visited)) {
            std::puts("error while matching");
            std::exit(-1);
        }
    }
    int pos = 0;
    for (int i = 0; i < n; ++i) {
        auto& r = rows[static_cast<size_t>(owner[static_cast<size_t>(i)])];
        std::sort(r.begin(), r.end());
        mat->row_ptr[i] = pos;
        bool diag = false;
        for (const auto& [j, a] : r) {
            mat->col_idx[pos] = j;
            mat->value[pos] = a;
            ++pos;
            if (j == i && std::abs(a) > kMinDiag) diag = true;
        }
        if (!diag) {
            std::puts("error: missing diag");
            std::exit(-1);
        }
    }
    mat->row_ptr[n] = pos;
}

void CsrMatVec(const CsrMatrix* mat, const double* vec, double* sol) {
    for (int r = 0; r < mat->size; ++r) {
        double s = 0.0;
        for (int p = mat->row_ptr[r]; p < mat->row_ptr[r + 1]; ++p) s += mat->value[p] * vec[mat->col_idx[p]];
        sol[r] = s;
    }
}

}  // extern "C"
