// A C++ host of libpar-ilu0-gmres.so checking that PreconditionedGmres keeps
// the reference's object layout (include/gmres.hpp:24-46 of the reference:
// param_, precon_, stat_, then a {bool, three pointers} block), so that code
// compiled against the reference's header -- whose inline
// SetIluPreconditioner stores only precon_ -- works with this library.
//
// It solves a 2D Poisson system (ILU(0), GMRES(30)) three ways and prints
// "iterations <a> <b> <c>" and "max_diff <d>":
//   a: SetIluPreconditioner (device sweeps);
//   b: precon_ written through the reference's layout, as the reference's
//      inline SetIluPreconditioner would (host callback path);
//   c: SetPreconditioner with a host callback.
// Exit status 0 when all three converge.
#include <cmath>
#include <cstdio>
#include <functional>
#include <new>
#include <vector>

#include "gmres.hpp"

namespace {

// The reference's members of PreconditionedGmres, in order.
struct ReferenceLayout {
    GmresParameters param_;
    std::function<void(const double*, double*)> precon_;
    GmresStat stat_;
    struct {
        bool setup = false;
        double* vv = nullptr;
        double* z = nullptr;
        double* hh = nullptr;
    } intermediate_;
};

// The reference's members of IluSolver (include/ilu_solver.hpp:22-29).
struct ReferenceIluLayout {
    int threads_;
    CsrMatrix aMatrix_;
    void* ext_;
    void* extt_;
};

static_assert(sizeof(lljbash::PreconditionedGmres) == sizeof(ReferenceLayout), "object size differs");
static_assert(alignof(lljbash::PreconditionedGmres) == alignof(ReferenceLayout), "alignment differs");
static_assert(sizeof(lljbash::IluSolver) == sizeof(ReferenceIluLayout), "IluSolver size differs");
static_assert(alignof(lljbash::IluSolver) == alignof(ReferenceIluLayout), "IluSolver alignment differs");

void poisson(int s, CsrMatrix* a) {
    const int n = s * s;
    std::vector<int> rp(1, 0), ci;
    std::vector<double> va;
    for (int y = 0; y < s; ++y)
        for (int x = 0; x < s; ++x) {
            auto add = [&](int c, double v) { ci.push_back(c); va.push_back(v); };
            const int i = x + s * y;
            if (y > 0) add(i - s, -1.0);
            if (x > 0) add(i - 1, -1.0);
            add(i, 4.0);
            if (x + 1 < s) add(i + 1, -1.0);
            if (y + 1 < s) add(i + s, -1.0);
            rp.push_back(static_cast<int>(ci.size()));
        }
    SetupCsrMatrix(a, n, static_cast<int>(ci.size()));
    for (int i = 0; i <= n; ++i) a->row_ptr[i] = rp[i];
    for (size_t k = 0; k < ci.size(); ++k) {
        a->col_idx[k] = ci[k];
        a->value[k] = va[k];
    }
}

}  // namespace

int main() {
    CsrMatrix a = CSR_MATRIX_DEFAULT;
    poisson(40, &a);
    const int n = a.size;
    std::vector<double> ones(n, 1.0), b(n);
    CsrMatVec(&a, ones.data(), b.data());
    lljbash::IluSolver ilu;
    CopyCsrMatrix(ilu.GetMatrix(), &a);
    ilu.SetupMatrix();
    if (!ilu.Factorize()) return 4;
    auto host_apply = [&ilu](const double* rhs, double* sol) { ilu.Substitute(rhs, sol); };

    int its[3] = {0, 0, 0};
    bool ok[3] = {false, false, false};
    std::vector<double> x[3];
    for (int form = 0; form < 3; ++form) {
        lljbash::PreconditionedGmres g;
        g.param().krylov_subspace_dimension = 30;
        if (form == 0) g.SetIluPreconditioner(&ilu);
        if (form == 1) reinterpret_cast<ReferenceLayout*>(&g)->precon_ = host_apply;
        if (form == 2) g.SetPreconditioner(host_apply);
        x[form].assign(n, 0.0);
        const auto r = g.Solve(&a, b.data(), x[form].data());
        ok[form] = r.first;
        its[form] = r.second;
    }
    double diff = 0.0;
    for (int i = 0; i < n; ++i) diff = std::fmax(diff, std::fmax(std::fabs(x[1][i] - x[0][i]), std::fabs(x[2][i] - x[0][i])));
    std::printf("iterations %d %d %d\nconverged %d %d %d\nmax_diff %.3e\n", its[0], its[1], its[2], ok[0], ok[1], ok[2],
                diff);
    DestroyCsrMatrix(&a);
    return ok[0] && ok[1] && ok[2] ? 0 : 2;
}
