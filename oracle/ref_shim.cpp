// ref_shim.cpp — C entry points around the reference's own Eigen-free CPU
// solvers, compiled from their sources where they lie:
//   Utilities/solvers/cpu/jacobi_solver.cpp  (solvers::cpu::solveMomentumJacobi)
//   Utilities/solvers/cpu/pcg_solver.cpp     (solvers::cpu::solvePCG_CSR)
// TEST INFRASTRUCTURE ONLY: used to write and check the golden solver
// fixtures. The reference's full SIMPLE path cannot be compiled (its
// vendored Eigen lacks Eigen/Core), so only these two leaves are built.
#include <vector>

namespace solvers {
namespace cpu {
int solveMomentumJacobi(int n, const std::vector<int>& row_ptr, const std::vector<int>& col_idx,
                        const std::vector<float>& values, const std::vector<float>& b, std::vector<float>& x,
                        const std::vector<int>& diagOffset, float tol, int maxIter);
int solvePCG_CSR(int n, const std::vector<int>& row_ptr, const std::vector<int>& col_idx,
                 const std::vector<float>& values, const std::vector<float>& b, std::vector<float>& x,
                 float tol, int maxIter);
}  // namespace cpu
}  // namespace solvers

extern "C" {

int ref_jacobi(int n, const int* row_ptr, const int* col_idx, const float* values, const float* b, float* x,
               const int* diagOff, float tol, int maxIter) {
    const int nnz = row_ptr[n];
    std::vector<int> rp(row_ptr, row_ptr + n + 1), ci(col_idx, col_idx + nnz), d(diagOff, diagOff + n);
    std::vector<float> va(values, values + nnz), bb(b, b + n), xx(x, x + n);
    const int it = solvers::cpu::solveMomentumJacobi(n, rp, ci, va, bb, xx, d, tol, maxIter);
    for (int i = 0; i < n; ++i) x[i] = xx[i];
    return it;
}

int ref_pcg(int n, const int* row_ptr, const int* col_idx, const float* values, const float* b, float* x,
            float tol, int maxIter) {
    const int nnz = row_ptr[n];
    std::vector<int> rp(row_ptr, row_ptr + n + 1), ci(col_idx, col_idx + nnz);
    std::vector<float> va(values, values + nnz), bb(b, b + n), xx(x, x + n);
    const int it = solvers::cpu::solvePCG_CSR(n, rp, ci, va, bb, xx, tol, maxIter);
    for (int i = 0; i < n; ++i) x[i] = xx[i];
    return it;
}

}  // extern "C"
