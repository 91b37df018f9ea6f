// Lanczos ground-state solver (lanczos_su2::lanczos1, lanczos_su2.cpp:125-195,
// with diatridiag 237-259, compute_eigenvector 262-279 and check_eigenvector
// 320-332).  Krylov vectors live on the device; each iteration is one H_eff
// matvec plus deterministic device dot products and axpys (full classical
// Gram-Schmidt against all previous vectors, as the reference does).
#pragma once

#include <functional>
#include <vector>

namespace su2 {

class Context;

struct LanczosResult {
  double energy = 0;   // lowest Ritz value of the final tridiagonal
  int matvecs = 0;
};

// apply(x, y): y = H x on device buffers of length n.  psi holds the start
// vector on entry and the normalised Ritz vector on exit.
LanczosResult lanczos_ground(Context& c, const std::function<void(const double*, double*)>& apply,
                             double* psi, long long n, int mlanc, double tol = 1e-6);

// Eigenpairs of a real symmetric tridiagonal matrix (host; diagonal d, off-
// diagonal e, size m): values ascending, vectors column-major m x m.
void tridiag_eig(const std::vector<double>& d, const std::vector<double>& e,
                 std::vector<double>& w, std::vector<double>& z);

}  // namespace su2
