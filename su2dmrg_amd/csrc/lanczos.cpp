#include "lanczos.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

#include "blas1.hpp"
#include "common.hpp"
#include "context.hpp"

namespace su2 {

void tridiag_eig(const std::vector<double>& d, const std::vector<double>& e,
                 std::vector<double>& w, std::vector<double>& z) {
  // cyclic Jacobi on the dense m x m matrix: m is at most the Krylov size
  const int m = (int)d.size();
  std::vector<double> a((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i) {
    a[i + (size_t)m * i] = d[i];
    if (i + 1 < m) {
      a[i + (size_t)m * (i + 1)] = e[i];
      a[i + 1 + (size_t)m * i] = e[i];
    }
  }
  z.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i) z[i + (size_t)m * i] = 1.0;
  auto A = [&](int r, int c) -> double& { return a[r + (size_t)m * c]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) off += A(p, q) * A(p, q);
    if (off < 1e-300) break;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = A(p, q);
        if (std::fabs(apq) < 1e-300) continue;
        const double tau = (A(q, q) - A(p, p)) / (2 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1 + tau * tau));
        const double cs = 1 / std::sqrt(1 + t * t), sn = t * cs;
        for (int r = 0; r < m; ++r) {
          const double x = A(r, p), y = A(r, q);
          A(r, p) = cs * x - sn * y;
          A(r, q) = sn * x + cs * y;
        }
        for (int col = 0; col < m; ++col) {
          const double x = A(p, col), y = A(q, col);
          A(p, col) = cs * x - sn * y;
          A(q, col) = sn * x + cs * y;
        }
        for (int r = 0; r < m; ++r) {
          const double x = z[r + (size_t)m * p], y = z[r + (size_t)m * q];
          z[r + (size_t)m * p] = cs * x - sn * y;
          z[r + (size_t)m * q] = sn * x + cs * y;
        }
      }
  }
  std::vector<int> perm(m);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return A(x, x) < A(y, y); });
  w.resize(m);
  std::vector<double> zs((size_t)m * m);
  for (int i = 0; i < m; ++i) {
    w[i] = A(perm[i], perm[i]);
    for (int r = 0; r < m; ++r) zs[r + (size_t)m * i] = z[r + (size_t)m * perm[i]];
  }
  z.swap(zs);
}

LanczosResult lanczos_ground(Context& c, const std::function<void(const double*, double*)>& apply,
                             double* psi, long long n, int mlanc, double tol) {
  SU2_CHECK(mlanc >= 2, kErrArgument, "lanczos: mlanc=%d", mlanc);
  LanczosResult res;
  std::vector<double*> ff(mlanc + 1, nullptr);
  double* tmp = nullptr;
  for (auto& p : ff) p = c.dev_alloc_n<double>(std::max<long long>(n, 1));
  tmp = c.dev_alloc_n<double>(std::max<long long>(n, 1));
  // alphas of the main loop stay on the device until a Ritz step needs them
  // (indices [1, dev_hi)); slot mlanc holds the current reorthogonalisation
  // coefficient.  Only the residual norm, which steers the loop, is waited for.
  double* d_coef = c.dev_alloc_n<double>(mlanc + 1);
  int dev_hi = 1;
  std::vector<double> aal(mlanc, 0.0), nnl(mlanc + 1, 0.0);
  auto hv = [&](const double* x, double* y) {
    apply(x, y);
    ++res.matvecs;
  };
  auto nrm = [&](const double* x) { return std::sqrt(dev_dot(c, x, x, n)); };
  // Ritz vector of the m x m tridiagonal into psi, its value returned
  auto ritz = [&](int m) {
    if (dev_hi > 1) {
      SU2_HIP(hipMemcpyAsync(aal.data() + 1, d_coef + 1, sizeof(double) * (dev_hi - 1),
                             hipMemcpyDeviceToHost, c.stream()));
      SU2_HIP(hipStreamSynchronize(c.stream()));
    }
    std::vector<double> d(aal.begin(), aal.begin() + m), e(nnl.begin() + 1, nnl.begin() + m), w, z;
    tridiag_eig(d, e, w, z);
    if (z[0] < 0)   // Ritz vector with a positive overlap on the start vector
      for (int j = 0; j < m; ++j) z[j] = -z[j];
    dev_fill(c, psi, n, 0.0);
    for (int j = 0; j < m; ++j) dev_axpy(c, psi, ff[j], n, z[j]);
    return w[0];
  };
  // check_eigenvector: normalise psi, one more matvec, overlap of H psi with psi
  auto check = [&]() {
    const double nr = nrm(psi);
    dev_scal(c, psi, n, 1.0 / nr);
    hv(psi, tmp);
    const double prod = nrm(tmp);
    return std::fabs(dev_dot(c, tmp, psi, n) / prod);
  };
  auto finish = [&]() {
    for (auto p : ff) c.dev_free(p);
    c.dev_free(tmp);
    c.dev_free(d_coef);
  };

  // ff[0] = normalised start vector (initialize_lanczos, lanczos_su2.cpp:36-47)
  const double n0 = nrm(psi);
  SU2_CHECK(n0 > 0, kErrNumeric, "lanczos: zero start vector");
  dev_copy(c, ff[0], psi, n);
  dev_scal(c, ff[0], n, 1.0 / n0);

  hv(ff[0], ff[1]);
  aal[0] = dev_dot(c, ff[0], ff[1], n);
  dev_axpy(c, ff[1], ff[0], n, -aal[0]);
  nnl[1] = nrm(ff[1]);
  // an exactly invariant start vector (a 1-dim H_eff) leaves nnl = 0: the
  // unnormalised residual is then never read again
  if (nnl[1] > 0) dev_scal(c, ff[1], n, 1.0 / nnl[1]);
  if (mlanc == 2) {
    if (nnl[1] < tol) {
      res.energy = ritz(1);
      finish();
      return res;
    }
    hv(ff[1], ff[2]);
    aal[1] = dev_dot(c, ff[1], ff[2], n);
    res.energy = ritz(2);
    const double nr = nrm(psi);
    dev_scal(c, psi, n, 1.0 / nr);
    finish();
    return res;
  }
  for (int m = 2; m < mlanc; ++m) {
    hv(ff[m - 1], ff[m]);
    dev_dot_to(c, ff[m - 1], ff[m], n, d_coef + (m - 1));
    dev_axpy_dev(c, ff[m], ff[m - 1], n, d_coef + (m - 1), -1.0);
    dev_axpy(c, ff[m], ff[m - 2], n, -nnl[m - 1]);
    for (int j = 0; j < m; ++j) {
      dev_dot_to(c, ff[m], ff[j], n, d_coef + mlanc);
      dev_axpy_dev(c, ff[m], ff[j], n, d_coef + mlanc, -1.0);
    }
    dev_hi = m;
    nnl[m] = nrm(ff[m]);
    if (nnl[m] > 0) dev_scal(c, ff[m], n, 1.0 / nnl[m]);
    if (nnl[m] < tol) {
      res.energy = ritz(m);
      check();
      finish();
      return res;
    }
    if (m % 4 == 0) {
      res.energy = ritz(m);
      const double ov = check();
      if (std::fabs(ov - 1.0) < 1e-8) {
        finish();
        return res;
      }
    }
  }
  hv(ff[mlanc - 1], ff[mlanc]);
  aal[mlanc - 1] = dev_dot(c, ff[mlanc - 1], ff[mlanc], n);
  res.energy = ritz(mlanc);
  check();
  finish();
  return res;
}

}  // namespace su2
