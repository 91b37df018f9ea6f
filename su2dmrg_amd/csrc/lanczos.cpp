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

// One solve through the batched state machine (a batch of one), so a single
// run and a batched one execute the same operations.
LanczosResult lanczos_ground(Context& c, const ApplyFn& apply, double* psi, long long n, int mlanc,
                             double tol, const ApplyFn* check_apply) {
  std::vector<LanczosJob> jobs(1);
  jobs[0].psi = psi;
  jobs[0].n = n;
  jobs[0].mlanc = mlanc;
  jobs[0].tol = tol;
  jobs[0].mixed = check_apply != nullptr;
  lanczos_batch(c, jobs, [&](const std::vector<LanczosReq>& reqs) {
    for (auto& r : reqs) (r.check && check_apply ? *check_apply : apply)(r.x, r.y);
  });
  return jobs[0].res;
}

// lanczos_batch: lanczos_ground's schedule as a state machine per job

namespace {

enum LzState { kStart, kAfterFirst, kM2Final, kLoop, kCheckCont, kCheckDone, kFinal, kDone };

struct LzJob {
  LanczosJob* job;
  int st = kStart;
  int m = 0;
  std::vector<DevBuf> ff;   // Krylov vectors 0..mlanc
  DevBuf tmp;
  std::vector<double> aal, nnl, z;
  int slot0 = 0;            // first of this job's slots in the shared scalar array
  bool pending = false;     // has a matvec request this round
  const double* x = nullptr;
  double* y = nullptr;
  bool check = false;
};

}  // namespace

void lanczos_batch(Context& c, std::vector<LanczosJob>& jobs, const BatchApply& apply) {
  const int nj = (int)jobs.size();
  if (nj == 0) return;
  // device scalars of every job: coef[0..mlanc] (alphas and the current
  // reorthogonalisation coefficient), then A and B (norms, overlaps)
  int per = 0;
  for (auto& j : jobs) {
    SU2_CHECK(j.mlanc >= 2, kErrArgument, "lanczos: mlanc=%d", j.mlanc);
    per = std::max(per, j.mlanc + 3);
  }
  DevBuf scal(c, (long long)per * nj);
  std::vector<double> host((size_t)per * nj, 0.0);
  dev_fill(c, scal.p, (long long)per * nj, 0.0);
  std::vector<LzJob> L(nj);
  for (int i = 0; i < nj; ++i) {
    LzJob& s = L[i];
    s.job = &jobs[i];
    s.job->res = LanczosResult();
    const long long n = std::max<long long>(s.job->n, 1);
    s.ff.reserve(s.job->mlanc + 1);
    for (int k = 0; k <= s.job->mlanc; ++k) s.ff.emplace_back(c, n);
    s.tmp = DevBuf(c, n);
    s.aal.assign(s.job->mlanc, 0.0);
    s.nnl.assign(s.job->mlanc + 1, 0.0);
    s.slot0 = i * per;
  }
  auto coef = [&](LzJob& s, int k) { return scal.p + s.slot0 + k; };
  auto A = [&](LzJob& s) { return scal.p + s.slot0 + s.job->mlanc + 1; };
  auto B = [&](LzJob& s) { return scal.p + s.slot0 + s.job->mlanc + 2; };
  auto hcoef = [&](LzJob& s, int k) { return host[s.slot0 + k]; };
  auto hA = [&](LzJob& s) { return host[s.slot0 + s.job->mlanc + 1]; };
  auto hB = [&](LzJob& s) { return host[s.slot0 + s.job->mlanc + 2]; };
  auto request = [&](LzJob& s, const double* x, double* y, bool check) {
    s.pending = true;
    s.x = x;
    s.y = y;
    s.check = check;
    ++s.job->res.matvecs;
    if (check && s.job->mixed) ++s.job->res.check_matvecs;
  };
  // Ritz vector of the m x m tridiagonal into psi, then psi normalised on
  // the device (the check's normalisation, lanczos_su2.cpp:320-332)
  auto ritz = [&](LzJob& s, int m, bool normalise) {
    const long long n = s.job->n;
    std::vector<double> d(s.aal.begin(), s.aal.begin() + m), e(s.nnl.begin() + 1, s.nnl.begin() + m),
        w, z;
    tridiag_eig(d, e, w, z);
    if (z[0] < 0)   // Ritz vector with a positive overlap on the start vector
      for (int j = 0; j < m; ++j) z[j] = -z[j];
    dev_fill(c, s.job->psi, n, 0.0);
    for (int j = 0; j < m; ++j) dev_axpy(c, s.job->psi, s.ff[j].p, n, z[j]);
    s.job->res.ritz = s.job->res.energy = w[0];
    if (normalise) {
      dev_dot_to(c, s.job->psi, s.job->psi, n, A(s));
      dev_scal_inv_sqrt(c, s.job->psi, n, A(s));
    }
  };
  auto check_request = [&](LzJob& s, int st) {
    ritz(s, s.m, true);
    request(s, s.job->psi, s.tmp.p, true);
    s.st = st;
  };
  // after iteration m: the next Krylov matvec, or the last one
  auto advance = [&](LzJob& s) {
    ++s.m;
    if (s.m < s.job->mlanc) {
      request(s, s.ff[s.m - 1].p, s.ff[s.m].p, false);
      s.st = kLoop;
    } else {
      request(s, s.ff[s.job->mlanc - 1].p, s.ff[s.job->mlanc].p, false);
      s.st = kFinal;
    }
  };

  // start: ff[0] = psi / |psi| (initialize_lanczos, lanczos_su2.cpp:36-47)
  for (auto& s : L) {
    const long long n = s.job->n;
    dev_dot_to(c, s.job->psi, s.job->psi, n, B(s));
    dev_copy(c, s.ff[0].p, s.job->psi, n);
    dev_scal_inv_sqrt(c, s.ff[0].p, n, B(s));
    request(s, s.ff[0].p, s.ff[1].p, false);
    s.st = kAfterFirst;
  }
  bool first = true;
  for (;;) {
    std::vector<LanczosReq> reqs;
    for (int i = 0; i < nj; ++i)
      if (L[i].pending) reqs.push_back(LanczosReq{i, L[i].x, L[i].y, L[i].check});
    if (reqs.empty()) break;
    apply(reqs);
    // each job's vector algebra after its matvec, scalars into its slots
    for (auto& s : L) {
      if (!s.pending) continue;
      s.pending = false;
      const long long n = s.job->n;
      const int m = s.m;
      switch (s.st) {
        case kAfterFirst:
          dev_dot_to(c, s.ff[0].p, s.ff[1].p, n, coef(s, 0));
          dev_axpy_dev(c, s.ff[1].p, s.ff[0].p, n, coef(s, 0), -1.0);
          dev_dot_to(c, s.ff[1].p, s.ff[1].p, n, A(s));
          break;
        case kM2Final:
          dev_dot_to(c, s.ff[1].p, s.ff[2].p, n, coef(s, 1));
          break;
        case kLoop:
          dev_dot_to(c, s.ff[m - 1].p, s.ff[m].p, n, coef(s, m - 1));
          dev_axpy_dev(c, s.ff[m].p, s.ff[m - 1].p, n, coef(s, m - 1), -1.0);
          dev_axpy(c, s.ff[m].p, s.ff[m - 2].p, n, -s.nnl[m - 1]);
          for (int j = 0; j < m; ++j) {
            dev_dot_to(c, s.ff[m].p, s.ff[j].p, n, coef(s, s.job->mlanc));
            dev_axpy_dev(c, s.ff[m].p, s.ff[j].p, n, coef(s, s.job->mlanc), -1.0);
          }
          dev_dot_to(c, s.ff[m].p, s.ff[m].p, n, A(s));
          break;
        case kCheckCont:
        case kCheckDone:
          dev_dot_to(c, s.tmp.p, s.tmp.p, n, A(s));
          dev_dot_to(c, s.tmp.p, s.job->psi, n, B(s));
          break;
        case kFinal:
          dev_dot_to(c, s.ff[s.job->mlanc - 1].p, s.ff[s.job->mlanc].p, n,
                     coef(s, s.job->mlanc - 1));
          break;
        default:
          break;
      }
    }
    // one wait for the scalars of every job
    SU2_HIP(hipMemcpyAsync(host.data(), scal.p, sizeof(double) * host.size(),
                           hipMemcpyDeviceToHost, c.stream()));
    SU2_HIP(hipStreamSynchronize(c.stream()));
    for (auto& s : L) {
      if (s.st == kDone || s.pending) continue;
      const long long n = s.job->n;
      const int mlanc = s.job->mlanc;
      switch (s.st) {
        case kAfterFirst: {
          if (first) SU2_CHECK(hB(s) > 0, kErrNumeric, "lanczos: zero start vector");
          s.aal[0] = hcoef(s, 0);
          s.nnl[1] = std::sqrt(hA(s));
          // an exactly invariant start vector (a 1-dim H_eff) leaves nnl = 0:
          // the unnormalised residual is then never read again
          if (s.nnl[1] > 0) dev_scal(c, s.ff[1].p, n, 1.0 / s.nnl[1]);
          if (mlanc == 2) {
            if (s.nnl[1] < s.job->tol) {
              ritz(s, 1, false);
              s.st = kDone;
            } else {
              request(s, s.ff[1].p, s.ff[2].p, false);
              s.st = kM2Final;
            }
          } else {
            s.m = 2;
            request(s, s.ff[1].p, s.ff[2].p, false);
            s.st = kLoop;
          }
          break;
        }
        case kM2Final:
          s.aal[1] = hcoef(s, 1);
          ritz(s, 2, true);
          s.st = kDone;
          break;
        case kLoop: {
          const int m = s.m;
          s.aal[m - 1] = hcoef(s, m - 1);
          s.nnl[m] = std::sqrt(hA(s));
          if (s.nnl[m] > 0) dev_scal(c, s.ff[m].p, n, 1.0 / s.nnl[m]);
          if (s.nnl[m] < s.job->tol) check_request(s, kCheckDone);
          else if (m % 4 == 0) check_request(s, kCheckCont);
          else advance(s);
          break;
        }
        case kCheckCont:
        case kCheckDone: {
          const double prod = std::sqrt(hA(s)), rq = hB(s);
          s.job->res.residual = std::sqrt(std::max(0.0, prod * prod - rq * rq));
          if (s.job->mixed) s.job->res.energy = rq;
          const double ov = std::fabs(rq / prod);
          if (s.st == kCheckDone || std::fabs(ov - 1.0) < 1e-8) s.st = kDone;
          else advance(s);
          break;
        }
        case kFinal:
          s.aal[mlanc - 1] = hcoef(s, mlanc - 1);
          s.m = mlanc;
          check_request(s, kCheckDone);
          break;
        default:
          break;
      }
    }
    first = false;
  }
  SU2_HIP(hipStreamSynchronize(c.stream()));
}

}  // namespace su2
