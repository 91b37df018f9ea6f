#include "svd.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

#include "blas1.hpp"
#include "common.hpp"
#include "context.hpp"
#include "jacobi.hpp"
#include "ops.hpp"

namespace su2 {


namespace {

struct Sec {
  int twoj, m, n, k, dc;
  const double* A;
  DevBuf V, U;
  std::vector<double> s;   // k values, ascending, zeros for cut
};

// Singular values, keep rule and column factors of one wavefunction's
// sectors (Jacobi done): the factors s^(p-1) / s^p of every sector go into
// `fac` (host), the sector's offset into foff.
void svd_values(std::vector<Sec>& secs, std::vector<std::vector<double>>& wv, double p1,
                double p2, int max_dcut, std::vector<double>& fac, std::vector<long long>& foff) {
  const double tol = 1e-100;
  // singular values and column scaling (tensor.cpp:1533-1568, 1620-1655)
  for (size_t i = 0; i < secs.size(); ++i) {
    Sec& s = secs[i];
    foff.push_back((long long)fac.size());
    if (s.k == 0) continue;
    std::vector<double>& w = wv[i];
    const int k = s.k;
    s.s.assign(k, 0.0);
    const double num = w[k - 1];
    int dc = 0;
    if (num > tol)
      for (int t = k - 1; t >= 0; --t)
        if (w[t] / num > tol) {
          s.s[t] = std::sqrt(w[t]);
          ++dc;
        }
    dc = std::min(dc, max_dcut);
    dc = std::min(dc, s.m >= s.n ? s.m : s.n);
    s.dc = dc;
    // factors for the long-side matrix (held in U) and the eigenvectors (V),
    // s^(p-1) and s^p as the reference scales u = A v and v.  A kept value at
    // or below 1e-11 of the sector's largest still counts as a multiplet (the
    // reference's rule) and keeps its eigenvector, an orthonormal direction
    // of the short side as dsyev gives it; only a long-side column that would
    // be divided by such a value is zeroed, since A v is rounding noise there
    // and normalising noise would put a near-duplicate state into the basis.
    std::vector<double> fl(k, 0.0), fe(k, 0.0);
    const double pl = s.m >= s.n ? p1 - 1.0 : p2 - 1.0;  // long side: U (m>=n) or V (m<n)
    const double pe = s.m >= s.n ? p2 : p1;
    const double noise = 1e-11 * s.s[k - 1];
    for (int t = k - dc; t < k; ++t) {
      fe[t] = std::pow(s.s[t], pe);
      if (s.s[t] > noise || pl >= 0) fl[t] = std::pow(s.s[t], pl);
    }
    fac.insert(fac.end(), fl.begin(), fl.end());
    fac.insert(fac.end(), fe.begin(), fe.end());
  }
}

// Global cut and output tensors of one wavefunction (columns already scaled).
SvdResult svd_cut(Context& c, const Tensor& psi, std::vector<Sec>& secs, int max_dcut) {
  std::vector<double> wsort;
  std::vector<int> isort;
  for (size_t i = 0; i < secs.size(); ++i) {
    const Sec& s = secs[i];
    for (int t = 0; t < s.dc; ++t) {
      wsort.push_back(s.s[s.k - s.dc + t]);
      isort.push_back((int)i);
    }
  }
  // global cut: shell sort ascending, keep the top max_dcut
  shell_sort_double(wsort, isort);
  const int ntot = (int)wsort.size();
  const int k0 = max_dcut > ntot ? 0 : ntot - max_dcut;
  std::vector<int> keep(secs.size(), 0);
  for (int t = ntot - 1; t >= k0; --t) keep[isort[t]]++;
  double disc = 0;
  for (int t = 0; t < k0; ++t) disc += wsort[t] * wsort[t];

  SvdResult res;
  std::vector<int> bt, bd;
  for (size_t i = 0; i < secs.size(); ++i)
    if (keep[i] > 0) {
      bt.push_back(secs[i].twoj);
      bd.push_back(keep[i]);
    }
  res.bond = make_leg(bt, bd, -1);
  res.discarded_weight = disc;
  for (size_t i = 0; i < secs.size(); ++i)
    for (int t = secs[i].k - keep[i]; t < secs[i].k; ++t) res.wout.push_back(secs[i].s[t]);
  for (size_t i = 0; i < secs.size(); ++i)
    for (int t = secs[i].k - secs[i].dc; t < secs[i].k - keep[i]; ++t)
      res.wout.push_back(secs[i].s[t]);
  if (bt.empty()) fail(kErrNumeric, "svd: every singular value was cut");
  Leg la = psi.leg(0), lb = psi.leg(1);
  la.dir = 1;
  lb.dir = 1;
  res.U.reset(new Tensor(&c, {la, res.bond}, 0, nullptr));
  res.V.reset(new Tensor(&c, {lb, res.bond}, 0, nullptr));
  for (size_t i = 0; i < secs.size(); ++i) {
    const Sec& s = secs[i];
    const int d = keep[i];
    if (d == 0) continue;
    const int k = s.k;
    // kept columns are the last d of the k (ascending singular values)
    const double* Ucols;
    const double* Vcols;
    if (s.m >= s.n) {
      Ucols = s.U.p + (long long)(k - d) * s.m;
      Vcols = s.V.p + (long long)(k - d) * k;
    } else {
      Ucols = s.V.p + (long long)(k - d) * k;
      Vcols = s.U.p + (long long)(k - d) * s.n;
    }
    const int bu = res.U->find2(s.twoj, s.twoj), bv = res.V->find2(s.twoj, s.twoj);
    SU2_CHECK(bu >= 0 && bv >= 0, kErrStructure, "svd: missing output block");
    dev_copy(c, res.U->data() + res.U->blocks()[bu].offset, Ucols, (long long)s.m * d);
    dev_copy(c, res.V->data() + res.V->blocks()[bv].offset, Vcols, (long long)s.n * d);
    std::vector<double> kv(s.s.begin() + (k - d), s.s.end());
    res.kept.push_back(kv);
  }
  return res;
}

}  // namespace

std::vector<SvdResult> truncated_svd_batch(Context& c, const std::vector<SvdJob>& jobs) {
  c.make_current();
  std::vector<std::vector<Sec>> secs(jobs.size());
  for (size_t q = 0; q < jobs.size(); ++q) {
    const SvdJob& J = jobs[q];
    SU2_CHECK(J.psi && J.psi->nleg() == 2, kErrStructure, "svd: wavefunction must have two legs");
    SU2_CHECK(J.max_dcut >= 1, kErrArgument, "svd: max_dcut=%d", J.max_dcut);
    for (auto& b : J.psi->blocks()) {
      Sec s;
      s.twoj = b.twoj[0];
      s.m = b.dim[0];
      s.n = b.dim[1];
      s.k = std::min(s.m, s.n);
      s.dc = 0;
      s.A = J.psi->data() + b.offset;
      secs[q].push_back(std::move(s));
    }
  }
  // Singular triplets of every sector of every wavefunction in one batched
  // one-sided Jacobi SVD.  The reference diagonalises the Gram matrix of the
  // short side and re-diagonalises its small-eigenvalue subspace until
  // accurate (tensor.cpp:1472-1529); rotating the sector's own columns
  // reaches the same triplets without forming the Gram matrix.  "U" receives
  // the long side (A V for m >= n, A^T U for m < n: the reference's u = A v),
  // "V" the short side's eigenvectors, both in ascending order of the
  // singular values.  Each block's rotations depend on that block alone, so
  // the batch gives every wavefunction the bits it has by itself.
  std::vector<std::vector<std::vector<double>>> wv(jobs.size());
  int jacobi_sweeps = 0;
  double t_jacobi = 0, jacobi_worst = 0;
  {
    std::vector<JacobiBlock> jb;
    std::vector<std::pair<size_t, size_t>> idx;
    for (size_t q = 0; q < jobs.size(); ++q) {
      wv[q].resize(secs[q].size());
      for (size_t i = 0; i < secs[q].size(); ++i) {
        Sec& s = secs[q][i];
        if (s.k == 0) continue;
        const int L = std::max(s.m, s.n);
        s.U = DevBuf(c, (long long)L * s.k);
        s.V = DevBuf(c, (long long)s.k * s.k);
        JacobiBlock b;
        b.A = s.A;
        b.m = s.m;
        b.n = s.n;
        b.W = s.U.p;
        b.V = s.V.p;
        b.group = (int)q;
        jb.push_back(b);
        idx.push_back({q, i});
      }
    }
    const auto tj = std::chrono::steady_clock::now();
    jacobi_sweeps = jacobi_svd_batched(c, jb, 1e-13, 60, c.svd_block_cols, c.svd_inner_sweeps,
                                       &jacobi_worst, c.svd_precondition_min_k);
    t_jacobi = std::chrono::duration<double>(std::chrono::steady_clock::now() - tj).count();
    for (size_t t = 0; t < jb.size(); ++t) {
      std::vector<double>& w = wv[idx[t].first][idx[t].second];
      w.resize(jb[t].s.size());
      for (size_t q = 0; q < w.size(); ++q) w[q] = jb[t].s[q] * jb[t].s[q];
    }
  }
  // column factors of every sector of every wavefunction: one upload, the
  // scalings queued, no wait until the outputs are in place
  std::vector<double> fac;
  std::vector<std::vector<long long>> foff(jobs.size());
  for (size_t q = 0; q < jobs.size(); ++q)
    svd_values(secs[q], wv[q], jobs[q].p1, jobs[q].p2, jobs[q].max_dcut, fac, foff[q]);
  DevBuf dfac(c, (long long)fac.size());
  if (!fac.empty())
    SU2_HIP(hipMemcpyAsync(dfac.p, fac.data(), fac.size() * sizeof(double), hipMemcpyHostToDevice,
                           c.stream()));
  for (size_t q = 0; q < jobs.size(); ++q)
    for (size_t i = 0; i < secs[q].size(); ++i) {
      Sec& s = secs[q][i];
      if (s.k == 0) continue;
      const int longrows = s.m >= s.n ? s.m : s.n;
      dev_scale_columns(c, s.U.p, longrows, s.k, dfac.p + foff[q][i]);
      dev_scale_columns(c, s.V.p, s.k, s.k, dfac.p + foff[q][i] + s.k);
    }
  std::vector<SvdResult> out;
  for (size_t q = 0; q < jobs.size(); ++q) {
    out.push_back(svd_cut(c, *jobs[q].psi, secs[q], jobs[q].max_dcut));
    out.back().jacobi_sweeps = jacobi_sweeps;
    out.back().t_jacobi = t_jacobi / jobs.size();
    out.back().jacobi_worst = jacobi_worst;
  }
  SU2_HIP(hipStreamSynchronize(c.stream()));
  return out;
}

SvdResult truncated_svd(Context& c, const Tensor& psi, double p1, double p2, int max_dcut) {
  std::vector<SvdJob> jobs{SvdJob{&psi, p1, p2, max_dcut}};
  std::vector<SvdResult> r = truncated_svd_batch(c, jobs);
  return std::move(r[0]);
}

}  // namespace su2
