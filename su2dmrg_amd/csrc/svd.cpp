#include "svd.hpp"

#include <algorithm>
#include <cmath>

#include "blas1.hpp"
#include "common.hpp"
#include "context.hpp"
#include "eig.hpp"
#include "ops.hpp"

namespace su2 {

namespace {

struct DevBuf {
  double* p = nullptr;
  DevBuf() = default;
  explicit DevBuf(long long n) {
    if (n > 0) SU2_HIP(hipMalloc(&p, n * sizeof(double)));
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p;
      o.p = nullptr;
    }
    return *this;
  }
};

// Eigenpairs of the k x k Gram G (device, destroyed) with the reference's
// refinement of the small-eigenvalue subspace.  V (k x k, ld k) receives the
// eigenvectors, w the eigenvalues (ascending).
void gram_eig(Context& c, double* G, int k, double* V, std::vector<double>& w) {
  DevBuf Gc((long long)k * k);
  dev_copy(c, Gc.p, G, (long long)k * k);
  sym_eig(c, G, k, V, w);
  int npp = 0;
  for (int i = k - 1; i >= 0; --i)
    if (w[i] / w[k - 1] < 0.1) {
      npp = i + 1;
      break;
    }
  if (npp <= 1) {
    dev_fix_signs(c, V, k, k, k);
    return;
  }
  // ap: current projected Gram (np x np), vp: its eigenvectors (np x np)
  DevBuf ap((long long)k * k), vp((long long)k * k);
  dev_copy(c, ap.p, Gc.p, (long long)k * k);
  dev_copy(c, vp.p, V, (long long)k * k);
  int np = k;
  while (true) {
    DevBuf X((long long)npp * np), app((long long)npp * npp), appc((long long)npp * npp);
    DevBuf vpp((long long)npp * npp), T((long long)k * npp);
    {
      GemmBatch b;  // X = vp[:, :npp]^T ap
      b.add(X.p, npp, npp, np, vp.p, np, 1, ap.p, np, 0, np);
      b.run(c);
    }
    {
      GemmBatch b;  // app = X vp[:, :npp]
      b.add(app.p, npp, npp, npp, X.p, npp, 0, vp.p, np, 0, np);
      b.run(c);
    }
    dev_copy(c, appc.p, app.p, (long long)npp * npp);
    std::vector<double> ws;
    sym_eig(c, app.p, npp, vpp.p, ws);
    for (int i = 0; i < npp; ++i) w[i] = ws[i];
    {
      GemmBatch b;  // V[:, :npp] = V[:, :npp] vpp
      b.add(T.p, k, k, npp, V, k, 0, vpp.p, npp, 0, npp);
      b.run(c);
    }
    dev_copy(c, V, T.p, (long long)k * npp);
    dev_copy(c, ap.p, appc.p, (long long)npp * npp);
    dev_copy(c, vp.p, vpp.p, (long long)npp * npp);
    np = npp;
    int j = 0;
    for (int i = npp - 1; i >= 0; --i)
      if (ws[i] / ws[npp - 1] < 0.1) {
        j = i + 1;
        break;
      }
    npp = j;
    if (npp <= 1) break;
  }
  dev_fix_signs(c, V, k, k, k);
  SU2_HIP(hipStreamSynchronize(c.stream()));
}

}  // namespace

SvdResult truncated_svd(Context& c, const Tensor& psi, double p1, double p2, int max_dcut) {
  c.make_current();
  SU2_CHECK(psi.nleg() == 2, kErrStructure, "svd: wavefunction must have two legs");
  SU2_CHECK(max_dcut >= 1, kErrArgument, "svd: max_dcut=%d", max_dcut);
  const double tol = 1e-100;
  struct Sec {
    int twoj, m, n, k, dc;
    const double* A;
    DevBuf G, V, U;
    std::vector<double> s;   // k values, ascending, zeros for cut
  };
  std::vector<Sec> secs;
  for (auto& b : psi.blocks()) {
    Sec s;
    s.twoj = b.twoj[0];
    s.m = b.dim[0];
    s.n = b.dim[1];
    s.k = std::min(s.m, s.n);
    s.dc = 0;
    s.A = psi.data() + b.offset;
    secs.push_back(std::move(s));
  }
  // Gram matrices of all sectors in one grouped launch
  {
    GemmBatch g;
    for (auto& s : secs) {
      if (s.k == 0) continue;
      s.G = DevBuf((long long)s.k * s.k);
      if (s.m >= s.n) g.add(s.G.p, s.n, s.n, s.n, s.A, s.m, 1, s.A, s.m, 0, s.m);
      else g.add(s.G.p, s.m, s.m, s.m, s.A, s.m, 0, s.A, s.m, 1, s.n);
    }
    g.run(c);
  }
  std::vector<std::vector<double>> wv(secs.size());
  for (size_t i = 0; i < secs.size(); ++i) {
    Sec& s = secs[i];
    if (s.k == 0) continue;
    s.V = DevBuf((long long)s.k * s.k);
    gram_eig(c, s.G.p, s.k, s.V.p, wv[i]);
  }
  // the other factor: U = A V (m >= n) or V = A^T U (m < n); both stored as
  // "X" (the long side) with the eigenvectors in "E" (the short side)
  {
    GemmBatch g;
    for (auto& s : secs) {
      if (s.k == 0) continue;
      if (s.m >= s.n) {
        s.U = DevBuf((long long)s.m * s.k);
        g.add(s.U.p, s.m, s.m, s.k, s.A, s.m, 0, s.V.p, s.k, 0, s.k);
      } else {
        s.U = DevBuf((long long)s.n * s.k);
        g.add(s.U.p, s.n, s.n, s.k, s.A, s.m, 1, s.V.p, s.k, 0, s.k);
      }
    }
    g.run(c);
  }
  // singular values and column scaling (tensor.cpp:1533-1568, 1620-1655)
  std::vector<double> wsort;
  std::vector<int> isort;
  double* dfac = c.reduce_scratch(1);  // grown below
  for (size_t i = 0; i < secs.size(); ++i) {
    Sec& s = secs[i];
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
    // factors for the long-side matrix (held in U) and the eigenvectors (V)
    std::vector<double> fl(k, 0.0), fe(k, 0.0);
    const double pl = s.m >= s.n ? p1 - 1.0 : p2 - 1.0;  // long side: U (m>=n) or V (m<n)
    const double pe = s.m >= s.n ? p2 : p1;
    for (int t = k - dc; t < k; ++t) {
      if (s.s[t] > 0) {
        fl[t] = std::pow(s.s[t], pl);
        fe[t] = std::pow(s.s[t], pe);
      }
    }
    dfac = c.reduce_scratch(2 * k);
    std::vector<double> both(fl);
    both.insert(both.end(), fe.begin(), fe.end());
    SU2_HIP(hipMemcpyAsync(dfac, both.data(), 2 * k * sizeof(double), hipMemcpyHostToDevice,
                           c.stream()));
    const int longrows = s.m >= s.n ? s.m : s.n;
    dev_scale_columns(c, s.U.p, longrows, k, dfac);
    dev_scale_columns(c, s.V.p, k, k, dfac + k);
    SU2_HIP(hipStreamSynchronize(c.stream()));
    for (int t = 0; t < dc; ++t) {
      wsort.push_back(s.s[k - dc + t]);
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
    long long ldu, ldv;
    if (s.m >= s.n) {
      Ucols = s.U.p + (long long)(k - d) * s.m;
      ldu = s.m;
      Vcols = s.V.p + (long long)(k - d) * k;
      ldv = k;
    } else {
      Ucols = s.V.p + (long long)(k - d) * k;
      ldu = k;
      Vcols = s.U.p + (long long)(k - d) * s.n;
      ldv = s.n;
    }
    const int bu = res.U->find2(s.twoj, s.twoj), bv = res.V->find2(s.twoj, s.twoj);
    SU2_CHECK(bu >= 0 && bv >= 0, kErrStructure, "svd: missing output block");
    (void)ldu;
    (void)ldv;
    dev_copy(c, res.U->data() + res.U->blocks()[bu].offset, Ucols, (long long)s.m * d);
    dev_copy(c, res.V->data() + res.V->blocks()[bv].offset, Vcols, (long long)s.n * d);
    std::vector<double> kv(s.s.begin() + (k - d), s.s.end());
    res.kept.push_back(kv);
  }
  SU2_HIP(hipStreamSynchronize(c.stream()));
  return res;
}

}  // namespace su2
