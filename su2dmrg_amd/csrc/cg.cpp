#include "cg.hpp"

#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

namespace su2 {

static double lfact(int n) { return std::lgamma((double)n + 1.0); }

std::vector<double> cg_tensor(int j1, int j2, int j3) {
  const int n1 = j1 + 1, n2 = j2 + 1, n3 = j3 + 1;
  std::vector<double> t((size_t)n1 * n2 * n3, 0.0);
  if (j3 < std::abs(j1 - j2) || j3 > j1 + j2 || ((j1 + j2 + j3) & 1)) return t;
  // log-space prefactors keep the factorials finite up to 2j = 38
  const double pre = 0.5 * (std::log((double)(j3 + 1)) + lfact((j3 + j1 - j2) / 2) +
                            lfact((j3 - j1 + j2) / 2) + lfact((j1 + j2 - j3) / 2) -
                            lfact((j1 + j2 + j3) / 2 + 1));
  for (int a = 0; a < n1; ++a) {
    const int m1 = 2 * a - j1;
    for (int b = 0; b < n2; ++b) {
      const int m2 = 2 * b - j2;
      const int m3 = m1 + m2;
      if (m3 < -j3 || m3 > j3) continue;
      const int c = (m3 + j3) / 2;
      const double fm = 0.5 * (lfact((j1 + m1) / 2) + lfact((j1 - m1) / 2) + lfact((j2 + m2) / 2) +
                               lfact((j2 - m2) / 2) + lfact((j3 + m3) / 2) + lfact((j3 - m3) / 2));
      const int k1 = (j1 + j2 - j3) / 2, k2 = (j1 - m1) / 2, k3 = (j2 + m2) / 2;
      const int k4 = (j3 - j2 + m1) / 2, k5 = (j3 - j1 - m2) / 2;
      double s = 0.0;
      for (int z = 0; z <= (j1 + j2 + j3) / 2; ++z) {
        if (k1 - z < 0 || k2 - z < 0 || k3 - z < 0 || k4 + z < 0 || k5 + z < 0) continue;
        const double lz = lfact(z) + lfact(k1 - z) + lfact(k2 - z) + lfact(k3 - z) +
                          lfact(k4 + z) + lfact(k5 + z);
        const double term = std::exp(pre + fm - lz);
        s += (z & 1) ? -term : term;
      }
      t[a + (size_t)n1 * (b + (size_t)n2 * c)] = s;
    }
  }
  return t;
}

std::vector<double> singlet_tensor(int j) {
  const int n = j + 1;
  std::vector<double> t((size_t)n * n, 0.0);
  const double v = 1.0 / std::sqrt((double)n);
  for (int a = 0; a < n; ++a) {
    const int m = 2 * a - j;
    t[a + (size_t)n * (j - a)] = (((j - m) / 2) % 2 == 0) ? v : -v;
  }
  return t;
}

double heff_recoupling(int i, int k, int f) {
  if (f < std::abs(i - k) || f > i + k || (i % 2) != (f % 2)) return 0.0;
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, double> cache;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find({i, k, f});
    if (it != cache.end()) return it->second;
  }
  const int ni = i + 1, nk = k + 1, nf = f + 1;
  std::vector<double> c = cg_tensor(i, k, f);
  std::vector<double> si = singlet_tensor(i), sk = singlet_tensor(k), sf = singlet_tensor(f);
  // turned(a',b',c') = sum C(a,b,c) s_i(a',a) s_k(b',b) s_f(c',c); each
  // singlet has one nonzero per row, so the sum collapses to one term.
  double acc = 0.0;
  for (int a2 = 0; a2 < ni; ++a2)
    for (int b2 = 0; b2 < nk; ++b2)
      for (int c2 = 0; c2 < nf; ++c2) {
        const int a = i - a2, b = k - b2, cc = f - c2;
        const double turned = c[a + (size_t)ni * (b + (size_t)nk * cc)] * si[a2 + (size_t)ni * a] *
                              sk[b2 + (size_t)nk * b] * sf[c2 + (size_t)nf * cc];
        acc += c[a2 + (size_t)ni * (b2 + (size_t)nk * c2)] * turned;
      }
  const double v = acc * std::sqrt((double)(k + 1));
  std::lock_guard<std::mutex> g(mu);
  cache[{i, k, f}] = v;
  return v;
}

}  // namespace su2
