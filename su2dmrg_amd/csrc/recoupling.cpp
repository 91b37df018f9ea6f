#include "recoupling.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "cg.hpp"
#include "common.hpp"

namespace su2 {

namespace {

bool tri(int a, int b, int c) { return std::abs(a - b) <= c && c <= a + b && ((a + b + c) % 2 == 0); }

// dense rank-3 tensor, column-major (index 0 fastest)
struct T3 {
  int n0, n1, n2;
  std::vector<double> v;
  T3(int a, int b, int c) : n0(a), n1(b), n2(c), v((size_t)a * b * c, 0.0) {}
  double& operator()(int i, int j, int k) { return v[i + (size_t)n0 * (j + (size_t)n1 * k)]; }
  double operator()(int i, int j, int k) const { return v[i + (size_t)n0 * (j + (size_t)n1 * k)]; }
};

T3 cg(int j1, int j2, int j3) {
  T3 t(j1 + 1, j2 + 1, j3 + 1);
  t.v = cg_tensor(j1, j2, j3);
  return t;
}

// tensor::is_proportional_to (tensor.cpp:1231-1269): the ratio at the first
// largest reference element.  Every element is checked against r * ref to
// 1e-10 of the largest element: the contraction of five CG tensors at spins
// 2j ~ 30 and above carries rounding of a few 1e-13 on elements that are
// themselves ~1e-12, which an element-wise relative test would reject.
double ratio(const std::vector<double>& t, const std::vector<double>& ref) {
  const double tol = 1e-12;
  bool zero = true;
  for (double x : t) zero = zero && std::fabs(x) <= tol;
  if (zero) return 0.0;
  size_t pos = 0;
  double mx = 0;
  for (size_t i = 0; i < ref.size(); ++i)
    if (std::fabs(ref[i]) > mx) {
      mx = std::fabs(ref[i]);
      pos = i;
    }
  SU2_CHECK(mx > tol, kErrNumeric, "recoupling: zero reference tensor");
  const double r = t[pos] / ref[pos];
  double tmax = 0;
  for (double x : t) tmax = std::max(tmax, std::fabs(x));
  for (size_t i = 0; i < t.size(); ++i)
    SU2_CHECK(std::fabs(t[i] - r * ref[i]) <= 1e-10 * tmax, kErrNumeric,
              "recoupling: contraction not proportional to the output CG");
  return r;
}

std::mutex g_mu;
std::map<std::array<int, 7>, double> g_cache;

template <class F>
double cached(std::array<int, 7> key, F&& f) {
  {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) return it->second;
  }
  double v = f();
  std::lock_guard<std::mutex> g(g_mu);
  g_cache[key] = v;
  return v;
}

// T[x,p,y] = sqrt(k+1) sum_q C(s0,k,s1)[x,q,y] singlet_k[p,q]  (dmrg_su2.cpp:924-926)
T3 turned_site(int s0, int k, int s1) {
  T3 c = cg(s0, k, s1), t(s0 + 1, k + 1, s1 + 1);
  std::vector<double> sk = singlet_tensor(k);
  const double f = std::sqrt((double)k + 1);
  for (int x = 0; x <= s0; ++x)
    for (int p = 0; p <= k; ++p)
      for (int y = 0; y <= s1; ++y) {
        double s = 0;
        for (int q = 0; q <= k; ++q) s += c(x, q, y) * sk[p + (size_t)(k + 1) * q];
        t(x, p, y) = s * f;
      }
  return t;
}

std::vector<double> identity_tensor(int j) {
  std::vector<double> v((size_t)(j + 1) * (j + 1), 0.0);
  for (int i = 0; i <= j; ++i) v[i + (size_t)(j + 1) * i] = 1.0;
  return v;
}

}  // namespace

double rc_op_transform_left(int jl, int ja, int jlp, int jap, int k, int s) {
  if (!(tri(jl, s, ja) && tri(jlp, s, jap) && tri(jl, k, jlp) && tri(ja, k, jap))) return 0.0;
  return cached({1, jl, ja, jlp, jap, k, s}, [&] {
    T3 c1 = cg(jl, s, ja), c2 = cg(jl, k, jlp), c3 = cg(jlp, s, jap), out(ja + 1, k + 1, jap + 1);
    for (int a = 0; a <= ja; ++a)
      for (int q = 0; q <= k; ++q)
        for (int b = 0; b <= jap; ++b) {
          double acc = 0;
          for (int l = 0; l <= jl; ++l)
            for (int x = 0; x <= s; ++x) {
              const double u = c1(l, x, a);
              if (u == 0) continue;
              for (int m = 0; m <= jlp; ++m) acc += u * c2(l, q, m) * c3(m, x, b);
            }
          out(a, q, b) = acc;
        }
    return ratio(out.v, cg_tensor(ja, k, jap));
  });
}

double rc_op_transform_right(int jr, int ja, int jrp, int jap, int k, int s) {
  if (!(tri(s, jr, ja) && tri(s, jrp, jap) && tri(jr, k, jrp) && tri(ja, k, jap))) return 0.0;
  return cached({2, jr, ja, jrp, jap, k, s}, [&] {
    T3 c1 = cg(s, jr, ja), c2 = cg(jr, k, jrp), c3 = cg(s, jrp, jap), out(ja + 1, k + 1, jap + 1);
    for (int a = 0; a <= ja; ++a)
      for (int q = 0; q <= k; ++q)
        for (int b = 0; b <= jap; ++b) {
          double acc = 0;
          for (int x = 0; x <= s; ++x)
            for (int r = 0; r <= jr; ++r) {
              const double u = c1(x, r, a);
              if (u == 0) continue;
              for (int m = 0; m <= jrp; ++m) acc += u * c2(r, q, m) * c3(x, m, b);
            }
          out(a, q, b) = acc;
        }
    return ratio(out.v, cg_tensor(ja, k, jap));
  });
}

double rc_op_initial_left(int jl, int ja, int jap, int k, int s0, int s1) {
  if (!(tri(jl, s0, ja) && tri(jl, s1, jap) && tri(s0, k, s1) && tri(ja, k, jap))) return 0.0;
  return cached({3, jl, ja, jap, k, s0, s1}, [&] {
    T3 c1 = cg(jl, s0, ja), c2 = cg(s0, k, s1), c3 = cg(jl, s1, jap), out(ja + 1, k + 1, jap + 1);
    for (int a = 0; a <= ja; ++a)
      for (int q = 0; q <= k; ++q)
        for (int b = 0; b <= jap; ++b) {
          double acc = 0;
          for (int l = 0; l <= jl; ++l)
            for (int x = 0; x <= s0; ++x)
              for (int y = 0; y <= s1; ++y) acc += c1(l, x, a) * c2(x, q, y) * c3(l, y, b);
          out(a, q, b) = acc;
        }
    return ratio(out.v, cg_tensor(ja, k, jap));
  });
}

double rc_op_initial_right(int jr, int ja, int jap, int k, int s0, int s1) {
  if (!(tri(s0, jr, ja) && tri(s1, jr, jap) && tri(s0, k, s1) && tri(ja, k, jap))) return 0.0;
  return cached({4, jr, ja, jap, k, s0, s1}, [&] {
    T3 c1 = cg(s0, jr, ja), c2 = cg(s0, k, s1), c3 = cg(s1, jr, jap), out(ja + 1, k + 1, jap + 1);
    for (int a = 0; a <= ja; ++a)
      for (int q = 0; q <= k; ++q)
        for (int b = 0; b <= jap; ++b) {
          double acc = 0;
          for (int x = 0; x <= s0; ++x)
            for (int r = 0; r <= jr; ++r)
              for (int y = 0; y <= s1; ++y) acc += c1(x, r, a) * c2(x, q, y) * c3(y, r, b);
          out(a, q, b) = acc;
        }
    return ratio(out.v, cg_tensor(ja, k, jap));
  });
}

double rc_pairup_left(int ja, int jl, int jlp, int k, int s0, int s1) {
  if (!(tri(jl, s0, ja) && tri(jlp, s1, ja) && tri(jl, k, jlp) && tri(s0, k, s1))) return 0.0;
  return cached({5, ja, jl, jlp, k, s0, s1}, [&] {
    T3 c1 = cg(jl, s0, ja), c2 = cg(jl, k, jlp), t = turned_site(s0, k, s1), c3 = cg(jlp, s1, ja);
    std::vector<double> out((size_t)(ja + 1) * (ja + 1), 0.0);
    for (int a = 0; a <= ja; ++a)
      for (int b = 0; b <= ja; ++b) {
        double acc = 0;
        for (int l = 0; l <= jl; ++l)
          for (int x = 0; x <= s0; ++x) {
            const double u = c1(l, x, a);
            if (u == 0) continue;
            for (int q = 0; q <= k; ++q)
              for (int m = 0; m <= jlp; ++m)
                for (int y = 0; y <= s1; ++y) acc += u * c2(l, q, m) * t(x, q, y) * c3(m, y, b);
          }
        out[a + (size_t)(ja + 1) * b] = acc;
      }
    return ratio(out, identity_tensor(ja));
  });
}

double rc_pairup_right(int ja, int jr, int jrp, int k, int s0, int s1) {
  if (!(tri(s0, jr, ja) && tri(s1, jrp, ja) && tri(jr, k, jrp) && tri(s0, k, s1))) return 0.0;
  return cached({6, ja, jr, jrp, k, s0, s1}, [&] {
    T3 c1 = cg(s0, jr, ja), c2 = cg(jr, k, jrp), t = turned_site(s0, k, s1), c3 = cg(s1, jrp, ja);
    std::vector<double> out((size_t)(ja + 1) * (ja + 1), 0.0);
    for (int a = 0; a <= ja; ++a)
      for (int b = 0; b <= ja; ++b) {
        double acc = 0;
        for (int x = 0; x <= s0; ++x)
          for (int r = 0; r <= jr; ++r) {
            const double u = c1(x, r, a);
            if (u == 0) continue;
            for (int q = 0; q <= k; ++q)
              for (int m = 0; m <= jrp; ++m)
                for (int y = 0; y <= s1; ++y) acc += u * c2(r, q, m) * t(x, q, y) * c3(y, m, b);
          }
        out[a + (size_t)(ja + 1) * b] = acc;
      }
    return ratio(out, identity_tensor(ja));
  });
}

namespace {

// Block operator (rank kb, CG C(j,kb,jp)) times the site spin (C(s,2,s)),
// their operator legs coupled by C(kb,2,ko), in the enlarged basis:
//   out[a,qo,b] = sum F(l,x,a) C(j,kb,jp)[l,qb,m] C(s,2,s)[x,p,y]
//                     C(kb,2,ko)[qb,p,qo] F'(m,y,b)
// with F = C(j,s,ja)[l,x,a] on the left and C(s,j,ja)[x,l,a] on the right.
double coupled_site(bool right, int j, int ja, int jp, int jap, int kb, int ko, int s) {
  T3 f1 = right ? cg(s, j, ja) : cg(j, s, ja);
  T3 f2 = right ? cg(s, jp, jap) : cg(jp, s, jap);
  T3 bo = cg(j, kb, jp), st = cg(s, 2, s), cc = cg(kb, 2, ko);
  // W[x,y,qb,qo] = sum_p C(s,2,s)[x,p,y] C(kb,2,ko)[qb,p,qo]
  const int ns = s + 1, nb = kb + 1, no = ko + 1;
  std::vector<double> W((size_t)ns * ns * nb * no, 0.0);
  for (int x = 0; x < ns; ++x)
    for (int y = 0; y < ns; ++y)
      for (int qb = 0; qb < nb; ++qb)
        for (int qo = 0; qo < no; ++qo) {
          double acc = 0;
          for (int p = 0; p <= 2; ++p) acc += st(x, p, y) * cc(qb, p, qo);
          W[x + ns * (y + ns * (qb + (size_t)nb * qo))] = acc;
        }
  auto F1 = [&](int l, int x, int a) { return right ? f1(x, l, a) : f1(l, x, a); };
  auto F2 = [&](int m, int y, int b) { return right ? f2(y, m, b) : f2(m, y, b); };
  T3 out(ja + 1, ko + 1, jap + 1);
  for (int a = 0; a <= ja; ++a)
    for (int qo = 0; qo < no; ++qo)
      for (int b = 0; b <= jap; ++b) {
        double acc = 0;
        for (int l = 0; l <= j; ++l)
          for (int x = 0; x < ns; ++x) {
            const double u = F1(l, x, a);
            if (u == 0) continue;
            for (int m = 0; m <= jp; ++m)
              for (int y = 0; y < ns; ++y) {
                const double v = F2(m, y, b);
                if (v == 0) continue;
                for (int qb = 0; qb < nb; ++qb)
                  acc += u * bo(l, qb, m) * W[x + ns * (y + ns * (qb + (size_t)nb * qo))] * v;
              }
          }
        out(a, qo, b) = acc;
      }
  return ratio(out.v, cg_tensor(ja, ko, jap));
}

}  // namespace

double rc_couple_left(int jl, int ja, int jlp, int jap, int K, int s) {
  if (!(tri(jl, s, ja) && tri(jlp, s, jap) && tri(jl, 2, jlp) && tri(ja, K, jap))) return 0.0;
  return cached({9, jl, ja, jlp, jap, K, s},
                [&] { return coupled_site(false, jl, ja, jlp, jap, 2, K, s); });
}

double rc_couple_right(int jr, int ja, int jrp, int jap, int K, int s) {
  if (!(tri(s, jr, ja) && tri(s, jrp, jap) && tri(jr, 2, jrp) && tri(ja, K, jap))) return 0.0;
  return cached({10, jr, ja, jrp, jap, K, s},
                [&] { return coupled_site(true, jr, ja, jrp, jap, 2, K, s); });
}

double rc_reduce_left(int jl, int ja, int jlp, int jap, int K, int s) {
  if (!(tri(jl, s, ja) && tri(jlp, s, jap) && tri(jl, K, jlp) && tri(ja, 2, jap))) return 0.0;
  return cached({11, jl, ja, jlp, jap, K, s},
                [&] { return coupled_site(false, jl, ja, jlp, jap, K, 2, s); });
}

double rc_reduce_right(int jr, int ja, int jrp, int jap, int K, int s) {
  if (!(tri(s, jr, ja) && tri(s, jrp, jap) && tri(jr, K, jrp) && tri(ja, 2, jap))) return 0.0;
  return cached({12, jr, ja, jrp, jap, K, s},
                [&] { return coupled_site(true, jr, ja, jrp, jap, K, 2, s); });
}

double rc_reduce_weight(int K) {
  switch (K) {
    case 0: return -1.0 / std::sqrt(3.0);
    case 2: return 1.0;
    case 4: return -std::sqrt(5.0 / 3.0);
    default: fail(kErrArgument, "reduce weight: operator rank 2K=%d", K);
  }
  return 0.0;
}

namespace {

// src: singlet between l and F=(s r)->l;  dst: (l s)->E=r then singlet to r
void psi_pair(int jl, int s, int jr, std::vector<double>& a, std::vector<double>& b) {
  T3 cf = cg(s, jr, jl), ce = cg(jl, s, jr);
  std::vector<double> sl = singlet_tensor(jl), sr = singlet_tensor(jr);
  const int nl = jl + 1, ns = s + 1, nr = jr + 1;
  a.assign((size_t)nl * ns * nr, 0.0);
  b.assign((size_t)nl * ns * nr, 0.0);
  for (int l = 0; l < nl; ++l)
    for (int x = 0; x < ns; ++x)
      for (int r = 0; r < nr; ++r) {
        double u = 0, v = 0;
        for (int f = 0; f < nl; ++f) u += sl[l + (size_t)nl * f] * cf(x, r, f);
        for (int e = 0; e < nr; ++e) v += ce(l, x, e) * sr[e + (size_t)nr * r];
        a[l + (size_t)nl * (x + (size_t)ns * r)] = u;
        b[l + (size_t)nl * (x + (size_t)ns * r)] = v;
      }
}

}  // namespace

double rc_psi_left_to_right(int jl, int s, int jr) {
  if (!tri(jl, s, jr)) return 0.0;
  return cached({7, jl, s, jr, 0, 0, 0}, [&] {
    std::vector<double> src, dst;
    psi_pair(jl, s, jr, src, dst);
    double num = 0, den = 0;
    for (size_t i = 0; i < src.size(); ++i) {
      num += src[i] * dst[i];
      den += dst[i] * dst[i];
    }
    return num / den;
  });
}

double rc_psi_right_to_left(int jl, int s, int jr) {
  if (!tri(jl, s, jr)) return 0.0;
  return cached({8, jl, s, jr, 0, 0, 0}, [&] {
    std::vector<double> src, dst;
    psi_pair(jl, s, jr, dst, src);
    double num = 0, den = 0;
    for (size_t i = 0; i < src.size(); ++i) {
      num += src[i] * dst[i];
      den += dst[i] * dst[i];
    }
    return num / den;
  });
}

}  // namespace su2
