// Plaquette (Q) terms of the J-Q-J2 model in the device environments.
//
// For every plaquette p0 < p1 < p2 < p3 (Lattice::plq) the reference adds
//   -Q [ (S_p0.S_p1)(S_p2.S_p3) + (S_p0.S_p3)(S_p1.S_p2) ]
// and grows each channel site by site as a partial operator plqopr
// (dmrg_su2_omp.cpp:411-447 from the left, 553-590 from the right; the
// bilinear parts of the singlet projectors are in coup[1] and coup[3]).
// Channel 0 uses the reference's own kernels: the pair-up of S_p0 with s_p1
// (a scalar), transformations, the scalar times s_p2 (a vector), and the
// pair-up with s_p3 that completes the term into the block Hamiltonian.
// Channel 1 couples S_p0 (x) s_p1 to the ranks K = 0, 1, 2 at p1 (the
// reference's permutation with the qterm tensors carries the same ranks on
// one operator leg), transforms each rank, and reduces them with s_p2 to the
// vector S_p0 (s_p1.s_p2) at p2.  The right side runs p3 -> p0 in mirror.
// Every step is a batched strided scale-add with the recoupling factors of
// recoupling.hpp; the sums into one destination run in a fixed order.
#include <cmath>

#include "blas1.hpp"
#include "common.hpp"
#include "context.hpp"
#include "dmrg.hpp"
#include "heff.hpp"
#include "ops.hpp"
#include "recoupling.hpp"

namespace su2 {

namespace {

constexpr int kPhys = 1;
constexpr int kRanks[3] = {0, 2, 4};
const double kSigmaLeft = std::sqrt(3.0) / 2;
const double kSigmaRight = -std::sqrt(3.0) / 2;

Leg with_dir(Leg l, int d) {
  l.dir = d;
  return l;
}

// zeroed by zb.run() (one launch for all of a step's new operators)
std::unique_ptr<Tensor> rank_op(Context& c, const Leg& l, int k, ZeroBatch& zb) {
  std::unique_ptr<Tensor> t(new Tensor(
      &c, {with_dir(l, 1), make_leg({k}, {1}, 1), with_dir(l, -1)}, 0, nullptr, false));
  zb.add(*t);
  return t;
}

int op_rank(const Tensor& t) { return t.nleg() == 3 ? t.leg(1).twoj[0] : 0; }

// block id of (jf, jfp) in a two-leg scalar or three-leg rank-k operator
int block_of(const Tensor& t, int jf, int jfp) {
  return t.nleg() == 2 ? t.find2(jf, jfp) : t.find3(jf, op_rank(t), jfp);
}

void add_part(ScatterBatch& sb, Tensor& dst, int db, long long o, long long op_, const double* src,
              int rows, int cols, double alpha, long long Df) {
  StridedAxpy a{};
  a.dst = dst.data() + dst.blocks()[db].offset + o + op_ * Df;
  a.src = src;
  a.rs = 1;
  a.cs = Df;
  a.ld = rows;
  a.rows = rows;
  a.cols = cols;
  a.alpha = alpha;
  sb.add(a.dst, a);
}

// dst[jf, jf] += scale * sig * pairup(V (x) s): a rank-0 result (operator_pairup)
void pair_into(ScatterBatch& sb, const Fused& E, const Tensor& V, double sig, double scale,
               bool left, Tensor& dst) {
  for (int jf : E.fused.twoj) {
    const int db = block_of(dst, jf, jf);
    if (db < 0) continue;
    const long long Df = E.dim(jf);
    auto parts = E.parts(jf);
    for (auto& p : parts)
      for (auto& q : parts) {
        const int jb = std::get<0>(p), jbp = std::get<0>(q);
        const double f = left ? rc_pairup_left(jf, jb, jbp, 2, kPhys, kPhys)
                              : rc_pairup_right(jf, jb, jbp, 2, kPhys, kPhys);
        if (std::fabs(f) < 1e-8) continue;
        const int vb = V.find3(jb, 2, jbp);
        if (vb < 0) continue;
        add_part(sb, dst, db, std::get<1>(p), std::get<1>(q), V.data() + V.blocks()[vb].offset,
                 std::get<2>(p), std::get<2>(q), scale * f * sig, Df);
      }
  }
}

// The rank-kd operator dst from src (rank ks) on the block and the site
// spin: factor(jb, jf, jbp, jfp) * alpha for every pair of parts.
template <class F>
void site_product_into(ScatterBatch& sb, const Fused& E, const Tensor& src, double alpha,
                       F&& factor, bool same_part, Tensor& dst) {
  const int ks = op_rank(src);
  for (int db = 0; db < (int)dst.blocks().size(); ++db) {
    const Block& B = dst.blocks()[db];
    const int jf = B.twoj[0], jfp = B.twoj[2];
    const long long Df = E.dim(jf);
    for (auto& p : E.parts(jf))
      for (auto& q : E.parts(jfp)) {
        const int jb = std::get<0>(p), jbp = std::get<0>(q);
        if (same_part && jb != jbp) continue;
        const double f = factor(jb, jf, jbp, jfp);
        if (std::fabs(f) < 1e-12) continue;
        const int sbk = src.nleg() == 2 ? src.find2(jb, jbp) : src.find3(jb, ks, jbp);
        if (sbk < 0) continue;
        add_part(sb, dst, db, std::get<1>(p), std::get<1>(q), src.data() + src.blocks()[sbk].offset,
                 std::get<2>(p), std::get<2>(q), alpha * f, Df);
      }
  }
}

}  // namespace

void Dmrg::grow_plaquettes(const Fused& E, const Env& env, int i, bool left, Env& out) {
  if (lat_.plq.empty()) return;
  Context& c = c_;
  const double Q = lat_.qcoup;
  const double sig_site = left ? kSigmaLeft : kSigmaRight;
  const double sig_pair = left ? kSigmaRight : kSigmaLeft;
  ScatterBatch sb;
  ZeroBatch zb;
  for (int m = 0; m < (int)lat_.plq.size(); ++m) {
    const auto& p = lat_.plq[m];
    if (i < p[0] || i > p[3]) continue;
    const int a = left ? p[0] : p[3], b = left ? p[1] : p[2];
    const int cc = left ? p[2] : p[1], d = left ? p[3] : p[0];
    const bool beyond_b = left ? i > b : i < b;
    if (i != b && !beyond_b) continue;
    for (int ch = 0; ch < 2; ++ch) {
      const std::pair<int, int> key{m, ch};
      if (i == b) {
        auto it = env.S.find(a);
        SU2_CHECK(it != env.S.end(), kErrStructure, "plaquette %d: operator of site %d missing", m,
                  a);
        const Tensor& Sa = *it->second;
        if (ch == 0) {
          auto T0 = rank_op(c, E.fused, 0, zb);
          pair_into(sb, E, Sa, sig_pair, 1.0, left, *T0);
          out.P[key][0] = std::move(T0);
        } else {
          for (int K : kRanks) {
            auto TK = rank_op(c, E.fused, K, zb);
            site_product_into(sb, E, Sa, sig_site,
                              [&](int jb, int jf, int jbp, int jfp) {
                                return left ? rc_couple_left(jb, jf, jbp, jfp, K, kPhys)
                                            : rc_couple_right(jb, jf, jbp, jfp, K, kPhys);
                              },
                              false, *TK);
            out.P[key][K] = std::move(TK);
          }
        }
        continue;
      }
      auto pit = env.P.find(key);
      SU2_CHECK(pit != env.P.end(), kErrStructure, "plaquette %d: partial operator missing at %d",
                m, i);
      const auto& prev = pit->second;
      if (i == cc) {
        auto T2 = rank_op(c, E.fused, 2, zb);
        if (ch == 0) {
          // scalar (x) s_p2: operator_initial with a block operator
          site_product_into(sb, E, *prev.at(0), sig_site,
                            [&](int jb, int jf, int, int jfp) {
                              return left ? rc_op_initial_left(jb, jf, jfp, 2, kPhys, kPhys)
                                          : rc_op_initial_right(jb, jf, jfp, 2, kPhys, kPhys);
                            },
                            true, *T2);
        } else {
          for (auto& kv : prev) {
            const int K = kv.first;
            site_product_into(sb, E, *kv.second, rc_reduce_weight(K) * sig_site,
                              [&](int jb, int jf, int jbp, int jfp) {
                                return left ? rc_reduce_left(jb, jf, jbp, jfp, K, kPhys)
                                            : rc_reduce_right(jb, jf, jbp, jfp, K, kPhys);
                              },
                              false, *T2);
          }
        }
        out.P[key][2] = std::move(T2);
      } else if (i == d) {
        // completed: -Q (...) into the block Hamiltonian
        pair_into(sb, E, *prev.at(2), sig_pair, -Q, left, *out.H);
      } else {
        for (auto& kv : prev) {
          const int K = kv.first;
          auto TK = rank_op(c, E.fused, K, zb);
          site_product_into(sb, E, *kv.second, 1.0,
                            [&](int jb, int jf, int jbp, int jfp) {
                              return left ? rc_op_transform_left(jb, jf, jbp, jfp, K, kPhys)
                                          : rc_op_transform_right(jb, jf, jbp, jfp, K, kPhys);
                            },
                            false, *TK);
          out.P[key][K] = std::move(TK);
        }
      }
    }
  }
  zb.run(c);
  sb.run(c);
}

// dmrg_su2.cpp:635-655: only on adjacent sites; the right side carries -Q
void Dmrg::plaquette_terms(int il, int ir, const Env& El, const Env& Er,
                           std::vector<HeffTerm>& terms) const {
  if (lat_.plq.empty() || il + 1 != ir) return;
  const double Q = lat_.qcoup;
  for (int m = 0; m < (int)lat_.plq.size(); ++m) {
    const auto& p = lat_.plq[m];
    if (!(p[0] <= il && ir <= p[3])) continue;
    for (int ch = 0; ch < 2; ++ch) {
      const std::pair<int, int> key{m, ch};
      if (ir <= p[1]) {
        HeffTerm t;
        t.left.push_back({1.0, El.S.at(p[0]).get()});
        t.right.push_back({-Q, Er.P.at(key).at(2).get()});
        terms.push_back(t);
      } else if (il >= p[2]) {
        HeffTerm t;
        t.left.push_back({1.0, El.P.at(key).at(2).get()});
        t.right.push_back({-Q, Er.S.at(p[3]).get()});
        terms.push_back(t);
      } else {
        for (auto& kv : El.P.at(key)) {
          HeffTerm t;
          t.k = kv.first;
          t.left.push_back({1.0, kv.second.get()});
          t.right.push_back({-Q, Er.P.at(key).at(kv.first).get()});
          terms.push_back(t);
        }
      }
    }
  }
}

}  // namespace su2
