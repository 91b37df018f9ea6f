#include "heff.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>

#include "blas1.hpp"
#include "cg.hpp"
#include "common.hpp"
#include "context.hpp"

namespace su2 {

namespace {

constexpr double kFacSkip = 1e-12;    // tensor_su2_dmrg_src.cpp:996
// work units per launch (~8 and ~16 per CU on 256 CUs): stage 2 sums long K
// ranges and splits them, and finer units balance its XCDs better (measured
// on the 16x8 D=2000 centre bond: 4.35 ms at 2048 units, 3.99 ms at 4096,
// flat up to 12288)
constexpr int kTargetUnits1 = 2048;
constexpr int kTargetUnits2 = 4096;

void check_op(const Tensor& op, const Leg& in, const Leg& out, int k, const char* what) {
  SU2_CHECK(op.nleg() == 3, kErrStructure, "%s operator must have 3 legs", what);
  SU2_CHECK(op.leg(0).same_sectors(in), kErrStructure,
            "%s operator leg 0 does not match the wavefunction leg", what);
  SU2_CHECK(op.leg(2).same_sectors(out), kErrStructure,
            "%s operator leg 2 does not match the wavefunction leg", what);
  SU2_CHECK(op.leg(1).nsec() == 1 && op.leg(1).twoj[0] == k && op.leg(1).dim[0] == 1,
            kErrStructure, "%s operator leg 1 must be a single multiplet 2j=%d", what, k);
}

void check_scalar(const Tensor& h, const Leg& l, const char* what) {
  SU2_CHECK(h.nleg() == 2, kErrStructure, "%s must have 2 legs", what);
  SU2_CHECK(h.leg(0).same_sectors(l) && h.leg(1).same_sectors(l), kErrStructure,
            "%s legs do not match the wavefunction leg", what);
}

}  // namespace

HeffPlan::HeffPlan(Context& c, const Tensor& psi, const Tensor* hl, const Tensor* hr,
                   const std::vector<HeffTerm>& terms)
    : ctx_(c), pool_(c.pool()) {
  c.make_current();
  SU2_CHECK(psi.nleg() == 2 && psi.locspin() == 0, kErrStructure,
            "H_eff wavefunction must be a two-leg singlet tensor");
  const Leg& la = psi.leg(0);
  const Leg& lb = psi.leg(1);
  if (hl) check_scalar(*hl, la, "left block Hamiltonian");
  if (hr) check_scalar(*hr, lb, "right block Hamiltonian");
  const int T = (int)terms.size();
  for (auto& t : terms) {
    SU2_CHECK(!t.left.empty() && !t.right.empty(), kErrArgument, "H_eff term with an empty side");
    SU2_CHECK(t.k >= 0 && t.k % 2 == 0 && t.k <= 8, kErrArgument, "H_eff term operator leg %d",
              t.k);
    for (auto& p : t.left) check_op(*p.second, la, la, t.k, "left");
    for (auto& p : t.right) check_op(*p.second, lb, lb, t.k, "right");
  }
  psi_size_ = psi.size();

  // psi block lookup by 2j
  std::map<int, const Block*> pb;
  for (auto& b : psi.blocks()) pb[b.twoj[0]] = &b;

  // terms by operator leg: a transition (j -> j', k) bundles every term of
  // leg k, at its position in that list
  std::map<int, std::vector<int>> by_k;
  for (int ti = 0; ti < T; ++ti) by_k[terms[ti].k].push_back(ti);

  // transitions (j -> j', k) present on both sides of at least one term
  struct Tr {
    int ja, jap, k, nt;
    double fac;
    int Da, Dap, Db, Dbp;
    long long lb_off, rb_off, x_off;
  };
  std::vector<Tr> trs;
  {
    std::map<std::tuple<int, int, int>, int> seen;
    for (auto& t : terms)
      for (auto& p : t.left)
        for (auto& blk : p.second->blocks()) {
          const int ja = blk.twoj[0], jap = blk.twoj[2];
          if (!pb.count(ja) || !pb.count(jap)) continue;
          if (seen.count({ja, jap, t.k})) continue;
          const double f = heff_recoupling(ja, t.k, jap);
          if (std::fabs(f) < kFacSkip) continue;
          bool right_has = false;
          for (int tj : by_k[t.k])
            for (auto& q : terms[tj].right)
              if (q.second->find3(ja, t.k, jap) >= 0) right_has = true;
          if (!right_has) continue;
          seen[{ja, jap, t.k}] = 1;
          Tr tr{};
          tr.ja = ja;
          tr.jap = jap;
          tr.k = t.k;
          tr.nt = (int)by_k[t.k].size();
          tr.fac = f;
          tr.Da = la.dim_of(ja);
          tr.Dap = la.dim_of(jap);
          tr.Db = lb.dim_of(ja);
          tr.Dbp = lb.dim_of(jap);
          trs.push_back(tr);
        }
    std::sort(trs.begin(), trs.end(), [](const Tr& x, const Tr& y) {
      if (x.jap != y.jap) return x.jap < y.jap;
      if (x.ja != y.ja) return x.ja < y.ja;
      return x.k < y.k;
    });
  }

  // algorithmic flops in the reference's order (per term, per block present
  // on both sides), plus the block-Hamiltonian products
  for (auto& t : terms) {
    for (auto& tr : trs) {
      if (tr.k != t.k) continue;
      bool l = false, r = false;
      for (auto& p : t.left) l = l || p.second->find3(tr.ja, t.k, tr.jap) >= 0;
      for (auto& p : t.right) r = r || p.second->find3(tr.ja, t.k, tr.jap) >= 0;
      if (l && r)
        alg_flops_ += 2.0 * tr.Db * tr.Da * tr.Dap + 2.0 * tr.Dap * tr.Db * tr.Dbp;
    }
  }
  for (auto& b : psi.blocks()) {
    const double da = b.dim[0], db = b.dim[1];
    if (hl && hl->find2(b.twoj[0], b.twoj[0]) >= 0) alg_flops_ += 2.0 * da * da * db;
    if (hr && hr->find2(b.twoj[1], b.twoj[1]) >= 0) alg_flops_ += 2.0 * da * db * db;
  }

  // workspace: bundles and stage-1 outputs
  long long work = 0;
  for (auto& tr : trs) {
    tr.lb_off = work;
    work += (long long)tr.Da * tr.Dap * tr.nt;
    tr.rb_off = work;
    work += (long long)tr.nt * tr.Db * tr.Dbp;
    tr.x_off = work;
    work += (long long)tr.Dap * tr.nt * tr.Db;
  }
  // H_L of each sector transposed: every stage-2 slice is then op(A) = A,
  // one layout, which the grouped GEMM runs with its straight-line kernel
  std::map<int, long long> hlt_off;
  if (hl)
    for (auto& b : psi.blocks()) {
      const int j = b.twoj[0];
      if (hl->find2(j, j) >= 0 && !hlt_off.count(j)) {
        hlt_off[j] = work;
        work += (long long)b.dim[0] * b.dim[0];
      }
    }
  d_in_ = c.dev_alloc_n<double>(std::max<long long>(psi_size_, 1));
  d_out_ = c.dev_alloc_n<double>(std::max<long long>(psi_size_, 1));
  if (work > 0) d_work_ = c.dev_alloc_n<double>(work);
  dev_fill(c, d_work_, work, 0.0);

  // bundle assembly: each bundle element is the sum of its coupling's
  // summands in the order given (deterministic), written once; then H_L of
  // each sector transposed (source row r lands in column r)
  std::vector<BundleDesc> bundles;
  std::vector<BundleSrc> table;
  for (auto& tr : trs) {
    const std::vector<int>& tk = by_k.at(tr.k);
    for (int side = 0; side < 2; ++side) {
      size_t npos = 0;
      for (int ti : tk) npos = std::max(npos, (side ? terms[ti].right : terms[ti].left).size());
      BundleDesc d{};
      d.dst = d_work_ + (side ? tr.rb_off : tr.lb_off);
      d.src0 = (int)table.size();
      d.nt = tr.nt;
      d.npos = (int)npos;
      d.rows = side ? tr.Db : tr.Da;
      d.cols = side ? tr.Dbp : tr.Dap;
      d.interleaved = side;   // Rb rows run (term, row) with the term fastest
      for (int ti : tk) {
        const auto& ops = side ? terms[ti].right : terms[ti].left;
        for (size_t pos = 0; pos < npos; ++pos) {
          BundleSrc e{nullptr, 0.0};
          if (pos < ops.size()) {
            const Tensor& op = *ops[pos].second;
            const int id = op.find3(tr.ja, tr.k, tr.jap);
            if (id >= 0) e = BundleSrc{op.data() + op.blocks()[id].offset, ops[pos].first};
          }
          table.push_back(e);
        }
      }
      bundles.push_back(d);
    }
  }
  SU2_CHECK(table.size() < (1u << 31), kErrCapacity, "H_eff: %zu bundle summands", table.size());
  dev_bundle_gather(c, bundles, table);
  std::vector<TransposeDesc> transposes;
  for (auto& kv : hlt_off) {
    const Block& hb = hl->blocks()[hl->find2(kv.first, kv.first)];
    transposes.push_back(TransposeDesc{d_work_ + kv.second, hl->data() + hb.offset, hb.dim[0]});
  }
  dev_transpose_batch(c, transposes);

  // stage 1: X(j->j') = f * Lb^T psi(j)   [T*D_a(j') x D_b(j)]
  for (auto& tr : trs) {
    const Block* bj = pb[tr.ja];
    int p = stage1_.add_problem(d_work_ + tr.x_off, (long long)tr.nt * tr.Dap, tr.nt * tr.Dap,
                                tr.Db);
    stage1_.add_slice(p, d_work_ + tr.lb_off, tr.Da, 1, d_in_ + bj->offset, tr.Da, 0, tr.Da,
                      tr.fac);
  }
  // stage 2: psi'(j') = sum_j X(j->j') Rb(j->j') + H_L^T psi + psi H_R
  for (auto& b : psi.blocks()) {
    const int jap = b.twoj[0];
    const int Dap = b.dim[0], Dbp = b.dim[1];
    bool any = false;
    int p = -1;
    auto ensure = [&]() {
      if (p < 0) p = stage2_.add_problem(d_out_ + b.offset, Dap, Dap, Dbp);
      any = true;
    };
    for (auto& tr : trs) {
      if (tr.jap != jap) continue;
      ensure();
      stage2_.add_slice(p, d_work_ + tr.x_off, Dap, 0, d_work_ + tr.rb_off,
                        (long long)tr.nt * tr.Db, 0, tr.nt * tr.Db, 1.0);
    }
    if (hlt_off.count(jap)) {
      ensure();
      stage2_.add_slice(p, d_work_ + hlt_off.at(jap), Dap, 0, d_in_ + b.offset, Dap, 0, Dap, 1.0);
    }
    if (hr) {
      int id = hr->find2(b.twoj[1], b.twoj[1]);
      if (id >= 0) {
        ensure();
        stage2_.add_slice(p, d_in_ + b.offset, Dap, 0, hr->data() + hr->blocks()[id].offset, Dbp,
                          0, Dbp, 1.0);
      }
    }
    if (!any) zero_blocks_.push_back({b.offset, b.size});
  }
  stage1_.finalize(c, kTargetUnits1);
  stage2_.finalize(c, kTargetUnits2);
  SU2_HIP(hipStreamSynchronize(c.stream()));
}

HeffPlan::~HeffPlan() {
  pool_->free(d_in_);
  pool_->free(d_out_);
  pool_->free(d_work_);
}

void HeffPlan::apply_raw(const double* psi, double* out) {
  ctx_.make_current();
  dev_copy(ctx_, d_in_, psi, psi_size_);
  for (auto& z : zero_blocks_) dev_fill(ctx_, d_out_ + z.first, z.second, 0.0);
  stage1_.launch(ctx_.stream());
  stage2_.launch(ctx_.stream());
  dev_copy(ctx_, out, d_out_, psi_size_);
}

void HeffPlan::set_variants(int v1, int v2) {
  ctx_.make_current();
  stage1_.set_variant(v1);
  stage2_.set_variant(v2);
}

void HeffPlan::set_uniform(bool on) {
  ctx_.make_current();
  stage1_.allow_uniform(on);
  stage2_.allow_uniform(on);
  stage1_.set_variant(stage1_.variant());
  stage2_.set_variant(stage2_.variant());
}

void HeffPlan::autotune(int reps, int* chosen, std::vector<float>* t1, std::vector<float>* t2) {
  ctx_.make_current();
  // stage 2 reads the stage-1 output; run once so its inputs are defined
  stage1_.launch(ctx_.stream());
  int a = stage1_.autotune(ctx_.stream(), reps, t1);
  int b = stage2_.autotune(ctx_.stream(), reps, t2);
  SU2_HIP(hipStreamSynchronize(ctx_.stream()));
  if (chosen) {
    chosen[0] = a;
    chosen[1] = b;
  }
}

void HeffPlan::apply(const Tensor& psi, Tensor& out) {
  SU2_CHECK(psi.size() == psi_size_ && out.size() == psi_size_, kErrArgument,
            "H_eff apply: vector layout differs from the plan");
  apply_raw(psi.data(), out.data());
}

}  // namespace su2
