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
// K splits of the two launches, per problem (GroupedGemm::finalize): stage 1
// has short K (one sector dimension) and many tiles and is rarely split;
// stage 2 sums long K ranges, and about 96 k-tiles per unit balance its XCDs
// (measured on the 16x8 D=2000 centre bond, where the launch then has about
// 4000 units: 4.35 ms at 2000 units, 3.99 ms at 4000, flat up to 12000)
constexpr int kMinUnits1 = 128, kKtPerUnit1 = 128;
constexpr int kMinUnits2 = 512, kKtPerUnit2 = 96;

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
  // H_R of each sector copied: the plan must not read the caller's tensor
  // after it is built (the caller may free it)
  std::map<int, long long> hr_off;
  if (hr)
    for (auto& b : psi.blocks()) {
      const int j = b.twoj[1];
      if (hr->find2(j, j) >= 0 && !hr_off.count(j)) {
        hr_off[j] = work;
        work += (long long)b.dim[1] * b.dim[1];
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
  for (auto& kv : hr_off) {
    const Block& hb = hr->blocks()[hr->find2(kv.first, kv.first)];
    dev_copy(c, d_work_ + kv.second, hr->data() + hb.offset, hb.size);
  }

  // The two launches as problems over three buffers (bundle workspace, input,
  // output) with element offsets, instantiated for FP64 now and for FP32
  // when a mixed-precision apply first asks for it.
  // stage 1: X(j->j') = f * Lb^T psi(j)   [T*D_a(j') x D_b(j)]
  for (auto& tr : trs) {
    const Block* bj = pb[tr.ja];
    GemmSpec g{kBufWork, tr.x_off, (long long)tr.nt * tr.Dap, tr.nt * tr.Dap, tr.Db, {}};
    g.slices.push_back(
        SliceSpec{kBufWork, tr.lb_off, tr.Da, 1, kBufIn, bj->offset, tr.Da, tr.Da, tr.fac});
    spec1_.push_back(g);
  }
  // stage 2: psi'(j') = sum_j X(j->j') Rb(j->j') + H_L^T psi + psi H_R
  for (auto& b : psi.blocks()) {
    const int jap = b.twoj[0];
    const int Dap = b.dim[0], Dbp = b.dim[1];
    GemmSpec g{kBufOut, b.offset, Dap, Dap, Dbp, {}};
    for (auto& tr : trs) {
      if (tr.jap != jap) continue;
      g.slices.push_back(SliceSpec{kBufWork, tr.x_off, Dap, 0, kBufWork, tr.rb_off,
                                   (long long)tr.nt * tr.Db, tr.nt * tr.Db, 1.0});
    }
    if (hlt_off.count(jap))
      g.slices.push_back(
          SliceSpec{kBufWork, hlt_off.at(jap), Dap, 0, kBufIn, b.offset, Dap, Dap, 1.0});
    if (hr_off.count(b.twoj[1]))
      g.slices.push_back(
          SliceSpec{kBufIn, b.offset, Dap, 0, kBufWork, hr_off.at(b.twoj[1]), Dbp, Dbp, 1.0});
    if (g.slices.empty()) zero_blocks_.push_back({b.offset, b.size});
    else spec2_.push_back(g);
  }
  work_size_ = work;
  void* bases[3] = {d_work_, d_in_, d_out_};
  instantiate(stage1_, spec1_, bases, sizeof(double));
  instantiate(stage2_, spec2_, bases, sizeof(double));
  stage1_.finalize(c, kMinUnits1, kKtPerUnit1);
  stage2_.finalize(c, kMinUnits2, kKtPerUnit2);
  // no wait: the host-side tables above were copied when their uploads were
  // queued, and every later use of the plan is ordered on the stream
}

void HeffPlan::instantiate(GroupedGemm& g, const std::vector<GemmSpec>& spec, void* const* bases,
                           size_t es) {
  auto ptr = [&](int buf, long long off) { return (void*)((char*)bases[buf] + off * es); };
  for (auto& q : spec) {
    const int p = g.add_problem(ptr(q.cbuf, q.coff), q.ldc, q.M, q.N);
    for (auto& sl : q.slices)
      g.add_slice(p, ptr(sl.abuf, sl.aoff), sl.lda, sl.ta, ptr(sl.bbuf, sl.boff), sl.ldb, 0, sl.K,
                  sl.alpha);
  }
}

void HeffPlan::ensure_fp32() {
  if (stage1_32_) return;
  ctx_.make_current();
  // FP32 copies of the bundles and block Hamiltonians (the stage-1 output
  // region is overwritten by every apply), input and output vectors
  d_work32_ = (float*)pool_->alloc(sizeof(float) * std::max<long long>(work_size_, 1));
  d_in32_ = (float*)pool_->alloc(sizeof(float) * std::max<long long>(psi_size_, 1));
  d_out32_ = (float*)pool_->alloc(sizeof(float) * std::max<long long>(psi_size_, 1));
  dev_convert_f64_f32(ctx_, d_work32_, d_work_, work_size_);
  std::unique_ptr<GroupedGemm> s1(new GroupedGemm(Prec::F32)), s2(new GroupedGemm(Prec::F32));
  void* bases[3] = {d_work32_, d_in32_, d_out32_};
  instantiate(*s1, spec1_, bases, sizeof(float));
  instantiate(*s2, spec2_, bases, sizeof(float));
  s1->finalize(ctx_, kMinUnits1, kKtPerUnit1, var32_[0]);
  s2->finalize(ctx_, kMinUnits2, kKtPerUnit2, var32_[1]);
  stage1_32_ = std::move(s1);
  stage2_32_ = std::move(s2);
  SU2_HIP(hipStreamSynchronize(ctx_.stream()));
}

HeffPlan::~HeffPlan() {
  pool_->free(d_in_);
  pool_->free(d_out_);
  pool_->free(d_work_);
  pool_->free(d_work32_);
  pool_->free(d_in32_);
  pool_->free(d_out32_);
}

void HeffPlan::apply_raw(const double* psi, double* out) {
  ctx_.make_current();
  if (prec_ == Prec::F32) {
    // FP32 MFMA with FP32 accumulation on FP32 copies; FP64 in and out
    ensure_fp32();
    dev_convert_f64_f32(ctx_, d_in32_, psi, psi_size_);
    for (auto& z : zero_blocks_) dev_fill_f32(ctx_, d_out32_ + z.first, z.second, 0.0f);
    stage1_32_->launch(ctx_.stream());
    stage2_32_->launch(ctx_.stream());
    dev_convert_f32_f64(ctx_, out, d_out32_, psi_size_);
    return;
  }
  dev_copy(ctx_, d_in_, psi, psi_size_);
  for (auto& z : zero_blocks_) dev_fill(ctx_, d_out_ + z.first, z.second, 0.0);
  stage1_.launch(ctx_.stream());
  stage2_.launch(ctx_.stream());
  dev_copy(ctx_, out, d_out_, psi_size_);
}

void HeffPlan::set_precision(Prec p) {
  prec_ = p;
  if (p == Prec::F32) ensure_fp32();
}

GroupedGemm& HeffPlan::stage(int i) {
  if (prec_ == Prec::F32) {
    ensure_fp32();
    return i == 0 ? *stage1_32_ : *stage2_32_;
  }
  return i == 0 ? stage1_ : stage2_;
}

void HeffPlan::set_variants(int v1, int v2) {
  ctx_.make_current();
  stage(0).set_variant(v1);
  stage(1).set_variant(v2);
  if (prec_ == Prec::F32) {
    var32_[0] = v1;
    var32_[1] = v2;
  }
}

void HeffPlan::set_uniform(bool on) {
  ctx_.make_current();
  for (int i = 0; i < 2; ++i) {
    stage(i).allow_uniform(on);
    stage(i).set_variant(stage(i).variant());
  }
}

void HeffPlan::autotune(int reps, int* chosen, std::vector<float>* t1, std::vector<float>* t2) {
  ctx_.make_current();
  // stage 2 reads the stage-1 output; run once so its inputs are defined
  stage(0).launch(ctx_.stream());
  int a = stage(0).autotune(ctx_.stream(), reps, t1);
  int b = stage(1).autotune(ctx_.stream(), reps, t2);
  SU2_HIP(hipStreamSynchronize(ctx_.stream()));
  if (prec_ == Prec::F32) {
    var32_[0] = a;
    var32_[1] = b;
  }
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
