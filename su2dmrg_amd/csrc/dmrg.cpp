#include "dmrg.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>

#include "blas1.hpp"
#include "common.hpp"
#include "context.hpp"
#include "batch.hpp"
#include "heff.hpp"
#include "lanczos.hpp"
#include "ops.hpp"
#include "recoupling.hpp"
#include "svd.hpp"

namespace su2 {

namespace {

constexpr int kPhys = 1;                       // spin-1/2 sites (dmrg_su2.cpp:37)
constexpr int kOp = 2;                         // spin-1 operator leg
constexpr int kNfree = 2;                      // dmrg_su2.cpp:54
constexpr int kMlanc = 5;                      // dmrg_su2.cpp:530
const double kSigmaLeft = std::sqrt(3.0) / 2;  // sigma[0] (dmrg_su2.cpp:504)
const double kSigmaRight = -std::sqrt(3.0) / 2;  // sigma[1] (dmrg_su2.cpp:505)

Leg site_leg() { return make_leg({kPhys}, {1}, 1); }
Leg op_leg() { return make_leg({kOp}, {1}, 1); }

Leg with_dir(Leg l, int d) {
  l.dir = d;
  return l;
}

// with zb, the new operator is zeroed by zb.run() instead of a memset
std::unique_ptr<Tensor> scalar_op(Context& c, const Leg& l, ZeroBatch* zb = nullptr) {
  std::unique_ptr<Tensor> t(
      new Tensor(&c, {with_dir(l, 1), with_dir(l, -1)}, 0, nullptr, zb == nullptr));
  if (zb) zb->add(*t);
  return t;
}
std::unique_ptr<Tensor> vector_op(Context& c, const Leg& l, ZeroBatch* zb = nullptr) {
  std::unique_ptr<Tensor> t(
      new Tensor(&c, {with_dir(l, 1), op_leg(), with_dir(l, -1)}, 0, nullptr, zb == nullptr));
  if (zb) zb->add(*t);
  return t;
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

Fused make_fused(const Leg& block, bool left) {
  Fused F;
  F.block = block;
  F.fused = left ? fuse_legs(block, site_leg(), -1) : fuse_legs(site_leg(), block, -1);
  std::map<int, long long> pos;
  for (int jf : F.fused.twoj) pos[jf] = 0;
  for (int s = 0; s < block.nsec(); ++s) {
    const int jb = block.twoj[s];
    for (int jf = std::abs(jb - kPhys); jf <= jb + kPhys; jf += 2) {
      auto it = pos.find(jf);
      if (it == pos.end()) continue;
      F.off[{jb, jf}] = it->second;
      it->second += block.dim[s];
    }
  }
  return F;
}

std::vector<std::tuple<int, long long, int>> Fused::parts(int jf) const {
  std::vector<std::tuple<int, long long, int>> out;
  for (auto& kv : off)
    if (kv.first.second == jf) out.emplace_back(kv.first.first, kv.second, block.dim_of(kv.first.first));
  std::sort(out.begin(), out.end(),
            [](const auto& a, const auto& b) { return std::get<1>(a) < std::get<1>(b); });
  return out;
}

Dmrg::Dmrg(Context& c, int lx, int ly, double jcoup, double qcoup, int bdry, int totspin,
           int max_dcut)
    : c_(c), lat_(lx, ly, jcoup, qcoup, bdry), S_(totspin), D_(max_dcut) {
  SU2_CHECK(totspin >= 0 && totspin <= 24, kErrArgument, "totspin 2S=%d", totspin);
  SU2_CHECK(max_dcut >= 1, kErrArgument, "max_dcut=%d", max_dcut);
  lb_[-1] = make_leg({0}, {1}, 1);
  rb_[lat_.ns] = make_leg({totspin}, {1}, 1);
  L_[-1].reset(new Env());
  L_[-1]->bond = lb_[-1];
  R_[lat_.ns].reset(new Env());
  R_[lat_.ns]->bond = rb_[lat_.ns];
}

Fused Dmrg::fused_left(int i) const { return make_fused(lb_.at(i - 1), true); }
Fused Dmrg::fused_right(int i) const { return make_fused(rb_.at(i + 1), false); }

// The four environment kernels on the enlarged basis E (the reference's
// uu1 = fusion isometry, dmrg_su2_omp.cpp:604-825), as strided scale-adds
// with the recoupling factors; each adds into its output in call order.

// H (x) 1: overlap_transformation, factor 1 (tensor_su2_dmrg_src.cpp:531-688)
void env_overlap(ScatterBatch& sb, const Fused& E, const Tensor& H, Tensor& Hh) {
  for (int jf : E.fused.twoj) {
    const int bh = Hh.find2(jf, jf);
    const long long Df = E.dim(jf);
    for (auto& p : E.parts(jf)) {
      const int jb = std::get<0>(p);
      const long long o = std::get<1>(p);
      const int d = std::get<2>(p);
      const int bs = H.find2(jb, jb);
      if (bs < 0 || bh < 0) continue;
      StridedAxpy a{};
      a.dst = Hh.data() + Hh.blocks()[bh].offset + o + o * Df;
      a.src = H.data() + H.blocks()[bs].offset;
      a.rs = 1;
      a.cs = Df;
      a.ld = d;
      a.rows = d;
      a.cols = d;
      a.alpha = 1.0;
      sb.add(a.dst, a);
    }
  }
}

// coef S_k . s: operator_pairup (tensor_su2_dmrg_src.cpp:691-890), with
// sigma[1] on the left and sigma[0] on the right (dmrg_su2_omp.cpp:645,757)
void env_pairup(ScatterBatch& sb, const Fused& E, const Tensor& Sk, double coef, bool left,
                Tensor& Hh) {
  const double sig_pair = left ? kSigmaRight : kSigmaLeft;
  for (int jf : E.fused.twoj) {
    const int bh = Hh.find2(jf, jf);
    if (bh < 0) continue;
    const long long Df = E.dim(jf);
    auto parts = E.parts(jf);
    for (auto& p : parts)
      for (auto& q : parts) {
        const int jb = std::get<0>(p), jbp = std::get<0>(q);
        const double f = left ? rc_pairup_left(jf, jb, jbp, kOp, kPhys, kPhys)
                              : rc_pairup_right(jf, jb, jbp, kOp, kPhys, kPhys);
        if (std::fabs(f) < 1e-8) continue;
        const int bs = Sk.find3(jb, kOp, jbp);
        if (bs < 0) continue;
        StridedAxpy a{};
        a.dst = Hh.data() + Hh.blocks()[bh].offset + std::get<1>(p) + std::get<1>(q) * Df;
        a.src = Sk.data() + Sk.blocks()[bs].offset;
        a.rs = 1;
        a.cs = Df;
        a.ld = std::get<2>(p);
        a.rows = std::get<2>(p);
        a.cols = std::get<2>(q);
        a.alpha = coef * f * sig_pair;
        sb.add(a.dst, a);
      }
  }
}

// S_k (x) 1: operator_transformation (tensor_su2_dmrg_src.cpp:221-402)
void env_op_transform(ScatterBatch& sb, const Fused& E, const Tensor& Sk, bool left, Tensor& T) {
  for (auto& blk : T.blocks()) {
    const int jf = blk.twoj[0], jfp = blk.twoj[2];
    const long long Df = E.dim(jf);
    for (auto& p : E.parts(jf))
      for (auto& q : E.parts(jfp)) {
        const int jb = std::get<0>(p), jbp = std::get<0>(q);
        const double f = left ? rc_op_transform_left(jb, jf, jbp, jfp, kOp, kPhys)
                              : rc_op_transform_right(jb, jf, jbp, jfp, kOp, kPhys);
        if (std::fabs(f) < 1e-8) continue;
        const int bs = Sk.find3(jb, kOp, jbp);
        if (bs < 0) continue;
        StridedAxpy a{};
        a.dst = T.data() + blk.offset + std::get<1>(p) + std::get<1>(q) * Df;
        a.src = Sk.data() + Sk.blocks()[bs].offset;
        a.rs = 1;
        a.cs = Df;
        a.ld = std::get<2>(p);
        a.rows = std::get<2>(p);
        a.cols = std::get<2>(q);
        a.alpha = f;
        sb.add(a.dst, a);
      }
  }
}

// 1 (x) s: operator_initial with sigma[0] left / sigma[1] right
// (tensor_su2_dmrg_src.cpp:27-218)
void env_op_initial(ScatterBatch& sb, const Fused& E, bool left, Tensor& T) {
  const double sig = left ? kSigmaLeft : kSigmaRight;
  for (auto& blk : T.blocks()) {
    const int jf = blk.twoj[0], jfp = blk.twoj[2];
    const long long Df = E.dim(jf);
    for (auto& p : E.parts(jf))
      for (auto& q : E.parts(jfp)) {
        if (std::get<0>(p) != std::get<0>(q)) continue;
        const int jb = std::get<0>(p);
        const double f = left ? rc_op_initial_left(jb, jf, jfp, kOp, kPhys, kPhys)
                              : rc_op_initial_right(jb, jf, jfp, kOp, kPhys, kPhys);
        if (std::fabs(f) < 1e-8) continue;
        StridedAxpy a{};
        a.dst = T.data() + blk.offset + std::get<1>(p) + std::get<1>(q) * Df;
        a.src = nullptr;     // identity on the block sector
        a.rs = Df + 1;
        a.cs = 0;
        a.ld = 0;
        a.rows = std::get<2>(p);
        a.cols = 1;
        a.alpha = f * sig;
        sb.add(a.dst, a);
      }
  }
}

std::unique_ptr<Env> Dmrg::enlarge(const Fused& E, const Env& env, const std::vector<int>& keep,
                                   const std::vector<std::pair<double, int>>& pair, int newsite,
                                   bool left) {
  Context& c = c_;
  std::unique_ptr<Env> out(new Env());
  out->bond = E.fused;
  ZeroBatch zb;
  out->H = scalar_op(c, E.fused, &zb);
  for (int k : keep) out->S[k] = vector_op(c, E.fused, &zb);
  out->S[newsite] = vector_op(c, E.fused, &zb);
  zb.run(c);
  ScatterBatch sb;
  Tensor& Hh = *out->H;
  if (env.H) env_overlap(sb, E, *env.H, Hh);
  for (auto& pr : pair) {
    auto it = env.S.find(pr.second);
    SU2_CHECK(it != env.S.end(), kErrStructure, "pair operator of site %d missing", pr.second);
    env_pairup(sb, E, *it->second, pr.first, left, Hh);
  }
  for (int k : keep) {
    auto it = env.S.find(k);
    SU2_CHECK(it != env.S.end(), kErrStructure, "operator of site %d missing", k);
    env_op_transform(sb, E, *it->second, left, *out->S[k]);
  }
  env_op_initial(sb, E, left, *out->S[newsite]);
  sb.run(c);
  grow_plaquettes(E, env, newsite, left, *out);
  return out;
}

std::unique_ptr<Env> Dmrg::enlarged_left(int i, const Fused& E) {
  const Env& env = *L_.at(i - 1);
  std::vector<int> keep;
  std::vector<std::pair<double, int>> pair;
  for (int k = 0; k < i; ++k) {
    if (lat_.keep_left(i, k) && env.S.count(k)) keep.push_back(k);
    if (lat_.h(i, k)) pair.push_back({lat_.J(i, k), k});
  }
  return enlarge(E, env, keep, pair, i, true);
}

std::unique_ptr<Env> Dmrg::enlarged_right(int i, const Fused& E) {
  const Env& env = *R_.at(i + 1);
  std::vector<int> keep;
  std::vector<std::pair<double, int>> pair;
  for (int k = i + 1; k < lat_.ns; ++k) {
    if (lat_.keep_right(i, k) && env.S.count(k)) keep.push_back(k);
    if (lat_.h(i, k)) pair.push_back({lat_.J(i, k), k});
  }
  return enlarge(E, env, keep, pair, i, false);
}

// O' = P^T O P for every (O, destination) pair (the truncated-basis update of
// dmrg_su2_omp.cpp:319-601): Y = O[a->b] P[b], then O'[a->b] = P[a]^T Y, two
// grouped launches for all operators.
void project_ops(Context& c, const Tensor& P,
                 const std::vector<std::pair<const Tensor*, Tensor*>>& ops) {
  struct Job {
    const double* O;
    const double* Pa;
    const double* Pb;
    double* dst;
    int Da, Db, da, db;
    long long y;
  };
  std::vector<Job> jobs;
  long long ytot = 0;
  for (auto& op : ops) {
    const Tensor& O = *op.first;
    Tensor& T = *op.second;
    const int last = T.nleg() - 1;
    for (auto& tb : T.blocks()) {
      const int a = tb.twoj[0], b = tb.twoj[last];
      const int ob = last == 1 ? O.find2(a, b) : O.find3(a, O.leg(1).twoj[0], b);
      const int pa = P.find2(a, a), pbk = P.find2(b, b);
      SU2_CHECK(ob >= 0 && pa >= 0 && pbk >= 0, kErrStructure, "project: missing block");
      Job j;
      j.O = O.data() + O.blocks()[ob].offset;
      j.Pa = P.data() + P.blocks()[pa].offset;
      j.Pb = P.data() + P.blocks()[pbk].offset;
      j.dst = T.data() + tb.offset;
      j.Da = P.blocks()[pa].dim[0];
      j.Db = P.blocks()[pbk].dim[0];
      j.da = P.blocks()[pa].dim[1];
      j.db = P.blocks()[pbk].dim[1];
      j.y = ytot;
      ytot += (long long)j.Da * j.db;
      jobs.push_back(j);
    }
  }
  DevBuf Y(c, ytot);
  {
    GemmBatch g;
    for (auto& j : jobs) g.add(Y.p + j.y, j.Da, j.Da, j.db, j.O, j.Da, 0, j.Pb, j.Db, 0, j.Db);
    g.run(c);
  }
  {
    GemmBatch g;
    for (auto& j : jobs) g.add(j.dst, j.da, j.da, j.db, j.Pa, j.Da, 1, Y.p + j.y, j.Da, 0, j.Da);
    g.run(c);
  }
}

std::unique_ptr<Env> Dmrg::project(const Tensor& P, const Env& hat) {
  Context& c = c_;
  const double t0 = now();
  std::unique_ptr<Env> out(new Env());
  out->bond = P.leg(1);
  std::vector<std::pair<const Tensor*, Tensor*>> ops;
  ZeroBatch zb;
  if (hat.H) {
    out->H = scalar_op(c, out->bond, &zb);
    ops.push_back({hat.H.get(), out->H.get()});
  }
  for (auto& kv : hat.S) {
    out->S[kv.first] = vector_op(c, out->bond, &zb);
    ops.push_back({kv.second.get(), out->S[kv.first].get()});
  }
  for (auto& kv : hat.P)
    for (auto& r : kv.second) {
      auto& dst = out->P[kv.first][r.first];
      dst.reset(new Tensor(&c, {with_dir(out->bond, 1), make_leg({r.first}, {1}, 1),
                                with_dir(out->bond, -1)},
                           0, nullptr, false));
      zb.add(*dst);
      ops.push_back({r.second.get(), dst.get()});
    }
  zb.run(c);
  project_ops(c, P, ops);
  stats_.t_project += now() - t0;
  return out;
}

// One environment kernel on one operator, alone (the C ABI's
// su2_env_transform): the kernel on the enlarged basis of `block`, then the
// projection with P when given.
std::unique_ptr<Tensor> env_transform(Context& c, int kind, bool left, const Leg& block,
                                      const Tensor* P, const Tensor* op, double coef) {
  c.make_current();
  Fused E = make_fused(block, left);
  if (P) {
    SU2_CHECK(P->nleg() == 2 && P->leg(0).same_sectors(E.fused), kErrStructure,
              "env transform: projector leg 0 is not the enlarged basis of the block");
  }
  auto check_op = [&](int nleg) {
    SU2_CHECK(op != nullptr, kErrArgument, "env transform: operator missing");
    SU2_CHECK(op->nleg() == nleg, kErrStructure, "env transform: operator with %d legs",
              op->nleg());
    SU2_CHECK(op->leg(0).same_sectors(block) && op->leg(nleg - 1).same_sectors(block),
              kErrStructure, "env transform: operator legs are not the block leg");
    if (nleg == 3)
      SU2_CHECK(op->leg(1).nsec() == 1 && op->leg(1).twoj[0] == kOp && op->leg(1).dim[0] == 1,
                kErrStructure, "env transform: operator leg 1 must be the spin-1 multiplet");
  };
  const bool scalar = kind == kEnvOverlap || kind == kEnvPairup;
  std::unique_ptr<Tensor> X = scalar ? scalar_op(c, E.fused) : vector_op(c, E.fused);
  ScatterBatch sb;
  switch (kind) {
    case kEnvOpTransform:
      check_op(3);
      env_op_transform(sb, E, *op, left, *X);
      break;
    case kEnvOverlap:
      check_op(2);
      env_overlap(sb, E, *op, *X);
      break;
    case kEnvPairup:
      check_op(3);
      env_pairup(sb, E, *op, coef, left, *X);
      break;
    case kEnvInitial:
      env_op_initial(sb, E, left, *X);
      break;
    default:
      fail(kErrArgument, "env transform: kind %d", kind);
  }
  sb.run(c);
  if (!P) {
    SU2_HIP(hipStreamSynchronize(c.stream()));
    return X;
  }
  const Leg bond = P->leg(1);
  std::unique_ptr<Tensor> out = scalar ? scalar_op(c, bond) : vector_op(c, bond);
  project_ops(c, *P, {{X.get(), out.get()}});
  SU2_HIP(hipStreamSynchronize(c.stream()));
  return out;
}

Dmrg::Solve::~Solve() = default;

// H_eff terms of the cut (il, ir) (dmrg_su2.cpp:571-610): couplings summed
// once per solve, the plaquette terms appended
std::vector<HeffTerm> Dmrg::heff_terms(int il, int ir, const Env& El, const Env& Er) const {
  std::vector<HeffTerm> terms;
  if (il < lat_.ns / 2) {
    for (int j = 0; j <= il; ++j) {
      auto lj = El.S.find(j);
      if (lj == El.S.end()) continue;
      HeffTerm t;
      for (int k = ir; k < lat_.ns; ++k)
        if (lat_.h(j, k) && Er.S.count(k)) t.right.push_back({lat_.J(j, k), Er.S.at(k).get()});
      if (t.right.empty()) continue;
      t.left.push_back({1.0, lj->second.get()});
      terms.push_back(t);
    }
  } else {
    for (int k = ir; k < lat_.ns; ++k) {
      auto rk = Er.S.find(k);
      if (rk == Er.S.end()) continue;
      HeffTerm t;
      for (int j = 0; j <= il; ++j)
        if (lat_.h(j, k) && El.S.count(j)) t.left.push_back({lat_.J(j, k), El.S.at(j).get()});
      if (t.left.empty()) continue;
      t.right.push_back({1.0, rk->second.get()});
      terms.push_back(t);
    }
  }
  plaquette_terms(il, ir, El, Er, terms);
  return terms;
}

// Everything of a two-site step before its Lanczos run: enlarged
// environments, the start vector in the two-site basis, the H_eff plan and
// the lower states (excited-state penalty).
std::unique_ptr<Dmrg::Solve> Dmrg::prepare(int il, int ir, std::unique_ptr<Tensor> psi0, double p1,
                                           double p2) {
  Context& c = c_;
  std::unique_ptr<Solve> sv(new Solve());
  Solve& S = *sv;
  S.il = il;
  S.ir = ir;
  S.p1 = p1;
  S.p2 = p2;
  S.t_step = now();
  double t0 = S.t_step;
  StepOut& o = S.o;
  o.FL = fused_left(il);
  o.FR = fused_right(ir);
  o.El = enlarged_left(il, o.FL);
  o.Er = enlarged_right(ir, o.FR);
  stats_.t_env += now() - t0;

  t0 = now();
  S.t_lanczos0 = t0;
  S.psi.reset(new Tensor(&c, {with_dir(o.FL.fused, 1), with_dir(o.FR.fused, 1)}, 0, nullptr));
  SU2_CHECK(S.psi->size() > 0, kErrStructure, "step %d,%d: no common sector for spin %d", il, ir,
            S_);
  if (psi0) {
    SU2_CHECK(psi0->size() == S.psi->size(), kErrStructure, "start vector layout mismatch");
    dev_copy(c, S.psi->data(), psi0->data(), S.psi->size());
  } else {
    dev_fill(c, S.psi->data(), S.psi->size(), 1.0);   // initialize_input_vector (tensor_su2.cpp:2354-2372)
  }
  std::vector<HeffTerm> terms = heff_terms(il, ir, *o.El, *o.Er);
  S.plan.reset(new HeffPlan(c, *S.psi, o.El->H.get(), o.Er->H.get(), terms));
  if (var1_ >= 0) S.plan->set_variants(var1_, var2_);
  // mixed precision: FP32 Krylov matvecs where the matvec is large enough
  // to gain from them; the Ritz-vector checks stay FP64
  S.fp32 = prec_ == Prec::F32 && S.plan->algorithmic_flops() >= min_fp32_flops_;
  if (S.fp32) {
    S.plan->set_precision(Prec::F32);
    if (var32_[0] >= 0) S.plan->set_variants(var32_[0], var32_[1]);
    S.plan->set_precision(Prec::F64);
  }
  S.mlanc = (int)std::min<long long>(kMlanc, S.psi->size() + 1);
  // lower states in this two-site basis (adjacent sites only, dmrg_su2.cpp:611)
  if (!lower_.empty() && il + 1 == ir) {
    S.phis = lower_vectors(il, ir, o.FL, o.FR, *S.psi);
    S.pen = DevBuf(c, 1);
  }
  return sv;
}

// - sum_j E_j <phi_j|x> phi_j (dmrg_su2.cpp:611-623), the overlaps kept on
// the device (no host wait)
void Dmrg::penalty(const Solve& S, const double* x, double* y) {
  for (size_t j = 0; j < S.phis.size(); ++j) {
    dev_dot_to(c_, S.phis[j]->data(), x, S.psi->size(), S.pen.p);
    dev_axpy_dev(c_, y, S.phis[j]->data(), S.psi->size(), S.pen.p, -lower_[j]->energy);
  }
}

// The solve's Lanczos run, as a batch of one (BatchDriver runs the same
// operations for many solves at once).
LanczosResult Dmrg::solve(Solve& S) {
  Context& c = c_;
  HeffBatch hb(c);
  hb.set_use_graph(use_graph_);
  std::vector<LanczosJob> jobs(1);
  jobs[0].psi = S.psi->data();
  jobs[0].n = S.psi->size();
  jobs[0].mlanc = S.mlanc;
  jobs[0].mixed = S.fp32;
  // matvec device time from an event pair per matvec, read once the solve
  // has finished (no wait per matvec); FP32 and FP64 matvecs apart
  std::vector<std::pair<hipEvent_t, int>> ev;
  auto release_events = [&]() {
    for (auto& e : ev) (void)hipEventDestroy(e.first);
    ev.clear();
  };
  auto apply = [&](const std::vector<LanczosReq>& reqs) {
    for (auto& r : reqs) {
      const Prec p = (S.fp32 && !r.check) ? Prec::F32 : Prec::F64;
      hipEvent_t e0, e1;
      SU2_HIP(hipEventCreate(&e0));
      ev.push_back({e0, (int)p});
      SU2_HIP(hipEventCreate(&e1));
      ev.push_back({e1, (int)p});
      SU2_HIP(hipEventRecord(e0, c.stream()));
      hb.apply({HeffBatch::Req{S.plan.get(), p, r.x, r.y}});
      SU2_HIP(hipEventRecord(e1, c.stream()));
      penalty(S, r.x, r.y);
    }
  };
  double ms[2] = {0, 0};
  try {
    lanczos_batch(c, jobs, apply);
    SU2_HIP(hipStreamSynchronize(c.stream()));
    for (size_t i = 0; i + 1 < ev.size(); i += 2) {
      float t = 0;
      SU2_HIP(hipEventElapsedTime(&t, ev[i].first, ev[i + 1].first));
      ms[ev[i].second == (int)Prec::F32 ? 1 : 0] += t;
    }
  } catch (...) {
    release_events();
    throw;
  }
  release_events();
  hb.clear();
  const LanczosResult& lr = jobs[0].res;
  HeffPlan& plan = *S.plan;
  const int n32 = S.fp32 ? lr.matvecs - lr.check_matvecs : 0;
  stats_.t_matvec += (ms[0] + ms[1]) * 1e-3;
  stats_.t_matvec32 += ms[1] * 1e-3;
  stats_.matvecs += lr.matvecs;
  stats_.matvecs32 += n32;
  stats_.heff_flops += plan.algorithmic_flops() * lr.matvecs;
  stats_.heff_flops32 += plan.algorithmic_flops() * n32;
  if (S.fp32)
    stats_.max_fp64_correction =
        std::max(stats_.max_fp64_correction, std::fabs(lr.ritz - lr.energy));
  stats_.t_lanczos += now() - S.t_lanczos0;
  release_plan(S);
  return jobs[0].res;
}

SvdResult Dmrg::svd(Solve& S) {
  const double t0 = now();
  SvdResult sv = truncated_svd(c_, *S.psi, S.p1, S.p2, D_);
  stats_.t_svd += now() - t0;
  return sv;
}

// The plan's bundles go back to the cache before the SVD allocates.
void Dmrg::release_plan(Solve& S) {
  S.plan.reset();
  S.phis.clear();
}

Dmrg::StepOut Dmrg::finish(Solve& S, const LanczosResult& lr, SvdResult&& sv) {
  StepOut o = std::move(S.o);
  o.tag = S.tag;
  o.energy = lr.energy;
  o.U = std::move(sv.U);
  o.V = std::move(sv.V);
  o.bond = sv.bond;
  energies_.push_back(lr.energy);
  history_.push_back({(int)sv.bond.total(), sv.discarded_weight, lr.matvecs});
  wout_ = sv.wout;
  last_bond_ = (int)sv.bond.total();
  stats_.t_jacobi += sv.t_jacobi;
  stats_.jacobi_sweeps += sv.jacobi_sweeps;
  stats_.svd_degraded += sv.jacobi_worst > 0 ? 1 : 0;
  if (verbose_) {
    std::fprintf(stderr, "step %lld (%d,%d) E=%.12f kept=%d discarded=%.3e matvecs=%d%s %.3fs\n",
                 (long long)energies_.size() - 1, S.il, S.ir, lr.energy, last_bond_,
                 sv.discarded_weight, lr.matvecs, S.fp32 ? " fp32" : "", now() - S.t_step);
    std::fflush(stderr);
  }
  stats_.steps++;
  return o;
}

// Lanczos start vector at (nl, nr) from the left factor A (E_nr x b) of the
// step at (nr, nr+1): psi0 = U[nl] . A re-coupled across site nr.
std::unique_ptr<Tensor> Dmrg::start_left(int nl, int nr, const Tensor& A, const Leg& b) {
  Context& c = c_;
  const double t0 = now();
  Fused E = fused_left(nr);               // parts: lb[nl] sectors
  Fused F = make_fused(b, false);         // F_nr = site x b
  Fused FL = fused_left(nl);
  const Tensor& P = *U_.at(nl);           // E_nl -> lb[nl]
  const Leg& lbl = lb_.at(nl);
  std::unique_ptr<Tensor> out(
      new Tensor(&c, {with_dir(FL.fused, 1), with_dir(F.fused, 1)}, 0, nullptr));
  ScatterBatch sb;
  GemmBatch g;
  std::vector<double*> tmp;
  for (int s = 0; s < lbl.nsec(); ++s) {
    const int jl = lbl.twoj[s];
    const int dl = lbl.dim[s];
    const int ob = out->find2(jl, jl), pb = P.find2(jl, jl);
    if (ob < 0 || pb < 0 || F.dim(jl) == 0) continue;
    const long long Df = F.dim(jl);
    double* blk = nullptr;
    blk = c.dev_alloc_n<double>((size_t)dl * Df);
    dev_fill(c, blk, (long long)dl * Df, 0.0);
    tmp.push_back(blk);
    for (auto& p : F.parts(jl)) {
      const int jb = std::get<0>(p);
      const int ab = A.find2(jb, jb);
      auto it = E.off.find({jl, jb});
      if (ab < 0 || it == E.off.end()) continue;
      StridedAxpy a{};
      a.dst = blk + std::get<1>(p) * dl;
      a.src = A.data() + A.blocks()[ab].offset + it->second;
      a.rs = 1;
      a.cs = dl;
      a.ld = A.blocks()[ab].dim[0];
      a.rows = dl;
      a.cols = std::get<2>(p);
      a.alpha = rc_psi_right_to_left(jl, kPhys, jb);
      sb.add(a.dst, a);
    }
    const int De = P.blocks()[pb].dim[0];
    g.add(out->data() + out->blocks()[ob].offset, De, De, (int)Df, P.data() + P.blocks()[pb].offset,
          De, 0, blk, dl, 0, dl);
  }
  sb.run(c);
  g.run(c);
  for (double* p : tmp) c.dev_free(p);
  stats_.t_start += now() - t0;
  return out;
}

// Start vector at (nl, nr) from the right factor B (F_nl x b) of the step at
// (nl-1, nl): psi0 = B re-coupled across site nl . V[nr]^T.
std::unique_ptr<Tensor> Dmrg::start_right(int nl, int nr, const Tensor& B, const Leg& b) {
  Context& c = c_;
  const double t0 = now();
  Fused F = fused_right(nl);              // parts: rb[nr] sectors
  Fused E = make_fused(b, true);          // E_nl = b x site
  Fused FR = fused_right(nr);
  const Tensor& P = *V_.at(nr);           // F_nr -> rb[nr]
  const Leg& rbr = rb_.at(nr);
  std::unique_ptr<Tensor> out(
      new Tensor(&c, {with_dir(E.fused, 1), with_dir(FR.fused, 1)}, 0, nullptr));
  ScatterBatch sb;
  GemmBatch g;
  std::vector<double*> tmp;
  for (int s = 0; s < rbr.nsec(); ++s) {
    const int jr = rbr.twoj[s];
    const int dr = rbr.dim[s];
    const int ob = out->find2(jr, jr), pb = P.find2(jr, jr);
    if (ob < 0 || pb < 0 || E.dim(jr) == 0) continue;
    const long long De = E.dim(jr);
    double* blk = nullptr;
    blk = c.dev_alloc_n<double>((size_t)De * dr);
    dev_fill(c, blk, De * dr, 0.0);
    tmp.push_back(blk);
    for (auto& p : E.parts(jr)) {
      const int jb = std::get<0>(p);
      const int bb = B.find2(jb, jb);
      auto it = F.off.find({jr, jb});
      if (bb < 0 || it == F.off.end()) continue;
      StridedAxpy a{};
      a.dst = blk + std::get<1>(p);
      a.src = B.data() + B.blocks()[bb].offset + it->second;
      a.rs = De;
      a.cs = 1;
      a.ld = B.blocks()[bb].dim[0];
      a.rows = dr;
      a.cols = std::get<2>(p);
      a.alpha = rc_psi_left_to_right(jb, kPhys, jr);
      sb.add(a.dst, a);
    }
    const int Dfr = P.blocks()[pb].dim[0];
    g.add(out->data() + out->blocks()[ob].offset, De, (int)De, Dfr, blk, De, 0,
          P.data() + P.blocks()[pb].offset, Dfr, 1, dr);
  }
  sb.run(c);
  g.run(c);
  for (double* p : tmp) c.dev_free(p);
  stats_.t_start += now() - t0;
  return out;
}

// Identity projector on a fused leg (the fusion isometry itself is the MPS
// tensor while no solve has truncated the site, dmrg_su2_omp.cpp:64-68).
static std::unique_ptr<Tensor> identity_projector(Context& c, const Leg& l) {
  std::unique_ptr<Tensor> P(new Tensor(&c, {with_dir(l, 1), with_dir(l, -1)}, 0, nullptr));
  ScatterBatch sb;
  for (auto& b : P->blocks()) {
    StridedAxpy a{};
    a.dst = P->data() + b.offset;
    a.src = nullptr;
    a.rs = b.dim[0] + 1;
    a.cs = 0;
    a.rows = b.dim[0];
    a.cols = 1;
    a.alpha = 1.0;
    sb.add(a.dst, a);
  }
  sb.run(c);
  return P;
}

void Dmrg::ensure_left(int i) {
  if (L_.count(i)) return;
  Fused E = fused_left(i);
  L_[i] = enlarged_left(i, E);     // untruncated: the projector is the identity
  lb_[i] = E.fused;
  U_[i] = identity_projector(c_, E.fused);
  overlap_left(i);
}

void Dmrg::ensure_right(int i) {
  if (R_.count(i)) return;
  Fused E = fused_right(i);
  R_[i] = enlarged_right(i, E);
  rb_[i] = E.fused;
  V_[i] = identity_projector(c_, E.fused);
  overlap_right(i);
}

// The two-site H_eff and wavefunction at the centre bond of the current state
// (after warmup() or sweep(), whose last solve is there): psi_j = A_j V_j^T
// per sector, A = U[ns/2 - 1] (carries the singular values), V = V[ns/2].
std::pair<std::unique_ptr<HeffPlan>, std::unique_ptr<Tensor>> Dmrg::centre() {
  Context& c = c_;
  c.make_current();
  const int il = lat_.ns / 2 - 1, ir = lat_.ns / 2;
  SU2_CHECK(U_.count(il) && V_.count(ir) && L_.count(il - 1) && R_.count(ir + 1) &&
                cur_.phase == kPhaseNone,
            kErrArgument, "centre: no state at the centre bond (run warmup or a sweep first)");
  Fused FL = fused_left(il), FR = fused_right(ir);
  std::unique_ptr<Env> El = enlarged_left(il, FL), Er = enlarged_right(ir, FR);
  std::unique_ptr<Tensor> psi(
      new Tensor(&c, {with_dir(FL.fused, 1), with_dir(FR.fused, 1)}, 0, nullptr));
  const Tensor& A = *U_.at(il);
  const Tensor& V = *V_.at(ir);
  GemmBatch g;
  for (auto& b : psi->blocks()) {
    const int j = b.twoj[0];
    const int ia = A.find2(j, j), iv = V.find2(j, j);
    if (ia < 0 || iv < 0) continue;   // sector cut away: zero block
    const Block& ba = A.blocks()[ia];
    const Block& bv = V.blocks()[iv];
    SU2_CHECK(ba.dim[0] == b.dim[0] && bv.dim[0] == b.dim[1] && ba.dim[1] == bv.dim[1],
              kErrStructure, "centre: factor dimensions");
    g.add(psi->data() + b.offset, b.dim[0], b.dim[0], b.dim[1], A.data() + ba.offset, ba.dim[0],
          0, V.data() + bv.offset, bv.dim[0], 1, ba.dim[1]);
  }
  g.run(c);
  std::vector<HeffTerm> terms = heff_terms(il, ir, *El, *Er);
  std::unique_ptr<HeffPlan> plan(new HeffPlan(c, *psi, El->H.get(), Er->H.get(), terms));
  SU2_HIP(hipStreamSynchronize(c.stream()));
  return {std::move(plan), std::move(psi)};
}

// The reference's step order as a cursor (dmrg_su2_omp.cpp:20-157), so that a
// driver can prepare one solve of each of several runs, run their Lanczos
// solves and SVDs together, and complete each.  warmup() and sweep() drive
// one run through the same cursor.

void Dmrg::begin_warmup() {
  c_.make_current();
  cur_ = Cursor{};
  cur_.phase = kPhaseWarmup;
  cur_.i = 0;
}

void Dmrg::begin_sweep() {
  c_.make_current();
  const int ns = lat_.ns;
  SU2_CHECK(U_.count(ns / 2 - 1), kErrArgument, "sweep before warm-up");
  cur_ = Cursor{};
  cur_.phase = kPhaseSweep;
  cur_.A = std::move(U_[ns / 2 - 1]);
  U_.erase(ns / 2 - 1);
  cur_.bond = lb_.at(ns / 2 - 1);
  cur_.energy = energies_.empty() ? 0 : energies_.back();
  cur_.pass = 1;
  cur_.i = ns / 2;
}

std::unique_ptr<Dmrg::Solve> Dmrg::next_solve() {
  const int ns = lat_.ns;
  const int nf = kNfree;
  if (cur_.phase == kPhaseWarmup) {
    // do_idmrg (dmrg_su2_omp.cpp:20-95)
    for (; cur_.i <= ns / 2 - 2; ++cur_.i) {
      const int i = cur_.i;
      ensure_left(i);
      ensure_right(ns - 1 - i);
      const bool final = (i == ns / 2 - 2);
      const bool solve = ((i + 2) * 2 * kPhys >= S_ && i >= 1) || final;
      if (!solve) continue;
      auto S = prepare(i + 1, ns - 2 - i, nullptr, final ? 1.0 : 0.0, 0.0);
      S->tag = final ? kTagWarmFinal : kTagWarm;
      return S;
    }
    cur_.phase = kPhaseNone;
    return nullptr;
  }
  if (cur_.phase != kPhaseSweep) return nullptr;
  // dmrg_su2::sweep (dmrg_su2_omp.cpp:98-157): pass 1 centre -> left edge,
  // pass 2 left edge -> right edge, pass 3 right edge -> centre
  if (cur_.pass == 1 && cur_.i < nf + 2) {
    cur_.pass = 2;
    cur_.i = nf;
  }
  if (cur_.pass == 2 && cur_.i >= ns - nf - 2) {
    cur_.pass = 3;
    cur_.i = ns - nf - 1;
  }
  if (cur_.pass == 3 && cur_.i <= ns / 2) {
    // end of the sweep: the centre's left factor carries the state
    if (cur_.il_last < 0) cur_.il_last = ns / 2 - 1;
    U_[cur_.il_last] = std::move(cur_.A);
    cur_.phase = kPhaseNone;
    return nullptr;
  }
  const int i = cur_.i;
  std::unique_ptr<Solve> S;
  if (cur_.pass == 1) {
    const int il = i - 2, ir = i - 1;
    const bool turn = (i == nf + 2);
    S = prepare(il, ir, start_left(il, ir, *cur_.A, cur_.bond), turn ? 0.0 : 1.0,
                turn ? 1.0 : 0.0);
    S->tag = turn ? kTagP1Turn : kTagP1;
  } else if (cur_.pass == 2) {
    const int il = i + 1, ir = i + 2;
    const bool turn = (i == ns - nf - 3);
    S = prepare(il, ir, start_right(il, ir, *cur_.B, cur_.bond), turn ? 1.0 : 0.0,
                turn ? 0.0 : 1.0);
    S->tag = turn ? kTagP2Turn : kTagP2;
  } else {
    const int il = i - 2, ir = i - 1;
    S = prepare(il, ir, start_left(il, ir, *cur_.A, cur_.bond), 1.0, 0.0);
    S->tag = kTagP3;
  }
  return S;
}

void Dmrg::complete(std::unique_ptr<Solve> S, const LanczosResult& lr, SvdResult&& sv) {
  StepOut o = finish(*S, lr, std::move(sv));
  const int il = S->il, ir = S->ir;
  S.reset();
  lb_[il] = o.bond;
  rb_[ir] = o.bond;
  auto right_env = [&]() {
    R_[ir] = project(*o.V, *o.Er);
    V_[ir] = std::move(o.V);
    overlap_right(ir);
  };
  auto left_env = [&]() {
    L_[il] = project(*o.U, *o.El);
    U_[il] = std::move(o.U);
    overlap_left(il);
  };
  switch (o.tag) {
    case kTagWarm:
      right_env();
      left_env();
      break;
    case kTagWarmFinal:
      right_env();
      U_[il] = std::move(o.U);
      break;
    case kTagP1:    // centre -> left: the right block grows, A carries on
    case kTagP2Turn:
    case kTagP3:
      right_env();
      cur_.A = std::move(o.U);
      if (o.tag == kTagP3) cur_.il_last = il;
      break;
    case kTagP1Turn:  // turn at the left edge: B carries on to the right
    case kTagP2:
      left_env();
      cur_.B = std::move(o.V);
      break;
    default:
      fail(kErrArgument, "dmrg: solve without a step tag");
  }
  cur_.bond = lb_.at(il);
  cur_.energy = o.energy;
  if (cur_.phase == kPhaseWarmup) {
    ++cur_.i;
  } else if (cur_.pass == 1 || cur_.pass == 3) {
    --cur_.i;
  } else {
    ++cur_.i;
  }
}

// do_idmrg (dmrg_su2_omp.cpp:20-95)
double Dmrg::warmup() {
  begin_warmup();
  double last = 0;
  while (std::unique_ptr<Solve> S = next_solve()) {
    LanczosResult lr = solve(*S);
    SvdResult sv = svd(*S);
    last = lr.energy;
    complete(std::move(S), lr, std::move(sv));
  }
  return last;
}

// dmrg_su2::sweep (dmrg_su2_omp.cpp:98-157)
double Dmrg::sweep() {
  begin_sweep();
  while (std::unique_ptr<Solve> S = next_solve()) {
    LanczosResult lr = solve(*S);
    SvdResult sv = svd(*S);
    complete(std::move(S), lr, std::move(sv));
  }
  return cur_.energy;
}

}  // namespace su2
