// Excited-state targeting on the device (exci > 0 of dmrg_su2).
//
// Lower states are read from their reference-format checkpoint files
// (read_mps into orth[j], dmrg_su2_omp.cpp:944-1068; energy by read_enr,
// 1100-1152).  For each lower state j the driver keeps the overlap
// environments ovlp[j][i] between the lower state's left (right) block
// states and this run's (overlap_initial / overlap_transformation,
// dmrg_su2_omp.cpp:327-328, 386, 470-471, 528; tensor_su2_dmrg_src.cpp:
// 531-688), updated wherever the operator environments are:
//   OL[i][jb] = sum_ja  Pj[jb][rows of ja]^T  OL[i-1][ja]  U[i][jb][rows of ja]
// and the mirror image on the right.  At a step on adjacent sites (il, il+1)
// the lower state in the two-site basis (overlapvec, dmrg_su2_omp.cpp:
// 219-228, 247-254) is
//   phi[jf] = (OL[il-1]^T A_il)[jf] (OR[il+2]^T B_il+1)[jf]^T,
// and every H_eff apply adds -E_j <phi_j|v> phi_j (dmrg_su2.cpp:611-623).
// All contractions are grouped FP64 MFMA GEMMs.
#include <cmath>
#include <map>
#include <memory>
#include <string>

#include "blas1.hpp"
#include "cg.hpp"
#include "common.hpp"
#include "context.hpp"
#include "dmrg.hpp"
#include "mpsio.hpp"
#include "ops.hpp"

namespace su2 {

namespace {

constexpr int kPhys = 1;

Leg with_dir(Leg l, int d) {
  l.dir = d;
  return l;
}

// device scratch from the context's cache, returned on scope exit
struct DevScratch {
  Context& c;
  double* p;
  DevScratch(Context& ctx, long long n) : c(ctx), p(ctx.dev_alloc_n<double>(n > 0 ? n : 0)) {}
  ~DevScratch() { c.dev_free(p); }
  DevScratch(const DevScratch&) = delete;
  DevScratch& operator=(const DevScratch&) = delete;
};

// Host blocks T[(ja, jb)] of one site of a lower state: d(ja) x d(jb),
// column-major, ja a sector of beta[k-1] and jb of beta[k], Clebsch-Gordan
// ratio folded in (as read_mps leaves them against make_cgc).
using SiteBlocks = std::map<std::pair<int, int>, std::vector<double>>;

SiteBlocks decode_site(const mpsio::File3& f, bool left_file, const std::string& what) {
  SiteBlocks T;
  const long long nten = mpsio::block_count(f.leg);
  long long o = 0, oc = 0;
  int s[3], tj[3];
  for (long long i = 0; i < nten; ++i) {
    if (!f.flag[i]) continue;
    mpsio::labels_of(f.leg, i, s, tj);
    const int d0 = f.leg[0].dim[s[0]], d1 = f.leg[1].dim[s[1]], d2 = f.leg[2].dim[s[2]];
    const std::vector<double> g = cg_tensor(tj[0], tj[1], tj[2]);
    const double r = mpsio::cg_ratio(f.cg.data() + oc, g, what);
    const double* src = f.red.data() + o;
    if (left_file) {
      // [beta[k-1], site, beta[k]]: block (ja, 1, jb) is d(ja) x 1 x d(jb)
      std::vector<double> b((size_t)d0 * d2);
      for (size_t e = 0; e < b.size(); ++e) b[e] = r * src[e];
      T[{tj[0], tj[2]}] = std::move(b);
    } else {
      // [site, beta[k], beta[k-1]]: block (1, jb, ja) is 1 x d(jb) x d(ja), T its transpose
      std::vector<double> b((size_t)d2 * d1);
      for (int c = 0; c < d2; ++c)
        for (int q = 0; q < d1; ++q) b[c + (size_t)q * d2] = r * src[q + (size_t)c * d1];
      T[{tj[2], tj[1]}] = std::move(b);
    }
    o += (long long)d0 * d1 * d2;
    oc += (long long)g.size();
  }
  return T;
}

// Projector view of one site: left = true gives fuse(beta[k-1], site) ->
// beta[k] (T(ja, jb) at the rows of ja inside jb); left = false gives
// fuse(site, beta[k]) -> beta[k-1] (T(ja, jb)^T at the rows of jb inside ja).
std::unique_ptr<Tensor> site_view(Context& c, const SiteBlocks& T, const Leg& prev,
                                  const Leg& next, bool left, const std::string& what) {
  const Leg& block = left ? prev : next;
  const Leg& bond = left ? next : prev;
  const Fused F = make_fused(block, left);
  std::unique_ptr<Tensor> P(new Tensor(&c, {with_dir(F.fused, 1), with_dir(bond, -1)}, 0, nullptr));
  std::vector<double> h(P->size(), 0.0);
  for (auto& kv : T) {
    const int ja = kv.first.first, jb = kv.first.second;
    const int da = prev.dim_of(ja), db = next.dim_of(jb);
    const int pb = left ? P->find2(jb, jb) : P->find2(ja, ja);
    auto it = left ? F.off.find({ja, jb}) : F.off.find({jb, ja});
    SU2_CHECK(pb >= 0 && it != F.off.end() && da > 0 && db > 0, kErrStructure,
              "lower state: %s block (%d,%d) does not fit its bonds", what.c_str(), ja, jb);
    const Block& B = P->blocks()[pb];
    const std::vector<double>& t = kv.second;
    if (left) {
      for (int col = 0; col < db; ++col)
        for (int r = 0; r < da; ++r)
          h[B.offset + it->second + r + (long long)col * B.dim[0]] = t[r + (size_t)col * da];
    } else {
      for (int col = 0; col < da; ++col)
        for (int r = 0; r < db; ++r)
          h[B.offset + it->second + r + (long long)col * B.dim[0]] = t[col + (size_t)r * da];
    }
  }
  if (!h.empty())
    SU2_HIP(hipMemcpyAsync(P->data(), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                           c.stream()));
  SU2_HIP(hipStreamSynchronize(c.stream()));
  return P;
}

std::unique_ptr<Tensor> unit_overlap(Context& c, const Leg& low, const Leg& cur) {
  std::unique_ptr<Tensor> t(new Tensor(&c, {with_dir(low, 1), with_dir(cur, -1)}, 0, nullptr));
  if (t->size()) dev_fill(c, t->data(), t->size(), 1.0);
  return t;
}

// One step of an overlap environment: prev pairs the lower state's and this
// run's bonds before the site, Pl / Pc are their projectors from the fused
// legs Fl / Fc (same orientation) to the new bonds.
//   out[jb] = Pl[jb]^T Y[jb],  Y[jb][rows of ja in Fl] = prev[ja] Pc[jb][rows of ja in Fc]
std::unique_ptr<Tensor> transfer(Context& c, const Tensor& prev, const Tensor& Pl, const Fused& Fl,
                                 const Tensor& Pc, const Fused& Fc) {
  std::unique_ptr<Tensor> out(
      new Tensor(&c, {with_dir(Pl.leg(1), 1), with_dir(Pc.leg(1), -1)}, 0, nullptr));
  struct Job {
    int ob, lb, cb, jb;
    long long y;
  };
  std::vector<Job> jobs;
  long long ytot = 0;
  for (int b = 0; b < (int)out->blocks().size(); ++b) {
    const int jb = out->blocks()[b].twoj[0];
    const int lb = Pl.find2(jb, jb), cb = Pc.find2(jb, jb);
    if (lb < 0 || cb < 0) continue;
    jobs.push_back({b, lb, cb, jb, ytot});
    ytot += (long long)Pl.blocks()[lb].dim[0] * Pc.blocks()[cb].dim[1];
  }
  if (out->size()) dev_fill(c, out->data(), out->size(), 0.0);
  DevScratch Y(c, ytot);
  if (ytot) dev_fill(c, Y.p, ytot, 0.0);
  GemmBatch g1, g2;
  for (const Job& j : jobs) {
    const Block& BL = Pl.blocks()[j.lb];
    const Block& BC = Pc.blocks()[j.cb];
    const int Dl = BL.dim[0], dcn = BC.dim[1];
    for (auto& p : Fc.parts(j.jb)) {
      const int ja = std::get<0>(p);
      const int pb = prev.find2(ja, ja);
      auto it = Fl.off.find({ja, j.jb});
      if (pb < 0 || it == Fl.off.end()) continue;
      const Block& BP = prev.blocks()[pb];
      SU2_CHECK(BP.dim[1] == std::get<2>(p) && BP.dim[0] == Fl.block.dim_of(ja), kErrStructure,
                "overlap: environment sector %d does not match the bonds", ja);
      g1.add(Y.p + j.y + it->second, Dl, BP.dim[0], dcn, prev.data() + BP.offset, BP.dim[0], 0,
             Pc.data() + BC.offset + std::get<1>(p), BC.dim[0], 0, BP.dim[1]);
    }
    g2.add(out->data() + out->blocks()[j.ob].offset, BL.dim[1], BL.dim[1], dcn,
           Pl.data() + BL.offset, Dl, 1, Y.p + j.y, Dl, 0, Dl);
  }
  if (!g1.empty()) g1.run(c);
  if (!g2.empty()) g2.run(c);
  SU2_HIP(hipStreamSynchronize(c.stream()));
  return out;
}

// Rows of one side of the two-site vector: X[rows of ja in Fc] = E[ja]^T P[jf][rows of ja in Fl]
// for every part ja of the current fused sector jf.
void lift(GemmBatch& g, double* X, int ldx, const Tensor& E, const Tensor& P, int pb,
          const Fused& Fl, const Fused& Fc, int jf) {
  const Block& BP = P.blocks()[pb];
  for (auto& p : Fc.parts(jf)) {
    const int ja = std::get<0>(p);
    const int eb = E.find2(ja, ja);
    auto it = Fl.off.find({ja, jf});
    if (eb < 0 || it == Fl.off.end()) continue;
    const Block& BE = E.blocks()[eb];
    SU2_CHECK(BE.dim[1] == std::get<2>(p), kErrStructure, "overlap vector: sector %d mismatch", ja);
    g.add(X + std::get<1>(p), ldx, BE.dim[1], BP.dim[1], E.data() + BE.offset, BE.dim[0], 1,
          P.data() + BP.offset + it->second, BP.dim[0], 0, BE.dim[0]);
  }
}

}  // namespace

void Dmrg::add_lower_state(const std::string& dir, int file_dcut, int ky, int exci) {
  Context& c = c_;
  c.make_current();
  const int ns = lat_.ns;
  SU2_CHECK(energies_.empty(), kErrArgument,
            "add_lower_state: add lower states before the first solve (or right after load)");
  SU2_CHECK(exci >= 0, kErrArgument, "add_lower_state: exci=%d", exci);
  const int fd = file_dcut > 0 ? file_dcut : D_;
  const std::string base = mpsio::uu_base(dir, ns, fd, S_, ky, exci);
  std::vector<mpsio::File3> files(ns);
  for (int i = 0; i < ns; ++i) files[i] = mpsio::read_file(base + std::to_string(i));
  std::unique_ptr<LowerState> st(new LowerState());
  st->energy = mpsio::read_energy(dir, mpsio::tag(ns, fd, S_, ky, exci));
  std::map<int, Leg>& beta = st->beta;
  beta[-1] = make_leg({0}, {1}, 1);
  for (int k = 0; k < ns / 2; ++k) {
    const mpsio::File3& f = files[k];
    SU2_CHECK(f.leg[1].nsec() == 1 && f.leg[1].twoj[0] == kPhys &&
                  f.leg[0].same_sectors(beta.at(k - 1)),
              kErrStructure, "lower state: site %d is not a left-site tensor of the chain", k);
    beta[k] = with_dir(f.leg[2], 1);
  }
  const Leg centre = beta.at(ns / 2 - 1);
  beta[ns - 1] = make_leg({S_}, {1}, 1);
  for (int k = ns - 1; k >= ns / 2; --k) {
    const mpsio::File3& f = files[k];
    SU2_CHECK(f.leg[0].nsec() == 1 && f.leg[0].twoj[0] == kPhys &&
                  f.leg[1].same_sectors(beta.at(k)),
              kErrStructure, "lower state: site %d is not a right-site tensor of the chain", k);
    beta[k - 1] = with_dir(f.leg[2], 1);
  }
  SU2_CHECK(beta.at(ns / 2 - 1).same_sectors(centre), kErrStructure,
            "lower state: the two halves meet on different centre bonds");
  st->left.resize(ns);
  st->right.resize(ns);
  for (int k = 0; k < ns; ++k) {
    const std::string what = "site " + std::to_string(k);
    const SiteBlocks T = decode_site(files[k], k < ns / 2, what);
    st->left[k] = site_view(c, T, beta.at(k - 1), beta.at(k), true, what);
    st->right[k] = site_view(c, T, beta.at(k - 1), beta.at(k), false, what);
  }
  st->OL[-1] = unit_overlap(c, beta.at(-1), lb_.at(-1));
  st->OR[ns] = unit_overlap(c, beta.at(ns - 1), rb_.at(ns));
  lower_.push_back(std::move(st));
  rebuild_overlaps();
}

void Dmrg::overlap_left(int i) {
  if (lower_.empty()) return;
  const Fused Fc = fused_left(i);
  const Tensor& P = *U_.at(i);
  for (auto& st : lower_)
    st->OL[i] = transfer(c_, *st->OL.at(i - 1), *st->left[i], make_fused(st->beta.at(i - 1), true),
                         P, Fc);
}

void Dmrg::overlap_right(int i) {
  if (lower_.empty()) return;
  const Fused Fc = fused_right(i);
  const Tensor& P = *V_.at(i);
  for (auto& st : lower_)
    st->OR[i] = transfer(c_, *st->OR.at(i + 1), *st->right[i], make_fused(st->beta.at(i), false),
                         P, Fc);
}

// overlaps of every site whose environment exists (after load, or a lower
// state added to a loaded run)
void Dmrg::rebuild_overlaps() {
  if (lower_.empty()) return;
  for (int i = 0; L_.count(i) && U_.count(i); ++i) overlap_left(i);
  for (int i = lat_.ns - 1; R_.count(i) && V_.count(i); --i) overlap_right(i);
}

std::vector<std::unique_ptr<Tensor>> Dmrg::lower_vectors(int il, int ir, const Fused& FL,
                                                         const Fused& FR, const Tensor& psi) {
  Context& c = c_;
  std::vector<std::unique_ptr<Tensor>> out;
  for (auto& st : lower_) {
    const Tensor& L = *st->OL.at(il - 1);
    const Tensor& R = *st->OR.at(ir + 1);
    const Tensor& A = *st->left[il];
    const Tensor& B = *st->right[ir];
    const Fused Fa = make_fused(st->beta.at(il - 1), true);
    const Fused Fb = make_fused(st->beta.at(ir), false);
    std::unique_ptr<Tensor> phi(new Tensor(&c, psi.legs(), 0, nullptr));
    if (phi->size()) dev_fill(c, phi->data(), phi->size(), 0.0);
    struct Job {
      int pb, ab, bb;
      long long at, bt;
    };
    std::vector<Job> jobs;
    long long tot = 0;
    for (int b = 0; b < (int)phi->blocks().size(); ++b) {
      const Block& PB = phi->blocks()[b];
      const int jf = PB.twoj[0];
      const int ab = A.find2(jf, jf), bb = B.find2(jf, jf);
      if (ab < 0 || bb < 0) continue;
      const int dlow = A.blocks()[ab].dim[1];
      SU2_CHECK(B.blocks()[bb].dim[1] == dlow, kErrStructure, "lower state: centre bond mismatch");
      jobs.push_back({b, ab, bb, tot, tot + (long long)PB.dim[0] * dlow});
      tot += (long long)(PB.dim[0] + PB.dim[1]) * dlow;
    }
    DevScratch X(c, tot);
    if (tot) dev_fill(c, X.p, tot, 0.0);
    GemmBatch g1, g2;
    for (const Job& j : jobs) {
      const Block& PB = phi->blocks()[j.pb];
      const int jf = PB.twoj[0];
      const int dlow = A.blocks()[j.ab].dim[1];
      lift(g1, X.p + j.at, PB.dim[0], L, A, j.ab, Fa, FL, jf);
      lift(g1, X.p + j.bt, PB.dim[1], R, B, j.bb, Fb, FR, jf);
      g2.add(phi->data() + PB.offset, PB.dim[0], PB.dim[0], PB.dim[1], X.p + j.at, PB.dim[0], 0,
             X.p + j.bt, PB.dim[1], 1, dlow);
    }
    if (!g1.empty()) g1.run(c);
    if (!g2.empty()) g2.run(c);
    SU2_HIP(hipStreamSynchronize(c.stream()));
    out.push_back(std::move(phi));
  }
  return out;
}

}  // namespace su2
