// Two-site SU(2) DMRG driver on the device (dmrg_su2, dmrg_su2_omp.cpp).
//
// State: for a left site i the projector U[i] from E_i = fuse(lb[i-1], site)
// to the kept bond lb[i] (the reference's uu[i] is this projector contracted
// with the fusion isometry, tensor_su2.cpp:1740); for a right site V[i] from
// F_i = fuse(site, rb[i+1]).  Environments: block Hamiltonian H and the spin
// operators S_k still needed (fll/frr).  A step at (il, ir):
//   * enlarged environments in E_il and F_ir: H (x) 1, S_k (x) 1, 1 (x) s and
//     the pair terms J S_k . s placed by batched device scatters with the
//     reference's recoupling factors (dmrg_su2_omp.cpp:604-825);
//   * H_eff plan + Lanczos on the device (dmrg_su2.cpp:527-664);
//   * truncated SVD (tensor_su2.cpp:1619-1758);
//   * the truncated environment U^T (enlarged) U by grouped GEMMs
//     (dmrg_su2_omp.cpp:319-601).
// Warm-up growth and the three-pass sweep follow dmrg_su2_omp.cpp:20-157
// with nfree = 2; the Lanczos start vector of a sweep step is the previous
// wavefunction carried across one site with the exact recoupling factor.
#pragma once

#include <map>
#include <memory>
#include <utility>
#include <string>
#include <tuple>
#include <vector>

#include "context.hpp"
#include "gemm.hpp"
#include "lanczos.hpp"
#include "lattice.hpp"
#include "structure.hpp"
#include "svd.hpp"
#include "tensor.hpp"

namespace su2 {

class Context;
class HeffPlan;
struct HeffTerm;

struct Env {
  Leg bond;
  std::unique_ptr<Tensor> H;                      // scalar block Hamiltonian (may be null)
  std::map<int, std::unique_ptr<Tensor>> S;       // site -> spin operator
  // partial plaquette operators plqopr (dmrg_su2_omp.cpp:411-447, 553-590):
  // (plaquette, channel) -> operator rank 2K -> three-leg operator whose
  // operator leg is the single multiplet 2K
  std::map<std::pair<int, int>, std::map<int, std::unique_ptr<Tensor>>> P;
};

struct Fused {
  Leg block, fused;
  std::map<std::pair<int, int>, long long> off;   // (2j block, 2j fused) -> offset
  // parts of fused sector jf: (2j block, offset, dim), by offset
  std::vector<std::tuple<int, long long, int>> parts(int jf) const;
  int dim(int jf) const { return fused.dim_of(jf); }
};

// One solve: kept multiplets of the new bond, squared singular values the
// cut dropped, H_eff applies of its Lanczos run.
struct StepRecord {
  int kept = 0;
  double discarded = 0;
  int matvecs = 0;
};

// One environment kernel applied to one block operator (the reference's
// tensor_su2 members called by dmrg_su2::prepare_site_operator_from_left /
// _right, dmrg_su2_omp.cpp:319-825, tensor_su2_dmrg_src.cpp:27-890), on the
// enlarged basis of `block` with one spin-1/2 site (left: block x site,
// right: site x block), then projected on P ([enlarged, new bond]: U or V of
// the truncated SVD) when P is given:
//   kEnvOpTransform  S_k (x) 1       operator_transformation   (op: 3 legs)
//   kEnvOverlap      H (x) 1         overlap_transformation    (op: 2 legs)
//   kEnvPairup       coef S_k . s    operator_pairup           (op: 3 legs)
//   kEnvInitial      1 (x) s         operator_initial          (no op)
enum EnvKind { kEnvOpTransform = 0, kEnvOverlap = 1, kEnvPairup = 2, kEnvInitial = 3 };
std::unique_ptr<Tensor> env_transform(Context& c, int kind, bool left, const Leg& block,
                                      const Tensor* P, const Tensor* op, double coef);

// Fused leg of a block with one spin-1/2 site: block x site (left) or
// site x block (right), with the offset of every block sector inside the
// fused sectors (tensor_su2::fuse(b1,b2), tensor_su2.cpp:839-909).
Fused make_fused(const Leg& block, bool left);

// A converged lower state for excited-state targeting (orth[j] of
// dmrg_su2, read by read_mps, dmrg_su2_omp.cpp:944-1068).  Each site k is
// held in both forms: left[k] maps fuse(beta[k-1], site) to beta[k] and
// right[k] maps fuse(site, beta[k]) to beta[k-1] (the reference re-labels
// orth between the two with left2right/right2left_vectran; for spin-1/2
// sites that is a per-block transpose with factor 1).  beta[k] is the bond
// right of site k (beta[-1] = {0}, beta[ns-1] = {2S}).  OL[i] / OR[i] are the
// overlap environments with this run (ovlp[j][i], dmrg_su2_omp.cpp:327-328,
// 386, 470-471, 528): legs (beta[i], lb[i]) and (beta[i-1], rb[i]).
struct LowerState {
  std::map<int, Leg> beta;
  std::vector<std::unique_ptr<Tensor>> left, right;
  double energy = 0;   // Ritz value of the lower run (gs_enr[j])
  std::map<int, std::unique_ptr<Tensor>> OL, OR;
};

struct DmrgStats {
  double t_env = 0, t_lanczos = 0, t_svd = 0, t_project = 0, t_start = 0;
  double heff_flops = 0;       // algorithmic FLOPs of all matvecs
  long long matvecs = 0;
  long long steps = 0;
  double t_matvec = 0;         // device time inside H_eff applies (events)
  double t_jacobi = 0;         // time inside the one-sided Jacobi SVD
  long long jacobi_sweeps = 0; // its sweeps, summed over solves
  // mixed precision: FP32 matvecs, their device time and algorithmic
  // FLOPs, and the largest |Ritz value - FP64 Rayleigh quotient| of a solve
  long long matvecs32 = 0;
  double t_matvec32 = 0;
  double heff_flops32 = 0;
  double max_fp64_correction = 0;
  long long svd_degraded = 0;  // SVDs accepted at the sweep ceiling (see jacobi.hpp)
};

class Dmrg {
 public:
  Dmrg(Context& c, int lx, int ly, double jcoup, double qcoup, int bdry, int totspin,
       int max_dcut);
  double warmup();
  double sweep();
  // Checkpoint in the reference's file format (dmrg_su2::save_mps2, save_ww,
  // save_enr, dmrg_su2_omp.cpp:864-941, 1070-1186): uu-*.dat/.bin per site,
  // ww-*.bin, enr-*.dat (appended), named after ns, max_dcut, 2S, ky, exci.
  void save(const std::string& dir, int ky, int exci) const;
  // read_mps + prepare_sweep (dmrg_su2_omp.cpp:944-1068, 160-193): projectors
  // of every site from the files of a run at file_dcut multiplets (0: this
  // run's), environments rebuilt from both edges; sweep() continues from the
  // centre.  Returns the saved energy (read_enr, 1100-1152).
  double load(const std::string& dir, int file_dcut, int ky, int exci);
  // Excited-state targeting (exci > 0): adds the state saved in dir as
  // uu/enr-{ns}-{file_dcut}-{2S}-{ky}-{exci}-* (file_dcut 0: this run's
  // max_dcut) to the states every two-site solve is kept orthogonal to, by
  // the penalty -E_j <phi_j|psi> phi_j of dmrg_su2.cpp:611-623.  Call before
  // warmup(), or after load() and before the next sweep.
  void add_lower_state(const std::string& dir, int file_dcut, int ky, int exci);
  int lower_states() const { return (int)lower_.size(); }
  // The two-site H_eff (without excited-state penalties) and the wavefunction
  // at the centre bond of the current state, after warmup() or sweep().
  std::pair<std::unique_ptr<HeffPlan>, std::unique_ptr<Tensor>> centre();

  // Step cursor over the reference's step order (dmrg_su2_omp.cpp:20-157):
  // begin a phase (the warm-up or one sweep), then take its two-site solves
  // one at a time with next_solve() (nullptr: the phase is over), run each
  // one's Lanczos (lanczos_ground / lanczos_batch with solve_spec) and SVD,
  // and hand both back to complete().  warmup() and sweep() are this loop
  // for one run; BatchDriver runs it for several runs in lockstep.
  struct Solve;
  void begin_warmup();
  void begin_sweep();
  std::unique_ptr<Solve> next_solve();
  void complete(std::unique_ptr<Solve> S, const LanczosResult& lr, SvdResult&& sv);
  double phase_energy() const { return cur_.energy; }
  int max_dcut() const { return D_; }
  Context& context() { return c_; }
  void set_max_dcut(int d) { D_ = d; }
  // 1: one line per solve on stderr (step, sites, energy, kept multiplets,
  // discarded weight, matvecs, seconds), as the reference logs each step
  void set_verbose(int v) { verbose_ = v; }
  const std::vector<double>& step_energies() const { return energies_; }
  const std::vector<StepRecord>& history() const { return history_; }
  const DmrgStats& stats() const { return stats_; }
  int ns() const { return lat_.ns; }
  int last_bond_dim() const { return last_bond_; }
  void set_heff_variants(int v1, int v2) {
    var1_ = v1;
    var2_ = v2;
  }
  void set_heff_variants32(int v1, int v2) {
    var32_[0] = v1;
    var32_[1] = v2;
  }
  // Matvec arithmetic of later solves: FP64, or mixed (FP32 MFMA Krylov
  // matvecs, FP64 Ritz-vector checks whose Rayleigh quotient is the energy;
  // Lanczos vectors, tridiagonal solve and SVD stay FP64)
  void set_matvec_precision(Prec p, double min_flops = -1) {
    prec_ = p;
    min_fp32_flops_ = min_flops < 0 ? kDefaultMinFp32Flops : min_flops;
  }
  // smallest matvec (algorithmic FLOPs) that a mixed-precision run does in
  // FP32: below it the FP32 copies cost more than the faster launches save
  static constexpr double kDefaultMinFp32Flops = 2e8;
  Prec matvec_precision() const { return prec_; }
  // replay each solve's H_eff launches from a HIP graph (default) or launch
  // them directly; the results are the same
  void set_use_graph(bool on) { use_graph_ = on; }

 private:
  // what completing a solve does with its factors (the branches of
  // dmrg_su2_omp.cpp:20-157)
  enum Tag { kTagNone, kTagWarm, kTagWarmFinal, kTagP1, kTagP1Turn, kTagP2, kTagP2Turn, kTagP3 };
  struct StepOut {
    int tag = kTagNone;
    double energy = 0;
    std::unique_ptr<Tensor> U, V;
    Leg bond;
    Fused FL, FR;
    std::unique_ptr<Env> El, Er;
  };
  Fused fused_left(int i) const;
  Fused fused_right(int i) const;
  std::unique_ptr<Env> enlarged_left(int i, const Fused& E);
  std::unique_ptr<Env> enlarged_right(int i, const Fused& E);
  std::unique_ptr<Env> enlarge(const Fused& E, const Env& env, const std::vector<int>& keep,
                               const std::vector<std::pair<double, int>>& pair, int newsite,
                               bool left);
  // the plaquette terms' partial operators of a block after adding site i
  // (out.H receives the plaquettes completed at i)
  void grow_plaquettes(const Fused& E, const Env& env, int i, bool left, Env& out);
  // H_eff terms of the plaquettes that the cut (il, il + 1) splits
  void plaquette_terms(int il, int ir, const Env& El, const Env& Er,
                       std::vector<HeffTerm>& terms) const;
  std::unique_ptr<Env> project(const Tensor& P, const Env& hat);
  enum Phase { kPhaseNone, kPhaseWarmup, kPhaseSweep };
  struct Cursor {
    int phase = kPhaseNone;
    int pass = 0, i = 0, il_last = -1;
    std::unique_ptr<Tensor> A, B;   // factors carried between sweep steps
    Leg bond;
    double energy = 0;
  };
  Cursor cur_;

 public:
  // one two-site solve, split so that several runs can advance in lockstep
  struct Solve {
    int il = 0, ir = 0;
    double p1 = 0, p2 = 0;
    double t_step = 0, t_lanczos0 = 0;
    StepOut o;                                 // fused legs, enlarged environments
    std::unique_ptr<Tensor> psi;               // start vector, then the Ritz vector
    std::unique_ptr<HeffPlan> plan;
    std::vector<std::unique_ptr<Tensor>> phis; // lower states in this basis
    int mlanc = 0;
    bool fp32 = false;
    int tag = kTagNone;
    DevBuf pen;                                // device overlap of the penalty
    ~Solve();
  };
  // adds the lower states' penalty -E_j <phi_j|x> phi_j to y (dmrg_su2.cpp:611-623)
  void penalty(const Solve& S, const double* x, double* y);
  LanczosResult solve(Solve& S);
  SvdResult svd(Solve& S);
  void release_plan(Solve& S);
  // the stats counters a driver of several runs fills in for this one
  DmrgStats& stats_mut() { return stats_; }

 private:
  std::vector<HeffTerm> heff_terms(int il, int ir, const Env& El, const Env& Er) const;
  std::unique_ptr<Solve> prepare(int il, int ir, std::unique_ptr<Tensor> psi0, double p1,
                                 double p2);
  StepOut finish(Solve& S, const LanczosResult& lr, SvdResult&& sv);
  std::unique_ptr<Tensor> start_left(int nl, int nr, const Tensor& A, const Leg& b);
  std::unique_ptr<Tensor> start_right(int nl, int nr, const Tensor& B, const Leg& b);
  void ensure_left(int i);
  void ensure_right(int i);
  // overlap environments of every lower state at site i from U_[i] / V_[i]
  void overlap_left(int i);
  void overlap_right(int i);
  void rebuild_overlaps();
  // each lower state in the two-site basis of (il, il + 1), psi's layout
  std::vector<std::unique_ptr<Tensor>> lower_vectors(int il, int ir, const Fused& FL,
                                                     const Fused& FR, const Tensor& psi);

  Context& c_;
  Lattice lat_;
  int S_, D_;
  std::map<int, Leg> lb_, rb_;
  std::map<int, std::unique_ptr<Env>> L_, R_;
  std::map<int, std::unique_ptr<Tensor>> U_, V_;   // projectors (U at the centre may carry s)
  std::vector<std::unique_ptr<LowerState>> lower_;
  std::vector<double> energies_;
  std::vector<StepRecord> history_;
  std::vector<double> wout_;                        // last solve's singular values
  DmrgStats stats_;
  int last_bond_ = 0;
  int var1_ = -1, var2_ = -1;
  int var32_[2] = {-1, -1};
  Prec prec_ = Prec::F64;
  double min_fp32_flops_ = kDefaultMinFp32Flops;
  bool use_graph_ = true;
  int verbose_ = 0;
};

}  // namespace su2
