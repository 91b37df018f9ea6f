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
#include <tuple>
#include <vector>

#include "lattice.hpp"
#include "structure.hpp"
#include "tensor.hpp"

namespace su2 {

class Context;

struct Env {
  Leg bond;
  std::unique_ptr<Tensor> H;                      // scalar block Hamiltonian (may be null)
  std::map<int, std::unique_ptr<Tensor>> S;       // site -> spin operator
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

struct DmrgStats {
  double t_env = 0, t_lanczos = 0, t_svd = 0, t_project = 0, t_start = 0;
  double heff_flops = 0;       // algorithmic FLOPs of all matvecs
  long long matvecs = 0;
  long long steps = 0;
  double t_matvec = 0;         // device time inside H_eff applies (events)
};

class Dmrg {
 public:
  Dmrg(Context& c, int lx, int ly, double jcoup, double qcoup, int bdry, int totspin,
       int max_dcut);
  double warmup();
  double sweep();
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

 private:
  struct StepOut {
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
  std::unique_ptr<Env> project(const Tensor& P, const Env& hat);
  StepOut step(int il, int ir, std::unique_ptr<Tensor> psi0, double p1, double p2);
  std::unique_ptr<Tensor> start_left(int nl, int nr, const Tensor& A, const Leg& b);
  std::unique_ptr<Tensor> start_right(int nl, int nr, const Tensor& B, const Leg& b);
  void ensure_left(int i);
  void ensure_right(int i);

  Context& c_;
  Lattice lat_;
  int S_, D_;
  std::map<int, Leg> lb_, rb_;
  std::map<int, std::unique_ptr<Env>> L_, R_;
  std::map<int, std::unique_ptr<Tensor>> U_, V_;   // projectors (U at the centre may carry s)
  std::vector<double> energies_;
  std::vector<StepRecord> history_;
  DmrgStats stats_;
  int last_bond_ = 0;
  int var1_ = -1, var2_ = -1;
  int verbose_ = 0;
};

}  // namespace su2
