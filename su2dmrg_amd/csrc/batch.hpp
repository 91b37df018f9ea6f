// Several independent solves advanced together on one device.
//
// The reference runs one DMRG job per process and one H_eff apply at a time
// (lanczos_su2::diag_op, lanczos_su2.cpp:335-338).  Here the H_eff applies of
// many concurrent Lanczos solves -- of one run or of many runs (a parameter
// scan, a served stream of jobs) -- go out as one grouped launch per GEMM
// stage, captured in a HIP graph and replayed every Lanczos round:
//
//   HeffBatch     y_i = H_i x_i for a list of (plan, precision, x, y)
//   lanczos_batch the reference's Lanczos schedule (lanczos_su2.cpp:125-195,
//                 237-338) for many solves in lockstep, one host wait per
//                 round for all of them
//   BatchDriver   several DMRG runs stepped through their warm-up and sweeps
//                 in lockstep: one batched Lanczos and one batched SVD per
//                 step for all runs
//
// Every problem keeps its own K split and every solve its own reduction
// order, so each result has the same bits alone (a batch of one, which is
// how a single run executes) and in any mix.
#pragma once

#include <hip/hip_runtime.h>

#include <exception>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "blas1.hpp"
#include "dmrg.hpp"
#include "gemm.hpp"
#include "lanczos.hpp"

namespace su2 {

class Context;
class HeffPlan;

class HeffBatch {
 public:
  struct Req {
    HeffPlan* plan;
    Prec prec;
    const double* x;
    double* y;
  };
  explicit HeffBatch(Context& c) : c_(c) {}
  ~HeffBatch();
  HeffBatch(const HeffBatch&) = delete;
  HeffBatch& operator=(const HeffBatch&) = delete;
  // y_i = H_i x_i for every request (plans must stay alive while this
  // object caches launches for them; call clear() before destroying plans).
  // The first apply for a given list of (plan, precision) builds its grouped
  // launches and captures them in a HIP graph; later applies for the same
  // list replay that graph with new x and y.  Does not wait.
  void apply(const std::vector<Req>& reqs);
  void clear();
  // Use the graph (default) or launch the same kernels directly.
  void set_use_graph(bool on) { use_graph_ = on; }
  int graphs_built() const { return graphs_built_; }
  long long replays() const { return replays_; }

 private:
  struct Pattern {
    GroupedGemm s1[2], s2[2];   // by precision: F64, F32
    std::vector<VecMove> fill;  // zero blocks of the outputs (fixed)
    int nreq = 0;
    long long max_in = 0, max_out = 0, max_fill = 0;
    VecMove* d_desc = nullptr;  // [nreq in | nreq out | fill]
    VecMove* h_desc = nullptr;  // pinned staging of the in/out descriptors
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;  // the last replay has read h_desc
    bool used = false;
    Pattern() : s1{GroupedGemm(Prec::F64), GroupedGemm(Prec::F32)},
                s2{GroupedGemm(Prec::F64), GroupedGemm(Prec::F32)} {}
  };
  Pattern& pattern(const std::vector<Req>& reqs);
  void launch_body(Pattern& P);
  void release(Pattern& P);
  Context& c_;
  std::map<std::vector<std::pair<const void*, int>>, std::unique_ptr<Pattern>> cache_;
  bool use_graph_ = true;
  int graphs_built_ = 0;
  long long replays_ = 0;
};

// A DMRG run's program: the warm-up, then sweeps, each with its kept
// multiplets and matvec arithmetic; optionally ending once the energy of two
// successive sweeps differs by less than stop (main_omp.cpp:184-202).
struct RunPhase {
  bool warmup = false;
  int max_dcut = 0;
  Prec prec = Prec::F64;
  double min_fp32_flops = -1;
};

class BatchDriver {
 public:
  explicit BatchDriver(Context& c) : c_(c), heff_(c) {}
  // Adds a run (not owned; it must outlive its completion).  Returns its id.
  int add(Dmrg* d, const std::vector<RunPhase>& program, double stop = -1);
  // One lockstep step: the next two-site solve of every unfinished run, one
  // batched Lanczos, one batched SVD, then each run's environment update.
  // Returns the number of runs that took a step.
  int tick();
  bool finished(int id) const { return runs_.at(id).done; }
  bool all_finished() const;
  int active() const;
  // energy after each finished phase (warm-up first)
  const std::vector<double>& phase_energies(int id) const { return runs_.at(id).energies; }
  void set_use_graph(bool on) { heff_.set_use_graph(on); }
  int graphs_built() const { return heff_.graphs_built(); }
  long long graph_replays() const { return heff_.replays(); }
  long long ticks() const { return ticks_; }
  Context& context() const { return c_; }
  double t_lanczos = 0, t_svd = 0, t_prepare = 0, t_complete = 0, t_matvec = 0;

 private:
  struct Run {
    Dmrg* d = nullptr;
    std::vector<RunPhase> program;
    double stop = -1;
    int phase = -1;            // index into program, -1 before the first
    bool done = false;
    std::vector<double> energies;
  };
  // the next solve of run r, starting its following phase when one ends
  std::unique_ptr<Dmrg::Solve> next(Run& r);
  Context& c_;
  HeffBatch heff_;
  std::vector<Run> runs_;
  long long ticks_ = 0;
};

// Runs several drivers to completion at once, one host thread each (lanes).
// Each driver must be on its own context, so on its own stream: the host
// work of one lane (plan builds, environment scatters, its waits for Lanczos
// scalars and Jacobi sweeps) overlaps the device work of the others, and
// latency-bound launches of different lanes share the chip.  A run's
// arithmetic is the same in any lane, so its results are bit for bit those
// of one lane.  The returned vector holds each lane's exception (null if it
// finished); after a failure the other lanes stop at their next tick.
std::vector<std::exception_ptr> run_batches_parallel(const std::vector<BatchDriver*>& drivers);

}  // namespace su2
