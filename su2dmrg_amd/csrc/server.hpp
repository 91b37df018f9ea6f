// Request loop serving DMRG jobs on one device with continuous batching.
//
// The reference runs one job per process invocation (main_omp.cpp:42-202,
// with -e/-dge for excited states, dmrg_su2_omp.cpp:944-1068).  Here jobs of
// different lattices, couplings, spin sectors, bond dimensions and target
// levels arrive over time and share the device:
//   * a job is admitted, in arrival order, once it has arrived, a slot is
//     free and its estimated device footprint fits what the admitted jobs
//     leave of the budget;
//   * admitted jobs keep their MPS and environments on the device and
//     advance together, one two-site solve each per BatchDriver tick (one
//     batched Lanczos and one batched SVD for all of them);
//   * a finished job frees its slot and memory at once; a job arriving while
//     others run joins at the next tick;
//   * with several lanes, the slots are split over lanes, each with its own
//     context (stream), BatchDriver and host thread; the next arrived job
//     that fits goes to a lane with a free slot and the fewest live jobs
//     (a lane's tick waits for its slowest job's step, so spreading the jobs
//     shortens every tick), and the lanes' host work overlaps each other's
//     device work (run_batches_parallel).
// An excited-state job (level k > 0) runs its ground state and the levels
// below k first, each stage saving its state in the reference's file format
// for the next (add_lower_state), all inside its slot.
#pragma once

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "batch.hpp"
#include "dmrg.hpp"

namespace su2 {

struct JobSpec {
  int lx = 4, ly = 4;
  double jcoup = 2.0, qcoup = 0.0;
  int bdry = 2, totspin = 0, max_dcut = 100;
  int exci = 0;        // target level: 0 ground state, k: k-th excited state
  int nsweep = 2;      // sweeps after the warm-up, per stage
  double stop = -1;    // |dE| early stop between sweeps (< 0: off)
};

struct JobResult {
  double energy = 0;         // final Ritz value of the target level
  double t_arrive = 0, t_admit = 0, t_finish = 0;   // seconds from the start of run()
  double footprint = 0;      // estimated device bytes the job was admitted with
  int stage = 0;             // levels completed
  bool done = false;
};

// Estimated device bytes of a job: left and right environments of every
// site (block Hamiltonian and the retained spin operators, about ly + 2
// operators of ~0.8 D^2 doubles each), the two-site H_eff bundles and Lanczos
// vectors of a step (enlarged dimension 2D), and SVD work space.
double job_footprint(const JobSpec& s);

class Server {
 public:
  Server(Context& c, double budget_bytes, int max_slots, const std::string& workdir);
  ~Server();
  int submit(const JobSpec& s, double arrival_s);
  // Runs the loop until every submitted job has finished; arrival times
  // count from the start of this call (an idle device waits for the next).
  void run();
  // Lanes (1..8, default 1); the extra lanes' contexts are made on the
  // device of the server's context for each run().
  void set_lanes(int n);
  int lanes() const { return lanes_; }
  const JobResult& result(int id) const { return res_.at(id); }
  int njobs() const { return (int)specs_.size(); }
  double wall() const { return wall_; }
  long long ticks() const { return ticks_; }
  int max_concurrent() const { return max_conc_; }
  void set_verbose(int v) { verbose_ = v; }

 private:
  struct Slot {
    int job = -1;
    int stage = 0;
    int run_id = -1;
    std::unique_ptr<Dmrg> d;
    std::string dir;
  };
  struct Lane;
  void start_stage(Lane& l, Slot& s);
  void lane_loop(Lane& l);
  Context& c_;
  double budget_;
  int max_slots_;
  int lanes_ = 1;
  std::string workdir_;
  std::vector<JobSpec> specs_;
  std::vector<double> arrive_;
  std::vector<JobResult> res_;
  double wall_ = 0;
  long long ticks_ = 0;
  int max_conc_ = 0;
  int verbose_ = 0;
  // shared by the lanes of one run(), under mu_
  std::mutex mu_;
  std::vector<int> order_;
  std::vector<int> lane_live_, lane_slots_;   // live jobs and slots per lane
  size_t next_ = 0;
  double used_ = 0, t0_ = 0;
  int finished_ = 0, live_ = 0;
  bool failed_ = false;
};

}  // namespace su2
