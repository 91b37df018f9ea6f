#include "server.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <thread>

#include "common.hpp"
#include "context.hpp"

namespace su2 {

namespace {

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

double job_footprint(const JobSpec& s) {
  const double D = s.max_dcut, De = 2.0 * s.max_dcut;
  const double ns = (double)s.lx * s.ly;
  const double ops = s.ly + 2 + (s.qcoup != 0 ? 2 * s.ly : 0);
  // left and right environments of every site, ~0.8 D^2 doubles per operator
  // (measured: the 16x8 D=2000 run peaks at 75 GB, this gives 82 GB)
  const double env = 2.0 * ns * ops * 0.8 * D * D;
  const double plan = 2.0 * ops * 3.0 * De * De / 4;   // H_eff bundles (both sides)
  const double krylov = 10.0 * De * De / 4;            // Lanczos vectors, SVD work space
  const double stages = s.exci > 0 ? 1.0 + 0.5 * s.exci : 1.0;   // lower states held
  return 8.0 * (env * stages + plan + krylov);
}

Server::Server(Context& c, double budget_bytes, int max_slots, const std::string& workdir)
    : c_(c), budget_(budget_bytes), max_slots_(max_slots), workdir_(workdir) {
  SU2_CHECK(budget_bytes > 0, kErrArgument, "server: budget %g bytes", budget_bytes);
  SU2_CHECK(max_slots >= 1, kErrArgument, "server: %d slots", max_slots);
}

Server::~Server() = default;

int Server::submit(const JobSpec& s, double arrival_s) {
  SU2_CHECK(s.exci >= 0 && s.exci <= 4, kErrArgument, "server: level %d", s.exci);
  SU2_CHECK(s.nsweep >= 1 && s.max_dcut >= 1, kErrArgument, "server: job program");
  SU2_CHECK(job_footprint(s) <= budget_, kErrCapacity,
            "server: job needs %.3g bytes, over the budget %.3g", job_footprint(s), budget_);
  specs_.push_back(s);
  arrive_.push_back(arrival_s);
  JobResult r;
  r.t_arrive = arrival_s;
  r.footprint = job_footprint(s);
  res_.push_back(r);
  return (int)specs_.size() - 1;
}

// a new run for the slot's current stage: the ground state, or level k kept
// orthogonal to the levels below it (read back from the job's directory)
void Server::start_stage(Slot& s) {
  const JobSpec& j = specs_[s.job];
  s.d.reset(new Dmrg(c_, j.lx, j.ly, j.jcoup, j.qcoup, j.bdry, j.totspin, j.max_dcut));
  for (int k = 0; k < s.stage; ++k) s.d->add_lower_state(s.dir, 0, 0, k);
  std::vector<RunPhase> prog;
  RunPhase w;
  w.warmup = true;
  w.max_dcut = j.max_dcut;
  prog.push_back(w);
  for (int k = 0; k < j.nsweep; ++k) {
    RunPhase p;
    p.max_dcut = j.max_dcut;
    prog.push_back(p);
  }
  s.run_id = drv_->add(s.d.get(), prog, j.stop);
}

void Server::run() {
  c_.make_current();
  drv_.reset(new BatchDriver(c_));
  const int nj = (int)specs_.size();
  std::vector<int> order(nj);
  for (int i = 0; i < nj; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return arrive_[a] < arrive_[b]; });
  std::vector<Slot> slots(max_slots_);
  size_t next = 0;
  double used = 0;
  int finished = 0;
  const double t0 = now();
  while (finished < nj) {
    const double t = now() - t0;
    // admit arrived jobs, in arrival order, while a slot and memory are free
    while (next < order.size() && arrive_[order[next]] <= t) {
      const int id = order[next];
      auto it = std::find_if(slots.begin(), slots.end(), [](const Slot& s) { return s.job < 0; });
      if (it == slots.end() || used + res_[id].footprint > budget_) break;
      ++next;
      Slot& s = *it;
      s.job = id;
      s.stage = 0;
      char buf[64];
      std::snprintf(buf, sizeof buf, "/job%d", id);
      s.dir = workdir_ + buf;
      if (specs_[id].exci > 0) mkdir(s.dir.c_str(), 0755);
      used += res_[id].footprint;
      res_[id].t_admit = now() - t0;
      start_stage(s);
      if (verbose_) std::fprintf(stderr, "server: t=%.2fs admit job %d\n", now() - t0, id);
    }
    int live = 0;
    for (auto& s : slots) live += s.job >= 0;
    max_conc_ = std::max(max_conc_, live);
    if (live == 0) {
      // idle until the next arrival
      if (next < order.size()) {
        const double wait = arrive_[order[next]] - (now() - t0);
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      }
      continue;
    }
    drv_->tick();
    ++ticks_;
    // retire finished stages: next level, or the job is done
    for (auto& s : slots) {
      if (s.job < 0 || !drv_->finished(s.run_id)) continue;
      const JobSpec& j = specs_[s.job];
      JobResult& r = res_[s.job];
      const std::vector<double>& en = drv_->phase_energies(s.run_id);
      r.energy = en.empty() ? 0 : en.back();
      r.stage = s.stage + 1;
      if (s.stage < j.exci) {
        s.d->save(s.dir, 0, s.stage);
        ++s.stage;
        start_stage(s);
        continue;
      }
      r.t_finish = now() - t0;
      r.done = true;
      used -= r.footprint;
      ++finished;
      if (verbose_)
        std::fprintf(stderr, "server: t=%.2fs job %d done, E=%.12f\n", r.t_finish, s.job,
                     r.energy);
      s.d.reset();
      s.job = -1;
      s.run_id = -1;
    }
  }
  wall_ = now() - t0;
}

}  // namespace su2
