#include "server.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <exception>
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

struct Server::Lane {
  int idx = 0;
  Context* c = nullptr;
  std::unique_ptr<Context> own;   // extra lanes' contexts
  std::unique_ptr<BatchDriver> drv;
  std::vector<Slot> slots;
  long long ticks = 0;
};

void Server::set_lanes(int n) {
  SU2_CHECK(n >= 1 && n <= 8, kErrArgument, "server: %d lanes", n);
  lanes_ = n;
}

// a new run for the slot's current stage: the ground state, or level k kept
// orthogonal to the levels below it (read back from the job's directory)
void Server::start_stage(Lane& l, Slot& s) {
  const JobSpec& j = specs_[s.job];
  s.d.reset(new Dmrg(*l.c, j.lx, j.ly, j.jcoup, j.qcoup, j.bdry, j.totspin, j.max_dcut));
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
  s.run_id = l.drv->add(s.d.get(), prog, j.stop);
}

// One lane: admit arrived jobs into its free slots (arrival order over all
// lanes, one memory budget), tick its driver, retire finished stages.
void Server::lane_loop(Lane& l) {
  l.c->make_current();
  const int nj = (int)specs_.size();
  for (;;) {
    std::vector<Slot*> admitted;
    double wait = 0;
    bool more = false;
    {
      std::lock_guard<std::mutex> g(mu_);
      if (failed_ || finished_ == nj) return;
      const double t = now() - t0_;
      while (next_ < order_.size() && arrive_[order_[next_]] <= t) {
        const int id = order_[next_];
        auto it = std::find_if(l.slots.begin(), l.slots.end(),
                               [](const Slot& s) { return s.job < 0; });
        if (it == l.slots.end() || used_ + res_[id].footprint > budget_) break;
        // the job goes to a lane with a free slot and the fewest live jobs
        int least = lane_live_[l.idx];
        for (size_t k = 0; k < lane_live_.size(); ++k)
          if (lane_live_[k] < lane_slots_[k]) least = std::min(least, lane_live_[k]);
        if (lane_live_[l.idx] > least) break;
        ++next_;
        Slot& s = *it;
        s.job = id;
        s.stage = 0;
        char buf[64];
        std::snprintf(buf, sizeof buf, "/job%d", id);
        s.dir = workdir_ + buf;
        used_ += res_[id].footprint;
        res_[id].t_admit = t;
        ++live_;
        ++lane_live_[l.idx];
        max_conc_ = std::max(max_conc_, live_);
        admitted.push_back(&s);
      }
      more = next_ < order_.size();
      if (more) wait = arrive_[order_[next_]] - t;
    }
    for (Slot* s : admitted) {
      if (specs_[s->job].exci > 0) mkdir(s->dir.c_str(), 0755);
      start_stage(l, *s);
      if (verbose_) std::fprintf(stderr, "server: t=%.2fs admit job %d\n", now() - t0_, s->job);
    }
    int live = 0;
    for (auto& s : l.slots) live += s.job >= 0;
    if (live == 0) {
      if (!more) return;   // nothing left to admit; the other lanes finish theirs
      // idle until the next arrival (or until another lane frees memory)
      const double d = std::min(std::max(wait, 0.0), 2e-3);
      std::this_thread::sleep_for(std::chrono::duration<double>(d > 0 ? d : 1e-4));
      continue;
    }
    l.drv->tick();
    ++l.ticks;
    // retire finished stages: next level, or the job is done
    for (auto& s : l.slots) {
      if (s.job < 0 || !l.drv->finished(s.run_id)) continue;
      const JobSpec& j = specs_[s.job];
      JobResult& r = res_[s.job];
      const std::vector<double>& en = l.drv->phase_energies(s.run_id);
      r.energy = en.empty() ? 0 : en.back();
      r.stage = s.stage + 1;
      if (s.stage < j.exci) {
        s.d->save(s.dir, 0, s.stage);
        ++s.stage;
        start_stage(l, s);
        continue;
      }
      s.d.reset();
      {
        std::lock_guard<std::mutex> g(mu_);
        r.t_finish = now() - t0_;
        r.done = true;
        used_ -= r.footprint;
        ++finished_;
        --live_;
        --lane_live_[l.idx];
      }
      if (verbose_)
        std::fprintf(stderr, "server: t=%.2fs job %d done, E=%.12f\n", r.t_finish, s.job,
                     r.energy);
      s.job = -1;
      s.run_id = -1;
    }
  }
}

void Server::run() {
  c_.make_current();
  const int nj = (int)specs_.size();
  order_.resize(nj);
  for (int i = 0; i < nj; ++i) order_[i] = i;
  std::stable_sort(order_.begin(), order_.end(),
                   [&](int a, int b) { return arrive_[a] < arrive_[b]; });
  next_ = 0;
  used_ = 0;
  finished_ = 0;
  live_ = 0;
  failed_ = false;
  max_conc_ = 0;
  const int nl = std::min(lanes_, max_slots_);
  std::vector<std::unique_ptr<Lane>> lanes;
  lane_live_.assign(nl, 0);
  lane_slots_.assign(nl, 0);
  for (int k = 0; k < nl; ++k) {
    std::unique_ptr<Lane> l(new Lane());
    l->idx = k;
    if (k == 0) {
      l->c = &c_;
    } else {
      l->own.reset(new Context(c_.device()));
      l->c = l->own.get();
    }
    l->drv.reset(new BatchDriver(*l->c));
    l->slots.resize(max_slots_ / nl + (k < max_slots_ % nl ? 1 : 0));
    lane_slots_[k] = (int)l->slots.size();
    lanes.push_back(std::move(l));
  }
  std::vector<std::exception_ptr> err(nl);
  auto go = [&](int k) {
    try {
      lane_loop(*lanes[k]);
    } catch (...) {
      err[k] = std::current_exception();
      std::lock_guard<std::mutex> g(mu_);
      failed_ = true;
    }
  };
  t0_ = now();
  std::vector<std::thread> th;
  try {
    for (int k = 1; k < nl; ++k) th.emplace_back(go, k);
  } catch (...) {
    {
      std::lock_guard<std::mutex> g(mu_);
      failed_ = true;
    }
    for (auto& t : th) t.join();
    throw;
  }
  go(0);
  for (auto& t : th) t.join();
  wall_ = now() - t0_;
  ticks_ = 0;
  for (auto& l : lanes) ticks_ += l->ticks;
  // each lane's runs, then its driver, before its context
  for (auto& l : lanes) {
    for (auto& s : l->slots) s.d.reset();
    l->drv.reset();
  }
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

}  // namespace su2
