#include "batch.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <set>
#include <thread>

#include "common.hpp"
#include "context.hpp"
#include "heff.hpp"
#include "svd.hpp"

namespace su2 {

namespace {

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// K splits of the batched launches: the same per-problem rules as a plan's
// own launches (HeffPlan), so every problem gets the split it has alone
constexpr int kMinUnits1 = 128, kKtPerUnit1 = 128;
constexpr int kMinUnits2 = 512, kKtPerUnit2 = 96;

}  // namespace

// ---------------------------------------------------------------------------
// HeffBatch

HeffBatch::~HeffBatch() { clear(); }

void HeffBatch::release(Pattern& P) {
  if (P.exec) (void)hipGraphExecDestroy(P.exec);
  if (P.graph) (void)hipGraphDestroy(P.graph);
  if (P.done) (void)hipEventDestroy(P.done);
  if (P.h_desc) (void)hipHostFree(P.h_desc);
  if (P.d_desc) c_.dev_free(P.d_desc);
  P.exec = nullptr;
  P.graph = nullptr;
  P.done = nullptr;
  P.h_desc = nullptr;
  P.d_desc = nullptr;
}

void HeffBatch::clear() {
  if (!cache_.empty()) (void)hipStreamSynchronize(c_.stream());
  for (auto& kv : cache_) release(*kv.second);
  cache_.clear();
}

HeffBatch::Pattern& HeffBatch::pattern(const std::vector<Req>& reqs) {
  std::vector<std::pair<const void*, int>> key;
  for (auto& r : reqs) key.push_back({r.plan, (int)r.prec});
  auto it = cache_.find(key);
  if (it != cache_.end()) return *it->second;
  std::unique_ptr<Pattern> P(new Pattern());
  P->nreq = (int)reqs.size();
  int var[2][2] = {{-1, -1}, {-1, -1}};   // [prec][stage]: the first plan's variants
  for (auto& r : reqs) {
    HeffPlan& h = *r.plan;
    const int pi = (int)r.prec;
    if (r.prec == Prec::F32) h.ensure_fp32();
    GroupedGemm& g1 = r.prec == Prec::F64 ? h.stage1_ : *h.stage1_32_;
    GroupedGemm& g2 = r.prec == Prec::F64 ? h.stage2_ : *h.stage2_32_;
    if (var[pi][0] < 0) {
      var[pi][0] = g1.variant();
      var[pi][1] = g2.variant();
    }
    P->s1[pi].append(g1);
    P->s2[pi].append(g2);
    P->max_in = std::max(P->max_in, h.psi_size_);
    for (auto& z : h.zero_blocks_) {
      VecMove m{};
      m.dst = r.prec == Prec::F64 ? (void*)(h.d_out_ + z.first) : (void*)(h.d_out32_ + z.first);
      m.n = z.second;
      m.mode = r.prec == Prec::F64 ? kZeroF64 : kZeroF32;
      P->fill.push_back(m);
      P->max_fill = std::max(P->max_fill, z.second);
    }
  }
  P->max_out = P->max_in;
  for (int pi = 0; pi < 2; ++pi) {
    if (P->s1[pi].nproblems() > 0) P->s1[pi].finalize(c_, kMinUnits1, kKtPerUnit1, var[pi][0]);
    if (P->s2[pi].nproblems() > 0) P->s2[pi].finalize(c_, kMinUnits2, kKtPerUnit2, var[pi][1]);
  }
  const size_t nd = 2 * reqs.size() + P->fill.size();
  P->d_desc = (VecMove*)c_.dev_alloc(nd * sizeof(VecMove));
  SU2_HIP(hipHostMalloc((void**)&P->h_desc, 2 * reqs.size() * sizeof(VecMove),
                        hipHostMallocDefault));
  if (!P->fill.empty())
    SU2_HIP(hipMemcpyAsync(P->d_desc + 2 * reqs.size(), P->fill.data(),
                           P->fill.size() * sizeof(VecMove), hipMemcpyHostToDevice, c_.stream()));
  SU2_HIP(hipEventCreateWithFlags(&P->done, hipEventDisableTiming));
  Pattern& ref = *P;
  cache_[key] = std::move(P);
  return ref;
}

// inputs in, zero blocks, both stages of both precisions, outputs out
void HeffBatch::launch_body(Pattern& P) {
  hipStream_t s = c_.stream();
  dev_move_batch(c_, P.d_desc, P.nreq, P.max_in);
  if (!P.fill.empty()) dev_move_batch(c_, P.d_desc + 2 * P.nreq, (int)P.fill.size(), P.max_fill);
  for (int pi = 0; pi < 2; ++pi)
    if (P.s1[pi].nproblems() > 0) P.s1[pi].launch(s);
  for (int pi = 0; pi < 2; ++pi)
    if (P.s2[pi].nproblems() > 0) P.s2[pi].launch(s);
  dev_move_batch(c_, P.d_desc + P.nreq, P.nreq, P.max_out);
}

void HeffBatch::apply(const std::vector<Req>& reqs) {
  if (reqs.empty()) return;
  c_.make_current();
  Pattern& P = pattern(reqs);
  // the previous replay must have read the staged descriptors
  if (P.used) SU2_HIP(hipEventSynchronize(P.done));
  for (size_t i = 0; i < reqs.size(); ++i) {
    const Req& r = reqs[i];
    HeffPlan& h = *r.plan;
    VecMove in{}, out{};
    in.src = r.x;
    in.n = h.psi_size_;
    out.dst = r.y;
    out.n = h.psi_size_;
    if (r.prec == Prec::F64) {
      in.dst = h.d_in_;
      in.mode = kMoveF64;
      out.src = h.d_out_;
      out.mode = kMoveF64;
    } else {
      in.dst = h.d_in32_;
      in.mode = kMoveF64toF32;
      out.src = h.d_out32_;
      out.mode = kMoveF32toF64;
    }
    P.h_desc[i] = in;
    P.h_desc[reqs.size() + i] = out;
  }
  SU2_HIP(hipMemcpyAsync(P.d_desc, P.h_desc, 2 * reqs.size() * sizeof(VecMove),
                         hipMemcpyHostToDevice, c_.stream()));
  if (!use_graph_) {
    launch_body(P);
  } else {
    if (!P.exec) {
      // the launch sequence is recorded once (kernels only: fills are
      // kernels, no memset nodes) and replayed for every round
      hipStream_t s = c_.stream();
      SU2_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        launch_body(P);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      SU2_HIP(hipStreamEndCapture(s, &P.graph));
      hipError_t e = hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0);
      if (e == hipErrorOutOfMemory) {
        // the device memory caches of every context on this device give
        // back their free blocks, then one more try
        (void)hipGetLastError();
        DevicePool::release_device(c_.device());
        e = hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0);
      }
      SU2_HIP(e);
      ++graphs_built_;
    }
    SU2_HIP(hipGraphLaunch(P.exec, c_.stream()));
    ++replays_;
  }
  SU2_HIP(hipEventRecord(P.done, c_.stream()));
  P.used = true;
}

// ---------------------------------------------------------------------------
// BatchDriver

int BatchDriver::add(Dmrg* d, const std::vector<RunPhase>& program, double stop) {
  SU2_CHECK(d != nullptr, kErrArgument, "batch: null run");
  SU2_CHECK(!program.empty(), kErrArgument, "batch: empty program");
  SU2_CHECK(&d->context() == &c_, kErrArgument, "batch: run of another context");
  Run r;
  r.d = d;
  r.program = program;
  r.stop = stop;
  runs_.push_back(std::move(r));
  return (int)runs_.size() - 1;
}

bool BatchDriver::all_finished() const {
  for (auto& r : runs_)
    if (!r.done) return false;
  return true;
}

int BatchDriver::active() const {
  int n = 0;
  for (auto& r : runs_) n += r.done ? 0 : 1;
  return n;
}

std::unique_ptr<Dmrg::Solve> BatchDriver::next(Run& r) {
  for (;;) {
    if (r.phase >= 0) {
      std::unique_ptr<Dmrg::Solve> S = r.d->next_solve();
      if (S) return S;
      // the phase is over: its energy, then the early-stop rule
      const double e = r.d->phase_energy();
      r.energies.push_back(e);
      const size_t k = r.energies.size();
      if (r.stop > 0 && k >= 2 && !r.program[r.phase].warmup &&
          std::fabs(r.energies[k - 1] - r.energies[k - 2]) < r.stop) {
        r.done = true;
        return nullptr;
      }
    }
    if (r.phase + 1 >= (int)r.program.size()) {
      r.done = true;
      return nullptr;
    }
    ++r.phase;
    const RunPhase& ph = r.program[r.phase];
    if (ph.max_dcut > 0) r.d->set_max_dcut(ph.max_dcut);
    r.d->set_matvec_precision(ph.prec, ph.min_fp32_flops);
    if (ph.warmup) r.d->begin_warmup();
    else r.d->begin_sweep();
  }
}

int BatchDriver::tick() {
  c_.make_current();
  double t0 = now();
  std::vector<int> ids;
  std::vector<std::unique_ptr<Dmrg::Solve>> solves;
  for (int i = 0; i < (int)runs_.size(); ++i) {
    Run& r = runs_[i];
    if (r.done) continue;
    std::unique_ptr<Dmrg::Solve> S = next(r);
    if (!S) continue;
    ids.push_back(i);
    solves.push_back(std::move(S));
  }
  t_prepare += now() - t0;
  if (solves.empty()) return 0;
  ++ticks_;
  const int ns = (int)solves.size();
  // one batched Lanczos: every solve's matvec of a round in one apply
  t0 = now();
  std::vector<LanczosJob> jobs(ns);
  for (int q = 0; q < ns; ++q) {
    Dmrg::Solve& S = *solves[q];
    jobs[q].psi = S.psi->data();
    jobs[q].n = S.psi->size();
    jobs[q].mlanc = S.mlanc;
    jobs[q].mixed = S.fp32;
  }
  // batched-apply device time from an event pair per round, read at the end
  std::vector<hipEvent_t> ev;
  auto release_events = [&]() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    ev.clear();
  };
  auto apply = [&](const std::vector<LanczosReq>& reqs) {
    std::vector<HeffBatch::Req> hr;
    for (auto& r : reqs) {
      Dmrg::Solve& S = *solves[r.job];
      const Prec p = (S.fp32 && !r.check) ? Prec::F32 : Prec::F64;
      hr.push_back(HeffBatch::Req{S.plan.get(), p, r.x, r.y});
    }
    hipEvent_t e0, e1;
    SU2_HIP(hipEventCreate(&e0));
    ev.push_back(e0);
    SU2_HIP(hipEventCreate(&e1));
    ev.push_back(e1);
    SU2_HIP(hipEventRecord(e0, c_.stream()));
    heff_.apply(hr);
    SU2_HIP(hipEventRecord(e1, c_.stream()));
    for (auto& r : reqs) runs_[ids[r.job]].d->penalty(*solves[r.job], r.x, r.y);
  };
  float mv_ms = 0;
  try {
    lanczos_batch(c_, jobs, apply);
    for (size_t i = 0; i + 1 < ev.size(); i += 2) {
      float ms = 0;
      SU2_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      mv_ms += ms;
    }
  } catch (...) {
    heff_.clear();
    release_events();
    throw;
  }
  heff_.clear();   // the plans go away below; their cached launches first
  release_events();
  t_matvec += mv_ms * 1e-3;
  for (int q = 0; q < ns; ++q) {
    Dmrg::Solve& S = *solves[q];
    DmrgStats& st = runs_[ids[q]].d->stats_mut();
    const LanczosResult& lr = jobs[q].res;
    const int n32 = S.fp32 ? lr.matvecs - lr.check_matvecs : 0;
    st.matvecs += lr.matvecs;
    st.matvecs32 += n32;
    st.heff_flops += S.plan->algorithmic_flops() * lr.matvecs;
    st.heff_flops32 += S.plan->algorithmic_flops() * n32;
    if (S.fp32)
      st.max_fp64_correction = std::max(st.max_fp64_correction, std::fabs(lr.ritz - lr.energy));
    runs_[ids[q]].d->release_plan(S);
  }
  t_lanczos += now() - t0;
  // one batched SVD over every solve's wavefunction
  t0 = now();
  std::vector<SvdJob> sj;
  for (int q = 0; q < ns; ++q)
    sj.push_back(SvdJob{solves[q]->psi.get(), solves[q]->p1, solves[q]->p2,
                        runs_[ids[q]].d->max_dcut()});
  std::vector<SvdResult> sv = truncated_svd_batch(c_, sj);
  t_svd += now() - t0;
  t0 = now();
  for (int q = 0; q < ns; ++q)
    runs_[ids[q]].d->complete(std::move(solves[q]), jobs[q].res, std::move(sv[q]));
  t_complete += now() - t0;
  return ns;
}

// ---------------------------------------------------------------------------
// lanes

std::vector<std::exception_ptr> run_batches_parallel(const std::vector<BatchDriver*>& drivers) {
  std::set<const Context*> seen;
  for (BatchDriver* d : drivers) {
    SU2_CHECK(d != nullptr, kErrArgument, "lanes: null batch");
    SU2_CHECK(seen.insert(&d->context()).second, kErrArgument,
              "lanes: two batches on one context (each lane needs its own stream)");
  }
  std::vector<std::exception_ptr> err(drivers.size());
  std::atomic<bool> failed{false};
  auto lane = [&](size_t i) {
    try {
      BatchDriver& d = *drivers[i];
      d.context().make_current();   // the device is per host thread
      while (!failed.load() && d.tick() > 0) {
      }
    } catch (...) {
      err[i] = std::current_exception();
      failed.store(true);
    }
  };
  std::vector<std::thread> th;
  try {
    for (size_t i = 1; i < drivers.size(); ++i) th.emplace_back(lane, i);
  } catch (...) {   // no thread: stop the started lanes before reporting
    failed.store(true);
    for (auto& t : th) t.join();
    throw;
  }
  if (!drivers.empty()) lane(0);
  for (auto& t : th) t.join();
  return err;
}

}  // namespace su2
