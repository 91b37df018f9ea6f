// extern "C" entry points declared in include/su2dmrg.h.
#include "../../include/su2dmrg.h"

#include <cmath>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "blas1.hpp"
#include "cg.hpp"
#include "common.hpp"
#include "context.hpp"
#include "precond.hpp"
#include "dmrg.hpp"
#include "heff.hpp"
#include "recoupling.hpp"
#include "batch.hpp"
#include "server.hpp"
#include "svd.hpp"
#include "tensor.hpp"

struct su2_ctx {
  su2::Context c;
  explicit su2_ctx(int dev) : c(dev) {}
};
struct su2_tensor {
  std::unique_ptr<su2::Tensor> own;
  su2::Tensor& t;
  explicit su2_tensor(su2::Tensor* p) : own(p), t(*p) {}
};
struct su2_dmrg {
  su2::Dmrg d;
  template <class... A>
  explicit su2_dmrg(A&&... a) : d(std::forward<A>(a)...) {}
};
struct su2_heff {
  std::unique_ptr<su2::HeffPlan> own;
  su2::HeffPlan& p;
  explicit su2_heff(su2::HeffPlan* q) : own(q), p(*q) {}
};
struct su2_batch {
  su2::BatchDriver b;
  explicit su2_batch(su2::Context& c) : b(c) {}
};
struct su2_server {
  su2::Server s;
  template <class... A>
  explicit su2_server(A&&... a) : s(std::forward<A>(a)...) {}
};

namespace {

thread_local std::string g_err;

template <class F>
int guard(su2_ctx* ctx, F&& f) {
  try {
    f();
    return SU2_OK;
  } catch (const su2::Error& e) {
    if (ctx) ctx->c.last_error = e.what();
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->c.last_error = "host allocation failed";
    g_err = "host allocation failed";
    return SU2_ERR_CAPACITY;
  } catch (const std::exception& e) {
    if (ctx) ctx->c.last_error = e.what();
    g_err = e.what();
    return SU2_ERR_ARGUMENT;
  }
}

void need(const void* p, const char* what) {
  SU2_CHECK(p != nullptr, su2::kErrArgument, "%s is NULL", what);
}

}  // namespace

extern "C" {

const char* su2_version(void) { return "su2dmrg-mi355x 0.1"; }

int su2_ctx_create(int device, su2_ctx** out) {
  return guard(nullptr, [&] {
    need(out, "out");
    *out = new su2_ctx(device);
  });
}

int su2_ctx_destroy(su2_ctx* ctx) {
  if (!ctx) return SU2_OK;
  // wait for the stream here, where a device error can be reported; the
  // context is freed either way (its destructor does not throw)
  const int rc = guard(nullptr, [&] {
    ctx->c.make_current();
    SU2_HIP(hipStreamSynchronize(ctx->c.stream()));
  });
  delete ctx;
  return rc;
}

const char* su2_last_error(const su2_ctx* ctx) {
  return ctx ? ctx->c.last_error.c_str() : g_err.c_str();
}

int su2_ctx_synchronize(su2_ctx* ctx) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    ctx->c.make_current();
    SU2_HIP(hipStreamSynchronize(ctx->c.stream()));
  });
}

int su2_ctx_memory(su2_ctx* ctx, long long* live, long long* peak, long long* cached,
                   int reset_peak) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    su2::DevicePool& p = *ctx->c.pool();
    if (live) *live = (long long)p.live_bytes();
    if (peak) *peak = (long long)p.peak_bytes();
    if (cached) *cached = (long long)p.cached_bytes();
    if (reset_peak) p.reset_peak();
  });
}

int su2_ctx_set_svd_tuning(su2_ctx* ctx, int block_cols, int inner_sweeps) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    SU2_CHECK(block_cols == 16 || block_cols == 32, su2::kErrArgument,
              "svd block width %d (16 or 32)", block_cols);
    SU2_CHECK(inner_sweeps >= 1 && inner_sweeps <= 3, su2::kErrArgument,
              "svd inner sweeps %d (1..3)", inner_sweeps);
    ctx->c.svd_block_cols = block_cols;
    ctx->c.svd_inner_sweeps = inner_sweeps;
  });
}

int su2_ctx_set_svd_precondition(su2_ctx* ctx, int min_k) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    SU2_CHECK(min_k == 0 || (min_k >= 2 && min_k <= su2::kPrecondMaxK), su2::kErrArgument,
              "svd preconditioning from block size %d", min_k);
    ctx->c.svd_precondition_min_k = min_k;
  });
}

int su2_tensor_create(su2_ctx* ctx, int nleg, const int* nsec, const int* twoj, const int* dims,
                      const int* dirs, int locspin, const int* present, su2_tensor** out) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(nsec, "nsec");
    need(twoj, "twoj");
    need(dims, "dims");
    need(dirs, "dirs");
    need(out, "out");
    SU2_CHECK(nleg >= 1 && nleg <= su2::kMaxLegs, su2::kErrArgument, "nleg=%d", nleg);
    std::vector<su2::Leg> legs;
    int k = 0;
    long long nten = 1;
    for (int i = 0; i < nleg; ++i) {
      SU2_CHECK(nsec[i] >= 1 && nsec[i] <= 64, su2::kErrArgument, "leg %d: %d sectors", i, nsec[i]);
      std::vector<int> tj(twoj + k, twoj + k + nsec[i]), dm(dims + k, dims + k + nsec[i]);
      for (size_t s = 1; s < tj.size(); ++s)
        SU2_CHECK(tj[s] > tj[s - 1], su2::kErrArgument, "leg %d: 2j not strictly ascending", i);
      legs.push_back(su2::make_leg(tj, dm, dirs[i]));
      k += nsec[i];
      nten *= nsec[i];
    }
    std::vector<char> pres;
    if (present) {
      pres.resize(nten);
      for (long long i = 0; i < nten; ++i) pres[i] = present[i] != 0;
    }
    *out = new su2_tensor(new su2::Tensor(&ctx->c, legs, locspin, present ? &pres : nullptr));
  });
}

int su2_tensor_like(su2_ctx* ctx, const su2_tensor* src, su2_tensor** out) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(src, "src");
    need(out, "out");
    std::vector<char> present(src->t.nten(), 0);
    for (auto& b : src->t.blocks()) present[b.index] = 1;
    *out = new su2_tensor(new su2::Tensor(&ctx->c, src->t.legs(), src->t.locspin(), &present));
  });
}

int su2_tensor_destroy(su2_ctx* ctx, su2_tensor* t) {
  return guard(ctx, [&] {
    if (ctx) ctx->c.make_current();
    delete t;
  });
}

int su2_tensor_size(const su2_tensor* t, long long* ndouble, int* nblock) {
  return guard(nullptr, [&] {
    need(t, "tensor");
    if (ndouble) *ndouble = t->t.size();
    if (nblock) *nblock = (int)t->t.blocks().size();
  });
}

int su2_tensor_block(const su2_tensor* t, int i, int* index, int* twoj, int* dims,
                     long long* offset) {
  return guard(nullptr, [&] {
    need(t, "tensor");
    SU2_CHECK(i >= 0 && i < (int)t->t.blocks().size(), su2::kErrArgument, "block %d", i);
    const su2::Block& b = t->t.blocks()[i];
    if (index) *index = b.index;
    for (int l = 0; l < t->t.nleg(); ++l) {
      if (twoj) twoj[l] = b.twoj[l];
      if (dims) dims[l] = b.dim[l];
    }
    if (offset) *offset = b.offset;
  });
}

int su2_tensor_upload(su2_ctx* ctx, su2_tensor* t, const double* host, long long n) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(t, "tensor");
    SU2_CHECK(n == t->t.size(), su2::kErrArgument, "upload: %lld doubles for a tensor of %lld", n,
              t->t.size());
    if (n == 0) return;
    need(host, "host");
    ctx->c.make_current();
    SU2_HIP(hipMemcpyAsync(t->t.data(), host, n * sizeof(double), hipMemcpyHostToDevice,
                           ctx->c.stream()));
    SU2_HIP(hipStreamSynchronize(ctx->c.stream()));
  });
}

int su2_tensor_download(su2_ctx* ctx, const su2_tensor* t, double* host, long long n) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(t, "tensor");
    SU2_CHECK(n == t->t.size(), su2::kErrArgument, "download: %lld doubles for a tensor of %lld",
              n, t->t.size());
    if (n == 0) return;
    need(host, "host");
    ctx->c.make_current();
    SU2_HIP(hipMemcpyAsync(host, t->t.data(), n * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->c.stream()));
    SU2_HIP(hipStreamSynchronize(ctx->c.stream()));
  });
}

static void same_layout(const su2::Tensor& a, const su2::Tensor& b, const char* op) {
  SU2_CHECK(a.same_layout(b), su2::kErrStructure, "%s: tensors have different SU(2) structure",
            op);
}

int su2_tensor_copy(su2_ctx* ctx, su2_tensor* dst, const su2_tensor* src) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(dst, "dst");
    need(src, "src");
    same_layout(dst->t, src->t, "copy");
    ctx->c.make_current();
    su2::dev_copy(ctx->c, dst->t.data(), src->t.data(), src->t.size());
  });
}

int su2_tensor_axpy(su2_ctx* ctx, su2_tensor* y, double a, const su2_tensor* x) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(y, "y");
    need(x, "x");
    same_layout(y->t, x->t, "axpy");
    ctx->c.make_current();
    su2::dev_axpy(ctx->c, y->t.data(), x->t.data(), x->t.size(), a);
  });
}

int su2_tensor_scal(su2_ctx* ctx, su2_tensor* x, double a) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(x, "x");
    ctx->c.make_current();
    su2::dev_scal(ctx->c, x->t.data(), x->t.size(), a);
  });
}

int su2_tensor_dot(su2_ctx* ctx, const su2_tensor* x, const su2_tensor* y, double* out) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(x, "x");
    need(y, "y");
    need(out, "out");
    same_layout(x->t, y->t, "dot");
    ctx->c.make_current();
    *out = su2::dev_dot(ctx->c, x->t.data(), y->t.data(), x->t.size());
  });
}

int su2_tensor_dot_noalloc(su2_ctx* ctx, const su2_tensor* x, const su2_tensor* y, double* out) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(x, "x");
    need(y, "y");
    need(out, "out");
    same_layout(x->t, y->t, "dot");
    ctx->c.make_current();
    *out = su2::dev_dot_noalloc(ctx->c, x->t.data(), y->t.data(), x->t.size());
  });
}

int su2_noalloc_scratch_bytes(long long* bytes) {
  return guard(nullptr, [&] {
    need(bytes, "bytes");
    *bytes = su2::noalloc_scratch_bytes();
  });
}

int su2_tensor_normalize(su2_ctx* ctx, su2_tensor* x, double* norm) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(x, "x");
    ctx->c.make_current();
    double n = std::sqrt(su2::dev_dot(ctx->c, x->t.data(), x->t.data(), x->t.size()));
    SU2_CHECK(n > 0, su2::kErrNumeric, "normalize: zero vector");
    su2::dev_scal(ctx->c, x->t.data(), x->t.size(), 1.0 / n);
    if (norm) *norm = n;
  });
}

int su2_heff_create(su2_ctx* ctx, const su2_tensor* psi_layout, const su2_tensor* hl,
                    const su2_tensor* hr, int nterm, const int* term_rank,
                    const int* left_count, const su2_tensor* const* left_ops,
                    const double* left_coef, const int* right_count,
                    const su2_tensor* const* right_ops, const double* right_coef,
                    su2_heff** out) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(psi_layout, "psi_layout");
    need(out, "out");
    SU2_CHECK(nterm >= 0, su2::kErrArgument, "nterm=%d", nterm);
    std::vector<su2::HeffTerm> terms(nterm);
    if (nterm > 0) {
      need(left_count, "left_count");
      need(right_count, "right_count");
      need(left_ops, "left_ops");
      need(right_ops, "right_ops");
      need(left_coef, "left_coef");
      need(right_coef, "right_coef");
    }
    int li = 0, ri = 0;
    for (int t = 0; t < nterm; ++t) {
      SU2_CHECK(left_count[t] >= 1 && right_count[t] >= 1, su2::kErrArgument,
                "term %d: empty side", t);
      terms[t].k = term_rank ? term_rank[t] : 2;
      for (int i = 0; i < left_count[t]; ++i, ++li) {
        need(left_ops[li], "left operator");
        terms[t].left.push_back({left_coef[li], &left_ops[li]->t});
      }
      for (int i = 0; i < right_count[t]; ++i, ++ri) {
        need(right_ops[ri], "right operator");
        terms[t].right.push_back({right_coef[ri], &right_ops[ri]->t});
      }
    }
    *out = new su2_heff(new su2::HeffPlan(ctx->c, psi_layout->t, hl ? &hl->t : nullptr,
                                          hr ? &hr->t : nullptr, terms));
  });
}

int su2_heff_apply(su2_ctx* ctx, su2_heff* plan, const su2_tensor* psi, su2_tensor* out) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(plan, "plan");
    need(psi, "psi");
    need(out, "out");
    SU2_CHECK(psi != out, su2::kErrArgument, "heff_apply: out aliases psi");
    plan->p.apply(psi->t, out->t);
  });
}

int su2_heff_flops(const su2_heff* plan, double* algorithmic, double* executed) {
  return guard(nullptr, [&] {
    need(plan, "plan");
    if (algorithmic) *algorithmic = plan->p.algorithmic_flops();
    if (executed) *executed = plan->p.executed_flops();
  });
}

int su2_heff_destroy(su2_ctx* ctx, su2_heff* plan) {
  return guard(ctx, [&] {
    if (ctx) ctx->c.make_current();
    delete plan;
  });
}

int su2_heff_time(su2_ctx* ctx, su2_heff* plan, const su2_tensor* psi, su2_tensor* out, int reps,
                  double* ms_apply, double* ms_stage1, double* ms_stage2) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(plan, "plan");
    need(psi, "psi");
    need(out, "out");
    SU2_CHECK(reps >= 1, su2::kErrArgument, "reps=%d", reps);
    ctx->c.make_current();
    hipStream_t s = ctx->c.stream();
    plan->p.apply(psi->t, out->t);
    hipEvent_t e0, e1, e2, e3;
    SU2_HIP(hipEventCreate(&e0));
    SU2_HIP(hipEventCreate(&e1));
    SU2_HIP(hipEventCreate(&e2));
    SU2_HIP(hipEventCreate(&e3));
    SU2_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) plan->p.apply(psi->t, out->t);
    SU2_HIP(hipEventRecord(e1, s));
    SU2_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SU2_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms_apply) *ms_apply = ms / reps;
    // the two stages alone
    float m1 = 0, m2 = 0;
    for (int r = 0; r < reps; ++r) {
      float a = 0, b = 0;
      SU2_HIP(hipEventRecord(e0, s));
      plan->p.stage(0).launch(s);
      SU2_HIP(hipEventRecord(e2, s));
      plan->p.stage(1).launch(s);
      SU2_HIP(hipEventRecord(e3, s));
      SU2_HIP(hipEventSynchronize(e3));
      SU2_HIP(hipEventElapsedTime(&a, e0, e2));
      SU2_HIP(hipEventElapsedTime(&b, e2, e3));
      m1 += a;
      m2 += b;
    }
    if (ms_stage1) *ms_stage1 = m1 / reps;
    if (ms_stage2) *ms_stage2 = m2 / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    (void)hipEventDestroy(e3);
  });
}

int su2_heff_autotune(su2_ctx* ctx, su2_heff* plan, int reps, int* chosen, float* times,
                      int* nvariants) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(plan, "plan");
    SU2_CHECK(reps >= 1, su2::kErrArgument, "reps=%d", reps);
    std::vector<float> t1, t2;
    int ch[2];
    plan->p.autotune(reps, ch, &t1, &t2);
    if (chosen) {
      chosen[0] = ch[0];
      chosen[1] = ch[1];
    }
    const int nv = su2::gemm_num_variants(plan->p.precision());
    if (nvariants) *nvariants = nv;
    if (times) {
      for (int v = 0; v < nv; ++v) {
        times[v] = t1[v];
        times[nv + v] = t2[v];
      }
    }
  });
}

int su2_heff_set_variants(su2_ctx* ctx, su2_heff* plan, int stage1, int stage2) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(plan, "plan");
    plan->p.set_variants(stage1, stage2);
  });
}

int su2_heff_set_precision(su2_ctx* ctx, su2_heff* plan, int bits) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(plan, "plan");
    SU2_CHECK(bits == 64 || bits == 32, su2::kErrArgument, "matvec precision %d bits", bits);
    plan->p.set_precision(bits == 64 ? su2::Prec::F64 : su2::Prec::F32);
  });
}

int su2_heff_set_kernel_path(su2_ctx* ctx, su2_heff* plan, int uniform, int* layouts) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(plan, "plan");
    plan->p.set_uniform(uniform != 0);
    if (layouts) {
      layouts[0] = plan->p.stage(0).layout();
      layouts[1] = plan->p.stage(1).layout();
    }
  });
}

int su2_env_transform(su2_ctx* ctx, int kind, int side, int nsec, const int* twoj,
                      const int* dims, const su2_tensor* P, const su2_tensor* op, double coef,
                      su2_tensor** out) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(out, "out");
    SU2_CHECK(side == 0 || side == 1, su2::kErrArgument, "env transform: side %d", side);
    SU2_CHECK(nsec >= 1 && twoj && dims, su2::kErrArgument, "env transform: block leg");
    su2::Leg block = su2::make_leg(std::vector<int>(twoj, twoj + nsec),
                                   std::vector<int>(dims, dims + nsec), 1);
    std::unique_ptr<su2::Tensor> t =
        su2::env_transform(ctx->c, kind, side == 0, block, P ? &P->t : nullptr,
                           op ? &op->t : nullptr, coef);
    *out = new su2_tensor(t.release());
  });
}

int su2_svd_truncate(su2_ctx* ctx, const su2_tensor* psi, double p1, double p2, int max_dcut,
                     su2_tensor** U, su2_tensor** V, double* kept, int cap, int* nkept,
                     double* discarded) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(psi, "psi");
    need(U, "U");
    need(V, "V");
    su2::SvdResult r = su2::truncated_svd(ctx->c, psi->t, p1, p2, max_dcut);
    int n = 0;
    for (auto& v : r.kept)
      for (double x : v) {
        if (kept && n < cap) kept[n] = x;
        ++n;
      }
    if (nkept) *nkept = n;
    if (discarded) *discarded = r.discarded_weight;
    *U = new su2_tensor(r.U.release());
    *V = new su2_tensor(r.V.release());
  });
}

int su2_recoupling(int kind, const int* labels, double* out) {
  return guard(nullptr, [&] {
    need(labels, "labels");
    need(out, "out");
    const int* l = labels;
    switch (kind) {
      case 0: *out = su2::rc_op_transform_left(l[0], l[1], l[2], l[3], l[4], l[5]); break;
      case 1: *out = su2::rc_op_transform_right(l[0], l[1], l[2], l[3], l[4], l[5]); break;
      case 2: *out = su2::rc_op_initial_left(l[0], l[1], l[2], l[3], l[4], l[5]); break;
      case 3: *out = su2::rc_op_initial_right(l[0], l[1], l[2], l[3], l[4], l[5]); break;
      case 4: *out = su2::rc_pairup_left(l[0], l[1], l[2], l[3], l[4], l[5]); break;
      case 5: *out = su2::rc_pairup_right(l[0], l[1], l[2], l[3], l[4], l[5]); break;
      case 6: *out = su2::rc_psi_left_to_right(l[0], l[1], l[2]); break;
      case 7: *out = su2::rc_psi_right_to_left(l[0], l[1], l[2]); break;
      case 8: *out = su2::heff_recoupling(l[0], l[1], l[2]); break;
      case 9: *out = su2::rc_couple_left(l[0], l[1], l[2], l[3], l[4], l[5]); break;
      case 10: *out = su2::rc_couple_right(l[0], l[1], l[2], l[3], l[4], l[5]); break;
      case 11: *out = su2::rc_reduce_left(l[0], l[1], l[2], l[3], l[4], l[5]); break;
      case 12: *out = su2::rc_reduce_right(l[0], l[1], l[2], l[3], l[4], l[5]); break;
      default: su2::fail(su2::kErrArgument, "recoupling kind %d", kind);
    }
  });
}

int su2_dmrg_create(su2_ctx* ctx, int lx, int ly, double jcoup, double qcoup, int bdry,
                    int totspin, int max_dcut, su2_dmrg** out) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(out, "out");
    ctx->c.make_current();
    *out = new su2_dmrg(ctx->c, lx, ly, jcoup, qcoup, bdry, totspin, max_dcut);
  });
}

int su2_dmrg_destroy(su2_ctx* ctx, su2_dmrg* d) {
  return guard(ctx, [&] {
    if (ctx) ctx->c.make_current();
    delete d;
  });
}

int su2_dmrg_warmup(su2_ctx* ctx, su2_dmrg* d, double* energy) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(d, "dmrg");
    const double e = d->d.warmup();
    if (energy) *energy = e;
  });
}

int su2_dmrg_sweep(su2_ctx* ctx, su2_dmrg* d, double* energy) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(d, "dmrg");
    const double e = d->d.sweep();
    if (energy) *energy = e;
  });
}

int su2_dmrg_set_max_dcut(su2_ctx* ctx, su2_dmrg* d, int max_dcut) {
  return guard(ctx, [&] {
    need(d, "dmrg");
    SU2_CHECK(max_dcut >= 1, su2::kErrArgument, "max_dcut=%d", max_dcut);
    d->d.set_max_dcut(max_dcut);
  });
}

int su2_dmrg_set_heff_variants(su2_ctx* ctx, su2_dmrg* d, int stage1, int stage2) {
  return guard(ctx, [&] {
    need(d, "dmrg");
    d->d.set_heff_variants(stage1, stage2);
  });
}

int su2_dmrg_save(su2_ctx* ctx, su2_dmrg* d, const char* dir, int ky, int exci) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(d, "dmrg");
    need(dir, "dir");
    d->d.save(dir, ky, exci);
  });
}

int su2_dmrg_load(su2_ctx* ctx, su2_dmrg* d, const char* dir, int file_dcut, int ky, int exci,
                  double* energy) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(d, "dmrg");
    need(dir, "dir");
    const double e = d->d.load(dir, file_dcut, ky, exci);
    if (energy) *energy = e;
  });
}

int su2_dmrg_add_lower_state(su2_ctx* ctx, su2_dmrg* d, const char* dir, int file_dcut, int ky,
                             int exci) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(d, "dmrg");
    need(dir, "dir");
    d->d.add_lower_state(dir, file_dcut, ky, exci);
  });
}

int su2_dmrg_set_verbose(su2_dmrg* d, int level) {
  return guard(nullptr, [&] {
    need(d, "dmrg");
    d->d.set_verbose(level);
  });
}

int su2_dmrg_energies(const su2_dmrg* d, double* out, int cap, int* n) {
  return guard(nullptr, [&] {
    need(d, "dmrg");
    const auto& e = d->d.step_energies();
    for (int i = 0; i < (int)e.size() && i < cap; ++i)
      if (out) out[i] = e[i];
    if (n) *n = (int)e.size();
  });
}

int su2_dmrg_history(const su2_dmrg* d, int* kept, double* discarded, int* matvecs, int cap,
                     int* n) {
  return guard(nullptr, [&] {
    need(d, "dmrg");
    const auto& h = d->d.history();
    for (int i = 0; i < (int)h.size() && i < cap; ++i) {
      if (kept) kept[i] = h[i].kept;
      if (discarded) discarded[i] = h[i].discarded;
      if (matvecs) matvecs[i] = h[i].matvecs;
    }
    if (n) *n = (int)h.size();
  });
}

int su2_dmrg_set_matvec_precision(su2_ctx* ctx, su2_dmrg* d, int bits, double min_flops) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(d, "dmrg");
    SU2_CHECK(bits == 64 || bits == 32, su2::kErrArgument, "matvec precision %d bits", bits);
    d->d.set_matvec_precision(bits == 64 ? su2::Prec::F64 : su2::Prec::F32, min_flops);
  });
}

int su2_batch_create(su2_ctx* ctx, su2_batch** out) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(out, "out");
    *out = new su2_batch(ctx->c);
  });
}

int su2_batch_destroy(su2_ctx* ctx, su2_batch* b) {
  return guard(ctx, [&] { delete b; });
}

int su2_batch_add(su2_ctx* ctx, su2_batch* b, su2_dmrg* d, int nsweep, const int* max_dcut,
                  const int* bits, double min_fp32_flops, double stop, int* id) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(b, "batch");
    need(d, "dmrg");
    SU2_CHECK(nsweep >= 0, su2::kErrArgument, "batch: %d sweeps", nsweep);
    std::vector<su2::RunPhase> prog;
    su2::RunPhase w;
    w.warmup = true;
    w.max_dcut = nsweep > 0 && max_dcut ? max_dcut[0] : 0;
    w.prec = nsweep > 0 && bits && bits[0] == 32 ? su2::Prec::F32 : su2::Prec::F64;
    w.min_fp32_flops = min_fp32_flops;
    prog.push_back(w);
    for (int k = 0; k < nsweep; ++k) {
      su2::RunPhase p;
      p.max_dcut = max_dcut ? max_dcut[k] : 0;
      SU2_CHECK(!bits || bits[k] == 64 || bits[k] == 32, su2::kErrArgument,
                "batch: matvec precision %d bits", bits ? bits[k] : 0);
      p.prec = bits && bits[k] == 32 ? su2::Prec::F32 : su2::Prec::F64;
      p.min_fp32_flops = min_fp32_flops;
      prog.push_back(p);
    }
    const int i = b->b.add(&d->d, prog, stop);
    if (id) *id = i;
  });
}

int su2_batch_tick(su2_ctx* ctx, su2_batch* b, int* stepped) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(b, "batch");
    const int n = b->b.tick();
    if (stepped) *stepped = n;
  });
}

int su2_batch_run(su2_ctx* ctx, su2_batch* b) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(b, "batch");
    while (b->b.tick() > 0) {
    }
  });
}

int su2_batch_run_lanes(su2_ctx* const* ctxs, su2_batch* const* bs, int n) {
  std::vector<su2::BatchDriver*> drivers;
  const int rc = guard(nullptr, [&] {
    SU2_CHECK(n >= 1, su2::kErrArgument, "lanes: %d batches", n);
    need(ctxs, "ctxs");
    need(bs, "batches");
    for (int i = 0; i < n; ++i) {
      need(ctxs[i], "ctx");
      need(bs[i], "batch");
      SU2_CHECK(&bs[i]->b.context() == &ctxs[i]->c, su2::kErrArgument,
                "lanes: batch %d is not on context %d", i, i);
      drivers.push_back(&bs[i]->b);
    }
  });
  if (rc != SU2_OK) return rc;
  std::vector<std::exception_ptr> err;
  const int rc2 = guard(nullptr, [&] { err = su2::run_batches_parallel(drivers); });
  if (rc2 != SU2_OK) return rc2;
  // each failed lane's message goes to its context; the first failure's
  // status is returned (and its message is the thread's global one)
  int first = SU2_OK;
  for (int i = 0; i < n; ++i) {
    if (!err[i]) continue;
    const int r = guard(ctxs[i], [&] { std::rethrow_exception(err[i]); });
    if (first == SU2_OK) first = r;
  }
  if (first != SU2_OK) {
    for (int i = 0; i < n; ++i)
      if (err[i]) {
        g_err = ctxs[i]->c.last_error;
        break;
      }
  }
  return first;
}

int su2_batch_set_graph(su2_ctx* ctx, su2_batch* b, int on) {
  return guard(ctx, [&] {
    need(b, "batch");
    b->b.set_use_graph(on != 0);
  });
}

int su2_batch_energies(const su2_batch* b, int id, int* finished, double* out, int cap, int* n) {
  return guard(nullptr, [&] {
    need(b, "batch");
    SU2_CHECK(id >= 0, su2::kErrArgument, "batch: run %d", id);
    const std::vector<double>& e = b->b.phase_energies(id);
    if (finished) *finished = b->b.finished(id) ? 1 : 0;
    for (int i = 0; i < (int)e.size() && i < cap && out; ++i) out[i] = e[i];
    if (n) *n = (int)e.size();
  });
}

int su2_batch_stats(const su2_batch* b, double* out, int cap) {
  return guard(nullptr, [&] {
    need(b, "batch");
    need(out, "out");
    const su2::BatchDriver& d = b->b;
    const double v[8] = {(double)d.ticks(),   (double)d.graphs_built(), (double)d.graph_replays(),
                         d.t_prepare,         d.t_lanczos,              d.t_matvec,
                         d.t_svd,             d.t_complete};
    for (int i = 0; i < 8 && i < cap; ++i) out[i] = v[i];
  });
}

int su2_dmrg_centre(su2_ctx* ctx, su2_dmrg* d, su2_heff** plan, su2_tensor** psi) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(d, "dmrg");
    need(plan, "plan");
    need(psi, "psi");
    auto r = d->d.centre();
    *psi = new su2_tensor(r.second.release());
    *plan = new su2_heff(r.first.release());
  });
}

int su2_dmrg_set_graph(su2_ctx* ctx, su2_dmrg* d, int on) {
  return guard(ctx, [&] {
    need(d, "dmrg");
    d->d.set_use_graph(on != 0);
  });
}

int su2_server_create(su2_ctx* ctx, double budget_bytes, int max_slots, const char* workdir,
                      su2_server** out) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(out, "out");
    need(workdir, "workdir");
    *out = new su2_server(ctx->c, budget_bytes, max_slots, std::string(workdir));
  });
}

int su2_server_destroy(su2_ctx* ctx, su2_server* s) {
  return guard(ctx, [&] { delete s; });
}

int su2_server_submit(su2_ctx* ctx, su2_server* s, const su2_job* job, double arrival_s,
                      int* id) {
  return guard(ctx, [&] {
    need(s, "server");
    need(job, "job");
    su2::JobSpec j;
    j.lx = job->lx;
    j.ly = job->ly;
    j.jcoup = job->jcoup;
    j.qcoup = job->qcoup;
    j.bdry = job->bdry;
    j.totspin = job->totspin;
    j.max_dcut = job->max_dcut;
    j.exci = job->exci;
    j.nsweep = job->nsweep;
    j.stop = job->stop;
    const int i = s->s.submit(j, arrival_s);
    if (id) *id = i;
  });
}

int su2_server_set_lanes(su2_ctx* ctx, su2_server* s, int lanes) {
  return guard(ctx, [&] {
    need(s, "server");
    s->s.set_lanes(lanes);
  });
}

int su2_server_run(su2_ctx* ctx, su2_server* s, int verbose) {
  return guard(ctx, [&] {
    need(ctx, "ctx");
    need(s, "server");
    s->s.set_verbose(verbose);
    s->s.run();
  });
}

int su2_server_result(const su2_server* s, int id, double* energy, double* times,
                      double* footprint) {
  return guard(nullptr, [&] {
    need(s, "server");
    SU2_CHECK(id >= 0 && id < s->s.njobs(), su2::kErrArgument, "server: job %d", id);
    const su2::JobResult& r = s->s.result(id);
    SU2_CHECK(r.done, su2::kErrArgument, "server: job %d has not finished", id);
    if (energy) *energy = r.energy;
    if (times) {
      times[0] = r.t_arrive;
      times[1] = r.t_admit;
      times[2] = r.t_finish;
    }
    if (footprint) *footprint = r.footprint;
  });
}

int su2_server_stats(const su2_server* s, double* out, int cap) {
  return guard(nullptr, [&] {
    need(s, "server");
    need(out, "out");
    const double v[3] = {s->s.wall(), (double)s->s.ticks(), (double)s->s.max_concurrent()};
    for (int i = 0; i < 3 && i < cap; ++i) out[i] = v[i];
  });
}

int su2_job_footprint(const su2_job* job, double* bytes) {
  return guard(nullptr, [&] {
    need(job, "job");
    need(bytes, "bytes");
    su2::JobSpec j;
    j.lx = job->lx;
    j.ly = job->ly;
    j.qcoup = job->qcoup;
    j.max_dcut = job->max_dcut;
    j.exci = job->exci;
    *bytes = su2::job_footprint(j);
  });
}

int su2_dmrg_stats(const su2_dmrg* d, double* out, int cap) {
  return guard(nullptr, [&] {
    need(d, "dmrg");
    need(out, "out");
    const su2::DmrgStats& s = d->d.stats();
    const double v[17] = {s.t_env,      s.t_lanczos,         s.t_svd,
                          s.t_project,  s.t_start,           s.t_matvec,
                          s.heff_flops, (double)s.matvecs,   (double)s.steps,
                          (double)d->d.last_bond_dim(), s.t_jacobi,
                          (double)s.jacobi_sweeps, (double)s.matvecs32, s.t_matvec32,
                          s.heff_flops32, s.max_fp64_correction, (double)s.svd_degraded};
    for (int i = 0; i < 17 && i < cap; ++i) out[i] = v[i];
  });
}

}  // extern "C"
