// H_eff matvec plan: the block-sparse two-site effective Hamiltonian applied
// to a two-leg SU(2) wavefunction, as dmrg_su2::hamiltonian_vector_
// multiplication_idmrg computes it (dmrg_su2.cpp:543-664):
//
//   psi'(j') = H_L(j')^T psi(j') + psi(j') H_R(j')
//            + sum_t sum_j f(j,j') L_t(j->j')^T psi(j) R_t(j->j')
//
// with f = the recoupling scalar of tensor_su2::hamiltonian_vector_
// multiplication (tensor_su2_dmrg_src.cpp:893-1041).  Each term side may be a
// coupling-weighted sum of operators (the reference re-sums the right
// operators on every call, dmrg_su2.cpp:585-599; here the sum is formed once
// when the plan is built).
//
// Execution: the T terms are packed into per-transition bundles
//   Lb(j->j') = [D_a(j) x D_a(j')*T]   (columns a' fastest, then t)
//   Rb(j->j') = [T*D_b(j) x D_b(j')]   (rows t fastest, then b)
// so that stage 1 is one GEMM per transition, X = f * Lb^T psi(j), and stage 2
// is one K-concatenated GEMM per output sector that also absorbs the two
// block-Hamiltonian products.  Both stages are single grouped launches.
#pragma once

#include <memory>
#include <utility>
#include <vector>

#include "gemm.hpp"
#include "tensor.hpp"

namespace su2 {

// One term sum_a c_a L_a (x) sum_b c_b R_b of H_eff; every operator has the
// operator leg 2j = k (2: spin operators; 0 and 4: the plaquette terms'
// rank-0 and rank-2 parts, contracted with fac_hamilt_vec[j][k][j'] like
// the spin operators, dmrg_su2.cpp:635-655).
struct HeffTerm {
  std::vector<std::pair<double, const Tensor*>> left, right;
  int k = 2;
};

class HeffPlan {
 public:
  HeffPlan(Context& c, const Tensor& psi_layout, const Tensor* hl, const Tensor* hr,
           const std::vector<HeffTerm>& terms);
  ~HeffPlan();
  HeffPlan(const HeffPlan&) = delete;
  HeffPlan& operator=(const HeffPlan&) = delete;

  // out = H psi; psi and out must have the layout the plan was built for.
  void apply(const Tensor& psi, Tensor& out);
  // Same, on raw device buffers of that layout (used by the Lanczos driver).
  void apply_raw(const double* psi, double* out);

  // Arithmetic of apply: FP64 (default), or FP32 MFMA with FP32 accumulation
  // on FP32 copies of the bundles and block Hamiltonians (made on first use),
  // the input rounded to FP32 and the result widened to FP64.
  void set_precision(Prec p);
  Prec precision() const { return prec_; }

  double algorithmic_flops() const { return alg_flops_; }
  double executed_flops() const { return stage1_.flops() + stage2_.flops(); }
  long long psi_size() const { return psi_size_; }
  // the two grouped launches of the current precision
  GroupedGemm& stage(int i);
  const GroupedGemm& stage1() const { return stage1_; }
  const GroupedGemm& stage2() const { return stage2_; }
  // Tile-variant choice of the two grouped launches of the current
  // precision: set explicitly, or time every compiled variant on this plan's
  // own problems and keep the fastest.
  void set_variants(int v1, int v2);
  // Straight-line GEMM kernel for the single-layout launches (default) or
  // the general kernel; both give the same bits.
  void set_uniform(bool on);
  void autotune(int reps, int* chosen, std::vector<float>* t1, std::vector<float>* t2);

 private:
  friend class HeffBatch;   // concatenates the launches of several plans
  // a launch's problems over the plan's buffers, offsets in elements
  enum { kBufWork = 0, kBufIn = 1, kBufOut = 2 };
  struct SliceSpec {
    int abuf;
    long long aoff, lda;
    int ta, bbuf;
    long long boff, ldb;
    int K;
    double alpha;
  };
  struct GemmSpec {
    int cbuf;
    long long coff, ldc;
    int M, N;
    std::vector<SliceSpec> slices;
  };
  static void instantiate(GroupedGemm& g, const std::vector<GemmSpec>& spec, void* const* bases,
                          size_t es);
  void ensure_fp32();

  Context& ctx_;
  std::shared_ptr<DevicePool> pool_;   // frees the buffers even after ctx_ is gone
  long long psi_size_ = 0;
  long long work_size_ = 0;
  double* d_in_ = nullptr;    // plan-owned copy of the input vector
  double* d_out_ = nullptr;   // plan-owned output vector
  double* d_work_ = nullptr;  // bundles + stage-1 intermediates
  float* d_work32_ = nullptr; // FP32 mirrors (mixed precision)
  float* d_in32_ = nullptr;
  float* d_out32_ = nullptr;
  std::vector<std::pair<long long, long long>> zero_blocks_;  // (offset, size) without terms
  std::vector<GemmSpec> spec1_, spec2_;
  GroupedGemm stage1_, stage2_;
  std::unique_ptr<GroupedGemm> stage1_32_, stage2_32_;
  int var32_[2] = {-1, -1};
  Prec prec_ = Prec::F64;
  double alg_flops_ = 0;
};

}  // namespace su2
