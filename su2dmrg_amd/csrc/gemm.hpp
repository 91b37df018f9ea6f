// Grouped GEMM on CDNA4 matrix cores: FP64 (v_mfma_f64_16x16x4_f64) or FP32
// with FP32 accumulation (v_mfma_f32_16x16x4_f32).
//
// One launch runs a list of independent column-major problems
//   C_p[M_p x N_p] = sum_s alpha_s * op(A_s)[M_p x K_s] * op(B_s)[K_s x N_p]
// where each problem owns a contiguous list of K-slices (the "sum over
// terms and sectors" of the H_eff matvec becomes one long K loop).  A problem
// may be split along its concatenated K range; split partials go to a
// workspace and a second kernel adds them in a fixed order, so results are
// deterministic (no atomics).  The recoupling scalar alpha_s is applied to A
// as it is staged through LDS.  Every operand and C of one plan has the
// plan's element type (double or float).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

namespace su2 {

class Context;
class DevicePool;

// element type of a plan's operands and results
enum class Prec : int { F64 = 0, F32 = 1 };

struct GemmSlice {
  const void* A;
  const void* B;
  long long lda, ldb;
  int K;
  int transA, transB;
  int kt_begin;      // first k-tile index of this slice inside its problem
  double alpha;
};

struct GemmProblem {
  void* C;
  long long ldc;
  int M, N;
  int slice_begin, slice_end;
  int tiles_m, tiles_n;
  int n_fast;        // 1: consecutive tiles walk N (A panel reused), 0: walk M
  int ksplit;
  int kt_total;      // k-tiles over all slices (each slice padded to the tile)
  long long ws_offset;  // elements into the split-K workspace (ksplit > 1)
  long long unit_begin;
};

// Unit ranges of the 8 XCDs: XCD x runs units [begin[x], begin[x+1]).
struct XcdSplit {
  long long begin[9];
};

struct ReduceProblem {
  void* C;
  long long ldc;
  int M, N, ksplit;
  long long ws_offset;
  long long elem_begin;
};

// Number of compiled tile variants (BM x BN x BK x waves) of a precision and
// the default (the same index for both).
int gemm_num_variants(Prec p = Prec::F64);
constexpr int kDefaultGemmVariant = 1;

// Host-side plan: build problems, then finalize once; launch many times.
class GroupedGemm {
 public:
  explicit GroupedGemm(Prec p = Prec::F64) : prec_(p) {}
  ~GroupedGemm();
  GroupedGemm(const GroupedGemm&) = delete;
  GroupedGemm& operator=(const GroupedGemm&) = delete;

  // Adds a problem; returns its index.  Slices are added right after.  The
  // pointers are of the plan's element type.
  int add_problem(void* C, long long ldc, int M, int N);
  void add_slice(int problem, const void* A, long long lda, int transA, const void* B,
                 long long ldb, int transB, int K, double alpha);
  // Appends every problem and slice of another (unfinalized or finalized)
  // plan of the same precision: a launch over both lists gives each problem
  // the bits it has alone, since K splits are per problem.
  void append(const GroupedGemm& other);
  Prec prec() const { return prec_; }
  // Chooses the K split of every problem and uploads the tables for tile
  // variant `variant` (-1: default), in device memory of `c` (the plan then
  // launches on any stream but must not outlive c).  A problem's split
  // follows from the problem alone, so its summation order, and its bits, do
  // not depend on what else shares the launch: enough splits for about
  // min_units work units of its own, or one split per kt_per_unit k-tiles,
  // whichever is more, with at least 8 k-tiles per split.
  void finalize(Context& c, int min_units, int kt_per_unit, int variant = -1);
  // Re-tiles for another variant (device tables rebuilt).
  void set_variant(int v);
  int variant() const { return variant_; }
  // Launches whose slices share one layout (op(B) = B, op(A) fixed) run the
  // straight-line uniform kernel unless this is switched off (before
  // finalize or set_variant); layout() is 0 (NN), 1 (TN) or -1 (mixed).
  void allow_uniform(bool on) { uniform_ = on; }
  int layout() const { return layout_; }
  // Times every variant (HIP events, `reps` launches each) and keeps the
  // fastest; returns it.  `times` receives ms per launch by variant.
  int autotune(hipStream_t stream, int reps, std::vector<float>* times = nullptr);
  float time_ms(hipStream_t stream, int reps) const;
  void launch(hipStream_t stream) const;

  double flops() const { return flops_; }
  long long units() const { return total_units_; }
  size_t nproblems() const { return probs_.size(); }

 private:
  void release();
  std::vector<GemmProblem> probs_;      // as added
  std::vector<GemmSlice> slices_;
  std::vector<GemmProblem> run_probs_;  // tiled for the current variant
  std::vector<ReduceProblem> reds_;
  GemmProblem* d_probs_ = nullptr;
  GemmSlice* d_slices_ = nullptr;
  ReduceProblem* d_reds_ = nullptr;
  void* d_ws_ = nullptr;
  long long total_units_ = 0;
  XcdSplit split_{};
  long long grid_ = 0;   // 8 x the longest XCD range
  long long total_red_elems_ = 0;
  double flops_ = 0;
  int min_units_ = 256;
  int kt_per_unit_ = 64;
  std::shared_ptr<DevicePool> pool_;   // device memory of the finalizing context
  hipStream_t stream_ = nullptr;       // the finalizing context's stream (table uploads)
  int variant_ = -1;
  int layout_ = -1;
  bool uniform_ = true;
  Prec prec_ = Prec::F64;
};

}  // namespace su2
