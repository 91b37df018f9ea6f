// Small device building blocks shared by the SVD, environment and driver code.
#pragma once

#include <map>
#include <tuple>
#include <vector>

#include "blas1.hpp"
#include "gemm.hpp"

namespace su2 {

class Context;

// A batch of independent GEMMs launched as one grouped FP64 MFMA launch.
//   C = alpha op(A) op(B)        (C overwritten)
class GemmBatch {
 public:
  void add(double* C, long long ldc, int M, int N, const double* A, long long lda, int tA,
           const double* B, long long ldb, int tB, int K, double alpha = 1.0);
  // launch on the context stream (zero-fills problems with K == 0)
  void run(Context& c, int variant = -1);
  bool empty() const { return n_ == 0; }

 private:
  GroupedGemm g_;
  std::vector<std::pair<double*, std::pair<long long, std::pair<int, int>>>> zero_;
  int n_ = 0;
};

// Strided scale-adds (StridedAxpy) collected with a destination key; entries
// that hit the same key go to successive launches ("waves"), so every launch
// writes disjoint elements and the sum order is the order of add().
class ScatterBatch {
 public:
  void add(const void* key, const StridedAxpy& d);
  void run(Context& c);

 private:
  std::map<const void*, int> count_;
  std::vector<std::vector<StridedAxpy>> waves_;
};

// x[:, j] *= f[j] for the n columns of an m x n column-major matrix (ld m).
void dev_scale_columns(Context& c, double* x, int m, int n, const double* d_f);

}  // namespace su2
