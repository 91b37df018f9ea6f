#include "ops.hpp"

#include "common.hpp"
#include "context.hpp"

namespace su2 {

namespace {

__global__ void scale_columns_kernel(double* __restrict__ x, int m, int n,
                                     const double* __restrict__ f) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)m * n) return;
  x[e] *= f[e / m];
}

}  // namespace

void GemmBatch::add(double* C, long long ldc, int M, int N, const double* A, long long lda, int tA,
                    const double* B, long long ldb, int tB, int K, double alpha) {
  if (M <= 0 || N <= 0) return;
  if (K <= 0) {
    zero_.push_back({C, {ldc, {M, N}}});
    return;
  }
  int p = g_.add_problem(C, ldc, M, N);
  g_.add_slice(p, A, lda, tA, B, ldb, tB, K, alpha);
  ++n_;
}

void GemmBatch::run(Context& c, int variant) {
  for (auto& z : zero_) {
    // a K == 0 product is a zero block; columns are contiguous of length M
    const long long ldc = z.second.first;
    const int M = z.second.second.first, N = z.second.second.second;
    if (ldc == M) dev_fill(c, z.first, (long long)M * N, 0.0);
    else
      for (int j = 0; j < N; ++j) dev_fill(c, z.first + (long long)j * ldc, M, 0.0);
  }
  if (n_ == 0) return;
  g_.finalize(c, 256, 64, variant);
  g_.launch(c.stream());
}

void ScatterBatch::add(const void* key, const StridedAxpy& d) {
  if (d.rows <= 0 || d.cols <= 0) return;
  int w = count_[key]++;
  if ((int)waves_.size() <= w) waves_.resize(w + 1);
  waves_[w].push_back(d);
}

void ScatterBatch::run(Context& c) { dev_strided_axpy_launches(c, waves_); }

void dev_scale_columns(Context& c, double* x, int m, int n, const double* d_f) {
  const long long tot = (long long)m * n;
  if (tot <= 0) return;
  hipLaunchKernelGGL(scale_columns_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     c.stream(), x, m, n, d_f);
  SU2_HIP(hipGetLastError());
}

}  // namespace su2
