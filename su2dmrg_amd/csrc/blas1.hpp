// Deterministic block-vector BLAS-1 on the device.
//
// Replaces the serial per-block loops behind tensor_su2's Lanczos algebra
// (operator=, +=, -=, *=, /=, inner_prod, normalize_vector:
// tensor_su2.cpp:96-259, 324-356, 663-684).  Reductions use a fixed grid and
// a fixed tree order, so a given vector length always sums in the same order.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace su2 {

class Context;

void dev_fill(Context& c, double* x, long long n, double v);
void dev_copy(Context& c, double* y, const double* x, long long n);
void dev_scal(Context& c, double* x, long long n, double a);
void dev_axpy(Context& c, double* y, const double* x, long long n, double a);       // y += a x
void dev_axpby(Context& c, double* y, const double* x, long long n, double a, double b);  // y = a x + b y
// sum_i x_i y_i (synchronises the context stream and returns on the host)
double dev_dot(Context& c, const double* x, const double* y, long long n);

// Batched strided scale-add: for each descriptor d,
//   dst_d[r*rs + c*cs] += alpha_d * src_d[r + c*ld],  r < rows_d, c < cols_d
// (src_d == nullptr adds alpha_d itself).  Descriptors of one launch must
// write disjoint elements.
struct StridedAxpy {
  double* dst;
  const double* src;
  long long rs, cs, ld;
  int rows, cols;   // rows * cols < 2^31
  double alpha;
};
// One launch per inner vector, in order on the context stream (a later
// launch may add into what an earlier one wrote); the descriptors of all
// launches are uploaded together and the call does not wait.  Every descriptor
// is cut into chunks of up to 4096 elements, one workgroup each, so a
// workgroup reads its descriptor once and indexes with 32-bit arithmetic.
void dev_strided_axpy_launches(Context& c, const std::vector<std::vector<StridedAxpy>>& launches);
void dev_strided_axpy(Context& c, const std::vector<StridedAxpy>& desc);

}  // namespace su2
