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
// Many vector moves in one launch (descriptors in device memory, ndesc of
// them, the longest max_n elements): copies, FP64 -> FP32 rounding, FP32 ->
// FP64 widening, zero fills.  Used by the batched H_eff apply, whose launch
// sequence is captured in a HIP graph (a fill kernel, never a memset node).
enum VecMoveMode : int { kMoveF64 = 0, kMoveF64toF32 = 1, kMoveF32toF64 = 2, kZeroF64 = 3,
                         kZeroF32 = 4 };
struct VecMove {
  const void* src;
  void* dst;
  long long n;
  int mode, pad;
};
void dev_move_batch(Context& c, const VecMove* d_desc, int ndesc, long long max_n);
// x *= 1 / sqrt(s[0]), s in device memory (no wait)
void dev_scal_inv_sqrt(Context& c, double* x, long long n, const double* d_s);
// FP32 buffers of the mixed-precision matvec: fill, round to FP32, widen
void dev_fill_f32(Context& c, float* x, long long n, float v);
void dev_convert_f64_f32(Context& c, float* y, const double* x, long long n);
void dev_convert_f32_f64(Context& c, double* y, const float* x, long long n);
void dev_scal(Context& c, double* x, long long n, double a);
void dev_axpy(Context& c, double* y, const double* x, long long n, double a);       // y += a x
void dev_axpby(Context& c, double* y, const double* x, long long n, double a, double b);  // y = a x + b y
// sum_i x_i y_i (synchronises the context stream and returns on the host)
double dev_dot(Context& c, const double* x, const double* y, long long n);
// Allocation-free variant of dev_dot: the same kernels and summation order,
// with the partials in a fixed per-device array (noalloc_scratch_bytes bytes)
// instead of the context's scratch; no device or pinned allocation happens.
// noalloc_scratch_init(device) must have run once with that device current
// (Context's constructor does it).  Calls on one device share the array, so
// overlapping calls from several threads or streams race.
void noalloc_scratch_init(int device);
double dev_dot_noalloc(Context& c, const double* x, const double* y, long long n);
long long noalloc_scratch_bytes();
// the same sum, left on the device in *res (stream-ordered, no wait)
void dev_dot_to(Context& c, const double* x, const double* y, long long n, double* res);
// y += (sign * a[0]) x, the coefficient read on the device (no wait)
void dev_axpy_dev(Context& c, double* y, const double* x, long long n, const double* a,
                  double sign);

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

// Operator bundles: each descriptor holds nt summands of rows x cols
// column-major blocks, element q = r + rows*c of summand t written once as
//   blocked:     dst[q + rows*cols*t]
//   interleaved: dst[t + nt*q]
// with the value sum_{p < npos} alpha[t][p] * src[t][p][q], accumulated in p
// order from zero (a null src is an absent summand).  The sources of a
// descriptor are src_table[src0 + t*npos + p].  Threads walk dst in order, so
// the writes coalesce and every element is written exactly once.  The call
// does not wait.
struct BundleDesc {
  double* dst;
  int src0, nt, npos, rows, cols, interleaved;   // nt * rows * cols < 2^31
};
struct BundleSrc {
  const double* src;
  double alpha;
};
void dev_bundle_gather(Context& c, const std::vector<BundleDesc>& desc,
                       const std::vector<BundleSrc>& src_table);

// Batched transpose of square n x n blocks: dst[r*n + c] = src[r + c*n],
// through 64 x 64 tiles in LDS (both sides coalesce).  The call does not wait.
struct TransposeDesc {
  double* dst;
  const double* src;
  int n;
};
void dev_transpose_batch(Context& c, const std::vector<TransposeDesc>& desc);

}  // namespace su2
