#include "blas1.hpp"

#include "common.hpp"
#include "context.hpp"

namespace su2 {

namespace {

constexpr int kThreads = 256;
constexpr int kDotBlocks = 512;   // fixed partial count -> fixed summation order

__global__ void fill_kernel(double* x, long long n, double v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] = v;
}

__global__ void copy_kernel(double* __restrict__ y, const double* __restrict__ x, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = x[i];
}

__global__ void scal_kernel(double* x, long long n, double a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] *= a;
}

__global__ void axpby_kernel(double* __restrict__ y, const double* __restrict__ x, long long n,
                             double a, double b) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = a * x[i] + b * y[i];
}

__global__ void axpy_kernel(double* __restrict__ y, const double* __restrict__ x, long long n,
                            double a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] += a * x[i];
}

__global__ __launch_bounds__(kThreads) void dot_partial_kernel(const double* __restrict__ x,
                                                               const double* __restrict__ y,
                                                               long long n,
                                                               double* __restrict__ part) {
  __shared__ double sh[kThreads];
  double s = 0.0;
  long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  const long long stride = (long long)kDotBlocks * kThreads;
  for (; i < n; i += stride) s += x[i] * y[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kThreads) void dot_final_kernel(double* __restrict__ part) {
  __shared__ double sh[kThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < kDotBlocks; i += kThreads) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[kDotBlocks] = sh[0];
}

__global__ void strided_axpy_kernel(const StridedAxpy* __restrict__ desc, int ndesc,
                                    long long total) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  int lo = 0, hi = ndesc - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (desc[mid].elem_begin <= e) lo = mid;
    else hi = mid - 1;
  }
  const StridedAxpy d = desc[lo];
  long long i = e - d.elem_begin;
  long long r = i % d.rows, c = i / d.rows;
  // a null source adds alpha (an identity block when rs/cs walk a diagonal)
  d.dst[r * d.rs + c * d.cs] += d.src ? d.alpha * d.src[r + c * d.ld] : d.alpha;
}

unsigned grid_for(long long n) {
  long long b = (n + kThreads - 1) / kThreads;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

void dev_fill(Context& c, double* x, long long n, double v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c.stream(), x, n, v);
  SU2_HIP(hipGetLastError());
}

void dev_copy(Context& c, double* y, const double* x, long long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c.stream(), y, x, n);
  SU2_HIP(hipGetLastError());
}

void dev_scal(Context& c, double* x, long long n, double a) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scal_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c.stream(), x, n, a);
  SU2_HIP(hipGetLastError());
}

void dev_axpy(Context& c, double* y, const double* x, long long n, double a) {
  if (n <= 0) return;
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c.stream(), y, x, n, a);
  SU2_HIP(hipGetLastError());
}

void dev_axpby(Context& c, double* y, const double* x, long long n, double a, double b) {
  if (n <= 0) return;
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c.stream(), y, x, n, a,
                     b);
  SU2_HIP(hipGetLastError());
}

double dev_dot(Context& c, const double* x, const double* y, long long n) {
  if (n <= 0) return 0.0;
  double* part = c.reduce_scratch(kDotBlocks + 1);
  hipLaunchKernelGGL(dot_partial_kernel, dim3(kDotBlocks), dim3(kThreads), 0, c.stream(), x, y, n,
                     part);
  SU2_HIP(hipGetLastError());
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, c.stream(), part);
  SU2_HIP(hipGetLastError());
  double* h = c.host_scratch(1);
  SU2_HIP(hipMemcpyAsync(h, part + kDotBlocks, sizeof(double), hipMemcpyDeviceToHost, c.stream()));
  SU2_HIP(hipStreamSynchronize(c.stream()));
  return h[0];
}

void dev_strided_axpy(Context& c, const StridedAxpy* d_desc, int ndesc, long long total) {
  if (total <= 0 || ndesc <= 0) return;
  long long blocks = (total + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(strided_axpy_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, c.stream(),
                     d_desc, ndesc, total);
  SU2_HIP(hipGetLastError());
}

}  // namespace su2
