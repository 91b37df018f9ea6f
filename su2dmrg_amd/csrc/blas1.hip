#include "blas1.hpp"

#include <algorithm>
#include <cstring>

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

constexpr int kAxpyChunk = 4096;   // elements of one descriptor per workgroup

struct AxpyChunk {
  int desc, first, count, pad;
};

// One workgroup per chunk: elements [first, first + count) of one descriptor
// in column-major order (r fastest), so the source reads coalesce.
__global__ __launch_bounds__(kThreads) void strided_axpy_kernel(
    const StridedAxpy* __restrict__ desc, const AxpyChunk* __restrict__ chunks) {
  const AxpyChunk ch = chunks[blockIdx.x];
  const StridedAxpy d = desc[ch.desc];
  const int rows = d.rows;
  for (int j = threadIdx.x; j < ch.count; j += kThreads) {
    const int i = ch.first + j;
    const int c = i / rows, r = i - c * rows;
    // a null source adds alpha (an identity block when rs/cs walk a diagonal)
    double* p = d.dst + (r * d.rs + c * d.cs);
    *p += d.src ? d.alpha * d.src[r + c * d.ld] : d.alpha;
  }
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

void dev_strided_axpy_launches(Context& c,
                               const std::vector<std::vector<StridedAxpy>>& launches) {
  std::vector<StridedAxpy> flat;
  std::vector<AxpyChunk> chunks;
  std::vector<std::pair<size_t, size_t>> span;   // chunk range of each launch
  for (auto& l : launches) {
    const size_t c0 = chunks.size();
    for (const StridedAxpy& d : l) {
      const long long n = (long long)d.rows * d.cols;
      if (n <= 0) continue;
      SU2_CHECK(n < (1LL << 31), kErrCapacity, "strided axpy: block of %lld elements", n);
      const int di = (int)flat.size();
      flat.push_back(d);
      for (long long f = 0; f < n; f += kAxpyChunk)
        chunks.push_back(AxpyChunk{di, (int)f, (int)std::min<long long>(kAxpyChunk, n - f), 0});
    }
    span.push_back({c0, chunks.size()});
  }
  if (chunks.empty()) return;
  const size_t dbytes = (flat.size() * sizeof(StridedAxpy) + 15) / 16 * 16;
  const size_t bytes = dbytes + chunks.size() * sizeof(AxpyChunk);
  std::vector<char> host(bytes);
  std::memcpy(host.data(), flat.data(), flat.size() * sizeof(StridedAxpy));
  std::memcpy(host.data() + dbytes, chunks.data(), chunks.size() * sizeof(AxpyChunk));
  char* dev = (char*)c.dev_alloc(bytes);
  SU2_HIP(hipMemcpyAsync(dev, host.data(), bytes, hipMemcpyHostToDevice, c.stream()));
  const StridedAxpy* d_desc = (const StridedAxpy*)dev;
  const AxpyChunk* d_chunks = (const AxpyChunk*)(dev + dbytes);
  for (auto& s : span) {
    if (s.second == s.first) continue;
    hipLaunchKernelGGL(strided_axpy_kernel, dim3((unsigned)(s.second - s.first)), dim3(kThreads),
                       0, c.stream(), d_desc, d_chunks + s.first);
    SU2_HIP(hipGetLastError());
  }
  c.dev_free(dev);
}

void dev_strided_axpy(Context& c, const std::vector<StridedAxpy>& desc) {
  dev_strided_axpy_launches(c, {desc});
}

}  // namespace su2
