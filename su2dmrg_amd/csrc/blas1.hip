#include "blas1.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>

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

__global__ void fill_f32_kernel(float* x, long long n, float v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] = v;
}

// rounding to nearest (the hardware conversion) and exact widening
__global__ void f64_to_f32_kernel(float* __restrict__ y, const double* __restrict__ x,
                                  long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = (float)x[i];
}

__global__ void f32_to_f64_kernel(double* __restrict__ y, const float* __restrict__ x,
                                  long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = (double)x[i];
}

// One launch for many vector moves: descriptor blockIdx.y, a grid-stride loop
// over its elements (a copy, a rounding to FP32, a widening to FP64, or a
// fill with zero).
__global__ void move_batch_kernel(const VecMove* __restrict__ desc) {
  const VecMove d = desc[blockIdx.y];
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  switch (d.mode) {
    case kMoveF64:
      for (; i < d.n; i += stride) ((double*)d.dst)[i] = ((const double*)d.src)[i];
      break;
    case kMoveF64toF32:
      for (; i < d.n; i += stride) ((float*)d.dst)[i] = (float)((const double*)d.src)[i];
      break;
    case kMoveF32toF64:
      for (; i < d.n; i += stride) ((double*)d.dst)[i] = (double)((const float*)d.src)[i];
      break;
    case kZeroF64:
      for (; i < d.n; i += stride) ((double*)d.dst)[i] = 0.0;
      break;
    case kZeroF32:
      for (; i < d.n; i += stride) ((float*)d.dst)[i] = 0.0f;
      break;
    default:
      break;
  }
}

// x *= 1 / sqrt(s[0]) with s read on the device (normalisation without a
// host round trip)
__global__ void scal_inv_sqrt_kernel(double* x, long long n, const double* __restrict__ s) {
  const double a = 1.0 / sqrt(s[0]);
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] *= a;
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

__global__ __launch_bounds__(kThreads) void dot_final_kernel(const double* __restrict__ part,
                                                             double* __restrict__ res) {
  __shared__ double sh[kThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < kDotBlocks; i += kThreads) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *res = sh[0];
}

// y += (sign * a[0]) x with the coefficient read on the device
__global__ void axpy_dev_kernel(double* __restrict__ y, const double* __restrict__ x, long long n,
                                const double* __restrict__ a, double sign) {
  const double av = sign * a[0];
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] += av * x[i];
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

constexpr int kGatherChunk = 4096;   // bundle elements per workgroup

constexpr int kGatherMaxPos = 16;    // summands held in registers (q-chunks)

// A chunk is either elements [first, first + count) of the bundle (pad 0) or,
// for an interleaved bundle of at most kThreads terms and kGatherMaxPos
// summands, source positions q in [first, first + count) of every term
// (pad 1): each thread then keeps one term, its sources and couplings in
// registers, and walks q; lanes still write consecutive elements.
__global__ __launch_bounds__(kThreads) void bundle_gather_kernel(
    const BundleDesc* __restrict__ desc, const BundleSrc* __restrict__ table,
    const AxpyChunk* __restrict__ chunks) {
  const AxpyChunk ch = chunks[blockIdx.x];
  const BundleDesc d = desc[ch.desc];
  if (ch.pad) {
    const int nt = d.nt, G = kThreads / nt;
    const int i = threadIdx.x;
    if (i >= G * nt) return;   // no barrier follows
    const int t = i % nt, g = i / nt;
    const BundleSrc* s = table + d.src0 + t * d.npos;
    const double* sp[kGatherMaxPos];
    double al[kGatherMaxPos];
#pragma unroll
    for (int p = 0; p < kGatherMaxPos; ++p) {
      sp[p] = p < d.npos ? s[p].src : nullptr;
      al[p] = p < d.npos ? s[p].alpha : 0.0;
    }
    for (int k = g; k < ch.count; k += G) {
      const int q = ch.first + k;
      double acc = 0.0;
#pragma unroll
      for (int p = 0; p < kGatherMaxPos; ++p)
        if (sp[p]) acc += al[p] * sp[p][q];
      d.dst[t + nt * q] = acc;
    }
    return;
  }
  const int blk = d.rows * d.cols;
  for (int j = threadIdx.x; j < ch.count; j += kThreads) {
    const int e = ch.first + j;
    int t, q;
    if (d.interleaved) {
      t = e % d.nt;
      q = e / d.nt;
    } else {
      t = e / blk;
      q = e - t * blk;
    }
    const BundleSrc* s = table + d.src0 + t * d.npos;
    double acc = 0.0;
    for (int p = 0; p < d.npos; ++p)
      if (s[p].src) acc += s[p].alpha * s[p].src[q];
    d.dst[e] = acc;
  }
}

struct TileRef {
  int desc, r0, c0, pad;
};

constexpr int kTile = 64;

__global__ __launch_bounds__(kThreads) void transpose_kernel(const TransposeDesc* __restrict__ desc,
                                                             const TileRef* __restrict__ tiles) {
  __shared__ double t[kTile][kTile + 1];
  const TileRef tr = tiles[blockIdx.x];
  const TransposeDesc d = desc[tr.desc];
  const int n = d.n;
  const int lx = threadIdx.x % kTile, ly = threadIdx.x / kTile;
  constexpr int kRowsPerPass = kThreads / kTile;
  // t[c - c0][r - r0] = src[r + c*n], reads along r
  for (int k = ly; k < kTile; k += kRowsPerPass) {
    const int r = tr.r0 + lx, c = tr.c0 + k;
    if (r < n && c < n) t[k][lx] = d.src[r + (long long)c * n];
  }
  __syncthreads();
  // dst[r*n + c], writes along c
  for (int k = ly; k < kTile; k += kRowsPerPass) {
    const int r = tr.r0 + k, c = tr.c0 + lx;
    if (r < n && c < n) d.dst[(long long)r * n + c] = t[lx][k];
  }
}

// Strided scale-add of a descriptor that writes along its columns (cs == 1,
// rs > 1): a 64 x 64 tile is read along r into LDS and added along c, so
// both the reads and the writes coalesce.
__global__ __launch_bounds__(kThreads) void strided_axpy_tiled_kernel(
    const StridedAxpy* __restrict__ desc, const TileRef* __restrict__ tiles) {
  __shared__ double t[kTile][kTile + 1];
  const TileRef tr = tiles[blockIdx.x];
  const StridedAxpy d = desc[tr.desc];
  const int lx = threadIdx.x % kTile, ly = threadIdx.x / kTile;
  constexpr int kRowsPerPass = kThreads / kTile;
  for (int k = ly; k < kTile; k += kRowsPerPass) {
    const int r = tr.r0 + lx, c = tr.c0 + k;
    if (r < d.rows && c < d.cols) t[k][lx] = d.src[r + (long long)c * d.ld];
  }
  __syncthreads();
  for (int k = ly; k < kTile; k += kRowsPerPass) {
    const int r = tr.r0 + k, c = tr.c0 + lx;
    if (r < d.rows && c < d.cols) {
      // rounded product, then the sum (not fused): the arithmetic of
      // strided_axpy_kernel
#pragma clang fp contract(off)
      double* p = d.dst + (r * d.rs + c * d.cs);
      const double v = d.alpha * t[lx][k];
      *p += v;
    }
  }
}

unsigned grid_for(long long n) {
  long long b = (n + kThreads - 1) / kThreads;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Scratch of the allocation-free dot: the kDotBlocks partial sums and the
// result.  One copy per GPU, held until the process ends; host code reaches
// it only through the address hipGetSymbolAddress returns for the current
// device (noalloc_scratch), never through this name.
constexpr int kNoAllocScratchDoubles = kDotBlocks + 1;
__device__ __align__(256) double scratch[kNoAllocScratchDoubles];

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;   // by device ordinal; null until resolved

}  // namespace

void noalloc_scratch_init(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return;
  int cur = -1;
  SU2_HIP(hipGetDevice(&cur));
  SU2_CHECK(cur == device, kErrArgument, "noalloc scratch: device %d is not current (%d)", device,
            cur);
  double* ptr = nullptr;
  SU2_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
  SU2_CHECK(ptr != nullptr, kErrDevice, "noalloc scratch: no address on device %d", device);
  g_scratch_ptr[device] = ptr;
}

double dev_dot_noalloc(Context& c, const double* x, const double* y, long long n) {
  double* part = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c.device() < (int)g_scratch_ptr.size()) part = g_scratch_ptr[c.device()];
  }
  SU2_CHECK(part != nullptr, kErrArgument,
            "noalloc dot: scratch of device %d not resolved (context not initialised)", c.device());
  // the partial count is fixed, so every length fits the scratch
  static_assert(kDotBlocks + 1 <= kNoAllocScratchDoubles, "dot scratch too small");
  double res = 0.0;
  if (n > 0) {
    hipLaunchKernelGGL(dot_partial_kernel, dim3(kDotBlocks), dim3(kThreads), 0, c.stream(), x, y,
                       n, part);
    SU2_HIP(hipGetLastError());
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, c.stream(), part,
                       part + kDotBlocks);
    SU2_HIP(hipGetLastError());
    // pageable destination: the copy completes before the call returns
    SU2_HIP(hipMemcpyAsync(&res, part + kDotBlocks, sizeof(double), hipMemcpyDeviceToHost,
                           c.stream()));
  }
  SU2_HIP(hipStreamSynchronize(c.stream()));
  return res;
}

long long noalloc_scratch_bytes() { return (long long)sizeof(double) * kNoAllocScratchDoubles; }

void dev_fill(Context& c, double* x, long long n, double v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c.stream(), x, n, v);
  SU2_HIP(hipGetLastError());
}

void dev_move_batch(Context& c, const VecMove* d_desc, int ndesc, long long max_n) {
  if (ndesc <= 0 || max_n <= 0) return;
  SU2_CHECK(ndesc < 65536, kErrCapacity, "vector moves: %d descriptors", ndesc);
  unsigned gx = grid_for(max_n);
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(move_batch_kernel, dim3(gx, ndesc), dim3(kThreads), 0, c.stream(), d_desc);
  SU2_HIP(hipGetLastError());
}

void dev_scal_inv_sqrt(Context& c, double* x, long long n, const double* d_s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scal_inv_sqrt_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c.stream(), x, n,
                     d_s);
  SU2_HIP(hipGetLastError());
}

void dev_fill_f32(Context& c, float* x, long long n, float v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_f32_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c.stream(), x, n, v);
  SU2_HIP(hipGetLastError());
}

void dev_convert_f64_f32(Context& c, float* y, const double* x, long long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(f64_to_f32_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c.stream(), y, x,
                     n);
  SU2_HIP(hipGetLastError());
}

void dev_convert_f32_f64(Context& c, double* y, const float* x, long long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(f32_to_f64_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c.stream(), y, x,
                     n);
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
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, c.stream(), part,
                     part + kDotBlocks);
  SU2_HIP(hipGetLastError());
  double* h = c.host_scratch(1);
  SU2_HIP(hipMemcpyAsync(h, part + kDotBlocks, sizeof(double), hipMemcpyDeviceToHost, c.stream()));
  SU2_HIP(hipStreamSynchronize(c.stream()));
  return h[0];
}

void dev_dot_to(Context& c, const double* x, const double* y, long long n, double* res) {
  if (n <= 0) {
    dev_fill(c, res, 1, 0.0);
    return;
  }
  double* part = c.reduce_scratch(kDotBlocks + 1);
  hipLaunchKernelGGL(dot_partial_kernel, dim3(kDotBlocks), dim3(kThreads), 0, c.stream(), x, y, n,
                     part);
  SU2_HIP(hipGetLastError());
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, c.stream(), part, res);
  SU2_HIP(hipGetLastError());
}

void dev_axpy_dev(Context& c, double* y, const double* x, long long n, const double* a,
                  double sign) {
  if (n <= 0) return;
  hipLaunchKernelGGL(axpy_dev_kernel, dim3(grid_for(n)), dim3(kThreads), 0, c.stream(), y, x, n, a,
                     sign);
  SU2_HIP(hipGetLastError());
}

void dev_strided_axpy_launches(Context& c,
                               const std::vector<std::vector<StridedAxpy>>& launches) {
  std::vector<StridedAxpy> flat;
  std::vector<AxpyChunk> chunks;
  std::vector<TileRef> tiles;
  // chunk and tile ranges of each launch
  std::vector<std::pair<size_t, size_t>> span, tspan;
  for (auto& l : launches) {
    const size_t c0 = chunks.size(), t0 = tiles.size();
    for (const StridedAxpy& d : l) {
      const long long n = (long long)d.rows * d.cols;
      if (n <= 0) continue;
      SU2_CHECK(n < (1LL << 31), kErrCapacity, "strided axpy: block of %lld elements", n);
      const int di = (int)flat.size();
      flat.push_back(d);
      if (d.src && d.cs == 1 && d.rs > 1 && d.rows > 1 && d.cols > 1) {
        for (int cc = 0; cc < d.cols; cc += kTile)
          for (int r0 = 0; r0 < d.rows; r0 += kTile) tiles.push_back(TileRef{di, r0, cc, 0});
      } else {
        for (long long f = 0; f < n; f += kAxpyChunk)
          chunks.push_back(AxpyChunk{di, (int)f, (int)std::min<long long>(kAxpyChunk, n - f), 0});
      }
    }
    span.push_back({c0, chunks.size()});
    tspan.push_back({t0, tiles.size()});
  }
  if (chunks.empty() && tiles.empty()) return;
  SU2_CHECK(tiles.size() < (1u << 31), kErrCapacity, "strided axpy: %zu tiles", tiles.size());
  const size_t dbytes = (flat.size() * sizeof(StridedAxpy) + 15) / 16 * 16;
  const size_t cbytes = (chunks.size() * sizeof(AxpyChunk) + 15) / 16 * 16;
  const size_t bytes = dbytes + cbytes + tiles.size() * sizeof(TileRef);
  std::vector<char> host(bytes, 0);
  std::memcpy(host.data(), flat.data(), flat.size() * sizeof(StridedAxpy));
  if (!chunks.empty())
    std::memcpy(host.data() + dbytes, chunks.data(), chunks.size() * sizeof(AxpyChunk));
  if (!tiles.empty())
    std::memcpy(host.data() + dbytes + cbytes, tiles.data(), tiles.size() * sizeof(TileRef));
  char* dev = (char*)c.dev_alloc(bytes);
  SU2_HIP(hipMemcpyAsync(dev, host.data(), bytes, hipMemcpyHostToDevice, c.stream()));
  const StridedAxpy* d_desc = (const StridedAxpy*)dev;
  const AxpyChunk* d_chunks = (const AxpyChunk*)(dev + dbytes);
  const TileRef* d_tiles = (const TileRef*)(dev + dbytes + cbytes);
  // the descriptors of one launch write disjoint elements, so its chunked
  // and tiled parts may run in either order
  for (size_t i = 0; i < span.size(); ++i) {
    const auto& s = span[i];
    const auto& ts = tspan[i];
    if (s.second > s.first) {
      hipLaunchKernelGGL(strided_axpy_kernel, dim3((unsigned)(s.second - s.first)),
                         dim3(kThreads), 0, c.stream(), d_desc, d_chunks + s.first);
      SU2_HIP(hipGetLastError());
    }
    if (ts.second > ts.first) {
      hipLaunchKernelGGL(strided_axpy_tiled_kernel, dim3((unsigned)(ts.second - ts.first)),
                         dim3(kThreads), 0, c.stream(), d_desc, d_tiles + ts.first);
      SU2_HIP(hipGetLastError());
    }
  }
  c.dev_free(dev);
}

void dev_strided_axpy(Context& c, const std::vector<StridedAxpy>& desc) {
  dev_strided_axpy_launches(c, {desc});
}

}  // namespace su2

namespace su2 {

namespace {

// Upload a descriptor table and a work table (each padded to 16 bytes) in one
// block of the context's cache; the block goes back to the cache once the
// stream has run past the launches that read it.
template <class A, class B, class C>
char* upload3(Context& c, const std::vector<A>& a, const std::vector<B>& b, const std::vector<C>& w,
              size_t& ob, size_t& ow) {
  auto pad = [](size_t x) { return (x + 15) / 16 * 16; };
  ob = pad(a.size() * sizeof(A));
  ow = ob + pad(b.size() * sizeof(B));
  const size_t bytes = ow + w.size() * sizeof(C);
  std::vector<char> host(bytes, 0);
  if (!a.empty()) std::memcpy(host.data(), a.data(), a.size() * sizeof(A));
  if (!b.empty()) std::memcpy(host.data() + ob, b.data(), b.size() * sizeof(B));
  std::memcpy(host.data() + ow, w.data(), w.size() * sizeof(C));
  char* dev = (char*)c.dev_alloc(bytes);
  SU2_HIP(hipMemcpyAsync(dev, host.data(), bytes, hipMemcpyHostToDevice, c.stream()));
  return dev;
}

}  // namespace

void dev_bundle_gather(Context& c, const std::vector<BundleDesc>& desc,
                       const std::vector<BundleSrc>& src_table) {
  std::vector<AxpyChunk> chunks;
  for (size_t i = 0; i < desc.size(); ++i) {
    const BundleDesc& d = desc[i];
    const long long n = (long long)d.nt * d.rows * d.cols;
    if (n <= 0) continue;
    SU2_CHECK(n < (1LL << 31), kErrCapacity, "bundle gather: bundle of %lld elements", n);
    SU2_CHECK(d.npos >= 0 && d.src0 >= 0 &&
                  (long long)d.src0 + (long long)d.nt * d.npos <= (long long)src_table.size(),
              kErrArgument, "bundle gather: source table range out of bounds");
    if (d.interleaved && d.nt <= kThreads && d.npos <= kGatherMaxPos) {
      const long long nq = (long long)d.rows * d.cols;
      const long long qc = std::max(1, kGatherChunk / d.nt);
      for (long long q0 = 0; q0 < nq; q0 += qc)
        chunks.push_back(AxpyChunk{(int)i, (int)q0, (int)std::min<long long>(qc, nq - q0), 1});
    } else {
      for (long long f = 0; f < n; f += kGatherChunk)
        chunks.push_back(
            AxpyChunk{(int)i, (int)f, (int)std::min<long long>(kGatherChunk, n - f), 0});
    }
  }
  if (chunks.empty()) return;
  size_t ob = 0, ow = 0;
  char* dev = upload3(c, desc, src_table, chunks, ob, ow);
  hipLaunchKernelGGL(bundle_gather_kernel, dim3((unsigned)chunks.size()), dim3(kThreads), 0,
                     c.stream(), (const BundleDesc*)dev, (const BundleSrc*)(dev + ob),
                     (const AxpyChunk*)(dev + ow));
  SU2_HIP(hipGetLastError());
  c.dev_free(dev);
}

void dev_transpose_batch(Context& c, const std::vector<TransposeDesc>& desc) {
  std::vector<TileRef> tiles;
  for (size_t i = 0; i < desc.size(); ++i) {
    const int n = desc[i].n;
    for (int c0 = 0; c0 < n; c0 += kTile)
      for (int r0 = 0; r0 < n; r0 += kTile) tiles.push_back(TileRef{(int)i, r0, c0, 0});
  }
  if (tiles.empty()) return;
  SU2_CHECK(tiles.size() < (1u << 31), kErrCapacity, "transpose: %zu tiles", tiles.size());
  size_t ob = 0, ow = 0;
  const std::vector<char> none;
  char* dev = upload3(c, desc, none, tiles, ob, ow);
  hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)tiles.size()), dim3(kThreads), 0, c.stream(),
                     (const TransposeDesc*)dev, (const TileRef*)(dev + ow));
  SU2_HIP(hipGetLastError());
  c.dev_free(dev);
}

}  // namespace su2
