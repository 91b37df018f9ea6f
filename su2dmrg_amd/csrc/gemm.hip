// Grouped FP64 / FP32 MFMA GEMM for gfx950.  See gemm.hpp for the problem
// model.
//
// Workgroup tile BM x BN (128 x 128), 4 waves in a 2 x 2 arrangement, each
// wave owning a 64 x 64 sub-tile = 4 x 4 blocks of 16 x 16 x 4 MFMAs
// (v_mfma_f64_16x16x4_f64, or v_mfma_f32_16x16x4_f32 with FP32
// accumulation).  K is staged through LDS in BK=16 slabs, double buffered:
// the next slab is fetched from global memory into registers while the
// current one feeds the matrix cores, then written to the other LDS buffer
// (one barrier per slab).
//
// MFMA operand layout (16x16x4, 64 lanes, both types):
//   A (16 x 4):  lane l holds A[l % 16][l / 16]
//   B (4 x 16):  lane l holds B[l / 16][l % 16]
// Result layout (4 values per lane, r = 0..3):
//   f64: lane l holds D[(l / 16) + 4*r][l % 16]
//        (confirmed on the MI355X with a one-wave probe)
//   f32: lane l holds D[4*(l / 16) + r][l % 16]
#include "gemm.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "context.hpp"

namespace su2 {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

// Matrix-core step of one element type: acc += a (16x4) * b (4x16), and the
// row of result register r held by a lane.
template <class T>
struct Mma;
template <>
struct Mma<double> {
  typedef f64x4 acc_t;
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mma<float> {
  typedef f32x4 acc_t;
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// Logical work unit of this workgroup, or -1: the hardware deals workgroups
// to the 8 XCDs round-robin (id % 8).  XCD x takes the contiguous units
// [begin[x], begin[x+1]), so that neighbouring tiles (which share A or B
// panels) meet in one L2; the ranges hold equal work, not equal unit counts
// (units of different problems differ in their K), and the grid is 8 times
// the longest range, the surplus workgroups leaving at once.
__device__ __forceinline__ long long xcd_unit(const XcdSplit& xs) {
  const int x = (int)(blockIdx.x % 8);
  const long long u = xs.begin[x] + blockIdx.x / 8;
  return u < xs.begin[x + 1] ? u : -1;
}

__device__ __forceinline__ int find_problem(const GemmProblem* __restrict__ probs, int nprob,
                                            long long unit) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (probs[mid].unit_begin <= unit) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Global-address-space view of a device pointer: loads through it compile to
// global_load (vmcnt only) instead of flat_load, whose lgkmcnt would make every
// LDS wait in the MFMA loop also wait for the next slab's memory traffic.
template <class T>
__device__ __forceinline__ T gload(const T* p, long long i) {
  return ((const T __attribute__((address_space(1)))*)p)[i];
}

template <class T_, int BM, int BN, int BK, int WM, int WN, int PAD>
struct Tile {
  typedef T_ T;
  static constexpr int NT = WM * WN * 64;
  static constexpr int WTM = BM / WM;
  static constexpr int WTN = BN / WN;
  static constexpr int FM = WTM / 16;
  static constexpr int FN = WTN / 16;
  static constexpr int APT = BM * BK / NT;   // A elements per thread per slab
  static constexpr int BPT = BN * BK / NT;   // B elements per thread per slab
  static constexpr int LDA_S = BM + PAD;     // LDS row pitch (elements)
  static constexpr int LDB_S = BN + PAD;
  static_assert(BM * BK % NT == 0 && BN * BK % NT == 0, "slab must split evenly");
  static_assert(WTM % 16 == 0 && WTN % 16 == 0 && BK % 4 == 0, "MFMA tiling");
  static_assert(APT <= 32 && BPT <= 32, "mask width");
};

template <class TL>
struct SlabRegs {
  typename TL::T a[TL::APT];
  typename TL::T b[TL::BPT];
  unsigned amask, bmask;   // bit i set: element i lies inside the matrix
};

// Issue the global loads of one BK slab.  Addresses are clamped into the
// matrix so every load is unconditional (a branch per load would make the
// compiler wait for each one); out-of-range elements are zeroed by the mask
// in store_slab, after the matrix cores have consumed the previous slab.
template <class E, int BM, int BN, int BK, int WM, int WN, int PAD>
__device__ __forceinline__ void load_slab(const GemmSlice& s, int kt_local, int m0, int n0, int M,
                                          int N, SlabRegs<Tile<E, BM, BN, BK, WM, WN, PAD>>& r) {
  using T = Tile<E, BM, BN, BK, WM, WN, PAD>;
  constexpr unsigned kEs = sizeof(E);
  const int t = threadIdx.x;
  const int k0 = kt_local * BK;
  r.amask = 0;
  r.bmask = 0;
  if (!s.transA) {
    // op(A) = A, column-major M x K: coalesced along m.
#pragma unroll
    for (int i = 0; i < T::APT; ++i) {
      int e = t + i * T::NT;
      int gm = m0 + e % BM, gk = k0 + e / BM;
      r.amask |= (unsigned)(gm < M && gk < s.K) << i;
      r.a[i] = gload((const E*)s.A, min(gm, M - 1) + (long long)min(gk, s.K - 1) * s.lda);
    }
  } else {
    // op(A) = A^T, A stored K x M: coalesced along k.
#pragma unroll
    for (int i = 0; i < T::APT; ++i) {
      int e = t + i * T::NT;
      int gk = k0 + e % BK, gm = m0 + e / BK;
      r.amask |= (unsigned)(gm < M && gk < s.K) << i;
      r.a[i] = gload((const E*)s.A, min(gk, s.K - 1) + (long long)min(gm, M - 1) * s.lda);
    }
  }
  if (!s.transB) {
    // B stored K x N: coalesced along k.
#pragma unroll
    for (int i = 0; i < T::BPT; ++i) {
      int e = t + i * T::NT;
      int gk = k0 + e % BK, gn = n0 + e / BK;
      r.bmask |= (unsigned)(gn < N && gk < s.K) << i;
      r.b[i] = gload((const E*)s.B, min(gk, s.K - 1) + (long long)min(gn, N - 1) * s.ldb);
    }
  } else {
    // op(B) = B^T, B stored N x K: coalesced along n.
#pragma unroll
    for (int i = 0; i < T::BPT; ++i) {
      int e = t + i * T::NT;
      int gn = n0 + e % BN, gk = k0 + e / BN;
      r.bmask |= (unsigned)(gn < N && gk < s.K) << i;
      r.b[i] = gload((const E*)s.B, min(gn, N - 1) + (long long)min(gk, s.K - 1) * s.ldb);
    }
  }
}

// Per-thread addressing of one slice for the current tile.  Every variant has
// NT divisible by BM, BN and BK, so element i of a thread's share of a slab
// sits at off0 + i * stride with a uniform stride, and a whole slab moves by a
// uniform step; only the M / N edge masks vary with i, and they are fixed for
// the tile.  Offsets are unsigned 32-bit byte offsets from the slice's
// (uniform) base pointer, so each load is one global_load with a scalar base
// and a vector offset; the host checks every slice matrix spans under 4 GB.
struct SliceAddr {
  unsigned a0, b0;          // byte offset of element 0 at k0 = 0
  unsigned a_safe, b_safe;  // in-matrix stand-in for elements outside the M / N edge
  unsigned amn, bmn;        // bit i: element i lies inside the M (A) / N (B) edge
};

template <class E>
__device__ __forceinline__ E gload_b(const void* p, unsigned byte_off) {
  typedef const char __attribute__((address_space(1)))* gcptr;
  return *(const E __attribute__((address_space(1)))*)((gcptr)p + byte_off);
}

template <class E, int BM, int BN, int BK, int WM, int WN, int PAD>
__device__ __forceinline__ SliceAddr slice_addr(const GemmSlice& s, int m0, int n0, int M, int N) {
  using T = Tile<E, BM, BN, BK, WM, WN, PAD>;
  constexpr unsigned kEs = sizeof(E);
  static_assert(T::NT % BM == 0 && T::NT % BN == 0 && T::NT % BK == 0, "affine slab layout");
  const unsigned t = threadIdx.x;
  const unsigned lda = (unsigned)s.lda, ldb = (unsigned)s.ldb;
  SliceAddr ad;
  if (!s.transA) {
    // column-major M x K: fixed row per thread, k = t / BM + i * NT / BM
    const int gm = m0 + (int)(t % BM);
    ad.a0 = kEs * ((unsigned)min(gm, M - 1) + (t / BM) * lda);
    ad.a_safe = ad.a0;
    ad.amn = gm < M ? 0xffffffffu : 0u;
  } else {
    // K x M storage: fixed k per thread, row = m0 + t / BK + i * NT / BK
    const int r0 = m0 + (int)(t / BK);
    ad.a0 = kEs * (t % BK + (unsigned)min(r0, M - 1) * lda);
    ad.a_safe = kEs * (t % BK + (unsigned)(M - 1) * lda);
    ad.amn = 0;
#pragma unroll
    for (int i = 0; i < T::APT; ++i) ad.amn |= (unsigned)(r0 + i * (T::NT / BK) < M) << i;
  }
  if (!s.transB) {
    // K x N storage: fixed k per thread, column = n0 + t / BK + i * NT / BK
    const int c0 = n0 + (int)(t / BK);
    ad.b0 = kEs * (t % BK + (unsigned)min(c0, N - 1) * ldb);
    ad.b_safe = kEs * (t % BK + (unsigned)(N - 1) * ldb);
    ad.bmn = 0;
#pragma unroll
    for (int i = 0; i < T::BPT; ++i) ad.bmn |= (unsigned)(c0 + i * (T::NT / BK) < N) << i;
  } else {
    // N x K storage: fixed column per thread, k = t / BN + i * NT / BN
    const int gn = n0 + (int)(t % BN);
    ad.b0 = kEs * ((unsigned)min(gn, N - 1) + (t / BN) * ldb);
    ad.b_safe = ad.b0;
    ad.bmn = gn < N ? 0xffffffffu : 0u;
  }
  return ad;
}

// Loads of a slab that lies wholly inside the slice's K range: uniform offset
// steps, masks from the tile edges only.
template <class E, int BM, int BN, int BK, int WM, int WN, int PAD>
__device__ __forceinline__ void load_slab_full(const GemmSlice& s, const SliceAddr& ad, int k0,
                                               SlabRegs<Tile<E, BM, BN, BK, WM, WN, PAD>>& r) {
  using T = Tile<E, BM, BN, BK, WM, WN, PAD>;
  constexpr unsigned kEs = sizeof(E);
  const unsigned lda8 = kEs * (unsigned)s.lda, ldb8 = kEs * (unsigned)s.ldb;
  const unsigned kk = (unsigned)k0;
  if (!s.transA) {
    const unsigned base = ad.a0 + kk * lda8, step = (T::NT / BM) * lda8;
#pragma unroll
    for (int i = 0; i < T::APT; ++i) r.a[i] = gload_b<E>(s.A, base + i * step);
  } else {
    const unsigned step = (T::NT / BK) * lda8;
#pragma unroll
    for (int i = 0; i < T::APT; ++i)
      r.a[i] = gload_b<E>(s.A, (((ad.amn >> i) & 1u) ? ad.a0 + i * step : ad.a_safe) + kEs * kk);
  }
  if (!s.transB) {
    const unsigned step = (T::NT / BK) * ldb8;
#pragma unroll
    for (int i = 0; i < T::BPT; ++i)
      r.b[i] = gload_b<E>(s.B, (((ad.bmn >> i) & 1u) ? ad.b0 + i * step : ad.b_safe) + kEs * kk);
  } else {
    const unsigned base = ad.b0 + kk * ldb8, step = (T::NT / BN) * ldb8;
#pragma unroll
    for (int i = 0; i < T::BPT; ++i) r.b[i] = gload_b<E>(s.B, base + i * step);
  }
  r.amask = ad.amn;
  r.bmask = ad.bmn;
}

// Slab kl of a slice: the uniform-step path when the slab is full in K, the
// clamped path for the slice's last, partial slab.
template <class E, int BM, int BN, int BK, int WM, int WN, int PAD>
__device__ __forceinline__ void load_slab_any(const GemmSlice& s, const SliceAddr& ad, int kl,
                                              int m0, int n0, int M, int N,
                                              SlabRegs<Tile<E, BM, BN, BK, WM, WN, PAD>>& r) {
  if ((kl + 1) * BK <= s.K) load_slab_full<E, BM, BN, BK, WM, WN, PAD>(s, ad, kl * BK, r);
  else load_slab<E, BM, BN, BK, WM, WN, PAD>(s, kl, m0, n0, M, N, r);
}

template <class E, int BM, int BN, int BK, int WM, int WN, int PAD>
__device__ __forceinline__ void store_slab(const GemmSlice& s,
                                           const SlabRegs<Tile<E, BM, BN, BK, WM, WN, PAD>>& r,
                                           E* As, E* Bs) {
  using T = Tile<E, BM, BN, BK, WM, WN, PAD>;
  constexpr unsigned kEs = sizeof(E);
  const int t = threadIdx.x;
  const E alpha = (E)s.alpha;
  if (!s.transA) {
#pragma unroll
    for (int i = 0; i < T::APT; ++i) {
      int e = t + i * T::NT;
      As[(e / BM) * T::LDA_S + e % BM] = ((r.amask >> i) & 1u) ? alpha * r.a[i] : E(0);
    }
  } else {
#pragma unroll
    for (int i = 0; i < T::APT; ++i) {
      int e = t + i * T::NT;
      As[(e % BK) * T::LDA_S + e / BK] = ((r.amask >> i) & 1u) ? alpha * r.a[i] : E(0);
    }
  }
  if (!s.transB) {
#pragma unroll
    for (int i = 0; i < T::BPT; ++i) {
      int e = t + i * T::NT;
      Bs[(e % BK) * T::LDB_S + e / BK] = ((r.bmask >> i) & 1u) ? r.b[i] : E(0);
    }
  } else {
#pragma unroll
    for (int i = 0; i < T::BPT; ++i) {
      int e = t + i * T::NT;
      Bs[(e / BN) * T::LDB_S + e % BN] = ((r.bmask >> i) & 1u) ? r.b[i] : E(0);
    }
  }
}

// Waves per SIMD: two for sub-tiles up to 64 x 64 per wave (at most 256
// registers per lane; one wave's MFMAs run while the other loads, stores its
// next slab or waits at the barrier), one for larger sub-tiles, whose
// accumulators take the 256 accumulation registers and whose MFMA work per
// slab covers the barrier by itself, and one where the two LDS slab buffers
// take more than half of the CU's 160 KB (a second workgroup cannot start).
template <int BM, int BN, int BK, int WM, int WN, int PAD, int ES>
struct WavesPerSimd {
  static constexpr int lds = 2 * BK * (BM + PAD + BN + PAD) * ES;
  static constexpr int value =
      ((BM / WM) * (BN / WN) * ES > 64 * 64 * 8 || lds > 80 * 1024) ? 1 : 2;
};

template <class E, int BM, int BN, int BK, int WM, int WN, int PAD>
__global__ __launch_bounds__(WM * WN * 64, (WavesPerSimd<BM, BN, BK, WM, WN, PAD, (int)sizeof(E)>::value)) void
grouped_gemm_f64_kernel(
    const GemmProblem* __restrict__ probs, int nprob, const GemmSlice* __restrict__ slices,
    void* __restrict__ ws, XcdSplit xs) {
  using T = Tile<E, BM, BN, BK, WM, WN, PAD>;
  constexpr unsigned kEs = sizeof(E);
  __shared__ E As[2][BK * T::LDA_S];
  __shared__ E Bs[2][BK * T::LDB_S];

  const long long unit = xcd_unit(xs);
  if (unit < 0) return;
  const int p = find_problem(probs, nprob, unit);
  const GemmProblem P = probs[p];
  long long u = unit - P.unit_begin;
  const int ntile = P.tiles_m * P.tiles_n;
  const int ks = (int)(u / ntile);
  const int tile = (int)(u % ntile);
  const int tm = P.n_fast ? tile / P.tiles_n : tile % P.tiles_m;
  const int tn = P.n_fast ? tile % P.tiles_n : tile / P.tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const int kt_lo = (int)((long long)P.kt_total * ks / P.ksplit);
  const int kt_hi = (int)((long long)P.kt_total * (ks + 1) / P.ksplit);

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave % WM) * T::WTM;
  const int wn0 = (wave / WM) * T::WTN;

  typename Mma<E>::acc_t acc[T::FM][T::FN];
#pragma unroll
  for (int i = 0; i < T::FM; ++i)
#pragma unroll
    for (int j = 0; j < T::FN; ++j) acc[i][j] = typename Mma<E>::acc_t{E(0), E(0), E(0), E(0)};

  if (kt_lo < kt_hi) {
    // locate the slice holding k-tile kt_lo
    int s = P.slice_begin;
    while (s + 1 < P.slice_end && slices[s + 1].kt_begin <= kt_lo) ++s;
    GemmSlice S = slices[s];
    int kl = kt_lo - S.kt_begin;
    SliceAddr ad = slice_addr<E, BM, BN, BK, WM, WN, PAD>(S, m0, n0, P.M, P.N);

    SlabRegs<T> regs;
    load_slab_any<E, BM, BN, BK, WM, WN, PAD>(S, ad, kl, m0, n0, P.M, P.N, regs);
    store_slab<E, BM, BN, BK, WM, WN, PAD>(S, regs, As[0], Bs[0]);
    __syncthreads();

    int buf = 0;
    for (int kt = kt_lo; kt < kt_hi; ++kt) {
      // next slab (possibly in the next slice)
      const bool more = (kt + 1 < kt_hi);
      GemmSlice Sn = S;
      int kln = kl + 1;
      if (more) {
        int nkt = (S.K + BK - 1) / BK;
        if (kln >= nkt) {
          ++s;
          Sn = slices[s];
          kln = 0;
          ad = slice_addr<E, BM, BN, BK, WM, WN, PAD>(Sn, m0, n0, P.M, P.N);
        }
        load_slab_any<E, BM, BN, BK, WM, WN, PAD>(Sn, ad, kln, m0, n0, P.M, P.N, regs);
      }
      const E* Ab = As[buf];
      const E* Bb = Bs[buf];
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        E af[T::FM], bf[T::FN];
        const int kr = kk + (lane >> 4);
#pragma unroll
        for (int i = 0; i < T::FM; ++i) af[i] = Ab[kr * T::LDA_S + wm0 + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < T::FN; ++j) bf[j] = Bb[kr * T::LDB_S + wn0 + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < T::FM; ++i)
#pragma unroll
          for (int j = 0; j < T::FN; ++j)
            acc[i][j] = Mma<E>::run(af[i], bf[j], acc[i][j]);
      }
      if (more) {
        store_slab<E, BM, BN, BK, WM, WN, PAD>(Sn, regs, As[buf ^ 1], Bs[buf ^ 1]);
        S = Sn;
        kl = kln;
      }
      __syncthreads();
      buf ^= 1;
    }
  }

  // epilogue: C directly, or this split's partial in the workspace
  E* dst;
  long long ld;
  if (P.ksplit == 1) {
    dst = (E*)P.C;
    ld = P.ldc;
  } else {
    dst = (E*)ws + P.ws_offset + (long long)ks * P.M * P.N;
    ld = P.M;
  }
#pragma unroll
  for (int i = 0; i < T::FM; ++i)
#pragma unroll
    for (int j = 0; j < T::FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = m0 + wm0 + i * 16 + Mma<E>::row(lane, r);
        int col = n0 + wn0 + j * 16 + (lane & 15);
        if (row < P.M && col < P.N) dst[row + (long long)col * ld] = acc[i][j][r];
      }
}

// ---------------------------------------------------------------------------
// Uniform-layout kernel: every slice of the launch has op(B) = B (stored
// K x N) and the same op(A) (TA = 0: A stored M x K; TA = 1: A stored K x M,
// used transposed).  With the layout fixed at compile time, a full slab's
// loads and LDS stores are straight-line code.  Elements past the M / N edge
// load a clamped in-matrix element rather than being masked: they only reach
// tile rows and columns that the epilogue does not store.  Only the K tail
// of a slice is zeroed, since its extra products would reach stored outputs;
// its out-of-range elements load a zero from device memory, so no value
// waits on its load.  The next slab's scale-and-store to LDS is issued
// inside the last group of MFMAs of the current slab, so the matrix cores
// keep working while it runs, and only the barrier separates two slabs.
// (Unrolling the K loop by two, to make the buffers compile-time, makes the
// register allocator spill the accumulators; the buffer is a runtime select.)

__device__ double g_gemm_zero = 0.0;
__device__ float g_gemm_zero_f = 0.0f;
template <class E>
__device__ __forceinline__ const E* gemm_zero();
template <>
__device__ __forceinline__ const double* gemm_zero<double>() {
  return &g_gemm_zero;
}
template <>
__device__ __forceinline__ const float* gemm_zero<float>() {
  return &g_gemm_zero_f;
}

template <class E, int BM, int BN, int BK, int WM, int WN, int PAD, int TA>
__device__ __forceinline__ void load_full_u(const GemmSlice& s, const SliceAddr& ad, int k0,
                                            SlabRegs<Tile<E, BM, BN, BK, WM, WN, PAD>>& r) {
  using T = Tile<E, BM, BN, BK, WM, WN, PAD>;
  constexpr unsigned kEs = sizeof(E);
  const unsigned lda8 = kEs * (unsigned)s.lda, ldb8 = kEs * (unsigned)s.ldb;
  const unsigned kk = (unsigned)k0;
  if (TA == 0) {
    const unsigned base = ad.a0 + kk * lda8, step = (T::NT / BM) * lda8;
#pragma unroll
    for (int i = 0; i < T::APT; ++i) r.a[i] = gload_b<E>(s.A, base + i * step);
  } else {
    const unsigned step = (T::NT / BK) * lda8;
#pragma unroll
    for (int i = 0; i < T::APT; ++i)
      r.a[i] = gload_b<E>(s.A, (((ad.amn >> i) & 1u) ? ad.a0 + i * step : ad.a_safe) + kEs * kk);
  }
  const unsigned step = (T::NT / BK) * ldb8;
#pragma unroll
  for (int i = 0; i < T::BPT; ++i)
    r.b[i] = gload_b<E>(s.B, (((ad.bmn >> i) & 1u) ? ad.b0 + i * step : ad.b_safe) + kEs * kk);
}

// The last, partial slab of a slice: elements with k >= K read g_gemm_zero.
template <class E, int BM, int BN, int BK, int WM, int WN, int PAD, int TA>
__device__ __forceinline__ void load_tail_u(const GemmSlice& s, int k0, int m0, int n0, int M,
                                            int N, SlabRegs<Tile<E, BM, BN, BK, WM, WN, PAD>>& r) {
  using T = Tile<E, BM, BN, BK, WM, WN, PAD>;
  constexpr unsigned kEs = sizeof(E);
  const int t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < T::APT; ++i) {
    const int e = t + i * T::NT;
    const int gm = min(m0 + (TA == 0 ? e % BM : e / BK), M - 1);
    const int gk = k0 + (TA == 0 ? e / BM : e % BK);
    const E* A = (const E*)s.A;
    const E* p = TA == 0 ? A + gm + (long long)gk * s.lda : A + gk + (long long)gm * s.lda;
    r.a[i] = gload(gk < s.K ? p : gemm_zero<E>(), 0);
  }
#pragma unroll
  for (int i = 0; i < T::BPT; ++i) {
    const int e = t + i * T::NT;
    const int gn = min(n0 + e / BK, N - 1);
    const int gk = k0 + e % BK;
    const E* p = (const E*)s.B + gk + (long long)gn * s.ldb;
    r.b[i] = gload(gk < s.K ? p : gemm_zero<E>(), 0);
  }
}

template <class E, int BM, int BN, int BK, int WM, int WN, int PAD, int TA>
__device__ __forceinline__ void store_u(const SlabRegs<Tile<E, BM, BN, BK, WM, WN, PAD>>& r,
                                        E alpha, E* As, E* Bs) {
  using T = Tile<E, BM, BN, BK, WM, WN, PAD>;
  constexpr unsigned kEs = sizeof(E);
  const int t = threadIdx.x;
  if (TA == 0) {
    E* a = As + (t / BM) * T::LDA_S + t % BM;
#pragma unroll
    for (int i = 0; i < T::APT; ++i) a[i * (T::NT / BM) * T::LDA_S] = alpha * r.a[i];
  } else {
    E* a = As + (t % BK) * T::LDA_S + t / BK;
#pragma unroll
    for (int i = 0; i < T::APT; ++i) a[i * (T::NT / BK)] = alpha * r.a[i];
  }
  E* b = Bs + (t % BK) * T::LDB_S + t / BK;
#pragma unroll
  for (int i = 0; i < T::BPT; ++i) b[i * (T::NT / BK)] = r.b[i];
}

template <class E, int BM, int BN, int BK, int WM, int WN, int PAD, int TA>
__global__ __launch_bounds__(WM * WN * 64, (WavesPerSimd<BM, BN, BK, WM, WN, PAD, (int)sizeof(E)>::value)) void
grouped_gemm_f64_uniform(const GemmProblem* __restrict__ probs, int nprob,
                         const GemmSlice* __restrict__ slices, void* __restrict__ ws,
                         XcdSplit xs) {
  using T = Tile<E, BM, BN, BK, WM, WN, PAD>;
  constexpr unsigned kEs = sizeof(E);
  __shared__ E As0[BK * T::LDA_S], As1[BK * T::LDA_S];
  __shared__ E Bs0[BK * T::LDB_S], Bs1[BK * T::LDB_S];

  const long long unit = xcd_unit(xs);
  if (unit < 0) return;
  const int p = find_problem(probs, nprob, unit);
  const GemmProblem P = probs[p];
  long long u = unit - P.unit_begin;
  const int ntile = P.tiles_m * P.tiles_n;
  const int ks = (int)(u / ntile);
  const int tile = (int)(u % ntile);
  const int tm = P.n_fast ? tile / P.tiles_n : tile % P.tiles_m;
  const int tn = P.n_fast ? tile % P.tiles_n : tile / P.tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const int kt_lo = (int)((long long)P.kt_total * ks / P.ksplit);
  const int kt_hi = (int)((long long)P.kt_total * (ks + 1) / P.ksplit);

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave % WM) * T::WTM;
  const int wn0 = (wave / WM) * T::WTN;

  typename Mma<E>::acc_t acc[T::FM][T::FN];
#pragma unroll
  for (int i = 0; i < T::FM; ++i)
#pragma unroll
    for (int j = 0; j < T::FN; ++j) acc[i][j] = typename Mma<E>::acc_t{E(0), E(0), E(0), E(0)};

  if (kt_lo < kt_hi) {
    int s = P.slice_begin;
    while (s + 1 < P.slice_end && slices[s + 1].kt_begin <= kt_lo) ++s;
    GemmSlice S = slices[s];
    int kl = kt_lo - S.kt_begin;
    int nkt = (S.K + BK - 1) / BK;
    SliceAddr ad = slice_addr<E, BM, BN, BK, WM, WN, PAD>(S, m0, n0, P.M, P.N);
    SlabRegs<T> regs;
    auto load = [&]() {
      if ((kl + 1) * BK <= S.K) load_full_u<E, BM, BN, BK, WM, WN, PAD, TA>(S, ad, kl * BK, regs);
      else load_tail_u<E, BM, BN, BK, WM, WN, PAD, TA>(S, kl * BK, m0, n0, P.M, P.N, regs);
    };
    load();
    store_u<E, BM, BN, BK, WM, WN, PAD, TA>(regs, (E)S.alpha, As0, Bs0);
    __syncthreads();

    // one slab from buffer CUR; the next slab (if any) is fetched into
    // registers first and stored to the other buffer before the last MFMAs
    auto step = [&](int cur, bool more) {
      const E* Ab = cur ? As1 : As0;
      const E* Bb = cur ? Bs1 : Bs0;
      E* An = cur ? As0 : As1;
      E* Bn = cur ? Bs0 : Bs1;
      if (more) {
        if (++kl >= nkt) {
          S = slices[++s];
          kl = 0;
          nkt = (S.K + BK - 1) / BK;
          ad = slice_addr<E, BM, BN, BK, WM, WN, PAD>(S, m0, n0, P.M, P.N);
        }
        load();
      }
      // Fragments of k-step kk + 1 are read while the MFMAs of kk run (two
      // register sets); the scheduling barriers keep the scheduler from
      // hoisting every step's reads to the top.  With no next slab the store
      // rewrites stale registers into the idle buffer, which no wave reads
      // again.
      E af[2][T::FM], bf[2][T::FN];
      auto frags = [&](int kk, int set) {
        const int kr = kk + (lane >> 4);
#pragma unroll
        for (int i = 0; i < T::FM; ++i)
          af[set][i] = Ab[kr * T::LDA_S + wm0 + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < T::FN; ++j)
          bf[set][j] = Bb[kr * T::LDB_S + wn0 + j * 16 + (lane & 15)];
      };
      frags(0, 0);
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const int set = (kk / 4) & 1;
        if (kk + 4 < BK) frags(kk + 4, set ^ 1);
        if (kk == BK - 4) store_u<E, BM, BN, BK, WM, WN, PAD, TA>(regs, (E)S.alpha, An, Bn);
#pragma unroll
        for (int i = 0; i < T::FM; ++i)
#pragma unroll
          for (int j = 0; j < T::FN; ++j)
            acc[i][j] = Mma<E>::run(af[set][i], bf[set][j], acc[i][j]);
        // issue order inside the step (masks: 0x8 MFMA, 0x100 LDS read,
        // 0x2 VALU, 0x200 LDS write): the next fragments right after the
        // first MFMA, so they have the whole group to arrive; in the last
        // step the scale-and-store is spread one MFMA apart
        if (kk + 4 < BK) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x100, (T::FM + T::FN + 1) / 2, 0);
          __builtin_amdgcn_sched_group_barrier(0x008, T::FM * T::FN - 1, 0);
        } else {
          constexpr int NQ = T::FM * T::FN < 8 ? T::FM * T::FN : 8;
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);
            __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
          }
          if (T::FM * T::FN > NQ)
            __builtin_amdgcn_sched_group_barrier(0x008, T::FM * T::FN - NQ, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      __syncthreads();
    };
    for (int kt = kt_lo; kt < kt_hi; ++kt) step((kt - kt_lo) & 1, kt + 1 < kt_hi);
  }

  E* dst;
  long long ld;
  if (P.ksplit == 1) {
    dst = (E*)P.C;
    ld = P.ldc;
  } else {
    dst = (E*)ws + P.ws_offset + (long long)ks * P.M * P.N;
    ld = P.M;
  }
#pragma unroll
  for (int i = 0; i < T::FM; ++i)
#pragma unroll
    for (int j = 0; j < T::FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = m0 + wm0 + i * 16 + Mma<E>::row(lane, r);
        int col = n0 + wn0 + j * 16 + (lane & 15);
        if (row < P.M && col < P.N) dst[row + (long long)col * ld] = acc[i][j][r];
      }
}

template <class E>
__global__ void splitk_reduce_kernel(const ReduceProblem* __restrict__ reds, int nred,
                                     const void* __restrict__ ws, long long total) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  int lo = 0, hi = nred - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (reds[mid].elem_begin <= e) lo = mid;
    else hi = mid - 1;
  }
  const ReduceProblem R = reds[lo];
  long long i = e - R.elem_begin;
  long long mn = (long long)R.M * R.N;
  const E* src = (const E*)ws + R.ws_offset + i;
  E s = E(0);
  for (int k = 0; k < R.ksplit; ++k) s += src[k * mn];
  int row = (int)(i % R.M), col = (int)(i / R.M);
  ((E*)R.C)[row + (long long)col * R.ldc] = s;
}

typedef void (*GemmKernel)(const GemmProblem*, int, const GemmSlice*, void*, XcdSplit);
typedef void (*ReduceKernel)(const ReduceProblem*, int, const void*, long long);

struct VariantInfo {
  GemmKernel fn;   // any mix of slice layouts
  GemmKernel fn_nn, fn_tn;  // every slice op(B) = B and op(A) = A / A^T
  int bm, bn, bk, threads;
};

#define SU2_VARIANT(E, BM, BN, BK, WM, WN, PAD)                                            \
  {grouped_gemm_f64_kernel<E, BM, BN, BK, WM, WN, PAD>,                                    \
   grouped_gemm_f64_uniform<E, BM, BN, BK, WM, WN, PAD, 0>,                                \
   grouped_gemm_f64_uniform<E, BM, BN, BK, WM, WN, PAD, 1>, BM, BN, BK, WM * WN * 64}

const VariantInfo kVariants64[] = {
    SU2_VARIANT(double, 128, 128, 16, 2, 2, 0),   // 0
    SU2_VARIANT(double, 128, 128, 16, 2, 2, 2),   // 1: padded LDS rows
    SU2_VARIANT(double, 256, 128, 16, 4, 2, 2),   // 2
    SU2_VARIANT(double, 128, 256, 16, 2, 4, 2),   // 3
    SU2_VARIANT(double, 64, 64, 16, 2, 2, 2),     // 4
    SU2_VARIANT(double, 128, 128, 32, 2, 2, 2),   // 5
    SU2_VARIANT(double, 128, 64, 16, 2, 2, 2),    // 6
    SU2_VARIANT(double, 256, 128, 16, 2, 2, 2),   // 7: 128 x 64 per wave, one wave per SIMD
    SU2_VARIANT(double, 128, 256, 16, 2, 2, 2),   // 8: 64 x 128 per wave
};
// FP32: LDS rows padded by 4 or 16 floats (16 floats shift consecutive
// k-rows of a fragment read by 16 banks, so a 64-lane ds_read_b32 that
// spans four rows is conflict-free)
const VariantInfo kVariants32[] = {
    SU2_VARIANT(float, 128, 128, 16, 2, 2, 4),    // 0
    SU2_VARIANT(float, 128, 128, 16, 2, 2, 16),   // 1
    SU2_VARIANT(float, 256, 128, 16, 4, 2, 16),   // 2
    SU2_VARIANT(float, 128, 256, 16, 2, 4, 16),   // 3
    SU2_VARIANT(float, 64, 64, 16, 2, 2, 16),     // 4
    SU2_VARIANT(float, 128, 128, 32, 2, 2, 16),   // 5
    SU2_VARIANT(float, 128, 64, 16, 2, 2, 16),    // 6
    SU2_VARIANT(float, 256, 128, 16, 2, 2, 16),   // 7: 128 x 64 per wave
    SU2_VARIANT(float, 128, 256, 16, 2, 2, 16),   // 8: 64 x 128 per wave
};
constexpr int kNumVariants64 = sizeof(kVariants64) / sizeof(kVariants64[0]);
constexpr int kNumVariants32 = sizeof(kVariants32) / sizeof(kVariants32[0]);

const VariantInfo& variant_info(Prec p, int v) {
  return p == Prec::F64 ? kVariants64[v] : kVariants32[v];
}
int num_variants(Prec p) { return p == Prec::F64 ? kNumVariants64 : kNumVariants32; }
size_t elem_size(Prec p) { return p == Prec::F64 ? sizeof(double) : sizeof(float); }

}  // namespace

int gemm_num_variants(Prec p) { return num_variants(p); }

GroupedGemm::~GroupedGemm() { release(); }

void GroupedGemm::release() {
  if (pool_) {
    pool_->free(d_probs_);   // one block holds the problem, slice and reduce tables
    pool_->free(d_ws_);
  }
  d_probs_ = nullptr;
  d_slices_ = nullptr;
  d_reds_ = nullptr;
  d_ws_ = nullptr;
  run_probs_.clear();
  reds_.clear();
  total_units_ = 0;
  total_red_elems_ = 0;
}

int GroupedGemm::add_problem(void* C, long long ldc, int M, int N) {
  SU2_CHECK(variant_ < 0, kErrArgument, "gemm plan already finalized");
  SU2_CHECK(M >= 0 && N >= 0 && ldc >= std::max(M, 1), kErrArgument, "gemm: bad problem shape");
  GemmProblem p{};
  p.C = C;
  p.ldc = ldc;
  p.M = M;
  p.N = N;
  p.slice_begin = p.slice_end = (int)slices_.size();
  probs_.push_back(p);
  return (int)probs_.size() - 1;
}

void GroupedGemm::add_slice(int problem, const void* A, long long lda, int transA, const void* B,
                            long long ldb, int transB, int K, double alpha) {
  SU2_CHECK(variant_ < 0, kErrArgument, "gemm plan already finalized");
  SU2_CHECK(problem == (int)probs_.size() - 1, kErrArgument,
            "gemm: slices must follow their problem");
  GemmProblem& p = probs_[problem];
  SU2_CHECK(K > 0, kErrArgument, "gemm: empty slice");
  SU2_CHECK(lda >= (transA ? K : std::max(p.M, 1)) && ldb >= (transB ? std::max(p.N, 1) : K),
            kErrArgument, "gemm: leading dimension too small");
  // the kernel addresses a slice with unsigned 32-bit byte offsets (plus one
  // tile of margin past the M / N edge)
  const long long lim = (1LL << 32) / (long long)elem_size(prec_) - 1;
  const long long a_span = transA ? (long long)(p.M + 256) * lda + K : (long long)(K + 256) * lda + p.M;
  const long long b_span = transB ? (long long)(K + 256) * ldb + p.N : (long long)(p.N + 256) * ldb + K;
  SU2_CHECK(a_span < lim && b_span < lim, kErrCapacity,
            "gemm: slice matrix over 4 GB (M=%d N=%d K=%d)", p.M, p.N, K);
  GemmSlice sl{};
  sl.A = A;
  sl.B = B;
  sl.lda = lda;
  sl.ldb = ldb;
  sl.K = K;
  sl.transA = transA;
  sl.transB = transB;
  sl.alpha = alpha;
  slices_.push_back(sl);
  p.slice_end = (int)slices_.size();
  flops_ += 2.0 * p.M * (double)p.N * K;
}

void GroupedGemm::append(const GroupedGemm& o) {
  SU2_CHECK(variant_ < 0, kErrArgument, "gemm plan already finalized");
  SU2_CHECK(o.prec_ == prec_, kErrArgument, "gemm: appending a plan of another precision");
  const int base = (int)slices_.size();
  for (const GemmProblem& q : o.probs_) {
    GemmProblem p = q;
    p.slice_begin += base;
    p.slice_end += base;
    probs_.push_back(p);
  }
  slices_.insert(slices_.end(), o.slices_.begin(), o.slices_.end());
  flops_ += o.flops_;
}

void GroupedGemm::finalize(Context& c, int min_units, int kt_per_unit, int variant) {
  SU2_CHECK(variant_ < 0, kErrArgument, "gemm plan already finalized");
  pool_ = c.pool();
  for (auto& p : probs_)
    SU2_CHECK(p.slice_end > p.slice_begin, kErrArgument, "gemm: problem without K slices");
  min_units_ = std::max(1, min_units);
  kt_per_unit_ = std::max(1, kt_per_unit);
  stream_ = c.stream();
  set_variant(variant < 0 ? kDefaultGemmVariant : variant);
}

void GroupedGemm::set_variant(int v) {
  SU2_CHECK(v >= 0 && v < num_variants(prec_), kErrArgument, "gemm variant %d out of range", v);
  const VariantInfo& V = variant_info(prec_, v);
  release();
  variant_ = v;
  // one layout for every slice selects the straight-line kernel
  layout_ = -1;
  if (uniform_ && !slices_.empty()) {
    bool nn = true, tn = true;
    for (auto& x : slices_) {
      nn = nn && x.transB == 0 && x.transA == 0;
      tn = tn && x.transB == 0 && x.transA == 1;
    }
    layout_ = nn ? 0 : tn ? 1 : -1;
  }
  std::vector<GemmSlice> sl = slices_;
  std::vector<GemmProblem> probs = probs_;
  for (auto& p : probs) {
    int kt = 0;
    for (int s = p.slice_begin; s < p.slice_end; ++s) {
      sl[s].kt_begin = kt;
      kt += (sl[s].K + V.bk - 1) / V.bk;
    }
    p.kt_total = kt;
    p.tiles_m = (p.M + V.bm - 1) / V.bm;
    p.tiles_n = (p.N + V.bn - 1) / V.bn;
    // keep the larger operand's panel fixed across neighbouring tiles, so it
    // streams from HBM about once while the smaller one is re-read from cache
    p.n_fast = p.M > p.N ? 1 : 0;
    // K split from this problem's own shape only (see finalize)
    const long long tiles = std::max(1LL, (long long)p.tiles_m * p.tiles_n);
    const long long by_units = (min_units_ + tiles - 1) / tiles;
    const long long by_depth = p.kt_total / kt_per_unit_;
    const long long ks = std::max(by_units, by_depth);
    p.ksplit = (int)std::max(1LL, std::min(ks, std::max(1LL, (long long)p.kt_total / 8)));
  }
  // units in order of decreasing cost (stable), so that the cheap units of a
  // launch come last in the XCD ranges and fill their tails
  std::stable_sort(probs.begin(), probs.end(), [](const GemmProblem& a, const GemmProblem& b) {
    return (a.kt_total + a.ksplit - 1) / a.ksplit > (b.kt_total + b.ksplit - 1) / b.ksplit;
  });
  long long units = 0, ws = 0;
  for (auto& p : probs) {
    if (p.M == 0 || p.N == 0) continue;
    const int ks = p.ksplit;
    p.unit_begin = units;
    units += (long long)p.tiles_m * p.tiles_n * ks;
    if (ks > 1) {
      p.ws_offset = ws;
      ReduceProblem r{};
      r.C = p.C;
      r.ldc = p.ldc;
      r.M = p.M;
      r.N = p.N;
      r.ksplit = ks;
      r.ws_offset = ws;
      r.elem_begin = total_red_elems_;
      total_red_elems_ += (long long)p.M * p.N;
      reds_.push_back(r);
      ws += (long long)ks * p.M * p.N;
    }
    run_probs_.push_back(p);
  }
  total_units_ = units;
  SU2_CHECK(total_units_ < (1ll << 31), kErrCapacity, "gemm: too many work units");
  // XCD ranges of equal work: a unit of problem p costs its k-tiles
  // ceil(kt_total / ksplit) plus about two for its prologue and epilogue
  {
    std::vector<double> w(run_probs_.size()), cum(run_probs_.size() + 1, 0.0);
    for (size_t i = 0; i < run_probs_.size(); ++i) {
      const GemmProblem& p = run_probs_[i];
      w[i] = std::ceil((double)p.kt_total / p.ksplit) + 2.0;
      cum[i + 1] = cum[i] + w[i] * p.tiles_m * p.tiles_n * p.ksplit;
    }
    split_.begin[0] = 0;
    size_t i = 0;
    for (int x = 1; x < 8; ++x) {
      const double target = cum.back() * x / 8.0;
      while (i + 1 < run_probs_.size() && cum[i + 1] <= target) ++i;
      long long u = total_units_;
      if (!run_probs_.empty())
        u = run_probs_[i].unit_begin + (long long)std::ceil((target - cum[i]) / w[i]);
      split_.begin[x] = std::max(split_.begin[x - 1], std::min(u, total_units_));
    }
    split_.begin[8] = total_units_;
    long long longest = 0;
    for (int x = 0; x < 8; ++x) longest = std::max(longest, split_.begin[x + 1] - split_.begin[x]);
    grid_ = 8 * longest;
    SU2_CHECK(grid_ < (1ll << 31), kErrCapacity, "gemm: grid of %lld workgroups", grid_);
  }
  // the three tables in one block, uploaded with one copy on the plan's
  // stream (ordered before every launch of the plan)
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t bp = al(run_probs_.size() * sizeof(GemmProblem));
  const size_t bs = al(sl.size() * sizeof(GemmSlice));
  const size_t br = reds_.size() * sizeof(ReduceProblem);
  if (bp + bs + br > 0) {
    std::vector<char> host(bp + bs + br, 0);
    std::memcpy(host.data(), run_probs_.data(), run_probs_.size() * sizeof(GemmProblem));
    std::memcpy(host.data() + bp, sl.data(), sl.size() * sizeof(GemmSlice));
    std::memcpy(host.data() + bp + bs, reds_.data(), br);
    char* d = (char*)pool_->alloc(host.size());
    SU2_HIP(hipMemcpyAsync(d, host.data(), host.size(), hipMemcpyHostToDevice, stream_));
    d_probs_ = run_probs_.empty() ? nullptr : (GemmProblem*)d;
    d_slices_ = sl.empty() ? nullptr : (GemmSlice*)(d + bp);
    d_reds_ = reds_.empty() ? nullptr : (ReduceProblem*)(d + bp + bs);
    if (!d_probs_) d_probs_ = (GemmProblem*)d;   // the block is freed through d_probs_
  }
  if (!reds_.empty()) d_ws_ = pool_->alloc(ws * elem_size(prec_));
}

void GroupedGemm::launch(hipStream_t stream) const {
  SU2_CHECK(variant_ >= 0, kErrArgument, "gemm plan not finalized");
  if (total_units_ == 0) return;
  const VariantInfo& V = variant_info(prec_, variant_);
  const GemmKernel fn = layout_ == 0 ? V.fn_nn : layout_ == 1 ? V.fn_tn : V.fn;
  hipLaunchKernelGGL(fn, dim3((unsigned)grid_), dim3(V.threads), 0, stream, d_probs_,
                     (int)run_probs_.size(), d_slices_, d_ws_, split_);
  SU2_HIP(hipGetLastError());
  if (!reds_.empty()) {
    long long blocks = (total_red_elems_ + 255) / 256;
    const ReduceKernel rk =
        prec_ == Prec::F64 ? splitk_reduce_kernel<double> : splitk_reduce_kernel<float>;
    hipLaunchKernelGGL(rk, dim3((unsigned)blocks), dim3(256), 0, stream, d_reds_,
                       (int)reds_.size(), (const void*)d_ws_, total_red_elems_);
    SU2_HIP(hipGetLastError());
  }
}

float GroupedGemm::time_ms(hipStream_t stream, int reps) const {
  hipEvent_t e0, e1;
  SU2_HIP(hipEventCreate(&e0));
  SU2_HIP(hipEventCreate(&e1));
  launch(stream);  // warm
  SU2_HIP(hipEventRecord(e0, stream));
  for (int r = 0; r < reps; ++r) launch(stream);
  SU2_HIP(hipEventRecord(e1, stream));
  SU2_HIP(hipEventSynchronize(e1));
  float ms = 0;
  SU2_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

int GroupedGemm::autotune(hipStream_t stream, int reps, std::vector<float>* times) {
  SU2_CHECK(variant_ >= 0, kErrArgument, "gemm plan not finalized");
  int best = variant_;
  float best_ms = 1e30f;
  if (times) times->assign(num_variants(prec_), 0.f);
  for (int v = 0; v < num_variants(prec_); ++v) {
    set_variant(v);
    float ms = time_ms(stream, reps);
    if (times) (*times)[v] = ms;
    if (ms < best_ms) {
      best_ms = ms;
      best = v;
    }
  }
  set_variant(best);
  return best;
}

}  // namespace su2
