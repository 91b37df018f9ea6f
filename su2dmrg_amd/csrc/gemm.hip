// Grouped FP64 MFMA GEMM for gfx950.  See gemm.hpp for the problem model.
//
// Workgroup tile BM x BN (128 x 128), 4 waves in a 2 x 2 arrangement, each
// wave owning a 64 x 64 sub-tile = 4 x 4 blocks of v_mfma_f64_16x16x4_f64.
// K is staged through LDS in BK=16 slabs, double buffered: the next slab is
// fetched from global memory into registers while the current one feeds the
// matrix cores, then written to the other LDS buffer (one barrier per slab).
//
// MFMA operand layout (16x16x4 f64, 64 lanes):
//   A (16 x 4):  lane l holds A[l % 16][l / 16]
//   B (4 x 16):  lane l holds B[l / 16][l % 16]
//   D (16 x 16): lane l holds D[(l / 16) + 4*r][l % 16], r = 0..3
// (the D layout was confirmed on the MI355X with a one-wave probe)
#include "gemm.hpp"

#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace su2 {

typedef double f64x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int BM = kGemmBM;
constexpr int BN = kGemmBN;
constexpr int BK = kGemmBK;
constexpr int WAVES_M = 2;
constexpr int WAVES_N = 2;
constexpr int NT = WAVES_M * WAVES_N * 64;
constexpr int WTM = BM / WAVES_M;  // 64
constexpr int WTN = BN / WAVES_N;  // 64
constexpr int FM = WTM / 16;       // 4
constexpr int FN = WTN / 16;       // 4
constexpr int A_PER_T = BM * BK / NT;  // 8
constexpr int B_PER_T = BN * BK / NT;  // 8

// Logical work-unit id of this workgroup: the hardware deals workgroups to
// the 8 XCDs round-robin (id % 8); give each XCD a contiguous run of logical
// units so neighbouring tiles (which share A or B panels) meet in one L2.
__device__ __forceinline__ long long xcd_remap(long long bid, long long nb) {
  long long q = nb / 8, r = nb % 8;
  long long x = bid % 8, local = bid / 8;
  return x * q + (x < r ? x : r) + local;
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

struct SlabRegs {
  double a[A_PER_T];
  double b[B_PER_T];
};

__device__ __forceinline__ void load_slab(const GemmSlice& s, int kt_local, int m0, int n0, int M,
                                          int N, SlabRegs& r) {
  const int t = threadIdx.x;
  const int k0 = kt_local * BK;
  const double alpha = s.alpha;
  if (!s.transA) {
    // op(A) = A, column-major M x K: coalesced along m.
#pragma unroll
    for (int i = 0; i < A_PER_T; ++i) {
      int e = t + i * NT;
      int m = e % BM, k = e / BM;
      int gm = m0 + m, gk = k0 + k;
      r.a[i] = (gm < M && gk < s.K) ? alpha * s.A[gm + (long long)gk * s.lda] : 0.0;
    }
  } else {
    // op(A) = A^T, A stored K x M: coalesced along k.
#pragma unroll
    for (int i = 0; i < A_PER_T; ++i) {
      int e = t + i * NT;
      int k = e % BK, m = e / BK;
      int gm = m0 + m, gk = k0 + k;
      r.a[i] = (gm < M && gk < s.K) ? alpha * s.A[gk + (long long)gm * s.lda] : 0.0;
    }
  }
  if (!s.transB) {
    // B stored K x N: coalesced along k.
#pragma unroll
    for (int i = 0; i < B_PER_T; ++i) {
      int e = t + i * NT;
      int k = e % BK, n = e / BK;
      int gn = n0 + n, gk = k0 + k;
      r.b[i] = (gn < N && gk < s.K) ? s.B[gk + (long long)gn * s.ldb] : 0.0;
    }
  } else {
    // op(B) = B^T, B stored N x K: coalesced along n.
#pragma unroll
    for (int i = 0; i < B_PER_T; ++i) {
      int e = t + i * NT;
      int n = e % BN, k = e / BN;
      int gn = n0 + n, gk = k0 + k;
      r.b[i] = (gn < N && gk < s.K) ? s.B[gn + (long long)gk * s.ldb] : 0.0;
    }
  }
}

__device__ __forceinline__ void store_slab(const GemmSlice& s, const SlabRegs& r,
                                           double (*As)[BM], double (*Bs)[BN]) {
  const int t = threadIdx.x;
  if (!s.transA) {
#pragma unroll
    for (int i = 0; i < A_PER_T; ++i) {
      int e = t + i * NT;
      As[e / BM][e % BM] = r.a[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < A_PER_T; ++i) {
      int e = t + i * NT;
      As[e % BK][e / BK] = r.a[i];
    }
  }
  if (!s.transB) {
#pragma unroll
    for (int i = 0; i < B_PER_T; ++i) {
      int e = t + i * NT;
      Bs[e % BK][e / BK] = r.b[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < B_PER_T; ++i) {
      int e = t + i * NT;
      Bs[e / BN][e % BN] = r.b[i];
    }
  }
}

__global__ __launch_bounds__(NT) void grouped_gemm_f64_kernel(
    const GemmProblem* __restrict__ probs, int nprob, const GemmSlice* __restrict__ slices,
    double* __restrict__ ws, long long total_units) {
  __shared__ double As[2][BK][BM];
  __shared__ double Bs[2][BK][BN];

  const long long unit = xcd_remap(blockIdx.x, gridDim.x);
  if (unit >= total_units) return;
  const int p = find_problem(probs, nprob, unit);
  const GemmProblem P = probs[p];
  long long u = unit - P.unit_begin;
  const int ntile = P.tiles_m * P.tiles_n;
  const int ks = (int)(u / ntile);
  const int tile = (int)(u % ntile);
  const int tm = tile % P.tiles_m;
  const int tn = tile / P.tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const int kt_lo = (int)((long long)P.kt_total * ks / P.ksplit);
  const int kt_hi = (int)((long long)P.kt_total * (ks + 1) / P.ksplit);

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave % WAVES_M) * WTM;
  const int wn0 = (wave / WAVES_M) * WTN;

  f64x4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

  if (kt_lo < kt_hi) {
    // locate the slice holding k-tile kt_lo
    int s = P.slice_begin;
    while (s + 1 < P.slice_end && slices[s + 1].kt_begin <= kt_lo) ++s;
    GemmSlice S = slices[s];
    int kl = kt_lo - S.kt_begin;

    SlabRegs regs;
    load_slab(S, kl, m0, n0, P.M, P.N, regs);
    store_slab(S, regs, As[0], Bs[0]);
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
        }
        load_slab(Sn, kln, m0, n0, P.M, P.N, regs);
      }
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        double af[FM], bf[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = As[buf][kk + (lane >> 4)][wm0 + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < FN; ++j) bf[j] = Bs[buf][kk + (lane >> 4)][wn0 + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
      }
      if (more) {
        store_slab(Sn, regs, As[buf ^ 1], Bs[buf ^ 1]);
        S = Sn;
        kl = kln;
      }
      __syncthreads();
      buf ^= 1;
    }
  }

  // epilogue: C directly, or this split's partial in the workspace
  double* dst;
  long long ld;
  if (P.ksplit == 1) {
    dst = P.C;
    ld = P.ldc;
  } else {
    dst = ws + P.ws_offset + (long long)ks * P.M * P.N;
    ld = P.M;
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = m0 + wm0 + i * 16 + (lane >> 4) + 4 * r;
        int col = n0 + wn0 + j * 16 + (lane & 15);
        if (row < P.M && col < P.N) dst[row + (long long)col * ld] = acc[i][j][r];
      }
}

__global__ void splitk_reduce_kernel(const ReduceProblem* __restrict__ reds, int nred,
                                     const double* __restrict__ ws, long long total) {
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
  const double* src = ws + R.ws_offset + i;
  double s = 0.0;
  for (int k = 0; k < R.ksplit; ++k) s += src[k * mn];
  int row = (int)(i % R.M), col = (int)(i / R.M);
  R.C[row + (long long)col * R.ldc] = s;
}

}  // namespace

GroupedGemm::~GroupedGemm() {
  if (d_probs_) (void)hipFree(d_probs_);
  if (d_slices_) (void)hipFree(d_slices_);
  if (d_reds_) (void)hipFree(d_reds_);
  if (d_ws_) (void)hipFree(d_ws_);
}

int GroupedGemm::add_problem(double* C, long long ldc, int M, int N) {
  SU2_CHECK(!finalized_, kErrArgument, "gemm plan already finalized");
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

void GroupedGemm::add_slice(int problem, const double* A, long long lda, int transA,
                            const double* B, long long ldb, int transB, int K, double alpha) {
  SU2_CHECK(problem == (int)probs_.size() - 1, kErrArgument,
            "gemm: slices must follow their problem");
  GemmProblem& p = probs_[problem];
  SU2_CHECK(K > 0, kErrArgument, "gemm: empty slice");
  SU2_CHECK(lda >= (transA ? K : std::max(p.M, 1)) && ldb >= (transB ? std::max(p.N, 1) : K),
            kErrArgument, "gemm: leading dimension too small");
  GemmSlice s{};
  s.A = A;
  s.B = B;
  s.lda = lda;
  s.ldb = ldb;
  s.K = K;
  s.transA = transA;
  s.transB = transB;
  s.alpha = alpha;
  slices_.push_back(s);
  p.slice_end = (int)slices_.size();
  flops_ += 2.0 * p.M * (double)p.N * K;
}

void GroupedGemm::finalize(int target_units) {
  SU2_CHECK(!finalized_, kErrArgument, "gemm plan already finalized");
  // drop empty problems (M or N zero); problems without slices are zero-filled
  // by the caller before launch, they are rejected here.
  double work = 0;
  for (auto& p : probs_) {
    SU2_CHECK(p.slice_end > p.slice_begin, kErrArgument, "gemm: problem without K slices");
    int kt = 0;
    for (int s = p.slice_begin; s < p.slice_end; ++s) {
      slices_[s].kt_begin = kt;
      kt += (slices_[s].K + BK - 1) / BK;
    }
    p.kt_total = kt;
    p.tiles_m = (p.M + BM - 1) / BM;
    p.tiles_n = (p.N + BN - 1) / BN;
    work += (double)p.tiles_m * p.tiles_n * kt;
  }
  // split K so that the launch has about target_units comparable work units
  double unit_work = std::max(1.0, work / std::max(1, target_units));
  long long units = 0, ws = 0;
  std::vector<GemmProblem> kept;
  for (auto& p : probs_) {
    if (p.M == 0 || p.N == 0) continue;
    int ks = (int)std::floor((double)p.kt_total / unit_work);
    ks = std::max(1, std::min(ks, std::max(1, p.kt_total / 8)));
    p.ksplit = ks;
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
    kept.push_back(p);
  }
  probs_.swap(kept);
  total_units_ = units;
  if (!probs_.empty()) {
    SU2_HIP(hipMalloc(&d_probs_, probs_.size() * sizeof(GemmProblem)));
    SU2_HIP(hipMemcpy(d_probs_, probs_.data(), probs_.size() * sizeof(GemmProblem),
                      hipMemcpyHostToDevice));
  }
  if (!slices_.empty()) {
    SU2_HIP(hipMalloc(&d_slices_, slices_.size() * sizeof(GemmSlice)));
    SU2_HIP(hipMemcpy(d_slices_, slices_.data(), slices_.size() * sizeof(GemmSlice),
                      hipMemcpyHostToDevice));
  }
  if (!reds_.empty()) {
    SU2_HIP(hipMalloc(&d_reds_, reds_.size() * sizeof(ReduceProblem)));
    SU2_HIP(hipMemcpy(d_reds_, reds_.data(), reds_.size() * sizeof(ReduceProblem),
                      hipMemcpyHostToDevice));
    SU2_HIP(hipMalloc(&d_ws_, ws * sizeof(double)));
  }
  finalized_ = true;
}

void GroupedGemm::launch(hipStream_t stream) const {
  SU2_CHECK(finalized_, kErrArgument, "gemm plan not finalized");
  if (total_units_ == 0) return;
  SU2_CHECK(total_units_ < (1ll << 31), kErrCapacity, "gemm: too many work units");
  hipLaunchKernelGGL(grouped_gemm_f64_kernel, dim3((unsigned)total_units_), dim3(NT), 0, stream,
                     d_probs_, (int)probs_.size(), d_slices_, d_ws_, total_units_);
  SU2_HIP(hipGetLastError());
  if (!reds_.empty()) {
    long long blocks = (total_red_elems_ + 255) / 256;
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                       d_reds_, (int)reds_.size(), d_ws_, total_red_elems_);
    SU2_HIP(hipGetLastError());
  }
}

}  // namespace su2
