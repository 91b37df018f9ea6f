// Preconditioner of the one-sided Jacobi SVD, see precond.hpp.
#include "precond.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "common.hpp"
#include "context.hpp"
#include "ops.hpp"

namespace su2 {

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int PB = 16;        // panel width
constexpr int PT = 1024;      // threads of the panel kernels
constexpr int PROWS = 2;      // rows (or columns) per thread of the panel kernels
static_assert(PT * PROWS == kPrecondMaxK, "panel register tile");

struct PDesc {
  double* G;   // K x K, column-major: Gram matrix, then the Cholesky factor Rc
               // (upper, zeros below).  Read row-major it is X = Rc^T (X[i][j]
               // at G[j + i K]), which the Householder QR then overwrites
               // with its vectors (below X's diagonal)
  double* Q;   // K x K, column-major: Q2
  double* T;   // 256 per panel: the panel's 16 x 16 upper triangular T (column-major)
  int K;
  int rk;      // rows of Rc up to its last nonzero pivot (X's later columns are zero)
  double tiny; // Cholesky pivot floor
};

// Compile-time loop: f(std::integral_constant<int, k>) for k = K0 .. K1-1, so
// that register arrays are indexed by constants only (a runtime index would
// send them to scratch memory).
template <int K0, int K1, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (K0 < K1) {
    f(std::integral_constant<int, K0>{});
    static_for<K0 + 1, K1>(f);
  }
}

// Upper Cholesky of G by one 16-row panel: rows c0..c0+bp-1, columns
// c0..K-1, column j of the row panel on thread j % 1024, in registers.  A
// pivot at or below tiny (or
// negative from rounding) gives a zero row.  One barrier per step: row k is
// published before its pivot is used, into alternating LDS buffers.
__global__ __launch_bounds__(PT) void chol_panel_kernel(const PDesc* __restrict__ desc, int p) {
  const PDesc D = desc[blockIdx.x];
  const int K = D.K, c0 = PB * p;
  if (c0 >= K) return;
  const int bp = min(PB, K - c0), nc = K - c0;
  const int t = threadIdx.x;
  __shared__ double s_raw[2][PB];
  double w[PROWS][PB];
  double* gcol[PROWS];
#pragma unroll
  for (int q = 0; q < PROWS; ++q) {
    const int j = t + PT * q;
    gcol[q] = D.G + (size_t)(c0 + min(j, nc - 1)) * K + c0;
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int j = t + PT * q;
#pragma unroll
    for (int i = 0; i < PB; ++i) w[q][i] = (j < nc && i < bp) ? gcol[q][i] : 0.0;
  }
  static_for<0, PB>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    __builtin_amdgcn_sched_barrier(0);
    if (k < bp) {
      double* raw = s_raw[k & 1];
      // row k as it stands (columns c0 + k .. c0 + 15 live on threads k..15)
      if (t >= k && t < PB) raw[t] = w[0][k];
      __syncthreads();
      const double dk = raw[k];
      const double inv = dk > D.tiny ? 1.0 / sqrt(dk) : 0.0;
      // R[k][j] for my columns (zero left of the diagonal), then the rows
      // below: w[i][j] -= R[k][i] R[k][j] for k < i <= j (row by row, so
      // that each R[k][i] is live for one row only)
      double rkj[PROWS];
#pragma unroll
      for (int q = 0; q < PROWS; ++q) {
        const int j = t + PT * q;
        rkj[q] = (j >= k) ? w[q][k] * inv : 0.0;
        w[q][k] = rkj[q];
      }
#pragma unroll
      for (int i = k + 1; i < PB; ++i) {
        if (i >= bp) continue;
        const double f = -raw[i] * inv;
#pragma unroll
        for (int q = 0; q < PROWS; ++q)
          if (i <= t + PT * q) w[q][i] = fma(f, rkj[q], w[q][i]);
      }
    }
  });
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int j = t + PT * q;
    if (j >= nc) continue;
#pragma unroll
    for (int i = 0; i < PB; ++i)
      if (i < bp) gcol[q][i] = (i > j) ? 0.0 : w[q][i];
  }
  // the panel's columns below its diagonal block: zero (Rc is upper
  // triangular, and read row-major they are X's entries above its diagonal)
  for (int e = t; e < bp * (nc - bp); e += PT) {
    const int i = e / (nc - bp), r = e % (nc - bp);
    D.G[(size_t)(c0 + bp + r) + (size_t)(c0 + i) * K] = 0.0;
  }
}

// Trailing update of the Cholesky: G[i][j] -= sum_k R[c0+k][i] R[c0+k][j]
// for c1 <= i <= j (upper triangle only; the panel kernels read nothing
// below it).  64 x 64 tiles, k summed in order.
__global__ __launch_bounds__(256) void chol_trail_kernel(const PDesc* __restrict__ desc, int p) {
  const PDesc D = desc[blockIdx.z];
  const int K = D.K, c0 = PB * p, c1 = c0 + PB;
  if (c1 >= K) return;
  const int ti = blockIdx.x, tj = blockIdx.y;
  const int n = K - c1;
  if (ti > tj || ti * 64 >= n || tj * 64 >= n) return;
  __shared__ double Ri[PB][64], Rj[PB][64];
  const int t = threadIdx.x;
  for (int e = t; e < PB * 64; e += 256) {
    const int k = e / 64, x = e % 64;
    const int gi = c1 + ti * 64 + x, gj = c1 + tj * 64 + x;
    Ri[k][x] = gi < K ? D.G[(size_t)(c0 + k) + (size_t)gi * K] : 0.0;
    Rj[k][x] = gj < K ? D.G[(size_t)(c0 + k) + (size_t)gj * K] : 0.0;
  }
  __syncthreads();
  const int x = t % 64;                 // row within the tile (contiguous in memory)
  const int gi = c1 + ti * 64 + x;
  for (int s = 0; s < 16; ++s) {
    const int y = t / 64 + 4 * s;       // column within the tile
    const int gj = c1 + tj * 64 + y;
    if (gi < K && gj < K && gi <= gj) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < PB; ++k) acc = fma(Ri[k][x], Rj[k][y], acc);
      D.G[(size_t)gi + (size_t)gj * K] -= acc;
    }
  }
}

// Householder QR of one 16-column panel of X (rows c0..K-1; row c0 + r on
// thread r % 1024, in registers).
// Column j becomes its Householder vector below the diagonal (the unit
// diagonal implicit); the panel's T factor of the compact WY form H_0 ...
// H_{b-1} = I - V T V^T goes to D.T.  Each step takes one block reduction of
// the 16 products d_c = sum_{r>j} x_j(r) x_c(r): with row j, they give
// sigma^2 = d_j, w_c = v_j^T x_c = x_c(j) + scale d_c (c > j) and
// y_c = v_c^T v_j = x_c(j) + scale d_c (c < j, for T).
__global__ __launch_bounds__(PT) void hh_panel_kernel(const PDesc* __restrict__ desc, int p) {
  const PDesc D = desc[blockIdx.x];
  const int K = D.K, c0 = PB * p;
  // a panel of zero columns (past the last nonzero pivot) is the identity
  if (c0 >= D.rk) return;
  const int bp = min(PB, K - c0), m = K - c0;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  __shared__ double s_red[2][PT / 64][PB + 1];
  __shared__ double s_row[2][PB];
  __shared__ double s_d[2][PB];
  __shared__ double s_T[PB][PB];
  double x[2][PB];
  double* xrow[PROWS];
#pragma unroll
  for (int s = 0; s < PROWS; ++s) {
    const int r = t + PT * s;
    xrow[s] = D.G + (size_t)(c0 + min(r, m - 1)) * K + c0;
  }
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int r = t + PT * s;
#pragma unroll
    for (int c = 0; c < PB; ++c) x[s][c] = (r < m && c < bp) ? xrow[s][c] : 0.0;
  }
  if (t < PB * PB) s_T[t % PB][t / PB] = 0.0;
  static_for<0, PB>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    __builtin_amdgcn_sched_barrier(0);
    if (j < bp) {
      const int b = j & 1;
      // d_c over my rows below the diagonal, summed over each wave by a
      // reduce-scatter (lanes that differ in bit 32, 16, 8, 4 swap halves of
      // their 16, 8, 4, 2 sums, then the last sum is reduced over bits 2, 1):
      // 17 exchanges instead of 16 x 6
      double v[PB];
#pragma unroll
      for (int c = 0; c < PB; ++c) {
        v[c] = 0.0;
        if (c < bp) {
#pragma unroll
          for (int s = 0; s < 2; ++s) {
            const int r = t + PT * s;
            if (r > j && r < m) v[c] = fma(x[s][j], x[s][c], v[c]);
          }
        }
      }
      static_assert(PB == 16, "reduce-scatter of 16 sums over a 64-lane wave");
#pragma unroll
      for (int h = 8, o = 32; h >= 1; h >>= 1, o >>= 1) {
        // keep the upper h sums on lanes with bit o set, the lower h otherwise
        const bool up = (lane & o) != 0;
#pragma unroll
        for (int i = 0; i < h; ++i) {
          const double send = up ? v[i] : v[h + i];
          const double keep = up ? v[h + i] : v[i];
          v[i] = keep + __shfl_xor(send, o);
        }
      }
      v[0] += __shfl_xor(v[0], 2);
      v[0] += __shfl_xor(v[0], 1);
      if ((lane & 3) == 0) {
        const int c = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                      ((lane >> 2) & 1);
        s_red[b][wave][c] = v[0];
      }
      if (t == j) {
#pragma unroll
        for (int c = 0; c < PB; ++c) s_row[b][c] = x[0][c];
      }
      __syncthreads();
      if (t < PB) {
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < PT / 64; ++w) v += s_red[b][w][t];
        s_d[b][t] = v;
      }
      __syncthreads();
      // the reflector, computed alike on every thread
      const double alpha = s_row[b][j], sig2 = s_d[b][j];
      double tau = 0.0, scale = 0.0;
      if (sig2 != 0.0) {
        const double beta = -copysign(sqrt(fma(alpha, alpha, sig2)), alpha);
        tau = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      // T column j (T[t][k], k < j, was written by this same thread)
      if (t < j) {
        double acc = 0.0;
        for (int k = t; k < j; ++k) acc = fma(s_T[t][k], fma(scale, s_d[b][k], s_row[b][k]), acc);
        s_T[t][j] = -tau * acc;
      }
      if (t == j) s_T[j][j] = tau;
      // v_j into column j, H_j applied to the later columns (rows >= j);
      // w_c = v_j^T x_c read from LDS once for all rows
      double wc[PB];
#pragma unroll
      for (int c = j + 1; c < PB; ++c) wc[c] = fma(scale, s_d[b][c], s_row[b][c]);
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int r = t + PT * s;
        if (r >= j && r < m) {
          const double vr = (r == j) ? 1.0 : x[s][j] * scale;
          if (r > j) x[s][j] = vr;
#pragma unroll
          for (int c = j + 1; c < PB; ++c)
            if (c < bp) x[s][c] = fma(-tau * vr, wc[c], x[s][c]);
        }
      }
    }
  });
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int r = t + PT * s;
    if (r >= m) continue;
#pragma unroll
    for (int c = 0; c < PB; ++c)
      if (c < bp) xrow[s][c] = x[s][c];
  }
  if (t < PB * PB) D.T[(size_t)p * PB * PB + t] = s_T[t % PB][t / PB];
}

// Householder vector i of panel p at panel row r (the unit diagonal implicit).
__device__ __forceinline__ double hh_v(const PDesc& D, int c0, int bp, int m, int i, int r) {
  if (i >= bp || r >= m || r < i) return 0.0;
  if (r == i) return 1.0;
  return D.G[(size_t)(c0 + r) * D.K + (c0 + i)];
}

// A <- A - V op(T) (V^T A) for 16 columns of A (rows c0..K-1): A = X from
// column c0 + bp (the QR's trailing columns, op(T) = T^T) or A = Q from
// column c0 (Q2's backward formation, op(T) = T).  One workgroup owns its 16
// columns over every row, so the products and the update need no other
// workgroup.  1024 threads: 256-row chunks of V and A staged through LDS,
// each of the 16 waves accumulating V^T A over its 16 rows of every chunk
// on the matrix cores (four 16x16x4 FP64 MFMAs), the 16 partial products
// summed in wave order, Z = op(T) sums in LDS; then A -= V Z by the same
// waves, a 16 x 16 tile each on the matrix cores.
__global__ __launch_bounds__(1024) void hh_apply_kernel(const PDesc* __restrict__ desc, int p,
                                                        int qr) {
  const PDesc D = desc[blockIdx.y];
  const int K = D.K, c0 = PB * p;
  if (c0 >= D.rk) return;   // identity reflectors
  const int bp = min(PB, K - c0), m = K - c0;
  // X's columns from rk on are zero and stay zero
  const int colstart = c0 + (qr ? bp : 0), ncols = (qr ? D.rk : K) - colstart;
  const int chunk = blockIdx.x * PB;
  if (chunk >= ncols) return;
  // A(row, col): X row-major (the QR's trailing columns) or Q column-major
  double* A = qr ? D.G : D.Q;
  const size_t rs = qr ? (size_t)K : 1, cs = qr ? 1 : (size_t)K;
  constexpr int NG = 4;
  __shared__ double sv[NG][64][PB + 1], sa[NG][64][PB + 1];
  __shared__ double sp[16][PB][PB + 1];   // the waves' partial V^T A
  __shared__ double sz[PB][PB + 1];
  static_assert(PB == 16 && NG * 4 == 16, "16 waves of 16 rows per 256-row chunk");
  const int t = threadIdx.x, g = t / 256, tg = t % 256;
  const int i = tg / PB, c = tg % PB;
  const int lane = t & 63, wave = t >> 6, gw = wave >> 2, rw = (wave & 3) * 16;
  f64x4 pacc = {0.0, 0.0, 0.0, 0.0};
  // every thread runs the same iterations (the barriers inside need that);
  // a group past the last row loads zeros
  for (int base = 0; base < m; base += 64 * NG) {
    const int r0 = base + 64 * g;
    for (int e = tg; e < 64 * PB; e += 256) {
      // X: the 16 columns of a row are contiguous, so k runs fastest
      const int rr = e / PB, k = e % PB;
      const int r = r0 + rr;
      sv[g][rr][k] = hh_v(D, c0, bp, m, k, r);
      if (qr) {
        const int col = chunk + k;
        sa[g][rr][k] = (r < m && col < ncols) ? A[(size_t)(c0 + r) * rs + (size_t)(colstart + col) * cs] : 0.0;
      }
    }
    if (!qr) {
      for (int e = tg; e < 64 * PB; e += 256) {
        // Q: the rows of a column are contiguous, so rr runs fastest
        const int rr = e % 64, k = e / 64;
        const int r = r0 + rr, col = chunk + k;
        sa[g][rr][k] = (r < m && col < ncols) ? A[(size_t)(c0 + r) * rs + (size_t)(colstart + col) * cs] : 0.0;
      }
    }
    __syncthreads();
    // (V^T A)[a][b] += sum over 4 rows: lane l feeds V[row l/16][l%16] as
    // the A operand and A[row l/16][l%16] as the B operand
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int rr = rw + 4 * j + (lane >> 4);
      pacc = __builtin_amdgcn_mfma_f64_16x16x4f64(sv[gw][rr][lane & 15], sa[gw][rr][lane & 15],
                                                  pacc, 0, 0, 0);
    }
    __syncthreads();
  }
  // result register r of lane l is row l/16 + 4 r, column l % 16
#pragma unroll
  for (int r = 0; r < 4; ++r) sp[wave][(lane >> 4) + 4 * r][lane & 15] = pacc[r];
  __syncthreads();
  if (g == 0) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < 16; ++w) v += sp[w][i][c];
    sp[0][i][c] = v;   // row i of wave 0's block is read by this thread only
  }
  __syncthreads();
  if (g == 0) {
    const double* T = D.T + (size_t)p * PB * PB;   // T[a][b] at a + 16 b
    double z = 0.0;
#pragma unroll
    for (int k = 0; k < PB; ++k) z = fma(qr ? T[k + PB * i] : T[i + PB * k], sp[0][k][c], z);
    sz[i][c] = (chunk + c < ncols) ? z : 0.0;
  }
  __syncthreads();
  // A -= V Z: each wave a 16 x 16 tile of the 256-row chunk on the matrix
  // cores (lane l feeds V[row l%16][k] and Z[k][l%16], k = 4 j + l/16), the
  // result subtracted from A (lane l: rows l/16 + 4 r, column l % 16)
  for (int base = 0; base < m; base += 64 * NG) {
    const int r0 = base + 64 * g;
    for (int e = tg; e < 64 * PB; e += 256) {
      const int rr = e / PB, k = e % PB;
      sv[g][rr][k] = hh_v(D, c0, bp, m, k, r0 + rr);
    }
    __syncthreads();
    f64x4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * j + (lane >> 4);
      y = __builtin_amdgcn_mfma_f64_16x16x4f64(sv[gw][rw + (lane & 15)][k], sz[k][lane & 15], y,
                                               0, 0, 0);
    }
    const int cc = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = base + 64 * gw + rw + (lane >> 4) + 4 * r;
      if (row < m && chunk + cc < ncols)
        A[(size_t)(c0 + row) * rs + (size_t)(colstart + chunk + cc) * cs] -= y[r];
    }
    __syncthreads();
  }
}

// rk of every block: one past the last nonzero diagonal entry of Rc (a
// dropped pivot leaves its whole row exactly zero)
__global__ __launch_bounds__(256) void rank_prefix_kernel(const PDesc* __restrict__ desc,
                                                          int* __restrict__ rk) {
  const PDesc D = desc[blockIdx.x];
  __shared__ int s[256];
  int last = -1;
  for (int i = threadIdx.x; i < D.K; i += 256)
    if (D.G[(size_t)i + (size_t)i * D.K] != 0.0) last = max(last, i);
  s[threadIdx.x] = last;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = max(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) rk[blockIdx.x] = s[0] + 1;
}

__global__ void eye_batched_kernel(const PDesc* __restrict__ desc) {
  const PDesc D = desc[blockIdx.y];
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)D.K * D.K;
  if (e < n) D.Q[e] = (e % D.K) == (e / D.K) ? 1.0 : 0.0;
}

unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

void precondition_jacobi(Context& c, std::vector<PrecondBlock>& blocks) {
  const int nb = (int)blocks.size();
  if (nb == 0) return;
  SU2_CHECK(nb < 65536, kErrCapacity, "jacobi preconditioner: %d blocks", nb);
  hipStream_t st = c.stream();
  int kmax = 0;
  long long scratch = 0;
  for (auto& b : blocks) {
    SU2_CHECK(b.K >= 1 && b.K <= kPrecondMaxK && b.L >= b.K, kErrArgument,
              "jacobi preconditioner: block %d x %d", b.L, b.K);
    kmax = std::max(kmax, b.K);
    const long long np = (b.K + PB - 1) / PB;
    scratch += 2LL * b.K * b.K + np * PB * PB;
  }
  DevBuf work(c, scratch);
  std::vector<PDesc> hd(nb);
  long long off = 0;
  for (int i = 0; i < nb; ++i) {
    const PrecondBlock& b = blocks[i];
    const long long kk = (long long)b.K * b.K, np = (b.K + PB - 1) / PB;
    PDesc& d = hd[i];
    d.G = work.p + off;
    d.Q = d.G + kk;
    d.T = d.Q + kk;
    off += 2 * kk + np * PB * PB;
    d.K = b.K;
    d.rk = b.K;
    d.tiny = 1e-15 * b.normF * b.normF;
  }
  DevBuf ddesc(c, (long long)((sizeof(PDesc) * nb + 7) / 8));
  SU2_HIP(hipMemcpyAsync(ddesc.p, hd.data(), sizeof(PDesc) * nb, hipMemcpyHostToDevice, st));
  const PDesc* dd = (const PDesc*)ddesc.p;
  // Gram matrices
  {
    GemmBatch g;
    for (int i = 0; i < nb; ++i) {
      const PrecondBlock& b = blocks[i];
      g.add(hd[i].G, b.K, b.K, b.K, b.W, b.L, 1, b.W, b.L, 0, b.L);
    }
    g.run(c);
  }
  const int npmax = (kmax + PB - 1) / PB;
  // Cholesky by row panels
  for (int p = 0; p < npmax; ++p) {
    hipLaunchKernelGGL(chol_panel_kernel, dim3(nb), dim3(PT), 0, st, dd, p);
    const int n = kmax - PB * (p + 1);
    if (n > 0) {
      const unsigned nt = cdiv(n, 64);
      hipLaunchKernelGGL(chol_trail_kernel, dim3(nt, nt, nb), dim3(256), 0, st, dd, p);
    }
  }
  // Rc's rank prefix: X = Rc^T has zero columns from rk on (a rank-deficient
  // block, such as a two-site wavefunction sector, drops its last pivots),
  // whose panels are identity reflectors and whose updates change nothing;
  // one wait brings the counts back to bound the panel loops
  {
    DevBuf drk(c, (nb + 1) / 2);
    hipLaunchKernelGGL(rank_prefix_kernel, dim3(nb), dim3(256), 0, st, dd, (int*)drk.p);
    std::vector<int> rk(nb);
    SU2_HIP(hipMemcpyAsync(rk.data(), drk.p, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
    SU2_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nb; ++i) hd[i].rk = rk[i];
    SU2_HIP(hipMemcpyAsync(ddesc.p, hd.data(), sizeof(PDesc) * nb, hipMemcpyHostToDevice, st));
  }
  int rkmax = 0;
  for (int i = 0; i < nb; ++i) rkmax = std::max(rkmax, hd[i].rk);
  const int npq = (rkmax + PB - 1) / PB;
  // Householder QR of Rc^T by column panels (trailing columns updated)
  for (int p = 0; p < npq; ++p) {
    hipLaunchKernelGGL(hh_panel_kernel, dim3(nb), dim3(PT), 0, st, dd, p);
    const int n = rkmax - PB * (p + 1);
    if (n > 0) {
      hipLaunchKernelGGL(hh_apply_kernel, dim3(cdiv(n, PB), nb), dim3(1024), 0, st, dd, p, 1);
    }
  }
  // Q2 = H_0 ... H_{np-1}, formed backwards from the identity
  hipLaunchKernelGGL(eye_batched_kernel, dim3(cdiv((long long)kmax * kmax, 256), nb), dim3(256),
                     0, st, dd);
  for (int p = npq - 1; p >= 0; --p) {
    const int n = kmax - PB * p;
    hipLaunchKernelGGL(hh_apply_kernel, dim3(cdiv(n, PB), nb), dim3(1024), 0, st, dd, p, 0);
  }
  SU2_HIP(hipGetLastError());
  // W Q2 and V Q2
  {
    GemmBatch g;
    for (int i = 0; i < nb; ++i) {
      const PrecondBlock& b = blocks[i];
      g.add(b.Wout, b.L, b.L, b.K, b.W, b.L, 0, hd[i].Q, b.K, 0, b.K);
      g.add(b.Vout, b.K, b.K, b.K, b.V, b.K, 0, hd[i].Q, b.K, 0, b.K);
    }
    g.run(c);
  }
}

}  // namespace su2
