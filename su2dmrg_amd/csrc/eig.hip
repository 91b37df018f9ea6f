#include "eig.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

#include "blas1.hpp"
#include "common.hpp"
#include "context.hpp"
#include "gemm.hpp"

namespace su2 {

namespace {

constexpr int NMAX = 64;      // largest LDS problem
constexpr int BLK = 16;       // block width of the block-Jacobi driver
constexpr int MAX_SWEEPS = 60;

__device__ __forceinline__ int pair_p(int r, int k, int m) {
  // circle method over m (even) players; player m-1 is fixed
  return k == 0 ? m - 1 : (r + k) % (m - 1);
}
__device__ __forceinline__ int pair_q(int r, int k, int m) {
  return k == 0 ? r : (r - k + (m - 1)) % (m - 1);
}

// One 64-lane wave per job: cyclic Jacobi with round-robin pair ordering.
// A rotation is applied when |a_pq| exceeds 1e-16 sqrt(|a_pp a_qq|) and an
// absolute floor of 1e-18 ||A||_F; sweeps stop when none is applied.
__global__ __launch_bounds__(64) void eig_small_kernel(const EigJob* __restrict__ jobs) {
  __shared__ double a[NMAX][NMAX + 1];
  __shared__ double v[NMAX][NMAX + 1];
  __shared__ double cs[NMAX / 2], sn[NMAX / 2];
  __shared__ int pp[NMAX / 2], qq[NMAX / 2];
  __shared__ double red[64];
  __shared__ int rot_count;
  __shared__ int rank_of[NMAX];

  const EigJob J = jobs[blockIdx.x];
  const int n = J.bp + J.bq;
  const int lane = threadIdx.x;
  if (n <= 0) return;

  for (int e = lane; e < n * n; e += 64) {
    const int r = e % n, c = e / n;
    const int gr = r < J.bp ? J.p0 + r : J.q0 + (r - J.bp);
    const int gc = c < J.bp ? J.p0 + c : J.q0 + (c - J.bp);
    a[r][c] = J.A[gr + (long long)gc * J.lda];
    v[r][c] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  // symmetrise and take the Frobenius norm
  double s2 = 0.0;
  for (int e = lane; e < n * n; e += 64) {
    const int r = e % n, c = e / n;
    if (r < c) {
      const double m = 0.5 * (a[r][c] + a[c][r]);
      a[r][c] = m;
      a[c][r] = m;
    }
  }
  __syncthreads();
  for (int e = lane; e < n * n; e += 64) {
    const int r = e % n, c = e / n;
    s2 += a[r][c] * a[r][c];
  }
  red[lane] = s2;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (lane < w) red[lane] += red[lane + w];
    __syncthreads();
  }
  const double floor_abs = 1e-18 * sqrt(red[0]);
  __syncthreads();

  const int m = n + (n & 1);
  const int npair = m / 2;
  for (int sweep = 0; sweep < MAX_SWEEPS && n > 1; ++sweep) {
    if (lane == 0) rot_count = 0;
    __syncthreads();
    for (int r = 0; r < m - 1; ++r) {
      if (lane < npair) {
        const int p = pair_p(r, lane, m), q = pair_q(r, lane, m);
        double c = 1.0, s = 0.0;
        if (p < n && q < n) {
          const double apq = a[p][q];
          const double app = a[p][p], aqq = a[q][q];
          if (fabs(apq) > fmax(1e-16 * sqrt(fabs(app * aqq)), floor_abs)) {
            const double tau = (aqq - app) / (2.0 * apq);
            double t;
            if (fabs(tau) > 1e150) t = 0.5 / tau;
            else t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
            atomicAdd(&rot_count, 1);
          }
        }
        pp[lane] = p;
        qq[lane] = q;
        cs[lane] = c;
        sn[lane] = s;
      }
      __syncthreads();
      // columns: A <- A J
      for (int e = lane; e < npair * n; e += 64) {
        const int k = e / n, row = e % n;
        const int p = pp[k], q = qq[k];
        if (p >= n || q >= n || sn[k] == 0.0) continue;
        const double c = cs[k], s = sn[k];
        const double x = a[row][p], y = a[row][q];
        a[row][p] = c * x - s * y;
        a[row][q] = s * x + c * y;
      }
      __syncthreads();
      // rows: A <- J^T A ; and V <- V J
      for (int e = lane; e < npair * n; e += 64) {
        const int k = e / n, col = e % n;
        const int p = pp[k], q = qq[k];
        if (p >= n || q >= n || sn[k] == 0.0) continue;
        const double c = cs[k], s = sn[k];
        const double x = a[p][col], y = a[q][col];
        a[p][col] = c * x - s * y;
        a[q][col] = s * x + c * y;
        const double vx = v[col][p], vy = v[col][q];
        v[col][p] = c * vx - s * vy;
        v[col][q] = s * vx + c * vy;
      }
      __syncthreads();
    }
    if (rot_count == 0) break;
    __syncthreads();
  }

  // ranks for the ascending order (ties by index), then write out
  for (int i = lane; i < n; i += 64) {
    int rk = i;
    if (J.sort) {
      rk = 0;
      const double di = a[i][i];
      for (int j = 0; j < n; ++j) {
        const double dj = a[j][j];
        if (dj < di || (dj == di && j < i)) ++rk;
      }
    }
    rank_of[i] = rk;
  }
  __syncthreads();
  for (int e = lane; e < n * n; e += 64) {
    const int row = e % n, col = e / n;
    J.V[row + (long long)rank_of[col] * J.ldv] = v[row][col];
  }
  if (J.W)
    for (int i = lane; i < n; i += 64) J.W[rank_of[i]] = a[i][i];
}

// sum of squares of the off-diagonal part and of the diagonal of an n x n matrix
__global__ void offnorm_kernel(const double* __restrict__ A, int n, double* __restrict__ part) {
  __shared__ double so[256], sd[256];
  double o = 0, d = 0;
  const long long nn = (long long)n * n;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < nn; e += (long long)gridDim.x * 256) {
    const int r = (int)(e % n), c = (int)(e / n);
    const double x = A[e];
    if (r == c) d += x * x;
    else o += x * x;
  }
  so[threadIdx.x] = o;
  sd[threadIdx.x] = d;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      so[threadIdx.x] += so[threadIdx.x + w];
      sd[threadIdx.x] += sd[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = so[0];
    part[2 * blockIdx.x + 1] = sd[0];
  }
}

__global__ void diag_kernel(const double* __restrict__ A, int n, double* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = A[i + (long long)i * n];
}

__global__ void identity_kernel(double* __restrict__ V, int n) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < (long long)n * n) V[e] = ((e % n) == (e / n)) ? 1.0 : 0.0;
}

__global__ void gather_cols_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                   const int* __restrict__ perm, int n) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int r = (int)(e % n), c = (int)(e / n);
  dst[e] = src[r + (long long)perm[c] * n];
}

// Column sign convention: the first entry with |v| >= max|v|/2 of each
// column becomes positive (robust to equal-magnitude entries).  One
// workgroup per column.
__global__ __launch_bounds__(256) void fix_signs_kernel(double* __restrict__ V, int rows, long long ld) {
  __shared__ double sv[256];
  __shared__ int si[256];
  double* col = V + (long long)blockIdx.x * ld;
  double best = 0.0;
  for (int r = threadIdx.x; r < rows; r += 256) best = fmax(best, fabs(col[r]));
  sv[threadIdx.x] = best;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sv[threadIdx.x] = fmax(sv[threadIdx.x], sv[threadIdx.x + w]);
    __syncthreads();
  }
  const double half = 0.5 * sv[0];
  int first = 0x7fffffff;
  for (int r = threadIdx.x; r < rows; r += 256)
    if (fabs(col[r]) >= half && r < first) first = r;
  si[threadIdx.x] = first;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) si[threadIdx.x] = min(si[threadIdx.x], si[threadIdx.x + w]);
    __syncthreads();
  }
  const int k = si[0];
  const bool flip = (k < rows) && col[k] < 0.0;
  __syncthreads();
  if (flip)
    for (int r = threadIdx.x; r < rows; r += 256) col[r] = -col[r];
}

unsigned blocks_for(long long n, int t) { return (unsigned)std::max<long long>(1, (n + t - 1) / t); }

}  // namespace

void dev_fix_signs(Context& c, double* V, int rows, int cols, long long ld) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(fix_signs_kernel, dim3(cols), dim3(256), 0, c.stream(), V, rows, ld);
  SU2_HIP(hipGetLastError());
}

void eig_small_batched(Context& c, const EigJob* d_jobs, int njob) {
  if (njob <= 0) return;
  hipLaunchKernelGGL(eig_small_kernel, dim3(njob), dim3(64), 0, c.stream(), d_jobs);
  SU2_HIP(hipGetLastError());
}

void sym_eig(Context& c, double* A, int n, double* V, std::vector<double>& w) {
  SU2_CHECK(n >= 1, kErrArgument, "sym_eig: n=%d", n);
  w.assign(n, 0.0);
  hipStream_t st = c.stream();
  if (n <= NMAX) {
    // single LDS solve, sorted, straight into V
    EigJob j{};
    j.A = A;
    j.lda = n;
    j.p0 = 0;
    j.bp = n;
    j.q0 = 0;
    j.bq = 0;
    j.V = V;
    j.ldv = n;
    double* dW = c.reduce_scratch(n);
    j.W = dW;
    j.sort = 1;
    EigJob* dj = nullptr;
    SU2_HIP(hipMalloc(&dj, sizeof(EigJob)));
    SU2_HIP(hipMemcpyAsync(dj, &j, sizeof(EigJob), hipMemcpyHostToDevice, st));
    eig_small_batched(c, dj, 1);
    SU2_HIP(hipMemcpyAsync(w.data(), dW, n * sizeof(double), hipMemcpyDeviceToHost, st));
    SU2_HIP(hipStreamSynchronize(st));
    SU2_HIP(hipFree(dj));
    return;
  }

  // ---- block Jacobi ------------------------------------------------------
  const int nb = (n + BLK - 1) / BLK;
  const int m = nb + (nb & 1);
  const int npair = m / 2;
  const int nround = m - 1;
  auto bsize = [&](int b) { return b >= nb ? 0 : std::min(BLK, n - b * BLK); };
  auto pair_of = [&](int r, int k, int& P, int& Q) {
    P = (k == 0) ? m - 1 : (r + k) % (m - 1);
    Q = (k == 0) ? r : (r - k + (m - 1)) % (m - 1);
    if (bsize(P) == 0) std::swap(P, Q);  // the dummy block always goes second
  };

  const long long nn = (long long)n * n;
  double *G2 = nullptr, *V2 = nullptr, *R = nullptr;
  SU2_HIP(hipMalloc(&G2, nn * sizeof(double)));
  SU2_HIP(hipMalloc(&V2, nn * sizeof(double)));
  const int RLD = 2 * BLK;
  SU2_HIP(hipMalloc(&R, (size_t)npair * RLD * RLD * sizeof(double)));
  hipLaunchKernelGGL(identity_kernel, dim3(blocks_for(nn, 256)), dim3(256), 0, st, V, n);

  // per-round pair jobs and the two GEMM passes, built once and replayed
  std::vector<EigJob> jobs((size_t)nround * npair);
  std::vector<GroupedGemm*> colp(nround), rowp(nround);
  for (int r = 0; r < nround; ++r) {
    colp[r] = new GroupedGemm();
    rowp[r] = new GroupedGemm();
    for (int k = 0; k < npair; ++k) {
      int P, Q;
      pair_of(r, k, P, Q);
      const int bp = bsize(P), bq = bsize(Q);
      const int p0 = P * BLK, q0 = bq ? Q * BLK : 0;
      double* Rk = R + (size_t)k * RLD * RLD;
      EigJob& j = jobs[(size_t)r * npair + k];
      j.A = A;
      j.lda = n;
      j.p0 = p0;
      j.bp = bp;
      j.q0 = q0;
      j.bq = bq;
      j.V = Rk;
      j.ldv = RLD;
      j.W = nullptr;
      j.sort = 0;
      if (bp == 0) continue;
      // column pass: X2[:, P] = X[:, P] R_PP + X[:, Q] R_QP (X = A and V), same for Q
      for (int which = 0; which < 2; ++which) {
        const double* X = which == 0 ? A : V;
        double* X2 = which == 0 ? G2 : V2;
        for (int half = 0; half < 2; ++half) {
          const int c0 = half == 0 ? p0 : q0, bc = half == 0 ? bp : bq, roff = half == 0 ? 0 : bp;
          if (bc == 0) continue;
          int pr = colp[r]->add_problem(X2 + (long long)c0 * n, n, n, bc);
          colp[r]->add_slice(pr, X + (long long)p0 * n, n, 0, Rk + (long long)roff * RLD, RLD, 0,
                             bp, 1.0);
          if (bq)
            colp[r]->add_slice(pr, X + (long long)q0 * n, n, 0, Rk + bp + (long long)roff * RLD,
                               RLD, 0, bq, 1.0);
        }
      }
      // row pass: A[P, :] = R_PP^T G2[P, :] + R_QP^T G2[Q, :], same for Q
      for (int half = 0; half < 2; ++half) {
        const int r0 = half == 0 ? p0 : q0, br = half == 0 ? bp : bq, coff = half == 0 ? 0 : bp;
        if (br == 0) continue;
        int pr = rowp[r]->add_problem(A + r0, n, br, n);
        rowp[r]->add_slice(pr, Rk + (long long)coff * RLD, RLD, 1, G2 + p0, n, 0, bp, 1.0);
        if (bq)
          rowp[r]->add_slice(pr, Rk + bp + (long long)coff * RLD, RLD, 1, G2 + q0, n, 0, bq, 1.0);
      }
    }
    colp[r]->finalize(512, 4);
    rowp[r]->finalize(512, 4);
  }
  EigJob* djobs = nullptr;
  SU2_HIP(hipMalloc(&djobs, jobs.size() * sizeof(EigJob)));
  SU2_HIP(hipMemcpyAsync(djobs, jobs.data(), jobs.size() * sizeof(EigJob), hipMemcpyHostToDevice,
                         st));

  const int nred = 64;
  double* part = c.reduce_scratch(2 * nred + n);
  double* hpart = c.host_scratch(2 * nred);
  for (int sweep = 0; sweep < MAX_SWEEPS; ++sweep) {
    hipLaunchKernelGGL(offnorm_kernel, dim3(nred), dim3(256), 0, st, A, n, part);
    SU2_HIP(hipMemcpyAsync(hpart, part, 2 * nred * sizeof(double), hipMemcpyDeviceToHost, st));
    SU2_HIP(hipStreamSynchronize(st));
    double off = 0, dia = 0;
    for (int i = 0; i < nred; ++i) {
      off += hpart[2 * i];
      dia += hpart[2 * i + 1];
    }
    if (off <= 1e-30 * (off + dia)) break;
    for (int r = 0; r < nround; ++r) {
      eig_small_batched(c, djobs + (size_t)r * npair, npair);
      colp[r]->launch(st);
      rowp[r]->launch(st);
      dev_copy(c, V, V2, nn);
    }
  }
  // eigenvalues, ascending order, eigenvector columns permuted to match
  double* dd = part + 2 * nred;
  hipLaunchKernelGGL(diag_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, A, n, dd);
  SU2_HIP(hipMemcpyAsync(w.data(), dd, n * sizeof(double), hipMemcpyDeviceToHost, st));
  SU2_HIP(hipStreamSynchronize(st));
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return w[x] < w[y]; });
  std::vector<double> ws(n);
  for (int i = 0; i < n; ++i) ws[i] = w[perm[i]];
  w.swap(ws);
  int* dperm = nullptr;
  SU2_HIP(hipMalloc(&dperm, n * sizeof(int)));
  SU2_HIP(hipMemcpyAsync(dperm, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(gather_cols_kernel, dim3(blocks_for(nn, 256)), dim3(256), 0, st, V, V2, dperm,
                     n);
  dev_copy(c, V, V2, nn);
  SU2_HIP(hipStreamSynchronize(st));
  for (auto* g : colp) delete g;
  for (auto* g : rowp) delete g;
  SU2_HIP(hipFree(dperm));
  SU2_HIP(hipFree(djobs));
  SU2_HIP(hipFree(G2));
  SU2_HIP(hipFree(V2));
  SU2_HIP(hipFree(R));
}

}  // namespace su2
