// One-sided block Jacobi SVD, see jacobi.hpp.
#include "jacobi.hpp"

#include <algorithm>
#include <map>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>

#include "blas1.hpp"
#include "common.hpp"
#include "context.hpp"
#include "precond.hpp"

namespace su2 {

namespace {

constexpr int RCH = 32;         // rows per staged chunk
constexpr int INNER_SWEEPS = 3; // most inner Jacobi sweeps per visit of a pair
constexpr int MAX_SLICES = 8;   // row slices per pair
constexpr int MIN_SLICE_ROWS = 64;
// Columns with norm at or below kNoiseRel ||A||_F are at the rounding level
// of the block: they carry no direction, no pair with one of them holds the
// iteration open, and the sweeps leave them out of the schedule.
constexpr double kNoiseRel = 1e-13;

typedef double f64x4 __attribute__((ext_vector_type(4)));

struct BlockDesc {
  double* W;        // L x k, ld L
  double* V;        // k x k, ld k
  int L, k;
  double normF;     // Frobenius norm of the block (absolute floor scale)
};

struct PairJob {
  int blk;          // index into the block table
  int p0, bp;       // first column block
  int q0, bq;       // second column block (bq = 0: the first block alone)
  int P;            // row slices (workgroups) that share the pair
};

__device__ __forceinline__ int rr_p(int r, int k, int m) { return k == 0 ? m - 1 : (r + k) % (m - 1); }
__device__ __forceinline__ int rr_q(int r, int k, int m) {
  return k == 0 ? r : (r - k + (m - 1)) % (m - 1);
}

// 1 / x and 1 / sqrt(x) for x > 0 (normal): the hardware estimates
// refined by two Newton-Raphson steps, y <- y (2 - x y) and
// y <- y (3 - x y^2) / 2, each of which squares the relative error.
__device__ __forceinline__ double recip_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  return fma(y, fma(-x, y, 1.0), y);
}
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  y = y * fma(-hx * y, y, 1.5);
  return y * fma(-hx * y, y, 1.5);
}

// rows [lo, hi) of slice s when L rows are cut into P slices
__device__ __forceinline__ void slice_rows(int L, int P, int s, int& lo, int& hi) {
  const int per = (L + P - 1) / P;
  lo = min(L, s * per);
  hi = min(L, lo + per);
}

// Upper-triangle tile u (row-major over tr <= tc) of a TT x TT tile grid.
__device__ __forceinline__ void upper_tile(int u, int TT, int& tr, int& tc) {
  tr = 0;
  while (u >= TT - tr) {
    u -= TT - tr;
    ++tr;
  }
  tc = tr + u;
}

// Partial Gram matrix of one pair over one row slice of W, on the matrix
// cores.  RCH-row chunks of the pair's NP = 2 CB columns are staged in LDS
// (columns >= n zero).  The Gram matrix is symmetric, so only the
// TT (TT + 1) / 2 tiles of 16 x 16 on and above the diagonal (TT = NP / 16)
// are computed, spread over the waves, each a v_mfma_f64_16x16x4f64 chain:
// A = X^T fragment (lane l: X[r + l/16][16 tr + l%16]), B = X fragment (lane
// l: X[r + l/16][16 tc + l%16]), D lane l: G[16 tr + l/16 + 4 q][16 tc +
// l%16].  Each product X[r][a] X[r][b] is the same number either way round,
// so the mirrored tile is the one its own chain would give.  The NP x NP
// result goes row-major to gpart[blockIdx.x] (workgroup = job * pmax +
// slice).
template <int CB, int NT, int RCH>
__global__ __launch_bounds__(NT) void gram_partial_kernel(const BlockDesc* __restrict__ blocks,
                                                          const PairJob* __restrict__ jobs,
                                                          int pmax, double* __restrict__ gpart) {
  constexpr int NP = 2 * CB, TT = NP / 16, NU = TT * (TT + 1) / 2, NW = NT / 64;
  constexpr int TPW = (NU + NW - 1) / NW;   // tiles per wave (the last waves may hold none)
  static_assert(NW * 64 == NT, "whole waves");
  __shared__ double X[RCH][NP + 1];
  __shared__ int colmap[NP];
  const int job = blockIdx.x / pmax, slice = blockIdx.x - job * pmax;
  const PairJob J = jobs[job];
  if (slice >= J.P) return;
  const BlockDesc B = blocks[J.blk];
  const int n = J.bp + J.bq;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int c = t; c < n; c += NT) colmap[c] = c < J.bp ? J.p0 + c : J.q0 + (c - J.bp);
  int lo, hi;
  slice_rows(B.L, J.P, slice, lo, hi);
  int ttr[TPW], ttc[TPW];
#pragma unroll
  for (int j = 0; j < TPW; ++j) upper_tile(min(wave * TPW + j, NU - 1), TT, ttr[j], ttc[j]);
  f64x4 acc[TPW];
#pragma unroll
  for (int j = 0; j < TPW; ++j) acc[j] = f64x4{0.0, 0.0, 0.0, 0.0};
  // the next chunk's global loads are issued before the current chunk's MFMAs
  constexpr int LPT = RCH * NP / NT;
  static_assert(LPT * NT == RCH * NP, "chunk elements per thread");
  double pre[LPT];
  auto fetch = [&](int r0) {
    const int nr = min(RCH, hi - r0);
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      const int e = t + NT * i, rr = e % RCH, c = e / RCH;
      pre[i] = (rr < nr && c < n) ? B.W[(r0 + rr) + (long long)colmap[c] * B.L] : 0.0;
    }
  };
  __syncthreads();
  if (lo < hi) fetch(lo);
  for (int r0 = lo; r0 < hi; r0 += RCH) {
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      const int e = t + NT * i;
      X[e % RCH][e / RCH] = pre[i];
    }
    __syncthreads();
    if (r0 + RCH < hi) fetch(r0 + RCH);
#pragma unroll
    for (int k4 = 0; k4 < RCH; k4 += 4) {
      const int row = k4 + (lane >> 4);
#pragma unroll
      for (int j = 0; j < TPW; ++j)
        if (wave * TPW + j < NU)
          acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(X[row][16 * ttr[j] + (lane & 15)],
                                                        X[row][16 * ttc[j] + (lane & 15)], acc[j],
                                                        0, 0, 0);
    }
    __syncthreads();
  }
  double* out = gpart + (size_t)blockIdx.x * NP * NP;
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    if (wave * TPW + j >= NU) continue;
    const int tr = ttr[j], tc = ttc[j];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int a = 16 * tr + (lane >> 4) + 4 * q, b = 16 * tc + (lane & 15);
      out[a * NP + b] = acc[j][q];
      if (tr != tc) out[b * NP + a] = acc[j][q];
    }
  }
}

// Orthogonalise one pair of column blocks: every row slice of the pair sums
// the same partial Gram matrices in the same order, so the P workgroups of a
// pair hold bit-identical Gram matrices, convergence decisions and rotations;
// each then rotates its own rows of W and V.  Slice 0 records the pair's
// largest scaled off-diagonal Gram entry (before rotating) in off[blk] (as
// ordered bits of a non-negative double).
//
// The pair's Gram matrix is diagonalised in LDS by parallel cyclic Jacobi
// (round-robin rounds of disjoint rotations, inner_sweeps <= INNER_SWEEPS
// sweeps: the outer iteration revisits the pair; one inner sweep per visit
// costs a few more outer sweeps but the least time).  A round updates
// G <- J^T G J by 2 x 2 blocks, with the rotated entry set to zero and the diagonal
// moved by t g_pq; a rotation is skipped below max(tol_in sqrt(g_pp g_qq),
// 0.2 tol_abs ||A||_F sqrt(max(g_pp, g_qq))), about ten times below the
// outer test.  Round latency, not arithmetic, bounds this loop (a 32 x 32
// round costs about as much as a 64 x 64 one), so the rotation parameters
// avoid the correctly rounded square root and division sequences, and the
// 64 x 64 case runs 512 threads (two waves per SIMD) in two groups: the
// lower half updates G, then wave 0 computes the next round's rotations
// while the upper half updates R with this round's (double-buffered)
// parameters.
template <int CB, int NT>
__global__ __launch_bounds__(NT) void jacobi_rotate_kernel(const BlockDesc* __restrict__ blocks,
                                                           const PairJob* __restrict__ jobs,
                                                           int pmax,
                                                           const double* __restrict__ gpart,
                                                           double tol_rel, double tol_abs,
                                                           unsigned long long* __restrict__ off,
                                                           int inner_sweeps) {
  constexpr int NP = 2 * CB, HALF = NP / 2, TT = NP / 16, NW = NT / 64;
  // 512 and more threads: G and R are updated by separate halves
  constexpr bool SPLIT = NT >= 512;
  constexpr int NTG = SPLIT ? NT / 2 : NT, NTR = SPLIT ? NT / 2 : NT;
  constexpr int RT = RCH / 16, TPW = RT * TT / NW;
  static_assert(TPW * NW == RT * TT, "tiles must split evenly over the waves");
  static_assert(NT >= HALF && NT % NP == 0, "thread mapping");
  __shared__ double G[NP][NP + 1];
  __shared__ double R[NP][NP + 1];
  __shared__ double X[RCH][NP + 1];
  __shared__ double cs[2][HALF], sn[2][HALF], tn[2][HALF];
  __shared__ int pp[2][HALF], qq[2][HALF];
  __shared__ int colmap[NP];
  __shared__ double red[NT];
  __shared__ int rotated[INNER_SWEEPS];

  const int job = blockIdx.x / pmax, slice = blockIdx.x - job * pmax;
  const PairJob J = jobs[job];
  if (slice >= J.P) return;
  const BlockDesc B = blocks[J.blk];
  const int n = J.bp + J.bq;
  const int t = threadIdx.x;
  if (n < 2) return;
  for (int c = t; c < n; c += NT) colmap[c] = c < J.bp ? J.p0 + c : J.q0 + (c - J.bp);
  if (t < INNER_SWEEPS) rotated[t] = 0;
  {
    // the slices' partial Gram matrices, summed in slice order; all loads of
    // a thread are issued before the sums
    const double* gp = gpart + (size_t)job * pmax * NP * NP;
    constexpr int PER = NP * NP / NT, GRP = 4;
    static_assert(PER * NT == NP * NP && PER % GRP == 0, "Gram elements per thread");
    for (int i0 = 0; i0 < PER; i0 += GRP) {
      double part[GRP][MAX_SLICES];
#pragma unroll
      for (int i = 0; i < GRP; ++i)
#pragma unroll
        for (int s = 0; s < MAX_SLICES; ++s)
          part[i][s] = s < J.P ? gp[(size_t)s * NP * NP + t + NT * (i0 + i)] : 0.0;
#pragma unroll
      for (int i = 0; i < GRP; ++i) {
        const int e = t + NT * (i0 + i), a = e / NP, b = e % NP;
        double g = 0.0;
#pragma unroll
        for (int s = 0; s < MAX_SLICES; ++s)
          if (s < J.P) g += part[i][s];
        G[a][b] = g;
        R[a][b] = a == b ? 1.0 : 0.0;
      }
    }
  }
  __syncthreads();

  // ---- convergence measure of this pair ----
  double worst = 0.0;
  bool rotate = false;
  for (int e = t; e < n * n; e += NT) {
    const int a = e % n, b = e / n;
    if (a >= b) continue;
    const double g = fabs(G[a][b]);
    const double na = sqrt(G[a][a]), nb = sqrt(G[b][b]);
    // columns at the rounding level of the block carry no direction (the SVD
    // zeroes them); they do not hold the iteration open
    if (fmin(na, nb) <= kNoiseRel * B.normF) continue;
    const double scale = na * nb;
    // relative orthogonality, or the rounding floor of columns that carry
    // absolute errors of order eps ||A||
    const double floor_ab = tol_abs * B.normF * (na + nb);
    if (g > fmax(tol_rel * scale, floor_ab)) {
      rotate = true;
      worst = fmax(worst, g / scale);
    }
  }
  red[t] = worst + (rotate ? 4.0 : 0.0);   // +4 marks "needs a rotation" (ratios are <= 1)
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (t < w) red[t] = fmax(red[t], red[t + w]);
    __syncthreads();
  }
  const double top = red[0];
  if (t == 0 && slice == 0) {
    const double ratio = top >= 4.0 ? top - 4.0 : top;
    // a pair that still needs a rotation reports at least 2 (above any ratio)
    const double rep = top >= 4.0 ? 2.0 + ratio : ratio;
    atomicMax(off + J.blk, (unsigned long long)__double_as_longlong(rep));
  }
  if (top < 4.0) return;   // orthogonal to tolerance: nothing to rotate

  // ---- eigenvectors of G by parallel cyclic Jacobi ----
  const int m = n + (n & 1);   // index n (odd n) is a zero row and column
  const int npair = m / 2;
  const int nround = m - 1;
  const double tol_in = 1e-14;
  const double fl = 0.1 * tol_abs * B.normF;
  // Rotation of slot k in round r (into parameter buffer b): the
  // round-robin pair for k < npair, an identity on the spare indices
  // (2k, 2k+1) >= m otherwise, so that the HALF slots partition all NP
  // indices and every updating thread owns a fixed set of 2 x 2 blocks and R
  // entries with no branches.
  auto rotation = [&](int r, int k, int b, int sweep) {
    int p, q;
    double c = 1.0, s = 0.0, tt = 0.0;
    if (k < npair) {
      p = rr_p(r, k, m);
      q = rr_q(r, k, m);
      if (p < n && q < n) {
        const double apq = G[p][q];
        const double gpp = G[p][p], gqq = G[q][q];
        const double app = fabs(gpp), aqq = fabs(gqq);
        // skip test on squares (no square roots): |g_pq| against
        // tol_in sqrt(g_pp g_qq) and the rounding floor 2 fl sqrt(max g)
        const double a2 = apq * apq;
        if (a2 > tol_in * tol_in * (app * aqq) && a2 > 4.0 * fl * fl * fmax(app, aqq) &&
            fabs(apq) > 1e-300) {
          // t = tan(theta) of the smaller rotation angle,
          //   t = 2 g_pq sign(d) / (|d| + sqrt(d^2 + 4 g_pq^2)), d = g_qq - g_pp
          // (the textbook tau form without its division); c = 1 / sqrt(1 +
          // t^2) keeps the rotation orthogonal to rounding for any t, so
          // the reciprocals need not be correctly rounded
          const double d = gqq - gpp;
          const double h2 = d * d + 4.0 * a2;
          const double den = fabs(d) + h2 * rsqrt_nr(h2);
          tt = (d >= 0 ? 2.0 : -2.0) * apq * recip_nr(den);
          c = rsqrt_nr(1.0 + tt * tt);
          s = tt * c;
          rotated[sweep] = 1;
        }
      }
    } else {
      p = 2 * k;
      q = 2 * k + 1;
    }
    pp[b][k] = p;
    qq[b][k] = q;
    cs[b][k] = c;
    sn[b][k] = s;
    tn[b][k] = tt;
  };
  // G <- J^T G J on the 2 x 2 blocks (k, k2) of a full-square partition
  // (each element has one owner; identity rotations store their values back
  // unchanged) by NTG threads, tg < NTG.  All loads of a thread precede its
  // stores.
  auto update_g = [&](int b, int tg) {
    constexpr int KS = NTG / HALF, MB = HALF / KS;
    static_assert(MB * KS == HALF, "block partition");
    const int k = tg % HALF, kb = tg / HALF;
    const int p = pp[b][k], q = qq[b][k];
    const double c = cs[b][k], s = sn[b][k], tk = tn[b][k];
    double g[MB][4], c2v[MB], s2v[MB];
    int p2v[MB], q2v[MB];
#pragma unroll
    for (int i = 0; i < MB; ++i) {
      const int k2 = kb + KS * i;
      p2v[i] = pp[b][k2];
      q2v[i] = qq[b][k2];
      c2v[i] = cs[b][k2];
      s2v[i] = sn[b][k2];
      g[i][0] = G[p][p2v[i]];
      g[i][1] = G[p][q2v[i]];
      g[i][2] = G[q][p2v[i]];
      g[i][3] = G[q][q2v[i]];
    }
#pragma unroll
    for (int i = 0; i < MB; ++i) {
      const double c2 = c2v[i], s2 = s2v[i];
      const double t00 = c2 * g[i][0] - s2 * g[i][1], t01 = s2 * g[i][0] + c2 * g[i][1];
      const double t10 = c2 * g[i][2] - s2 * g[i][3], t11 = s2 * g[i][2] + c2 * g[i][3];
      double n00 = c * t00 - s * t10, n01 = c * t01 - s * t11;
      double n10 = s * t00 + c * t10, n11 = s * t01 + c * t11;
      // the rotated pair's own block: annihilated exactly, diagonal moved by t g_pq
      const bool own = (kb + KS * i == k) && s != 0.0;
      n00 = own ? g[i][0] - tk * g[i][1] : n00;
      n11 = own ? g[i][3] + tk * g[i][1] : n11;
      n01 = own ? 0.0 : n01;
      n10 = own ? 0.0 : n10;
      G[p][p2v[i]] = n00;
      G[p][q2v[i]] = n01;
      G[q][p2v[i]] = n10;
      G[q][q2v[i]] = n11;
    }
  };
  // R <- R J by NTR threads, tr < NTR, each on fixed rows and pair slots
  // (rows >= n hold zeros in the rotated columns)
  auto update_r = [&](int b, int tr) {
    constexpr int RS = NTR / NP, MR = HALF / RS;
    static_assert(MR * RS == HALF, "row partition");
    const int row = tr % NP, rb = tr / NP;
    double ru[MR], rv[MR], rc[MR], rsn[MR];
    int rp[MR], rq[MR];
#pragma unroll
    for (int i = 0; i < MR; ++i) {
      const int kk = rb + RS * i;
      rp[i] = pp[b][kk];
      rq[i] = qq[b][kk];
      rc[i] = cs[b][kk];
      rsn[i] = sn[b][kk];
      ru[i] = R[row][rp[i]];
      rv[i] = R[row][rq[i]];
    }
#pragma unroll
    for (int i = 0; i < MR; ++i) {
      R[row][rp[i]] = rc[i] * ru[i] - rsn[i] * rv[i];
      R[row][rq[i]] = rsn[i] * ru[i] + rc[i] * rv[i];
    }
  };
  if (!SPLIT) {
    // one group: rotations, barrier, G and R updates, barrier
    for (int sweep = 0; sweep < inner_sweeps; ++sweep) {
      for (int r = 0; r < nround; ++r) {
        if (t < HALF) rotation(r, t, 0, sweep);
        __syncthreads();
        update_g(0, t);
        update_r(0, t);
        __syncthreads();
      }
      if (!rotated[sweep]) break;
    }
  } else {
    // Two groups: threads below NTG update G; then, while wave 0 computes
    // the next round's rotations from the updated G, the other NTR threads
    // update R with this round's (double-buffered parameters).
    if (t < HALF) rotation(0, t, 0, 0);
    __syncthreads();
    int b = 0;
    for (int sweep = 0; sweep < inner_sweeps; ++sweep) {
      for (int r = 0; r < nround; ++r, b ^= 1) {
        if (t < NTG) update_g(b, t);
        __syncthreads();
        const bool last = r + 1 == nround;
        const int rn = last ? 0 : r + 1, sw = last ? sweep + 1 : sweep;
        if (t < HALF && sw < inner_sweeps) rotation(rn, t, b ^ 1, sw);
        if (t >= NTG) update_r(b, t - NTG);
        __syncthreads();
      }
      if (!rotated[sweep]) break;
    }
  }

  // ---- rotate this slice's rows: W[:, cols] <- W[:, cols] R, same for V ----
  // Y = X R per 32-row chunk on the matrix cores, the (RCH/16) x (NP/16)
  // output tiles spread over the waves: A = X fragment (lane l:
  // X[16 rt + l%16][k + l/16]), B = R fragment (lane l: R[k + l/16][16 ct + l%16]).
  // The result goes back through LDS so each column segment is written whole.
  // The W chunks of this slice and then its V chunks form one stream; the
  // next chunk's global loads are issued before the current chunk's MFMAs.
  const int lane = t & 63, wave = t >> 6;
  constexpr int LPT = RCH * NP / NT;   // elements of a chunk per thread
  int wlo, whi, vlo, vhi;
  slice_rows(B.L, J.P, slice, wlo, whi);
  slice_rows(B.k, J.P, slice, vlo, vhi);
  const int nw = (whi - wlo + RCH - 1) / RCH, nv = (vhi - vlo + RCH - 1) / RCH;
  auto chunk = [&](int ci, double*& M, long long& ld, int& r0, int& nr) {
    if (ci < nw) {
      M = B.W;
      ld = B.L;
      r0 = wlo + ci * RCH;
      nr = min(RCH, whi - r0);
    } else {
      M = B.V;
      ld = B.k;
      r0 = vlo + (ci - nw) * RCH;
      nr = min(RCH, vhi - r0);
    }
  };
  double pre[LPT];
  auto fetch = [&](int ci) {
    double* M;
    long long ld;
    int r0, nr;
    chunk(ci, M, ld, r0, nr);
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      const int e = t + NT * i, rr = e % RCH, c = e / RCH;
      pre[i] = (rr < nr && c < n) ? M[(r0 + rr) + (long long)colmap[c] * ld] : 0.0;
    }
  };
  if (nw + nv > 0) fetch(0);
  for (int ci = 0; ci < nw + nv; ++ci) {
    double* M;
    long long ld;
    int r0, nr;
    chunk(ci, M, ld, r0, nr);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      const int e = t + NT * i;
      X[e % RCH][e / RCH] = pre[i];
    }
    __syncthreads();
    if (ci + 1 < nw + nv) fetch(ci + 1);
    f64x4 y[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) y[u] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NP; k += 4) {
#pragma unroll
      for (int u = 0; u < TPW; ++u) {
        const int tau = wave * TPW + u, rt = tau % RT, ct = tau / RT;
        y[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(X[16 * rt + (lane & 15)][k + (lane >> 4)],
                                                    R[k + (lane >> 4)][16 * ct + (lane & 15)],
                                                    y[u], 0, 0, 0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int tau = wave * TPW + u, rt = tau % RT, ct = tau / RT;
#pragma unroll
      for (int q = 0; q < 4; ++q) X[16 * rt + (lane >> 4) + 4 * q][16 * ct + (lane & 15)] = y[u][q];
    }
    __syncthreads();
    for (int e = t; e < RCH * n; e += NT) {
      const int rr = e % RCH, c = e / RCH;
      if (rr < nr) M[(r0 + rr) + (long long)colmap[c] * ld] = X[rr][c];
    }
  }
}

// W = A (transpose = 0) or A^T (transpose = 1); A is m x n with ld m
__global__ void load_work_kernel(const double* __restrict__ A, int m, int n, int transpose,
                                 double* __restrict__ W) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;   // 32 x 8
  if (!transpose) {
    for (int j = ty; j < 32; j += 8) {
      const int r = bx + tx, c = by + j;
      if (r < m && c < n) W[r + (long long)c * m] = A[r + (long long)c * m];
    }
    return;
  }
  // W (n x m) = A^T: read A tile (rows bx.., cols by..), write transposed
  for (int j = ty; j < 32; j += 8) {
    const int r = bx + tx, c = by + j;
    tile[j][tx] = (r < m && c < n) ? A[r + (long long)c * m] : 0.0;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int wr = by + tx, wc = bx + j;   // W row = A column, W column = A row
    if (wr < n && wc < m) W[wr + (long long)wc * n] = tile[tx][j];
  }
}

__global__ void eye_kernel(double* __restrict__ V, int k) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < (long long)k * k) V[e] = (e % k) == (e / k) ? 1.0 : 0.0;
}

// squared column norms of W (one workgroup per column)
__global__ __launch_bounds__(256) void colnorm_kernel(const double* __restrict__ W, int L,
                                                      double* __restrict__ out) {
  __shared__ double s[256];
  const double* col = W + (long long)blockIdx.x * L;
  double a = 0.0;
  for (int r = threadIdx.x; r < L; r += 256) a += col[r] * col[r];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// dst[:, c] = sign[c] * src[:, perm[c]]
__global__ void permute_cols_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                    const int* __restrict__ perm, const double* __restrict__ sign,
                                    int rows, int cols) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)rows * cols) return;
  const int r = (int)(e % rows), c = (int)(e / rows);
  dst[e] = sign[c] * src[r + (long long)perm[c] * rows];
}

__global__ void fill_u64_kernel(unsigned long long* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// sign of the first entry with |v| >= max|v|/2 of each column (as +-1)
__global__ __launch_bounds__(256) void col_sign_kernel(const double* __restrict__ V, int rows,
                                                       double* __restrict__ sign) {
  __shared__ double sv[256];
  __shared__ int si[256];
  const double* col = V + (long long)blockIdx.x * rows;
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
  if (threadIdx.x == 0) sign[blockIdx.x] = (si[0] < rows && col[si[0]] < 0.0) ? -1.0 : 1.0;
}

// ---- batched helpers: one launch for every block of the batch ----

// block of global column gc (columns of all blocks concatenated)
__device__ __forceinline__ int col_block(const long long* __restrict__ noff, int nb, long long gc) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (noff[mid] <= gc) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// squared column norms of every block's current W into out[0, ncol) and of
// its V into out[ncol, 2 ncol) (one workgroup per column)
__global__ __launch_bounds__(256) void colnorm_batched_kernel(const BlockDesc* __restrict__ desc,
                                                              const long long* __restrict__ noff,
                                                              int nb, long long ncol,
                                                              double* __restrict__ out) {
  __shared__ double s[2][256];
  const long long gc = blockIdx.x;
  const int b = col_block(noff, nb, gc);
  const BlockDesc B = desc[b];
  const double* col = B.W + (gc - noff[b]) * B.L;
  const double* vcol = B.V + (gc - noff[b]) * B.k;
  double a = 0.0, v = 0.0;
  for (int r = threadIdx.x; r < B.L; r += 256) a += col[r] * col[r];
  for (int r = threadIdx.x; r < B.k; r += 256) v += vcol[r] * vcol[r];
  s[0][threadIdx.x] = a;
  s[1][threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s[0][threadIdx.x] += s[0][threadIdx.x + w];
      s[1][threadIdx.x] += s[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[gc] = s[0][0];
    out[ncol + gc] = s[1][0];
  }
}

// dst[:, c] = sign[c] * src[:, perm[c]] for every job (perm and sign hold
// the batch's columns, a job's at its block's column offset)
struct PermJob {
  const double* src;
  double* dst;
  int rows, cols;
  long long elem_begin, col_off;
};

__global__ void permute_batched_kernel(const PermJob* __restrict__ jobs, int nj, long long total,
                                       const int* __restrict__ perm,
                                       const double* __restrict__ sign) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  int lo = 0, hi = nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].elem_begin <= e) lo = mid;
    else hi = mid - 1;
  }
  const PermJob J = jobs[lo];
  const long long i = e - J.elem_begin;
  const int r = (int)(i % J.rows), c = (int)(i / J.rows);
  J.dst[i] = sign[J.col_off + c] * J.src[r + (long long)perm[J.col_off + c] * J.rows];
}

// col_sign_kernel for the V of every block (desc[b].V, k x k), divided by
// the column's norm: the accumulated rotations leave V's columns off unit
// length by the rounding of every rotation applied (a few 1e-13 after tens of
// sweeps over hundreds of columns), and W = A V carries the same factor, so
// one scaling of both restores the triplet
__global__ __launch_bounds__(256) void col_sign_batched_kernel(const BlockDesc* __restrict__ desc,
                                                               const long long* __restrict__ noff,
                                                               int nb, double* __restrict__ sign) {
  __shared__ double sv[256];
  __shared__ int si[256];
  const long long gc = blockIdx.x;
  const int b = col_block(noff, nb, gc);
  const int rows = desc[b].k;
  const double* col = desc[b].V + (gc - noff[b]) * rows;
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
  double n2 = 0.0;
  for (int r = threadIdx.x; r < rows; r += 256) {
    if (fabs(col[r]) >= half && r < first) first = r;
    n2 += col[r] * col[r];
  }
  si[threadIdx.x] = first;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) si[threadIdx.x] = min(si[threadIdx.x], si[threadIdx.x + w]);
    __syncthreads();
  }
  sv[threadIdx.x] = n2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sv[threadIdx.x] += sv[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double inv = sv[0] > 0.0 ? 1.0 / sqrt(sv[0]) : 1.0;
    sign[gc] = (si[0] < rows && col[si[0]] < 0.0) ? -inv : inv;
  }
}

unsigned nblk(long long n, int t) { return (unsigned)std::max<long long>(1, (n + t - 1) / t); }

// device memory from the context's cache, returned on scope exit
struct DevMem {
  Context& c;
  void* p = nullptr;
  DevMem(Context& ctx, size_t bytes) : c(ctx), p(ctx.dev_alloc(bytes)) {}
  ~DevMem() { c.dev_free(p); }
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
};

}  // namespace

int jacobi_svd_batched(Context& c, std::vector<JacobiBlock>& blocks, double tol, int max_sweeps,
                       int block_cols, int inner_sweeps, double* worst_out,
                       int precondition_min_k) {
  if (worst_out) *worst_out = 0.0;
  hipStream_t st = c.stream();
  const int nb = (int)blocks.size();
  if (nb == 0) return 0;
  SU2_CHECK(block_cols == 16 || block_cols == 32, kErrArgument, "jacobi: block width %d",
            block_cols);
  SU2_CHECK(inner_sweeps >= 1 && inner_sweeps <= INNER_SWEEPS, kErrArgument,
            "jacobi: inner sweeps %d", inner_sweeps);
  // pairs of 16-column blocks run one wave per workgroup (several per CU);
  // pairs of 32-column blocks run four waves
  const int CB = block_cols, NP = 2 * CB;
  const int slice_target = CB == 16 ? 1024 : 256;
  std::vector<int> L(nb), K(nb);
  std::vector<long long> noff(nb + 1, 0);
  for (int b = 0; b < nb; ++b) {
    SU2_CHECK(blocks[b].m >= 1 && blocks[b].n >= 1, kErrArgument, "jacobi: empty block");
    L[b] = std::max(blocks[b].m, blocks[b].n);
    K[b] = std::min(blocks[b].m, blocks[b].n);
    noff[b + 1] = noff[b] + K[b];
  }
  const long long ncol = noff[nb];
  SU2_CHECK(ncol < (1LL << 31), kErrCapacity, "jacobi: %lld columns", ncol);
  // second buffers: each sweep starts by moving the columns into descending
  // norm order (de Rijk's ordering; it cuts the sweep count severalfold)
  std::vector<std::unique_ptr<DevMem>> W2(nb), V2(nb);
  std::vector<double*> Wc(nb), Vc(nb), Wo(nb), Vo(nb);
  for (int b = 0; b < nb; ++b) {
    JacobiBlock& J = blocks[b];
    W2[b].reset(new DevMem(c, sizeof(double) * (size_t)L[b] * K[b]));
    V2[b].reset(new DevMem(c, sizeof(double) * (size_t)K[b] * K[b]));
    Wc[b] = J.W;
    Vc[b] = J.V;
    Wo[b] = (double*)W2[b]->p;
    Vo[b] = (double*)V2[b]->p;
    hipLaunchKernelGGL(load_work_kernel, dim3(nblk(J.m, 32), nblk(J.n, 32)), dim3(32, 8), 0, st,
                       J.A, J.m, J.n, J.m < J.n ? 1 : 0, J.W);
    hipLaunchKernelGGL(eye_kernel, dim3(nblk((long long)K[b] * K[b], 256)), dim3(256), 0, st, J.V,
                       K[b]);
  }
  SU2_HIP(hipGetLastError());
  DevMem dn(c, sizeof(double) * 2 * ncol), dperm(c, sizeof(int) * ncol),
      done(c, sizeof(double) * ncol), dsign(c, sizeof(double) * ncol),
      dnoff(c, sizeof(long long) * (nb + 1)), dpj(c, sizeof(PermJob) * 2 * nb);
  std::vector<double> hn(2 * ncol);   // W then V squared column norms
  std::vector<int> hp(ncol);
  {
    std::vector<double> ones(ncol, 1.0);
    SU2_HIP(hipMemcpyAsync(done.p, ones.data(), sizeof(double) * ncol, hipMemcpyHostToDevice, st));
    SU2_HIP(hipMemcpyAsync(dnoff.p, noff.data(), sizeof(long long) * (nb + 1),
                           hipMemcpyHostToDevice, st));
  }
  std::vector<BlockDesc> desc(nb);
  DevMem ddesc(c, sizeof(BlockDesc) * nb);
  auto upload_desc = [&]() {
    for (int b = 0; b < nb; ++b) {
      desc[b].W = Wc[b];
      desc[b].V = Vc[b];
    }
    SU2_HIP(hipMemcpyAsync(ddesc.p, desc.data(), sizeof(BlockDesc) * nb, hipMemcpyHostToDevice,
                           st));
  };
  // squared column norms of the current W of every block, to the host
  auto norms = [&]() {
    hipLaunchKernelGGL(colnorm_batched_kernel, dim3((unsigned)ncol), dim3(256), 0, st,
                       (const BlockDesc*)ddesc.p, (const long long*)dnoff.p, nb, (long long)ncol,
                       (double*)dn.p);
    SU2_HIP(hipMemcpyAsync(hn.data(), dn.p, sizeof(double) * 2 * ncol, hipMemcpyDeviceToHost,
                           st));
    SU2_HIP(hipStreamSynchronize(st));
  };
  // src -> dst column moves of W and V (or the given jobs) in one launch
  auto permute = [&](const std::vector<PermJob>& pj0, const double* sign) {
    std::vector<PermJob> pj = pj0;
    long long tot = 0;
    for (auto& j : pj) {
      j.elem_begin = tot;
      tot += (long long)j.rows * j.cols;
    }
    if (tot == 0) return;
    SU2_HIP(hipMemcpyAsync(dpj.p, pj.data(), sizeof(PermJob) * pj.size(), hipMemcpyHostToDevice,
                           st));
    hipLaunchKernelGGL(permute_batched_kernel, dim3(nblk(tot, 256)), dim3(256), 0, st,
                       (const PermJob*)dpj.p, (int)pj.size(), tot, (const int*)dperm.p, sign);
  };
  for (int b = 0; b < nb; ++b) desc[b] = BlockDesc{Wc[b], Vc[b], L[b], K[b], 0.0};
  upload_desc();
  norms();
  for (int b = 0; b < nb; ++b) {
    double g = 0;
    for (int i = 0; i < K[b]; ++i) g += hn[noff[b] + i];
    desc[b].normF = std::sqrt(g);
  }
  // descending column norms of every block (de Rijk's ordering), W and V
  // permuted alike
  // blocks still iterating: a block leaves after its first sweep without a
  // rotation (later sweeps would leave it unchanged: its columns stay in
  // order and no pair rotates), so converged blocks cost no further rounds
  std::vector<char> active(nb, 1);
  auto sort_permute = [&]() {
    std::vector<PermJob> pj;
    for (int b = 0; b < nb; ++b) {
      if (!active[b]) continue;
      int* p = hp.data() + noff[b];
      const double* w = hn.data() + noff[b];
      std::iota(p, p + K[b], 0);
      std::stable_sort(p, p + K[b], [&](int x, int y) { return w[x] > w[y]; });
      pj.push_back(PermJob{Wc[b], Wo[b], L[b], K[b], 0, noff[b]});
      pj.push_back(PermJob{Vc[b], Vo[b], K[b], K[b], 0, noff[b]});
    }
    if (pj.empty()) return;
    SU2_HIP(hipMemcpyAsync(dperm.p, hp.data(), sizeof(int) * ncol, hipMemcpyHostToDevice, st));
    permute(pj, (const double*)done.p);
    for (int b = 0; b < nb; ++b) {
      if (!active[b]) continue;
      std::swap(Wc[b], Wo[b]);
      std::swap(Vc[b], Vo[b]);
    }
    upload_desc();
  };
  // QR preconditioning (precond.hpp) of every block with precondition_min_k
  // <= K <= kPrecondMaxK: columns sorted, then W <- W P Q2 and V <- P Q2.  A block's
  // own size decides, so its bits do not depend on the rest of the batch (a
  // block left out is sorted here instead of at the first sweep: the same
  // permutation, and the first sweep's sort is then the identity).
  if (precondition_min_k > 0) {
    std::vector<PrecondBlock> pb;
    std::vector<int> pre;
    for (int b = 0; b < nb; ++b)
      if (K[b] >= precondition_min_k && K[b] <= kPrecondMaxK) pre.push_back(b);
    if (!pre.empty()) {
      sort_permute();
      for (int b : pre)
        pb.push_back(PrecondBlock{Wc[b], Vc[b], Wo[b], Vo[b], L[b], K[b], desc[b].normF});
      precondition_jacobi(c, pb);
      for (int b : pre) {
        std::swap(Wc[b], Wo[b]);
        std::swap(Vc[b], Vo[b]);
      }
      upload_desc();
      norms();
    }
  }
  // Deflation: the columns of a block above the rounding level, kept[b] of
  // them, lead the block after the descending sort, and only they are
  // scheduled.  A column at the rounding level is never rotated again (the
  // rotate launch skips its pairs), so kept[b] only falls; the QR
  // preconditioner puts a rank-deficient block's null directions into such
  // columns at once (a 16x8 D=2000 centre sector: 626 of 1383).  The count
  // comes from the block's own norms, so the bits do not depend on the batch.
  std::vector<int> kept(K.begin(), K.end());
  auto count_kept = [&]() {
    bool changed = false;
    for (int b = 0; b < nb; ++b) {
      if (!active[b]) continue;
      const double thr = kNoiseRel * desc[b].normF * (kNoiseRel * desc[b].normF);
      int k = 0;
      for (int i = 0; i < K[b]; ++i) k += hn[noff[b] + i] > thr;
      changed = changed || k != kept[b];
      kept[b] = k;
    }
    return changed;
  };
  count_kept();
  // round-robin schedule of column-block pairs of the active blocks' kept
  // columns, merged across blocks by round
  std::vector<std::vector<PairJob>> rounds;
  std::vector<int> round_off, round_pmax;
  std::vector<PairJob> all;
  size_t gpart_max = 0;
  auto build_schedule = [&]() {
    rounds.clear();
    all.clear();
    gpart_max = 0;
    for (int b = 0; b < nb; ++b) {
      if (!active[b]) continue;
      const int kb = kept[b];
      const int ncb = (kb + CB - 1) / CB;
      auto width = [&](int i) { return i >= ncb ? 0 : std::min(CB, kb - i * CB); };
      if (ncb <= 1) {
        if (rounds.empty()) rounds.resize(1);
        if (kb >= 2) rounds[0].push_back(PairJob{b, 0, kb, 0, 0, 1});
        continue;
      }
      const int mp = ncb + (ncb & 1);
      if ((int)rounds.size() < mp - 1) rounds.resize(mp - 1);
      for (int r = 0; r < mp - 1; ++r)
        for (int k = 0; k < mp / 2; ++k) {
          int P = (k == 0) ? mp - 1 : (r + k) % (mp - 1);
          int Q = (k == 0) ? r : (r - k + (mp - 1)) % (mp - 1);
          if (width(P) == 0) std::swap(P, Q);
          if (width(P) == 0) continue;
          rounds[r].push_back(PairJob{b, P * CB, width(P), width(Q) ? Q * CB : 0, width(Q), 1});
        }
    }
    // Row slices: a group (the sectors of one wavefunction) with few pairs in
    // a round spreads each pair over up to MAX_SLICES workgroups (at least 64
    // rows each), about slice_target workgroups per group and round, so that
    // the round fills the chip.  The slice count of a pair follows from its own
    // group alone, never from the other groups of the batch: the slices'
    // partial Gram matrices are summed in slice order, so the slicing decides
    // the bits, and a wavefunction gets the same bits alone and in a batch.
    round_off.assign(rounds.size() + 1, 0);
    round_pmax.assign(rounds.size(), 1);
    for (size_t r = 0; r < rounds.size(); ++r) {
      round_off[r] = (int)all.size();
      const int nj = (int)rounds[r].size();
      std::map<int, int> per_group;
      for (const PairJob& j : rounds[r]) per_group[blocks[j.blk].group]++;
      int used = 1;
      for (PairJob& j : rounds[r]) {
        const int pmax =
            std::min(MAX_SLICES, std::max(1, slice_target / per_group[blocks[j.blk].group]));
        j.P = std::max(1, std::min(pmax, L[j.blk] / MIN_SLICE_ROWS));
        used = std::max(used, j.P);
      }
      round_pmax[r] = used;
      gpart_max = std::max(gpart_max, (size_t)nj * used);
      all.insert(all.end(), rounds[r].begin(), rounds[r].end());
    }
    round_off[rounds.size()] = (int)all.size();
  };
  // the full schedule sizes the buffers; a smaller one replaces it in place
  // (a pair of a smaller schedule may take more slices, never more than
  // MAX_SLICES, and no round has more pairs than the full one's)
  build_schedule();
  size_t gpart_cap = 1;
  for (size_t r = 0; r + 1 < round_off.size(); ++r)
    gpart_cap = std::max(gpart_cap, (size_t)(round_off[r + 1] - round_off[r]) * MAX_SLICES);
  DevMem dgpart(c, sizeof(double) * NP * NP * gpart_cap);
  DevMem djobs(c, sizeof(PairJob) * std::max<size_t>(1, all.size())),
      doff(c, sizeof(unsigned long long) * nb);
  auto upload_jobs = [&]() {
    if (!all.empty())
      SU2_HIP(hipMemcpyAsync(djobs.p, all.data(), sizeof(PairJob) * all.size(),
                             hipMemcpyHostToDevice, st));
  };
  upload_jobs();
  std::vector<unsigned long long> hoff(nb);
  bool converged = all.empty();
  double worst = 0;
  int sweeps_done = 0;
  for (int sweep = 0; sweep < max_sweeps && !converged; ++sweep) {
    ++sweeps_done;
    // descending column norms (hn: read back with the previous sweep's
    // convergence counters, or before the first sweep)
    sort_permute();
    hipLaunchKernelGGL(fill_u64_kernel, dim3(nblk(nb, 256)), dim3(256), 0, st,
                       (unsigned long long*)doff.p, nb);
    for (size_t r = 0; r < rounds.size(); ++r) {
      const int nj = round_off[r + 1] - round_off[r];
      if (nj == 0) continue;
      const int pm = round_pmax[r];
      const PairJob* jr = (const PairJob*)djobs.p + round_off[r];
      if (CB == 16) {
        hipLaunchKernelGGL((gram_partial_kernel<16, 64, RCH>), dim3(nj * pm), dim3(64), 0, st,
                           (const BlockDesc*)ddesc.p, jr, pm, (double*)dgpart.p);
        hipLaunchKernelGGL((jacobi_rotate_kernel<16, 64>), dim3(nj * pm), dim3(64), 0, st,
                           (const BlockDesc*)ddesc.p, jr, pm, (const double*)dgpart.p, tol, 1e-14,
                           (unsigned long long*)doff.p, inner_sweeps);
      } else {
        hipLaunchKernelGGL((gram_partial_kernel<32, 1024, 64>), dim3(nj * pm), dim3(1024), 0, st,
                           (const BlockDesc*)ddesc.p, jr, pm, (double*)dgpart.p);
        hipLaunchKernelGGL((jacobi_rotate_kernel<32, 512>), dim3(nj * pm), dim3(512), 0, st,
                           (const BlockDesc*)ddesc.p, jr, pm, (const double*)dgpart.p, tol, 1e-14,
                           (unsigned long long*)doff.p, inner_sweeps);
      }
    }
    SU2_HIP(hipGetLastError());
    SU2_HIP(hipMemcpyAsync(hoff.data(), doff.p, sizeof(unsigned long long) * nb,
                           hipMemcpyDeviceToHost, st));
    // the column norms for the next sweep's order (or the final one) come
    // back with the counters: one wait per sweep
    hipLaunchKernelGGL(colnorm_batched_kernel, dim3((unsigned)ncol), dim3(256), 0, st,
                       (const BlockDesc*)ddesc.p, (const long long*)dnoff.p, nb, (long long)ncol,
                       (double*)dn.p);
    SU2_HIP(hipMemcpyAsync(hn.data(), dn.p, sizeof(double) * 2 * ncol, hipMemcpyDeviceToHost,
                           st));
    SU2_HIP(hipStreamSynchronize(st));
    converged = true;
    worst = 0;
    bool left = false;
    for (int b = 0; b < nb; ++b) {
      if (!active[b]) continue;
      double v;
      std::memcpy(&v, &hoff[b], sizeof v);
      if (v >= 2.0) {   // some pair still rotated this sweep
        converged = false;
        worst = std::max(worst, v - 2.0);
      } else {
        active[b] = 0;
        left = true;
      }
    }
    // columns that fell to the rounding level leave the schedule too
    const bool shrunk = count_kept();
    if ((left || shrunk) && !converged) {
      build_schedule();
      SU2_CHECK(gpart_max <= gpart_cap, kErrCapacity, "jacobi: %zu partial Gram matrices",
                gpart_max);
      upload_jobs();
    }
  }
  // Out of sweeps: the columns are accepted when orthogonal to 1e-6 of their
  // norms -- the singular values then carry errors of order 1e-12 relative
  // (second order in the residual overlaps), harmless for a truncation --
  // and the shortfall is reported through *worst_out; beyond that it is an
  // error.
  SU2_CHECK(converged || worst <= 1e-6, kErrNumeric,
            "jacobi svd: no convergence after %d sweeps (worst |w_i.w_j|/|w_i||w_j| = %.2e)",
            max_sweeps, worst);
  if (worst_out) *worst_out = converged ? 0.0 : worst;

  // singular values in ascending order, sign convention on V (W follows),
  // results into the caller's buffers -- for every block in a few launches;
  // hn holds the final column norms
  std::vector<int> ident(ncol);
  for (int b = 0; b < nb; ++b) {
    JacobiBlock& J = blocks[b];
    const int k = K[b];
    // s_i = |W e_i| / |V e_i|: the singular value of the normalised triplet
    std::vector<double> s2(k);
    for (int i = 0; i < k; ++i) {
      const double v2 = hn[ncol + noff[b] + i];
      s2[i] = v2 > 0.0 ? hn[noff[b] + i] / v2 : hn[noff[b] + i];
    }
    int* perm = hp.data() + noff[b];
    std::iota(perm, perm + k, 0);
    std::iota(ident.data() + noff[b], ident.data() + noff[b] + k, 0);
    std::stable_sort(perm, perm + k, [&](int x, int y) { return s2[x] < s2[y]; });
    J.s.resize(k);
    for (int i = 0; i < k; ++i) J.s[i] = std::sqrt(s2[perm[i]]);
  }
  DevMem dident(c, sizeof(int) * ncol);
  SU2_HIP(hipMemcpyAsync(dperm.p, hp.data(), sizeof(int) * ncol, hipMemcpyHostToDevice, st));
  SU2_HIP(hipMemcpyAsync(dident.p, ident.data(), sizeof(int) * ncol, hipMemcpyHostToDevice, st));
  // V: sorted copy in the spare buffer, its signs
  {
    std::vector<PermJob> pj;
    for (int b = 0; b < nb; ++b) pj.push_back(PermJob{Vc[b], Vo[b], K[b], K[b], 0, noff[b]});
    permute(pj, (const double*)done.p);
  }
  for (int b = 0; b < nb; ++b) {
    desc[b].W = Wc[b];
    desc[b].V = Vo[b];
  }
  SU2_HIP(hipMemcpyAsync(ddesc.p, desc.data(), sizeof(BlockDesc) * nb, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(col_sign_batched_kernel, dim3((unsigned)ncol), dim3(256), 0, st,
                     (const BlockDesc*)ddesc.p, (const long long*)dnoff.p, nb, (double*)dsign.p);
  // V with its signs into J.V (or the other buffer, when J.V is the spare),
  // W with the same order and signs into J.W (or the spare)
  std::vector<PermJob> vj, wj;
  std::vector<VecMove> moves;
  for (int b = 0; b < nb; ++b) {
    JacobiBlock& J = blocks[b];
    double* vdst = J.V == Vo[b] ? Vc[b] : J.V;
    vj.push_back(PermJob{Vo[b], vdst, K[b], K[b], 0, noff[b]});
    double* wdst = J.W == Wc[b] ? Wo[b] : J.W;
    wj.push_back(PermJob{Wc[b], wdst, L[b], K[b], 0, noff[b]});
    if (wdst != J.W) moves.push_back(VecMove{wdst, J.W, (long long)L[b] * K[b], kMoveF64, 0});
    if (vdst != J.V) moves.push_back(VecMove{vdst, J.V, (long long)K[b] * K[b], kMoveF64, 0});
  }
  {
    // the V moves read the identity permutation, the W moves the sort order
    // (the second upload of the job table is queued behind the first launch)
    std::swap(dperm.p, dident.p);
    permute(vj, (const double*)dsign.p);
    std::swap(dperm.p, dident.p);
    permute(wj, (const double*)dsign.p);
  }
  if (!moves.empty()) {
    long long mx = 0;
    for (auto& m : moves) mx = std::max(mx, m.n);
    DevMem dm(c, sizeof(VecMove) * moves.size());
    SU2_HIP(hipMemcpyAsync(dm.p, moves.data(), sizeof(VecMove) * moves.size(),
                           hipMemcpyHostToDevice, st));
    dev_move_batch(c, (const VecMove*)dm.p, (int)moves.size(), mx);
  }
  // no wait: the outputs are used in stream order, and the host results (s)
  // are already on the host
  SU2_HIP(hipGetLastError());
  return sweeps_done;
}

}  // namespace su2
