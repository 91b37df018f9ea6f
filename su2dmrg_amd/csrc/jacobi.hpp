// One-sided block Jacobi SVD (Hestenes) of many dense blocks at once.
//
// The reference obtains each sector's singular vectors from the Gram matrix
// with dsyev and then re-diagonalises the small-eigenvalue subspace again and
// again to recover their accuracy (tensor.cpp:1472-1529).  Here the columns
// of the block itself are rotated until mutually orthogonal: for a block A
// (m x n) with k = min(m, n) the working matrix W is A (m >= n) or A^T, and
// the accumulated rotation V (k x k) converges to the Gram eigenvectors, so
// that W = A V holds the reference's u = A v and the column norms of W are
// the singular values, accurate relative to each value rather than to the
// largest one.
//
// Columns are cut into blocks of 16 (or 32).  A sweep visits every pair of blocks
// once, in round-robin rounds of disjoint pairs; two kernel launches run all
// pairs of a round for every block of the batch.  The first forms partial
// Gram matrices (up to 64 x 64) of each pair over row slices of W, the second
// sums them, diagonalises the pair's Gram matrix in LDS by parallel cyclic
// Jacobi (one to three inner sweeps per visit) and applies the eigenvector
// matrix to the pair's columns of W and V, each slice to its own rows.  A
// round with few pairs cuts every pair into up to 8 slices so that it spans
// the chip.  Before each sweep the columns are put in descending norm order.
// Sweeps stop when every pair is orthogonal to
// |g_ij| <= max(tol * sqrt(g_ii g_jj), 1e-14 ||A||_F (sqrt g_ii + sqrt g_jj))
// (the absolute floor ends the iteration on columns that carry rounding
// errors of order eps ||A||); columns with norm <= 1e-13 ||A||_F are left
// out of the test.  A block that fits one column block is one self-pair per
// sweep.
#pragma once

#include <vector>

namespace su2 {

class Context;

struct JacobiBlock {
  const double* A;   // m x n, column-major, leading dimension m (device)
  int m, n;
  double* W;         // out: L x k (L = max(m, n), k = min(m, n)), ld L, columns by
                     //      ascending singular value (device, caller-allocated)
  double* V;         // out: k x k, ld k, matching columns (device, caller-allocated)
  std::vector<double> s;   // out: k singular values, ascending
  int group = 0;           // blocks of one group (one wavefunction) share the
                           // row slicing of their pairs; groups never affect
                           // each other's bits
};

// Decomposes every block and returns the number of sweeps it took.  W and V columns are sorted ascending by singular value; each
// column of V has its first entry with |v| >= max|v|/2 made positive (the
// LAPACK routine the reference calls leaves signs open), W's columns follow.
// block_cols: 16 (pairs of 32 columns, one wave per pair) or 32 (pairs of 64
// columns, eight waves per pair); inner_sweeps: 1..3 inner Jacobi sweeps per
// visit of a pair.
// worst_out (nullable) receives 0 when every block converged, else the
// largest remaining |w_i.w_j| / |w_i||w_j| (accepted up to 1e-6, beyond it
// the call throws).
// precondition_min_k: QR-precondition the blocks with precondition_min_k <=
// min(m, n) <= kPrecondMaxK first (precond.hpp; 0: none); the results differ
// from the plain iteration's by rounding only.
int jacobi_svd_batched(Context& c, std::vector<JacobiBlock>& blocks, double tol = 1e-13,
                       int max_sweeps = 60, int block_cols = 32, int inner_sweeps = 1,
                       double* worst_out = nullptr, int precondition_min_k = 0);

}  // namespace su2
