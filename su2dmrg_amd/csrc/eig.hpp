// FP64 symmetric eigensolver on the device (replaces dsyev_ behind
// obtain_symmetric_matrix_eigenvector, tensor.cpp:731-759).
//
// Small matrices (n <= 64) are diagonalised by cyclic Jacobi in LDS, one
// 64-lane wave per matrix, many matrices per launch.  Larger ones use block
// Jacobi: the matrix is cut into 16-wide blocks, each round pairs the blocks
// round-robin, diagonalises every 32x32 pair problem with the LDS solver, and
// applies the rotations to the matrix and the eigenvector basis with the
// grouped FP64 MFMA GEMM (two launches per round, no host sync).  Sweeps
// repeat until the off-diagonal norm falls below 1e-15 of the total.
// Eigenvalues come back ascending with matching eigenvector columns, the
// order dsyev_ uses.
#pragma once

#include <vector>

namespace su2 {

class Context;

struct EigJob {
  const double* A;   // input matrix (column-major)
  long long lda;
  int p0, bp;        // rows/cols p0..p0+bp-1 ...
  int q0, bq;        // ... followed by q0..q0+bq-1 form the problem (bq may be 0)
  double* V;         // output eigenvectors (bp+bq square, leading dim ldv)
  long long ldv;
  double* W;         // output eigenvalues (nullable)
  int sort;          // 1: ascending order
};

// Batched small solves (bp+bq <= 64) on device job descriptors.
void eig_small_batched(Context& c, const EigJob* d_jobs, int njob);

// Sign convention of eigenvector columns: the first entry with |v| >= max|v|/2
// is made positive (the LAPACK routine the reference calls leaves signs to the
// implementation).
void dev_fix_signs(Context& c, double* V, int rows, int cols, long long ld);

// Full eigendecomposition of the n x n symmetric matrix A (device, leading
// dimension n; destroyed).  V (device, n x n, ld n) receives eigenvectors,
// w (host) eigenvalues ascending.
void sym_eig(Context& c, double* A, int n, double* V, std::vector<double>& w);

}  // namespace su2
