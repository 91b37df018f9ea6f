// Preconditioner of the one-sided Jacobi SVD (after Drmač and Veselić's QR
// preconditioning, with the first triangular factor taken from the Gram
// matrix).
//
// For a block W (L x K, L >= K, columns already in descending norm order):
//   G  = W^T W                          (grouped FP64 GEMM)
//   G ~ Rc^T Rc                         (upper Cholesky by 16-row panels; a
//                                        pivot at or below `tiny` gives a
//                                        zero row, so a semidefinite G is
//                                        handled)
//   Rc^T = Q2 R2                        (Householder QR by 16-column panels,
//                                        compact WY form I - V T V^T)
//   W <- W Q2,  V <- V Q2               (grouped FP64 GEMMs)
// Q2 is a product of Householder reflectors, so it is orthogonal to rounding
// whatever the accuracy of Rc; the Jacobi iteration that follows runs on the
// true W Q2, so the SVD's accuracy is that of the plain iteration.  Only the
// number of sweeps depends on Rc: W Q2 = Q R2^T with R2^T lower triangular
// and graded, whose columns the Jacobi rotations orthogonalise in a few
// sweeps (on 16x8 D=2000 centre-bond sectors: 8 sweeps against 22 to 26
// unpreconditioned, in a CPU model of this block Jacobi).
#pragma once

#include <vector>

namespace su2 {

class Context;

struct PrecondBlock {
  const double* W;  // L x K, ld L
  const double* V;  // K x K, ld K
  double* Wout;     // L x K, ld L: receives W Q2
  double* Vout;     // K x K, ld K: receives V Q2
  int L, K;
  double normF;     // Frobenius norm of W (the Cholesky pivot floor is 1e-15 normF^2)
};

// Largest K the panel kernels take (two rows per thread of 1024, held in
// registers; a third row would not fit them and spilled).
constexpr int kPrecondMaxK = 2048;

// Preconditions every block (each must have 1 <= K <= kPrecondMaxK and
// L >= K): Wout = W Q2, Vout = V Q2.  Blocks never affect each other's bits.
void precondition_jacobi(Context& c, std::vector<PrecondBlock>& blocks);

}  // namespace su2
