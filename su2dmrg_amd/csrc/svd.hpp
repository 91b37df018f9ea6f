// Truncated SVD of a two-leg singlet wavefunction across its sectors
// (tensor_su2::svd, tensor_su2.cpp:1619-1758, with tensor::svd per sector,
// tensor.cpp:1472-1700).
//
// Per sector j with block A (m x n): the Gram matrix A^T A (m >= n) or A A^T
// is formed with the grouped FP64 MFMA GEMM and diagonalised on the device;
// eigenpairs with w/w_max < 0.1 are re-diagonalised in their projected Gram
// until one remains (tensor.cpp:1498-1529); the other factor is A V (or
// A^T U); singular values are sqrt(w) for w/w_max > 1e-100; kept columns are
// scaled by s^(p1-1)/s^p2 (or s^p1/s^(p2-1)).  The largest max_dcut values
// over all sectors survive, ordered by the reference's shell sort so ties at
// the cut fall the same way (sort2.f:76-111).
#pragma once

#include <memory>
#include <vector>

#include "structure.hpp"
#include "tensor.hpp"

namespace su2 {

struct SvdResult {
  Leg bond;                              // kept sectors (2j ascending), dir -1
  std::unique_ptr<Tensor> U;             // [psi leg 0, bond]
  std::unique_ptr<Tensor> V;             // [psi leg 1, bond]
  std::vector<std::vector<double>> kept; // ascending singular values per bond sector
  double discarded_weight = 0;           // sum of squares of the cut values
  // the reference's wout (tensor_su2.cpp:1707-1731): kept values sector by
  // sector, then the cut ones sector by sector, each ascending
  std::vector<double> wout;
  int jacobi_sweeps = 0;      // one-sided Jacobi sweeps the sectors took
  double t_jacobi = 0;        // seconds inside the Jacobi SVD
  double jacobi_worst = 0;    // 0, or the residual overlap accepted after the
                              // sweep ceiling (no convergence to 1e-13)
};

SvdResult truncated_svd(Context& c, const Tensor& psi, double p1, double p2, int max_dcut);

// Several wavefunctions at once: every sector of every one in one batched
// Jacobi SVD, then each one's keep rule and cut.  Each result equals the
// single call's (its blocks' rotations depend on those blocks alone).
struct SvdJob {
  const Tensor* psi;
  double p1, p2;
  int max_dcut;
};
std::vector<SvdResult> truncated_svd_batch(Context& c, const std::vector<SvdJob>& jobs);

}  // namespace su2
