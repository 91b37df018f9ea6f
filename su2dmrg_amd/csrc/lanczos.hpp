// Lanczos ground-state solver (lanczos_su2::lanczos1, lanczos_su2.cpp:125-195,
// with diatridiag 237-259, compute_eigenvector 262-279 and check_eigenvector
// 320-332).  Krylov vectors live on the device; each iteration is one H_eff
// matvec plus deterministic device dot products and axpys (full classical
// Gram-Schmidt against all previous vectors, as the reference does).
#pragma once

#include <functional>
#include <vector>

namespace su2 {

class Context;

struct LanczosResult {
  double energy = 0;   // lowest Ritz value of the final tridiagonal, or the
                       // Rayleigh quotient of check_apply (see below)
  double ritz = 0;     // the Ritz value itself
  double residual = -1;  // ||H psi - E psi|| from the last check (-1: none)
  int matvecs = 0;
  int check_matvecs = 0;  // of them, through check_apply
};

typedef std::function<void(const double*, double*)> ApplyFn;

// apply(x, y): y = H x on device buffers of length n.  psi holds the start
// vector on entry and the normalised Ritz vector on exit.  check_apply
// (nullable) replaces apply in the Ritz-vector check (check_eigenvector,
// lanczos_su2.cpp:320-332): a mixed-precision solve builds its Krylov space
// with an FP32 apply and checks with an FP64 one, whose Rayleigh quotient
// <psi|H psi> then is the energy returned (an FP64 correction of the Ritz
// value, accurate to second order in the Ritz vector's error).
LanczosResult lanczos_ground(Context& c, const ApplyFn& apply, double* psi, long long n, int mlanc,
                             double tol = 1e-6, const ApplyFn* check_apply = nullptr);

// One Lanczos solve of a batch: the start vector in psi (overwritten with the
// normalised Ritz vector), its length, the Krylov size, and whether its
// Krylov matvecs run in FP32 with FP64 checks (mixed precision).
struct LanczosJob {
  double* psi = nullptr;
  long long n = 0;
  int mlanc = 5;
  double tol = 1e-6;
  bool mixed = false;
  LanczosResult res;
};

// The matvecs of one round: request r asks y = H_job x with the job's Krylov
// arithmetic (check = false) or FP64 (check = true).
struct LanczosReq {
  int job;
  const double* x;
  double* y;
  bool check;
};
typedef std::function<void(const std::vector<LanczosReq>&)> BatchApply;

// The reference's Lanczos schedule (as lanczos_ground) for every job in
// lockstep: each round one batched apply for all unfinished jobs, their
// vector algebra queued, one host wait for all of their scalars.
void lanczos_batch(Context& c, std::vector<LanczosJob>& jobs, const BatchApply& apply);

// Eigenpairs of a real symmetric tridiagonal matrix (host; diagonal d, off-
// diagonal e, size m): values ascending, vectors column-major m x m.
void tridiag_eig(const std::vector<double>& d, const std::vector<double>& e,
                 std::vector<double>& w, std::vector<double>& z);

}  // namespace su2
