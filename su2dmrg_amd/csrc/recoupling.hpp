// Recoupling scalars of the environment kernels.
//
// The reference fills its fac_* tables by contracting dense CG tensors in the
// leg pattern of each tensor::contract_dmrg_* routine and taking the ratio to
// the output CG tensor (dmrg_su2.cpp:777-943).  Each function below performs
// that contraction for one label tuple (cached), with the leg bookkeeping of:
//   op_transform_*  contract_dmrg_operator_transformation (tensor_dmrg_src.cpp:275-360)
//   op_initial_*    contract_dmrg_operator_initial        (tensor_dmrg_src.cpp:69-101)
//   pairup_*        contract_dmrg_operator_pairup         (tensor_dmrg_src.cpp:169-210)
// Labels are doubled spins: l/l' block bond, s site, a/a' fused leg, r right
// bond, k operator leg.  psi_left_to_right/psi_right_to_left move a singlet
// wavefunction's coupling across one site (used for Lanczos start vectors).
#pragma once

namespace su2 {

double rc_op_transform_left(int jl, int ja, int jlp, int jap, int k, int s);
double rc_op_transform_right(int jr, int ja, int jrp, int jap, int k, int s);
double rc_op_initial_left(int jl, int ja, int jap, int k, int s0, int s1);
double rc_op_initial_right(int jr, int ja, int jap, int k, int s0, int s1);
double rc_pairup_left(int ja, int jl, int jlp, int k, int s0, int s1);
double rc_pairup_right(int ja, int jr, int jrp, int k, int s0, int s1);
// Plaquette (Q) terms: a block vector operator times the site spin coupled
// to rank K (2K = 0, 2, 4), and a block rank-K operator times the site spin
// coupled back to rank 1 (the reference's permutation with the qterm
// tensors, tensor_su2_hamiltonian.cpp:19-76, 871-1108).  The rank-1 result is
// sum_K rc_reduce_weight(K) * reduce(X^K): with these weights it is the block
// operator times (s_b . s_c) in the normalisation of the spin operators.
double rc_couple_left(int jl, int ja, int jlp, int jap, int K, int s);
double rc_couple_right(int jr, int ja, int jrp, int jap, int K, int s);
double rc_reduce_left(int jl, int ja, int jlp, int jap, int K, int s);
double rc_reduce_right(int jr, int ja, int jrp, int jap, int K, int s);
double rc_reduce_weight(int K);
double rc_psi_left_to_right(int jl, int s, int jr);
double rc_psi_right_to_left(int jl, int s, int jr);

}  // namespace su2
