// Clebsch-Gordan tensors and recoupling scalars (host, FP64).
//
// cgc(j1,j2,j3) is <j1 m1; j2 m2 | j3 m3> with doubled spins from Racah's
// closed form, dense column-major [j1+1][j2+1][j3+1] with index i <-> m=2i-j
// (the layout of tensor::make_cgc, tensor.cpp:1353-1407).  singlet(j) is
// (-1)^((j-m)/2)/sqrt(j+1) on m1+m2=0 (tensor::make_singlet, 1327-1350).
#pragma once

#include <vector>

namespace su2 {

std::vector<double> cg_tensor(int j1, int j2, int j3);
std::vector<double> singlet_tensor(int j);

// Recoupling scalar of one H_eff term block (ket 2j=i, operator 2j=k, bra
// 2j=f): <C, C turned on every leg by the singlet> * sqrt(k+1), as defined in
// dmrg_su2.cpp:759-775.  Zero outside the triangle or for parity mismatch.
double heff_recoupling(int i, int k, int f);

}  // namespace su2
