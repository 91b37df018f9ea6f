// Reference-format checkpoint files (dmrg_su2::save_mps2 / read_mps /
// save_enr / read_enr, dmrg_su2_omp.cpp:864-1152): helpers shared by the
// checkpoint code and the excited-state loader.
#pragma once

#include <string>
#include <vector>

#include "structure.hpp"

namespace su2 {
namespace mpsio {

// one three-leg MPS file: legs, present-block flags, reduced and CG bodies
struct File3 {
  Leg leg[3];
  std::vector<int> flag;
  std::vector<double> red, cg;
};

// "{dir}/uu-{ns}-{D}-{2S}-{ky}-{exci}-" (the site index follows)
std::string uu_base(const std::string& dir, int ns, int d, int s, int ky, int exci);
// "{ns}-{D}-{2S}-{ky}-{exci}"
std::string tag(int ns, int d, int s, int ky, int exci);
long long block_count(const Leg* l);
// sector indices and 2j labels of block index idx (leg 0 fastest)
void labels_of(const Leg* l, long long idx, int* s, int* tj);
void write_file(const std::string& path, const File3& f);
// reads path.dat / path.bin; throws kErrArgument / kErrStructure
File3 read_file(const std::string& path);
// r with stored = r * ref; throws unless the stored CG is such a multiple
double cg_ratio(const double* stored, const std::vector<double>& ref, const std::string& what);
// last value of enr-{tag}.dat plus 1000 (read_enr, dmrg_su2_omp.cpp:1140-1146)
double read_energy(const std::string& dir, const std::string& tag);

}  // namespace mpsio
}  // namespace su2
