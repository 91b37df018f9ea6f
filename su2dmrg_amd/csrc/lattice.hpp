// Square-lattice cylinder tables of dmrg_su2::dmrg_su2 (dmrg_su2.cpp:31-303):
// snake site order (open in x, periodic in y), bond classes hmap[i][j] in
// {0,1,2,3}, couplings coup[1..3], the plaquettes of the Q term and the
// operator retention flags fll/frr.
#pragma once

#include <array>
#include <cstddef>
#include <vector>

namespace su2 {

struct Lattice {
  int lx = 0, ly = 0, ns = 0, bdry = 2;
  double jcoup = 0, qcoup = 0;
  std::vector<int> hmap;     // ns*ns, row-major [i*ns+j]
  double coup[4] = {0, 0, 0, 0};
  std::vector<char> fll, frr;  // ns*ns
  // plaquettes (x,y), (x,y+1), (x+1,y+1), (x+1,y) for x < lx-1 as sorted
  // snake positions p0 < p1 < p2 < p3 (plqpos, dmrg_su2.cpp:200-232); empty
  // unless |qcoup| > 1e-5 (plqflg, 263-281)
  std::vector<std::array<int, 4>> plq;

  Lattice(int lx, int ly, double jcoup, double qcoup, int bdry);
  int h(int i, int j) const { return hmap[(size_t)i * ns + j]; }
  double J(int i, int j) const { return h(i, j) ? coup[h(i, j)] : 0.0; }
  bool keep_left(int i, int k) const { return fll[(size_t)i * ns + k] != 0; }
  bool keep_right(int i, int k) const { return frr[(size_t)i * ns + k] != 0; }
};

}  // namespace su2
