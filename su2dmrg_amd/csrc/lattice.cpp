#include "lattice.hpp"

#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace su2 {

Lattice::Lattice(int lx_, int ly_, double jc, double qc, int bd)
    : lx(lx_), ly(ly_), ns(lx_ * ly_), bdry(bd), jcoup(jc), qcoup(qc) {
  SU2_CHECK(lx >= 1 && ly >= 1 && ns >= 4, kErrArgument, "lattice %dx%d too small", lx, ly);
  SU2_CHECK(bdry >= 0 && bdry <= 3, kErrArgument, "boundary code %d not supported", bdry);
  SU2_CHECK(std::fabs(qcoup) <= 1e-5 || ly >= 2, kErrArgument,
            "plaquette (Q) terms need ly >= 2 (qcoup=%g)", qcoup);
  hmap.assign((size_t)ns * ns, 0);
  auto site = [&](int x, int y) { return x * ly + ((x % 2 == 0) ? y : ly - 1 - y); };
  for (int k = 0; k <= 4; ++k) {
    for (int i = 0; i < ns; ++i) {
      const int x1 = i / ly;
      const int y1 = (x1 % 2 == 0) ? i % ly : ly - 1 - (i % ly);
      int x2 = 0, y2 = 0;
      bool flag = true;
      if (bdry == 0) {
        if (k >= 2) continue;
        if (k == 0) { x2 = (x1 + 1) % lx; y2 = y1; }
        else { x2 = x1; y2 = (y1 + 1) % ly; }
      } else if (bdry == 2) {
        if (k == 0) { x2 = x1 + 1; y2 = y1; }
        else if (k == 1) { x2 = x1; y2 = (y1 + 1) % ly; }
        else if (k == 2) { x2 = x1 + 1; y2 = (y1 + 1) % ly; }
        else if (k == 3) { x2 = x1 - 1; y2 = (y1 + 1) % ly; }
        else flag = false;
      } else if (bdry == 1) {  // J1-J3
        if (k == 0) { x2 = x1 + 1; y2 = y1; }
        else if (k == 1) { x2 = x1; y2 = (y1 + 1) % ly; }
        else if (k == 2) { x2 = x1 + 2; y2 = y1; }
        else if (k == 3) { x2 = x1; y2 = (y1 + 2) % ly; }
        else flag = false;
      } else {  // bdry == 3, Shastry-Sutherland (dmrg_su2.cpp:154-170)
        // square bonds, plus one diagonal on every other plaquette with the
        // diagonals of neighbouring plaquettes orthogonal; the reference's
        // k = 4 pair gets no coupling class there, so it is left out
        if (k == 0) { x2 = x1 + 1; y2 = y1; }
        else if (k == 1) { x2 = x1; y2 = (y1 + 1) % ly; }
        else if (k == 2 && x1 % 2 == 0 && y1 % 2 == 0) { x2 = x1 + 1; y2 = (y1 + 1) % ly; }
        else if (k == 3 && x1 % 2 == 1 && y1 % 2 == 0) { x2 = x1 + 1; y2 = (y1 + ly - 1) % ly; }
        else flag = false;
      }
      if (!flag || x2 < 0 || x2 >= lx || y2 < 0 || y2 >= ly) continue;
      const int j = site(x2, y2);
      // coupling class (dmrg_su2.cpp:180-197): on the Shastry-Sutherland
      // lattice the square bonds take coup[2] and the dimers coup[1]
      int v = 0;
      if (bdry == 2 && k == 1 && (x1 == 0 || x1 == lx - 1)) v = 3;
      else if ((bdry == 3 && k < 2) || ((bdry == 2 || bdry == 1) && k >= 2 && k < 4)) v = 2;
      else if ((bdry == 3 && k >= 2 && k < 4) || (bdry != 3 && k < 2)) v = 1;
      if (!v) continue;
      hmap[(size_t)i * ns + j] = v;
      hmap[(size_t)j * ns + i] = v;
    }
  }
  coup[1] = jcoup + qcoup / 2.0;
  coup[2] = 1.0 - std::fabs(qcoup);
  coup[3] = jcoup + qcoup / 4.0;
  if (std::fabs(qcoup) > 1e-5) {
    for (int m = 0; m < ns - ly; ++m) {
      const int x1 = m / ly, y1 = m % ly;
      std::array<int, 4> p = {site(x1, y1), site(x1, (y1 + 1) % ly),
                              site(x1 + 1, (y1 + 1) % ly), site(x1 + 1, y1)};
      std::sort(p.begin(), p.end());
      plq.push_back(p);
    }
  }
  fll.assign((size_t)ns * ns, 0);
  frr.assign((size_t)ns * ns, 0);
  for (int i = 0; i < ns; ++i) {
    for (int j = 0; j < i; ++j)
      for (int k = i + 1; k < ns; ++k)
        if (h(j, k)) fll[(size_t)i * ns + j] = 1;
    for (int j = i + 1; j < ns; ++j)
      for (int k = 0; k < i; ++k)
        if (h(j, k)) frr[(size_t)i * ns + j] = 1;
    // S_p0 stays until p1 and S_p3 until p2 of a plaquette (dmrg_su2.cpp:293-302)
    for (auto& p : plq) {
      if (p[0] < i && i < p[1]) fll[(size_t)i * ns + p[0]] = 1;
      if (p[2] < i && i < p[3]) frr[(size_t)i * ns + p[3]] = 1;
    }
  }
}

}  // namespace su2
