// SU(2) leg and block structure (host, integer, bit-exact with the reference).
//
// A Leg is the reference's su2bond (su2bond.hpp:3-29): sectors with doubled
// spin 2j ascending, a multiplet count per sector, a direction +-1.
// Block validity and the block index follow su2struct (su2struct.cpp:243-332):
// the block index is mixed radix over the legs with leg 0 fastest, and a
// label tuple is allowed when sequential fusion of its legs reaches locspin.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace su2 {

constexpr int kFuseCap = 38;   // fused 2j is capped at 38 (su2bond.cpp:175-177)
constexpr int kMaxLegs = 4;

struct Leg {
  std::vector<int> twoj;   // ascending
  std::vector<int> dim;    // multiplets per sector
  int dir = 1;

  int nsec() const { return (int)twoj.size(); }
  int find(int tj) const {
    for (int i = 0; i < nsec(); ++i)
      if (twoj[i] == tj) return i;
    return -1;
  }
  int dim_of(int tj) const {
    int i = find(tj);
    return i < 0 ? 0 : dim[i];
  }
  long total() const {
    long s = 0;
    for (int d : dim) s += d;
    return s;
  }
  bool same_sectors(const Leg& o) const { return twoj == o.twoj && dim == o.dim; }
  bool operator==(const Leg& o) const { return dir == o.dir && same_sectors(o); }
};

// Shell sort with gap n/2, n/4, ... and a strict '>' exchange (sort2.f:38-73),
// applied to integer keys carrying a payload.
void shell_sort_int(std::vector<int>& key, std::vector<int>& payload);
// Same gap sequence over double keys (sort2.f:76-111): decides truncation ties.
void shell_sort_double(std::vector<double>& key, std::vector<int>& payload);

// Leg from unsorted (2j, dim) lists: sorted like su2bond::set_su2bond.
Leg make_leg(std::vector<int> twoj, std::vector<int> dim, int dir);

// su2bond::fuse(b1,b2) (su2bond.cpp:155-212) when dir==0 (legs must share a
// direction and the result points the other way), su2bond::fuse(b1,b2,dir)
// (su2bond.cpp:215-260) otherwise.
Leg fuse_legs(const Leg& a, const Leg& b, int dir = 0);

// Triangle-chain test of su2struct::check_angularmoments (su2struct.cpp:264-289).
bool labels_allowed(const int* twoj, int nleg, int locspin);

// Position of sector (2j of leg a, 2j of leg b) inside the fused sector 2j_f
// when leg a and leg b are fused by tensor_su2::fuse(b1,b2) (tensor_su2.cpp:
// 839-909): leg-b sectors outer, leg-a sectors inner, offset += d_a*d_b.
// Returns -1 if the pair does not fuse to 2j_f.
long fused_offset(const Leg& a, const Leg& b, int ia, int ib, int twoj_f);

}  // namespace su2
