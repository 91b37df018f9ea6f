#include "structure.hpp"

#include <cstdlib>

#include "common.hpp"

namespace su2 {

template <class K>
static void shell_sort_impl(std::vector<K>& key, std::vector<int>& pay) {
  const int n = (int)key.size();
  for (int gap = n / 2; gap > 0; gap /= 2) {
    for (int i = gap; i < n; ++i) {
      int j = i - gap;
      while (true) {
        if (key[j] > key[j + gap]) {
          std::swap(key[j], key[j + gap]);
          std::swap(pay[j], pay[j + gap]);
        } else {
          break;
        }
        if (j >= gap) j -= gap;
        else break;
      }
    }
  }
}

void shell_sort_int(std::vector<int>& key, std::vector<int>& payload) {
  shell_sort_impl(key, payload);
}

void shell_sort_double(std::vector<double>& key, std::vector<int>& payload) {
  shell_sort_impl(key, payload);
}

Leg make_leg(std::vector<int> twoj, std::vector<int> dim, int dir) {
  SU2_CHECK(twoj.size() == dim.size(), kErrArgument, "leg: %zu labels, %zu dims",
            twoj.size(), dim.size());
  shell_sort_int(twoj, dim);
  Leg l;
  l.twoj = std::move(twoj);
  l.dim = std::move(dim);
  l.dir = dir;
  for (size_t i = 1; i < l.twoj.size(); ++i)
    SU2_CHECK(l.twoj[i] != l.twoj[i - 1], kErrArgument, "leg: repeated 2j=%d", l.twoj[i]);
  for (int t : l.twoj) SU2_CHECK(t >= 0, kErrArgument, "leg: negative 2j=%d", t);
  for (int d : l.dim) SU2_CHECK(d >= 0, kErrArgument, "leg: negative dim %d", d);
  return l;
}

Leg fuse_legs(const Leg& a, const Leg& b, int dir) {
  if (dir == 0) {
    SU2_CHECK(a.dir == b.dir, kErrStructure, "fuse: legs of different directions");
    dir = -a.dir;
  }
  std::vector<int> ang, dim;
  for (int i = 0; i < a.nsec(); ++i) {
    for (int j = 0; j < b.nsec(); ++j) {
      int lo = std::abs(a.twoj[i] - b.twoj[j]);
      int hi = a.twoj[i] + b.twoj[j];
      if (hi > kFuseCap) hi = kFuseCap;
      for (int k = lo; k <= hi; k += 2) {
        bool found = false;
        for (size_t l = 0; l < ang.size(); ++l) {
          if (ang[l] == k) {
            dim[l] += a.dim[i] * b.dim[j];
            found = true;
            break;
          }
        }
        if (!found) {
          ang.push_back(k);
          dim.push_back(a.dim[i] * b.dim[j]);
        }
      }
    }
  }
  return make_leg(ang, dim, dir);
}

bool labels_allowed(const int* twoj, int nleg, int locspin) {
  if (nleg == 1) return twoj[0] == locspin;
  // reachable set of 2J after fusing legs 0..k sequentially (capped at 38)
  std::vector<char> reach(kFuseCap + 1, 0), next(kFuseCap + 1, 0);
  if (twoj[0] > kFuseCap) return false;
  reach[twoj[0]] = 1;
  for (int l = 1; l < nleg; ++l) {
    std::fill(next.begin(), next.end(), 0);
    for (int s = 0; s <= kFuseCap; ++s) {
      if (!reach[s]) continue;
      int lo = std::abs(s - twoj[l]);
      int hi = s + twoj[l];
      if (hi > kFuseCap) hi = kFuseCap;
      for (int k = lo; k <= hi; k += 2) next[k] = 1;
    }
    reach.swap(next);
  }
  return locspin <= kFuseCap && reach[locspin];
}

long fused_offset(const Leg& a, const Leg& b, int ia, int ib, int twoj_f) {
  long pos = 0;
  for (int j = 0; j < b.nsec(); ++j) {
    for (int i = 0; i < a.nsec(); ++i) {
      int lo = std::abs(a.twoj[i] - b.twoj[j]);
      int hi = a.twoj[i] + b.twoj[j];
      if (twoj_f < lo || twoj_f > hi || ((twoj_f - lo) & 1)) continue;
      if (i == ia && j == ib) return pos;
      pos += (long)a.dim[i] * b.dim[j];
    }
  }
  return -1;
}

}  // namespace su2
