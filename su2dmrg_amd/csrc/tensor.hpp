// Device-resident SU(2) block-sparse tensor.
//
// The reference keeps one heap array per block (tarr/parr) plus an explicit
// dense CG tensor per block (tcgc/pcgc, tensor_su2.hpp:6-124).  Here all
// reduced blocks of a tensor live in one HBM allocation, laid out in block-
// index order (leg 0 fastest over sector indices, su2struct.cpp:243-261),
// each block column-major with leg 0 fastest (tensor.cpp:1101-1119).  The CG
// part is implicit: 2-leg tensors carry the unit singlet (wavefunctions,
// tensor_su2.cpp:2334-2351) or the identity (scalar operators), 3-leg tensors
// carry make_cgc(j0, j1, j2) (tensor_su2_dmrg_src.cpp:163-165).  This is the
// same serialisation the reference writes with get_telement(tele1,..,iflag)
// (tensor_su2.cpp:532-549): present-block flags, then reduced doubles.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <vector>

#include "structure.hpp"

namespace su2 {

struct Block {
  int twoj[kMaxLegs] = {0, 0, 0, 0};
  int sec[kMaxLegs] = {0, 0, 0, 0};   // sector index on each leg
  int dim[kMaxLegs] = {1, 1, 1, 1};
  long long offset = 0;               // doubles into the tensor's data
  long long size = 0;
  int index = 0;                      // block index (mixed radix)
};

class Context;
class DevicePool;

class Tensor {
 public:
  // zero = false leaves the blocks unset: the caller zeroes them (ZeroBatch)
  // before any kernel reads or accumulates into them
  Tensor(Context* ctx, std::vector<Leg> legs, int locspin, const std::vector<char>* present,
         bool zero = true);
  ~Tensor();
  Tensor(const Tensor&) = delete;
  Tensor& operator=(const Tensor&) = delete;

  int nleg() const { return (int)legs_.size(); }
  const Leg& leg(int i) const { return legs_[i]; }
  const std::vector<Leg>& legs() const { return legs_; }
  int locspin() const { return locspin_; }
  long long size() const { return size_; }
  double* data() { return d_; }
  const double* data() const { return d_; }
  const std::vector<Block>& blocks() const { return blocks_; }
  int nten() const { return nten_; }
  // block id for labels, or -1
  int find(const int* twoj) const;
  int find2(int a, int b) const { int t[2] = {a, b}; return find(t); }
  int find3(int a, int b, int c) const { int t[3] = {a, b, c}; return find(t); }
  bool same_layout(const Tensor& o) const;
  Context* ctx() const { return ctx_; }

 private:
  Context* ctx_;
  std::vector<Leg> legs_;
  int locspin_;
  int nten_ = 1;
  std::vector<Block> blocks_;
  std::map<long long, int> by_key_;
  double* d_ = nullptr;
  std::shared_ptr<DevicePool> pool_;   // outlives the context if need be
  long long size_ = 0;
};

// Zero fills of many tensors as one upload and one launch (in place of a
// memset per tensor), queued on the context's stream by run().
class ZeroBatch {
 public:
  void add(Tensor& t);
  void run(Context& c);

 private:
  std::vector<std::pair<double*, long long>> z_;
};

}  // namespace su2
