// Library context: one device, one stream, the last error message, and a
// small pool of device scratch for reductions.
//
// The reference reads its configuration from globals (psize, max_dcut, the
// fac_* tables: tensor_su2.cpp:14-16, tensor.cpp:25); here everything a call
// needs hangs off this object, so calls on different contexts are independent.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace su2 {

class Context {
 public:
  explicit Context(int device);
  ~Context();
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  void set_stream(hipStream_t s) { stream_ = s; }
  void make_current() const;

  std::string last_error;

  // device scratch for block-reduction partials (grown on demand, never shrunk)
  double* reduce_scratch(size_t n);
  // pinned host scratch for reading back reduction results
  double* host_scratch(size_t n);

  int max_dcut = 8000;  // reference default (main_omp.cpp:15)
  int svd_block_cols = 32;   // one-sided Jacobi column block width (16 or 32)
  int svd_inner_sweeps = 1;  // its inner sweeps per visit of a pair (1..3)

 private:
  int device_;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  double* d_red_ = nullptr;
  size_t d_red_n_ = 0;
  double* h_red_ = nullptr;
  size_t h_red_n_ = 0;
};

}  // namespace su2
