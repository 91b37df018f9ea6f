// Common status handling for the SU(2) DMRG HIP library.
//
// The reference prints to cout and calls exit(0) on every structural error
// (e.g. tensor_su2.cpp:165-168); this library instead returns an int status
// through the C ABI and records a message that su2_last_error() returns.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace su2 {

enum Status : int {
  kOk = 0,
  kErrArgument = 1,   // malformed block table, shape mismatch, bad leg
  kErrStructure = 2,  // SU(2) labels that cannot be contracted/fused
  kErrDevice = 3,     // HIP runtime error
  kErrCapacity = 4,   // exceeds a documented maximum
  kErrNumeric = 5,    // eigensolver did not converge etc.
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

#define SU2_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      ::su2::fail(::su2::kErrDevice, "%s failed: %s (%s:%d)", #call,         \
                  hipGetErrorString(e_), __FILE__, __LINE__);                \
  } while (0)

#define SU2_CHECK(cond, code, ...)                                           \
  do {                                                                       \
    if (!(cond)) ::su2::fail(code, __VA_ARGS__);                             \
  } while (0)

}  // namespace su2
