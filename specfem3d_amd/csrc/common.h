// common.h -- shared definitions for the MI355X solver-loop library (gfx950, FP32).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

typedef float realw;

#define NGLLX 5
#define NGLL2 25
#define NGLL3 125
#define N_SLS 3

// Fortran-callable symbols: lower case + one trailing underscore (the reference's configure
// detected "lower case, underscore, no extra underscore" for FC_FUNC_).
#define FC_FUNC_(name, NAME) name##_

namespace sf {

// Fatal error: mirrors the reference's exit_on_error (print, append to an error file, exit(1)),
// but first records the message for sf_last_error() so a host harness can read it.
[[noreturn]] void fatal(const char* where, const char* msg);
void set_error(const char* msg);

inline void check(hipError_t e, const char* where) {
  if (e != hipSuccess) fatal(where, hipGetErrorString(e));
}

}  // namespace sf

#define SF_CHECK(x) ::sf::check((x), #x " @ " __FILE__)
#define SF_LAUNCH_CHECK(where) ::sf::check(hipGetLastError(), where)

// grid for n work items with `block` threads per block (1-D; HIP allows 2^31-1 blocks in x)
inline dim3 sf_grid(int64_t n, int block) {
  int64_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  return dim3((unsigned)b, 1, 1);
}
