// Common definitions for the gfx950 AFQMC walker-step library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace afq {

// Error raised by host-side launchers; the C-ABI layer turns it into a status code plus the text
// returned by afq_last_error().
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
  kOk = 0,
  kErrArg = -1,
  kErrHip = -2,
  kErrShape = -3,
  kErrAlloc = -4,
  kErrState = -5,
};

#define AFQ_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::afq::Error(::afq::kErrHip, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define AFQ_CHECK(cond, code, msg)                                             \
  do {                                                                         \
    if (!(cond)) throw ::afq::Error((code), std::string(__func__) + ": " + (msg)); \
  } while (0)

#define AFQ_LAUNCH_CHECK() AFQ_HIP(hipGetLastError())

// Interleaved complex scalars as stored in HBM (std::complex layout).
template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T re, im;
};
using z64 = cplx<double>;  // complex128
using c32 = cplx<float>;   // complex64

template <typename T>
__host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) {
  return {a.re + b.re, a.im + b.im};
}

// Wave64 reductions (CDNA wavefronts are 64 lanes wide).
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace afq
