// Common helpers for the MI355X (gfx950) kernels of mrirecon_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>

// ------------------------------------------------------------------ status codes
// Returned by every C-ABI entry point; mrh_last_error() gives the message.
enum {
  MRH_OK = 0,
  MRH_EINVAL = 1,      // bad argument / shape            -> ValueError
  MRH_EHIP = 2,        // HIP runtime failure              -> RuntimeError
  MRH_ENOMEM = 3,      // device allocation failed         -> MemoryError
  MRH_EUNSUP = 4,      // unsupported size / configuration -> NotImplementedError
};

void mrh_set_error(const char* fmt, ...);

#define MRH_HIP(call)                                                          \
  do {                                                                         \
    hipError_t _e = (call);                                                    \
    if (_e != hipSuccess) {                                                    \
      mrh_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e),     \
                    __FILE__, __LINE__);                                       \
      return (_e == hipErrorOutOfMemory) ? MRH_ENOMEM : MRH_EHIP;              \
    }                                                                          \
  } while (0)

#define MRH_CHECK_LAUNCH() MRH_HIP(hipGetLastError())

#define MRH_REQUIRE(cond, code, ...)                                           \
  do {                                                                         \
    if (!(cond)) { mrh_set_error(__VA_ARGS__); return (code); }                \
  } while (0)

// ------------------------------------------------------------------ complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}

// ------------------------------------------------------------------ reductions
__device__ __forceinline__ double warp_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ float warp_sum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
