// Shared device and host helpers for the MI355X variant library.
//
// Everything here is gfx950-specific: 64-wide wavefronts, 16-byte vector
// memory operations and FP64 arithmetic without implicit contraction (the
// library is compiled with -ffp-contract=off; fused multiply-adds are
// written out with fma() where they are wanted).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rajaperf_mi355x.h"

namespace rpx {

constexpr int kWave = 64;

// ---- host-side status plumbing -------------------------------------------

inline int hip_status(hipError_t e) { return e == hipSuccess ? RPX_SUCCESS : -static_cast<int>(e); }

#define RPX_TRY(expr)                                   \
  do {                                                  \
    hipError_t rpx_e_ = (expr);                         \
    if (rpx_e_ != hipSuccess) return -static_cast<int>(rpx_e_); \
  } while (0)

// Launch status: returns the launch error of the most recent launch.
inline int launch_status() { return hip_status(hipGetLastError()); }

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <typename T>
inline T tune_or(const rpx_tuning* t, T rpx_tuning::*field, T dflt) {
  return (t && (t->*field) > 0) ? (t->*field) : dflt;
}
inline int tune_flags(const rpx_tuning* t) { return t ? t->flags : 0; }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---- device helpers -------------------------------------------------------

// Wavefront sum of an FP64 value; lane 0 holds the result.  Butterfly over
// the full 64 lanes (DPP/permute lowering is left to the compiler).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Inclusive wavefront scan (Hillis-Steele over 64 lanes).
__device__ __forceinline__ double wave_inclusive_scan(double v, int lane) {
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    double o = __shfl_up(v, off, kWave);
    if (lane >= off) v += o;
  }
  return v;
}

// SplitMix64 output function (Steele, Lea & Flood, OOPSLA 2014).
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kStreamMul = 0xD1B54A32D192ED03ull;

__host__ __device__ __forceinline__ uint64_t seeded_state0(uint64_t seed, uint64_t stream_id) {
  return mix64(seed ^ (stream_id * kStreamMul));
}
// u in [0,1) with 53 random bits.
__host__ __device__ __forceinline__ double seeded_unit(uint64_t state0, uint64_t i) {
  return static_cast<double>(mix64(state0 + (i + 1) * kGamma) >> 11) * 0x1.0p-53;
}

}  // namespace rpx
