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

// Host tables sent to device memory on a stream without a host
// synchronisation (runtime.hip): the bytes are copied into library-owned
// pinned staging memory and sent with hipMemcpyAsync, and a staging slot is
// reused only once the copies that read it have completed (one event per
// slot).  Under stream capture the copies would read the slot at every
// replay, so there it returns RPX_ERR_UNSUPPORTED.
struct HostChunk {
  void* dst;
  const void* src;
  size_t bytes;
};
int upload_async(const HostChunk* chunks, int n, hipStream_t s);

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <typename T>
inline T tune_or(const rpx_tuning* t, T rpx_tuning::*field, T dflt) {
  return (t && (t->*field) > 0) ? (t->*field) : dflt;
}
inline int tune_flags(const rpx_tuning* t) { return t ? t->flags : 0; }

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

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

// XCD-major numbering: XCD x (workgroups b with b % 8 == x) gets the
// contiguous range of logical ids [x*H, (x+1)*H), H = floor(G/8); the last
// G % 8 workgroups keep their own ids.  A bijection of [0, G) for every G.
// For streams it also keeps each XCD's address translations to its own
// eighth of the arrays: measured on 2 GiB arrays, reads 7.08 -> 7.31 TB/s,
// writes 6.69 -> 6.85 TB/s, TRIAD 6.44 -> 6.52 TB/s.
__host__ __device__ __forceinline__ unsigned xcd_major(unsigned b, unsigned G) {
  const unsigned h = G / 8;
  return b < 8 * h ? (b % 8) * h + b / 8 : b;
}

// Output tile (tm, tn) of hardware workgroup b for a GEMM-shaped grid.  With
// one workgroup per CU, the grid runs in rounds of 256 workgroups, 32 per XCD
// (workgroup b runs on XCD b % 8).  When both tile counts are multiples of 16, each round covers
// a 16 x 16 patch of tiles and each XCD a 4 x 8 part of it, so an XCD's L2
// holds the 4 A tiles and 8 B tiles its 32 workgroups read at every K step
// (not 1 A and 32 B tiles, as a row-major patch would).  Otherwise tiles
// are numbered XCD-major.
__device__ __forceinline__ void tile_of(unsigned b, unsigned G, int tiles_m, int tiles_n, int64_t& tm, int64_t& tn) {
  if (tiles_m % 16 == 0 && tiles_n % 16 == 0) {
    const unsigned x = b % 8, slot = b / 8;  // XCD, order within the XCD
    const unsigned round = slot / 32, pos = slot % 32;
    const unsigned pn = unsigned(tiles_n) / 16;  // patches per patch row
    const unsigned patch_m = round / pn, patch_n = round % pn;
    tm = int64_t(patch_m) * 16 + (x / 2) * 4 + pos / 8;
    tn = int64_t(patch_n) * 16 + (x % 2) * 8 + pos % 8;
    return;
  }
  const unsigned t = xcd_major(b, G);
  tm = t / unsigned(tiles_n);
  tn = t % unsigned(tiles_n);
}

}  // namespace rpx
