// Stream group: COPY, MUL, ADD, TRIAD.
//
// Reference bodies: src/stream/COPY.hpp:24-25, MUL.hpp:25-26, ADD.hpp:25-26,
// TRIAD.hpp:26-27; the suite's HIP launch (one element per thread, scalar
// 8-byte accesses) is TRIAD-Hip.cpp:24-61.
//
// MI355X design: each work-item moves U 16-byte vectors per operand
// (global_load_dwordx4 / global_store_dwordx4); the U vectors of one
// work-item are a workgroup-width apart, so every vector instruction of a
// wavefront covers 1 KiB of contiguous memory.  All loads of a work-item
// are issued before any arithmetic.  Loads and stores are nontemporal by
// default: no operand is touched twice, and streaming them past L2/MALL
// measured 6.4 TB/s for TRIAD at 2^28 against 5.6 TB/s with regular
// accesses (block 256, one vector per work-item; larger U measured slower).
// RPX_TUNE_CACHED selects regular accesses.  Arithmetic follows
// the Base_Seq association with no contraction (TRIAD is b + (alpha*c)
// rounded twice), so results are bit-identical to Base_Seq.
#include "rpx_common.hpp"

using namespace rpx;

namespace {

typedef double v2d __attribute__((ext_vector_type(2)));

enum Op { kCopy = 0, kMul = 1, kAdd = 2, kTriad = 3 };

template <int OP>
__device__ __forceinline__ double apply(double x, double y, double alpha) {
  if constexpr (OP == kCopy) return x;
  else if constexpr (OP == kMul) return alpha * x;
  else if constexpr (OP == kAdd) return x + y;
  else {
    const double t = alpha * y;  // TRIAD: out = x + alpha*y, x=b, y=c
    return x + t;
  }
}

template <int OP>
__device__ __forceinline__ v2d apply2(v2d x, v2d y, double alpha) {
  v2d r;
  r.x = apply<OP>(x.x, y.x, alpha);
  r.y = apply<OP>(x.y, y.y, alpha);
  return r;
}

// Operand roles: out = f(x, y).  COPY/MUL read x only.
template <int OP, int U, int BS, bool NT>
__global__ __launch_bounds__(BS) void stream_vec_kernel(double* __restrict__ out,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ y, double alpha,
                                                        int64_t nvec, int64_t n) {
  constexpr bool kTwoInputs = (OP == kAdd || OP == kTriad);
  const v2d* __restrict__ xv = reinterpret_cast<const v2d*>(x);
  const v2d* __restrict__ yv = reinterpret_cast<const v2d*>(y);
  v2d* __restrict__ ov = reinterpret_cast<v2d*>(out);

  const int64_t base = int64_t(xcd_major(blockIdx.x, gridDim.x)) * (BS * U) + threadIdx.x;
  if (base + (U - 1) * BS < nvec) {
    // Full tile: no bounds checks.
    v2d a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = NT ? __builtin_nontemporal_load(&xv[base + u * BS]) : xv[base + u * BS];
    if constexpr (kTwoInputs) {
#pragma unroll
      for (int u = 0; u < U; ++u) b[u] = NT ? __builtin_nontemporal_load(&yv[base + u * BS]) : yv[base + u * BS];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const v2d r = apply2<OP>(a[u], kTwoInputs ? b[u] : a[u], alpha);
      if constexpr (NT) __builtin_nontemporal_store(r, &ov[base + u * BS]);
      else ov[base + u * BS] = r;
    }
  } else {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t v = base + u * BS;
      if (v < nvec) {
        const v2d a = xv[v];
        const v2d b = kTwoInputs ? yv[v] : a;
        ov[v] = apply2<OP>(a, b, alpha);
      }
    }
  }
  // Odd length: the last element is handled by the first work-item.
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const int64_t i = n - 1;
    const double a = x[i];
    const double b = kTwoInputs ? y[i] : a;
    out[i] = apply<OP>(a, b, alpha);
  }
}

// Fallback for operands that are not 16-byte aligned: one element per
// work-item with a grid-stride loop.
template <int OP>
__global__ __launch_bounds__(256) void stream_scalar_kernel(double* __restrict__ out,
                                                            const double* __restrict__ x,
                                                            const double* __restrict__ y,
                                                            double alpha, int64_t n) {
  constexpr bool kTwoInputs = (OP == kAdd || OP == kTriad);
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double a = x[i];
    const double b = kTwoInputs ? y[i] : a;
    out[i] = apply<OP>(a, b, alpha);
  }
}

template <int OP, int U, int BS>
int launch_vec(double* out, const double* x, const double* y, double alpha, int64_t n, bool nt,
               hipStream_t s) {
  const int64_t nvec = n / 2;
  int64_t grid = ceil_div(nvec, int64_t(BS) * U);
  if (grid < 1) grid = 1;
  if (grid > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  if (nt)
    stream_vec_kernel<OP, U, BS, true><<<dim3(unsigned(grid)), dim3(BS), 0, s>>>(out, x, y, alpha, nvec, n);
  else
    stream_vec_kernel<OP, U, BS, false><<<dim3(unsigned(grid)), dim3(BS), 0, s>>>(out, x, y, alpha, nvec, n);
  return launch_status();
}

template <int OP, int BS>
int dispatch_u(int items, double* out, const double* x, const double* y, double alpha, int64_t n,
               bool nt, hipStream_t s) {
  switch (items) {
    case 2: return launch_vec<OP, 1, BS>(out, x, y, alpha, n, nt, s);
    case 4: return launch_vec<OP, 2, BS>(out, x, y, alpha, n, nt, s);
    case 8: return launch_vec<OP, 4, BS>(out, x, y, alpha, n, nt, s);
    case 16: return launch_vec<OP, 8, BS>(out, x, y, alpha, n, nt, s);
    default: return RPX_ERR_UNSUPPORTED;
  }
}

// Defaults chosen by a sweep on MI355X at 2^28 doubles (block 256, one
// 16-byte vector per work-item, nontemporal loads and stores).
constexpr int kDefaultBlock = 256;
constexpr int kDefaultItems = 2;

template <int OP>
int run_stream(double* out, const double* x, const double* y, double alpha, int64_t n,
               const rpx_tuning* tune, void* stream) {
  constexpr bool kTwoInputs = (OP == kAdd || OP == kTriad);
  if (n < 0) return RPX_ERR_INVALID_ARGUMENT;
  if (n == 0) return RPX_SUCCESS;
  if (!out || !x || (kTwoInputs && !y)) return RPX_ERR_INVALID_ARGUMENT;
  hipStream_t s = as_stream(stream);
  const int bs = tune_or(tune, &rpx_tuning::block_size, kDefaultBlock);
  const int items = tune_or(tune, &rpx_tuning::items_per_thread, kDefaultItems);
  const bool nt = (tune_flags(tune) & RPX_TUNE_CACHED) == 0;
  const bool vec_ok = aligned16(out) && aligned16(x) && (!kTwoInputs || aligned16(y));
  if (!vec_ok) {
    int64_t grid = ceil_div(n, 256);
    if (grid > 65536) grid = 65536;
    stream_scalar_kernel<OP><<<dim3(unsigned(grid)), dim3(256), 0, s>>>(out, x, y, alpha, n);
    return launch_status();
  }
  switch (bs) {
    case 64: return dispatch_u<OP, 64>(items, out, x, y, alpha, n, nt, s);
    case 128: return dispatch_u<OP, 128>(items, out, x, y, alpha, n, nt, s);
    case 256: return dispatch_u<OP, 256>(items, out, x, y, alpha, n, nt, s);
    case 512: return dispatch_u<OP, 512>(items, out, x, y, alpha, n, nt, s);
    case 1024: return dispatch_u<OP, 1024>(items, out, x, y, alpha, n, nt, s);
    default: return RPX_ERR_UNSUPPORTED;
  }
}

}  // namespace

extern "C" {

int rpx_copy_f64(double* c, const double* a, int64_t n, const rpx_tuning* tune, void* stream) {
  return run_stream<kCopy>(c, a, nullptr, 0.0, n, tune, stream);
}
int rpx_mul_f64(double* b, const double* c, double alpha, int64_t n, const rpx_tuning* tune, void* stream) {
  return run_stream<kMul>(b, c, nullptr, alpha, n, tune, stream);
}
int rpx_add_f64(double* c, const double* a, const double* b, int64_t n, const rpx_tuning* tune, void* stream) {
  return run_stream<kAdd>(c, a, b, 0.0, n, tune, stream);
}
int rpx_triad_f64(double* a, const double* b, const double* c, double alpha, int64_t n, const rpx_tuning* tune,
                  void* stream) {
  return run_stream<kTriad>(a, b, c, alpha, n, tune, stream);
}

}  // extern "C"
