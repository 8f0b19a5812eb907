// DOT and REDUCE_SUM: single-launch deterministic FP64 reductions.
//
// Reference: src/stream/DOT.hpp:24-25, DOT-Seq.cpp:34-51 (per-rep
// m_dot += dot), DOT-Hip.cpp:28-94 (LDS tree + one atomicAdd per block,
// order non-deterministic); src/algorithm/REDUCE_SUM.hpp:31-32,
// REDUCE_SUM-Seq.cpp:34-51, REDUCE_SUM-Hip.cpp:36-198.
//
// MI355X design: a fixed grid streams the input with 16-byte loads, U
// vectors per work-item per step and 2U independent FP64 accumulators; the
// wavefront is reduced with a 64-lane butterfly and the waves of a
// workgroup through LDS.  Each workgroup writes one partial; the workgroup
// that takes the last ticket sums all partials in index order with a fixed
// tree and resets the ticket.  No floating-point atomics, so a given
// (n, grid) always produces the same bits.
#include "rpx_common.hpp"

using namespace rpx;

namespace {

typedef double v2d __attribute__((ext_vector_type(2)));

constexpr size_t kHeaderBytes = 256;  // ticket counter, padded to a cache line
constexpr int kMaxGrid = 8192;

template <int BS>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  constexpr int kWaves = BS / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  v = wave_sum(v);
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double r = 0.0;
  if (wave == 0) {
    r = (lane < kWaves) ? lds[lane] : 0.0;
    r = wave_sum(r);
  }
  return r;  // valid in thread 0
}

template <bool DOT, int BS, int U>
__global__ __launch_bounds__(BS) void reduce_kernel(const double* __restrict__ a,
                                                    const double* __restrict__ b, int64_t n,
                                                    double* __restrict__ partials,
                                                    unsigned* __restrict__ ticket,
                                                    double* __restrict__ result, int accumulate) {
  __shared__ double lds[BS / kWave];
  __shared__ int is_last;

  const v2d* av = reinterpret_cast<const v2d*>(a);
  const v2d* bv = reinterpret_cast<const v2d*>(b);
  const int64_t nvec = n / 2;
  const int64_t step = int64_t(gridDim.x) * BS * U;

  double acc[2 * U];
#pragma unroll
  for (int k = 0; k < 2 * U; ++k) acc[k] = 0.0;

  int64_t base = int64_t(blockIdx.x) * BS * U + threadIdx.x;
  // Full steps without bounds checks.
  for (; base + (U - 1) * BS < nvec; base += step) {
    v2d x[U], y[U];
#pragma unroll
    for (int u = 0; u < U; ++u) x[u] = av[base + u * BS];
    if constexpr (DOT) {
#pragma unroll
      for (int u = 0; u < U; ++u) y[u] = bv[base + u * BS];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        acc[2 * u] = fma(x[u].x, y[u].x, acc[2 * u]);
        acc[2 * u + 1] = fma(x[u].y, y[u].y, acc[2 * u + 1]);
      }
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        acc[2 * u] += x[u].x;
        acc[2 * u + 1] += x[u].y;
      }
    }
  }
  // Ragged last step.
  if (base < nvec) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t v = base + u * BS;
      if (v < nvec) {
        const v2d x = av[v];
        if constexpr (DOT) {
          const v2d y = bv[v];
          acc[2 * u] = fma(x.x, y.x, acc[2 * u]);
          acc[2 * u + 1] = fma(x.y, y.y, acc[2 * u + 1]);
        } else {
          acc[2 * u] += x.x;
          acc[2 * u + 1] += x.y;
        }
      }
    }
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const double x = a[n - 1];
    acc[0] += DOT ? x * b[n - 1] : x;
  }

  double t = 0.0;
#pragma unroll
  for (int k = 0; k < 2 * U; ++k) t += acc[k];
  const double bsum = block_sum<BS>(t, lds);

  // Partials travel through agent-scope atomic stores/loads, coherent across
  // the XCDs' L2s; the partial store completes (vmcnt) before the ticket is
  // taken, so the workgroup that takes the last ticket sees every partial
  // without an L2 write-back or invalidate.
  unsigned long long* pbits = reinterpret_cast<unsigned long long*>(partials);
  if (threadIdx.x == 0) {
    __hip_atomic_store(&pbits[blockIdx.x], static_cast<unsigned long long>(__double_as_longlong(bsum)),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    const unsigned prev = atomicAdd(ticket, 1u);
    is_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return;

  // Last workgroup: fixed-order sum of all partials.
  double s = 0.0;
  for (int i = threadIdx.x; i < int(gridDim.x); i += BS)
    s += __longlong_as_double(
        static_cast<long long>(__hip_atomic_load(&pbits[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
  __syncthreads();  // lds reuse
  const double total = block_sum<BS>(s, lds);
  if (threadIdx.x == 0) {
    if (accumulate) result[0] = result[0] + total;
    else result[0] = total;
    atomicExch(ticket, 0u);  // ready for the next call
  }
}

__global__ void zero_u32_kernel(unsigned* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

constexpr int kDefaultBlock = 256;
constexpr int kDefaultItems = 8;  // doubles per work-item per step
constexpr int kDefaultGrid = 2048;

int grid_for(int64_t n, int bs, int items, const rpx_tuning* tune) {
  const int64_t nvec = n / 2;
  const int64_t per_block = int64_t(bs) * (items / 2);
  int64_t g = ceil_div(nvec > 0 ? nvec : 1, per_block);
  const int64_t cap = tune_or(tune, &rpx_tuning::grid_size, kDefaultGrid);
  if (g > cap) g = cap;
  if (g > kMaxGrid) g = kMaxGrid;
  if (g < 1) g = 1;
  return int(g);
}

template <bool DOT, int BS>
int launch_u(int items, int grid, const double* a, const double* b, int64_t n, double* partials,
             unsigned* ticket, double* result, int accumulate, hipStream_t s) {
  switch (items) {
    case 2: reduce_kernel<DOT, BS, 1><<<grid, BS, 0, s>>>(a, b, n, partials, ticket, result, accumulate); break;
    case 4: reduce_kernel<DOT, BS, 2><<<grid, BS, 0, s>>>(a, b, n, partials, ticket, result, accumulate); break;
    case 8: reduce_kernel<DOT, BS, 4><<<grid, BS, 0, s>>>(a, b, n, partials, ticket, result, accumulate); break;
    case 16: reduce_kernel<DOT, BS, 8><<<grid, BS, 0, s>>>(a, b, n, partials, ticket, result, accumulate); break;
    default: return RPX_ERR_UNSUPPORTED;
  }
  return launch_status();
}

template <bool DOT>
int run_reduce(const double* a, const double* b, int64_t n, double* result, int accumulate, void* scratch,
               size_t scratch_bytes, const rpx_tuning* tune, void* stream) {
  if (n < 0 || !result || (n > 0 && (!a || (DOT && !b)))) return RPX_ERR_INVALID_ARGUMENT;
  if (!scratch) return RPX_ERR_INVALID_ARGUMENT;
  if (!aligned16(a) || (DOT && !aligned16(b))) return RPX_ERR_INVALID_ARGUMENT;
  const int bs = tune_or(tune, &rpx_tuning::block_size, kDefaultBlock);
  const int items = tune_or(tune, &rpx_tuning::items_per_thread, kDefaultItems);
  const int grid = grid_for(n, bs, items, tune);
  if (scratch_bytes < kHeaderBytes + size_t(grid) * sizeof(double)) return RPX_ERR_SCRATCH_TOO_SMALL;
  unsigned* ticket = reinterpret_cast<unsigned*>(scratch);
  double* partials = reinterpret_cast<double*>(reinterpret_cast<char*>(scratch) + kHeaderBytes);
  hipStream_t s = as_stream(stream);
  switch (bs) {
    case 128: return launch_u<DOT, 128>(items, grid, a, b, n, partials, ticket, result, accumulate, s);
    case 256: return launch_u<DOT, 256>(items, grid, a, b, n, partials, ticket, result, accumulate, s);
    case 512: return launch_u<DOT, 512>(items, grid, a, b, n, partials, ticket, result, accumulate, s);
    case 1024: return launch_u<DOT, 1024>(items, grid, a, b, n, partials, ticket, result, accumulate, s);
    default: return RPX_ERR_UNSUPPORTED;
  }
}

}  // namespace

extern "C" {

size_t rpx_reduce_scratch_bytes(int64_t n, const rpx_tuning* tune) {
  const int bs = tune_or(tune, &rpx_tuning::block_size, kDefaultBlock);
  const int items = tune_or(tune, &rpx_tuning::items_per_thread, kDefaultItems);
  if (n < 0 || bs <= 0 || items < 2) return 0;
  return kHeaderBytes + size_t(grid_for(n, bs, items, tune)) * sizeof(double);
}

int rpx_reduce_scratch_init(void* scratch, size_t scratch_bytes, void* stream) {
  if (!scratch || scratch_bytes < kHeaderBytes) return RPX_ERR_INVALID_ARGUMENT;
  zero_u32_kernel<<<1, 64, 0, as_stream(stream)>>>(reinterpret_cast<unsigned*>(scratch),
                                                    int(kHeaderBytes / sizeof(unsigned)));
  return launch_status();
}

int rpx_dot_f64(const double* a, const double* b, int64_t n, double* result, int accumulate, void* scratch,
                size_t scratch_bytes, const rpx_tuning* tune, void* stream) {
  return run_reduce<true>(a, b, n, result, accumulate, scratch, scratch_bytes, tune, stream);
}

int rpx_reduce_sum_f64(const double* x, int64_t n, double* result, void* scratch, size_t scratch_bytes,
                       const rpx_tuning* tune, void* stream) {
  return run_reduce<false>(x, nullptr, n, result, 0, scratch, scratch_bytes, tune, stream);
}

}  // extern "C"
