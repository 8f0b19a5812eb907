// DOT and REDUCE_SUM: single-launch deterministic FP64 reductions.
//
// Reference: src/stream/DOT.hpp:24-25, DOT-Seq.cpp:34-51 (per-rep
// m_dot += dot), DOT-Hip.cpp:28-94 (LDS tree + one atomicAdd per block,
// order non-deterministic); src/algorithm/REDUCE_SUM.hpp:31-32,
// REDUCE_SUM-Seq.cpp:34-51, REDUCE_SUM-Hip.cpp:36-198.
//
// MI355X design: the input is cut into fixed chunks of K steps of BS x U
// 16-byte vectors, one chunk per workgroup (many short workgroups balance
// across the 8 XCDs better than a persistent grid: 7.1 TB/s against
// 6.6 TB/s measured for DOT at 2^28).  Loads are nontemporal, 2U
// independent FP64 accumulators per work-item, a 64-lane butterfly per
// wavefront and LDS across wavefronts.  Each workgroup publishes one
// partial; partials are combined in fixed order through two ticket levels
// -- the last workgroup of each group of 64 sums its group, the last group
// sums the group partials.  Device-scope atomics on one cache line
// serialise (about 80 ns each on MI355X), so every group ticket has a
// 256-byte line of its own and no line takes more than max(64, G/64)
// atomics; with tickets packed 32 to a line, DOT at 2^28 lost 0.17 ms.
// No floating-point atomics: a given (n, tuning) always produces the same
// bits.  Every ticket is reset and every partial zeroed once consumed, so a
// complete call leaves all the scratch it used zero: the ticket slots of a
// later call with another n or tuning (a larger layout reaches into what
// were partials) start at zero.
#include "rpx_common.hpp"

#include <mutex>

using namespace rpx;

// Library-owned scratch of the allocation-free DOT entry point
// (rpx_dot_f64_noalloc).  Each device holds its own copy of this array from
// module load until the process ends; host code reaches a device's copy only
// through the address hipGetSymbolAddress returns for it.  1 MiB holds the
// reduction layout of the default tuning up to about 1.2e9 doubles.
static __device__ __align__(256) unsigned char scratch[1u << 20];

namespace {

typedef double v2d __attribute__((ext_vector_type(2)));

constexpr size_t kHeaderBytes = 256;  // ticket counter, padded to a cache line

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

constexpr int kGroup = 64;         // workgroups per first-level ticket
constexpr int kTicketStride = 64;  // unsigneds: one 256-byte line per group ticket

struct ReduceLayout {
  int64_t grid, steps, ngroups;
  size_t gtickets_off, gpartials_off, partials_off, total;
};

__device__ __forceinline__ void publish(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                     static_cast<unsigned long long>(__double_as_longlong(v)), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // compiler ordering
  __builtin_amdgcn_s_waitcnt(0x0F70);                      // vmcnt(0): the store has completed
}
__device__ __forceinline__ double fetch(const double* p) {
  return __longlong_as_double(static_cast<long long>(
      __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
}

template <bool DOT, int BS, int U, bool NT>
__global__ __launch_bounds__(BS) void reduce_kernel(const double* __restrict__ a,
                                                    const double* __restrict__ b, int64_t n, int64_t steps,
                                                    int64_t ngroups, double* __restrict__ partials,
                                                    double* __restrict__ gpartials,
                                                    unsigned* __restrict__ gtickets,
                                                    unsigned* __restrict__ ticket,
                                                    double* __restrict__ result, int accumulate) {
  __shared__ double lds[BS / kWave];
  __shared__ int is_last;

  const v2d* av = reinterpret_cast<const v2d*>(a);
  const v2d* bv = reinterpret_cast<const v2d*>(b);
  const int64_t nvec = n / 2;
  const int64_t chunk = steps * (BS * U);
  // Chunks in workgroup order: numbering them XCD-major (as the streams do)
  // measured 1-2% slower here at the default 32 items per work-item, and no
  // faster than this order at its own best chunk size.
  const unsigned tile = blockIdx.x;
  const int64_t cbeg = int64_t(tile) * chunk;
  const int64_t cend = (cbeg + chunk < nvec) ? cbeg + chunk : nvec;

  double acc[2 * U];
#pragma unroll
  for (int k = 0; k < 2 * U; ++k) acc[k] = 0.0;

  int64_t base = cbeg + threadIdx.x;
  for (; base + (U - 1) * BS < cend; base += BS * U) {
    v2d x[U], y[U];
#pragma unroll
    for (int u = 0; u < U; ++u) x[u] = NT ? __builtin_nontemporal_load(&av[base + u * BS]) : av[base + u * BS];
    if constexpr (DOT) {
#pragma unroll
      for (int u = 0; u < U; ++u) y[u] = NT ? __builtin_nontemporal_load(&bv[base + u * BS]) : bv[base + u * BS];
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
  // Ragged last step of the chunk.
  if (base < cend) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t v = base + u * BS;
      if (v < cend) {
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
  // the XCDs' L2s; each partial store completes (vmcnt) before its ticket is
  // taken, so the workgroup that takes a group's last ticket sees all of the
  // group's partials without an L2 write-back or invalidate.
  const int64_t g = tile / kGroup;
  const int64_t gfirst = g * kGroup;
  const int gsize = int((int64_t(gridDim.x) - gfirst < kGroup) ? int64_t(gridDim.x) - gfirst : kGroup);
  if (threadIdx.x == 0) {
    publish(&partials[tile], bsum);
    const unsigned prev = atomicAdd(&gtickets[g * kTicketStride], 1u);
    is_last = (prev == unsigned(gsize - 1));
  }
  __syncthreads();
  if (!is_last) return;

  // Last workgroup of group g: fixed-order sum of the group's partials.
  if (threadIdx.x < kWave) {
    const int lane = threadIdx.x;
    double v = 0.0;
    if (lane < gsize) {
      v = fetch(&partials[gfirst + lane]);
      partials[gfirst + lane] = 0.0;  // consumed: the region is left zero for any later layout
    }
    v = wave_sum(v);
    if (lane == 0) {
      publish(&gpartials[g], v);
      atomicExch(&gtickets[g * kTicketStride], 0u);  // ready for the next call
      const unsigned prev = atomicAdd(ticket, 1u);
      is_last = (prev == unsigned(ngroups - 1));
    }
  }
  __syncthreads();
  if (!is_last) return;

  // Last group: fixed-order sum of the group partials.
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < ngroups; i += BS) {
    s += fetch(&gpartials[i]);
    gpartials[i] = 0.0;
  }
  const double total = block_sum<BS>(s, lds);
  if (threadIdx.x == 0) {
    if (accumulate) result[0] = result[0] + total;
    else result[0] = total;
    atomicExch(ticket, 0u);  // ready for the next call
  }
}

__global__ void zero_u32_kernel(unsigned* p, int64_t n) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

// Defaults from a sweep on MI355X at 2^28 doubles: 256 work-items, 4
// vectors per step, 4 steps (8192 doubles per workgroup for both DOT and
// REDUCE_SUM).
// 512 work-items per workgroup: against 256, DOT 0.628 -> 0.619 ms and
// REDUCE_SUM 0.317 -> 0.311 ms at 2^28 doubles (interleaved, six repeats).
constexpr int kDefaultBlock = 512;
constexpr int kDefaultItems = 32;  // doubles per work-item over its workgroup's chunk

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// items = 2*U*steps doubles per work-item; U = min(4, items/2).
inline int vec_per_step(int items) { return items >= 8 ? 4 : items / 2; }

bool layout_for(int64_t n, int bs, int items, const rpx_tuning* tune, ReduceLayout* l) {
  if (n < 0 || items < 2 || (items & (items - 1)) != 0) return false;
  if (bs != 128 && bs != 256 && bs != 512 && bs != 1024) return false;
  const int u = vec_per_step(items);
  const int64_t nvec = n / 2 > 0 ? n / 2 : 1;
  int64_t steps = items / (2 * u);
  int64_t grid = ceil_div(nvec, steps * bs * u);
  const int64_t cap = tune_or(tune, &rpx_tuning::grid_size, 0);
  if (cap > 0 && grid > cap) {  // a grid cap lengthens each workgroup's chunk
    steps = ceil_div(nvec, cap * bs * u);
    grid = ceil_div(nvec, steps * bs * u);
  }
  if (grid > INT32_MAX) return false;
  l->grid = grid;
  l->steps = steps;
  l->ngroups = ceil_div(grid, kGroup);
  l->gtickets_off = kHeaderBytes;
  l->gpartials_off = l->gtickets_off + size_t(l->ngroups) * kTicketStride * sizeof(unsigned);
  l->partials_off = l->gpartials_off + align256(size_t(l->ngroups) * sizeof(double));
  l->total = l->partials_off + align256(size_t(grid) * sizeof(double));
  return true;
}

template <bool DOT, int BS, int U>
void launch(bool nt, const ReduceLayout& l, const double* a, const double* b, int64_t n, char* base,
            double* result, int accumulate, hipStream_t s) {
  auto* ticket = reinterpret_cast<unsigned*>(base);
  auto* gtickets = reinterpret_cast<unsigned*>(base + l.gtickets_off);
  auto* gpartials = reinterpret_cast<double*>(base + l.gpartials_off);
  auto* partials = reinterpret_cast<double*>(base + l.partials_off);
  const dim3 grid(unsigned(l.grid)), block(BS);
  if (nt)
    reduce_kernel<DOT, BS, U, true><<<grid, block, 0, s>>>(a, b, n, l.steps, l.ngroups, partials, gpartials,
                                                            gtickets, ticket, result, accumulate);
  else
    reduce_kernel<DOT, BS, U, false><<<grid, block, 0, s>>>(a, b, n, l.steps, l.ngroups, partials, gpartials,
                                                             gtickets, ticket, result, accumulate);
}

template <bool DOT, int BS>
void launch_u(int u, bool nt, const ReduceLayout& l, const double* a, const double* b, int64_t n, char* base,
              double* result, int accumulate, hipStream_t s) {
  switch (u) {
    case 1: launch<DOT, BS, 1>(nt, l, a, b, n, base, result, accumulate, s); break;
    case 2: launch<DOT, BS, 2>(nt, l, a, b, n, base, result, accumulate, s); break;
    default: launch<DOT, BS, 4>(nt, l, a, b, n, base, result, accumulate, s); break;
  }
}

template <bool DOT>
int run_reduce(const double* a, const double* b, int64_t n, double* result, int accumulate, void* scratch,
               size_t scratch_bytes, const rpx_tuning* tune, void* stream) {
  if (n < 0 || !result || (n > 0 && (!a || (DOT && !b)))) return RPX_ERR_INVALID_ARGUMENT;
  if (!scratch) return RPX_ERR_INVALID_ARGUMENT;
  if (!aligned16(a) || (DOT && !aligned16(b))) return RPX_ERR_INVALID_ARGUMENT;
  const int bs = tune_or(tune, &rpx_tuning::block_size, kDefaultBlock);
  const int items = tune_or(tune, &rpx_tuning::items_per_thread, kDefaultItems);
  ReduceLayout l;
  if (!layout_for(n, bs, items, tune, &l)) return RPX_ERR_UNSUPPORTED;
  if (scratch_bytes < l.total) return RPX_ERR_SCRATCH_TOO_SMALL;
  const bool nt = (tune_flags(tune) & RPX_TUNE_CACHED) == 0;
  const int u = vec_per_step(items);
  char* base = reinterpret_cast<char*>(scratch);
  hipStream_t s = as_stream(stream);
  switch (bs) {
    case 128: launch_u<DOT, 128>(u, nt, l, a, b, n, base, result, accumulate, s); break;
    case 256: launch_u<DOT, 256>(u, nt, l, a, b, n, base, result, accumulate, s); break;
    case 512: launch_u<DOT, 512>(u, nt, l, a, b, n, base, result, accumulate, s); break;
    default: launch_u<DOT, 1024>(u, nt, l, a, b, n, base, result, accumulate, s); break;
  }
  return launch_status();
}

// Per-device address of `scratch`, looked up on a device's first
// allocation-free call with that device current.  The tickets in the array
// are zeroed then, on the caller's stream; every complete call leaves them
// zero again.
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};

int noalloc_scratch(hipStream_t s, void** out) {
  int dev = -1;
  RPX_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return RPX_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    RPX_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    const int64_t words = int64_t(sizeof(scratch) / sizeof(unsigned));
    zero_u32_kernel<<<dim3(unsigned(ceil_div(words, 256))), dim3(256), 0, s>>>(reinterpret_cast<unsigned*>(p),
                                                                                 words);
    const int st = launch_status();
    if (st != RPX_SUCCESS) return st;
    g_noalloc_ptr[dev] = p;
  }
  *out = g_noalloc_ptr[dev];
  return RPX_SUCCESS;
}

}  // namespace

extern "C" {

size_t rpx_reduce_scratch_bytes(int64_t n, const rpx_tuning* tune) {
  const int bs = tune_or(tune, &rpx_tuning::block_size, kDefaultBlock);
  const int items = tune_or(tune, &rpx_tuning::items_per_thread, kDefaultItems);
  ReduceLayout l;
  if (!layout_for(n, bs, items, tune, &l)) return 0;
  return l.total;
}

// Zeroes the whole scratch buffer (all ticket levels); required once.
int rpx_reduce_scratch_init(void* scratch, size_t scratch_bytes, void* stream) {
  if (!scratch || scratch_bytes < kHeaderBytes) return RPX_ERR_INVALID_ARGUMENT;
  const int64_t words = int64_t(scratch_bytes / sizeof(unsigned));
  int64_t g = ceil_div(words, 256);
  if (g > 4096) g = 4096;
  zero_u32_kernel<<<dim3(unsigned(g)), dim3(256), 0, as_stream(stream)>>>(reinterpret_cast<unsigned*>(scratch),
                                                                         words);
  return launch_status();
}

int rpx_dot_f64(const double* a, const double* b, int64_t n, double* result, int accumulate, void* scratch,
                size_t scratch_bytes, const rpx_tuning* tune, void* stream) {
  return run_reduce<true>(a, b, n, result, accumulate, scratch, scratch_bytes, tune, stream);
}

size_t rpx_dot_noalloc_max_scratch_bytes(void) { return sizeof(scratch); }

int rpx_dot_f64_noalloc(const double* a, const double* b, int64_t n, double* result, int accumulate,
                        const rpx_tuning* tune, void* stream) {
  if (n < 0 || !result || (n > 0 && (!a || !b))) return RPX_ERR_INVALID_ARGUMENT;
  const size_t need = rpx_reduce_scratch_bytes(n, tune);
  if (need == 0) return RPX_ERR_UNSUPPORTED;
  if (need > sizeof(scratch)) return RPX_ERR_SCRATCH_TOO_SMALL;
  void* p = nullptr;
  const int st = noalloc_scratch(as_stream(stream), &p);
  if (st != RPX_SUCCESS) return st;
  return run_reduce<true>(a, b, n, result, accumulate, p, sizeof(scratch), tune, stream);
}

int rpx_reduce_sum_f64(const double* x, int64_t n, double* result, void* scratch, size_t scratch_bytes,
                       const rpx_tuning* tune, void* stream) {
  return run_reduce<false>(x, nullptr, n, result, 0, scratch, scratch_bytes, tune, stream);
}

}  // extern "C"
