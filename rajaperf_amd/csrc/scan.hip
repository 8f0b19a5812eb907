// SCAN: exclusive FP64 prefix sum in one pass with decoupled look-back.
//
// Reference: src/algorithm/SCAN.hpp:27-32 (y[i] = s; s += x[i]),
// SCAN-Seq.cpp:31-45, SCAN-OMP.cpp:33-90, SCAN-Hip.cpp:39-326 and the
// suite's grid scan src/common/HipGridScan.hpp:34-197, which memsets its
// ready flags every rep and spins on predecessors indexed by blockIdx.x.
//
// MI355X design:
//  * a tile is 256 work-items x 16 doubles = 32 KiB, loaded with striped
//    16-byte loads (each wavefront instruction covers 1 KiB), transposed
//    through padded LDS to a blocked layout for the per-thread serial scan,
//    and transposed back for striped 16-byte stores;
//  * the tile id is the workgroup id: each XCD dispatches its workgroups in
//    id order, so the lowest unfinished tile is always resident and the
//    look-back always makes progress.  (A ticket counter taken with one
//    atomic per tile serialises on its single address -- measured 5.5 ms
//    against 0.8 ms for 2^28 doubles on MI355X.)
//  * the look-back is done by one wavefront over windows of 64
//    predecessors; flags carry an epoch so the scratch needs no clearing
//    between calls.  The last tile advances the epoch after its look-back:
//    by then every tile has published a flag, so every tile has read the
//    epoch;
//  * every spin is bounded; a wait that expires records RPX_ERR_WAIT_EXPIRED
//    in the scratch header and the kernel still runs to completion.
#include "rpx_common.hpp"

using namespace rpx;

namespace {

typedef double v2d __attribute__((ext_vector_type(2)));

struct ScanHeader {
  unsigned reserved0;
  unsigned reserved1;
  unsigned long long epoch;
  unsigned error;
};
constexpr size_t kHeaderBytes = 256;
constexpr int kBS = 256;
constexpr int kIPT = 16;
constexpr int kTile = kBS * kIPT;
constexpr int kPadShift = 4;  // one pad double per 16
constexpr int kLdsDoubles = kTile + (kTile >> kPadShift);
constexpr unsigned kSpinLimit = 1u << 22;

__device__ __forceinline__ int pad(int e) { return e + (e >> kPadShift); }

// All look-back state (flags and values) moves through agent-scope atomic
// loads and stores, which gfx950 performs coherently across the XCDs' L2s
// (sc1).  Publishing a value before its flag therefore needs only the value
// store to have completed (s_waitcnt vmcnt(0)) -- not the L2 write-back that
// an agent-scope release fence would issue, which would also flush every
// dirty output line of the XCD.  On the reading side the value load is only
// issued after the flag load has returned and been tested.
__device__ __forceinline__ unsigned long long load_flag(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_flag(unsigned long long* p, unsigned long long v) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // compiler ordering
  __builtin_amdgcn_s_waitcnt(0x0F70);                      // vmcnt(0): value store complete
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_value(const double* p) {
  unsigned long long bits =
      __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double(static_cast<long long>(bits));
}
__device__ __forceinline__ void store_value(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                     static_cast<unsigned long long>(__double_as_longlong(v)), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kBS) void scan_kernel(const double* x, double* y, int64_t n, int64_t ntiles,
                                                   ScanHeader* __restrict__ hdr,
                                                   unsigned long long* __restrict__ flags,
                                                   double* __restrict__ aggs, double* __restrict__ incls) {
  __shared__ double tile[kLdsDoubles];
  __shared__ double wave_tot[kBS / kWave];
  __shared__ unsigned long long s_epoch;
  __shared__ double s_prefix;

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;

  if (tid == 0) s_epoch = __hip_atomic_load(&hdr->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int64_t t = blockIdx.x;
  const unsigned long long E = s_epoch;
  const int64_t start = t * kTile;
  const int64_t valid = (n - start < kTile) ? (n - start) : kTile;

  // Striped load -> LDS.
  const bool full = (valid == kTile) && ((reinterpret_cast<uintptr_t>(x + start) & 15u) == 0);
  if (full) {
    const v2d* xv = reinterpret_cast<const v2d*>(x + start);
    v2d r[kIPT / 2];
#pragma unroll
    for (int k = 0; k < kIPT / 2; ++k) r[k] = xv[k * kBS + tid];
#pragma unroll
    for (int k = 0; k < kIPT / 2; ++k) {
      const int e = 2 * (k * kBS + tid);
      tile[pad(e)] = r[k].x;
      tile[pad(e + 1)] = r[k].y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
      const int e = k * kBS + tid;
      tile[pad(e)] = (e < valid) ? x[start + e] : 0.0;
    }
  }
  __syncthreads();

  // Blocked serial scan of kIPT items per work-item.
  double v[kIPT];
  double tsum = 0.0;
#pragma unroll
  for (int i = 0; i < kIPT; ++i) {
    v[i] = tile[pad(tid * kIPT + i)];
    tsum += v[i];
  }
  // Work-item totals: wave scan, then across waves.
  const double incl = wave_inclusive_scan(tsum, lane);
  if (lane == kWave - 1) wave_tot[wave] = incl;
  __syncthreads();
  double wave_off = 0.0, tile_agg = 0.0;
#pragma unroll
  for (int w = 0; w < kBS / kWave; ++w) {
    const double wt = wave_tot[w];
    if (w < wave) wave_off += wt;
    tile_agg += wt;
  }
  const double thread_excl = wave_off + (incl - tsum);

  // Publish the aggregate, then look back (wavefront 0).
  if (wave == 0) {
    if (t == 0) {
      if (lane == 0) {
        store_value(&incls[0], tile_agg);
        store_flag(&flags[0], (E << 2) | 2ull);
        s_prefix = 0.0;
      }
    } else {
      if (lane == 0) {
        store_value(&aggs[t], tile_agg);
        store_flag(&flags[t], (E << 2) | 1ull);
      }
      double excl = 0.0;
      int64_t top = t - 1;
      bool expired = false;
      while (true) {
        const int64_t k = top - lane;
        unsigned status = 2;  // before tile 0: an inclusive prefix of zero
        if (k >= 0) {
          unsigned spins = 0;
          while (true) {
            const unsigned long long f = load_flag(&flags[k]);
            status = static_cast<unsigned>(f & 3ull);
            if ((f >> 2) == E && status != 0) break;
            if (++spins > kSpinLimit) {
              expired = true;
              status = 2;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");  // value loads stay after the flag tests
        const unsigned long long inc_mask = __ballot(status == 2);
        const int first_inc = inc_mask ? (__ffsll(static_cast<long long>(inc_mask)) - 1) : kWave;
        double val = 0.0;
        if (k >= 0 && !expired) {
          if (lane < first_inc) val = load_value(&aggs[k]);
          else if (lane == first_inc) val = load_value(&incls[k]);
        }
        excl += wave_sum(val);
        if (inc_mask) break;
        top -= kWave;
      }
      if (__ballot(expired) != 0 && lane == 0) atomicOr(&hdr->error, 1u);
      if (lane == 0) {
        store_value(&incls[t], excl + tile_agg);
        store_flag(&flags[t], (E << 2) | 2ull);
        s_prefix = excl;
      }
    }
    // Every tile has read the epoch once the last tile's look-back is done.
    if (lane == 0 && t == ntiles - 1) atomicAdd(&hdr->epoch, 1ull);
  }
  __syncthreads();

  // Exclusive values: running sum from the tile prefix plus the work-item offset.
  double run = s_prefix + thread_excl;
#pragma unroll
  for (int i = 0; i < kIPT; ++i) {
    const double xi = v[i];
    v[i] = run;
    run += xi;
  }
  __syncthreads();  // everyone has read its blocked inputs
#pragma unroll
  for (int i = 0; i < kIPT; ++i) tile[pad(tid * kIPT + i)] = v[i];
  __syncthreads();

  const bool full_out = (valid == kTile) && ((reinterpret_cast<uintptr_t>(y + start) & 15u) == 0);
  if (full_out) {
    v2d* yv = reinterpret_cast<v2d*>(y + start);
#pragma unroll
    for (int k = 0; k < kIPT / 2; ++k) {
      const int e = 2 * (k * kBS + tid);
      v2d r;
      r.x = tile[pad(e)];
      r.y = tile[pad(e + 1)];
      yv[k * kBS + tid] = r;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
      const int e = k * kBS + tid;
      if (e < valid) y[start + e] = tile[pad(e)];
    }
  }

}

__global__ void zero_kernel(unsigned* p, int64_t nwords) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nwords; i += stride) p[i] = 0u;
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct Layout {
  size_t flags_off, aggs_off, incls_off, total;
};
Layout layout_for(int64_t ntiles) {
  Layout l;
  l.flags_off = kHeaderBytes;
  l.aggs_off = l.flags_off + align256(size_t(ntiles) * 8);
  l.incls_off = l.aggs_off + align256(size_t(ntiles) * 8);
  l.total = l.incls_off + align256(size_t(ntiles) * 8);
  return l;
}

int scan_one(const double* x, double* y, int64_t n, void* scratch, size_t scratch_bytes, hipStream_t s) {
  if (n == 0) return RPX_SUCCESS;
  const int64_t ntiles = ceil_div(n, kTile);
  if (ntiles > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  const Layout l = layout_for(ntiles);
  if (scratch_bytes < l.total) return RPX_ERR_SCRATCH_TOO_SMALL;
  char* base = reinterpret_cast<char*>(scratch);
  scan_kernel<<<dim3(unsigned(ntiles)), dim3(kBS), 0, s>>>(
      x, y, n, ntiles, reinterpret_cast<ScanHeader*>(base), reinterpret_cast<unsigned long long*>(base + l.flags_off),
      reinterpret_cast<double*>(base + l.aggs_off), reinterpret_cast<double*>(base + l.incls_off));
  return launch_status();
}

}  // namespace

extern "C" {

size_t rpx_scan_scratch_bytes(int64_t n, const rpx_tuning* /*tune*/) {
  if (n < 0) return 0;
  return layout_for(ceil_div(n > 0 ? n : 1, kTile)).total;
}

int rpx_scan_exclusive_f64(const double* x, double* y, int64_t n, void* scratch, size_t scratch_bytes,
                           const rpx_tuning* /*tune*/, void* stream) {
  if (n < 0 || (n > 0 && (!x || !y || !scratch))) return RPX_ERR_INVALID_ARGUMENT;
  return scan_one(x, y, n, scratch, scratch_bytes, as_stream(stream));
}

size_t rpx_scan_segmented_scratch_bytes(const int64_t* offsets_host, int64_t nseg) {
  if (!offsets_host || nseg < 0) return 0;
  int64_t maxlen = 1;
  for (int64_t s = 0; s < nseg; ++s) {
    const int64_t len = offsets_host[s + 1] - offsets_host[s];
    if (len > maxlen) maxlen = len;
  }
  return rpx_scan_scratch_bytes(maxlen, nullptr);
}

int rpx_scan_exclusive_segmented_f64(const double* x, double* y, const int64_t* offsets_host, int64_t nseg,
                                     void* scratch, size_t scratch_bytes, void* stream) {
  if (!offsets_host || nseg < 0 || !scratch) return RPX_ERR_INVALID_ARGUMENT;
  for (int64_t s = 0; s < nseg; ++s)
    if (offsets_host[s + 1] < offsets_host[s] || offsets_host[s] < 0) return RPX_ERR_INVALID_ARGUMENT;
  if (nseg > 0 && offsets_host[nseg] > 0 && (!x || !y)) return RPX_ERR_INVALID_ARGUMENT;
  // Segments run back to back on the stream; each reuses the same scratch,
  // which is legal because consecutive launches on one stream are ordered.
  for (int64_t s = 0; s < nseg; ++s) {
    const int64_t b = offsets_host[s], len = offsets_host[s + 1] - b;
    const int rc = scan_one(x + b, y + b, len, scratch, scratch_bytes, as_stream(stream));
    if (rc != RPX_SUCCESS) return rc;
  }
  return RPX_SUCCESS;
}

// Zeroes a scan scratch buffer; required once before its first use.
int rpx_scan_scratch_init(void* scratch, size_t scratch_bytes, void* stream) {
  if (!scratch || scratch_bytes < kHeaderBytes) return RPX_ERR_INVALID_ARGUMENT;
  const int64_t words = int64_t(scratch_bytes / 4);
  int64_t g = ceil_div(words, 256);
  if (g > 4096) g = 4096;
  zero_kernel<<<dim3(unsigned(g)), dim3(256), 0, as_stream(stream)>>>(reinterpret_cast<unsigned*>(scratch), words);
  return launch_status();
}

// Reads the scratch header's error word (synchronous).
int rpx_scan_scratch_status(const void* scratch) {
  if (!scratch) return RPX_ERR_INVALID_ARGUMENT;
  ScanHeader h;
  RPX_TRY(hipMemcpy(&h, scratch, sizeof(h), hipMemcpyDeviceToHost));
  return h.error ? RPX_ERR_WAIT_EXPIRED : RPX_SUCCESS;
}

}  // extern "C"
