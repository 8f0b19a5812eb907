// SCAN: exclusive FP64 prefix sum in one pass with decoupled look-back (and
// the inclusive form as an extra output).
//
// Reference: src/algorithm/SCAN.hpp:27-32 (y[i] = s; s += x[i]),
// SCAN-Seq.cpp:31-45, SCAN-OMP.cpp:33-90, SCAN-Hip.cpp:39-326 and the
// suite's grid scan src/common/HipGridScan.hpp:34-197, which memsets its
// ready flags every rep and spins on predecessors indexed by blockIdx.x.
//
// MI355X design:
//  * a tile is 256 work-items x IPT doubles (IPT 16, 32 or 64: 32-128 KiB).
//    Each wavefront owns a contiguous quarter of the tile and reads it as
//    IPT/2 rows of 1 KiB (one 16-byte load per lane per row); it scans the
//    rows in order in registers -- lane pair sum, a 64-lane scan on DPP row
//    shifts and row broadcasts, and a running row carry -- so no LDS buffer
//    limits residency, and the result leaves through the same 16-byte row
//    stores.  Row loads and stores are nontemporal unless RPX_TUNE_CACHED is
//    set.  Larger tiles mean fewer look-back hops: 2^28 doubles scan at 4.2,
//    5.1 and 6.0 TB/s with IPT 16, 32 and 64 on MI355X (4.0, 4.9 and 5.5
//    TB/s with the 64-lane scan on ds_bpermute shuffles);
//  * the tile id is the workgroup id: each XCD dispatches its workgroups in
//    id order, so the lowest unfinished tile is always resident and the
//    look-back always makes progress.  (A ticket counter taken with one
//    atomic per tile serialises on its single address -- measured 5.5 ms
//    against 0.8 ms for 2^28 doubles on MI355X.)
//  * the look-back is done by one wavefront over windows of 64
//    predecessors; flags carry a tag (header epoch + 1) so the scratch needs
//    no clearing between calls.  The last tile advances the epoch after its
//    look-back:
//    by then every tile has published a flag, so every tile has read the
//    epoch;
//  * every spin is bounded; a wait that expires records RPX_ERR_WAIT_EXPIRED
//    in the scratch header and the kernel still runs to completion.
#include "rpx_common.hpp"

#include <vector>

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
constexpr int kDefaultIPT = 64;
constexpr int kDefaultScanBS = 256;  // work-items per tile (256 or 512)
constexpr int kMinTile = kBS * 16;  // scratch is sized for the smallest tile
constexpr unsigned kSpinLimit = 1u << 22;

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

// Wavefront scans on DPP lane moves (VALU only, no LDS round trips).  A
// 64-bit value moves as two 32-bit halves; lanes whose source is outside
// the pattern, or in a row the mask leaves out, receive 0.
template <int CTRL, int ROW_MASK, bool ZERO_FILL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, ROW_MASK, 0xF, ZERO_FILL);
  const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, ROW_MASK, 0xF, ZERO_FILL);
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(unsigned(hi)) << 32) |
                                                     unsigned(lo)));
}

// Inclusive scan over the 64 lanes: shifts by 1, 2, 4 and 8 within each row
// of 16 lanes, then lane 15's row total into rows 1 and 3 (row_bcast:15)
// and lane 31's half total into rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ double wave_inclusive_scan_dpp(double v) {
  v += dpp_move<0x111, 0xF, false>(v);  // row_shr:1
  v += dpp_move<0x112, 0xF, false>(v);  // row_shr:2
  v += dpp_move<0x114, 0xF, false>(v);  // row_shr:4
  v += dpp_move<0x118, 0xF, false>(v);  // row_shr:8
  v += dpp_move<0x142, 0xA, false>(v);  // row_bcast:15
  v += dpp_move<0x143, 0xC, false>(v);  // row_bcast:31
  return v;
}

// Lane l receives lane l-1's value, lane 0 receives 0 (wave_shr:1).
__device__ __forceinline__ double wave_shift_up_dpp(double v) { return dpp_move<0x138, 0xF, true>(v); }
// Lane l receives lane l+1's value, lane 63 receives 0 (wave_shl:1).
__device__ __forceinline__ double wave_shift_down_dpp(double v) { return dpp_move<0x130, 0xF, true>(v); }

// Lane L's value, wavefront-uniform.
template <int L>
__device__ __forceinline__ double lane_value(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(b), L);
  const int hi = __builtin_amdgcn_readlane(int(b >> 32), L);
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(unsigned(hi)) << 32) |
                                                     unsigned(lo)));
}
__device__ __forceinline__ double last_lane(double v) { return lane_value<kWave - 1>(v); }

// Segment tables of a segmented scan: segment s covers [off[s], off[s+1])
// and owns the global tiles [tile0[s], tile0[s+1]); tile t belongs to
// segment tile_seg[t].  A segment's tiles start at even elements (the
// segment's start rounded down to even, then whole tiles), so every tile
// after its first reads and writes 16-byte rows; the first tile skips the
// element before an odd start.
struct SegTables {
  const int64_t* off;
  const int64_t* tile0;
  const int* tile_seg;
};

// SEG = false: one array of n elements, tile t covers [t*T, t*T + T), flags
// carry the epoch read from the header.  SEG = true: tiles come from the
// segment tables, the look-back stops at the segment's first tile, and the
// flags were cleared before the launch (epoch 0, no epoch advance).
template <int BS, int ROWS, bool NT, bool SEG, bool INCL = false>  // 16-byte rows per lane: 128*ROWS doubles a wave
__global__ __launch_bounds__(BS) void scan_kernel(const double* x, double* y, int64_t n, int64_t ntiles,
                                                   ScanHeader* __restrict__ hdr,
                                                   unsigned long long* __restrict__ flags,
                                                   double* __restrict__ aggs, double* __restrict__ incls,
                                                   int vec, SegTables seg) {
  constexpr int kRowLen = 2 * kWave;            // doubles per wavefront row
  constexpr int kChunk = kRowLen * ROWS;        // doubles per wavefront
  constexpr int kTileN = kChunk * (BS / kWave);
  __shared__ double wave_tot[BS / kWave];
  __shared__ unsigned long long s_epoch;
  __shared__ double s_prefix;

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  if (!SEG && tid == 0) s_epoch = __hip_atomic_load(&hdr->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

  const int64_t t = blockIdx.x;
  // The tile spans [tal, tal + kTileN) from an even tal and owns the
  // elements [lo, tend) of it; first is its segment's first tile.
  int64_t tal, lo, tend, first;
  if (SEG) {
    const int sg = seg.tile_seg[t];
    first = seg.tile0[sg];
    const int64_t sbeg = seg.off[sg], send = seg.off[sg + 1];
    tal = (sbeg & ~int64_t(1)) + (t - first) * kTileN;
    lo = sbeg > tal ? sbeg : tal;
    tend = (tal + kTileN < send) ? tal + kTileN : send;
  } else {
    first = 0;
    tal = lo = t * kTileN;
    tend = (tal + kTileN < n) ? tal + kTileN : n;
  }
  const int64_t base = tal + int64_t(wave) * kChunk + 2 * lane;
  // x and y are 16-byte aligned (vec) and tal is even, so the rows are too.
  const bool full = vec && lo == tal && tend - tal == kTileN;
  n = tend;  // bounds of the partial-tile paths below

  // Each wavefront owns a contiguous chunk, read as ROWS coalesced 1 KiB rows.
  v2d v[ROWS];
  if (full) {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const v2d* p = reinterpret_cast<const v2d*>(x + base + r * kRowLen);
      v[r] = NT ? __builtin_nontemporal_load(p) : *p;
    }
  } else {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const int64_t i0 = base + r * kRowLen;
      v[r].x = (i0 >= lo && i0 < n) ? x[i0] : 0.0;
      v[r].y = (i0 + 1 >= lo && i0 + 1 < n) ? x[i0 + 1] : 0.0;
    }
  }

  // Exclusive scan inside the wavefront, row by row: lane pair sums, a
  // 64-lane scan on DPP moves, and a running carry of the row totals.  The
  // pair sum replaces the pair's second element, so each row keeps two
  // values; the inclusive form (INCL) is built from these exclusive values
  // when it is stored.
  double carry = 0.0;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const double p = v[r].x + v[r].y;
    const double inc = wave_inclusive_scan_dpp(p);
    const double exc = wave_shift_up_dpp(inc);
    const double row_total = last_lane(inc);
    const double e0 = carry + exc;
    v[r] = v2d{e0, e0 + v[r].x};
    carry += row_total;
  }
  if (lane == 0) wave_tot[wave] = carry;
  __syncthreads();
  // Flag tag: 0 for the segmented form (its flags are cleared before every
  // launch), epoch + 1 for the single-array form, so a single-array scan
  // never accepts a flag a segmented call left in a shared scratch.
  const unsigned long long E = SEG ? 0ull : s_epoch + 1;
  double wave_off = 0.0, tile_agg = 0.0;
#pragma unroll
  for (int w = 0; w < BS / kWave; ++w) {
    const double wt = wave_tot[w];
    if (w < wave) wave_off += wt;
    tile_agg += wt;
  }

  // Publish the aggregate, then look back (wavefront 0).
  if (wave == 0) {
    if (t == first) {
      if (lane == 0) {
        store_value(&incls[t], tile_agg);
        store_flag(&flags[t], (E << 2) | 2ull);
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
        unsigned status = 2;  // before the first tile: an inclusive prefix of zero
        if (k >= first) {
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
        if (k >= first && !expired) {
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
    if (!SEG && lane == 0 && t == ntiles - 1) atomicAdd(&hdr->epoch, 1ull);
  }
  __syncthreads();

  const double off = s_prefix + wave_off;
  // The inclusive value of an element is the exclusive value of the element
  // after it: for lane l's pair, its own second value and lane l+1's first
  // (for lane 63 the next row's first, after the last row the wavefront's
  // total).  Element n-1 of a partial tile takes the same rule, since the
  // padding beyond n is zero.
  if constexpr (INCL) {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const double next = (r + 1 < ROWS) ? lane_value<0>(v[r + 1].x) : carry;
      const double shifted = wave_shift_down_dpp(v[r].x);
      v[r] = v2d{v[r].y, lane == kWave - 1 ? next : shifted};
    }
  }
  if (full) {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      v2d* p = reinterpret_cast<v2d*>(y + base + r * kRowLen);
      const v2d o = v2d{off + v[r].x, off + v[r].y};
      if (NT) __builtin_nontemporal_store(o, p);
      else *p = o;
    }
  } else {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const int64_t i0 = base + r * kRowLen;
      if (i0 >= lo && i0 < n) y[i0] = off + v[r].x;
      if (i0 + 1 >= lo && i0 + 1 < n) y[i0 + 1] = off + v[r].y;
    }
  }
}

__global__ void zero_kernel(unsigned* p, int64_t nwords) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nwords; i += stride) p[i] = 0u;
}

// Before a segmented scan: clears the flags and fills tile_seg from tile0
// (the last segment whose first tile is <= t owns tile t; empty segments
// share their first tile with the next segment and so are never chosen).
__global__ __launch_bounds__(256) void seg_prep_kernel(unsigned* __restrict__ flag_words, int64_t nwords,
                                                       const int64_t* __restrict__ tile0, int nseg,
                                                       int* __restrict__ tile_seg, int64_t ntiles) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int64_t end = nwords > ntiles ? nwords : ntiles;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < end; i += stride) {
    if (i < nwords) flag_words[i] = 0u;
    if (i < ntiles) {
      int a = 0, b = nseg - 1;
      while (a < b) {
        const int mid = (a + b + 1) >> 1;
        if (tile0[mid] <= i) a = mid;
        else b = mid - 1;
      }
      tile_seg[i] = a;
    }
  }
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

template <int ROWS, bool SEG, int BS = kBS, bool INCL = false>
void launch_rows(bool nt, unsigned grid, const double* x, double* y, int64_t n, int64_t ntiles, ScanHeader* hdr,
                 unsigned long long* flags, double* aggs, double* incls, int vec, SegTables seg, hipStream_t s) {
  if (nt)
    scan_kernel<BS, ROWS, true, SEG, INCL><<<dim3(grid), dim3(BS), 0, s>>>(x, y, n, ntiles, hdr, flags, aggs, incls,
                                                                            vec, seg);
  else
    scan_kernel<BS, ROWS, false, SEG, INCL><<<dim3(grid), dim3(BS), 0, s>>>(x, y, n, ntiles, hdr, flags, aggs, incls,
                                                                             vec, seg);
}

template <bool INCL>
int scan_one(const double* x, double* y, int64_t n, int ipt, int bs, bool nt, void* scratch, size_t scratch_bytes,
             hipStream_t s) {
  if (n == 0) return RPX_SUCCESS;
  if (ipt != 16 && ipt != 32 && ipt != 64) return RPX_ERR_INVALID_ARGUMENT;
  if (bs != 256 && bs != 512) return RPX_ERR_INVALID_ARGUMENT;
  const int64_t ntiles = ceil_div(n, int64_t(bs) * ipt);
  if (ntiles > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  const Layout l = layout_for(ntiles);
  if (scratch_bytes < l.total) return RPX_ERR_SCRATCH_TOO_SMALL;
  char* base = reinterpret_cast<char*>(scratch);
  auto* hdr = reinterpret_cast<ScanHeader*>(base);
  auto* flags = reinterpret_cast<unsigned long long*>(base + l.flags_off);
  auto* aggs = reinterpret_cast<double*>(base + l.aggs_off);
  auto* incls = reinterpret_cast<double*>(base + l.incls_off);
  const int vec = aligned16(x) && aligned16(y);
  const unsigned grid = unsigned(ntiles);
  const SegTables none{nullptr, nullptr, nullptr};
  if (bs == 512) {
    if (ipt == 16) launch_rows<8, false, 512, INCL>(nt, grid, x, y, n, ntiles, hdr, flags, aggs, incls, vec, none, s);
    else if (ipt == 32)
      launch_rows<16, false, 512, INCL>(nt, grid, x, y, n, ntiles, hdr, flags, aggs, incls, vec, none, s);
    else launch_rows<32, false, 512, INCL>(nt, grid, x, y, n, ntiles, hdr, flags, aggs, incls, vec, none, s);
  } else {
    if (ipt == 16) launch_rows<8, false, kBS, INCL>(nt, grid, x, y, n, ntiles, hdr, flags, aggs, incls, vec, none, s);
    else if (ipt == 32)
      launch_rows<16, false, kBS, INCL>(nt, grid, x, y, n, ntiles, hdr, flags, aggs, incls, vec, none, s);
    else launch_rows<32, false, kBS, INCL>(nt, grid, x, y, n, ntiles, hdr, flags, aggs, incls, vec, none, s);
  }
  return launch_status();
}

// Segmented scan: every tile of every segment in one launch (tiles of
// 256 x 64 doubles from the segment's even-rounded start; a segment's first
// and last tiles may be partial).  Scratch: the single-array layout for the
// total tile count, then off[nseg+1] and tile0[nseg+1] (one upload), then
// tile_seg[ntiles] (filled on the device).
constexpr int64_t kSegTile = int64_t(kBS) * kDefaultIPT;

inline int64_t seg_tiles(int64_t sbeg, int64_t send) {
  return send > sbeg ? ceil_div(send - (sbeg & ~int64_t(1)), kSegTile) : 0;
}

struct SegLayout {
  Layout l;
  int64_t ntiles;
  size_t tab_off, tile_seg_off, total;
};
SegLayout seg_layout(const int64_t* offsets_host, int64_t nseg) {
  SegLayout g;
  g.ntiles = 0;
  for (int64_t s = 0; s < nseg; ++s) g.ntiles += seg_tiles(offsets_host[s], offsets_host[s + 1]);
  g.l = layout_for(g.ntiles > 0 ? g.ntiles : 1);
  g.tab_off = g.l.total;
  g.tile_seg_off = g.tab_off + align256(size_t(2 * (nseg + 1)) * sizeof(int64_t));
  g.total = g.tile_seg_off + align256(size_t(g.ntiles) * sizeof(int));
  return g;
}

}  // namespace

extern "C" {

size_t rpx_scan_scratch_bytes(int64_t n, const rpx_tuning* /*tune*/) {
  if (n < 0) return 0;
  return layout_for(ceil_div(n > 0 ? n : 1, kMinTile)).total;
}

int rpx_scan_exclusive_f64(const double* x, double* y, int64_t n, void* scratch, size_t scratch_bytes,
                           const rpx_tuning* tune, void* stream) {
  if (n < 0 || (n > 0 && (!x || !y || !scratch))) return RPX_ERR_INVALID_ARGUMENT;
  const int ipt = tune_or(tune, &rpx_tuning::items_per_thread, kDefaultIPT);
  const int bs = tune_or(tune, &rpx_tuning::block_size, kDefaultScanBS);
  const bool nt = (tune_flags(tune) & RPX_TUNE_CACHED) == 0;
  return scan_one<false>(x, y, n, ipt, bs, nt, scratch, scratch_bytes, as_stream(stream));
}

int rpx_scan_inclusive_f64(const double* x, double* y, int64_t n, void* scratch, size_t scratch_bytes,
                           const rpx_tuning* tune, void* stream) {
  if (n < 0 || (n > 0 && (!x || !y || !scratch))) return RPX_ERR_INVALID_ARGUMENT;
  const int ipt = tune_or(tune, &rpx_tuning::items_per_thread, kDefaultIPT);
  const int bs = tune_or(tune, &rpx_tuning::block_size, kDefaultScanBS);
  const bool nt = (tune_flags(tune) & RPX_TUNE_CACHED) == 0;
  return scan_one<true>(x, y, n, ipt, bs, nt, scratch, scratch_bytes, as_stream(stream));
}

size_t rpx_scan_segmented_scratch_bytes(const int64_t* offsets_host, int64_t nseg) {
  if (!offsets_host || nseg < 0 || nseg > INT32_MAX) return 0;
  for (int64_t s = 0; s < nseg; ++s)
    if (offsets_host[s + 1] < offsets_host[s]) return 0;
  return seg_layout(offsets_host, nseg).total;
}

int rpx_scan_exclusive_segmented_f64(const double* x, double* y, const int64_t* offsets_host, int64_t nseg,
                                     void* scratch, size_t scratch_bytes, void* stream) {
  if (!offsets_host || nseg < 0 || nseg > INT32_MAX || !scratch) return RPX_ERR_INVALID_ARGUMENT;
  for (int64_t s = 0; s < nseg; ++s)
    if (offsets_host[s + 1] < offsets_host[s] || offsets_host[s] < 0) return RPX_ERR_INVALID_ARGUMENT;
  if (nseg > 0 && offsets_host[nseg] > 0 && (!x || !y)) return RPX_ERR_INVALID_ARGUMENT;
  const SegLayout g = seg_layout(offsets_host, nseg);
  if (scratch_bytes < g.total) return RPX_ERR_SCRATCH_TOO_SMALL;
  if (g.ntiles == 0) return RPX_SUCCESS;
  if (g.ntiles > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  hipStream_t st = as_stream(stream);
  char* base = reinterpret_cast<char*>(scratch);
  // Tables: the offsets, then each segment's first tile (one copy).
  std::vector<int64_t> tab(size_t(2 * (nseg + 1)));
  int64_t t = 0;
  for (int64_t s = 0; s <= nseg; ++s) {
    tab[size_t(s)] = offsets_host[s];
    tab[size_t(nseg + 1 + s)] = t;
    if (s < nseg) t += seg_tiles(offsets_host[s], offsets_host[s + 1]);
  }
  auto* d_off = reinterpret_cast<int64_t*>(base + g.tab_off);
  auto* d_tile0 = d_off + (nseg + 1);
  auto* d_tile_seg = reinterpret_cast<int*>(base + g.tile_seg_off);
  const HostChunk table{d_off, tab.data(), tab.size() * sizeof(int64_t)};
  if (const int rc = upload_async(&table, 1, st); rc != RPX_SUCCESS) return rc;
  // Flags start at zero for every call (no epochs across segments).
  auto* flags = reinterpret_cast<unsigned long long*>(base + g.l.flags_off);
  const int64_t words = g.ntiles * 2;
  const int64_t work = words > g.ntiles ? words : g.ntiles;
  seg_prep_kernel<<<dim3(unsigned(ceil_div(work, 256) < 4096 ? ceil_div(work, 256) : 4096)), dim3(256), 0, st>>>(
      reinterpret_cast<unsigned*>(flags), words, d_tile0, int(nseg), d_tile_seg, g.ntiles);
  if (const int rc = launch_status(); rc != RPX_SUCCESS) return rc;
  const SegTables seg{d_off, d_tile0, d_tile_seg};
  const int vec = aligned16(x) && aligned16(y);
  launch_rows<32, true>(true, unsigned(g.ntiles), x, y, offsets_host[nseg], g.ntiles,
                        reinterpret_cast<ScanHeader*>(base), flags, reinterpret_cast<double*>(base + g.l.aggs_off),
                        reinterpret_cast<double*>(base + g.l.incls_off), vec, seg, st);
  return launch_status();
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
