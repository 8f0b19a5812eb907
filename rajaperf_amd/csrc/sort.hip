// SORT: segmented ascending sort of FP64 keys, in place.
//
// Reference: src/algorithm/SORT.hpp:18-25 (rep irep sorts the length-N
// segment x + irep*N), SORT-Seq.cpp:31-42 (std::sort), SORT-Hip.cpp:35-43
// (RAJA::sort on the device; the suite has no hand-written GPU sort).
//
// MI355X design: least-significant-digit radix sort over 8 passes of 8-bit
// digits of an order-preserving 64-bit image of the double (sign bit set for
// non-negative values, all bits flipped for negative ones), so unsigned order
// is numeric order with -0.0 before +0.0.  Every segment is cut into tiles of
// 2048 keys; each pass is three launches over all tiles of all segments:
//   hist:    per-tile digit counts (LDS atomics) -> counts[seg][tile][digit]
//   offsets: one workgroup per segment scans its counts digit-major, giving
//            every (digit, tile) its output offset inside the segment (the
//            tile-major layout makes each step's 256 loads one coalesced row)
//   scatter: per tile, a stable rank of each key among equal digits (the
//            lanes of a row with equal digits found through a per-wavefront
//            LDS mask table, row by row, then a prefix over the wavefronts);
//            the tile is first reordered by digit in LDS, then written out so
//            that the keys of one digit go to consecutive addresses from
//            consecutive work-items.
// A wavefront owns 512 consecutive keys of its tile and reads them as 8
// coalesced rows of 64, which keeps the ranking order equal to the tile order
// (LSD stability).  2048-key tiles (22.5 KB of LDS, seven wavefronts per
// SIMD), the mask table in place of 8 ballots per row, and a write-out that
// reads all of a work-item's slots first took 256 ragged segments (35.8 M
// keys) from 2.32 to 2.20 ms together; alone, each gained 1-2%.  Pass 0 converts doubles to keys on load; pass 7 converts
// back on store.  The passes ping-pong between x and a scratch copy.
#include <vector>

#include "rpx_common.hpp"

using namespace rpx;

namespace {

constexpr int kBS = 256;
constexpr int kRows = 8;
constexpr int kTile = kBS * kRows;  // 2048 keys
constexpr int kRadix = 256;
constexpr int kPasses = 8;

typedef unsigned long long u64;

struct TileInfo {
  long long seg_x;       // first key of the tile's segment in x
  long long seg_alt;     // first key of the segment in the packed scratch copy
  long long count_base;  // first counter of the segment
  int len;               // keys in the tile
  int local;             // tile index inside the segment
  int ntiles;            // tiles of the segment
  int seg_len;           // keys in the segment
};

struct SegInfo {
  long long count_base;
  int ntiles;
  int pad;
};

// One per non-empty segment of a ragged layout, built on the host (O(nseg))
// and expanded on the device into the per-tile TileInfo table.
struct SegDesc {
  long long seg_x, seg_alt, count_base;
  int first_tile, ntiles, seg_len, pad;
};

// Uniform segmentation (equal lengths), or an explicit table in scratch.
struct Plan {
  const TileInfo* tiles;  // null => uniform
  const SegInfo* segs;
  long long seg_len;
  int tiles_per_seg;
};

__device__ __forceinline__ TileInfo get_tile(const Plan& p, int t) {
  if (p.tiles) return p.tiles[t];
  TileInfo ti;
  const int seg = t / p.tiles_per_seg;
  ti.local = t % p.tiles_per_seg;
  ti.ntiles = p.tiles_per_seg;
  ti.seg_x = (long long)seg * p.seg_len;
  ti.seg_alt = ti.seg_x;
  const long long rem = p.seg_len - (long long)ti.local * kTile;
  ti.len = int(rem < kTile ? rem : kTile);
  ti.count_base = (long long)seg * p.tiles_per_seg * kRadix;
  ti.seg_len = int(p.seg_len);
  return ti;
}

__device__ __forceinline__ u64 to_key(u64 b) { return (b >> 63) ? ~b : (b | 0x8000000000000000ull); }
__device__ __forceinline__ u64 from_key(u64 u) { return (u >> 63) ? (u & 0x7FFFFFFFFFFFFFFFull) : ~u; }

__device__ __forceinline__ int tile_pos(int wave, int r, int lane) { return wave * (kRows * kWave) + r * kWave + lane; }

// Per-wavefront sub-histograms: a digit shared by many keys (the top byte
// of doubles of one sign and binade) then contends within one wavefront only.
__global__ __launch_bounds__(kBS) void hist_kernel(const u64* __restrict__ src, unsigned* __restrict__ counts, Plan plan,
                                                   int shift, int convert, int src_is_x) {
  __shared__ unsigned h[kBS / kWave][kRadix];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const TileInfo ti = get_tile(plan, xcd_major(blockIdx.x, gridDim.x));
  const long long sbase = (src_is_x ? ti.seg_x : ti.seg_alt) + (long long)ti.local * kTile;
#pragma unroll
  for (int w = 0; w < kBS / kWave; ++w) h[w][tid] = 0u;
  __syncthreads();
  u64 k[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int p = tile_pos(wave, r, lane);
    k[r] = p < ti.len ? src[sbase + p] : 0ull;
  }
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    if (tile_pos(wave, r, lane) < ti.len) {
      const u64 key = convert ? to_key(k[r]) : k[r];
      atomicAdd(&h[wave][(key >> shift) & 0xFF], 1u);
    }
  }
  __syncthreads();
  counts[ti.count_base + (long long)ti.local * kRadix + tid] = h[0][tid] + h[1][tid] + h[2][tid] + h[3][tid];
}

// Exclusive scan over the 256 work-items of a workgroup (one value each).
__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned* wsum) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  unsigned incl = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const unsigned o = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += o;
  }
  if (lane == kWave - 1) wsum[wave] = incl;
  __syncthreads();
  unsigned run = incl - v;
  for (int w = 0; w < wave; ++w) run += wsum[w];
  return run;
}

// One workgroup per segment: exclusive scan of counts in digit-major order.
// kWalkers work-items share each digit: walker w of digit d visits the
// segment's tiles [w*q, (w+1)*q), q = ceil(nt / kWalkers), in batches of
// kBatch coalesced loads.  The largest segment's workgroup sets the kernel's
// time, and its walk is kWalkers times shorter than with one work-item per
// digit.  The walkers' sums give each digit's total (scanned over the
// digits) and each walker's starting offset; the walkers then write their
// tiles' offsets.
constexpr int kBatch = 32;
constexpr int kWalkers = 4;
constexpr int kOffThreads = kWalkers * kRadix;
__global__ __launch_bounds__(kOffThreads) void offsets_kernel(unsigned* __restrict__ counts, Plan plan) {
  __shared__ unsigned part[kWalkers][kRadix];
  __shared__ unsigned wsum[kRadix / kWave];
  long long base;
  int nt;
  if (plan.segs) {
    base = plan.segs[blockIdx.x].count_base;
    nt = plan.segs[blockIdx.x].ntiles;
  } else {
    base = (long long)blockIdx.x * plan.tiles_per_seg * kRadix;
    nt = plan.tiles_per_seg;
  }
  const int d = threadIdx.x % kRadix, w = threadIdx.x / kRadix;  // w is wavefront-uniform
  const int q = (nt + kWalkers - 1) / kWalkers;
  const int t0 = w * q < nt ? w * q : nt;
  const int t1 = (w + 1) * q < nt ? (w + 1) * q : nt;
  unsigned* col = counts + base + d;  // col[t * kRadix]: digit d of tile t
  unsigned s = 0;
  int t = t0;
  for (; t + kBatch <= t1; t += kBatch) {
    unsigned v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) v[u] = col[(long long)(t + u) * kRadix];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) s += v[u];
  }
  for (; t < t1; ++t) s += col[(long long)t * kRadix];
  part[w][d] = s;
  __syncthreads();
  // Walker 0's wavefronts: digit totals, scanned over the 256 digits.
  unsigned run = 0;
  if (w == 0) {
    unsigned total = 0;
#pragma unroll
    for (int k = 0; k < kWalkers; ++k) total += part[k][d];
    const int lane = d & (kWave - 1), wave = d / kWave;
    unsigned incl = total;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const unsigned o = __shfl_up(incl, off, kWave);
      if (lane >= off) incl += o;
    }
    if (lane == kWave - 1) wsum[wave] = incl;
    run = incl - total;
  }
  __syncthreads();
  if (w == 0) {
    for (int k = 0; k < d / kWave; ++k) run += wsum[k];
#pragma unroll
    for (int k = 0; k < kWalkers; ++k) {  // each walker's first offset
      const unsigned c = part[k][d];
      part[k][d] = run;
      run += c;
    }
  }
  __syncthreads();
  run = part[w][d];
  for (t = t0; t + kBatch <= t1; t += kBatch) {
    unsigned v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) v[u] = col[(long long)(t + u) * kRadix];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      col[(long long)(t + u) * kRadix] = run;
      run += v[u];
    }
  }
  for (; t < t1; ++t) {
    const unsigned c = col[(long long)t * kRadix];
    col[(long long)t * kRadix] = run;
    run += c;
  }
}

// Tiles are numbered XCD-major (each XCD takes a contiguous range of a
// segment's tiles): the partial 128-byte lines at the ends of adjacent tiles'
// digit runs are then written from one L2, close together in time.  Tile
// order does not change the result (every position comes from the offsets);
// measured 2.47 -> 2.33 ms on 256 ragged segments (35.8 M keys).
__global__ __launch_bounds__(kBS) void scatter_kernel(const u64* __restrict__ src, u64* __restrict__ dst,
                                                      const unsigned* __restrict__ counts, Plan plan, int shift,
                                                      int convert_in, int convert_out, int src_is_x) {
  __shared__ unsigned cnt[kBS / kWave][kRadix];
  __shared__ unsigned goff[kRadix];
  __shared__ unsigned dstart[kRadix];  // first LDS slot of each digit in the reordered tile
  __shared__ unsigned wsum[kBS / kWave];
  __shared__ u64 sorted[kTile];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const TileInfo ti = get_tile(plan, xcd_major(blockIdx.x, gridDim.x));
  const long long sbase = (src_is_x ? ti.seg_x : ti.seg_alt) + (long long)ti.local * kTile;
  const long long dbase = src_is_x ? ti.seg_alt : ti.seg_x;
#pragma unroll
  for (int w = 0; w < kBS / kWave; ++w) cnt[w][tid] = 0u;
  goff[tid] = counts[ti.count_base + (long long)ti.local * kRadix + tid];
  // Match masks of the ranking, one per wavefront and digit, in the part of
  // the reorder buffer that is not written before the ranking ends.
  static_assert(kTile >= (kBS / kWave) * kRadix, "match masks fit in the reorder buffer");
  u64(*match)[kRadix] = reinterpret_cast<u64(*)[kRadix]>(sorted);
#pragma unroll
  for (int w = 0; w < kBS / kWave; ++w) match[w][tid] = 0ull;
  __syncthreads();

  u64 key[kRows];
  unsigned rank[kRows];
  const u64 lane_below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int p = tile_pos(wave, r, lane);
    const bool valid = p < ti.len;
    u64 k = 0;
    if (valid) {
      k = src[sbase + p];
      if (convert_in) k = to_key(k);
    }
    key[r] = k;
    const unsigned d = unsigned(k >> shift) & 0xFFu;
    // The row's lanes of digit d: each valid lane ORs its bit into the
    // wavefront's mask of d and reads the mask back.  One wavefront's LDS
    // operations complete in issue order, each over all its lanes, so the
    // read sees every lane's bit, and the group's highest lane clears the
    // mask after its read, before the next row's ORs.
    u64 m = 0;
    if (valid) {
      atomicOr(&match[wave][d], 1ull << lane);
      m = match[wave][d];
    }
    unsigned base = 0;
    if (valid) base = cnt[wave][d];
    rank[r] = base + unsigned(__popcll(m & lane_below));
    // highest lane of the group advances the wavefront's digit counter
    if (valid && (m >> lane) == 1ull) {
      cnt[wave][d] = base + unsigned(__popcll(m));
      match[wave][d] = 0ull;
    }
  }
  __syncthreads();
  // Prefix over wavefronts for every digit (thread = digit), then over the
  // digits: each key's slot in the tile reordered by digit.
  unsigned total = 0;
#pragma unroll
  for (int w = 0; w < kBS / kWave; ++w) {
    const unsigned c = cnt[w][tid];
    cnt[w][tid] = total;
    total += c;
  }
  dstart[tid] = block_exclusive_scan(total, wsum);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int p = tile_pos(wave, r, lane);
    if (p < ti.len) {
      const unsigned d = unsigned(key[r] >> shift) & 0xFFu;
      sorted[dstart[d] + cnt[wave][d] + rank[r]] = key[r];
    }
  }
  __syncthreads();
  // Slot i holds the (i - dstart[d])-th key of digit d in the tile, which
  // goes to goff[d] + i - dstart[d]: runs of one digit are contiguous.
  if (ti.len == kTile) {
    // Full tile: all of the work-item's slots are read before any is
    // written out, so the LDS reads overlap instead of waiting in turn.
    u64 kk[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) kk[r] = sorted[tid + r * kBS];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int i = tid + r * kBS;
      const unsigned d = unsigned(kk[r] >> shift) & 0xFFu;
      const long long pos = (long long)goff[d] + (i - dstart[d]);
      if (pos < ti.seg_len) dst[dbase + pos] = convert_out ? from_key(kk[r]) : kk[r];
    }
  } else {
    for (int i = tid; i < ti.len; i += kBS) {
      const u64 k = sorted[i];
      const unsigned d = unsigned(k >> shift) & 0xFFu;
      const long long pos = (long long)goff[d] + (i - dstart[d]);
      if (pos < ti.seg_len) dst[dbase + pos] = convert_out ? from_key(k) : k;
    }
  }
}

// The ragged layout's tables from the per-segment descriptors: work-item t
// writes TileInfo t (its segment found by binary search over the
// descriptors' first tiles) and, for t < nsd, SegInfo t.
__global__ __launch_bounds__(256) void tile_table_kernel(const SegDesc* __restrict__ d, int nsd,
                                                         TileInfo* __restrict__ tiles, SegInfo* __restrict__ segs,
                                                         int total_tiles) {
  const int t = int(blockIdx.x) * 256 + int(threadIdx.x);
  if (t < nsd) segs[t] = SegInfo{d[t].count_base, d[t].ntiles, 0};
  if (t >= total_tiles) return;
  int lo = 0, hi = nsd - 1;  // the last descriptor whose first tile is <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].first_tile <= t) lo = mid;
    else hi = mid - 1;
  }
  const SegDesc sd = d[lo];
  TileInfo ti;
  ti.seg_x = sd.seg_x;
  ti.seg_alt = sd.seg_alt;
  ti.count_base = sd.count_base;
  ti.local = t - sd.first_tile;
  ti.ntiles = sd.ntiles;
  ti.seg_len = sd.seg_len;
  const int rem = sd.seg_len - ti.local * kTile;
  ti.len = rem < kTile ? rem : kTile;
  tiles[t] = ti;
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct Sizes {
  long long total_keys = 0, total_tiles = 0;
  bool uniform = true;
  long long seg_len = 0;
};

Sizes measure(const int64_t* off, int64_t nseg) {
  Sizes s;
  for (int64_t i = 0; i < nseg; ++i) {
    const long long len = off[i + 1] - off[i];
    s.total_keys += len;
    s.total_tiles += (len + kTile - 1) / kTile;
    if (i == 0) s.seg_len = len;
    else if (len != s.seg_len) s.uniform = false;
  }
  if (nseg > 0 && off[0] != 0) s.uniform = false;
  if (s.uniform && s.seg_len == 0) s.uniform = false;
  return s;
}

size_t scratch_for(const Sizes& s, int64_t nseg) {
  size_t b = align256(size_t(s.total_keys) * 8);                      // ping-pong keys
  b += align256(size_t(s.total_tiles) * kRadix * sizeof(unsigned));  // counters
  if (!s.uniform) {
    b += align256(size_t(s.total_tiles) * sizeof(TileInfo));
    b += align256(size_t(nseg) * sizeof(SegInfo));
    b += align256(size_t(nseg) * sizeof(SegDesc));
  }
  return b;
}

}  // namespace

extern "C" {

size_t rpx_sort_segmented_scratch_bytes(const int64_t* offsets_host, int64_t nseg) {
  if (!offsets_host || nseg < 0) return 0;
  return scratch_for(measure(offsets_host, nseg), nseg);
}

int rpx_sort_segmented_f64(double* x, const int64_t* offsets_host, int64_t nseg, void* scratch, size_t scratch_bytes,
                           void* stream) {
  if (!offsets_host || nseg < 0) return RPX_ERR_INVALID_ARGUMENT;
  for (int64_t i = 0; i < nseg; ++i) {
    const long long len = offsets_host[i + 1] - offsets_host[i];
    if (offsets_host[i] < 0 || len < 0 || len > INT32_MAX) return RPX_ERR_INVALID_ARGUMENT;
  }
  const Sizes s = measure(offsets_host, nseg);
  if (s.total_keys == 0) return RPX_SUCCESS;
  if (!x || !scratch) return RPX_ERR_INVALID_ARGUMENT;
  if (scratch_bytes < scratch_for(s, nseg)) return RPX_ERR_SCRATCH_TOO_SMALL;
  if (s.total_tiles > INT32_MAX || nseg > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  hipStream_t st = as_stream(stream);

  // Segments with no keys contribute no tiles; non-uniform layouts (and
  // segments not starting at x[0]) use an explicit tile table.
  char* base = reinterpret_cast<char*>(scratch);
  u64* alt = reinterpret_cast<u64*>(base);
  size_t off = align256(size_t(s.total_keys) * 8);
  unsigned* counts = reinterpret_cast<unsigned*>(base + off);
  off += align256(size_t(s.total_tiles) * kRadix * sizeof(unsigned));

  Plan plan{nullptr, nullptr, s.seg_len, int((s.seg_len + kTile - 1) / kTile)};
  int64_t nseg_launch = nseg;
  if (!s.uniform) {
    // Ragged layout: explicit tile and segment tables.  Segments keep their
    // positions in x; in the scratch copy they are packed back to back.  The
    // host describes each non-empty segment; a kernel expands the
    // descriptors into one TileInfo per tile.
    std::vector<SegDesc> desc;
    desc.reserve(size_t(nseg));
    long long cbase = 0, packed = 0;
    int first = 0;
    for (int64_t i = 0; i < nseg; ++i) {
      const long long len = offsets_host[i + 1] - offsets_host[i];
      const int nt = int((len + kTile - 1) / kTile);
      if (nt == 0) continue;
      desc.push_back(SegDesc{offsets_host[i], packed, cbase, first, nt, int(len), 0});
      cbase += (long long)nt * kRadix;
      packed += len;
      first += nt;
    }
    TileInfo* dt = reinterpret_cast<TileInfo*>(base + off);
    off += align256(size_t(s.total_tiles) * sizeof(TileInfo));
    SegInfo* ds = reinterpret_cast<SegInfo*>(base + off);
    off += align256(size_t(nseg) * sizeof(SegInfo));
    SegDesc* dd = reinterpret_cast<SegDesc*>(base + off);
    const HostChunk table{dd, desc.data(), desc.size() * sizeof(SegDesc)};
    if (const int rc = upload_async(&table, 1, st); rc != RPX_SUCCESS) return rc;
    tile_table_kernel<<<dim3(unsigned(ceil_div(s.total_tiles, 256))), dim3(256), 0, st>>>(
        dd, int(desc.size()), dt, ds, int(s.total_tiles));
    if (const int rc = launch_status(); rc != RPX_SUCCESS) return rc;
    plan.tiles = dt;
    plan.segs = ds;
    nseg_launch = int64_t(desc.size());
  }

  u64* xs = reinterpret_cast<u64*>(x);
  for (int pass = 0; pass < kPasses; ++pass) {
    const u64* src = (pass % 2 == 0) ? xs : alt;
    u64* dst = (pass % 2 == 0) ? alt : xs;
    const int shift = 8 * pass;
    const int src_is_x = (pass % 2 == 0);
    hist_kernel<<<dim3(unsigned(s.total_tiles)), dim3(kBS), 0, st>>>(src, counts, plan, shift, pass == 0, src_is_x);
    offsets_kernel<<<dim3(unsigned(nseg_launch)), dim3(kOffThreads), 0, st>>>(counts, plan);
    scatter_kernel<<<dim3(unsigned(s.total_tiles)), dim3(kBS), 0, st>>>(src, dst, counts, plan, shift, pass == 0,
                                                                        pass == kPasses - 1, src_is_x);
    const int rc = launch_status();
    if (rc != RPX_SUCCESS) return rc;
  }
  return RPX_SUCCESS;
}

}  // extern "C"
