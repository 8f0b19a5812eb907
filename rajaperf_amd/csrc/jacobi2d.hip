// JACOBI_2D: 5-point Jacobi sweeps on an N x N row-major FP64 grid.
//
// Reference: src/polybench/POLYBENCH_JACOBI_2D.hpp:41-45 (bodies),
// POLYBENCH_JACOBI_2D-Seq.cpp:30-54 (tsteps x two sweeps over the interior),
// POLYBENCH_JACOBI_2D-Hip.cpp:27-114 (one thread per point, 32 x bs/32
// blocks, no data reuse beyond the caches).
//
// MI355X design: a workgroup owns a tile of 512 columns x R rows.  Each
// work-item holds a column pair as one 16-byte vector; it issues the loads of
// all R+2 rows of its pair before any arithmetic (R+2 16-byte loads in flight
// per work-item), takes its west/east neighbours from the adjacent lanes by
// cross-lane shuffles (only the two edge lanes of a wavefront load a scalar
// halo), and stores the R output rows as 16-byte vectors.  Tiles are
// numbered so that each XCD walks a horizontal band of the grid top to
// bottom: the two halo rows a tile shares with its vertical neighbour are
// then read from the same XCD's L2.  Each output is
// 0.2*((((c + w) + e) + s) + n), the Base_Seq association, so sweeps are
// bit-identical to Base_Seq.
#include "rpx_common.hpp"


using namespace rpx;

namespace {

typedef double v2d __attribute__((ext_vector_type(2)));

constexpr int kBS = 256;             // work-items per tile row
constexpr int kCols = 2 * kBS;       // columns per tile

__device__ __forceinline__ double point(double c, double w, double e, double s, double n) {
  return 0.2 * ((((c + w) + e) + s) + n);
}

// XCD-aware tile id: hardware block b runs on XCD b % 8; XCD x walks a
// contiguous range of logical tiles.
__device__ __forceinline__ unsigned logical_tile(unsigned b, unsigned G) { return xcd_major(b, G); }

// lane l receives v from lane l-1; lane 0 receives 0.
__device__ __forceinline__ double lane_shr_zero(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
// lane l receives v from lane l+1; lane 63 receives 0.
__device__ __forceinline__ double lane_shl_zero(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

template <int R>
__global__ __launch_bounds__(kBS) void jacobi_vec_kernel(double* __restrict__ dst, const double* __restrict__ src,
                                                         int64_t N, int strips, int chunks) {
  const unsigned tile = logical_tile(blockIdx.x, gridDim.x);
  const int strip = int(tile % unsigned(strips));
  const int chunk = int(tile / unsigned(strips));
  const int lane = threadIdx.x & (kWave - 1);

  const int64_t pair = int64_t(strip) * kBS + threadIdx.x;  // column pair index
  const int64_t j0 = 2 * pair;                               // first column of the pair
  const bool in_grid = j0 < N;
  const int64_t r0 = 1 + int64_t(chunk) * R;                // first output row
  int64_t rend = r0 + R;
  if (rend > N - 1) rend = N - 1;
  const int rows = int(rend - r0);

  // Loads: rows r0-1 .. r0+R (clamped), the pair and the edge halos.
  v2d v[R + 2];
  double wh[R], eh[R];
#pragma unroll
  for (int k = 0; k < R + 2; ++k) {
    v[k] = v2d{0.0, 0.0};
    if (in_grid && k <= rows + 1) v[k] = *reinterpret_cast<const v2d*>(src + (r0 - 1 + k) * N + j0);
  }
  const bool left_edge = (lane == 0) && in_grid && j0 >= 1;
  const bool right_edge = (lane == kWave - 1) && in_grid && j0 + 2 < N;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    wh[k] = 0.0;
    eh[k] = 0.0;
    if (k < rows) {
      if (left_edge) wh[k] = src[(r0 + k) * N + j0 - 1];
      if (right_edge) eh[k] = src[(r0 + k) * N + j0 + 2];
    }
  }

  const bool w_ok = in_grid && j0 >= 1;                  // column j0 is interior
  const bool e_ok = in_grid && j0 + 1 <= N - 2;          // column j0+1 is interior
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const v2d c = v[k + 1];
    // West of j0 is the previous lane's .y; east of j0+1 is the next lane's .x.
    double w = lane_shr_zero(c.y);
    double e = lane_shl_zero(c.x);
    if (lane == 0) w = wh[k];
    if (lane == kWave - 1) e = eh[k];
    if (k < rows) {
      const int64_t row = r0 + k;
      const double o0 = point(c.x, w, c.y, v[k + 2].x, v[k].x);
      const double o1 = point(c.y, c.x, e, v[k + 2].y, v[k].y);
      double* out = dst + row * N + j0;
      if (w_ok && e_ok) {
        *reinterpret_cast<v2d*>(out) = v2d{o0, o1};
      } else {
        if (w_ok) out[0] = o0;
        if (e_ok) out[1] = o1;
      }
    }
  }
}

// General fallback (odd N or unaligned grid): one point per work-item.
__global__ __launch_bounds__(256) void jacobi_scalar_kernel(double* __restrict__ dst, const double* __restrict__ src,
                                                            int64_t N) {
  const int64_t inner = N - 2;
  const int64_t total = inner * inner;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += stride) {
    const int64_t i = 1 + t / inner, j = 1 + t % inner;
    dst[j + i * N] = point(src[j + i * N], src[j - 1 + i * N], src[j + 1 + i * N], src[j + (i + 1) * N],
                           src[j + (i - 1) * N]);
  }
}

// ---- fused time steps: T steps (2T sweeps) per launch -----------------------
//
// One wavefront owns a strip of 128 columns (its lanes hold column pairs) and
// a chunk of R output rows.  With H = 2T sweeps per launch the strip starts at
// j0 = 112*s - H and writes the 112 columns [j0+H, j0+H+112), whole 128-byte
// lines when rows start on one (writing 128-2H columns per strip instead, so
// that stores start mid-line, made a 50-step run at N = 8192 3.4% slower);
// the columns on either side are halo.  The wavefront loads R+2H rows of A (level 0)
// and runs a cascade of levels 1..2T: at step q, level L produces its row
// rbase+q-L from the last three rows of level L-1, held in a 3-row rolling
// register window.  West/east neighbours come from the adjacent lanes.  Odd
// levels are B's of the time steps, even levels the new A's; on the grid
// boundary a level takes B's boundary values (odd levels) or A's (even
// levels, which never change), exactly as the two-sweep update reads them.
// Level 2T-1 (the final B) and level 2T (the final A) are written for the
// owned rows and columns only, each by exactly one wavefront; the new A goes
// to a different grid than the one read, so no wavefront reads a value
// another one has already replaced.
//
// Per wavefront and (step, level) the interior path issues 8 FP64 adds, 2
// multiplies and 4 DPP moves and nothing else: row offsets are scalar (the
// wavefront index is made uniform with readfirstlane) and the wavefronts
// that touch the grid boundary take a separate path.  Measured on MI355X at
// N = 8192, a launch of depth 3 moves the ideal 1.08 GB (read A, write the
// new A); R = 16 rows per wavefront is the fastest tile of this kernel, and
// the streaming kernel below (chunks of 99 rows) is about 5% faster again.
constexpr int kStripCols = 128;
// Output columns per strip, at every depth: 112 columns are 7 whole 128-byte
// lines, and strip s writes columns [112 s, 112 s + 112), so where rows start
// on a line (N a multiple of 16) every store covers whole lines and no line
// is written by two strips.  A strip loads columns [112 s - H, 112 s - H + 128),
// which leaves at least H halo columns on both sides for H = 2T <= 8.
constexpr int kOwnCols = 112;


typedef unsigned int v4u __attribute__((ext_vector_type(4)));
constexpr int kAhead = 6;  // level-0 rows loaded ahead of their use
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 raw buffer: untyped dword access
constexpr int kStoreNT = 2;              // buffer cache policy: nontemporal (the next launch reads it, not this one)

// The cascade of one wavefront.  EDGE = false is the interior case: the
// wavefront's R+2H rows and 128 columns lie strictly inside the grid, so no
// level row is a boundary row and no lane holds a boundary column.  The two
// edge lanes then take 0 for their missing west/east neighbour; that error
// moves one column inward per level and stays inside the H halo columns.
// Interior rows are addressed through buffer resources based at the
// wavefront's first row and column: each access is one SGPR row offset plus
// the lane's 16-byte offset.  EDGE = true handles any wavefront: loads
// outside the grid give 0, and each level row re-reads the boundary values
// it must keep (B's at odd levels, A's -- unchanged in Ain -- at even ones),
// as the two-sweep update leaves them.  Level-0 rows stream in kAhead rows
// ahead of their use; the empty asm statement keeps the compiler from
// hoisting every load to the top, which would hold all R+4T rows in
// registers at once.
template <int R, int T, bool WRITE_B, bool EDGE>
__device__ __forceinline__ void fused_tile(double* __restrict__ Aout, double* __restrict__ B,
                                           const double* __restrict__ Ain, int64_t N, int64_t rbase, int64_t j0,
                                           int rows, int lane) {
  constexpr int H = 2 * T;       // sweeps per launch = halo width
  constexpr int NR = R + 2 * H;  // level-0 rows
  const int pitch = int(N) * int(sizeof(double));
  const int64_t origin = rbase * N + j0;
  const int span = NR * pitch;
  const int voff = lane * 16;
  const int64_t jl = j0 + 2 * lane;       // first column of this lane's pair
  const bool col_in = jl >= 0 && jl < N;  // N even and jl even: jl+1 < N too
  static_assert(2 * H <= kStripCols - kOwnCols, "halo columns on both sides of the owned ones");
  const bool own = lane >= T && lane < T + kOwnCols / 2;  // columns [j0+H, j0+H+112)
  const bool x_out = own && jl >= 1 && jl <= N - 2;
  const bool y_out = own && jl + 1 >= 1 && jl + 1 <= N - 2;

  auto load_row = [&](int k) -> v2d {
    if constexpr (!EDGE) {
      const auto in = __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(Ain + origin), 0, span, kRsrcWord3);
      return __builtin_bit_cast(v2d, __builtin_amdgcn_raw_buffer_load_b128(in, voff, k * pitch, 0));
    } else {
      const int64_t row = rbase + k;
      v2d v{0.0, 0.0};
      if (col_in && row >= 0 && row < N && k < rows + 2 * H) v = *reinterpret_cast<const v2d*>(Ain + row * N + jl);
      return v;
    }
  };

  v2d a[NR];
#pragma unroll
  for (int k = 0; k < NR && k <= kAhead + 1; ++k) a[k] = load_row(k);

  v2d win[H][3];  // win[L-1]: the last three rows of level L
#pragma unroll
  for (int q = 2; q < NR; ++q) {
    asm volatile("" ::: "memory");
    if (q + kAhead < NR) a[q + kAhead] = load_row(q + kAhead);
#pragma unroll
    for (int L = 1; L <= H; ++L) {
      if (q < 2 * L) continue;  // level L starts once level L-1 has three rows
      const v2d up = (L == 1) ? a[q - 2] : win[L - 2][0];
      const v2d c = (L == 1) ? a[q - 1] : win[L - 2][1];
      const v2d dn = (L == 1) ? a[q] : win[L - 2][2];
      const double w = lane_shr_zero(c.y);
      const double e = lane_shl_zero(c.x);
      v2d r;
      r.x = point(c.x, w, c.y, dn.x, up.x);
      r.y = point(c.y, c.x, e, dn.y, up.y);
      const int k = q - L;  // row of this level row, relative to rbase
      if constexpr (EDGE) {
        const int64_t g = rbase + k;
        const double* keep = (L % 2 == 1) ? B : Ain;  // the boundary this level keeps
        if (g == 0 || g == N - 1) {
          if (col_in) r = *reinterpret_cast<const v2d*>(keep + g * N + jl);
        } else if (g > 0 && g < N - 1) {
          if (jl == 0) r.x = keep[g * N];
          if (jl + 1 == N - 1) r.y = keep[g * N + N - 1];
        }
      }
      win[L - 1][0] = win[L - 1][1];
      win[L - 1][1] = win[L - 1][2];
      win[L - 1][2] = r;
      if ((L == H || (WRITE_B && L == H - 1)) && k >= H && k < H + R) {
        if constexpr (!EDGE) {
          const auto out =
              __builtin_amdgcn_make_buffer_rsrc((L == H ? Aout : B) + origin, 0, span, kRsrcWord3);
          if (own) __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, r), out, voff, k * pitch, kStoreNT);
        } else if (k < H + rows) {
          double* o = (L == H ? Aout : B) + (rbase + k) * N + jl;
          if (x_out && y_out) {
            *reinterpret_cast<v2d*>(o) = r;
          } else {
            if (x_out) o[0] = r.x;
            if (y_out) o[1] = r.y;
          }
        }
      }
    }
  }
}

template <int R, int T, bool WRITE_B>
__global__ __launch_bounds__(256) void jacobi_fused_kernel(double* __restrict__ Aout, double* __restrict__ B,
                                                           const double* __restrict__ Ain, int64_t N, int strips,
                                                           int chunks) {
  constexpr int H = 2 * T;
  constexpr int NR = R + 2 * H;
  const int lane = threadIdx.x & (kWave - 1);
  // XCD-banded workgroup order (see logical_tile): each XCD walks a band of
  // row chunks, so the halo rows shared by vertically adjacent chunks are
  // re-read from the same L2.
  // readfirstlane: the wavefront index is uniform, and saying so keeps all
  // row/column arithmetic below on the scalar unit.
  const unsigned wave_global = __builtin_amdgcn_readfirstlane(
      logical_tile(blockIdx.x, gridDim.x) * (blockDim.x / kWave) + (threadIdx.x / kWave));
  const unsigned total = unsigned(strips) * unsigned(chunks);
  if (wave_global >= total) return;
  const int strip = int(wave_global % unsigned(strips));  // a workgroup spans 4 adjacent strips
  const int chunk = int(wave_global / unsigned(strips));
  const int64_t j0 = int64_t(strip) * kOwnCols - H;
  const int64_t r0 = 1 + int64_t(chunk) * R;
  int64_t rend = r0 + R;
  if (rend > N - 1) rend = N - 1;
  const int rows = int(rend - r0);
  const int64_t rbase = r0 - H;  // global row of level-0 index 0
  const bool interior = rows == R && rbase >= 0 && rbase + NR <= N && j0 >= 1 && j0 + kStripCols <= N - 1;
  if (interior)
    fused_tile<R, T, WRITE_B, false>(Aout, B, Ain, N, rbase, j0, rows, lane);
  else
    fused_tile<R, T, WRITE_B, true>(Aout, B, Ain, N, rbase, j0, rows, lane);
}

// ---- fused time steps over tall chunks -------------------------------------
//
// The same cascade as fused_tile, for chunks of many rows (R >= 48): a
// wavefront runs down its chunk in a loop, so the 2H halo rows above and
// below are loaded and recomputed once per chunk instead of once per 16 rows.
// Level-0 rows sit in a ring of kRing registers (rows q-2 .. q+kAhead at
// step q) and level L's last three rows in win[L-1][row % 3]; the loop body
// is a block of kBlk steps starting at a multiple of kBlk, so every ring and
// window slot is a compile-time index.  Steps 0 .. kWarm-1 (the levels
// starting up) have compile-time step numbers.  A block runs the interior
// form (no boundary code, buffer-resource addressing) when every row it
// loads, computes or stores lies inside the grid and its stores stay inside
// the chunk; otherwise, and for the chunk's last block, it runs the edge
// form of fused_tile.  Each output is computed by the same operations in the
// same order as in fused_tile, so the results are bit-identical.
constexpr int kRing = kAhead + 3;
constexpr int kBlk = 9;
constexpr int kWarm = 2 * kBlk;
constexpr int kMinStreamRows = 48;
constexpr int kMaxStreamRows = 4096;
static_assert(kBlk % kRing == 0 && kBlk % 3 == 0, "block steps keep the ring and window slots in place");

struct StreamGeo {
  double* Aout;
  double* B;
  const double* Ain;
  int64_t N, rbase, j0, jl;
  int rows, NR, pitch, voff;
  bool col_in, own, x_out, y_out;
};

// Step q = QB + U of the cascade (QB = q0, a multiple of kBlk; WARM: q0 = 0
// and the step number is the compile-time U).  Interior accesses are
// relative to level-0 row `base` (base = 0 in the warm-up, q0 - H after).
template <int T, bool WRITE_B, bool EDGE, bool WARM, int U>
__device__ __forceinline__ void stream_step(const StreamGeo& g, int q0, v2d (&a)[kRing], v2d (&win)[2 * T][3]) {
  constexpr int H = 2 * T;
  const int q = WARM ? U : q0 + U;
  const int base = WARM ? 0 : q0 - H;
  const int64_t origin = (g.rbase + base) * g.N + g.j0;
  const int span = (kWarm + kAhead + H + 2) * g.pitch;
  {
    const int k = q + kAhead;  // level-0 row into the slot of row q - 3
    v2d v{0.0, 0.0};
    if constexpr (!EDGE) {
      const auto in = __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(g.Ain + origin), 0, span, kRsrcWord3);
      v = __builtin_bit_cast(v2d, __builtin_amdgcn_raw_buffer_load_b128(in, g.voff, (k - base) * g.pitch, 0));
    } else {
      const int64_t row = g.rbase + k;
      if (g.col_in && row >= 0 && row < g.N && k < g.NR) v = *reinterpret_cast<const v2d*>(g.Ain + row * g.N + g.jl);
    }
    a[(U + kAhead) % kRing] = v;
  }
#pragma unroll
  for (int L = 1; L <= H; ++L) {
    if (WARM && U < 2 * L) continue;  // level L starts once level L-1 has three rows
    v2d up, c, dn;
    if (L == 1) {
      up = a[(U + kRing - 2) % kRing];
      c = a[(U + kRing - 1) % kRing];
      dn = a[U % kRing];
    } else {
      up = win[L - 2][(U - L - 1 + 3 * kBlk) % 3];
      c = win[L - 2][(U - L + 3 * kBlk) % 3];
      dn = win[L - 2][(U - L + 1 + 3 * kBlk) % 3];
    }
    const double w = lane_shr_zero(c.y);
    const double e = lane_shl_zero(c.x);
    v2d r;
    r.x = point(c.x, w, c.y, dn.x, up.x);
    r.y = point(c.y, c.x, e, dn.y, up.y);
    const int k = q - L;  // row of this level row, relative to rbase
    if constexpr (EDGE) {
      const int64_t gr = g.rbase + k;
      const double* keep = (L % 2 == 1) ? g.B : g.Ain;  // the boundary this level keeps
      if (gr == 0 || gr == g.N - 1) {
        if (g.col_in) r = *reinterpret_cast<const v2d*>(keep + gr * g.N + g.jl);
      } else if (gr > 0 && gr < g.N - 1) {
        if (g.jl == 0) r.x = keep[gr * g.N];
        if (g.jl + 1 == g.N - 1) r.y = keep[gr * g.N + g.N - 1];
      }
    }
    win[L - 1][(U - L + 3 * kBlk) % 3] = r;
    if (L == H || (WRITE_B && L == H - 1)) {
      if constexpr (!EDGE) {
        // Interior blocks store only rows inside the chunk (checked by the caller).
        if (!WARM || U - L >= H) {
          const auto out =
              __builtin_amdgcn_make_buffer_rsrc((L == H ? g.Aout : g.B) + origin, 0, span, kRsrcWord3);
          if (g.own)
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, r), out, g.voff, (k - base) * g.pitch,
                                                   kStoreNT);
        }
      } else if (k >= H && k < H + g.rows) {
        double* o = (L == H ? g.Aout : g.B) + (g.rbase + k) * g.N + g.jl;
        if (g.x_out && g.y_out) {
          *reinterpret_cast<v2d*>(o) = r;
        } else {
          if (g.x_out) o[0] = r.x;
          if (g.y_out) o[1] = r.y;
        }
      }
    }
  }
}

template <int T, bool WRITE_B, bool EDGE, bool WARM, int U0, int U1>
__device__ __forceinline__ void stream_steps(const StreamGeo& g, int q0, v2d (&a)[kRing], v2d (&win)[2 * T][3]) {
  if constexpr (U0 < U1) {
    asm volatile("" ::: "memory");
    stream_step<T, WRITE_B, EDGE, WARM, U0>(g, q0, a, win);
    stream_steps<T, WRITE_B, EDGE, WARM, U0 + 1, U1>(g, q0, a, win);
  }
}

template <int T, bool EDGE>
__device__ __forceinline__ void stream_first_loads(const StreamGeo& g, v2d (&a)[kRing]) {
#pragma unroll
  for (int k = 0; k <= kAhead + 1; ++k) {
    v2d v{0.0, 0.0};
    if constexpr (!EDGE) {
      const auto in = __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(g.Ain + g.rbase * g.N + g.j0), 0,
                                                        (kWarm + kAhead + 2 * T + 2) * g.pitch, kRsrcWord3);
      v = __builtin_bit_cast(v2d, __builtin_amdgcn_raw_buffer_load_b128(in, g.voff, k * g.pitch, 0));
    } else {
      const int64_t row = g.rbase + k;
      if (g.col_in && row >= 0 && row < g.N && k < g.NR) v = *reinterpret_cast<const v2d*>(g.Ain + row * g.N + g.jl);
    }
    a[k % kRing] = v;
  }
}

// The first and last strips (the grid's west and east columns) run the edge
// form in every block, which waits on each boundary-value load; they take
// chunks of kEdgeRows rows and the first workgroups of the launch, which the
// dispatcher spreads over all XCDs, so they finish early.  The other strips
// take chunks of R rows in the XCD-banded order.
constexpr int kEdgeRows = 24;

struct StreamGrid {
  int strips, chunks, R;      // tall chunks of the inner strips 1 .. strips-2
  int edge_strips, edge_chunks, edge_blocks;
};

template <int T, bool WRITE_B>
__global__ __launch_bounds__(256) void jacobi_stream_kernel(double* __restrict__ Aout, double* __restrict__ B,
                                                            const double* __restrict__ Ain, int64_t N,
                                                            const StreamGrid sg) {
  constexpr int H = 2 * T;
  static_assert(2 * H <= kStripCols - kOwnCols, "halo columns on both sides of the owned ones");
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned wave = threadIdx.x / kWave;
  int strip, chunk, R;
  if (blockIdx.x < unsigned(sg.edge_blocks)) {
    const unsigned e = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + wave);
    if (e >= unsigned(sg.edge_strips) * unsigned(sg.edge_chunks)) return;
    const int side = int(e % unsigned(sg.edge_strips));
    strip = side == 0 ? 0 : sg.strips - 1;
    chunk = int(e / unsigned(sg.edge_strips));
    R = kEdgeRows;
  } else {
    const unsigned inner = unsigned(sg.strips - sg.edge_strips);
    const unsigned i = __builtin_amdgcn_readfirstlane(
        logical_tile(blockIdx.x - sg.edge_blocks, gridDim.x - sg.edge_blocks) * (blockDim.x / kWave) + wave);
    if (i >= inner * unsigned(sg.chunks)) return;
    strip = 1 + int(i % inner);
    chunk = int(i / inner);
    R = sg.R;
  }
  StreamGeo g;
  g.Aout = Aout;
  g.B = B;
  g.Ain = Ain;
  g.N = N;
  g.j0 = int64_t(strip) * kOwnCols - H;
  const int64_t r0 = 1 + int64_t(chunk) * R;
  int64_t rend = r0 + R;
  if (rend > N - 1) rend = N - 1;
  g.rows = int(rend - r0);
  g.rbase = r0 - H;
  g.NR = g.rows + 2 * H;
  g.pitch = int(N) * int(sizeof(double));
  g.voff = lane * 16;
  g.jl = g.j0 + 2 * lane;
  g.col_in = g.jl >= 0 && g.jl < N;
  g.own = lane >= T && lane < T + kOwnCols / 2;
  g.x_out = g.own && g.jl >= 1 && g.jl <= N - 2;
  g.y_out = g.own && g.jl + 1 >= 1 && g.jl + 1 <= N - 2;
  const bool strip_in = g.j0 >= 1 && g.j0 + kStripCols <= N - 1;

  v2d a[kRing];
  v2d win[H][3];
  // Warm-up, steps 2 .. kWarm-1: interior when its rows (loads up to row
  // kWarm-1+kAhead, stores up to row kWarm-H) lie inside the grid and the chunk.
  if (strip_in && g.rbase >= 0 && g.rbase + kWarm + kAhead <= N - 1 && g.NR > kWarm + 1) {
    stream_first_loads<T, false>(g, a);
    stream_steps<T, WRITE_B, false, true, 2, kWarm>(g, 0, a, win);
  } else {
    stream_first_loads<T, true>(g, a);
    stream_steps<T, WRITE_B, true, true, 2, kWarm>(g, 0, a, win);
  }
  // A block runs the interior form when its stores (A rows up to
  // q0+kBlk-1-H, B rows one further) stay inside the chunk and its rows lie
  // inside the grid.  Three loops, each with one form, so that the loop that
  // carries the interior blocks sees only its own loads and stores in flight
  // (the compiler's waits at a loop edge that joins both forms drained every
  // block).
  auto rows_in = [&](int q) {
    return q + kBlk + 1 <= g.NR && g.rbase + q - H >= 1 && g.rbase + q + kBlk - 1 + kAhead <= N - 1;
  };
  int q0 = kWarm;
  if (strip_in) {
    while (q0 + kBlk + 1 <= g.NR && !rows_in(q0)) {
      stream_steps<T, WRITE_B, true, false, 0, kBlk>(g, q0, a, win);
      q0 += kBlk;
    }
    while (rows_in(q0)) {
      stream_steps<T, WRITE_B, false, false, 0, kBlk>(g, q0, a, win);
      q0 += kBlk;
    }
  }
  while (q0 < g.NR) {
    stream_steps<T, WRITE_B, true, false, 0, kBlk>(g, q0, a, win);
    q0 += kBlk;
  }
}

// Copies the boundary rows and columns of src into dst.
__global__ __launch_bounds__(256) void boundary_copy_kernel(double* __restrict__ dst, const double* __restrict__ src,
                                                            int64_t N) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < N) {
    dst[t] = src[t];                              // row 0
    dst[(N - 1) * N + t] = src[(N - 1) * N + t];  // row N-1
    dst[t * N] = src[t * N];                      // column 0
    dst[t * N + N - 1] = src[t * N + N - 1];      // column N-1
  }
}

template <int R, int T>
int launch_fused(double* Aout, double* B, const double* Ain, int64_t N, bool write_b, hipStream_t s) {
  const int64_t strips = ceil_div(N - 1, kOwnCols);  // output columns [0, N-1)
  const int64_t chunks = ceil_div(N - 2, R);
  const int64_t grid = ceil_div(strips * chunks, 4);
  if (grid > INT32_MAX || strips > INT32_MAX || chunks > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  if (write_b)
    jacobi_fused_kernel<R, T, true><<<dim3(unsigned(grid)), dim3(256), 0, s>>>(Aout, B, Ain, N, int(strips),
                                                                               int(chunks));
  else
    jacobi_fused_kernel<R, T, false><<<dim3(unsigned(grid)), dim3(256), 0, s>>>(Aout, B, Ain, N, int(strips),
                                                                                int(chunks));
  return launch_status();
}

template <int T>
int launch_stream(double* Aout, double* B, const double* Ain, int64_t N, int rows, bool write_b, hipStream_t s) {
  const int64_t strips = ceil_div(N - 1, kOwnCols);
  const int64_t chunks = ceil_div(N - 2, rows);
  const int64_t edge_strips = strips < 2 ? strips : 2;
  const int64_t edge_chunks = ceil_div(N - 2, kEdgeRows);
  const int64_t edge_blocks = ceil_div(edge_strips * edge_chunks, 4);
  const int64_t grid = edge_blocks + ceil_div((strips - edge_strips) * chunks, 4);
  if (grid > INT32_MAX || strips > INT32_MAX || edge_chunks > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  const StreamGrid sg{int(strips), int(chunks), rows, int(edge_strips), int(edge_chunks), int(edge_blocks)};
  if (write_b)
    jacobi_stream_kernel<T, true><<<dim3(unsigned(grid)), dim3(256), 0, s>>>(Aout, B, Ain, N, sg);
  else
    jacobi_stream_kernel<T, false><<<dim3(unsigned(grid)), dim3(256), 0, s>>>(Aout, B, Ain, N, sg);
  return launch_status();
}

template <int T>
int fused_rows(double* Aout, double* B, const double* Ain, int64_t N, int rows, bool write_b, hipStream_t s) {
  if (rows >= kMinStreamRows) return launch_stream<T>(Aout, B, Ain, N, rows, write_b, s);
  switch (rows) {
    case 4: return launch_fused<4, T>(Aout, B, Ain, N, write_b, s);
    case 8: return launch_fused<8, T>(Aout, B, Ain, N, write_b, s);
    case 12: return launch_fused<12, T>(Aout, B, Ain, N, write_b, s);
    case 16: return launch_fused<16, T>(Aout, B, Ain, N, write_b, s);
    case 32:  // deeper 32-row tiles would not fit in registers
      if constexpr (T == 1) return launch_fused<32, 1>(Aout, B, Ain, N, write_b, s);
      return RPX_ERR_UNSUPPORTED;
    default: return RPX_ERR_UNSUPPORTED;
  }
}

// `depth` time steps in one launch, reading Ain and writing B and Aout.
// write_b = false leaves B's interior unwritten.
int fused_launch(double* Aout, double* B, const double* Ain, int64_t N, int rows, int depth, bool write_b,
                 hipStream_t s) {
  switch (depth) {
    case 1: return fused_rows<1>(Aout, B, Ain, N, rows, write_b, s);
    case 2: return fused_rows<2>(Aout, B, Ain, N, rows, write_b, s);
    case 3: return fused_rows<3>(Aout, B, Ain, N, rows, write_b, s);
    case 4:  // streaming chunks only: the 16-row tile at depth 4 was slower than depth 3
      return rows >= kMinStreamRows ? launch_stream<4>(Aout, B, Ain, N, rows, write_b, s) : RPX_ERR_UNSUPPORTED;
    default: return RPX_ERR_UNSUPPORTED;
  }
}

constexpr int kDefaultRows = 16;
// Default rows per wavefront: streaming chunks of 99 rows for the launches
// that store no B on grids of N >= 3072, 16-row tiles otherwise.  With B's
// stores as well the tile kernel is faster (0.321 against 0.337 ms per
// depth-2 launch at N = 8192), and on smaller grids tall chunks leave too few
// wavefronts (depth 3 at N = 1026 and 2050: 0.027 and 0.035 ms with tiles,
// 0.039 and 0.046 ms at best with chunks; at N = 4098: 0.069 against
// 0.059 ms).  An explicit `rows` applies to every launch.
// Four time steps per launch (tune->depth = 4: streaming chunks only, 144
// rows by default) are faster still per time step, 0.060 against 0.073 ms
// at N = 8192; the default depth stays 3.
constexpr int kDefaultFusedRows = 99;
constexpr int kDefaultFusedRowsDepth4 = 144;
constexpr int kDefaultFusedRowsTile = 16;
constexpr int64_t kMinStreamN = 3072;
inline int fused_rows_for(int rows_req, bool write_b, int64_t N, int depth) {
  if (rows_req) return rows_req;
  if (depth == 4) return kDefaultFusedRowsDepth4;
  return (write_b || N < kMinStreamN) ? kDefaultFusedRowsTile : kDefaultFusedRows;
}
inline int default_depth(int64_t) { return 3; }
constexpr int kMaxDepth = 3;  // tiles; streaming chunks also run depth 4
bool fused_shape_ok(int rows, int depth) {
  if (depth == 4) return rows >= kMinStreamRows && rows <= kMaxStreamRows;
  if (depth < 1 || depth > kMaxDepth) return false;
  return rows == 4 || rows == 8 || rows == 12 || rows == 16 || (rows == 32 && depth == 1) ||
         (rows >= kMinStreamRows && rows <= kMaxStreamRows);
}

// Buffer offsets within a wavefront's rows are 32-bit: (R + 4T) * N * 8
// bytes must stay below 2^31.
constexpr int64_t kMaxFusedN = int64_t(1) << 22;

template <int R>
int launch_vec(double* dst, const double* src, int64_t N, hipStream_t s) {
  const int64_t strips = ceil_div(N, kCols);
  const int64_t chunks = ceil_div(N - 2, R);
  const int64_t grid = strips * chunks;
  if (grid > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  jacobi_vec_kernel<R><<<dim3(unsigned(grid)), dim3(kBS), 0, s>>>(dst, src, N, int(strips), int(chunks));
  return launch_status();
}

int sweep(double* dst, const double* src, int64_t N, const rpx_tuning* tune, hipStream_t s) {
  if (N < 3) return RPX_SUCCESS;  // no interior
  const bool vec = (N % 2 == 0) && aligned16(dst) && aligned16(src);
  if (!vec) {
    int64_t grid = ceil_div((N - 2) * (N - 2), 256);
    if (grid > 65536) grid = 65536;
    jacobi_scalar_kernel<<<dim3(unsigned(grid)), dim3(256), 0, s>>>(dst, src, N);
    return launch_status();
  }
  switch (tune_or(tune, &rpx_tuning::items_per_thread, kDefaultRows)) {
    case 4: return launch_vec<4>(dst, src, N, s);
    case 8: return launch_vec<8>(dst, src, N, s);
    case 16: return launch_vec<16>(dst, src, N, s);
    case 32: return launch_vec<32>(dst, src, N, s);
    default: return RPX_ERR_UNSUPPORTED;
  }
}

}  // namespace

extern "C" {

int rpx_jacobi2d_sweep_f64(double* dst, const double* src, int64_t N, const rpx_tuning* tune, void* stream) {
  if (N < 0 || (N > 0 && (!dst || !src)) || dst == src) return RPX_ERR_INVALID_ARGUMENT;
  return sweep(dst, src, N, tune, as_stream(stream));
}

size_t rpx_jacobi2d_fused_scratch_bytes(int64_t N) { return N > 0 ? size_t(N) * size_t(N) * sizeof(double) : 0; }

int rpx_jacobi2d_fused_launch_f64(double* Aout, double* B, const double* Ain, int64_t N, int depth,
                                  const rpx_tuning* tune, void* stream) {
  if (N < 0 || !Aout || !B || !Ain || Aout == Ain || Aout == B || B == Ain) return RPX_ERR_INVALID_ARGUMENT;
  if (N % 2 != 0 || N > kMaxFusedN || !aligned16(Aout) || !aligned16(B) || !aligned16(Ain))
    return RPX_ERR_UNSUPPORTED;
  if (N < 3) return RPX_SUCCESS;
  const bool write_b = (tune_flags(tune) & RPX_TUNE_JACOBI_SKIP_B) == 0;
  const int rows = fused_rows_for(tune_or(tune, &rpx_tuning::items_per_thread, 0), write_b, N, depth);
  if (!fused_shape_ok(rows, depth))
    return RPX_ERR_UNSUPPORTED;
  return fused_launch(Aout, B, Ain, N, rows, depth, write_b, as_stream(stream));
}

int rpx_jacobi2d_copy_boundary_f64(double* dst, const double* src, int64_t N, void* stream) {
  if (N < 0 || (N > 0 && (!dst || !src))) return RPX_ERR_INVALID_ARGUMENT;
  if (N == 0) return RPX_SUCCESS;
  boundary_copy_kernel<<<dim3(unsigned(ceil_div(N, 256))), dim3(256), 0, as_stream(stream)>>>(dst, src, N);
  return launch_status();
}

int rpx_jacobi2d_fused_f64(double* A, double* B, int64_t N, int64_t tsteps, void* scratch, size_t scratch_bytes,
                           const rpx_tuning* tune, void* stream) {
  if (N < 0 || tsteps < 0 || (N > 0 && (!A || !B)) || A == B) return RPX_ERR_INVALID_ARGUMENT;
  if (N < 3 || tsteps == 0) return RPX_SUCCESS;
  hipStream_t s = as_stream(stream);
  const bool vec = (N % 2 == 0) && N <= kMaxFusedN && aligned16(A) && aligned16(B) && aligned16(scratch);
  if (!vec) return rpx_jacobi2d_f64(A, B, N, tsteps, tune, stream);
  if (!scratch || scratch == A || scratch == B) return RPX_ERR_INVALID_ARGUMENT;
  if (scratch_bytes < rpx_jacobi2d_fused_scratch_bytes(N)) return RPX_ERR_SCRATCH_TOO_SMALL;
  const int rows_req = tune_or(tune, &rpx_tuning::items_per_thread, 0);
  const int depth = tune_or(tune, &rpx_tuning::depth, default_depth(N));
  if (!fused_shape_ok(fused_rows_for(rows_req, false, N, depth), depth) ||
      !fused_shape_ok(fused_rows_for(rows_req, true, N, depth), depth))
    return RPX_ERR_UNSUPPORTED;

  // Launch plan: depths summing to tsteps, an even number of launches so the
  // A <-> scratch ping-pong ends in A.  A lone single step runs as the
  // two in-place sweeps instead.
  int64_t nfull = tsteps / depth;
  int rem = int(tsteps % depth);
  int64_t t_done = 0;
  int first_split[2] = {0, 0};  // an extra pair of launches taken from one full or remainder launch
  int64_t count = nfull + (rem > 0 ? 1 : 0);
  if (count % 2 == 1) {
    if (rem >= 2) {  // split the remainder
      first_split[0] = rem - 1;
      first_split[1] = 1;
      rem = 0;
    } else if (nfull >= 1 && depth >= 2) {  // split one full launch
      first_split[0] = depth - 1;
      first_split[1] = 1;
      --nfull;
    } else {  // one single step left over: two in-place sweeps
      int rc = sweep(B, A, N, nullptr, s);
      if (rc == RPX_SUCCESS) rc = sweep(A, B, N, nullptr, s);
      if (rc != RPX_SUCCESS) return rc;
      t_done = 1;
      if (rem == 1) rem = 0;
      else --nfull;
    }
  }
  if (t_done == tsteps) return RPX_SUCCESS;
  double* Q = reinterpret_cast<double*>(scratch);
  boundary_copy_kernel<<<dim3(unsigned(ceil_div(N, 256))), dim3(256), 0, s>>>(Q, A, N);
  int rc = launch_status();
  double* src = A;
  double* dst = Q;
  // Only the last launch stores B: each launch's B is overwritten by the
  // next before anything reads it (a launch reads only B's boundary, which
  // never changes).
  int64_t left = (first_split[0] > 0) + (first_split[1] > 0) + nfull + (rem > 0);
  auto go = [&](int d) {
    if (rc != RPX_SUCCESS || d == 0) return;
    const bool last = --left == 0;
    rc = fused_launch(dst, B, src, N, fused_rows_for(rows_req, last, N, d), d, last, s);
    double* t = src;
    src = dst;
    dst = t;
  };
  go(first_split[0]);
  go(first_split[1]);
  for (int64_t i = 0; i < nfull; ++i) go(depth);
  go(rem);
  return rc;
}

int rpx_jacobi2d_f64(double* A, double* B, int64_t N, int64_t tsteps, const rpx_tuning* tune, void* stream) {
  if (N < 0 || tsteps < 0 || (N > 0 && (!A || !B)) || A == B) return RPX_ERR_INVALID_ARGUMENT;
  hipStream_t s = as_stream(stream);
  for (int64_t t = 0; t < tsteps; ++t) {
    int rc = sweep(B, A, N, tune, s);
    if (rc != RPX_SUCCESS) return rc;
    rc = sweep(A, B, N, tune, s);
    if (rc != RPX_SUCCESS) return rc;
  }
  return RPX_SUCCESS;
}

}  // extern "C"
