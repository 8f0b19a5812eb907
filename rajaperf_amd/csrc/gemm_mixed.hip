// Mixed-precision GEMM with FP64 accuracy from int8 matrix cores.
//
// Same result semantics as the suite's GEMM (src/polybench/POLYBENCH_GEMM.hpp
// :12-50: C = alpha*A*B, beta never reaches C).  The products run on
// v_mfma_i32_16x16x64_i8 over error-free fixed-point slices of A and B
// (the Ozaki splitting):
//   * every row i of A is scaled by 2^-e_i (e_i from frexp of the row's max
//     |A|), every column j of B by 2^-f_j, so all entries lie in (-1, 1);
//   * each scaled entry x is cut into s signed 7-bit digits,
//       x = sum_{p<s} d_p 2^{-7(p+1)} + r,  |d_p| <= 127,  |r| < 2^{-7s},
//     by exact FP64 operations (scale by 128, truncate, subtract);
//   * S_t = sum_{p+q=t} sum_k dA_p[i][k] dB_q[k][j] is an exact int32 sum for
//     every t < s: level t has at most s digit pairs, so |S_t| <= s*127^2*nk,
//     below 2^31 for nk <= (2^31-1)/(s*127^2) (19020 at s = 7, 16643 at 8);
//   * C[i][j] = alpha * 2^(e_i+f_j) * sum_{t<s} S_t 2^{-7(t+2)}, summed in
//     FP64 from the smallest term up.
// The dropped digit products are below (s+3) 2^{-7s} relative to
// 2^(e_i+f_j) per k term; with the default s = 7 the error is below 1e-12
// relative to |alpha|(|A||B|)_ij.  There are s(s+1)/2 slice products.
//
// Scratch holds the int8 slices (A: [p][i][k], B: [q][j][k], K contiguous,
// zero-padded to tile multiples), the per-row/column exponents and the
// int32 partial sums S_t ([t][i][j] over the padded grid).
#include "rpx_common.hpp"

using namespace rpx;

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kMaxSlices = 8;
constexpr int kTM = 256, kTN = 256, kTK = 128;  // block tile, K bytes per stage
constexpr int kChunks = kTK / 16;               // 16-byte chunks per tile row and stage
// Largest nk for which every S_t stays inside int32 with s slices.
constexpr int64_t max_k(int s) { return int64_t(INT32_MAX) / (int64_t(s) * 127 * 127); }

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct MixedLayout {
  int64_t nip, njp, nkp;
  size_t a_off, b_off, ea_off, fb_off, s_off, total;
};
MixedLayout mixed_layout(int64_t ni, int64_t nj, int64_t nk, int s) {
  MixedLayout l;
  l.nip = round_up(ni, kTM);
  l.njp = round_up(nj, kTN);
  l.nkp = round_up(nk > 0 ? nk : 1, kTK);
  l.a_off = 0;
  l.b_off = align256(size_t(s) * l.nip * l.nkp);
  l.ea_off = l.b_off + align256(size_t(s) * l.njp * l.nkp);
  l.fb_off = l.ea_off + align256(size_t(l.nip) * 4);
  l.s_off = l.fb_off + align256(size_t(l.njp) * 4);
  l.total = l.s_off + align256(size_t(s) * l.nip * l.njp * 4);  // S_t, int32, [t][i][j]
  return l;
}

__device__ __forceinline__ int exponent_of(double m) {
  int e = 0;
  if (m != 0.0) frexp(m, &e);
  return e;
}

// Cuts x (|x| < 1) into s digits; digits[p] in [-127, 127].
template <int S>
__device__ __forceinline__ void cut(double x, int* digits) {
#pragma unroll
  for (int p = 0; p < S; ++p) {
    const double t = x * 128.0;
    const double d = trunc(t);
    digits[p] = int(d);
    x = t - d;
  }
}

// One workgroup per (padded) row of A: row exponent, then the slices.  Loads
// are unrolled and unconditional wherever a whole group lies inside the row,
// so each work-item keeps 8 or 16 loads in flight instead of waiting on one
// load at a time.
template <int S>
__global__ __launch_bounds__(256) void split_rows_kernel(const double* __restrict__ A, int64_t ni, int64_t nk,
                                                         int64_t nkp, int64_t nip, signed char* __restrict__ As,
                                                         int* __restrict__ ea) {
  __shared__ double red[4];
  const int64_t i = blockIdx.x;
  const int tid = threadIdx.x;
  const bool live = i < ni;
  const double* row = A + (live ? i * nk : 0);
  double m = 0.0;
  if (live) {
    int64_t k = tid;
    for (; k + 7 * 256 < nk; k += 8 * 256) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = row[k + u * 256];
#pragma unroll
      for (int u = 0; u < 8; ++u) m = fmax(m, fabs(v[u]));
    }
    for (; k < nk; k += 256) m = fmax(m, fabs(row[k]));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((tid & 63) == 0) red[tid >> 6] = m;
  __syncthreads();
  m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  const int e = exponent_of(m);
  if (tid == 0) ea[i] = e;
  // 16 consecutive k per work-item per step; one 16-byte store per slice.
  for (int64_t k0 = int64_t(tid) * 16; k0 < nkp; k0 += 256 * 16) {
    double x[16];
    if (live && k0 + 16 <= nk) {
#pragma unroll
      for (int u = 0; u < 16; ++u) x[u] = row[k0 + u];
    } else {
#pragma unroll
      for (int u = 0; u < 16; ++u) x[u] = (live && k0 + u < nk) ? row[k0 + u] : 0.0;
    }
    int dig[16][S];
#pragma unroll
    for (int u = 0; u < 16; ++u) cut<S>(ldexp(x[u], -e), dig[u]);
#pragma unroll
    for (int p = 0; p < S; ++p) {
      v4i w;
#pragma unroll
      for (int c = 0; c < 4; ++c)
        w[c] = (dig[4 * c][p] & 0xFF) | ((dig[4 * c + 1][p] & 0xFF) << 8) | ((dig[4 * c + 2][p] & 0xFF) << 16) |
               ((dig[4 * c + 3][p] & 0xFF) << 24);
      *reinterpret_cast<v4i*>(As + (int64_t(p) * nip + i) * nkp + k0) = w;
    }
  }
}

// Column maxima of B as exponents: atomicMax over K chunks, 8 loads in
// flight per work-item.
__global__ __launch_bounds__(256) void col_exponent_kernel(const double* __restrict__ B, int64_t nk, int64_t nj,
                                                           int64_t kchunk, int* __restrict__ fb) {
  const int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (j >= nj) return;
  const int64_t kb = int64_t(blockIdx.y) * kchunk;
  const int64_t ke = (kb + kchunk < nk) ? kb + kchunk : nk;
  double m = 0.0;
  int64_t k = kb;
  for (; k + 8 <= ke; k += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = B[(k + u) * nj + j];
#pragma unroll
    for (int u = 0; u < 8; ++u) m = fmax(m, fabs(v[u]));
  }
  for (; k < ke; ++k) m = fmax(m, fabs(B[k * nj + j]));
  if (m != 0.0) atomicMax(&fb[j], exponent_of(m));
}

__global__ void fill_int_kernel(int* p, int64_t n, int v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// B -> Bs[q][j][k] through an LDS transpose: a workgroup owns 64 columns x 64 k.
template <int S>
__global__ __launch_bounds__(256) void split_cols_kernel(const double* __restrict__ B, int64_t nk, int64_t nj,
                                                         int64_t nkp, int64_t njp, const int* __restrict__ fb,
                                                         signed char* __restrict__ Bs) {
  __shared__ __attribute__((aligned(16))) signed char lds[S][64][64 + 16];
  const int64_t j0 = int64_t(blockIdx.x) * 64, k0 = int64_t(blockIdx.y) * 64;
  const int tid = threadIdx.x;
  const int jj = tid & 63;
  const int64_t j = j0 + jj;
  const int f = (j < nj) ? fb[j] : 0;
  const int sh = -(f == INT32_MIN ? 0 : f);
  double x[16];
  if (j0 + 64 <= nj && k0 + 64 <= nk) {
#pragma unroll
    for (int u = 0; u < 16; ++u) x[u] = B[(k0 + (tid >> 6) + 4 * u) * nj + j];
  } else {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int64_t k = k0 + (tid >> 6) + 4 * u;
      x[u] = (j < nj && k < nk) ? B[k * nj + j] : 0.0;
    }
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    int dig[S];
    cut<S>(ldexp(x[u], sh), dig);
#pragma unroll
    for (int q = 0; q < S; ++q) lds[q][jj][(tid >> 6) + 4 * u] = (signed char)dig[q];
  }
  __syncthreads();
  // 64 columns x 64 bytes per slice: 16 bytes per work-item.
  const int row = tid >> 2, part = (tid & 3) * 16;
#pragma unroll
  for (int q = 0; q < S; ++q) {
    const v4i w = *reinterpret_cast<const v4i*>(&lds[q][row][part]);
    *reinterpret_cast<v4i*>(Bs + (int64_t(q) * njp + j0 + row) * nkp + k0 + part) = w;
  }
}

// LDS slot (in 16-byte chunks) of chunk c of tile row r; the XOR stays
// inside the row's kChunks slots.  It spreads the 16 rows an MFMA fragment
// reads at one chunk over all 16 bank groups (16-byte groups of a 256-byte
// bank cycle), and the chunks a loader wavefront writes likewise.
__device__ __forceinline__ int swz(int r, int c) {
  if constexpr (kChunks == 16) return r * kChunks + (c ^ (r & 15));
  else return r * kChunks + (c ^ ((r >> 1) & (kChunks - 1)));
}
static_assert(kChunks == 16 || kChunks == 8, "swizzle covers 8 or 16 chunks per row");

// 256x256 output tile per workgroup of eight wavefronts (64x128 each, a 4x8
// grid of 16x16 int32 accumulators).  For t = s-1 down to 0 the K loop runs
// over the concatenation of the slice pairs (p, t-p), and the exact int32
// sum S_t of the tile goes to scratch; ozaki_combine_kernel folds the S_t
// into C.  The tile size is set by L2 bandwidth: a stage brings 128 bytes of
// K for 256 rows of A and of B (64 KiB) for 8.4 M MACs, 128 MACs per byte,
// where a 128x128 tile's 64 left the MFMAs waiting on L2.  Keeping only the
// int32 accumulators in registers (an FP64 fold would take two more per
// output) is what lets a wavefront own 64x128 outputs at two wavefronts per
// SIMD.  Two stages of LDS (128 KiB), one workgroup per CU; the next stage's
// global loads are in flight in registers while one computes.
constexpr int kThreads = 512;
constexpr int kWaveM = 64, kWaveN = 128;
constexpr int kFm = kWaveM / 16, kFn = kWaveN / 16;  // 16x16 tiles per wavefront
constexpr int kLoads = kTM * kChunks / kThreads;       // 16-byte chunks per work-item per operand
static_assert(kTM == kTN, "one loader mapping for both operands");

template <int S>
__global__ __launch_bounds__(kThreads) void ozaki_gemm_kernel(int* __restrict__ Sout, const signed char* __restrict__ As,
                                                              const signed char* __restrict__ Bs, int64_t nip,
                                                              int64_t njp, int64_t nkp, int tiles_n) {
  __shared__ v4i Al[2][kTM * kChunks];
  __shared__ v4i Bl[2][kTN * kChunks];
  int64_t tm, tn;
  tile_of(blockIdx.x, gridDim.x, int(nip / kTM), tiles_n, tm, tn);
  const int64_t m0 = tm * kTM, n0 = tn * kTN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * kWaveM, wn = (wave & 1) * kWaveN;
  const int64_t ksteps = nkp / kTK;
  // Loader mapping: work-item tid moves chunk lc of rows lr + kRowStep*u, u < kLoads.
  constexpr int kRowStep = kThreads / kChunks;
  const int lc = tid % kChunks, lr = tid / kChunks;
  const int fr = lane & 15, fk = lane >> 4;  // MFMA operand row and k quarter

  for (int t = S - 1; t >= 0; --t) {
    const int p_lo = t - (S - 1) > 0 ? t - (S - 1) : 0;
    const int p_hi = t < S - 1 ? t : S - 1;
    const int64_t nsteps = int64_t(p_hi - p_lo + 1) * ksteps;
    v4i acc[kFm][kFn];
#pragma unroll
    for (int i = 0; i < kFm; ++i)
#pragma unroll
      for (int j = 0; j < kFn; ++j) acc[i][j] = v4i{0, 0, 0, 0};

    v4i ga[kLoads], gb[kLoads];
    auto gload = [&](int64_t st) {
      const int p = p_lo + int(st / ksteps);
      const int q = t - p;
      const int64_t k0 = (st % ksteps) * kTK;
      const signed char* a = As + (int64_t(p) * nip + m0 + lr) * nkp + k0 + 16 * lc;
      const signed char* b = Bs + (int64_t(q) * njp + n0 + lr) * nkp + k0 + 16 * lc;
#pragma unroll
      for (int u = 0; u < kLoads; ++u) {
        ga[u] = *reinterpret_cast<const v4i*>(a + int64_t(kRowStep * u) * nkp);
        gb[u] = *reinterpret_cast<const v4i*>(b + int64_t(kRowStep * u) * nkp);
      }
    };
    auto lstore = [&](int buf) {
#pragma unroll
      for (int u = 0; u < kLoads; ++u) {
        Al[buf][swz(lr + kRowStep * u, lc)] = ga[u];
        Bl[buf][swz(lr + kRowStep * u, lc)] = gb[u];
      }
    };

    gload(0);
    __syncthreads();  // previous t's readers are done with buffer 0
    lstore(0);
    __syncthreads();
    for (int64_t st = 0; st < nsteps; ++st) {
      const int buf = int(st & 1);
      if (st + 1 < nsteps) gload(st + 1);
      // Raised wave priority over the MFMA block: with two wavefronts per
      // SIMD the arbiter then keeps feeding the matrix pipe from one while
      // the other issues its loads and LDS stores (7 slices at 4096^3:
      // 2.05-2.10 -> 2.00-2.01 ms; the FP64 MFMA kernel was slower with it).
      __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int kk = 0; kk < kChunks / 4; ++kk) {
        const int c = kk * 4 + fk;
        v4i af[kFm], bf[kFn];
#pragma unroll
        for (int i = 0; i < kFm; ++i) af[i] = Al[buf][swz(wm + i * 16 + fr, c)];
#pragma unroll
        for (int j = 0; j < kFn; ++j) bf[j] = Bl[buf][swz(wn + j * 16 + fr, c)];
#pragma unroll
        for (int i = 0; i < kFm; ++i)
#pragma unroll
          for (int j = 0; j < kFn; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
      }
      __builtin_amdgcn_s_setprio(0);
      if (st + 1 < nsteps) {
        lstore(buf ^ 1);
        __syncthreads();
      }
    }
    // int32 16x16 accumulator layout: element r of lane l is row 4*(l/16)+r,
    // column l%16.  The padded S_t grid covers every tile.
    int* st_out = Sout + int64_t(t) * nip * njp;
#pragma unroll
    for (int i = 0; i < kFm; ++i)
#pragma unroll
      for (int j = 0; j < kFn; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t gi = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
          const int64_t gj = n0 + wn + j * 16 + (lane & 15);
          st_out[gi * njp + gj] = acc[i][j][r];
        }
  }
}

// C[i][j] = alpha * 2^(e_i+f_j) * sum_t S_t[i][j] 2^{-7(t+2)}, the terms
// added in FP64 from t = s-1 (the smallest) down to 0.
template <int S>
__global__ __launch_bounds__(256) void ozaki_combine_kernel(double* __restrict__ C, const int* __restrict__ Sin,
                                                            const int* __restrict__ ea, const int* __restrict__ fb,
                                                            double alpha, int64_t ni, int64_t nj, int64_t nip,
                                                            int64_t njp) {
  const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= ni * nj) return;
  const int64_t i = idx / nj, j = idx % nj;
  double v = 0.0;
#pragma unroll
  for (int t = S - 1; t >= 0; --t) v += double(Sin[(int64_t(t) * nip + i) * njp + j]) * ldexp(1.0, -7 * (t + 2));
  const int fj = fb[j];
  C[idx] = alpha * ldexp(v, ea[i] + (fj == INT32_MIN ? 0 : fj));
}

template <int S>
int run_mixed(double* C, const double* A, const double* B, double alpha, int64_t ni, int64_t nj, int64_t nk,
              const MixedLayout& l, char* base, hipStream_t s) {
  signed char* As = reinterpret_cast<signed char*>(base + l.a_off);
  signed char* Bs = reinterpret_cast<signed char*>(base + l.b_off);
  int* ea = reinterpret_cast<int*>(base + l.ea_off);
  int* fb = reinterpret_cast<int*>(base + l.fb_off);
  split_rows_kernel<S><<<dim3(unsigned(l.nip)), dim3(256), 0, s>>>(A, ni, nk, l.nkp, l.nip, As, ea);
  fill_int_kernel<<<dim3(unsigned(ceil_div(l.njp, 256))), dim3(256), 0, s>>>(fb, l.njp, INT32_MIN);
  const int64_t kchunk = 64;
  if (nk > 0)
    col_exponent_kernel<<<dim3(unsigned(ceil_div(nj, 256)), unsigned(ceil_div(nk, kchunk))), dim3(256), 0, s>>>(
        B, nk, nj, kchunk, fb);
  split_cols_kernel<S><<<dim3(unsigned(l.njp / 64), unsigned(l.nkp / 64)), dim3(256), 0, s>>>(B, nk, nj, l.nkp,
                                                                                              l.njp, fb, Bs);
  int* St = reinterpret_cast<int*>(base + l.s_off);
  const int64_t tm = l.nip / kTM, tn = l.njp / kTN;
  ozaki_gemm_kernel<S><<<dim3(unsigned(tm * tn)), dim3(kThreads), 0, s>>>(St, As, Bs, l.nip, l.njp, l.nkp, int(tn));
  ozaki_combine_kernel<S><<<dim3(unsigned(ceil_div(ni * nj, 256))), dim3(256), 0, s>>>(C, St, ea, fb, alpha, ni, nj,
                                                                                       l.nip, l.njp);
  return launch_status();
}

}  // namespace

extern "C" {

size_t rpx_gemm_mixed_scratch_bytes(int64_t ni, int64_t nj, int64_t nk, int nslices) {
  if (ni < 0 || nj < 0 || nk < 0) return 0;
  const int s = nslices <= 0 ? 7 : nslices;
  if (s > kMaxSlices) return 0;
  return mixed_layout(ni, nj, nk, s).total;
}

int rpx_gemm_mixed_f64(double* C, const double* A, const double* B, double alpha, double /*beta*/, int64_t ni,
                       int64_t nj, int64_t nk, int nslices, void* scratch, size_t scratch_bytes, void* stream) {
  if (ni < 0 || nj < 0 || nk < 0 || nslices < 0 || nslices > kMaxSlices) return RPX_ERR_INVALID_ARGUMENT;
  if (ni == 0 || nj == 0) return RPX_SUCCESS;
  if (!C || !scratch || (nk > 0 && (!A || !B))) return RPX_ERR_INVALID_ARGUMENT;
  const int s = nslices == 0 ? 7 : nslices;
  if (nk > max_k(s)) return RPX_ERR_UNSUPPORTED;
  const MixedLayout l = mixed_layout(ni, nj, nk, s);
  if (scratch_bytes < l.total) return RPX_ERR_SCRATCH_TOO_SMALL;
  if (l.nip > INT32_MAX || (l.nip / kTM) * (l.njp / kTN) > INT32_MAX || ceil_div(ni * nj, 256) > INT32_MAX)
    return RPX_ERR_UNSUPPORTED;
  char* base = reinterpret_cast<char*>(scratch);
  hipStream_t st = as_stream(stream);
  switch (s) {
    case 1: return run_mixed<1>(C, A, B, alpha, ni, nj, nk, l, base, st);
    case 2: return run_mixed<2>(C, A, B, alpha, ni, nj, nk, l, base, st);
    case 3: return run_mixed<3>(C, A, B, alpha, ni, nj, nk, l, base, st);
    case 4: return run_mixed<4>(C, A, B, alpha, ni, nj, nk, l, base, st);
    case 5: return run_mixed<5>(C, A, B, alpha, ni, nj, nk, l, base, st);
    case 6: return run_mixed<6>(C, A, B, alpha, ni, nj, nk, l, base, st);
    case 7: return run_mixed<7>(C, A, B, alpha, ni, nj, nk, l, base, st);
    case 8: return run_mixed<8>(C, A, B, alpha, ni, nj, nk, l, base, st);
    default: return RPX_ERR_UNSUPPORTED;
  }
}

}  // extern "C"
