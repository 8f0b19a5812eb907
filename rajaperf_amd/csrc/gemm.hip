// GEMM (polybench): C[i][j] = sum_k (alpha*A[i][k]) * B[k][j].  With
// RPX_TUNE_GEMM_BETA the textbook form C = alpha*A*B + beta*C instead (the
// old C read once; the MFMA kernel starts its sums at beta*C, the
// comparator adds beta*C to its sum).
//
// Reference: src/polybench/POLYBENCH_GEMM.hpp:12-50 (C *= beta, then C is
// overwritten by the dot, so beta never reaches the result),
// POLYBENCH_GEMM-Seq.cpp:30-50, POLYBENCH_GEMM-Hip.cpp:27-102 (one thread
// per C element, no LDS, no matrix cores).
//
// Two FP64 paths:
//  * gemm_naive_kernel: the one-thread-per-element shape of the suite's
//    Base_HIP variant, kept as the "plain FP64 GPU" comparator;
//  * gemm_mfma_f64_kernel: 128x128 workgroup tiles, four wavefronts of 64x64,
//    each a 4x4 grid of v_mfma_f64_16x16x4f64 accumulators; A (pre-scaled by
//    alpha, as the suite rounds alpha*A[i][k]) and B staged through LDS in
//    16-deep K slabs with the next slab's global loads in flight during the
//    current slab's MFMAs.  Tiles are mapped XCD-major so neighbouring tiles
//    that share A rows / B columns run on the same XCD's L2.
#include "rpx_common.hpp"

using namespace rpx;

namespace {

typedef double v2d __attribute__((ext_vector_type(2)));
typedef double v4d __attribute__((ext_vector_type(4)));

// ---- plain FP64 comparator ---------------------------------------------------
__global__ __launch_bounds__(256) void gemm_naive_kernel(double* __restrict__ C, const double* __restrict__ A,
                                                         const double* __restrict__ B, double alpha, double beta,
                                                         int use_beta, int64_t ni, int64_t nj, int64_t nk) {
  const int64_t j = int64_t(blockIdx.x) * 32 + (threadIdx.x & 31);
  const int64_t i = int64_t(blockIdx.y) * 8 + (threadIdx.x >> 5);
  if (i >= ni || j >= nj) return;
  double dot = 0.0;
  for (int64_t k = 0; k < nk; ++k) dot = fma(alpha * A[k + i * nk], B[j + k * nj], dot);
  C[j + i * nj] = use_beta ? dot + beta * C[j + i * nj] : dot;
}

// ---- FP64 MFMA ---------------------------------------------------------------
//
// 128x128 output tile per workgroup of four wavefronts (64x64 each, a 4x4
// grid of v_mfma_f64_16x16x4f64 tiles), K in slabs of 16, double-buffered
// in LDS.  Both operands sit in LDS k-contiguous ([m][k] for A, [n][k] for
// B, rows padded by two doubles), and in each pair of MFMA steps lane (r, q)
// of a fragment takes the adjacent k = 8h + 2q and 8h + 2q + 1, so one
// 16-byte LDS read feeds two MFMAs.  The k order inside a slab is thereby
// permuted; the result is the same sum in another association.  Interior
// tiles load full slabs with unconditional 16-byte loads (A rows as they
// are, B as 2x4 blocks transposed in registers); edge tiles, partial slabs
// and odd shapes take guarded 8-byte loads.
constexpr int kTM = 128, kTN = 128, kTK = 16;
constexpr int kThreads = 256;
constexpr int kLdsRow = kTK + 2;  // doubles per LDS row of either operand

// BETA: the textbook C = beta*C + the products (a separate kernel; the
// suite's form reads no C).
template <bool BETA>
__device__ __forceinline__ void gemm_mfma_f64_tile(double* __restrict__ C, const double* __restrict__ A,
                                                   const double* __restrict__ B, double alpha, double beta,
                                                   int64_t ni, int64_t nj, int64_t nk, int tiles_n, int vec) {
  __shared__ __attribute__((aligned(16))) double As[2][kTM * kLdsRow];  // As[m][k]
  __shared__ __attribute__((aligned(16))) double Bs[2][kTN * kLdsRow];  // Bs[n][k]

  const unsigned tile = xcd_major(blockIdx.x, gridDim.x);
  const int64_t m0 = int64_t(tile / unsigned(tiles_n)) * kTM;
  const int64_t n0 = int64_t(tile % unsigned(tiles_n)) * kTN;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 64;  // wave's row offset in the tile
  const int wn = (wave & 1) * 64;   // wave's column offset
  const bool interior = vec && m0 + kTM <= ni && n0 + kTN <= nj;

  // Loader mapping.  A: row a_m, k a_k .. a_k+7.  B: k rows b_k, b_k+1,
  // columns b_n .. b_n+3.
  const int a_m = tid >> 1, a_k = (tid & 1) * 8;
  const int b_k = (tid >> 5) * 2, b_n = (tid & 31) * 4;

  double ra[8], rb[8];  // rb[c*2 + t]: B(k = b_k + t, n = b_n + c)
  auto load_slab = [&](int64_t k0) {
    if (interior && k0 + kTK <= nk) {
      const double* pa = A + (m0 + a_m) * nk + k0 + a_k;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const v2d x = *reinterpret_cast<const v2d*>(pa + 2 * u);
        ra[2 * u] = alpha * x.x;
        ra[2 * u + 1] = alpha * x.y;
      }
      const double* pb = B + (k0 + b_k) * nj + n0 + b_n;
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const v2d x0 = *reinterpret_cast<const v2d*>(pb + t * nj);
        const v2d x1 = *reinterpret_cast<const v2d*>(pb + t * nj + 2);
        rb[0 * 2 + t] = x0.x;
        rb[1 * 2 + t] = x0.y;
        rb[2 * 2 + t] = x1.x;
        rb[3 * 2 + t] = x1.y;
      }
    } else {
      const int64_t gi = m0 + a_m;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int64_t gk = k0 + a_k + u;
        ra[u] = (gi < ni && gk < nk) ? alpha * A[gi * nk + gk] : 0.0;
      }
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int64_t gk = k0 + b_k + t, gj = n0 + b_n + c;
          rb[c * 2 + t] = (gk < nk && gj < nj) ? B[gk * nj + gj] : 0.0;
        }
    }
  };
  auto store_slab = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
      *reinterpret_cast<v2d*>(&As[buf][a_m * kLdsRow + a_k + 2 * u]) = v2d{ra[2 * u], ra[2 * u + 1]};
#pragma unroll
    for (int c = 0; c < 4; ++c)
      *reinterpret_cast<v2d*>(&Bs[buf][(b_n + c) * kLdsRow + b_k]) = v2d{rb[c * 2], rb[c * 2 + 1]};
  };

  // With BETA the accumulators start at beta*C (the old C read once, before
  // the K loop), so the result is beta*C plus the products; otherwise at 0.
  // The C reads go one fragment row at a time (the empty asm statement keeps
  // the compiler from issuing all 64 together, which took 100 more VGPRs
  // and a wavefront per SIMD).
  v4d acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if constexpr (BETA) {
        if (j == 0) asm volatile("" ::: "memory");
      }
      acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
      if constexpr (BETA) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t gi = m0 + wm + i * 16 + 4 * r + (lane >> 4);
          const int64_t gj = n0 + wn + j * 16 + (lane & 15);
          // Clamped into C: the accumulators of rows or columns outside C
          // are never stored, so what they start from does not matter.
          const int64_t ci = gi < ni ? gi : ni - 1, cj = gj < nj ? gj : nj - 1;
          acc[i][j][r] = beta * C[ci * nj + cj];
        }
      }
    }

  const int64_t nslabs = (nk + kTK - 1) / kTK;
  load_slab(0);
  store_slab(0);
  __syncthreads();
  const int fr = lane & 15;  // fragment row (A) / column (B)
  const int fq = lane >> 4;  // k pair within an 8-deep half slab
  for (int64_t s = 0; s < nslabs; ++s) {
    const int buf = int(s & 1);
    if (s + 1 < nslabs) load_slab((s + 1) * kTK);
#pragma unroll
    for (int h = 0; h < kTK / 8; ++h) {
      v2d af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
        af[i] = *reinterpret_cast<const v2d*>(&As[buf][(wm + i * 16 + fr) * kLdsRow + 8 * h + 2 * fq]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        bf[j] = *reinterpret_cast<const v2d*>(&Bs[buf][(wn + j * 16 + fr) * kLdsRow + 8 * h + 2 * fq]);
      // All 16 tiles take their first k before any takes its second, so
      // each accumulator's two MFMAs are 16 instructions apart.
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i].x, bf[j].x, acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i].y, bf[j].y, acc[i][j], 0, 0, 0);
      // The backend's LDS-read/MFMA interleave for small GEMM loops: 2.228
      // against 2.248 ms at 4096^3 (strategy 1 was slower, 2.305 ms).
      __builtin_amdgcn_iglp_opt(0);
    }
    if (s + 1 < nslabs) {
      store_slab(buf ^ 1);
      __syncthreads();
    }
  }

  // v_mfma_f64_16x16x4f64 accumulator layout (measured on gfx950): element r
  // of lane l is row 4*r + l/16, column l%16.
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = m0 + wm + i * 16 + 4 * r + (lane >> 4);
        const int64_t gj = n0 + wn + j * 16 + (lane & 15);
        if (gi < ni && gj < nj) C[gi * nj + gj] = acc[i][j][r];
      }
  }
}

__global__ __launch_bounds__(kThreads) void gemm_mfma_f64_kernel(double* __restrict__ C,
                                                                 const double* __restrict__ A,
                                                                 const double* __restrict__ B, double alpha,
                                                                 int64_t ni, int64_t nj, int64_t nk, int tiles_n,
                                                                 int vec) {
  gemm_mfma_f64_tile<false>(C, A, B, alpha, 0.0, ni, nj, nk, tiles_n, vec);
}

// The textbook form.  Left to itself the compiler gives this instance 256
// VGPRs and 148 AGPRs (one wavefront per SIMD, 1.6x the time); capped at
// two wavefronts per SIMD it spills 20 bytes per work-item.
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void gemm_mfma_f64_beta_kernel(
    double* __restrict__ C, const double* __restrict__ A, const double* __restrict__ B, double alpha, double beta,
    int64_t ni, int64_t nj, int64_t nk, int tiles_n, int vec) {
  gemm_mfma_f64_tile<true>(C, A, B, alpha, beta, ni, nj, nk, tiles_n, vec);
}

}  // namespace

extern "C" {

int rpx_gemm_f64(double* C, const double* A, const double* B, double alpha, double beta, int64_t ni,
                 int64_t nj, int64_t nk, const rpx_tuning* tune, void* stream) {
  if (ni < 0 || nj < 0 || nk < 0) return RPX_ERR_INVALID_ARGUMENT;
  if (ni == 0 || nj == 0) return RPX_SUCCESS;
  if (!C || (nk > 0 && (!A || !B))) return RPX_ERR_INVALID_ARGUMENT;
  hipStream_t s = as_stream(stream);
  const int use_beta = (tune_flags(tune) & RPX_TUNE_GEMM_BETA) != 0;
  if (tune_flags(tune) & RPX_TUNE_NAIVE) {
    const int64_t gx = ceil_div(nj, 32), gy = ceil_div(ni, 8);
    if (gx > INT32_MAX || gy > 65535) return RPX_ERR_UNSUPPORTED;
    gemm_naive_kernel<<<dim3(unsigned(gx), unsigned(gy)), dim3(256), 0, s>>>(C, A, B, alpha, beta, use_beta, ni,
                                                                              nj, nk);
    return launch_status();
  }
  const int64_t tm = ceil_div(ni, kTM), tn = ceil_div(nj, kTN);
  if (tm * tn > INT32_MAX || tn > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  const int vec = (nk % 2 == 0) && (nj % 2 == 0) && aligned16(A) && aligned16(B);
  if (use_beta)
    gemm_mfma_f64_beta_kernel<<<dim3(unsigned(tm * tn)), dim3(kThreads), 0, s>>>(C, A, B, alpha, beta, ni, nj, nk,
                                                                                int(tn), vec);
  else
    gemm_mfma_f64_kernel<<<dim3(unsigned(tm * tn)), dim3(kThreads), 0, s>>>(C, A, B, alpha, ni, nj, nk, int(tn),
                                                                           vec);
  return launch_status();
}

}  // extern "C"
