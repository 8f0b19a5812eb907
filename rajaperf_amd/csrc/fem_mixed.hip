// MASS3DPA with its basis contractions on int8 matrix cores, FP64 accuracy.
//
// Same operator as the suite's MASS3DPA (src/apps/MASS3DPA.hpp:158-339,
// MASS3DPA-Seq.cpp:29-94): Y_e += Bt' x Bt' x Bt' (D_e . (B x B x B X_e)),
// with B(q,d) = B[q + Q1D*d] and the suite's literal Bt'(d,q) = Bt[q+D1D*d].
//
// MI355X design.  For D1D <= 4 and Q1D <= 5 the three forward contractions
// together are one dense Q^3 x D^3 Kronecker matrix Kf (<= 125 x 64), shared
// by all elements, and the three backward ones one D^3 x Q^3 matrix Kb; each
// element's action is then two small GEMMs whose K (64, then 128) fits the
// K = 64 of v_mfma_i32_16x16x64_i8.  Both run as error-free int8 slice
// products (the Ozaki splitting used by the mixed GEMM):
//   * every row of Kf and Kb is scaled by a power of two from its max |entry|,
//     every element's X and every element's D . (Kf X) by one from their
//     max, so all entries lie in (-1, 1), and cut into S signed 7-bit digits
//     by exact FP64 operations (scale by 128, truncate, subtract);
//   * for each level t < S the slice pairs (p, t-p) accumulate exactly in
//     int32 (|sum| <= 7 * 128 * 127^2 < 2^31), and the levels are summed in
//     FP64 with weight 2^-7(t+2), smallest first;
//   * the quadrature scaling by D and the final Y += are FP64.
// The dropped digit products bound the error of each contraction by about
// (S+3) 2^-7S relative to the scaled magnitudes: with S = 7 each output is
// within 1e-13 of the FP64 result relative to |Bt'|x3 (|D| . |B|x3 |X|).
//
// A workgroup of four wavefronts owns 32 elements.  X's slices, the block's
// quadrature data (loaded straight into LDS at the start) and then the
// slices of D . (Kf X) live in LDS (about 50 KB).  The forward product gives
// each wavefront two 16-row bands of the 128 quadrature rows for all 32
// elements, the backward product one 16-row band of the 64 dof rows; the
// Kronecker slices of its bands (S x 2 KB) are loaded into the wavefront's
// registers once, from the scratch buffer that every workgroup shares
// through L2 (loading them inside the slice loop left the MFMAs waiting on
// L2 latency: 0.58 ms at NE = 262144).
#include "rpx_common.hpp"
#include "fem_common.hpp"

using namespace rpx;

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kMaxSlices = 8;
constexpr int kQP = 128;  // Q^3 padded: forward rows, backward K
constexpr int kDP = 64;   // D^3 padded: forward K, backward rows
constexpr int kEB = 32;   // elements per workgroup
constexpr int kXRow = kDP + 16;  // LDS row strides in bytes (16 B of padding)
constexpr int kURow = kQP + 16;

struct MixedFemLayout {
  size_t kf_off, kb_off, ea_off, eb_off, total;
};
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
MixedFemLayout fem_layout(int s) {
  MixedFemLayout l;
  l.kf_off = 0;
  l.kb_off = align256(size_t(s) * kQP * kDP);
  l.ea_off = l.kb_off + align256(size_t(s) * kDP * kQP);
  l.eb_off = l.ea_off + align256(kQP * sizeof(int));
  l.total = l.eb_off + align256(kDP * sizeof(int));
  return l;
}

__device__ __forceinline__ int exponent_of(double m) {
  int e = 0;
  if (m != 0.0) frexp(m, &e);
  return e;
}

// x (|x| < 1) = sum_p d_p 2^-7(p+1) + r, d_p in [-127, 127], |r| < 2^-7S.
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

// The same cut one digit at a time: returns the next digit, keeps the rest.
__device__ __forceinline__ int next_digit(double& x) {
  const double t = x * 128.0;
  const double d = trunc(t);
  x = t - d;
  return int(d);
}

__device__ __forceinline__ int pack4(int a, int b, int c, int d) {
  return (a & 0xFF) | ((b & 0xFF) << 8) | ((c & 0xFF) << 16) | ((d & 0xFF) << 24);
}

// The two Kronecker matrices, row-scaled and sliced:
//   Kf[p][m][k], m = (qz*Q + qy)*Q + qx, k = (dz*D + dy)*D + dx,
//     Kf(m,k) = (B(qz,dz) * B(qy,dy)) * B(qx,dx);
//   Kb[p][j][k], j = (dz*D + dy)*D + dx, k = (qz*Q + qy)*Q + qx,
//     Kb(j,k) = (Bt'(dz,qz) * Bt'(dy,qy)) * Bt'(dx,qx).
// One work-item per row: 128 rows of Kf, then 64 rows of Kb.
template <int S>
__global__ __launch_bounds__(256) void kron_slices_kernel(const double* __restrict__ B, const double* __restrict__ Bt,
                                                          int D, int Q, signed char* __restrict__ Kf,
                                                          signed char* __restrict__ Kb, int* __restrict__ ea,
                                                          int* __restrict__ eb) {
  const int r = threadIdx.x;
  const int D3 = D * D * D, Q3 = Q * Q * Q;
  auto bq = [&](int q, int d) { return B[q + Q * d]; };
  auto bt = [&](int d, int q) { return Bt[q + D * d]; };
  if (r < kQP) {
    const int m = r;
    const int qz = m / (Q * Q), qy = (m / Q) % Q, qx = m % Q;
    double mx = 0.0;
    for (int k = 0; k < D3; ++k) {
      const int dz = k / (D * D), dy = (k / D) % D, dx = k % D;
      if (m < Q3) mx = fmax(mx, fabs((bq(qz, dz) * bq(qy, dy)) * bq(qx, dx)));
    }
    const int e = exponent_of(mx);
    ea[m] = e;
    for (int k = 0; k < kDP; k += 4) {
      int dig[4][S];
      for (int u = 0; u < 4; ++u) {
        const int kk = k + u;
        const int dz = kk / (D * D), dy = (kk / D) % D, dx = kk % D;
        const double v = (m < Q3 && kk < D3) ? (bq(qz, dz) * bq(qy, dy)) * bq(qx, dx) : 0.0;
        cut<S>(ldexp(v, -e), dig[u]);
      }
      for (int p = 0; p < S; ++p)
        *reinterpret_cast<int*>(Kf + (p * kQP + m) * kDP + k) = pack4(dig[0][p], dig[1][p], dig[2][p], dig[3][p]);
    }
  } else if (r < kQP + kDP) {
    const int j = r - kQP;
    const int dz = j / (D * D), dy = (j / D) % D, dx = j % D;
    double mx = 0.0;
    for (int k = 0; k < Q3; ++k) {
      const int qz = k / (Q * Q), qy = (k / Q) % Q, qx = k % Q;
      if (j < D3) mx = fmax(mx, fabs((bt(dz, qz) * bt(dy, qy)) * bt(dx, qx)));
    }
    const int e = exponent_of(mx);
    eb[j] = e;
    for (int k = 0; k < kQP; k += 4) {
      int dig[4][S];
      for (int u = 0; u < 4; ++u) {
        const int kk = k + u;
        const int qz = kk / (Q * Q), qy = (kk / Q) % Q, qx = kk % Q;
        const double v = (j < D3 && kk < Q3) ? (bt(dz, qz) * bt(dy, qy)) * bt(dx, qx) : 0.0;
        cut<S>(ldexp(v, -e), dig[u]);
      }
      for (int p = 0; p < S; ++p)
        *reinterpret_cast<int*>(Kb + (p * kDP + j) * kQP + k) = pack4(dig[0][p], dig[1][p], dig[2][p], dig[3][p]);
    }
  }
}

template <int D, int Q, int S>
__global__ __launch_bounds__(256) void mass3dpa_mixed_kernel(const signed char* __restrict__ Kf,
                                                             const signed char* __restrict__ Kb,
                                                             const int* __restrict__ ea, const int* __restrict__ eb,
                                                             const double* __restrict__ Dg,
                                                             const double* __restrict__ Xg, double* __restrict__ Yg,
                                                             int64_t NE) {
  constexpr int D3 = D * D * D, Q3 = Q * Q * Q;
  // LDS: X's slices and the block's quadrature data D during the forward
  // product, then the slices of D . (Kf X) over both.
  constexpr int kXsBytes = S * kEB * kXRow;
  constexpr int kDlBytes = (kEB * Q3 * 8 + 15) / 16 * 16;
  constexpr int kUsBytes = S * kEB * kURow;
  constexpr int kLdsBytes = (kXsBytes + kDlBytes) > kUsBytes ? (kXsBytes + kDlBytes) : kUsBytes;
  __shared__ __attribute__((aligned(16))) signed char lds[kLdsBytes];
  __shared__ int eX[kEB];
  __shared__ double umax[4][kEB];
  auto Xs = reinterpret_cast<signed char(*)[kEB][kXRow]>(lds);
  auto Us = reinterpret_cast<signed char(*)[kEB][kURow]>(lds);
  double* Dl = reinterpret_cast<double*>(lds + kXsBytes);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int fr = lane & 15, fk = (lane >> 4) * 16;  // MFMA operand row / first K byte
  const int64_t e0 = int64_t(blockIdx.x) * kEB;
  const int ne = int((NE - e0) < kEB ? (NE - e0) : kEB);

  // The block's D (ne * Q3 contiguous doubles) goes straight to LDS with
  // global_load_lds (each lane's 16 bytes land at base + 16*lane) and is in
  // flight during the slicing and the forward product.  A trailing odd
  // double is copied by one work-item.
  {
    const int count = ne * Q3;
    const double* src = Dg + e0 * Q3;
    for (int c = wave; c * 128 < count; c += 4) {
      const int i = c * 128 + 2 * lane;
      if (i + 1 < count)
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + i),
                                         (__attribute__((address_space(3))) void*)(Dl + c * 128), 16, 0, 0);
    }
    if ((count & 1) && tid == 0) Dl[count - 1] = src[count - 1];
  }

  // X -> exponent and slices.  Eight work-items per element, eight values each.
  {
    const int n = tid >> 3, k0 = (tid & 7) * 8;
    const int64_t e = e0 + n;
    double x[8];
    double mx = 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = k0 + u;
      x[u] = (e < NE && k < D3) ? Xg[e * D3 + k] : 0.0;
      mx = fmax(mx, fabs(x[u]));
    }
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) mx = fmax(mx, __shfl_xor(mx, off, kWave));
    const int ex = exponent_of(mx);
    if ((tid & 7) == 0) eX[n] = ex;
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = ldexp(x[u], -ex);
    // One digit of every value per pass keeps few registers live.
#pragma unroll
    for (int p = 0; p < S; ++p) {
      int d[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) d[u] = next_digit(x[u]);
      int2 w;
      w.x = pack4(d[0], d[1], d[2], d[3]);
      w.y = pack4(d[4], d[5], d[6], d[7]);
      *reinterpret_cast<int2*>(&Xs[p][n][k0]) = w;
    }
  }
  __syncthreads();

  // Forward: T(m, n) = sum_k Kf(m, k) X_n(k) for the wavefront's rows
  // m in [32*wave, 32*wave + 32) and all 32 elements.  The wavefront's Kf
  // slices (2 x 16 rows x 64 bytes each) stay in registers; X's slices come
  // from LDS.
  v4d t_acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) t_acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
  {
    v4i kf[S][2];
#pragma unroll
    for (int p = 0; p < S; ++p)
#pragma unroll
      for (int i = 0; i < 2; ++i)
        kf[p][i] = *reinterpret_cast<const v4i*>(Kf + (p * kQP + wave * 32 + i * 16 + fr) * kDP + fk);
#pragma unroll
    for (int t = S - 1; t >= 0; --t) {
      v4i acc[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = v4i{0, 0, 0, 0};
#pragma unroll
      for (int p = 0; p <= t; ++p) {
        const int q = t - p;
        // Load X's slices per pair: a compiler barrier keeps the unrolled
        // pairs from hoisting (and holding) all of their LDS loads at once.
        asm volatile("" ::: "memory");
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const v4i bf = *reinterpret_cast<const v4i*>(&Xs[q][j * 16 + fr][fk]);
#pragma unroll
          for (int i = 0; i < 2; ++i)
            acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(kf[p][i], bf, acc[i][j], 0, 0, 0);
        }
      }
      const double w = ldexp(1.0, -7 * (t + 2));
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) t_acc[i][j][r] += double(acc[i][j][r]) * w;
    }
  }

  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wavefront's D loads are in LDS
  __syncthreads();                      // ... and every other wavefront's

  // U = D . T, and each element's max |U| over the wavefront's rows.
  // Accumulator layout: element r of lane l is row 4*(l/16)+r, column l%16.
  double u[2][2][4];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int n = j * 16 + fr;
    const int64_t e = e0 + n;
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = wave * 32 + i * 16 + 4 * (lane >> 4) + r;
        double v = 0.0;
        if (m < Q3 && e < NE) v = ldexp(t_acc[i][j][r], ea[m] + eX[n]) * Dl[n * Q3 + m];
        u[i][j][r] = v;
        mx = fmax(mx, fabs(v));
      }
    mx = fmax(mx, __shfl_xor(mx, 16, kWave));
    mx = fmax(mx, __shfl_xor(mx, 32, kWave));
    if (lane < 16) umax[wave][n] = mx;
  }
  __syncthreads();

  // U -> slices, scaled per element.
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int n = j * 16 + fr;
    const int eu = exponent_of(fmax(fmax(umax[0][n], umax[1][n]), fmax(umax[2][n], umax[3][n])));
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      double v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = ldexp(u[i][j][r], -eu);
      const int m0 = wave * 32 + i * 16 + 4 * (lane >> 4);
#pragma unroll
      for (int p = 0; p < S; ++p) {
        const int d0 = next_digit(v[0]), d1 = next_digit(v[1]), d2 = next_digit(v[2]), d3 = next_digit(v[3]);
        *reinterpret_cast<int*>(&Us[p][n][m0]) = pack4(d0, d1, d2, d3);
      }
    }
  }
  __syncthreads();

  // Backward: Y_n(jd) += sum_m Kb(jd, m) U_n(m) for the wavefront's dof rows
  // jd in [16*wave, 16*wave + 16) and all 32 elements; the wavefront's Kb
  // slices (16 rows x 128 bytes each) stay in registers.
  v4d y_acc[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) y_acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
  {
    v4i kb[S][kQP / 64];
#pragma unroll
    for (int p = 0; p < S; ++p)
#pragma unroll
      for (int ks = 0; ks < kQP / 64; ++ks)
        kb[p][ks] = *reinterpret_cast<const v4i*>(Kb + (p * kDP + wave * 16 + fr) * kQP + ks * 64 + fk);
#pragma unroll
    for (int t = S - 1; t >= 0; --t) {
      v4i acc[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[j] = v4i{0, 0, 0, 0};
#pragma unroll
      for (int p = 0; p <= t; ++p) {
        const int q = t - p;
        asm volatile("" ::: "memory");
#pragma unroll
        for (int ks = 0; ks < kQP / 64; ++ks)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            const v4i bf = *reinterpret_cast<const v4i*>(&Us[q][j * 16 + fr][ks * 64 + fk]);
            acc[j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(kb[p][ks], bf, acc[j], 0, 0, 0);
          }
      }
      const double w = ldexp(1.0, -7 * (t + 2));
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) y_acc[j][r] += double(acc[j][r]) * w;
    }
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int n = j * 16 + fr;
    const int64_t e = e0 + n;
    if (e >= NE) continue;
    const int eu = exponent_of(fmax(fmax(umax[0][n], umax[1][n]), fmax(umax[2][n], umax[3][n])));
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int jd = wave * 16 + 4 * (lane >> 4) + r;
      if (jd < D3) Yg[e * D3 + jd] += ldexp(y_acc[j][r], eb[jd] + eu);
    }
  }
}

template <int D, int Q, int S>
int launch_mixed(const double* Dq, const double* X, double* Y, int64_t NE, const MixedFemLayout& l, char* base,
                 hipStream_t s) {
  const int64_t grid = ceil_div(NE, kEB);
  if (grid > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  mass3dpa_mixed_kernel<D, Q, S><<<dim3(unsigned(grid)), dim3(256), 0, s>>>(
      reinterpret_cast<const signed char*>(base + l.kf_off), reinterpret_cast<const signed char*>(base + l.kb_off),
      reinterpret_cast<const int*>(base + l.ea_off), reinterpret_cast<const int*>(base + l.eb_off), Dq, X, Y, NE);
  return launch_status();
}

template <int S>
int run_slices(const double* B, const double* Bt, const double* Dq, const double* X, double* Y, int64_t NE, int d1d,
               int q1d, const MixedFemLayout& l, char* base, hipStream_t s) {
  kron_slices_kernel<S><<<1, 256, 0, s>>>(B, Bt, d1d, q1d, reinterpret_cast<signed char*>(base + l.kf_off),
                                          reinterpret_cast<signed char*>(base + l.kb_off),
                                          reinterpret_cast<int*>(base + l.ea_off),
                                          reinterpret_cast<int*>(base + l.eb_off));
  int rc = launch_status();
  if (rc != RPX_SUCCESS) return rc;
  const int key = d1d * 10 + q1d;
  switch (key) {
    case 11: return launch_mixed<1, 1, S>(Dq, X, Y, NE, l, base, s);
    case 12: return launch_mixed<1, 2, S>(Dq, X, Y, NE, l, base, s);
    case 13: return launch_mixed<1, 3, S>(Dq, X, Y, NE, l, base, s);
    case 22: return launch_mixed<2, 2, S>(Dq, X, Y, NE, l, base, s);
    case 23: return launch_mixed<2, 3, S>(Dq, X, Y, NE, l, base, s);
    case 24: return launch_mixed<2, 4, S>(Dq, X, Y, NE, l, base, s);
    case 33: return launch_mixed<3, 3, S>(Dq, X, Y, NE, l, base, s);
    case 34: return launch_mixed<3, 4, S>(Dq, X, Y, NE, l, base, s);
    case 35: return launch_mixed<3, 5, S>(Dq, X, Y, NE, l, base, s);
    case 44: return launch_mixed<4, 4, S>(Dq, X, Y, NE, l, base, s);
    case 45: return launch_mixed<4, 5, S>(Dq, X, Y, NE, l, base, s);
    default: return RPX_ERR_UNSUPPORTED;
  }
}

// ---- DIFFUSION3DPA ------------------------------------------------------------
//
// The diffusion operator (src/apps/DIFFUSION3DPA.hpp:216-458,
// DIFFUSION3DPA-Seq.cpp:28-110) in Kronecker form.  With the effective
// half-storage tables Be(q,d), Ge(q,d) of diffusion_tables (so that bases
// breaking the MFEM symmetry give the Base_Seq operator), element e's
// gradient at the quadrature points is
//   T_c = K_c X_e,  K_0 = Be(x)Be(x)Ge,  K_1 = Be(x)Ge(x)Be,  K_2 = Ge(x)Be(x)Be
// (factors for z, y, x: K_c(m, k) with m = (qz*Q + qy)*Q + qx and
// k = (dz*D + dy)*D + dx), the flux is U_c = sum_c' O_cc' T_c' at every
// point (O symmetric, stored as (11,12,13,22,23,33), or nine components
// row-major, DIFFUSION3DPA.hpp:384-392), and Y_e += sum_c K_c^T U_c.
// Forward and backward products are the MASS3DPA scheme above, per
// component: rows of K_c and of K_c^T scaled by powers of two, X and U per
// element, S digits each, exact int32 levels, FP64 recombination; the flux
// and Y += are FP64.
//
// A workgroup of four wavefronts owns 16 elements.  Wavefront w computes the
// quadrature rows [32w, 32w + 32) of all three components (so the flux at a
// point needs no exchange), then the dof rows [16w, 16w + 16) of the
// backward product.  The slices of K_c (forward) and of K_c^T (backward) of
// the wavefront's rows are loaded into registers one component at a time.
// LDS holds X's slices, then U's slices for the three components (S x 16 x
// 432 bytes, 48 KB at S = 7).  O is read straight from global memory in
// the accumulator layout: 16 consecutive points of one component and
// element per quarter-wavefront, 128 contiguous bytes.
constexpr int kDEB = 16;                // elements per workgroup
constexpr int kDUComp = kQP + 16;       // LDS bytes per component of an element's U slice row
constexpr int kDURow = 3 * kDUComp;

struct DiffMixedLayout {
  size_t kf_off, kb_off, ea_off, eb_off, total;
};
DiffMixedLayout diff_layout(int s) {
  DiffMixedLayout l;
  l.kf_off = 0;                                          // Kf[p][c][m][k]
  l.kb_off = align256(size_t(s) * 3 * kQP * kDP);        // Kb[p][c][j][m]
  l.ea_off = l.kb_off + align256(size_t(s) * 3 * kDP * kQP);
  l.eb_off = l.ea_off + align256(3 * kQP * sizeof(int));  // ea[c][m]
  l.total = l.eb_off + align256(3 * kDP * sizeof(int));   // eb[c][j]
  return l;
}

// One workgroup per component c: work-items 0..127 cut the rows of K_c,
// 128..191 the rows of K_c^T.
template <int S>
__global__ __launch_bounds__(256) void diff_kron_slices_kernel(const double* __restrict__ Bg,
                                                               const double* __restrict__ Gg, int D, int Q,
                                                               signed char* __restrict__ Kf,
                                                               signed char* __restrict__ Kb, int* __restrict__ ea,
                                                               int* __restrict__ eb) {
  __shared__ double sB[64], sG[64];  // effective tables, [q*D + d]
  const int c = blockIdx.x, r = threadIdx.x;
  const int D3 = D * D * D, Q3 = Q * Q * Q;
  if (r < Q * D) diffusion_tables(r / D, r % D, Q, D, Bg, Gg, &sB[r], &sG[r]);
  __syncthreads();
  // factor of axis a (0 = z, 1 = y, 2 = x): Ge on axis 2 - c, Be elsewhere
  auto f = [&](int a, int q, int d) { return a == 2 - c ? sG[q * D + d] : sB[q * D + d]; };
  auto kval = [&](int m, int k) {
    const int qz = m / (Q * Q), qy = (m / Q) % Q, qx = m % Q;
    const int dz = k / (D * D), dy = (k / D) % D, dx = k % D;
    return (f(0, qz, dz) * f(1, qy, dy)) * f(2, qx, dx);
  };
  if (r < kQP) {
    const int m = r;
    double mx = 0.0;
    if (m < Q3)
      for (int k = 0; k < D3; ++k) mx = fmax(mx, fabs(kval(m, k)));
    const int e = exponent_of(mx);
    ea[c * kQP + m] = e;
    for (int k = 0; k < kDP; k += 4) {
      int dig[4][S];
      for (int u = 0; u < 4; ++u) {
        const int kk = k + u;
        cut<S>(ldexp((m < Q3 && kk < D3) ? kval(m, kk) : 0.0, -e), dig[u]);
      }
      for (int p = 0; p < S; ++p)
        *reinterpret_cast<int*>(Kf + ((size_t(p) * 3 + c) * kQP + m) * kDP + k) =
            pack4(dig[0][p], dig[1][p], dig[2][p], dig[3][p]);
    }
  } else if (r < kQP + kDP) {
    const int j = r - kQP;
    double mx = 0.0;
    if (j < D3)
      for (int m = 0; m < Q3; ++m) mx = fmax(mx, fabs(kval(m, j)));
    const int e = exponent_of(mx);
    eb[c * kDP + j] = e;
    for (int m = 0; m < kQP; m += 4) {
      int dig[4][S];
      for (int u = 0; u < 4; ++u) {
        const int mm = m + u;
        cut<S>(ldexp((j < D3 && mm < Q3) ? kval(mm, j) : 0.0, -e), dig[u]);
      }
      for (int p = 0; p < S; ++p)
        *reinterpret_cast<int*>(Kb + ((size_t(p) * 3 + c) * kDP + j) * kQP + m) =
            pack4(dig[0][p], dig[1][p], dig[2][p], dig[3][p]);
    }
  }
}

template <int D, int Q, int S, bool SYMM>
__global__ __launch_bounds__(256) void diffusion3dpa_mixed_kernel(const signed char* __restrict__ Kf,
                                                                  const signed char* __restrict__ Kb,
                                                                  const int* __restrict__ ea,
                                                                  const int* __restrict__ eb,
                                                                  const double* __restrict__ Dg,
                                                                  const double* __restrict__ Xg,
                                                                  double* __restrict__ Yg, int64_t NE) {
  constexpr int D3 = D * D * D, Q3 = Q * Q * Q;
  constexpr int NC = SYMM ? 6 : 9;
  constexpr int kXsBytes = S * kDEB * kXRow;
  constexpr int kUsBytes = S * kDEB * kDURow;
  constexpr int kLdsBytes = kXsBytes > kUsBytes ? kXsBytes : kUsBytes;
  __shared__ __attribute__((aligned(16))) signed char lds[kLdsBytes];
  __shared__ int eX[kDEB];
  __shared__ double umax[4][kDEB];
  auto Xs = reinterpret_cast<signed char(*)[kDEB][kXRow]>(lds);
  auto Us = reinterpret_cast<signed char(*)[kDEB][kDURow]>(lds);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int fr = lane & 15, fk = (lane >> 4) * 16;  // MFMA operand row (element) / first K byte
  const int64_t e0 = int64_t(blockIdx.x) * kDEB;

  // X -> exponent and slices: sixteen work-items per element, four values each.
  {
    const int n = tid >> 4, k0 = (tid & 15) * 4;
    const int64_t e = e0 + n;
    double x[4];
    double mx = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + u;
      x[u] = (e < NE && k < D3) ? Xg[e * D3 + k] : 0.0;
      mx = fmax(mx, fabs(x[u]));
    }
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) mx = fmax(mx, __shfl_xor(mx, off, kWave));
    const int ex = exponent_of(mx);
    if ((tid & 15) == 0) eX[n] = ex;
#pragma unroll
    for (int u = 0; u < 4; ++u) x[u] = ldexp(x[u], -ex);
#pragma unroll
    for (int p = 0; p < S; ++p) {
      const int d0 = next_digit(x[0]), d1 = next_digit(x[1]), d2 = next_digit(x[2]), d3 = next_digit(x[3]);
      *reinterpret_cast<int*>(&Xs[p][n][k0]) = pack4(d0, d1, d2, d3);
    }
  }
  __syncthreads();

  // Forward: T_c(m, n) = sum_k K_c(m, k) X_n(k), rows m in [32w, 32w + 32).
  v4d t_acc[3][2];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int i = 0; i < 2; ++i) t_acc[c][i] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v4i kf[S][2];
#pragma unroll
    for (int p = 0; p < S; ++p)
#pragma unroll
      for (int i = 0; i < 2; ++i)
        kf[p][i] = *reinterpret_cast<const v4i*>(Kf + ((size_t(p) * 3 + c) * kQP + wave * 32 + i * 16 + fr) * kDP +
                                                 fk);
#pragma unroll
    for (int t = S - 1; t >= 0; --t) {
      v4i acc[2] = {v4i{0, 0, 0, 0}, v4i{0, 0, 0, 0}};
#pragma unroll
      for (int p = 0; p <= t; ++p) {
        asm volatile("" ::: "memory");
        const v4i bf = *reinterpret_cast<const v4i*>(&Xs[t - p][fr][fk]);
#pragma unroll
        for (int i = 0; i < 2; ++i) acc[i] = __builtin_amdgcn_mfma_i32_16x16x64_i8(kf[p][i], bf, acc[i], 0, 0, 0);
      }
      const double w = ldexp(1.0, -7 * (t + 2));
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) t_acc[c][i][r] += double(acc[i][r]) * w;
    }
  }

  // Flux U_c = sum_c' O_cc' T_c' at the wavefront's points, and each
  // element's max |U|.  Accumulator element r of lane l is row 4*(l/16)+r,
  // column (element) l%16.
  const int n = fr;
  const int64_t e = e0 + n;
  double u[3][2][4];
  {
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = wave * 32 + i * 16 + 4 * (lane >> 4) + r;
        double u0 = 0.0, u1 = 0.0, u2 = 0.0;
        if (m < Q3 && e < NE) {
          const double g0 = ldexp(t_acc[0][i][r], ea[m] + eX[n]);
          const double g1 = ldexp(t_acc[1][i][r], ea[kQP + m] + eX[n]);
          const double g2 = ldexp(t_acc[2][i][r], ea[2 * kQP + m] + eX[n]);
          const double* o = Dg + e * (NC * Q3) + m;
          const double O11 = o[0], O12 = o[Q3], O13 = o[2 * Q3];
          const double O21 = SYMM ? O12 : o[3 * Q3];
          const double O22 = SYMM ? o[3 * Q3] : o[4 * Q3];
          const double O23 = SYMM ? o[4 * Q3] : o[5 * Q3];
          const double O31 = SYMM ? O13 : o[6 * Q3];
          const double O32 = SYMM ? O23 : o[7 * Q3];
          const double O33 = SYMM ? o[5 * Q3] : o[8 * Q3];
          u0 = ((O11 * g0) + (O12 * g1)) + (O13 * g2);
          u1 = ((O21 * g0) + (O22 * g1)) + (O23 * g2);
          u2 = ((O31 * g0) + (O32 * g1)) + (O33 * g2);
        }
        u[0][i][r] = u0;
        u[1][i][r] = u1;
        u[2][i][r] = u2;
        mx = fmax(mx, fmax(fabs(u0), fmax(fabs(u1), fabs(u2))));
      }
    mx = fmax(mx, __shfl_xor(mx, 16, kWave));
    mx = fmax(mx, __shfl_xor(mx, 32, kWave));
    if (lane < 16) umax[wave][n] = mx;
  }
  __syncthreads();  // every wavefront is done with X's slices and has its max

  // U -> slices, one power-of-two scale per element.
  const int eu = exponent_of(fmax(fmax(umax[0][n], umax[1][n]), fmax(umax[2][n], umax[3][n])));
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      double v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = ldexp(u[c][i][r], -eu);
      const int m0 = wave * 32 + i * 16 + 4 * (lane >> 4);
#pragma unroll
      for (int p = 0; p < S; ++p) {
        const int d0 = next_digit(v[0]), d1 = next_digit(v[1]), d2 = next_digit(v[2]), d3 = next_digit(v[3]);
        *reinterpret_cast<int*>(&Us[p][n][c * kDUComp + m0]) = pack4(d0, d1, d2, d3);
      }
    }
  __syncthreads();

  // Backward: Y_n(j) += sum_c sum_m K_c(m, j) U_c,n(m), dof rows j in [16w, 16w + 16).
  v4d y_acc = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v4i kb[S][kQP / 64];
#pragma unroll
    for (int p = 0; p < S; ++p)
#pragma unroll
      for (int ks = 0; ks < kQP / 64; ++ks)
        kb[p][ks] = *reinterpret_cast<const v4i*>(Kb + ((size_t(p) * 3 + c) * kDP + wave * 16 + fr) * kQP +
                                                  ks * 64 + fk);
    v4d yc = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = S - 1; t >= 0; --t) {
      v4i acc = v4i{0, 0, 0, 0};
#pragma unroll
      for (int p = 0; p <= t; ++p) {
        asm volatile("" ::: "memory");
#pragma unroll
        for (int ks = 0; ks < kQP / 64; ++ks) {
          const v4i bf = *reinterpret_cast<const v4i*>(&Us[t - p][fr][c * kDUComp + ks * 64 + fk]);
          acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(kb[p][ks], bf, acc, 0, 0, 0);
        }
      }
      const double w = ldexp(1.0, -7 * (t + 2));
#pragma unroll
      for (int r = 0; r < 4; ++r) yc[r] += double(acc[r]) * w;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) y_acc[r] += ldexp(yc[r], eb[c * kDP + wave * 16 + 4 * (lane >> 4) + r]);
  }
  if (e < NE) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = wave * 16 + 4 * (lane >> 4) + r;
      if (j < D3) Yg[e * D3 + j] += ldexp(y_acc[r], eu);
    }
  }
}

template <int D, int S, bool SYMM>
int launch_diff_mixed(const double* Dq, const double* X, double* Y, int64_t NE, const DiffMixedLayout& l,
                      char* base, hipStream_t s) {
  const int64_t grid = ceil_div(NE, kDEB);
  if (grid > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  diffusion3dpa_mixed_kernel<D, D + 1, S, SYMM><<<dim3(unsigned(grid)), dim3(256), 0, s>>>(
      reinterpret_cast<const signed char*>(base + l.kf_off), reinterpret_cast<const signed char*>(base + l.kb_off),
      reinterpret_cast<const int*>(base + l.ea_off), reinterpret_cast<const int*>(base + l.eb_off), Dq, X, Y, NE);
  return launch_status();
}

template <int S, bool SYMM>
int run_diff_order(const double* Dq, const double* X, double* Y, int64_t NE, int d1d, const DiffMixedLayout& l,
                   char* base, hipStream_t s) {
  switch (d1d) {
    case 1: return launch_diff_mixed<1, S, SYMM>(Dq, X, Y, NE, l, base, s);
    case 2: return launch_diff_mixed<2, S, SYMM>(Dq, X, Y, NE, l, base, s);
    case 3: return launch_diff_mixed<3, S, SYMM>(Dq, X, Y, NE, l, base, s);
    case 4: return launch_diff_mixed<4, S, SYMM>(Dq, X, Y, NE, l, base, s);
    default: return RPX_ERR_UNSUPPORTED;
  }
}

template <int S>
int run_diff_slices(const double* Bg, const double* Gg, const double* Dq, const double* X, double* Y, int64_t NE,
                    int d1d, int symmetric, const DiffMixedLayout& l, char* base, hipStream_t s) {
  diff_kron_slices_kernel<S><<<3, 256, 0, s>>>(Bg, Gg, d1d, d1d + 1, reinterpret_cast<signed char*>(base + l.kf_off),
                                               reinterpret_cast<signed char*>(base + l.kb_off),
                                               reinterpret_cast<int*>(base + l.ea_off),
                                               reinterpret_cast<int*>(base + l.eb_off));
  if (const int rc = launch_status(); rc != RPX_SUCCESS) return rc;
  return symmetric ? run_diff_order<S, true>(Dq, X, Y, NE, d1d, l, base, s)
                   : run_diff_order<S, false>(Dq, X, Y, NE, d1d, l, base, s);
}

}  // namespace

extern "C" {

size_t rpx_mass3dpa_mixed_scratch_bytes(int d1d, int q1d, int nslices) {
  const int s = nslices == 0 ? 7 : nslices;
  if (d1d < 1 || d1d > 4 || q1d < d1d || q1d > 5 || q1d > d1d + 2 || s < 1 || s > kMaxSlices) return 0;
  return fem_layout(s).total;
}

int rpx_mass3dpa_mixed_f64(const double* B, const double* Bt, const double* D, const double* X, double* Y,
                           int64_t NE, int d1d, int q1d, int nslices, void* scratch, size_t scratch_bytes,
                           void* stream) {
  if (NE < 0 || nslices < 0 || nslices > kMaxSlices) return RPX_ERR_INVALID_ARGUMENT;
  if (NE == 0) return RPX_SUCCESS;
  if (!B || !Bt || !D || !X || !Y || !scratch) return RPX_ERR_INVALID_ARGUMENT;
  if (d1d < 1 || d1d > 4 || q1d < d1d || q1d > 5 || q1d > d1d + 2) return RPX_ERR_UNSUPPORTED;
  const int s = nslices == 0 ? 7 : nslices;
  const MixedFemLayout l = fem_layout(s);
  if (scratch_bytes < l.total) return RPX_ERR_SCRATCH_TOO_SMALL;
  char* base = reinterpret_cast<char*>(scratch);
  hipStream_t st = as_stream(stream);
  switch (s) {
    case 1: return run_slices<1>(B, Bt, D, X, Y, NE, d1d, q1d, l, base, st);
    case 2: return run_slices<2>(B, Bt, D, X, Y, NE, d1d, q1d, l, base, st);
    case 3: return run_slices<3>(B, Bt, D, X, Y, NE, d1d, q1d, l, base, st);
    case 4: return run_slices<4>(B, Bt, D, X, Y, NE, d1d, q1d, l, base, st);
    case 5: return run_slices<5>(B, Bt, D, X, Y, NE, d1d, q1d, l, base, st);
    case 6: return run_slices<6>(B, Bt, D, X, Y, NE, d1d, q1d, l, base, st);
    case 7: return run_slices<7>(B, Bt, D, X, Y, NE, d1d, q1d, l, base, st);
    case 8: return run_slices<8>(B, Bt, D, X, Y, NE, d1d, q1d, l, base, st);
    default: return RPX_ERR_UNSUPPORTED;
  }
}

size_t rpx_diffusion3dpa_mixed_scratch_bytes(int d1d, int q1d, int nslices) {
  const int s = nslices == 0 ? 7 : nslices;
  if (d1d < 1 || d1d > 4 || q1d != d1d + 1 || s < 5 || s > kMaxSlices) return 0;
  return diff_layout(s).total;
}

int rpx_diffusion3dpa_mixed_f64(const double* Basis, const double* dBasis, const double* D, const double* X,
                                double* Y, int64_t NE, int d1d, int q1d, int symmetric, int nslices, void* scratch,
                                size_t scratch_bytes, void* stream) {
  if (NE < 0 || nslices < 0 || nslices > kMaxSlices) return RPX_ERR_INVALID_ARGUMENT;
  if (NE == 0) return RPX_SUCCESS;
  if (!Basis || !dBasis || !D || !X || !Y || !scratch) return RPX_ERR_INVALID_ARGUMENT;
  if (d1d < 1 || d1d > 4 || q1d != d1d + 1) return RPX_ERR_UNSUPPORTED;
  const int s = nslices == 0 ? 7 : nslices;
  const DiffMixedLayout l = diff_layout(s);
  if (scratch_bytes < l.total) return RPX_ERR_SCRATCH_TOO_SMALL;
  char* base = reinterpret_cast<char*>(scratch);
  hipStream_t st = as_stream(stream);
  switch (s) {
    case 5: return run_diff_slices<5>(Basis, dBasis, D, X, Y, NE, d1d, symmetric, l, base, st);
    case 6: return run_diff_slices<6>(Basis, dBasis, D, X, Y, NE, d1d, symmetric, l, base, st);
    case 7: return run_diff_slices<7>(Basis, dBasis, D, X, Y, NE, d1d, symmetric, l, base, st);
    case 8: return run_diff_slices<8>(Basis, dBasis, D, X, Y, NE, d1d, symmetric, l, base, st);
    default: return RPX_ERR_UNSUPPORTED;
  }
}

}  // extern "C"
