// Mixed-precision GEMM with FP64 accuracy from int8 matrix cores.
//
// Same result semantics as the suite's GEMM (src/polybench/POLYBENCH_GEMM.hpp
// :12-50: C = alpha*A*B, beta never reaches C).  The product is emulated
// with the modular integer scheme of Ozaki, Uchino & Imamura ("Ozaki scheme
// II", 2025): one int8 GEMM per modulus instead of one per pair of
// fixed-point slices.
//   * Row i of A is scaled by 2^(b-e_i) and column j of B by 2^(b-f_j),
//     where 2^(e_i-1) <= max_k |A[i][k]| < 2^e_i (likewise f_j), and rounded
//     to integers: |A'|, |B'| <= 2^b.  X = A'B' is an exact integer with
//     |X| <= nk 4^b.
//   * T pairwise coprime moduli m_t <= 256 (the largest such numbers) with
//     M = prod m_t > 2 nk 4^b.  For each t the symmetric residues of A' and
//     B' lie in [-128, 128] (128 only for m_t = 256, where it is stored as
//     the byte -128, the same residue) and one GEMM on
//     v_mfma_i32_16x16x64_i8 gives their product exactly in int32 (nk <=
//     131071); its epilogue keeps d_t = that product modulo m_t, a
//     symmetric residue of X, one byte per output.
//   * Chinese remainder theorem: X = sum_t d_t y_t M/m_t (mod M), y_t =
//     (M/m_t)^-1 mod m_t, so X/M = sum_t d_t y_t/m_t minus the nearest
//     integer, |X/M| < 1/2.  With y_t/m_t = h_t + l_t, h_t a multiple of
//     2^-44, the sum of d_t h_t is kept exactly in FP64 modulo 1 and the sum
//     of d_t l_t (below T 2^-36) is added after the integer part is removed,
//     so X comes out with a relative error of a few 2^-53, plus an absolute
//     error below 2^-84 M, however much it cancels.
//   * C[i][j] = alpha 2^(e_i+f_j-2b) X.
// The only approximation is the rounding of A and B to b bits below their
// row/column maxima: |C - alpha AB|_ij <= |alpha| 2^(e_i+f_j-b) nk (1+2^-b)
// plus FP64 rounding of the result.  b is the largest value with
// 2 nk 4^b < M, at most 46 (so every quotient below is exact, see
// sym_mod); by default T is the smallest count giving b >= 42 (13 moduli,
// b = 44 at nk = 4096), an error below 2^-42 relative to
// |alpha| max_k|A[i][k]| max_k|B[k][j]| nk.
//
// Scratch holds the int8 residues (A: [t][i][k], B: [t][j][k], K
// contiguous, zero-padded to tile multiples), the per-row/column exponents
// and the residue bytes d_t ([t] then the GEMM's accumulator fragments).
#include "rpx_common.hpp"

#include <mutex>

using namespace rpx;

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kMaxMods = 16;
// The 16 largest pairwise coprime integers not above 256.
constexpr int kMods[kMaxMods] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193};
constexpr int kMinBits = 42;       // automatic modulus count: the smallest T with b >= 42
constexpr int kMaxBits = 46;       // |A'|, |B'| <= 2^46: exact quotients in sym_mod
constexpr int64_t kMaxK = 131071;  // 128 * 128 * nk < 2^31

constexpr int kTM = 256, kTN = 256, kTK = 128;  // block tile, K bytes per stage
constexpr int kChunks = kTK / 16;               // 16-byte chunks per tile row and stage

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Per-call constants of the modular scheme, passed by value to the kernels.
struct CrtParams {
  int T, b;
  double m[kMaxMods];    // moduli
  double inv[kMaxMods];  // fl(1/m_t), used only to pick a quotient that is then corrected
  double h[kMaxMods];    // floor(2^44 y_t/m_t) 2^-44, y_t = (M/m_t)^-1 mod m_t
  double l[kMaxMods];    // y_t/m_t - h_t, rounded
  double M_hi, M_lo;     // prod m_t = M_hi + M_lo, M_lo rounded (a 2^-106 relative error)
};

typedef unsigned __int128 u128;

int64_t inverse_mod(int64_t a, int64_t m) {
  int64_t t = 0, nt = 1, r = m, nr = a % m;
  while (nr != 0) {
    const int64_t q = r / nr;
    int64_t tmp = t - q * nt; t = nt; nt = tmp;
    tmp = r - q * nr; r = nr; nr = tmp;
  }
  return t < 0 ? t + m : t;
}

// Largest b <= kMaxBits with 2 nk 4^b < M, or -1.
int bits_for(int64_t nk, int T) {
  u128 M = 1;
  for (int t = 0; t < T; ++t) M *= u128(kMods[t]);
  const u128 k = u128(nk > 0 ? nk : 1);
  for (int b = kMaxBits; b >= 0; --b)
    if (2 * k * (u128(1) << (2 * b)) < M) return b;
  return -1;
}

int moduli_for(int64_t nk, int nmoduli) {
  if (nmoduli > 0) return nmoduli;
  for (int T = 1; T <= kMaxMods; ++T)
    if (bits_for(nk, T) >= kMinBits) return T;
  return kMaxMods;
}

bool crt_setup(int64_t nk, int nmoduli, CrtParams* p) {
  const int T = moduli_for(nk, nmoduli);
  if (T < 1 || T > kMaxMods) return false;
  const int b = bits_for(nk, T);
  if (b < 1) return false;
  p->T = T;
  p->b = b;
  u128 M = 1;
  for (int t = 0; t < T; ++t) M *= u128(kMods[t]);
  // M < 2^126 fits long double's range; hi is an integer, so hiu = hi exactly
  // and M - hi (below 2^73 in magnitude) is exact in 128-bit integers
  // before its rounding to FP64.
  const double hi = double(static_cast<long double>(M));
  const u128 hiu = static_cast<u128>(static_cast<long double>(hi));
  p->M_hi = hi;
  p->M_lo = hiu >= M ? -double(static_cast<long double>(hiu - M)) : double(static_cast<long double>(M - hiu));
  for (int t = 0; t < kMaxMods; ++t) {
    const int64_t m = kMods[t];
    p->m[t] = double(m);
    p->inv[t] = 1.0 / double(m);
    int64_t mt = 1;  // (M / m_t) mod m_t
    for (int s = 0; s < T; ++s)
      if (s != t) mt = mt * kMods[s] % m;
    const uint64_t y = t < T ? uint64_t(inverse_mod(mt, m)) : 0;
    const uint64_t num = y << 44;  // y < 256: below 2^52
    p->h[t] = std::ldexp(double(num / uint64_t(m)), -44);
    p->l[t] = std::ldexp(double(num % uint64_t(m)) / double(m), -44);
  }
  return true;
}

struct MixedLayout {
  int64_t nip, njp, nkp;
  size_t a_off, b_off, ea_off, fb_off, r_off, total;
};
MixedLayout mixed_layout(int64_t ni, int64_t nj, int64_t nk, int T) {
  MixedLayout l;
  l.nip = round_up(ni, kTM);
  l.njp = round_up(nj, kTN);
  l.nkp = round_up(nk > 0 ? nk : 1, 2 * kTK);  // the K loop takes two stages per step
  l.a_off = 0;
  l.b_off = align256(size_t(T) * l.nip * l.nkp);
  l.ea_off = l.b_off + align256(size_t(T) * l.njp * l.nkp);
  l.fb_off = l.ea_off + align256(size_t(l.nip) * 4);
  l.r_off = l.fb_off + align256(size_t(l.njp) * 4);
  l.total = l.r_off + align256(size_t(T) * l.nip * l.njp);  // d_t, one byte per output
  return l;
}

__device__ __forceinline__ int exponent_of(double m) {
  int e = 0;
  if (m != 0.0) frexp(m, &e);
  return e;
}

// Symmetric residue of the integer-valued x modulo m, |x| <= 2^46 (or
// |x| < 2^31): fl(x fl(1/m)) is within 2^-13 of x/m, and x/m is at least
// 1/(2m) >= 2^-9 from a half-integer unless m = 256 and x = 128 (mod 256).
// So q = rint(x/m) exactly, the fma is exact, and the result lies in
// [-(m-1)/2, (m-1)/2], or in [-128, 128] for m = 256, where the byte kept
// (r & 0xFF) makes 128 the same residue as -128.
__device__ __forceinline__ int sym_mod(double x, double m, double inv) {
  return int(fma(-rint(x * inv), m, x));
}

// One workgroup per (padded) row of A: row exponent, then the residues.
// Each work-item takes 16 consecutive k and writes one 16-byte chunk per
// modulus.
__global__ __launch_bounds__(256) void split_rows_kernel(const double* __restrict__ A, int64_t ni, int64_t nk,
                                                         int64_t nkp, int64_t nip, signed char* __restrict__ As,
                                                         int* __restrict__ ea, const CrtParams P) {
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
  for (int64_t k0 = int64_t(tid) * 16; k0 < nkp; k0 += 256 * 16) {
    double x[16];
    if (live && k0 + 16 <= nk) {
#pragma unroll
      for (int u = 0; u < 16; ++u) x[u] = row[k0 + u];
    } else {
#pragma unroll
      for (int u = 0; u < 16; ++u) x[u] = (live && k0 + u < nk) ? row[k0 + u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) x[u] = rint(ldexp(x[u], P.b - e));
    for (int t = 0; t < P.T; ++t) {
      v4i w;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        int packed = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = sym_mod(x[4 * c + q], P.m[t], P.inv[t]);
          packed |= (r & 0xFF) << (8 * q);
        }
        w[c] = packed;
      }
      *reinterpret_cast<v4i*>(As + (int64_t(t) * nip + i) * nkp + k0) = w;
    }
  }
}

// Column maxima of B as exponents.  A workgroup takes 64 columns x
// kColExpRows rows: wavefront w reads rows w, w + 4, ... (each load is 64
// consecutive columns of one row, 16 loads in flight per work-item), the
// four wavefronts' maxima meet in LDS, and one atomicMax per column and
// workgroup (nk / 512 per column) publishes the exponent.
constexpr int kColExpRows = 512;
__global__ __launch_bounds__(256) void col_exponent_kernel(const double* __restrict__ B, int64_t nk, int64_t nj,
                                                           int* __restrict__ fb) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t j = int64_t(blockIdx.x) * 64 + lane;
  const int64_t k0 = int64_t(blockIdx.y) * kColExpRows;
  const int64_t ke = (k0 + kColExpRows < nk) ? k0 + kColExpRows : nk;
  double m = 0.0;
  if (j < nj) {
    int64_t k = k0 + w;
    for (; k + 4 * 15 < ke; k += 4 * 16) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = B[(k + 4 * u) * nj + j];
#pragma unroll
      for (int u = 0; u < 16; ++u) m = fmax(m, fabs(v[u]));
    }
    for (; k < ke; k += 4) m = fmax(m, fabs(B[k * nj + j]));
  }
  red[w][lane] = m;
  __syncthreads();
  if (w == 0 && j < nj) {
    m = fmax(fmax(red[0][lane], red[1][lane]), fmax(red[2][lane], red[3][lane]));
    if (m != 0.0) atomicMax(&fb[j], exponent_of(m));
  }
}

__global__ void fill_int_kernel(int* p, int64_t n, int v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// B -> Bs[t][j][k]: a workgroup owns 64 columns x 128 k; work-item (w, j)
// reads k0 + 16w .. k0 + 16w + 15 of column j (each load is 64 consecutive
// columns of one row) and, per modulus, packs those 16 residues into one
// 16-byte LDS write at row j, chunk w (rows padded to 144 bytes: the 16
// writes of a quarter-wavefront fall in distinct bank groups).  Eight
// work-items then read one 128-byte row back and store it as a whole line
// of Bs, so every store instruction writes eight whole lines.
constexpr int kColPitch = 144;  // bytes per LDS row of the B transpose
__global__ __launch_bounds__(512) void split_cols_kernel(const double* __restrict__ B, int64_t nk, int64_t nj,
                                                         int64_t nkp, int64_t njp, const int* __restrict__ fb,
                                                         signed char* __restrict__ Bs, const CrtParams P) {
  __shared__ __attribute__((aligned(16))) signed char lds[2][64 * kColPitch];
  const int tid = threadIdx.x;
  const int64_t j0 = int64_t(blockIdx.x) * 64;
  const int64_t j = j0 + (tid & 63);
  const int64_t kb = int64_t(blockIdx.y) * 128;
  const int64_t k0 = kb + 16 * (tid >> 6);
  const int f = (j < nj) ? fb[j] : 0;
  const int sh = P.b - (f == INT32_MIN ? 0 : f);
  double x[16];
  if (j < nj && k0 + 16 <= nk) {
#pragma unroll
    for (int u = 0; u < 16; ++u) x[u] = B[(k0 + u) * nj + j];
  } else {
#pragma unroll
    for (int u = 0; u < 16; ++u) x[u] = (j < nj && k0 + u < nk) ? B[(k0 + u) * nj + j] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) x[u] = rint(ldexp(x[u], sh));
  const int jj = tid & 63, w16 = 16 * (tid >> 6);  // this work-item's LDS row and byte offset in it
  const int rrow = tid >> 3, rpart = 16 * (tid & 7);  // the row chunk it stores
  for (int t = 0; t < P.T; ++t) {
    v4i w;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      int packed = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = sym_mod(x[4 * c + q], P.m[t], P.inv[t]);
        packed |= (r & 0xFF) << (8 * q);
      }
      w[c] = packed;
    }
    signed char* buf = lds[t & 1];
    *reinterpret_cast<v4i*>(buf + jj * kColPitch + w16) = w;
    __syncthreads();  // the row is complete; buffer t & 1 was last read two moduli ago, before the previous barrier
    const v4i row = *reinterpret_cast<const v4i*>(buf + rrow * kColPitch + rpart);
    *reinterpret_cast<v4i*>(Bs + (int64_t(t) * njp + j0 + rrow) * nkp + kb + rpart) = row;
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
// grid of 16x16 int32 accumulators).  For each modulus the K loop runs over
// that modulus's residues, and the epilogue reduces each exact int32 sum
// modulo m_t and stores the byte.  The tile size is set by L2 bandwidth: a
// stage brings 128 bytes of K for 256 rows of A and of B (64 KiB) for
// 8.4 M MACs, 128 MACs per byte, where a 128x128 tile's 64 left the MFMAs
// waiting on L2.  Two stages of LDS (128 KiB), one workgroup per CU.
constexpr int kThreads = 512;
constexpr int kWaveM = 64, kWaveN = 128;
constexpr int kFm = kWaveM / 16, kFn = kWaveN / 16;  // 16x16 tiles per wavefront
constexpr int kFrags = (kThreads / kWave) * kFm * kFn;  // accumulator fragments per workgroup tile
static_assert(kTM == kTN && kTM == 256 && kChunks == 8, "the stage loader covers 256 rows of 8 chunks");

typedef __attribute__((address_space(3))) void lds_void;
typedef const __attribute__((address_space(1))) void global_void;

// Stages go from global memory straight into LDS (global_load_lds, 16
// bytes per lane: each instruction fills 64 consecutive slots, i.e. 8
// swizzled rows, so lane l loads the logical chunk that swz() puts in slot
// l), one stage ahead of the one being computed.  All stages of all moduli
// form one sequence g = t * ksteps + st, so the next modulus's first stage
// is loaded during the current modulus's last.  The two stage buffers are
// separate arrays and the loop is unrolled by two, so every LDS access
// names its buffer: the compiler then sees that the loads in flight cannot
// write what the fragment reads read, and adds no wait for them; the wait
// for a stage is explicit (vmcnt(0), then the barrier that also frees the
// other buffer).  Against register-staged stages (loads into registers,
// then LDS stores a step later) this saves 32 VGPRs and the LDS store
// instructions: 1.05-1.06 against 1.08-1.10 ms per 4096^3 call.
__global__ __launch_bounds__(kThreads) void ozaki_gemm_kernel(unsigned* __restrict__ R,
                                                              const signed char* __restrict__ As,
                                                              const signed char* __restrict__ Bs, int64_t nip,
                                                              int64_t njp, int64_t nkp, int tiles_n,
                                                              const CrtParams P) {
  __shared__ v4i A0[kTM * kChunks], A1[kTM * kChunks];
  __shared__ v4i B0[kTN * kChunks], B1[kTN * kChunks];
  int64_t tm, tn;
  tile_of(blockIdx.x, gridDim.x, int(nip / kTM), tiles_n, tm, tn);
  const int64_t m0 = tm * kTM, n0 = tn * kTN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * kWaveM, wn = (wave & 1) * kWaveN;
  const int ksteps = int(nkp / kTK);  // even: nkp is a multiple of 2 kTK
  const int total = P.T * ksteps;     // stages over all moduli
  const int fr = lane & 15, fk = lane >> 4;  // MFMA operand row and k quarter
  // This wavefront's fragments in the residue array: [tile][wave][i][j][lane].
  const int64_t frag0 = ((tm * tiles_n + tn) * (kThreads / kWave) + wave) * (kFm * kFn);
  const int64_t per_mod = nip * njp / 4;  // 32-bit words per modulus
  // Loader: instruction u (0..3) of wavefront w fills slots 64 (4w + u) ..
  // + 63: row drow0 + 8u (drow0 = 32w + lane/8), physical chunk lane%8,
  // which holds logical chunk (lane%8) ^ ((row >> 1) & 7).
  const int drow0 = 32 * wave + (lane >> 3);
  int dchunk[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) dchunk[u] = (lane & 7) ^ (((drow0 + 8 * u) >> 1) & 7);
  const signed char* a0 = As + (m0 + drow0) * nkp;
  const signed char* b0 = Bs + (n0 + drow0) * nkp;
  const int64_t a_mod = nip * nkp, b_mod = njp * nkp;
  auto load_stage = [&](int g, v4i* Ad, v4i* Bd) {  // stage g: modulus g / ksteps, K step g % ksteps
    const int tt = g / ksteps, st = g - tt * ksteps;
    const signed char* at = a0 + tt * a_mod + st * kTK;
    const signed char* bt = b0 + tt * b_mod + st * kTK;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      __builtin_amdgcn_global_load_lds((global_void*)(at + int64_t(8 * u) * nkp + 16 * dchunk[u]),
                                       (lds_void*)(Ad + 64 * (4 * wave + u)), 16, 0, 0);
      __builtin_amdgcn_global_load_lds((global_void*)(bt + int64_t(8 * u) * nkp + 16 * dchunk[u]),
                                       (lds_void*)(Bd + 64 * (4 * wave + u)), 16, 0, 0);
    }
  };
  v4i acc[kFm][kFn];
  auto compute = [&](const v4i* Ar, const v4i* Br) {
    // Raised wave priority over the MFMA block: with two wavefronts per
    // SIMD the arbiter then keeps feeding the matrix pipe from one while
    // the other issues its loads.
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int kk = 0; kk < kChunks / 4; ++kk) {
      const int c = kk * 4 + fk;
      v4i af[kFm], bf[kFn];
#pragma unroll
      for (int i = 0; i < kFm; ++i) af[i] = Ar[swz(wm + i * 16 + fr, c)];
#pragma unroll
      for (int j = 0; j < kFn; ++j) bf[j] = Br[swz(wn + j * 16 + fr, c)];
#pragma unroll
      for (int i = 0; i < kFm; ++i)
#pragma unroll
        for (int j = 0; j < kFn; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
      // The backend's LDS-read/MFMA interleave for small GEMM loops: 1.056-1.063
      // against 1.085-1.093 ms per 4096^3 call, alternating runs (an explicit
      // one LDS read per two MFMAs was slower, 1.10 ms).
      __builtin_amdgcn_iglp_opt(0);
    }
    __builtin_amdgcn_s_setprio(0);
  };
  // This wavefront's stage loads have landed (vmcnt(0)); after the barrier
  // every wavefront's have, and every wavefront has finished reading the
  // other buffer (its reads fed the MFMAs issued before the barrier).
  auto stage_ready = [&]() {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  };

  load_stage(0, A0, B0);
  int g = 0;
  for (int t = 0; t < P.T; ++t) {
#pragma unroll
    for (int i = 0; i < kFm; ++i)
#pragma unroll
      for (int j = 0; j < kFn; ++j) acc[i][j] = v4i{0, 0, 0, 0};
    for (int st = 0; st < ksteps; st += 2, g += 2) {
      stage_ready();  // stage g in buffer 0
      if (g + 1 < total) load_stage(g + 1, A1, B1);
      compute(A0, B0);
      stage_ready();  // stage g + 1 in buffer 1
      if (g + 2 < total) load_stage(g + 2, A0, B0);
      compute(A1, B1);
    }
    // d_t = the sum's symmetric residue modulo m_t (|sum| < 2^31, see
    // sym_mod); the four rows of a lane's accumulator go to one 32-bit word.
    const double m = P.m[t], inv = P.inv[t];
    unsigned* rt = R + int64_t(t) * per_mod + frag0 * kWave + lane;
#pragma unroll
    for (int i = 0; i < kFm; ++i)
#pragma unroll
      for (int j = 0; j < kFn; ++j) {
        unsigned w = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r) w |= unsigned(sym_mod(double(acc[i][j][r]), m, inv) & 0xFF) << (8 * r);
        rt[(i * kFn + j) * kWave] = w;
      }
  }
}

// One work-item per (fragment, lane): the four outputs of that accumulator
// slot.  X/M = sum_t d_t (h_t + l_t) mod 1; the h_t part is exact (every
// term is a multiple of 2^-44 below 128 in magnitude and the sum is reduced
// modulo 1 after each add).
__global__ __launch_bounds__(256) void crt_combine_kernel(double* __restrict__ C, const unsigned* __restrict__ R,
                                                          const int* __restrict__ ea, const int* __restrict__ fb,
                                                          double alpha, double beta, int use_beta, int64_t ni,
                                                          int64_t nj, int64_t nip, int64_t njp, int tiles_n,
                                                          const CrtParams P) {
  const int64_t g = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int64_t per_mod = nip * njp / 4;
  if (g >= per_mod) return;
  const int lane = int(g % kWave);
  const int64_t frag = g / kWave;
  const int64_t tile = frag / kFrags;
  const int f = int(frag % kFrags);
  const int wave = f / (kFm * kFn), i = (f / kFn) % kFm, j = f % kFn;
  const int64_t row0 = (tile / tiles_n) * kTM + (wave >> 1) * kWaveM + i * 16 + 4 * (lane >> 4);
  const int64_t col = (tile % tiles_n) * kTN + (wave & 1) * kWaveN + j * 16 + (lane & 15);
  // All residue words first (T <= 16), so their loads are in flight together.
  unsigned wv[kMaxMods];
#pragma unroll
  for (int t = 0; t < kMaxMods; ++t) wv[t] = t < P.T ? R[int64_t(t) * per_mod + g] : 0u;
  double fh[4] = {0.0, 0.0, 0.0, 0.0}, fl[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < kMaxMods; ++t) {
    if (t >= P.T) break;
    const double h = P.h[t], l = P.l[t];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double d = double(int(wv[t] << (24 - 8 * r)) >> 24);  // the signed byte r
      const double s = fma(d, h, fh[r]);  // exact: multiples of 2^-44 below 129 in magnitude
      fh[r] = s - floor(s);
      fl[r] = fma(d, l, fl[r]);
    }
  }
  if (col >= nj) return;
  const int fj = fb[col];
  const int sc = (fj == INT32_MIN ? 0 : fj) - 2 * P.b;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int64_t row = row0 + r;
    if (row >= ni) break;
    // X = (fx + fl) M with fx = fh - nearest integer (exact, a multiple of
    // 2^-52 in [-1/2, 1/2]) and M = M_hi + M_lo: the product fx M_hi is
    // split exactly by an fma, the small terms are added to its error.
    double fx = fh[r] - rint(fh[r]);
    if (fx + fl[r] > 0.5) fx -= 1.0;
    if (fx + fl[r] < -0.5) fx += 1.0;
    const double p = fx * P.M_hi;
    const double e = fma(fx, P.M_hi, -p);
    const double x = p + (e + fma(fx, P.M_lo, fl[r] * P.M_hi));
    const double v = alpha * ldexp(x, ea[row] + sc);
    double* c = C + row * nj + col;
    *c = use_beta ? v + beta * *c : v;
  }
}

// A's row split does not depend on B's column exponents or B's split, so it
// runs on a library-owned auxiliary stream (non-blocking, default
// priority), forked from and joined back to the caller's stream with events;
// stream capture follows the fork and join.  One auxiliary stream and event
// pair per device, created on first use; the mutex keeps concurrent callers
// from recording the shared events at the same time.
struct AuxStream {
  hipStream_t aux = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
constexpr int kMaxDevices = 64;
std::mutex g_aux_mu;
AuxStream g_aux[kMaxDevices];

int aux_for_device(AuxStream** out) {
  int dev = -1;
  RPX_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return RPX_ERR_UNSUPPORTED;
  AuxStream& a = g_aux[dev];
  if (!a.aux) {
    RPX_TRY(hipStreamCreateWithFlags(&a.aux, hipStreamNonBlocking));
    RPX_TRY(hipEventCreateWithFlags(&a.fork, hipEventDisableTiming));
    RPX_TRY(hipEventCreateWithFlags(&a.join, hipEventDisableTiming));
  }
  *out = &a;
  return RPX_SUCCESS;
}

}  // namespace

extern "C" {

size_t rpx_gemm_mixed_scratch_bytes(int64_t ni, int64_t nj, int64_t nk, int nmoduli) {
  if (ni < 0 || nj < 0 || nk < 0 || nmoduli < 0 || nmoduli > kMaxMods || nk > kMaxK) return 0;
  CrtParams P;
  if (!crt_setup(nk, nmoduli, &P)) return 0;
  return mixed_layout(ni, nj, nk, P.T).total;
}

int rpx_gemm_mixed_bits(int64_t nk, int nmoduli) {
  if (nk < 0 || nk > kMaxK || nmoduli < 0 || nmoduli > kMaxMods) return RPX_ERR_INVALID_ARGUMENT;
  CrtParams P;
  if (!crt_setup(nk, nmoduli, &P)) return RPX_ERR_UNSUPPORTED;
  return P.b;
}

int rpx_gemm_mixed_f64(double* C, const double* A, const double* B, double alpha, double beta, int64_t ni,
                       int64_t nj, int64_t nk, int nmoduli, void* scratch, size_t scratch_bytes, void* stream) {
  return rpx_gemm_mixed_tuned_f64(C, A, B, alpha, beta, ni, nj, nk, nmoduli, nullptr, scratch, scratch_bytes,
                                  stream);
}

int rpx_gemm_mixed_tuned_f64(double* C, const double* A, const double* B, double alpha, double beta, int64_t ni,
                             int64_t nj, int64_t nk, int nmoduli, const rpx_tuning* tune, void* scratch,
                             size_t scratch_bytes, void* stream) {
  if (ni < 0 || nj < 0 || nk < 0 || nmoduli < 0 || nmoduli > kMaxMods) return RPX_ERR_INVALID_ARGUMENT;
  if (ni == 0 || nj == 0) return RPX_SUCCESS;
  if (!C || !scratch || (nk > 0 && (!A || !B))) return RPX_ERR_INVALID_ARGUMENT;
  if (nk > kMaxK) return RPX_ERR_UNSUPPORTED;
  CrtParams P;
  if (!crt_setup(nk, nmoduli, &P)) return RPX_ERR_UNSUPPORTED;
  const MixedLayout l = mixed_layout(ni, nj, nk, P.T);
  if (scratch_bytes < l.total) return RPX_ERR_SCRATCH_TOO_SMALL;
  const int64_t tm = l.nip / kTM, tn = l.njp / kTN;
  if (l.nip > INT32_MAX || tm * tn > INT32_MAX || ceil_div(l.nip * l.njp / 4, 256) > INT32_MAX)
    return RPX_ERR_UNSUPPORTED;
  char* base = reinterpret_cast<char*>(scratch);
  signed char* As = reinterpret_cast<signed char*>(base + l.a_off);
  signed char* Bs = reinterpret_cast<signed char*>(base + l.b_off);
  int* ea = reinterpret_cast<int*>(base + l.ea_off);
  int* fb = reinterpret_cast<int*>(base + l.fb_off);
  unsigned* R = reinterpret_cast<unsigned*>(base + l.r_off);
  hipStream_t s = as_stream(stream);
  std::lock_guard<std::mutex> lock(g_aux_mu);
  AuxStream* ax = nullptr;
  if (const int st = aux_for_device(&ax); st != RPX_SUCCESS) return st;
  RPX_TRY(hipEventRecord(ax->fork, s));
  RPX_TRY(hipStreamWaitEvent(ax->aux, ax->fork, 0));
  split_rows_kernel<<<dim3(unsigned(l.nip)), dim3(256), 0, ax->aux>>>(A, ni, nk, l.nkp, l.nip, As, ea, P);
  if (const int st = launch_status(); st != RPX_SUCCESS) return st;
  RPX_TRY(hipEventRecord(ax->join, ax->aux));
  fill_int_kernel<<<dim3(unsigned(ceil_div(l.njp, 256))), dim3(256), 0, s>>>(fb, l.njp, INT32_MIN);
  if (nk > 0)
    col_exponent_kernel<<<dim3(unsigned(ceil_div(nj, 64)), unsigned(ceil_div(nk, kColExpRows))), dim3(256), 0, s>>>(
        B, nk, nj, fb);
  split_cols_kernel<<<dim3(unsigned(l.njp / 64), unsigned(l.nkp / 128)), dim3(512), 0, s>>>(B, nk, nj, l.nkp, l.njp,
                                                                                           fb, Bs, P);
  RPX_TRY(hipStreamWaitEvent(s, ax->join, 0));
  ozaki_gemm_kernel<<<dim3(unsigned(tm * tn)), dim3(kThreads), 0, s>>>(R, As, Bs, l.nip, l.njp, l.nkp, int(tn), P);
  const int use_beta = (tune_flags(tune) & RPX_TUNE_GEMM_BETA) != 0;
  crt_combine_kernel<<<dim3(unsigned(ceil_div(l.nip * l.njp / 4, 256))), dim3(256), 0, s>>>(
      C, R, ea, fb, alpha, beta, use_beta, ni, nj, l.nip, l.njp, int(tn), P);
  return launch_status();
}

}  // extern "C"
