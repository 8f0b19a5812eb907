// MASS3DPA and DIFFUSION3DPA: partial-assembly operator actions on hex
// elements, sum-factorised, FP64.
//
// Reference: src/apps/MASS3DPA.hpp:158-339 with MASS3DPA-Seq.cpp:29-94, and
// src/apps/DIFFUSION3DPA.hpp:216-458 with DIFFUSION3DPA-Seq.cpp:28-110.  The
// suite's HIP variants run one element per workgroup of (Q1D,Q1D[,Q1D])
// work-items with 8-9 barriers (MASS3DPA-Hip.cpp:25-111,
// DIFFUSION3DPA-Hip.cpp:25-134).
//
// MI355X design: a workgroup of 256 work-items owns E elements.  Each of the
// six contraction stages maps one work-item to one 1D "pencil" of one element
// (e.g. the x-contraction to the (e,dz,dy) pencils), keeps the pencil in
// registers and writes the stage result to LDS; stages are separated by
// workgroup barriers, so a workgroup takes 6 barriers per E elements.  X and
// the quadrature data are read once with coalesced pencil loads, Y is
// updated in place.  The contraction order of every stage and the
// association of every sum are those of the Base_Seq loops and no
// multiply-add is fused, so results are bit-identical to Base_Seq.
//
// Stage order (forward x, y, z; backward x, y, z):
//   P1 (e,dz,dy): DDQ[dz][dy][qx] = sum_dx X[dz][dy][dx] B(qx,dx)
//   P2 (e,dz,qx): DQQ[dz][qy][qx] = sum_dy DDQ[dz][dy][qx] B(qy,dy)
//   P3 (e,qy,qx): QQQ[qz][qy][qx] = (sum_dz DQQ[dz][qy][qx] B(qz,dz)) * D
//   P4 (e,qz,qy): QQD[qz][qy][dx] = sum_qx QQQ[qz][qy][qx] Bt'(dx,qx)
//   P5 (e,qz,dx): QDD[qz][dy][dx] = sum_qy QQD[qz][qy][dx] Bt'(dy,qy)
//   P6 (e,dy,dx): Y[dz][dy][dx]  += sum_qz QDD[qz][dy][dx] Bt'(dz,qz)
// with B(q,d) = B[q + Q*d] and the suite's literal Bt'(d,q) = Bt[q + D*d]
// (MASS3DPA.hpp:175).  DIFFUSION3DPA carries three components per stage.
#include "rpx_common.hpp"
#include "fem_common.hpp"

#include <algorithm>
#include <mutex>
#include <vector>

using namespace rpx;

namespace {

constexpr int kBS = 256;
constexpr int kLdsBudget = 48 * 1024;

// MASS3DPA's LDS bytes per workgroup by order, which sets its elements per
// workgroup: below 48 KB a CU holds four workgroups where the registers
// allow it.  Timed with 24, 32, 36, 40 and 48 KB at NE = 262144 (65536 for
// D1D >= 5): 24 KB is ~10% faster at D1D = 3, 36 KB 2-6% at 4, 40 KB ~6% at
// 5; D1D = 2 and >= 6 show no gain.
constexpr int mass_lds_budget(int D) {
  return D == 3 ? 24 * 1024 : D == 4 ? 36 * 1024 : D == 5 ? 40 * 1024 : kLdsBudget;
}

// A compiler barrier at the top of each pencil iteration (orders D1D >= 7):
// the basis values are then re-read from LDS (a wavefront-uniform
// broadcast) for every pencil instead of all D*Q of them being hoisted into
// registers, which at these orders cost more VGPRs than the rest of the
// kernel together (246 against 164 at (8,9)).  Lower orders keep them in
// registers: re-reading cost 27% at (4,5).
__device__ __forceinline__ void reload_basis() { asm volatile("" ::: "memory"); }

// Global-memory pointers: with the address space in the type, accesses are
// global (not FLAT) instructions even where the pointer comes from memory.
typedef const __attribute__((address_space(1))) double gcdouble;
typedef __attribute__((address_space(1))) double gdouble;

// Doubles of D a work-item may hold in registers from the start of the
// element code to its third stage (measured at (4,5): 0.150 -> 0.137 ms).
constexpr int kMassPrefetch = 24;

template <int D, int Q>
struct MassShape {
  static constexpr int kA = Q * Q * Q;  // DDQ / QQQ / QDD per element
  static constexpr int kB = D * Q * Q;  // DQQ / QQD per element
  static constexpr int kBytes = 8 * (kA + kB);
  static constexpr int E0 = mass_lds_budget(D) / kBytes;
  static constexpr int E = E0 < 1 ? 1 : (E0 > 64 ? 64 : E0);
  static constexpr int kLds = E * (kA + kB) + 2 * Q * D;  // doubles of LDS
};

// Elements [e0, e0+ne) of one operator application, ne <= MassShape::E, by
// the whole workgroup, with four disjoint LDS buffers: bufA [E][kA],
// bufB [E][kB], sB[q*D + d] = B(q,d) and sBt[d*Q + q] = Bt'(d,q) =
// Bt[q + D*d] (Q*D each).  They are separate restrict parameters so the
// compiler may reorder accesses between them.  Ends with the workgroup's
// last LDS reads still pending: a caller that reuses the buffers must
// synchronise first.
template <int D, int Q, bool kEarlyXY = true>
__device__ __forceinline__ void mass3dpa_elements(gcdouble* __restrict__ Bg, gcdouble* __restrict__ Btg,
                                                  gcdouble* __restrict__ Dg, gcdouble* __restrict__ Xg,
                                                  gdouble* __restrict__ Yg, int64_t e0, int ne,
                                                  double* __restrict__ bufA, double* __restrict__ bufB,
                                                  double* __restrict__ sB, double* __restrict__ sBt) {
  using S = MassShape<D, Q>;
  constexpr int D2 = D * D, D3 = D * D * D, Q2 = Q * Q, Q3 = Q * Q * Q;
  const int tid = threadIdx.x;

  for (int t = tid; t < Q * D; t += kBS) {
    const int q = t / D, d = t % D;
    sB[t] = Bg[q + Q * d];
    const int d2 = t / Q, q2 = t % Q;
    sBt[t] = Btg[q2 + D * d2];
  }
  constexpr int kIt3 = (S::E * Q2 + kBS - 1) / kBS;  // P3 pencils per work-item
  constexpr bool kPre = kIt3 * Q <= kMassPrefetch;
  double dpre[kPre ? kIt3 : 1][Q];
  // X (P1) and Y (P6) pencils go to registers early where they fit in eight
  // doubles: X before the first barrier, Y at the start of P4 (when D's
  // registers are free), so neither load waits at the head of its stage.
  // Not at Q = D + 2 from D = 5, and not in the persistent loops (the
  // single-launch batch kernel and the per-order loop), where the extra live
  // registers cost a wavefront per SIMD.
  constexpr int kIt1 = (S::E * D2 + kBS - 1) / kBS;  // P1 and P6 pencils per work-item
  constexpr bool kPreXY = kEarlyXY && kIt1 * D <= 8 && !(D >= 5 && Q >= D + 2);
  double xpre[kPreXY ? kIt1 : 1][D];
  if constexpr (kPreXY) {
#pragma unroll
    for (int it = 0; it < kIt1; ++it) {
      const int w = tid + it * kBS;
      if (w < ne * D2) {
        gcdouble* xp = Xg + (e0 + w / D2) * D3 + (w % D2) * D;  // X[e][dz][dy][:]
#pragma unroll
        for (int dx = 0; dx < D; ++dx) xpre[it][dx] = xp[dx];
      }
    }
  }
  if constexpr (kPre) {
#pragma unroll
    for (int it = 0; it < kIt3; ++it) {
      const int w = tid + it * kBS;
      if (w < ne * Q2) {
        const int e = w / Q2, r = w % Q2;
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) dpre[it][qz] = Dg[(e0 + e) * Q3 + r + qz * Q2];
      }
    }
  }
  __syncthreads();

  // P1: x-contraction of X pencils.
  for (int w = tid, it = 0; w < ne * D2; w += kBS, ++it) {
    if constexpr (D >= 7) reload_basis();
    const int e = w / D2, r = w % D2, dz = r / D, dy = r % D;
    double x[D];
    if constexpr (kPreXY) {
#pragma unroll
      for (int dx = 0; dx < D; ++dx) x[dx] = xpre[it][dx];
    } else {
      gcdouble* xp = Xg + (e0 + e) * D3 + dz * D2 + dy * D;
#pragma unroll
      for (int dx = 0; dx < D; ++dx) x[dx] = xp[dx];
    }
    double* out = bufA + e * S::kA + (dz * D + dy) * Q;
#pragma unroll
    for (int qx = 0; qx < Q; ++qx) {
      double u = 0.0;
#pragma unroll
      for (int dx = 0; dx < D; ++dx) u += x[dx] * sB[qx * D + dx];
      out[qx] = u;
    }
  }
  __syncthreads();
  // P2: y-contraction.
  for (int w = tid; w < ne * D * Q; w += kBS) {
    if constexpr (D >= 7) reload_basis();
    const int e = w / (D * Q), r = w % (D * Q), dz = r / Q, qx = r % Q;
    const double* in = bufA + e * S::kA + dz * D * Q + qx;  // DDQ[dz][dy][qx], stride Q in dy
    double v[D];
#pragma unroll
    for (int dy = 0; dy < D; ++dy) v[dy] = in[dy * Q];
    double* out = bufB + e * S::kB + dz * Q2 + qx;  // DQQ[dz][qy][qx], stride Q in qy
#pragma unroll
    for (int qy = 0; qy < Q; ++qy) {
      double u = 0.0;
#pragma unroll
      for (int dy = 0; dy < D; ++dy) u += v[dy] * sB[qy * D + dy];
      out[qy * Q] = u;
    }
  }
  __syncthreads();
  // P3: z-contraction and the quadrature-point scaling.  The D values were
  // loaded before P1 when they fit in a few registers (kPre), so their
  // latency overlaps the first two stages.
#pragma unroll
  for (int it = 0; it < kIt3; ++it) {
    const int w = tid + it * kBS;
    if (w < ne * Q2) {
      if constexpr (D >= 7) reload_basis();
      const int e = w / Q2, r = w % Q2;  // r = qy*Q + qx
      const double* in = bufB + e * S::kB + r;  // DQQ[dz][qy][qx], stride Q2 in dz
      double v[D];
#pragma unroll
      for (int dz = 0; dz < D; ++dz) v[dz] = in[dz * Q2];
      double* out = bufA + e * S::kA + r;  // QQQ[qz][qy][qx], stride Q2 in qz
#pragma unroll
      for (int qz = 0; qz < Q; ++qz) {
        double u = 0.0;
#pragma unroll
        for (int dz = 0; dz < D; ++dz) u += v[dz] * sB[qz * D + dz];
        if constexpr (kPre)
          out[qz * Q2] = u * dpre[it][qz];
        else
          out[qz * Q2] = u * Dg[(e0 + e) * Q3 + r + qz * Q2];
      }
    }
  }
  __syncthreads();
  double ypre[kPreXY ? kIt1 : 1][D];
  if constexpr (kPreXY) {
#pragma unroll
    for (int it = 0; it < kIt1; ++it) {
      const int w = tid + it * kBS;
      if (w < ne * D2) {
        gcdouble* yp = Yg + (e0 + w / D2) * D3 + w % D2;  // Y[e][:][dy][dx]
#pragma unroll
        for (int dz = 0; dz < D; ++dz) ypre[it][dz] = yp[dz * D2];
      }
    }
  }
  // P4: backward x-contraction.
  for (int w = tid; w < ne * Q2; w += kBS) {
    if constexpr (D >= 7) reload_basis();
    const int e = w / Q2, r = w % Q2;  // r = qz*Q + qy
    const double* in = bufA + e * S::kA + r * Q;  // QQQ[qz][qy][qx], contiguous in qx
    double v[Q];
#pragma unroll
    for (int qx = 0; qx < Q; ++qx) v[qx] = in[qx];
    double* out = bufB + e * S::kB + r * D;  // QQD[qz][qy][dx]
#pragma unroll
    for (int dx = 0; dx < D; ++dx) {
      double u = 0.0;
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) u += v[qx] * sBt[dx * Q + qx];
      out[dx] = u;
    }
  }
  __syncthreads();
  // P5: backward y-contraction.
  for (int w = tid; w < ne * Q * D; w += kBS) {
    if constexpr (D >= 7) reload_basis();
    const int e = w / (Q * D), r = w % (Q * D), qz = r / D, dx = r % D;
    const double* in = bufB + e * S::kB + qz * Q * D + dx;  // QQD[qz][qy][dx], stride D in qy
    double v[Q];
#pragma unroll
    for (int qy = 0; qy < Q; ++qy) v[qy] = in[qy * D];
    double* out = bufA + e * S::kA + qz * D2 + dx;  // QDD[qz][dy][dx], stride D in dy
#pragma unroll
    for (int dy = 0; dy < D; ++dy) {
      double u = 0.0;
#pragma unroll
      for (int qy = 0; qy < Q; ++qy) u += v[qy] * sBt[dy * Q + qy];
      out[dy * D] = u;
    }
  }
  __syncthreads();
  // P6: backward z-contraction, accumulate into Y.
  for (int w = tid, it = 0; w < ne * D2; w += kBS, ++it) {
    if constexpr (D >= 7) reload_basis();
    const int e = w / D2, r = w % D2;  // r = dy*D + dx
    const double* in = bufA + e * S::kA + r;  // QDD[qz][dy][dx], stride D2 in qz
    double v[Q];
#pragma unroll
    for (int qz = 0; qz < Q; ++qz) v[qz] = in[qz * D2];
    gdouble* yp = Yg + (e0 + e) * D3 + r;
    // All of the pencil's Y values are loaded before the first store: where
    // the Y pointer comes from memory the compiler otherwise keeps each
    // load behind the previous store and waits on it D times per pencil.
    double y[D];
    if constexpr (kPreXY) {
#pragma unroll
      for (int dz = 0; dz < D; ++dz) y[dz] = ypre[it][dz];
    } else {
#pragma unroll
      for (int dz = 0; dz < D; ++dz) y[dz] = yp[dz * D2];
    }
#pragma unroll
    for (int dz = 0; dz < D; ++dz) {
      double u = 0.0;
#pragma unroll
      for (int qz = 0; qz < Q; ++qz) u += v[qz] * sBt[dz * Q + qz];
      yp[dz * D2] = y[dz] + u;
    }
  }
}

template <int D, int Q>
__global__ __launch_bounds__(kBS) void mass3dpa_kernel(const double* __restrict__ Bg,
                                                       const double* __restrict__ Btg,
                                                       const double* __restrict__ Dg,
                                                       const double* __restrict__ Xg, double* __restrict__ Yg,
                                                       int64_t NE) {
  using S = MassShape<D, Q>;
  __shared__ double bufA[S::E * S::kA];
  __shared__ double bufB[S::E * S::kB];
  __shared__ double sB[Q * D];
  __shared__ double sBt[D * Q];
  const int64_t e0 = int64_t(blockIdx.x) * S::E;
  const int ne = int((NE - e0) < S::E ? (NE - e0) : S::E);
  mass3dpa_elements<D, Q>((gcdouble*)Bg, (gcdouble*)Btg, (gcdouble*)Dg, (gcdouble*)Xg, (gdouble*)Yg, e0, ne, bufA, bufB,
                          sB, sBt);
}

// ---- ragged batches, one persistent launch ----------------------------------
//
// A batch is one operator application (its own bases, data and order, with
// Q1D = D1D + 1).  Batch b is cut into chunks of MassShape<D,Q>::E elements;
// chunk c belongs to batch chunk_batch[c], whose first chunk is
// chunk0[batch].  A grid of a few workgroups per CU strides over the chunks;
// each chunk dispatches on its batch's order to the same element code as the
// single-order kernel, so results are bit-identical to it.
constexpr int kBatchLds = kLdsBudget / 8 + 2 * 9 * 8;  // doubles: the largest kLds over the orders

typedef __attribute__((address_space(3))) double lds_double;

// One chunk of a batch of order D.  Not inlined: each order's element code
// keeps its own register allocation (inlined into one kernel the eight
// orders together needed 256 VGPRs, one wavefront per SIMD); the kernel is
// bounded to three wavefronts per SIMD, which the LDS allows.  The LDS
// pointer keeps its address space across the call so its accesses stay LDS
// instructions.  Measured on 64 batches of orders 2..8 (1.0 M elements):
// 3.3 ms, against 2.3 ms for one single-order launch per batch.
template <int D>
__device__ __attribute__((noinline)) void mass_batch_chunk(rpx_mass_batch b, int64_t e0, lds_double* lds) {
  using S = MassShape<D, D + 1>;
  static_assert(S::kLds <= kBatchLds, "LDS budget");
  const int ne = int((b.ne - e0) < S::E ? (b.ne - e0) : S::E);
  double* bufA = (double*)lds;
  double* bufB = bufA + S::E * S::kA;
  double* sB = bufB + S::E * S::kB;
  // The descriptors' pointers are global memory: typed so, the element code
  // issues global (not FLAT) accesses, which do not hold up its LDS waits.
  mass3dpa_elements<D, D + 1, false>((gcdouble*)b.B, (gcdouble*)b.Bt, (gcdouble*)b.D, (gcdouble*)b.X, (gdouble*)b.Y, e0, ne,
                              bufA, bufB, sB, sB + (D + 1) * D);
}

// Elements per chunk of a batch of order d1d (Q1D = d1d + 1).
__host__ __device__ inline int64_t batch_chunk_elems(int d1d) {
  switch (d1d) {
    case 1: return MassShape<1, 2>::E;
    case 2: return MassShape<2, 3>::E;
    case 3: return MassShape<3, 4>::E;
    case 4: return MassShape<4, 5>::E;
    case 5: return MassShape<5, 6>::E;
    case 6: return MassShape<6, 7>::E;
    case 7: return MassShape<7, 8>::E;
    default: return MassShape<8, 9>::E;
  }
}

// Scratch of a batched plan: a 256-byte header (the chunk count), the batch
// descriptors, each chunk's batch, each batch's first chunk.
struct BatchHeader {
  int64_t nchunks;
  int64_t nbatch;
  int64_t order_begin[9];  // chunks of order d1d are [order_begin[d1d], order_end[d1d])
  int64_t order_end[9];
};
static_assert(sizeof(BatchHeader) <= 256, "plan header");

__host__ __device__ inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
__host__ __device__ inline size_t chunk_batch_offset(int64_t nbatch) {
  return 256 + align256(size_t(nbatch) * sizeof(rpx_mass_batch));
}
__host__ __device__ inline size_t chunk0_offset(int64_t nbatch, int64_t nchunks) {
  return chunk_batch_offset(nbatch) + align256(size_t(nchunks) * sizeof(int));
}

// Chunks already run, per order: the first done[d] chunks of order d.
struct ChunkSkip {
  int64_t done[9];
};

// Every order's chunks after its first skip.done[d] (all of them for a zero
// skip), highest order first; plan order puts the orders' ranges in that
// order, so c only grows.
__global__ __launch_bounds__(kBS, 3) void mass3dpa_batched_kernel(const char* __restrict__ plan, ChunkSkip skip) {
  __shared__ double lds[kBatchLds];
  const BatchHeader* hdr = reinterpret_cast<const BatchHeader*>(plan);
  const int64_t nchunks = hdr->nchunks, nbatch = hdr->nbatch;
  const auto* batches = reinterpret_cast<const rpx_mass_batch*>(plan + 256);
  const auto* chunk_batch = reinterpret_cast<const int*>(plan + chunk_batch_offset(nbatch));
  const auto* chunk0 = reinterpret_cast<const int64_t*>(plan + chunk0_offset(nbatch, nchunks));
  rpx_mass_batch b{};
  int64_t cb = 0, ce = 0;  // chunk range of the batch held in b
  for (int d = 8; d >= 1; --d)
  for (int64_t c = hdr->order_begin[d] + skip.done[d] + blockIdx.x; c < hdr->order_end[d]; c += gridDim.x) {
    if (c >= ce) {  // the stride left b's chunks (each batch's are contiguous)
      const int bi = chunk_batch[c];
      b = batches[bi];
      cb = chunk0[bi];
      ce = cb + ceil_div(b.ne, batch_chunk_elems(b.d1d));
    }
    const int64_t k = c - cb;
    lds_double* l = (lds_double*)lds;
    switch (b.d1d) {
      case 1: mass_batch_chunk<1>(b, k * MassShape<1, 2>::E, l); break;
      case 2: mass_batch_chunk<2>(b, k * MassShape<2, 3>::E, l); break;
      case 3: mass_batch_chunk<3>(b, k * MassShape<3, 4>::E, l); break;
      case 4: mass_batch_chunk<4>(b, k * MassShape<4, 5>::E, l); break;
      case 5: mass_batch_chunk<5>(b, k * MassShape<5, 6>::E, l); break;
      case 6: mass_batch_chunk<6>(b, k * MassShape<6, 7>::E, l); break;
      case 7: mass_batch_chunk<7>(b, k * MassShape<7, 8>::E, l); break;
      default: mass_batch_chunk<8>(b, k * MassShape<8, 9>::E, l); break;
    }
    __syncthreads();  // the next chunk rewrites the LDS
  }
}

// The batches of one order, as one persistent launch over that order's range
// of the chunk table (the plan sorts batches by order, so the range is
// contiguous).  Same element code, LDS sizes and register allocation as the
// single-order kernel; an order with no batches exits at once.
template <int D>
__global__ __launch_bounds__(kBS) void mass3dpa_order_kernel(const char* __restrict__ plan) {
  using S = MassShape<D, D + 1>;
  __shared__ double bufA[S::E * S::kA];
  __shared__ double bufB[S::E * S::kB];
  __shared__ double sB[(D + 1) * D];
  __shared__ double sBt[D * (D + 1)];
  const BatchHeader* hdr = reinterpret_cast<const BatchHeader*>(plan);
  const int64_t nchunks = hdr->nchunks, nbatch = hdr->nbatch;
  const int64_t begin = hdr->order_begin[D], end = hdr->order_end[D];
  const auto* batches = reinterpret_cast<const rpx_mass_batch*>(plan + 256);
  const auto* chunk_batch = reinterpret_cast<const int*>(plan + chunk_batch_offset(nbatch));
  const auto* chunk0 = reinterpret_cast<const int64_t*>(plan + chunk0_offset(nbatch, nchunks));
  // Each batch's chunks are contiguous and c only grows, so the batch
  // descriptor is looked up again only when the stride leaves its range: a
  // lookup is a chain of three dependent loads, about 1-2 us, against a few
  // microseconds of work per chunk at the high orders.
  rpx_mass_batch b{};
  int64_t cb = 0, ce = 0;  // chunk range of the batch held in b
  for (int64_t c = begin + blockIdx.x; c < end; c += gridDim.x) {
    if (c >= ce) {
      const int bi = chunk_batch[c];
      b = batches[bi];
      cb = chunk0[bi];
      ce = cb + ceil_div(b.ne, int64_t(S::E));
    }
    const int64_t e0 = (c - cb) * S::E;
    const int ne = int((b.ne - e0) < S::E ? (b.ne - e0) : S::E);
    mass3dpa_elements<D, D + 1, false>((gcdouble*)b.B, (gcdouble*)b.Bt, (gcdouble*)b.D, (gcdouble*)b.X, (gdouble*)b.Y, e0,
                                ne, bufA, bufB, sB, sBt);
    __syncthreads();  // the next chunk rewrites the LDS
  }
}

// The batches of one order with one workgroup per chunk, for a grid of the
// order's chunk count.  Without the persistent loop around it the element
// code compiles as in the single-order kernel (164 against 118 VGPRs at
// D1D = 8, where the looped form waits on each LDS basis read of its last
// stages and runs 12-15% slower at D1D = 7, 8).
template <int D>
__global__ __launch_bounds__(kBS) void mass3dpa_order_chunk_kernel(const char* __restrict__ plan) {
  using S = MassShape<D, D + 1>;
  __shared__ double bufA[S::E * S::kA];
  __shared__ double bufB[S::E * S::kB];
  __shared__ double sB[(D + 1) * D];
  __shared__ double sBt[D * (D + 1)];
  const BatchHeader* hdr = reinterpret_cast<const BatchHeader*>(plan);
  const int64_t nchunks = hdr->nchunks, nbatch = hdr->nbatch;
  const int64_t c = hdr->order_begin[D] + blockIdx.x;
  if (c >= hdr->order_end[D]) return;
  const auto* batches = reinterpret_cast<const rpx_mass_batch*>(plan + 256);
  const auto* chunk_batch = reinterpret_cast<const int*>(plan + chunk_batch_offset(nbatch));
  const auto* chunk0 = reinterpret_cast<const int64_t*>(plan + chunk0_offset(nbatch, nchunks));
  const int bi = chunk_batch[c];
  const rpx_mass_batch b = batches[bi];
  const int64_t e0 = (c - chunk0[bi]) * S::E;
  const int ne = int((b.ne - e0) < S::E ? (b.ne - e0) : S::E);
  mass3dpa_elements<D, D + 1>((gcdouble*)b.B, (gcdouble*)b.Bt, (gcdouble*)b.D, (gcdouble*)b.X, (gdouble*)b.Y, e0,
                              ne, bufA, bufB, sB, sBt);
}

template <int D>
int launch_order_chunks(const char* plan, int64_t chunks, hipStream_t s) {
  if (chunks <= 0) return RPX_SUCCESS;
  if (chunks > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  mass3dpa_order_chunk_kernel<D><<<dim3(unsigned(chunks)), dim3(kBS), 0, s>>>(plan);
  return launch_status();
}

// Chunk counts per order of the plans written by rpx_mass3dpa_batched_plan,
// kept on the host by scratch address so that a run sizes its grids without
// reading the device.  At most kMaxPlanRecords are kept (oldest dropped); a
// scratch without a record runs on the persistent kernels.
struct PlanRecord {
  const void* scratch;
  int64_t chunks[9];
};
constexpr size_t kMaxPlanRecords = 64;
std::mutex g_plan_mu;
std::vector<PlanRecord> g_plans;

void forget_plan(const void* scratch) {
  std::lock_guard<std::mutex> lock(g_plan_mu);
  g_plans.erase(std::remove_if(g_plans.begin(), g_plans.end(),
                               [&](const PlanRecord& r) { return r.scratch == scratch; }),
                g_plans.end());
}

void record_plan(const void* scratch, const int64_t* chunks) {
  std::lock_guard<std::mutex> lock(g_plan_mu);
  g_plans.erase(std::remove_if(g_plans.begin(), g_plans.end(),
                               [&](const PlanRecord& r) { return r.scratch == scratch; }),
                g_plans.end());
  if (g_plans.size() >= kMaxPlanRecords) g_plans.erase(g_plans.begin());
  PlanRecord r;
  r.scratch = scratch;
  std::copy(chunks, chunks + 9, r.chunks);
  g_plans.push_back(r);
}

bool find_plan(const void* scratch, int64_t* chunks) {
  std::lock_guard<std::mutex> lock(g_plan_mu);
  for (const PlanRecord& r : g_plans)
    if (r.scratch == scratch) {
      std::copy(r.chunks, r.chunks + 9, chunks);
      return true;
    }
  return false;
}

// Workgroups of mass3dpa_order_kernel<D> resident per CU (queried once).
template <int D>
int order_blocks_per_cu() {
  static int cached = 0;
  if (cached == 0) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mass3dpa_order_kernel<D>, kBS, 0) != hipSuccess || n < 1)
      n = 1;
    cached = n;
  }
  return cached;
}

template <int D>
int launch_order(const char* plan, int cus, int64_t grid_override, hipStream_t s) {
  const int64_t grid = grid_override > 0 ? grid_override : int64_t(order_blocks_per_cu<D>()) * cus;
  if (grid > INT32_MAX) return RPX_ERR_INVALID_ARGUMENT;
  mass3dpa_order_kernel<D><<<dim3(unsigned(grid)), dim3(kBS), 0, s>>>(plan);
  return launch_status();
}

struct BatchLayout {
  int64_t nchunks;
  size_t batches_off, chunk_batch_off, chunk0_off, total;
};
BatchLayout batch_layout(const rpx_mass_batch* batches, int nbatch) {
  BatchLayout l;
  l.nchunks = 0;
  for (int i = 0; i < nbatch; ++i) l.nchunks += ceil_div(batches[i].ne, batch_chunk_elems(batches[i].d1d));
  l.batches_off = 256;
  l.chunk_batch_off = chunk_batch_offset(nbatch);
  l.chunk0_off = chunk0_offset(nbatch, l.nchunks);
  l.total = l.chunk0_off + align256(size_t(nbatch + 1) * sizeof(int64_t));
  return l;
}

// The plain FP64 comparator in the shape of the suite's Base_HIP variant
// (MASS3DPA-Hip.cpp:25-117): one element per workgroup of Q x Q work-items,
// two Q^3 LDS buffers, a barrier between stages.  Same sums in the same order
// as mass3dpa_kernel, so both are bit-identical to Base_Seq.
template <int D, int Q>
__global__ __launch_bounds__(Q* Q) void mass3dpa_naive_kernel(const double* __restrict__ Bg,
                                                               const double* __restrict__ Btg,
                                                               const double* __restrict__ Dg,
                                                               const double* __restrict__ Xg,
                                                               double* __restrict__ Yg) {
  constexpr int D2 = D * D, D3 = D * D * D, Q3 = Q * Q * Q;
  __shared__ double sB[Q * D];   // sB[q*D + d]  = B(q,d)
  __shared__ double sBt[D * Q];  // sBt[d*Q + q] = Bt'(d,q)
  __shared__ double s0[Q3];
  __shared__ double s1[Q3];
  const int tx = threadIdx.x % Q, ty = threadIdx.x / Q;
  const int64_t e = blockIdx.x;
  for (int t = threadIdx.x; t < Q * D; t += Q * Q) {
    sB[t] = Bg[t / D + Q * (t % D)];
    sBt[t] = Btg[t % Q + D * (t / Q)];
  }
  // X[dz][dy][dx] -> s0
  if (tx < D && ty < D)
    for (int dz = 0; dz < D; ++dz) s0[(dz * D + ty) * D + tx] = Xg[e * D3 + dz * D2 + ty * D + tx];
  __syncthreads();
  // x: s1[dz][dy][qx]   (tx = qx, ty = dy)
  if (ty < D)
    for (int dz = 0; dz < D; ++dz) {
      double u = 0.0;
      for (int dx = 0; dx < D; ++dx) u += s0[(dz * D + ty) * D + dx] * sB[tx * D + dx];
      s1[(dz * D + ty) * Q + tx] = u;
    }
  __syncthreads();
  // y: s0[dz][qy][qx]   (tx = qx, ty = qy)
  for (int dz = 0; dz < D; ++dz) {
    double u = 0.0;
    for (int dy = 0; dy < D; ++dy) u += s1[(dz * D + dy) * Q + tx] * sB[ty * D + dy];
    s0[(dz * Q + ty) * Q + tx] = u;
  }
  __syncthreads();
  // z and D: s1[qz][qy][qx]
  for (int qz = 0; qz < Q; ++qz) {
    double u = 0.0;
    for (int dz = 0; dz < D; ++dz) u += s0[(dz * Q + ty) * Q + tx] * sB[qz * D + dz];
    s1[(qz * Q + ty) * Q + tx] = u * Dg[e * Q3 + (qz * Q + ty) * Q + tx];
  }
  __syncthreads();
  // backward x: s0[qz][qy][dx]   (tx = dx, ty = qy)
  if (tx < D)
    for (int qz = 0; qz < Q; ++qz) {
      double u = 0.0;
      for (int qx = 0; qx < Q; ++qx) u += s1[(qz * Q + ty) * Q + qx] * sBt[tx * Q + qx];
      s0[(qz * Q + ty) * D + tx] = u;
    }
  __syncthreads();
  // backward y: s1[qz][dy][dx]   (tx = dx, ty = dy)
  if (tx < D && ty < D)
    for (int qz = 0; qz < Q; ++qz) {
      double u = 0.0;
      for (int qy = 0; qy < Q; ++qy) u += s0[(qz * Q + qy) * D + tx] * sBt[ty * Q + qy];
      s1[(qz * D + ty) * D + tx] = u;
    }
  __syncthreads();
  // backward z: Y[dz][dy][dx] +=
  if (tx < D && ty < D)
    for (int dz = 0; dz < D; ++dz) {
      double u = 0.0;
      for (int qz = 0; qz < Q; ++qz) u += s1[(qz * D + ty) * D + tx] * sBt[dz * Q + qz];
      Yg[e * D3 + dz * D2 + ty * D + tx] += u;
    }
}

// ---- DIFFUSION3DPA ----------------------------------------------------------

template <int D, int Q>
struct DiffShape {
  static constexpr int kA = 3 * Q * Q * Q;
  static constexpr int kB = 3 * D * Q * Q;
  static constexpr int kBytes = 8 * (kA + kB);
  static constexpr int E0 = kLdsBudget / kBytes;
  static constexpr int E = E0 < 1 ? 1 : (E0 > 64 ? 64 : E0);
};

template <int D, int Q, bool SYMM>
__global__ __launch_bounds__(kBS) void diffusion3dpa_kernel(const double* __restrict__ Bg,
                                                            const double* __restrict__ Gg,
                                                            const double* __restrict__ Dg,
                                                            const double* __restrict__ Xg,
                                                            double* __restrict__ Yg, int64_t NE) {
  using S = DiffShape<D, Q>;
  constexpr int E = S::E;
  constexpr int D2 = D * D, D3 = D * D * D, Q2 = Q * Q, Q3 = Q * Q * Q;
  constexpr int NC = SYMM ? 6 : 9;
  constexpr int A1 = Q3, B1 = D * Q2;  // per-component strides within an element
  __shared__ double bufA[E * S::kA];
  __shared__ double bufB[E * S::kB];
  __shared__ double sB[Q * D];  // sB[q*D + d] = effective B(q,d)
  __shared__ double sG[Q * D];  // sG[q*D + d] = effective G(q,d) with the reader's sign

  const int tid = threadIdx.x;
  const int64_t e0 = int64_t(blockIdx.x) * E;
  const int ne = int((NE - e0) < E ? (NE - e0) : E);

  for (int t = tid; t < Q * D; t += kBS) {
    const int q = t / D, d = t % D;
    double be, ge;
    diffusion_tables(q, d, Q, D, Bg, Gg, &be, &ge);
    sB[t] = be;
    sG[t] = ge;
  }
  // Register budget.  Hoisting the 2*Q*D basis values out of the pencil
  // loops costs more VGPRs than it saves LDS reads here (each basis read
  // feeds three components), so every pencil iteration re-reads them from
  // LDS (a wavefront-uniform broadcast); the registers go to the work-item's
  // D values, loaded before P1 so that their latency overlaps two stages.
  // Measured at NE = 262144 (65536 for D1D >= 5): (4,5) 0.447 -> 0.408 ms
  // with uniform D, 0.49 -> 0.47 ms with the symmetric-tensor D of the
  // roofline report (160 VGPRs, three wavefronts per SIMD; prefetching D
  // alone took 232); (5,6) 0.404 -> 0.237 ms; (6,7) 0.651 -> 0.438 ms.  The
  // non-symmetric (4,5) case, which has no prefetch, was 2% slower
  // re-reading, so it keeps the hoist below D1D = 5.
  constexpr bool kReload = SYMM || D >= 5;
  constexpr bool kPreD = SYMM && E * Q2 <= kBS;  // one P3 pencil per work-item
  double dpre[kPreD ? 6 : 1][Q];
  if constexpr (kPreD) {
    if (tid < ne * Q2) {
      const int e = tid / Q2, r = tid % Q2;
      const double* dp = Dg + (e0 + e) * (NC * Q3) + r;
#pragma unroll
      for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) dpre[c][qz] = dp[c * Q3 + qz * Q2];
    }
  }
  __syncthreads();

  // P1 (DIFFUSION3DPA_3): DDQ0 = sum X*B, DDQ1 = sum X*G over dx.
  for (int w = tid; w < ne * D2; w += kBS) {
    if constexpr (kReload) reload_basis();
    const int e = w / D2, r = w % D2, dz = r / D, dy = r % D;
    const double* xp = Xg + (e0 + e) * D3 + dz * D2 + dy * D;
    double x[D];
#pragma unroll
    for (int dx = 0; dx < D; ++dx) x[dx] = xp[dx];
    double* out = bufA + e * S::kA + (dz * D + dy) * Q;
#pragma unroll
    for (int qx = 0; qx < Q; ++qx) {
      double u = 0.0, v = 0.0;
#pragma unroll
      for (int dx = 0; dx < D; ++dx) {
        u += x[dx] * sB[qx * D + dx];
        v += x[dx] * sG[qx * D + dx];
      }
      out[qx] = u;
      out[A1 + qx] = v;
    }
  }
  __syncthreads();
  // P2 (DIFFUSION3DPA_4): DQQ0 = sum DDQ1*B, DQQ1 = sum DDQ0*G, DQQ2 = sum DDQ0*B over dy.
  for (int w = tid; w < ne * D * Q; w += kBS) {
    if constexpr (kReload) reload_basis();
    const int e = w / (D * Q), r = w % (D * Q), dz = r / Q, qx = r % Q;
    const double* in = bufA + e * S::kA + dz * D * Q + qx;
    double a0[D], a1[D];
#pragma unroll
    for (int dy = 0; dy < D; ++dy) {
      a0[dy] = in[dy * Q];
      a1[dy] = in[A1 + dy * Q];
    }
    double* out = bufB + e * S::kB + dz * Q2 + qx;
#pragma unroll
    for (int qy = 0; qy < Q; ++qy) {
      double u = 0.0, v = 0.0, ww = 0.0;
#pragma unroll
      for (int dy = 0; dy < D; ++dy) {
        u += a1[dy] * sB[qy * D + dy];
        v += a0[dy] * sG[qy * D + dy];
        ww += a0[dy] * sB[qy * D + dy];
      }
      out[qy * Q] = u;
      out[B1 + qy * Q] = v;
      out[2 * B1 + qy * Q] = ww;
    }
  }
  __syncthreads();
  // P3 (DIFFUSION3DPA_5): gradient at the point, times the symmetric tensor.
  for (int w = tid; w < ne * Q2; w += kBS) {
    if constexpr (kReload) reload_basis();
    const int e = w / Q2, r = w % Q2;
    const double* in = bufB + e * S::kB + r;
    double c0[D], c1[D], c2[D];
#pragma unroll
    for (int dz = 0; dz < D; ++dz) {
      c0[dz] = in[dz * Q2];
      c1[dz] = in[B1 + dz * Q2];
      c2[dz] = in[2 * B1 + dz * Q2];
    }
    const double* dp = Dg + (e0 + e) * (NC * Q3) + r;
    double* out = bufA + e * S::kA + r;
#pragma unroll
    for (int qz = 0; qz < Q; ++qz) {
      double u = 0.0, v = 0.0, ww = 0.0;
#pragma unroll
      for (int dz = 0; dz < D; ++dz) {
        u += c0[dz] * sB[qz * D + dz];
        v += c1[dz] * sB[qz * D + dz];
        ww += c2[dz] * sG[qz * D + dz];
      }
      const double* o = dp + qz * Q2;
      // Symmetric storage order (11, 12, 13, 22, 23, 33), DIFFUSION3DPA.hpp:384-392.
      const double O11 = kPreD ? dpre[0][qz] : o[0];
      const double O12 = kPreD ? dpre[1][qz] : o[Q3];
      const double O13 = kPreD ? dpre[2][qz] : o[2 * Q3];
      const double O21 = SYMM ? O12 : o[3 * Q3];
      const double O22 = kPreD ? dpre[3][qz] : (SYMM ? o[3 * Q3] : o[4 * Q3]);
      const double O23 = kPreD ? dpre[4][qz] : (SYMM ? o[4 * Q3] : o[5 * Q3]);
      const double O31 = SYMM ? O13 : o[6 * Q3];
      const double O32 = SYMM ? O23 : o[7 * Q3];
      const double O33 = kPreD ? dpre[5][qz] : (SYMM ? o[5 * Q3] : o[8 * Q3]);
      out[qz * Q2] = ((O11 * u) + (O12 * v)) + (O13 * ww);
      out[A1 + qz * Q2] = ((O21 * u) + (O22 * v)) + (O23 * ww);
      out[2 * A1 + qz * Q2] = ((O31 * u) + (O32 * v)) + (O33 * ww);
    }
  }
  __syncthreads();
  // P4 (DIFFUSION3DPA_7): QQD0 = sum QQQ0*Gt, QQD1 = sum QQQ1*Bt, QQD2 = sum QQQ2*Bt over qx.
  for (int w = tid; w < ne * Q2; w += kBS) {
    if constexpr (kReload) reload_basis();
    const int e = w / Q2, r = w % Q2;  // r = qz*Q + qy
    const double* in = bufA + e * S::kA + r * Q;
    double c0[Q], c1[Q], c2[Q];
#pragma unroll
    for (int qx = 0; qx < Q; ++qx) {
      c0[qx] = in[qx];
      c1[qx] = in[A1 + qx];
      c2[qx] = in[2 * A1 + qx];
    }
    double* out = bufB + e * S::kB + r * D;
#pragma unroll
    for (int dx = 0; dx < D; ++dx) {
      double u = 0.0, v = 0.0, ww = 0.0;
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) {
        u += c0[qx] * sG[qx * D + dx];
        v += c1[qx] * sB[qx * D + dx];
        ww += c2[qx] * sB[qx * D + dx];
      }
      out[dx] = u;
      out[B1 + dx] = v;
      out[2 * B1 + dx] = ww;
    }
  }
  __syncthreads();
  // P5 (DIFFUSION3DPA_8): QDD0 = sum QQD0*Bt, QDD1 = sum QQD1*Gt, QDD2 = sum QQD2*Bt over qy.
  for (int w = tid; w < ne * Q * D; w += kBS) {
    if constexpr (kReload) reload_basis();
    const int e = w / (Q * D), r = w % (Q * D), qz = r / D, dx = r % D;
    const double* in = bufB + e * S::kB + qz * Q * D + dx;
    double c0[Q], c1[Q], c2[Q];
#pragma unroll
    for (int qy = 0; qy < Q; ++qy) {
      c0[qy] = in[qy * D];
      c1[qy] = in[B1 + qy * D];
      c2[qy] = in[2 * B1 + qy * D];
    }
    double* out = bufA + e * S::kA + qz * D2 + dx;
#pragma unroll
    for (int dy = 0; dy < D; ++dy) {
      double u = 0.0, v = 0.0, ww = 0.0;
#pragma unroll
      for (int qy = 0; qy < Q; ++qy) {
        u += c0[qy] * sB[qy * D + dy];
        v += c1[qy] * sG[qy * D + dy];
        ww += c2[qy] * sB[qy * D + dy];
      }
      out[dy * D] = u;
      out[A1 + dy * D] = v;
      out[2 * A1 + dy * D] = ww;
    }
  }
  __syncthreads();
  // P6 (DIFFUSION3DPA_9): sum over qz, Y += (u + v) + w.
  for (int w = tid; w < ne * D2; w += kBS) {
    if constexpr (kReload) reload_basis();
    const int e = w / D2, r = w % D2;
    const double* in = bufA + e * S::kA + r;
    double c0[Q], c1[Q], c2[Q];
#pragma unroll
    for (int qz = 0; qz < Q; ++qz) {
      c0[qz] = in[qz * D2];
      c1[qz] = in[A1 + qz * D2];
      c2[qz] = in[2 * A1 + qz * D2];
    }
    double* yp = Yg + (e0 + e) * D3 + r;
#pragma unroll
    for (int dz = 0; dz < D; ++dz) {
      double u = 0.0, v = 0.0, ww = 0.0;
#pragma unroll
      for (int qz = 0; qz < Q; ++qz) {
        u += c0[qz] * sB[qz * D + dz];
        v += c1[qz] * sB[qz * D + dz];
        ww += c2[qz] * sG[qz * D + dz];
      }
      yp[dz * D2] += (u + v) + ww;
    }
  }
}

template <int D, int Q>
int launch_mass(const double* B, const double* Bt, const double* Dq, const double* X, double* Y, int64_t NE,
                hipStream_t s) {
  const int64_t grid = ceil_div(NE, MassShape<D, Q>::E);
  if (grid > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  mass3dpa_kernel<D, Q><<<dim3(unsigned(grid)), dim3(kBS), 0, s>>>(B, Bt, Dq, X, Y, NE);
  return launch_status();
}

template <int D, int Q>
int launch_mass_naive(const double* B, const double* Bt, const double* Dq, const double* X, double* Y, int64_t NE,
                      hipStream_t s) {
  if (NE > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  mass3dpa_naive_kernel<D, Q><<<dim3(unsigned(NE)), dim3(Q * Q), 0, s>>>(B, Bt, Dq, X, Y);
  return launch_status();
}

template <int D>
int mass_q(int q1d, bool naive, const double* B, const double* Bt, const double* Dq, const double* X, double* Y,
           int64_t NE, hipStream_t s) {
  if (q1d == D)
    return naive ? launch_mass_naive<D, D>(B, Bt, Dq, X, Y, NE, s) : launch_mass<D, D>(B, Bt, Dq, X, Y, NE, s);
  if (q1d == D + 1)
    return naive ? launch_mass_naive<D, D + 1>(B, Bt, Dq, X, Y, NE, s)
                 : launch_mass<D, D + 1>(B, Bt, Dq, X, Y, NE, s);
  if constexpr (D + 2 <= 9) {
    if (q1d == D + 2)
      return naive ? launch_mass_naive<D, D + 2>(B, Bt, Dq, X, Y, NE, s)
                   : launch_mass<D, D + 2>(B, Bt, Dq, X, Y, NE, s);
  }
  return RPX_ERR_UNSUPPORTED;
}

template <int D, bool SYMM>
int launch_diff(const double* B, const double* G, const double* Dq, const double* X, double* Y, int64_t NE,
                hipStream_t s) {
  constexpr int Q = D + 1;
  const int64_t grid = ceil_div(NE, DiffShape<D, Q>::E);
  if (grid > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  diffusion3dpa_kernel<D, Q, SYMM><<<dim3(unsigned(grid)), dim3(kBS), 0, s>>>(B, G, Dq, X, Y, NE);
  return launch_status();
}

template <bool SYMM>
int diff_d(int d1d, const double* B, const double* G, const double* Dq, const double* X, double* Y, int64_t NE,
           hipStream_t s) {
  switch (d1d) {
    case 1: return launch_diff<1, SYMM>(B, G, Dq, X, Y, NE, s);
    case 2: return launch_diff<2, SYMM>(B, G, Dq, X, Y, NE, s);
    case 3: return launch_diff<3, SYMM>(B, G, Dq, X, Y, NE, s);
    case 4: return launch_diff<4, SYMM>(B, G, Dq, X, Y, NE, s);
    case 5: return launch_diff<5, SYMM>(B, G, Dq, X, Y, NE, s);
    case 6: return launch_diff<6, SYMM>(B, G, Dq, X, Y, NE, s);
    case 7: return launch_diff<7, SYMM>(B, G, Dq, X, Y, NE, s);
    case 8: return launch_diff<8, SYMM>(B, G, Dq, X, Y, NE, s);
    default: return RPX_ERR_UNSUPPORTED;
  }
}

}  // namespace

extern "C" {

int rpx_mass3dpa_f64(const double* B, const double* Bt, const double* D, const double* X, double* Y, int64_t NE,
                     int d1d, int q1d, const rpx_tuning* tune, void* stream) {
  if (NE < 0) return RPX_ERR_INVALID_ARGUMENT;
  if (NE == 0) return RPX_SUCCESS;
  if (!B || !Bt || !D || !X || !Y) return RPX_ERR_INVALID_ARGUMENT;
  hipStream_t s = as_stream(stream);
  const bool naive = (tune_flags(tune) & RPX_TUNE_NAIVE) != 0;
  switch (d1d) {
    case 1: return mass_q<1>(q1d, naive, B, Bt, D, X, Y, NE, s);
    case 2: return mass_q<2>(q1d, naive, B, Bt, D, X, Y, NE, s);
    case 3: return mass_q<3>(q1d, naive, B, Bt, D, X, Y, NE, s);
    case 4: return mass_q<4>(q1d, naive, B, Bt, D, X, Y, NE, s);
    case 5: return mass_q<5>(q1d, naive, B, Bt, D, X, Y, NE, s);
    case 6: return mass_q<6>(q1d, naive, B, Bt, D, X, Y, NE, s);
    case 7: return mass_q<7>(q1d, naive, B, Bt, D, X, Y, NE, s);
    case 8: return mass_q<8>(q1d, naive, B, Bt, D, X, Y, NE, s);
    default: return RPX_ERR_UNSUPPORTED;
  }
}

size_t rpx_mass3dpa_batched_scratch_bytes(const rpx_mass_batch* batches_host, int nbatch) {
  if (!batches_host || nbatch < 0) return 0;
  for (int i = 0; i < nbatch; ++i)
    if (batches_host[i].ne < 0 || batches_host[i].d1d < 1) return 0;
  return batch_layout(batches_host, nbatch).total;
}

int rpx_mass3dpa_batched_plan(const rpx_mass_batch* batches_host, int nbatch, void* scratch, size_t scratch_bytes,
                              void* stream) {
  if (!batches_host || nbatch < 0 || !scratch) return RPX_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < nbatch; ++i) {
    const rpx_mass_batch& b = batches_host[i];
    if (b.ne < 0 || (b.ne > 0 && (!b.B || !b.Bt || !b.D || !b.X || !b.Y))) return RPX_ERR_INVALID_ARGUMENT;
    if (b.d1d < 1 || b.d1d > 8 || b.q1d != b.d1d + 1) return RPX_ERR_UNSUPPORTED;
  }
  const BatchLayout l = batch_layout(batches_host, nbatch);
  if (scratch_bytes < l.total) return RPX_ERR_SCRATCH_TOO_SMALL;
  if (l.nchunks > INT32_MAX) return RPX_ERR_UNSUPPORTED;
  forget_plan(scratch);  // recorded again once the new tables are uploaded
  // Chunks of the highest orders (the costliest) come first, so that the
  // persistent workgroups end on cheap chunks; a batch's chunks stay
  // contiguous.
  std::vector<int> order(static_cast<size_t>(nbatch));
  for (int i = 0; i < nbatch; ++i) order[size_t(i)] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return batches_host[a].d1d > batches_host[b].d1d; });
  std::vector<int> chunk_batch(size_t(l.nchunks));
  std::vector<int64_t> chunk0(size_t(nbatch + 1));
  int64_t c = 0;
  for (int i : order) {
    chunk0[size_t(i)] = c;
    const int64_t k = ceil_div(batches_host[i].ne, batch_chunk_elems(batches_host[i].d1d));
    for (int64_t j = 0; j < k; ++j) chunk_batch[size_t(c + j)] = i;
    c += k;
  }
  chunk0[size_t(nbatch)] = c;
  BatchHeader h{};
  h.nchunks = l.nchunks;
  h.nbatch = nbatch;
  for (int i : order) {
    const int d = batches_host[i].d1d;
    const int64_t first = chunk0[size_t(i)];
    const int64_t last = first + ceil_div(batches_host[i].ne, batch_chunk_elems(d));
    if (h.order_end[d] == h.order_begin[d]) h.order_begin[d] = first;  // first batch of this order
    h.order_end[d] = last;
  }
  hipStream_t st = as_stream(stream);
  char* base = reinterpret_cast<char*>(scratch);
  RPX_TRY(hipMemcpyAsync(base, &h, sizeof(h), hipMemcpyHostToDevice, st));
  if (nbatch > 0)
    RPX_TRY(hipMemcpyAsync(base + l.batches_off, batches_host, size_t(nbatch) * sizeof(rpx_mass_batch),
                           hipMemcpyHostToDevice, st));
  if (l.nchunks > 0)
    RPX_TRY(hipMemcpyAsync(base + l.chunk_batch_off, chunk_batch.data(), chunk_batch.size() * sizeof(int),
                           hipMemcpyHostToDevice, st));
  RPX_TRY(hipMemcpyAsync(base + l.chunk0_off, chunk0.data(), chunk0.size() * sizeof(int64_t),
                         hipMemcpyHostToDevice, st));
  RPX_TRY(hipStreamSynchronize(st));  // the host tables must outlive the copies
  int64_t chunks[9];
  for (int d = 0; d < 9; ++d) chunks[d] = h.order_end[d] - h.order_begin[d];
  record_plan(scratch, chunks);
  return RPX_SUCCESS;
}

int rpx_mass3dpa_batched_run(const void* scratch, const rpx_tuning* tune, void* stream) {
  if (!scratch) return RPX_ERR_INVALID_ARGUMENT;
  int dev = 0, cus = 0;
  RPX_TRY(hipGetDevice(&dev));
  RPX_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const char* plan = reinterpret_cast<const char*>(scratch);
  hipStream_t s = as_stream(stream);
  const int64_t grid_override = tune_or(tune, &rpx_tuning::grid_size, 0);
  if ((tune_flags(tune) & RPX_TUNE_MASS_SINGLE_LAUNCH) == 0) {
    int rc = RPX_SUCCESS;
    int64_t chunks[9];
    if (grid_override == 0 && find_plan(scratch, chunks)) {
      // One launch per order present, one workgroup per chunk, highest first.
      if (rc == RPX_SUCCESS) rc = launch_order_chunks<8>(plan, chunks[8], s);
      if (rc == RPX_SUCCESS) rc = launch_order_chunks<7>(plan, chunks[7], s);
      if (rc == RPX_SUCCESS) rc = launch_order_chunks<6>(plan, chunks[6], s);
      if (rc == RPX_SUCCESS) rc = launch_order_chunks<5>(plan, chunks[5], s);
      if (rc == RPX_SUCCESS) rc = launch_order_chunks<4>(plan, chunks[4], s);
      if (rc == RPX_SUCCESS) rc = launch_order_chunks<3>(plan, chunks[3], s);
      if (rc == RPX_SUCCESS) rc = launch_order_chunks<2>(plan, chunks[2], s);
      if (rc == RPX_SUCCESS) rc = launch_order_chunks<1>(plan, chunks[1], s);
      if (rc != RPX_SUCCESS) return rc;
      // The chunks beyond those grids: none when the record matches the
      // plan in scratch; the rest of a larger plan written there since by
      // other means (a copy into a reused address).
      ChunkSkip skip;
      std::copy(chunks, chunks + 9, skip.done);
      mass3dpa_batched_kernel<<<dim3(unsigned(cus)), dim3(kBS), 0, s>>>(plan, skip);
      return launch_status();
    }
    // One persistent launch per order, highest first (the plan's order).
    if (rc == RPX_SUCCESS) rc = launch_order<8>(plan, cus, grid_override, s);
    if (rc == RPX_SUCCESS) rc = launch_order<7>(plan, cus, grid_override, s);
    if (rc == RPX_SUCCESS) rc = launch_order<6>(plan, cus, grid_override, s);
    if (rc == RPX_SUCCESS) rc = launch_order<5>(plan, cus, grid_override, s);
    if (rc == RPX_SUCCESS) rc = launch_order<4>(plan, cus, grid_override, s);
    if (rc == RPX_SUCCESS) rc = launch_order<3>(plan, cus, grid_override, s);
    if (rc == RPX_SUCCESS) rc = launch_order<2>(plan, cus, grid_override, s);
    if (rc == RPX_SUCCESS) rc = launch_order<1>(plan, cus, grid_override, s);
    return rc;
  }
  // One persistent launch for all orders: workgroups per CU times CUs
  // (tune->grid_size overrides); workgroups beyond the chunk count exit at
  // once.  The kernel finds the tables from the counts in the plan's header.
  const int64_t grid = grid_override > 0 ? grid_override : 3 * int64_t(cus);
  if (grid > INT32_MAX) return RPX_ERR_INVALID_ARGUMENT;
  mass3dpa_batched_kernel<<<dim3(unsigned(grid)), dim3(kBS), 0, s>>>(plan, ChunkSkip{});
  return launch_status();
}

int rpx_mass3dpa_batched_f64(const rpx_mass_batch* batches_host, int nbatch, void* scratch, size_t scratch_bytes,
                             const rpx_tuning* tune, void* stream) {
  const int rc = rpx_mass3dpa_batched_plan(batches_host, nbatch, scratch, scratch_bytes, stream);
  if (rc != RPX_SUCCESS) return rc;
  return rpx_mass3dpa_batched_run(scratch, tune, stream);
}

int rpx_diffusion3dpa_f64(const double* Basis, const double* dBasis, const double* D, const double* X, double* Y,
                          int64_t NE, int d1d, int q1d, int symmetric, const rpx_tuning* /*tune*/, void* stream) {
  if (NE < 0) return RPX_ERR_INVALID_ARGUMENT;
  if (NE == 0) return RPX_SUCCESS;
  if (!Basis || !dBasis || !D || !X || !Y) return RPX_ERR_INVALID_ARGUMENT;
  if (q1d != d1d + 1) return RPX_ERR_UNSUPPORTED;
  hipStream_t s = as_stream(stream);
  return symmetric ? diff_d<true>(d1d, Basis, dBasis, D, X, Y, NE, s)
                   : diff_d<false>(d1d, Basis, dBasis, D, X, Y, NE, s);
}

}  // extern "C"
