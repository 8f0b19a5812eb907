// Device helpers shared by the FP64 and the mixed-precision FEM kernels.
#pragma once

#include "rpx_common.hpp"

namespace rpx {

// Effective basis tables of the suite's half-storage scheme
// (DIFFUSION3DPA.hpp:243-268, DIFFUSION3DPA_2 / _6).  With Q = D+1 the B and
// G slots of the shared table are disjoint, every slot has exactly two
// writers -- the reader's own (q,d) pair and its mirror (Q-1-q, D-1-d) -- and
// the Base_Seq loops write them in (d outer, q inner) order, identically for
// the forward (B, G) and the backward (Bt, Gt) tables.  The value a reader
// (q,d) sees is therefore that of whichever of the two writers comes later in
// that order; G values carry the writer's sign and are multiplied by the
// reader's sign on use.
__device__ __forceinline__ void diffusion_tables(int q, int d, int Q, int D, const double* b, const double* g,
                                                 double* Beff, double* Geff) {
  auto key = [&](int qq, int dd) { return dd * Q + qq; };
  auto sgn = [](int qq, int dd) { return qq <= dd ? -1.0 : 1.0; };
  // B slot read by (q,d) and its two writers.
  const int bi = q <= d ? q : Q - 1 - q, bj = q <= d ? d : D - 1 - d;
  const int mi = Q - 1 - bi, mj = D - 1 - bj;  // mirror writer
  const bool self_last = key(bi, bj) > key(mi, mj);
  *Beff = self_last ? b[bi + Q * bj] : b[mi + Q * mj];
  // G slot read by (q,d): writers (k,l) and its mirror.
  const int gk = q <= d ? Q - 1 - q : q, gl = q <= d ? D - 1 - d : d;
  const int nk = Q - 1 - gk, nl = D - 1 - gl;
  const bool g_self_last = key(gk, gl) > key(nk, nl);
  const double stored = g_self_last ? g[gk + Q * gl] * sgn(gk, gl) : g[nk + Q * nl] * sgn(nk, nl);
  *Geff = stored * sgn(q, d);
}

}  // namespace rpx
