// Apps group: constructors and setUp/updateChecksum/tearDown follow
// src/apps/MASS3DPA.cpp:22-89 and DIFFUSION3DPA.cpp:22-93.
#include "apps/Apps.hpp"

#include <algorithm>
#include <vector>

#include "stream/Stream.hpp"

namespace rajaperf {
namespace apps {

// ---- MASS3DPA ----------------------------------------------------------------------
MASS3DPA::MASS3DPA(const RunParams& params) : KernelBase(Apps_MASS3DPA, params) {
  if (params.getD1D() > 0) m_D1D = params.getD1D();
  m_Q1D = params.getQ1D() > 0 ? params.getQ1D() : (params.getD1D() > 0 ? m_D1D + 1 : 5);
  const Index_type D = m_D1D, Q = m_Q1D, Q3 = Q * Q * Q, D3 = D * D * D;
  const Index_type NE_default = 8000;
  setDefaultProblemSize(NE_default * Q3);
  setDefaultReps(50);
  m_NE = std::max((getTargetProblemSize() + Q3 / 2) / Q3, Index_type(1));
  setActualProblemSize(m_NE * Q3);
  setItsPerRep(getActualProblemSize());
  setKernelsPerRep(1);
  setBytesReadPerRep(2 * sizeof(Real_type) * Q * D + 2 * sizeof(Real_type) * D3 * m_NE +
                     1 * sizeof(Real_type) * Q3 * m_NE);
  setBytesWrittenPerRep(1 * sizeof(Real_type) * D3 * m_NE);
  setBytesAtomicModifyWrittenPerRep(0);
  setFLOPsPerRep(m_NE * (2 * D3 * Q + 2 * D * D * Q * Q + 2 * D * Q3 + Q3 + 2 * Q3 * D + 2 * Q * Q * D * D +
                         2 * Q * D3 + D3));
  setComplexity(Complexity::N);
  setVariantDefined(MI355X_Native);
}
void MASS3DPA::setUp(VariantID, size_t) {
  const Size_type QD = Size_type(m_Q1D) * m_D1D;
  const Size_type Q3 = Size_type(m_Q1D) * m_Q1D * m_Q1D, D3 = Size_type(m_D1D) * m_D1D * m_D1D;
  if (seededInputs()) {
    // B: Lagrange basis on Gauss-Lobatto nodes at Gauss-Legendre points;
    // Bt: B^T stored d1d x q1d column-major (read through the suite's
    // literal Bt[q + D1D*d] indexing by the kernel).
    std::vector<Real_type> B, G, Bt(QD);
    detail::femBasis(m_D1D, m_Q1D, B, G);
    for (int q = 0; q < m_Q1D; ++q)
      for (int d = 0; d < m_D1D; ++d) Bt[d + m_D1D * q] = B[q + m_Q1D * d];
    allocData(m_B, QD);
    detail::initDataHostToDevice(m_B, B);
    allocData(m_Bt, QD);
    detail::initDataHostToDevice(m_Bt, Bt);
    allocData(m_D, Q3 * m_NE);
    detail::initDataSeededDevice(m_D, Q3 * m_NE, run_params.getSeed(), 0.5, 2.0);
    allocData(m_X, D3 * m_NE);
    detail::initDataSeededDevice(m_X, D3 * m_NE, run_params.getSeed(), -1.0, 1.0);
    allocAndInitDataConst(m_Y, D3 * m_NE, 0.0);
  } else {  // MASS3DPA.cpp:76-84: all ones, Y zero
    allocAndInitDataConst(m_B, QD, 1.0);
    allocAndInitDataConst(m_Bt, QD, 1.0);
    allocAndInitDataConst(m_D, Q3 * m_NE, 1.0);
    allocAndInitDataConst(m_X, D3 * m_NE, 1.0);
    allocAndInitDataConst(m_Y, D3 * m_NE, 0.0);
  }
}
void MASS3DPA::updateChecksum(VariantID vid, size_t t) {
  checksum[vid][t] += calcChecksum(m_Y, Size_type(m_D1D) * m_D1D * m_D1D * m_NE);
}
void MASS3DPA::tearDown(VariantID, size_t) {
  deallocData(m_B);
  deallocData(m_Bt);
  deallocData(m_D);
  deallocData(m_X);
  deallocData(m_Y);
}
void MASS3DPA::setMI355XTuningDefinitions(VariantID vid) {
  addVariantTuningName(vid, "default");
  addVariantTuningName(vid, "naive");  // one element per Q1D x Q1D workgroup, the plain FP64 comparator
  // int8 MFMA on 7 Ozaki slices of the Kronecker form, where the order has it
  if (rpx_mass3dpa_mixed_scratch_bytes(m_D1D, m_Q1D, 7) > 0) addVariantTuningName(vid, "mixed");
}
void MASS3DPA::runMI355XVariant(VariantID, size_t tune_idx) {
  const Index_type reps = getRunReps();
  if (tune_idx == 2) {  // the sliced bases' scratch is allocated outside the timed reps
    const size_t bytes = rpx_mass3dpa_mixed_scratch_bytes(m_D1D, m_Q1D, 7);
    char* scratch = nullptr;
    allocData(scratch, Size_type(bytes));
    timedReps(reps, [&](RepIndex_type) {
      rpxCheck(rpx_mass3dpa_mixed_f64(m_B, m_Bt, m_D, m_X, m_Y, m_NE, m_D1D, m_Q1D, 7, scratch, bytes, stream()),
               "MASS3DPA mixed");
    });
    deallocData(scratch);
    return;
  }
  rpx_tuning tune = defaultTuning();
  if (tune_idx == 1) tune.flags |= RPX_TUNE_NAIVE;
  timedReps(reps, [&](RepIndex_type) {
    rpxCheck(rpx_mass3dpa_f64(m_B, m_Bt, m_D, m_X, m_Y, m_NE, m_D1D, m_Q1D, &tune, stream()), "MASS3DPA");
  });
}

// ---- DIFFUSION3DPA --------------------------------------------------------------
DIFFUSION3DPA::DIFFUSION3DPA(const RunParams& params) : KernelBase(Apps_DIFFUSION3DPA, params) {
  if (params.getD1D() > 0) m_D1D = params.getD1D();
  m_Q1D = params.getQ1D() > 0 ? params.getQ1D() : m_D1D + 1;
  const Index_type D = m_D1D, Q = m_Q1D, Q3 = Q * Q * Q, D3 = D * D * D;
  const Index_type NE_default = 15625;
  setDefaultProblemSize(NE_default * Q3);
  setDefaultReps(50);
  m_NE = std::max((getTargetProblemSize() + Q3 / 2) / Q3, Index_type(1));
  setActualProblemSize(m_NE * Q3);
  setItsPerRep(getActualProblemSize());
  setKernelsPerRep(1);
  setBytesReadPerRep(2 * sizeof(Real_type) * Q * D + 2 * sizeof(Real_type) * D3 * m_NE +
                     SYM * sizeof(Real_type) * Q3 * m_NE);
  setBytesWrittenPerRep(1 * sizeof(Real_type) * D3 * m_NE);
  setBytesAtomicModifyWrittenPerRep(0);
  setFLOPsPerRep(m_NE * (Q * D + 5 * D * D * Q * D + 7 * D * D * Q * Q + 7 * Q * D * Q * Q + 15 * Q3 + Q * D +
                         7 * Q * Q * D * Q + 7 * Q * Q * D * D + 7 * D * Q * D * D + 3 * D3));
  setComplexity(Complexity::N);
  setVariantDefined(MI355X_Native);
}
void DIFFUSION3DPA::setUp(VariantID, size_t) {
  const Size_type QD = Size_type(m_Q1D) * m_D1D;
  const Size_type Q3 = Size_type(m_Q1D) * m_Q1D * m_Q1D, D3 = Size_type(m_D1D) * m_D1D * m_D1D;
  if (seededInputs()) {
    std::vector<Real_type> B, G;
    detail::femBasis(m_D1D, m_Q1D, B, G);
    allocData(m_B, QD);
    detail::initDataHostToDevice(m_B, B);
    allocData(m_G, QD);
    detail::initDataHostToDevice(m_G, G);
    allocData(m_D, Q3 * SYM * m_NE);
    detail::initDataSeededSymTensorDevice(m_D, m_NE, int(Q3), SYM, run_params.getSeed());
    allocData(m_X, D3 * m_NE);
    detail::initDataSeededDevice(m_X, D3 * m_NE, run_params.getSeed(), -1.0, 1.0);
    allocAndInitDataConst(m_Y, D3 * m_NE, 0.0);
  } else {  // DIFFUSION3DPA.cpp:80-88: all ones, Y zero
    allocAndInitDataConst(m_B, QD, 1.0);
    allocAndInitDataConst(m_G, QD, 1.0);
    allocAndInitDataConst(m_D, Q3 * SYM * m_NE, 1.0);
    allocAndInitDataConst(m_X, D3 * m_NE, 1.0);
    allocAndInitDataConst(m_Y, D3 * m_NE, 0.0);
  }
}
void DIFFUSION3DPA::updateChecksum(VariantID vid, size_t t) {
  checksum[vid][t] += calcChecksum(m_Y, Size_type(m_D1D) * m_D1D * m_D1D * m_NE);
}
void DIFFUSION3DPA::tearDown(VariantID, size_t) {
  deallocData(m_B);
  deallocData(m_G);
  deallocData(m_D);
  deallocData(m_X);
  deallocData(m_Y);
}
void DIFFUSION3DPA::setMI355XTuningDefinitions(VariantID vid) {
  addVariantTuningName(vid, "default");
  // int8 MFMA on 7 Ozaki slices of the Kronecker form, where the order has it
  if (rpx_diffusion3dpa_mixed_scratch_bytes(m_D1D, m_Q1D, 7) > 0) addVariantTuningName(vid, "mixed");
}
void DIFFUSION3DPA::runMI355XVariant(VariantID, size_t tune_idx) {
  const Index_type reps = getRunReps();
  if (tune_idx == 1) {  // the sliced bases' scratch is allocated outside the timed reps
    const size_t bytes = rpx_diffusion3dpa_mixed_scratch_bytes(m_D1D, m_Q1D, 7);
    char* scratch = nullptr;
    allocData(scratch, Size_type(bytes));
    timedReps(reps, [&](RepIndex_type) {
      rpxCheck(rpx_diffusion3dpa_mixed_f64(m_B, m_G, m_D, m_X, m_Y, m_NE, m_D1D, m_Q1D, 1, 7, scratch, bytes,
                                           stream()),
               "DIFFUSION3DPA mixed");
    });
    deallocData(scratch);
    return;
  }
  const rpx_tuning tune = defaultTuning();
  timedReps(reps, [&](RepIndex_type) {
    rpxCheck(rpx_diffusion3dpa_f64(m_B, m_G, m_D, m_X, m_Y, m_NE, m_D1D, m_Q1D, 1, &tune, stream()),
             "DIFFUSION3DPA");
  });
}

}  // namespace apps
}  // namespace rajaperf
