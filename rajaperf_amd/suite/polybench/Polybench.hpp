// Polybench kernels (src/polybench/POLYBENCH_{GEMM,JACOBI_2D}.hpp/.cpp) with
// the MI355X_Native variant.
#pragma once

#include "KernelBase.hpp"

namespace rajaperf {
namespace polybench {

class POLYBENCH_GEMM : public KernelBase {
 public:
  explicit POLYBENCH_GEMM(const RunParams& params);
  void setUp(VariantID vid, size_t tune_idx) override;
  void updateChecksum(VariantID vid, size_t tune_idx) override;
  void tearDown(VariantID vid, size_t tune_idx) override;
  void runMI355XVariant(VariantID vid, size_t tune_idx) override;
  void setMI355XTuningDefinitions(VariantID vid) override;

  Index_type ni() const { return m_ni; }
  Index_type nk() const { return m_nk; }

 private:
  Index_type m_ni = 0, m_nj = 0, m_nk = 0;
  Real_type m_alpha = 0.0, m_beta = 0.0;
  Real_ptr m_A = nullptr, m_B = nullptr, m_C = nullptr;
};

class POLYBENCH_JACOBI_2D : public KernelBase {
 public:
  explicit POLYBENCH_JACOBI_2D(const RunParams& params);
  void setUp(VariantID vid, size_t tune_idx) override;
  void updateChecksum(VariantID vid, size_t tune_idx) override;
  void tearDown(VariantID vid, size_t tune_idx) override;
  void runMI355XVariant(VariantID vid, size_t tune_idx) override;
  void setMI355XTuningDefinitions(VariantID vid) override;

 private:
  Index_type m_N = 0, m_tsteps = 0;
  Real_ptr m_A = nullptr, m_B = nullptr, m_Ainit = nullptr, m_Binit = nullptr;
};

}  // namespace polybench
}  // namespace rajaperf
