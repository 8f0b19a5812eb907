// Apps kernels (src/apps/{MASS3DPA,DIFFUSION3DPA}.hpp/.cpp) with the
// MI355X_Native variant.  The polynomial orders are runtime values here
// (--d1d/--q1d); the suite compiles MASS3DPA at (4,5) and DIFFUSION3DPA at
// (3,4), which are the defaults.
#pragma once

#include "KernelBase.hpp"

namespace rajaperf {
namespace apps {

class MASS3DPA : public KernelBase {
 public:
  explicit MASS3DPA(const RunParams& params);
  void setMI355XTuningDefinitions(VariantID vid) override;
  void setUp(VariantID vid, size_t tune_idx) override;
  void updateChecksum(VariantID vid, size_t tune_idx) override;
  void tearDown(VariantID vid, size_t tune_idx) override;
  void runMI355XVariant(VariantID vid, size_t tune_idx) override;

 private:
  int m_D1D = 4, m_Q1D = 5;
  Index_type m_NE = 0;
  Real_ptr m_B = nullptr, m_Bt = nullptr, m_D = nullptr, m_X = nullptr, m_Y = nullptr;
};

class DIFFUSION3DPA : public KernelBase {
 public:
  explicit DIFFUSION3DPA(const RunParams& params);
  void setUp(VariantID vid, size_t tune_idx) override;
  void updateChecksum(VariantID vid, size_t tune_idx) override;
  void tearDown(VariantID vid, size_t tune_idx) override;
  void runMI355XVariant(VariantID vid, size_t tune_idx) override;
  void setMI355XTuningDefinitions(VariantID vid) override;

 private:
  static constexpr int SYM = 6;
  int m_D1D = 3, m_Q1D = 4;
  Index_type m_NE = 0;
  Real_ptr m_B = nullptr, m_G = nullptr, m_D = nullptr, m_X = nullptr, m_Y = nullptr;
};

}  // namespace apps
}  // namespace rajaperf
