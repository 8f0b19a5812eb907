// Algorithm group kernels (src/algorithm/{REDUCE_SUM,SCAN,SORT}.hpp/.cpp)
// with the MI355X_Native variant.
#pragma once

#include "KernelBase.hpp"

namespace rajaperf {
namespace algorithm {

class REDUCE_SUM : public KernelBase {
 public:
  explicit REDUCE_SUM(const RunParams& params);
  void setUp(VariantID vid, size_t tune_idx) override;
  void updateChecksum(VariantID vid, size_t tune_idx) override;
  void tearDown(VariantID vid, size_t tune_idx) override;
  void runMI355XVariant(VariantID vid, size_t tune_idx) override;
  void setMI355XTuningDefinitions(VariantID vid) override;

 private:
  Real_ptr m_x = nullptr;
  Real_type m_sum = 0.0;
};

class SCAN : public KernelBase {
 public:
  explicit SCAN(const RunParams& params);
  void setMI355XTuningDefinitions(VariantID vid) override;
  void setUp(VariantID vid, size_t tune_idx) override;
  void updateChecksum(VariantID vid, size_t tune_idx) override;
  void tearDown(VariantID vid, size_t tune_idx) override;
  void runMI355XVariant(VariantID vid, size_t tune_idx) override;

 private:
  Real_ptr m_x = nullptr, m_y = nullptr;
};

class SORT : public KernelBase {
 public:
  explicit SORT(const RunParams& params);
  // Each call uploads its segment table from host staging memory, which a
  // graph replay would not redo: no "-graph" tunings.
  bool graphCapturable() const override { return false; }
  void setUp(VariantID vid, size_t tune_idx) override;
  void updateChecksum(VariantID vid, size_t tune_idx) override;
  void tearDown(VariantID vid, size_t tune_idx) override;
  void runMI355XVariant(VariantID vid, size_t tune_idx) override;

 private:
  Real_ptr m_x = nullptr;
};

}  // namespace algorithm
}  // namespace rajaperf
