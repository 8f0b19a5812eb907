// Stream group kernels (src/stream/{COPY,MUL,ADD,TRIAD,DOT}.hpp/.cpp) with the
// MI355X_Native variant.
#pragma once

#include <functional>

#include "KernelBase.hpp"

namespace rajaperf {

// Throws like the suite's hipErrchk aborts (KernelBase.hpp:25-27).
void rpxCheck(int status, const char* what);

namespace stream {

// Tunings shared by the element-wise kernels; the last one, "auto", asks
// the autotuner, which times the others with `trial` (one rep).
void setStreamTunings(KernelBase& k, VariantID vid);
rpx_tuning streamTuning(const KernelBase& k, size_t tune_idx, const std::function<void(const rpx_tuning&)>& trial);

// Tunings of DOT and REDUCE_SUM (chunk sizes and grid caps, then "auto"),
// and the scratch size that serves all of them.
void setReduceTunings(KernelBase& k, VariantID vid);
size_t reduceScratchBytes(const KernelBase& k, Index_type n);
rpx_tuning reduceTuning(const KernelBase& k, size_t tune_idx, const std::function<void(const rpx_tuning&)>& trial);

class COPY : public KernelBase {
 public:
  explicit COPY(const RunParams& params);
  void setUp(VariantID vid, size_t tune_idx) override;
  void updateChecksum(VariantID vid, size_t tune_idx) override;
  void tearDown(VariantID vid, size_t tune_idx) override;
  void runMI355XVariant(VariantID vid, size_t tune_idx) override;
  void setMI355XTuningDefinitions(VariantID vid) override { setStreamTunings(*this, vid); }

 private:
  Real_ptr m_a = nullptr, m_c = nullptr;
};

class MUL : public KernelBase {
 public:
  explicit MUL(const RunParams& params);
  void setUp(VariantID vid, size_t tune_idx) override;
  void updateChecksum(VariantID vid, size_t tune_idx) override;
  void tearDown(VariantID vid, size_t tune_idx) override;
  void runMI355XVariant(VariantID vid, size_t tune_idx) override;
  void setMI355XTuningDefinitions(VariantID vid) override { setStreamTunings(*this, vid); }

 private:
  Real_ptr m_b = nullptr, m_c = nullptr;
  Real_type m_alpha = 0.0;
};

class ADD : public KernelBase {
 public:
  explicit ADD(const RunParams& params);
  void setUp(VariantID vid, size_t tune_idx) override;
  void updateChecksum(VariantID vid, size_t tune_idx) override;
  void tearDown(VariantID vid, size_t tune_idx) override;
  void runMI355XVariant(VariantID vid, size_t tune_idx) override;
  void setMI355XTuningDefinitions(VariantID vid) override { setStreamTunings(*this, vid); }

 private:
  Real_ptr m_a = nullptr, m_b = nullptr, m_c = nullptr;
};

class TRIAD : public KernelBase {
 public:
  explicit TRIAD(const RunParams& params);
  void setUp(VariantID vid, size_t tune_idx) override;
  void updateChecksum(VariantID vid, size_t tune_idx) override;
  void tearDown(VariantID vid, size_t tune_idx) override;
  void runMI355XVariant(VariantID vid, size_t tune_idx) override;
  void setMI355XTuningDefinitions(VariantID vid) override { setStreamTunings(*this, vid); }

 private:
  Real_ptr m_a = nullptr, m_b = nullptr, m_c = nullptr;
  Real_type m_alpha = 0.0;
};

class DOT : public KernelBase {
 public:
  explicit DOT(const RunParams& params);
  void setUp(VariantID vid, size_t tune_idx) override;
  void updateChecksum(VariantID vid, size_t tune_idx) override;
  void tearDown(VariantID vid, size_t tune_idx) override;
  void runMI355XVariant(VariantID vid, size_t tune_idx) override;
  void setMI355XTuningDefinitions(VariantID vid) override;

 private:
  Real_ptr m_a = nullptr, m_b = nullptr;
  Real_type m_dot = 0.0;
};

}  // namespace stream
}  // namespace rajaperf
