#include "KernelBase.hpp"

#include <algorithm>
#include <iostream>

namespace rajaperf {

KernelBase::KernelBase(KernelID kid, const RunParams& params)
    : run_params(params), kernel_id(kid), name(getFullKernelName(kid)) {}

Index_type KernelBase::getTargetProblemSize() const {  // KernelBase.cpp:115-125
  if (run_params.getSizeMeaning() == RunParams::Direct) return static_cast<Index_type>(run_params.getSize());
  return static_cast<Index_type>(default_prob_size * run_params.getSizeFactor());
}

Index_type KernelBase::getRunReps() const {  // KernelBase.cpp:127-136
  if (run_params.getInputState() == RunParams::CheckRun) return static_cast<Index_type>(run_params.getCheckRunReps());
  return static_cast<Index_type>(default_reps * run_params.getRepFactor());
}

void KernelBase::setVariantDefined(VariantID vid) {  // KernelBase.cpp:138-232
  if (!isVariantAvailable(vid)) return;
  if (vid == MI355X_Native) {
    setMI355XTuningDefinitions(vid);
    num_base_tunings = getNumVariantTunings(vid);
    if (run_params.getGraphMode() && graphCapturable())
      for (size_t t = 0; t < num_base_tunings; ++t)
        addVariantTuningName(vid, getVariantTuningName(vid, t) + "-graph");
  }
  const size_t n = getNumVariantTunings(vid);
  checksum[vid].assign(n, 0.0L);
  iweighted[vid].assign(n, 0.0);
  num_exec[vid].assign(n, 0);
  min_time[vid].assign(n, std::numeric_limits<double>::max());
  max_time[vid].assign(n, -std::numeric_limits<double>::max());
  tot_time[vid].assign(n, 0.0);
}

size_t KernelBase::getVariantTuningIndex(VariantID vid, const std::string& tname) const {
  const auto& names = variant_tuning_names[vid];
  for (size_t t = 0; t < names.size(); ++t)
    if (names[t] == tname) return t;
  return getUnknownTuningIdx();
}

void KernelBase::execute(VariantID vid, size_t tune_idx) {  // KernelBase.cpp:359-377
  running_variant = vid;
  running_tuning = tune_idx;
  last_time = 0.0;
  detail::resetDataInitCount();
  this->setUp(vid, tune_idx);
  this->runKernel(vid, tune_idx);
  iw_pass = 0.0;
  this->updateChecksum(vid, tune_idx);
  iweighted[vid].at(tune_idx) += iw_pass;
  this->tearDown(vid, tune_idx);
  running_variant = NumVariants;
  running_tuning = getUnknownTuningIdx();
}

void KernelBase::restoreRun(VariantID vid, size_t t, Checksum_type cs, double iw, double tot, double mn, double mx,
                            int nexec) {
  checksum[vid].at(t) = cs;
  iweighted[vid].at(t) = iw;
  tot_time[vid].at(t) = tot;
  min_time[vid].at(t) = mn;
  max_time[vid].at(t) = mx;
  num_exec[vid].at(t) = nexec;
}

void KernelBase::recordExecTime() {  // KernelBase.cpp:379-389
  num_exec[running_variant].at(running_tuning)++;
  min_time[running_variant].at(running_tuning) = std::min(min_time[running_variant].at(running_tuning), last_time);
  max_time[running_variant].at(running_tuning) = std::max(max_time[running_variant].at(running_tuning), last_time);
  tot_time[running_variant].at(running_tuning) += last_time;
}

void KernelBase::runKernel(VariantID vid, size_t tune_idx) {  // KernelBase.cpp:391-489
  if (!hasVariantDefined(vid)) return;
  switch (vid) {
    case MI355X_Native: {
      // "-graph" tunings run their base tuning with the rep loop captured
      graph_active = tune_idx >= num_base_tunings;
      const size_t base = graph_active ? tune_idx - num_base_tunings : tune_idx;
      try {
        runMI355XVariant(vid, base);
      } catch (...) {
        graph_active = false;
        throw;
      }
      graph_active = false;
      break;
    }
    default: getCout() << "\n  " << name << " : variant " << getVariantName(vid) << " is not built here\n"; break;
  }
}

void KernelBase::allocAndInitData(Real_ptr& ptr, Size_type len, Real_type lo, Real_type hi) {
  allocData(ptr, len);
  if (seededInputs()) detail::initDataSeededDevice(ptr, len, run_params.getSeed(), lo, hi);
  else detail::initDataDevice(ptr, len);
}

void KernelBase::allocAndInitDataConst(Real_ptr& ptr, Size_type len, Real_type val) {
  allocData(ptr, len);
  detail::initDataConstDevice(ptr, len, val);
}

void KernelBase::allocAndInitDataRandValue(Real_ptr& ptr, Size_type len, Real_type lo, Real_type hi) {
  allocData(ptr, len);
  if (seededInputs()) detail::initDataSeededDevice(ptr, len, run_params.getSeed(), lo, hi);
  else detail::initDataRandValueDevice(ptr, len);
}

Real_type KernelBase::initScalar(Real_type seeded_value) {
  Real_type d;
  detail::initData(d);
  return seededInputs() ? seeded_value : d;
}

rpx_tuning KernelBase::defaultTuning() const {
  rpx_tuning t{run_params.getGPUBlockSize(), run_params.getItemsPerThread(), 0, 0, 0};
  return t;
}

}  // namespace rajaperf
