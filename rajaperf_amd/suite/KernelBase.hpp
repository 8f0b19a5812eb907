// KernelBase: the suite's kernel plugin API (src/common/KernelBase.hpp:81-600,
// KernelBase.cpp), restated for the MI355X_Native variant.
//
// Same contract as the suite: the constructor sets problem shape, reps,
// bytes/FLOPs per rep and the defined variants/tunings; execute() runs
// resetDataInitCount -> setUp -> runKernel -> updateChecksum -> tearDown
// (KernelBase.cpp:359-377); runKernel dispatches on the VariantID to
// runMI355XVariant (the counterpart of runHipVariant); startTimer/stopTimer
// bracket the timed rep loop with a device synchronisation
// (KernelBase.hpp:271-294, 423-442).  Per-kernel data lives in HBM.  The
// rep loops go through timedReps, which can also replay them as HIP graphs.
#pragma once

#include <chrono>
#include <limits>
#include <string>
#include <vector>

#include "DataUtils.hpp"
#include "RAJAPerfSuite.hpp"
#include "RPTypes.hpp"
#include "RunParams.hpp"
#include "rajaperf_mi355x.h"

namespace rajaperf {

class KernelBase {
 public:
  static constexpr size_t getUnknownTuningIdx() { return std::numeric_limits<size_t>::max(); }
  static std::string getDefaultTuningName() { return "default"; }

  KernelBase(KernelID kid, const RunParams& params);
  virtual ~KernelBase() = default;

  KernelID getKernelID() const { return kernel_id; }
  const std::string& getName() const { return name; }

  // Setters used by kernel constructors (KernelBase.hpp:100-116).
  void setDefaultProblemSize(Index_type size) { default_prob_size = size; }
  void setActualProblemSize(Index_type size) { actual_prob_size = size; }
  void setDefaultReps(Index_type reps) { default_reps = reps; }
  void setItsPerRep(Index_type its) { its_per_rep = its; }
  void setKernelsPerRep(Index_type nk) { kernels_per_rep = nk; }
  void setBytesReadPerRep(Index_type b) { bytes_read_per_rep = b; }
  void setBytesWrittenPerRep(Index_type b) { bytes_written_per_rep = b; }
  void setBytesAtomicModifyWrittenPerRep(Index_type b) { bytes_atomic_modify_written_per_rep = b; }
  void setFLOPsPerRep(Index_type f) { FLOPs_per_rep = f; }
  void setComplexity(Complexity ac) { complexity = ac; }
  void setVariantDefined(VariantID vid);
  void addVariantTuningName(VariantID vid, std::string n) { variant_tuning_names[vid].emplace_back(std::move(n)); }

  // Tunings of the new variant; the default is the library's tuned choice.
  virtual void setMI355XTuningDefinitions(VariantID vid) { addVariantTuningName(vid, getDefaultTuningName()); }
  // Whether the rep loop can be captured in a HIP graph (under --graph each
  // tuning then also runs as "<tuning>-graph").  Kernels whose calls upload
  // per-call host tables say no: a graph would replay the first call's copy.
  virtual bool graphCapturable() const { return true; }

  Index_type getDefaultProblemSize() const { return default_prob_size; }
  Index_type getActualProblemSize() const { return actual_prob_size; }
  Index_type getDefaultReps() const { return default_reps; }
  Index_type getItsPerRep() const { return its_per_rep; }
  Index_type getKernelsPerRep() const { return kernels_per_rep; }
  Index_type getBytesPerRep() const {
    return bytes_read_per_rep + bytes_written_per_rep + 2 * bytes_atomic_modify_written_per_rep;
  }
  Index_type getBytesReadPerRep() const { return bytes_read_per_rep; }
  Index_type getBytesWrittenPerRep() const { return bytes_written_per_rep; }
  Index_type getFLOPsPerRep() const { return FLOPs_per_rep; }
  Complexity getComplexity() const { return complexity; }

  Index_type getTargetProblemSize() const;  // KernelBase.cpp:115-125
  Index_type getRunReps() const;            // KernelBase.cpp:127-136

  bool hasVariantDefined(VariantID vid) const { return !variant_tuning_names[vid].empty(); }
  size_t getNumVariantTunings(VariantID vid) const { return variant_tuning_names[vid].size(); }
  const std::string& getVariantTuningName(VariantID vid, size_t t) const { return variant_tuning_names[vid].at(t); }
  const std::vector<std::string>& getVariantTuningNames(VariantID vid) const { return variant_tuning_names[vid]; }
  size_t getVariantTuningIndex(VariantID vid, const std::string& tname) const;

  bool wasVariantTuningRun(VariantID vid, size_t t) const {
    return t != getUnknownTuningIdx() && t < num_exec[vid].size() && num_exec[vid][t] > 0;
  }
  double getLastTime() const { return last_time; }
  double getMinTime(VariantID vid, size_t t) const { return min_time[vid].at(t); }
  double getMaxTime(VariantID vid, size_t t) const { return max_time[vid].at(t); }
  double getTotTime(VariantID vid, size_t t) const { return tot_time[vid].at(t); }
  int getNumExec(VariantID vid, size_t t) const { return num_exec[vid].at(t); }
  Checksum_type getChecksum(VariantID vid, size_t t) const { return checksum[vid].at(t); }
  // The i-weighted FP64 sums of the checksummed arrays, over passes.
  double getIWeighted(VariantID vid, size_t t) const { return iweighted[vid].at(t); }

  void execute(VariantID vid, size_t tune_idx);
  // A run recorded by an earlier invocation (restored under --resume).
  void restoreRun(VariantID vid, size_t t, Checksum_type cs, double iw, double tot, double mn, double mx,
                  int nexec);

  // The stream of the timed rep loop's device work: the null stream, or the
  // capture stream in a "-graph" tuning.
  void* stream() const { return graph_active ? detail::captureStream() : nullptr; }

  // The timed rep loop: startTimer, `reps` calls of body(irep), stopTimer
  // (KernelBase.hpp:423-442 around the loops of every run<Backend>Variant).
  // In a "-graph" tuning all reps are captured into one HIP graph before the
  // timer starts (so capture and instantiation are not timed) and the timer
  // brackets one replay.  Either way the loop is a named roctx range,
  // "<kernel>.<variant>.<tuning>", for profilers.
  template <typename Body>
  void timedReps(Index_type reps, Body&& body) {
    const std::string range = name + "." + getVariantName(running_variant) + "." +
                              getVariantTuningName(running_variant, running_tuning);
    if (!graph_active) {
      detail::rangePush(range);
      startTimer();
      for (RepIndex_type irep = 0; irep < reps; ++irep) body(irep);
      stopTimer();
      detail::rangePop();
      return;
    }
    void* s = detail::captureStream();
    void* graph = nullptr;
    detail::beginCapture(s);
    try {
      for (RepIndex_type irep = 0; irep < reps; ++irep) body(irep);
    } catch (...) {
      detail::abandonCapture(s);
      throw;
    }
    graph = detail::endCapture(s);
    detail::rangePush(range);
    startTimer();
    detail::launchGraph(graph, s);
    stopTimer();
    detail::rangePop();
    detail::destroyGraph(graph);
  }

  void synchronize() {
    if (running_variant == MI355X_Native) detail::deviceSynchronize();
  }
  void startTimer() {
    synchronize();
    t0 = std::chrono::steady_clock::now();
  }
  void stopTimer() {
    synchronize();
    last_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    recordExecTime();
  }

  // Data helpers for the variant's data space (device memory).
  template <typename T>
  void allocData(T*& ptr, Size_type len) {
    ptr = static_cast<T*>(detail::deviceAlloc(len * sizeof(T)));
  }
  template <typename T>
  void deallocData(T*& ptr) {
    detail::deviceFree(ptr);
    ptr = nullptr;
  }
  // allocAndInitData: the suite formula in reference mode, uniform [lo,hi) in
  // seeded mode (DataUtils.hpp:317-327).
  void allocAndInitData(Real_ptr& ptr, Size_type len, Real_type lo, Real_type hi);
  void allocAndInitDataConst(Real_ptr& ptr, Size_type len, Real_type val);
  void allocAndInitDataRandValue(Real_ptr& ptr, Size_type len, Real_type lo, Real_type hi);
  // Scalar initialiser: the suite formula, or `seeded_value` in seeded mode.
  Real_type initScalar(Real_type seeded_value);
  bool seededInputs() const { return run_params.getInputMode() == RunParams::SeededInputs; }

  long double calcChecksum(const Real_type* dptr, Size_type len, Real_type scale = 1.0) {
    return detail::calcChecksumDevice(dptr, len, scale, &iw_pass);
  }
  // A host-side result's contribution to the i-weighted sum (scalars).
  void addIWeightedHost(const Real_type* h, Size_type len) { iw_pass += detail::iWeightedSumHost(h, len); }

  // Tuning for the variant: explicit values from --gpu_block_size /
  // --items_per_thread apply to the default tuning.
  rpx_tuning defaultTuning() const;

  virtual void runKernel(VariantID vid, size_t tune_idx);
  virtual void setUp(VariantID vid, size_t tune_idx) = 0;
  virtual void updateChecksum(VariantID vid, size_t tune_idx) = 0;
  virtual void tearDown(VariantID vid, size_t tune_idx) = 0;
  virtual void runMI355XVariant(VariantID vid, size_t tune_idx) = 0;

 protected:
  const RunParams& run_params;
  std::vector<Checksum_type> checksum[NumVariants];
  std::vector<double> iweighted[NumVariants];
  double iw_pass = 0.0;  // this pass's i-weighted sum, added after updateChecksum
  Checksum_type checksum_scale_factor = 1.0;

 private:
  void recordExecTime();

  KernelID kernel_id;
  std::string name;
  Index_type default_prob_size = -1, default_reps = -1, actual_prob_size = -1;
  Index_type its_per_rep = -1, kernels_per_rep = -1;
  Index_type bytes_read_per_rep = -1, bytes_written_per_rep = -1, bytes_atomic_modify_written_per_rep = -1;
  Index_type FLOPs_per_rep = -1;
  Complexity complexity = NumComplexities;
  std::vector<std::string> variant_tuning_names[NumVariants];
  size_t num_base_tunings = 0;  // tunings before the "-graph" copies
  bool graph_active = false;
  VariantID running_variant = NumVariants;
  size_t running_tuning = getUnknownTuningIdx();
  std::vector<int> num_exec[NumVariants];
  std::vector<double> min_time[NumVariants], max_time[NumVariants], tot_time[NumVariants];
  std::chrono::steady_clock::time_point t0;
  double last_time = 0.0;
};

}  // namespace rajaperf
