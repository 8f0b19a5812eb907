// Algorithm group: constructors and setUp/updateChecksum/tearDown follow
// src/algorithm/{REDUCE_SUM,SCAN,SORT}.cpp; runMI355XVariant holds the timed
// rep loop over the C ABI.
#include "algorithm/Algorithm.hpp"

#include "Autotune.hpp"

#include "stream/Stream.hpp"

namespace rajaperf {
namespace algorithm {

// ---- REDUCE_SUM (src/algorithm/REDUCE_SUM.cpp:22-76) ---------------------------
REDUCE_SUM::REDUCE_SUM(const RunParams& params) : KernelBase(Algorithm_REDUCE_SUM, params) {
  setDefaultProblemSize(1000000);
  setDefaultReps(50);
  setActualProblemSize(getTargetProblemSize());
  setItsPerRep(getActualProblemSize());
  setKernelsPerRep(1);
  setBytesReadPerRep(1 * sizeof(Real_type) * (1 + getActualProblemSize()));
  setBytesWrittenPerRep(1 * sizeof(Real_type));
  setBytesAtomicModifyWrittenPerRep(0);
  setFLOPsPerRep(getActualProblemSize());
  setComplexity(Complexity::N);
  setVariantDefined(MI355X_Native);
}
void REDUCE_SUM::setMI355XTuningDefinitions(VariantID vid) { stream::setReduceTunings(*this, vid); }
void REDUCE_SUM::setUp(VariantID, size_t) {
  allocAndInitData(m_x, getActualProblemSize(), -1.0, 1.0);
  m_sum = 0.0;
}
void REDUCE_SUM::updateChecksum(VariantID vid, size_t t) {
  checksum[vid].at(t) += detail::calcChecksumHost(&m_sum, 1, 1.0);  // REDUCE_SUM.cpp:73-76
  addIWeightedHost(&m_sum, 1);
}
void REDUCE_SUM::tearDown(VariantID, size_t) { deallocData(m_x); }
void REDUCE_SUM::runMI355XVariant(VariantID, size_t tune_idx) {
  const Index_type reps = getRunReps(), n = getActualProblemSize();
  const size_t sb = stream::reduceScratchBytes(*this, n);
  void* scratch = detail::deviceAlloc(sb + 256);
  rpxCheck(rpx_reduce_scratch_init(scratch, sb, nullptr), "reduce scratch init");
  Real_ptr d_sum = reinterpret_cast<Real_ptr>(static_cast<char*>(scratch) + sb);
  // Trials write the same sum the timed reps will.
  const rpx_tuning tune = stream::reduceTuning(*this, tune_idx, [&](const rpx_tuning& t) {
    rpxCheck(rpx_reduce_sum_f64(m_x, n, d_sum, scratch, sb, &t, nullptr), "REDUCE_SUM");
  });
  // m_sum = sum every rep (REDUCE_SUM-Seq.cpp:38-45): the last rep's value.
  timedReps(reps, [&](RepIndex_type) {
    rpxCheck(rpx_reduce_sum_f64(m_x, n, d_sum, scratch, sb, &tune, stream()), "REDUCE_SUM");
  });
  if (reps > 0) detail::copyToHost(&m_sum, d_sum, sizeof(Real_type));
  detail::deviceFree(scratch);
}

// ---- SCAN (src/algorithm/SCAN.cpp:22-77) ---------------------------------------
SCAN::SCAN(const RunParams& params) : KernelBase(Algorithm_SCAN, params) {
  setDefaultProblemSize(1000000);
  setDefaultReps(100);
  setActualProblemSize(getTargetProblemSize());
  setItsPerRep(getActualProblemSize());
  setKernelsPerRep(1);
  setBytesReadPerRep(1 * sizeof(Real_type) * getActualProblemSize());
  setBytesWrittenPerRep(1 * sizeof(Real_type) * getActualProblemSize());
  setBytesAtomicModifyWrittenPerRep(0);
  setFLOPsPerRep(1 * getActualProblemSize());
  checksum_scale_factor = 1e-2 * (static_cast<Checksum_type>(getDefaultProblemSize()) / getActualProblemSize()) /
                          getActualProblemSize();
  setComplexity(Complexity::N);
  setVariantDefined(MI355X_Native);
}
void SCAN::setMI355XTuningDefinitions(VariantID vid) {
  addVariantTuningName(vid, "default");  // 64 doubles per work-item
  addVariantTuningName(vid, "items_16");
  addVariantTuningName(vid, "items_32");
  addVariantTuningName(vid, "auto");
}
void SCAN::setUp(VariantID, size_t) {
  allocAndInitDataRandValue(m_x, getActualProblemSize(), -1.0, 1.0);
  allocAndInitDataConst(m_y, getActualProblemSize(), 0.0);
}
void SCAN::updateChecksum(VariantID vid, size_t t) {
  checksum[vid][t] += calcChecksum(m_y, getActualProblemSize(), Real_type(checksum_scale_factor));
}
void SCAN::tearDown(VariantID, size_t) {
  deallocData(m_x);
  deallocData(m_y);
}
void SCAN::runMI355XVariant(VariantID, size_t tune_idx) {
  const Index_type reps = getRunReps(), n = getActualProblemSize();
  const size_t sb = rpx_scan_scratch_bytes(n, nullptr);  // serves every tile size
  void* scratch = detail::deviceAlloc(sb);
  rpxCheck(rpx_scan_scratch_init(scratch, sb, nullptr), "scan scratch init");
  auto with_items = [&](int items) {
    rpx_tuning t = defaultTuning();
    t.items_per_thread = items;
    return t;
  };
  rpx_tuning tune = defaultTuning();
  if (tune_idx == 1) tune.items_per_thread = 16;
  if (tune_idx == 2) tune.items_per_thread = 32;
  if (tune_idx == 3)  // trials write the same y the timed reps will
    tune = Autotuner::instance()
               .choose(getName(), "n=" + std::to_string(n),
                       {{"default", with_items(64)}, {"items_16", with_items(16)}, {"items_32", with_items(32)}},
                       [&](const rpx_tuning& t) {
                         rpxCheck(rpx_scan_exclusive_f64(m_x, m_y, n, scratch, sb, &t, nullptr), "SCAN");
                       })
               .tuning;
  timedReps(reps, [&](RepIndex_type) {
    rpxCheck(rpx_scan_exclusive_f64(m_x, m_y, n, scratch, sb, &tune, stream()), "SCAN");
  });
  rpxCheck(rpx_scan_scratch_status(scratch), "SCAN look-back");
  detail::deviceFree(scratch);
}

// ---- SORT (src/algorithm/SORT.cpp:22-62, SORT-Seq.cpp:31-42) --------------------
SORT::SORT(const RunParams& params) : KernelBase(Algorithm_SORT, params) {
  setDefaultProblemSize(1000000);
  setDefaultReps(20);
  setActualProblemSize(getTargetProblemSize());
  setItsPerRep(getActualProblemSize());
  setKernelsPerRep(1);
  setBytesReadPerRep(1 * sizeof(Real_type) * getActualProblemSize());
  setBytesWrittenPerRep(1 * sizeof(Real_type) * getActualProblemSize());
  setBytesAtomicModifyWrittenPerRep(0);
  setFLOPsPerRep(0);
  setComplexity(Complexity::N_logN);
  setVariantDefined(MI355X_Native);
}
void SORT::setUp(VariantID, size_t) {
  allocAndInitDataRandValue(m_x, Size_type(getActualProblemSize() * getRunReps()), -1.0, 1.0);
}
void SORT::updateChecksum(VariantID vid, size_t t) {
  checksum[vid][t] += calcChecksum(m_x, Size_type(getActualProblemSize() * getRunReps()));
}
void SORT::tearDown(VariantID, size_t) { deallocData(m_x); }
void SORT::runMI355XVariant(VariantID, size_t) {
  const Index_type reps = getRunReps(), n = getActualProblemSize();
  const int64_t seg[2] = {0, n};
  const size_t sb = rpx_sort_segmented_scratch_bytes(seg, 1);
  void* scratch = sb ? detail::deviceAlloc(sb) : nullptr;
  // One segment per rep at x + n*irep (SORT.hpp:21-25).
  timedReps(reps, [&](RepIndex_type irep) {
    rpxCheck(rpx_sort_segmented_f64(m_x + n * irep, seg, 1, scratch, sb, stream()), "SORT");
  });
  if (scratch) detail::deviceFree(scratch);
}

}  // namespace algorithm
}  // namespace rajaperf
