// Stream group: constructors and setUp/updateChecksum/tearDown follow
// src/stream/{COPY,MUL,ADD,TRIAD,DOT}.cpp; runMI355XVariant holds the timed
// rep loop that calls the C ABI (the counterpart of <KERNEL>-Hip.cpp).
#include "stream/Stream.hpp"

#include "Autotune.hpp"

#include <stdexcept>
#include <string>

namespace rajaperf {

void rpxCheck(int status, const char* what) {
  if (status != RPX_SUCCESS) throw std::runtime_error(std::string(what) + ": " + rpx_status_string(status));
}

namespace stream {

namespace {
struct StreamTune {
  const char* name;
  int block, items, flags;
};
const StreamTune kStreamTunes[] = {
    {"default", 0, 0, 0},        {"block_128", 128, 0, 0}, {"block_512", 512, 0, 0},
    {"block_1024", 1024, 0, 0},  {"items_16", 0, 16, 0},   {"cached", 0, 0, RPX_TUNE_CACHED},
};
}  // namespace

constexpr size_t kNumStreamTunes = sizeof(kStreamTunes) / sizeof(kStreamTunes[0]);

// DOT / REDUCE_SUM: doubles per work-item and grid caps.
struct ReduceTune {
  const char* name;
  int items, grid;
};
const ReduceTune kReduceTunes[] = {
    {"default", 0, 0}, {"grid_1024", 0, 1024}, {"grid_4096", 0, 4096}, {"items_16", 16, 0}, {"items_64", 64, 0},
};
constexpr size_t kNumReduceTunes = sizeof(kReduceTunes) / sizeof(kReduceTunes[0]);

rpx_tuning fixedReduceTuning(const KernelBase& k, size_t idx) {
  rpx_tuning t = k.defaultTuning();
  if (kReduceTunes[idx].items) t.items_per_thread = kReduceTunes[idx].items;
  if (kReduceTunes[idx].grid) t.grid_size = kReduceTunes[idx].grid;
  return t;
}

rpx_tuning fixedStreamTuning(const KernelBase& k, size_t idx) {
  rpx_tuning t = k.defaultTuning();
  const StreamTune& s = kStreamTunes[idx];
  if (s.block) t.block_size = s.block;
  if (s.items) t.items_per_thread = s.items;
  t.flags |= s.flags;
  return t;
}

void setReduceTunings(KernelBase& k, VariantID vid) {
  for (const auto& t : kReduceTunes) k.addVariantTuningName(vid, t.name);
  k.addVariantTuningName(vid, "auto");
}

size_t reduceScratchBytes(const KernelBase& k, Index_type n) {
  size_t sb = 0;
  for (size_t i = 0; i < kNumReduceTunes; ++i) {
    const rpx_tuning t = fixedReduceTuning(k, i);
    const size_t b = rpx_reduce_scratch_bytes(n, &t);
    if (b > sb) sb = b;
  }
  return sb;
}

rpx_tuning reduceTuning(const KernelBase& k, size_t tune_idx, const std::function<void(const rpx_tuning&)>& trial) {
  if (tune_idx < kNumReduceTunes) return fixedReduceTuning(k, tune_idx);
  std::vector<TuneCandidate> cands;
  for (size_t i = 0; i < kNumReduceTunes; ++i) cands.push_back({kReduceTunes[i].name, fixedReduceTuning(k, i)});
  return Autotuner::instance()
      .choose(k.getName(), "n=" + std::to_string(k.getActualProblemSize()), cands, trial)
      .tuning;
}

void setStreamTunings(KernelBase& k, VariantID vid) {
  for (const auto& t : kStreamTunes) k.addVariantTuningName(vid, t.name);
  k.addVariantTuningName(vid, "auto");
}

rpx_tuning streamTuning(const KernelBase& k, size_t tune_idx, const std::function<void(const rpx_tuning&)>& trial) {
  if (tune_idx < kNumStreamTunes) return fixedStreamTuning(k, tune_idx);
  std::vector<TuneCandidate> cands;
  for (size_t i = 0; i < kNumStreamTunes; ++i) cands.push_back({kStreamTunes[i].name, fixedStreamTuning(k, i)});
  return Autotuner::instance()
      .choose(k.getName(), "n=" + std::to_string(k.getActualProblemSize()), cands, trial)
      .tuning;
}

// ---- COPY (src/stream/COPY.cpp:22-78) ------------------------------------------
COPY::COPY(const RunParams& params) : KernelBase(Stream_COPY, params) {
  setDefaultProblemSize(1000000);
  setDefaultReps(1800);
  setActualProblemSize(getTargetProblemSize());
  setItsPerRep(getActualProblemSize());
  setKernelsPerRep(1);
  setBytesReadPerRep(1 * sizeof(Real_type) * getActualProblemSize());
  setBytesWrittenPerRep(1 * sizeof(Real_type) * getActualProblemSize());
  setBytesAtomicModifyWrittenPerRep(0);
  setFLOPsPerRep(0);
  setComplexity(Complexity::N);
  setVariantDefined(MI355X_Native);
}
void COPY::setUp(VariantID, size_t) {
  allocAndInitData(m_a, getActualProblemSize(), 0.0, 1.0);
  allocAndInitDataConst(m_c, getActualProblemSize(), 0.0);
}
void COPY::updateChecksum(VariantID vid, size_t t) {
  checksum[vid][t] += calcChecksum(m_c, getActualProblemSize());
}
void COPY::tearDown(VariantID, size_t) {
  deallocData(m_a);
  deallocData(m_c);
}
void COPY::runMI355XVariant(VariantID, size_t tune_idx) {
  const Index_type reps = getRunReps(), n = getActualProblemSize();
  const rpx_tuning tune = streamTuning(*this, tune_idx, [&](const rpx_tuning& t) {
    rpxCheck(rpx_copy_f64(m_c, m_a, n, &t, nullptr), "COPY");
  });
  timedReps(reps, [&](RepIndex_type) {
    rpxCheck(rpx_copy_f64(m_c, m_a, n, &tune, stream()), "COPY");
  });
}

// ---- MUL (src/stream/MUL.cpp:22-78) --------------------------------------------
MUL::MUL(const RunParams& params) : KernelBase(Stream_MUL, params) {
  setDefaultProblemSize(1000000);
  setDefaultReps(1800);
  setActualProblemSize(getTargetProblemSize());
  setItsPerRep(getActualProblemSize());
  setKernelsPerRep(1);
  setBytesReadPerRep(1 * sizeof(Real_type) * getActualProblemSize());
  setBytesWrittenPerRep(1 * sizeof(Real_type) * getActualProblemSize());
  setBytesAtomicModifyWrittenPerRep(0);
  setFLOPsPerRep(1 * getActualProblemSize());
  setComplexity(Complexity::N);
  setVariantDefined(MI355X_Native);
}
void MUL::setUp(VariantID, size_t) {
  allocAndInitDataConst(m_b, getActualProblemSize(), 0.0);
  allocAndInitData(m_c, getActualProblemSize(), 0.0, 1.0);
  m_alpha = initScalar(0.75);
}
void MUL::updateChecksum(VariantID vid, size_t t) {
  checksum[vid][t] += calcChecksum(m_b, getActualProblemSize());
}
void MUL::tearDown(VariantID, size_t) {
  deallocData(m_b);
  deallocData(m_c);
}
void MUL::runMI355XVariant(VariantID, size_t tune_idx) {
  const Index_type reps = getRunReps(), n = getActualProblemSize();
  const rpx_tuning tune = streamTuning(*this, tune_idx, [&](const rpx_tuning& t) {
    rpxCheck(rpx_mul_f64(m_b, m_c, m_alpha, n, &t, nullptr), "MUL");
  });
  timedReps(reps, [&](RepIndex_type) {
    rpxCheck(rpx_mul_f64(m_b, m_c, m_alpha, n, &tune, stream()), "MUL");
  });
}

// ---- ADD (src/stream/ADD.cpp:22-78) --------------------------------------------
ADD::ADD(const RunParams& params) : KernelBase(Stream_ADD, params) {
  setDefaultProblemSize(1000000);
  setDefaultReps(1000);
  setActualProblemSize(getTargetProblemSize());
  setItsPerRep(getActualProblemSize());
  setKernelsPerRep(1);
  setBytesReadPerRep(2 * sizeof(Real_type) * getActualProblemSize());
  setBytesWrittenPerRep(1 * sizeof(Real_type) * getActualProblemSize());
  setBytesAtomicModifyWrittenPerRep(0);
  setFLOPsPerRep(1 * getActualProblemSize());
  setComplexity(Complexity::N);
  setVariantDefined(MI355X_Native);
}
void ADD::setUp(VariantID, size_t) {
  allocAndInitData(m_a, getActualProblemSize(), 0.0, 1.0);
  allocAndInitData(m_b, getActualProblemSize(), 0.0, 1.0);
  allocAndInitDataConst(m_c, getActualProblemSize(), 0.0);
}
void ADD::updateChecksum(VariantID vid, size_t t) {
  checksum[vid][t] += calcChecksum(m_c, getActualProblemSize());
}
void ADD::tearDown(VariantID, size_t) {
  deallocData(m_a);
  deallocData(m_b);
  deallocData(m_c);
}
void ADD::runMI355XVariant(VariantID, size_t tune_idx) {
  const Index_type reps = getRunReps(), n = getActualProblemSize();
  const rpx_tuning tune = streamTuning(*this, tune_idx, [&](const rpx_tuning& t) {
    rpxCheck(rpx_add_f64(m_c, m_a, m_b, n, &t, nullptr), "ADD");
  });
  timedReps(reps, [&](RepIndex_type) {
    rpxCheck(rpx_add_f64(m_c, m_a, m_b, n, &tune, stream()), "ADD");
  });
}

// ---- TRIAD (src/stream/TRIAD.cpp:22-84) ----------------------------------------
TRIAD::TRIAD(const RunParams& params) : KernelBase(Stream_TRIAD, params) {
  setDefaultProblemSize(1000000);
  setDefaultReps(1000);
  setActualProblemSize(getTargetProblemSize());
  setItsPerRep(getActualProblemSize());
  setKernelsPerRep(1);
  setBytesReadPerRep(2 * sizeof(Real_type) * getActualProblemSize());
  setBytesWrittenPerRep(1 * sizeof(Real_type) * getActualProblemSize());
  setBytesAtomicModifyWrittenPerRep(0);
  setFLOPsPerRep(2 * getActualProblemSize());
  checksum_scale_factor = 0.001 * (static_cast<Checksum_type>(getDefaultProblemSize()) / getActualProblemSize());
  setComplexity(Complexity::N);
  setVariantDefined(MI355X_Native);
}
void TRIAD::setUp(VariantID, size_t) {
  allocAndInitDataConst(m_a, getActualProblemSize(), 0.0);
  allocAndInitData(m_b, getActualProblemSize(), 0.0, 1.0);
  allocAndInitData(m_c, getActualProblemSize(), 0.0, 1.0);
  m_alpha = initScalar(0.75);
}
void TRIAD::updateChecksum(VariantID vid, size_t t) {
  checksum[vid][t] += calcChecksum(m_a, getActualProblemSize(), Real_type(checksum_scale_factor));
}
void TRIAD::tearDown(VariantID, size_t) {
  deallocData(m_a);
  deallocData(m_b);
  deallocData(m_c);
}
void TRIAD::runMI355XVariant(VariantID, size_t tune_idx) {
  const Index_type reps = getRunReps(), n = getActualProblemSize();
  const rpx_tuning tune = streamTuning(*this, tune_idx, [&](const rpx_tuning& t) {
    rpxCheck(rpx_triad_f64(m_a, m_b, m_c, m_alpha, n, &t, nullptr), "TRIAD");
  });
  timedReps(reps, [&](RepIndex_type) {
    rpxCheck(rpx_triad_f64(m_a, m_b, m_c, m_alpha, n, &tune, stream()), "TRIAD");
  });
}

// ---- DOT (src/stream/DOT.cpp:22-81, DOT-Seq.cpp:34-51) -------------------------
DOT::DOT(const RunParams& params) : KernelBase(Stream_DOT, params) {
  setDefaultProblemSize(1000000);
  setDefaultReps(2000);
  setActualProblemSize(getTargetProblemSize());
  setItsPerRep(getActualProblemSize());
  setKernelsPerRep(1);
  setBytesReadPerRep(1 * sizeof(Real_type) + 2 * sizeof(Real_type) * getActualProblemSize());
  setBytesWrittenPerRep(1 * sizeof(Real_type));
  setBytesAtomicModifyWrittenPerRep(0);
  setFLOPsPerRep(2 * getActualProblemSize());
  setComplexity(Complexity::N);
  setVariantDefined(MI355X_Native);
}
void DOT::setMI355XTuningDefinitions(VariantID vid) { setReduceTunings(*this, vid); }
void DOT::setUp(VariantID, size_t) {
  allocAndInitData(m_a, getActualProblemSize(), 0.0, 1.0);
  allocAndInitData(m_b, getActualProblemSize(), 0.0, 1.0);
  m_dot = 0.0;
}
void DOT::updateChecksum(VariantID vid, size_t t) {
  checksum[vid][t] += m_dot;
  addIWeightedHost(&m_dot, 1);
}
void DOT::tearDown(VariantID, size_t) {
  deallocData(m_a);
  deallocData(m_b);
}
void DOT::runMI355XVariant(VariantID, size_t tune_idx) {
  const Index_type reps = getRunReps(), n = getActualProblemSize();
  // Scratch sized for the largest chunk count of any tuning (the smallest
  // chunks); a trial's result goes to its own scalar.
  const size_t sb = reduceScratchBytes(*this, n);
  void* scratch = detail::deviceAlloc(sb + 512);
  rpxCheck(rpx_reduce_scratch_init(scratch, sb, nullptr), "reduce scratch init");
  Real_ptr d_dot = reinterpret_cast<Real_ptr>(static_cast<char*>(scratch) + sb);
  Real_ptr d_trial = d_dot + 32;
  rpxCheck(rpx_fill_const_f64(d_dot, 1, 0.0, nullptr), "DOT init");
  const rpx_tuning tune = reduceTuning(*this, tune_idx, [&](const rpx_tuning& t) {
    rpxCheck(rpx_dot_f64(m_a, m_b, n, d_trial, 0, scratch, sb, &t, nullptr), "DOT");
  });
  // m_dot += dot every rep (DOT-Seq.cpp:39-45), accumulated on the device in
  // rep order and read back once after the timed loop.
  timedReps(reps, [&](RepIndex_type) {
    rpxCheck(rpx_dot_f64(m_a, m_b, n, d_dot, 1, scratch, sb, &tune, stream()), "DOT");
  });
  Real_type dot = 0.0;
  detail::copyToHost(&dot, d_dot, sizeof(Real_type));
  m_dot += dot;
  detail::deviceFree(scratch);
}

}  // namespace stream
}  // namespace rajaperf
