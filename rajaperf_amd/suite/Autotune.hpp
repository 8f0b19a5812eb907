// Autotuner for the MI355X_Native variant: replaces the suite's static
// block-size tuning lists (CMake RAJA_PERFSUITE_GPU_BLOCKSIZES,
// rajaperf_config.hpp.in:121-128, GPUUtils.hpp:146-179, surfaced as
// block_<n> tunings, GPUUtils.hpp:345-373) by measurement.  A kernel that
// offers the "auto" tuning lists its candidate rpx_tuning values; the
// first time a (kernel, problem shape, device architecture) is seen the
// autotuner times every candidate with HIP events on the kernel's own data
// (before the timed region) and keeps the fastest.  Decisions are cached in
// memory and, with --tuning-cache <file>, in a text file that later runs
// read instead of measuring again.  Each decision is reported under the
// kernel's "auto" tuning in the JSON output.
#pragma once

#include <functional>
#include <string>
#include <vector>

#include "rajaperf_mi355x.h"

namespace rajaperf {

struct TuneCandidate {
  std::string name;
  rpx_tuning tuning;
};

struct TuneDecision {
  std::string kernel, shape, arch, choice;
  double ms_per_rep;  // the chosen candidate's trial time
  bool from_cache;
};

class Autotuner {
 public:
  static Autotuner& instance();

  // Forgets every decision and the cache file (each suite run starts here).
  void reset();

  // Reads decisions from `path` (if it exists) and writes every new one to it.
  void setCacheFile(const std::string& path);

  // The fastest of `cands` for this kernel and shape.  `trial` runs one rep
  // with a given tuning and must leave the kernel's results as the timed
  // reps will (it runs before them).
  const TuneCandidate& choose(const std::string& kernel, const std::string& shape,
                              const std::vector<TuneCandidate>& cands,
                              const std::function<void(const rpx_tuning&)>& trial);

  const std::vector<TuneDecision>& decisions() const { return made; }

 private:
  Autotuner() = default;
  void save() const;
  std::string arch();

  std::string cache_path, arch_name;
  std::vector<TuneDecision> known;  // every decision, from the file or measured
  std::vector<TuneDecision> made;   // the decisions used by this run
};

}  // namespace rajaperf
