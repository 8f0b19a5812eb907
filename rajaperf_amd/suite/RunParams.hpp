// Run parameters and command-line parsing (src/common/RunParams.hpp,
// RunParams.cpp:31-90 defaults, :296-1268 parser, :1893-2490 selection).
//
// Kept from the suite: -k/--kernels, -ek/--exclude-kernels, -v/--variants,
// -ev/--exclude-variants, -t/--tunings, -et/--exclude-tunings, --size,
// --sizefact, --repfact, --npasses, --checkrun, --gpu_block_size,
// --items_per_thread, -od/--outdir, -of/--outfile, --refvar, --dryrun,
// --disable-warmup, --show-progress, --print-kernels, --print-variants,
// -h/--help.  Added for this variant: --input-mode {reference,seeded},
// --seed, --tsteps (JACOBI_2D), --nk/--alpha/--beta (GEMM), --d1d/--q1d (FEM orders),
// --json <file> (machine-readable results), --tuning-cache <file> (the
// autotuner's decisions for the "auto" tunings), --graph (graph-replayed rep
// loops).  --npasses-combiners is the suite's.
#pragma once

#include <cstdint>
#include <set>
#include <string>
#include <vector>

#include "RAJAPerfSuite.hpp"

namespace rajaperf {

class RunParams {
 public:
  enum InputOpt { InfoRequest, DryRun, CheckRun, PerfRun, BadInput, Undefined };
  enum SizeMeaning { Unset, Factor, Direct };
  enum InputMode { ReferenceInputs, SeededInputs };

  RunParams(int argc, const char** argv);

  InputOpt getInputState() const { return input_state; }
  SizeMeaning getSizeMeaning() const { return size_meaning; }
  double getSizeFactor() const { return size_factor; }
  std::int64_t getSize() const { return size; }
  double getRepFactor() const { return rep_fact; }
  int getNumPasses() const { return npasses; }
  // How the per-pass times of a run are combined in the timing reports:
  // "Average", "Minimum", "Maximum" (RunParams.hpp CombinerOpt).
  const std::vector<std::string>& getNpassesCombiners() const { return npasses_combiners; }
  // --graph: every MI355X tuning is also run as "<tuning>-graph", its timed
  // rep loop captured in a HIP graph outside the timer and replayed inside.
  bool getGraphMode() const { return graph_mode; }
  int getCheckRunReps() const { return checkrun_reps; }
  bool showProgress() const { return show_progress; }
  bool getDisableWarmup() const { return disable_warmup; }
  const std::string& getOutputDirName() const { return outdir; }
  const std::string& getOutputFilePrefix() const { return outfile_prefix; }
  const std::string& getJsonFile() const { return json_file; }
  const std::string& getTuningCache() const { return tuning_cache; }
  // One JSON line per finished (kernel, variant, tuning) run, appended as it
  // ends; with --resume, runs already in the file are restored, not rerun.
  const std::string& getProgressFile() const { return progress_file; }
  bool getResume() const { return resume; }
  // rocprofv3 counter collection (CSV) joined into the roofline report.
  const std::string& getCountersFile() const { return counters_file; }
  InputMode getInputMode() const { return input_mode; }
  std::uint64_t getSeed() const { return seed; }
  std::int64_t getTsteps() const { return tsteps; }
  std::int64_t getNk() const { return nk; }
  // GEMM scalars from --alpha / --beta; has* false = the suite's (or the
  // seeded workload's) values.
  bool hasGemmAlpha() const { return has_alpha; }
  bool hasGemmBeta() const { return has_beta; }
  double getGemmAlpha() const { return gemm_alpha; }
  double getGemmBeta() const { return gemm_beta; }
  int getD1D() const { return d1d; }
  int getQ1D() const { return q1d; }
  int getGPUBlockSize() const { return gpu_block_size; }
  int getItemsPerThread() const { return items_per_thread; }

  const std::set<KernelID>& getKernelIDsToRun() const { return run_kernels; }
  const std::set<VariantID>& getVariantIDsToRun() const { return run_variants; }
  VariantID getReferenceVariantID() const { return reference_vid; }
  // Tuning selection: empty = all.
  const std::vector<std::string>& getTuningInput() const { return tuning_input; }
  const std::vector<std::string>& getExcludeTuningInput() const { return exclude_tuning_input; }
  const std::vector<std::string>& getInvalidInput() const { return invalid_input; }

  void printHelpMessage(std::ostream& str) const;
  void print(std::ostream& str) const;

 private:
  void parse(int argc, const char** argv);
  void processKernelInput();
  void processVariantInput();

  InputOpt input_state = Undefined;
  SizeMeaning size_meaning = Unset;
  double size_factor = 1.0;
  std::int64_t size = 0;
  double rep_fact = 1.0;
  int npasses = 1;
  std::vector<std::string> npasses_combiners;
  bool graph_mode = false;
  int checkrun_reps = 1;
  bool show_progress = false;
  bool disable_warmup = false;
  bool print_kernels = false;
  bool print_variants = false;
  std::string outdir = ".";
  std::string outfile_prefix = "RAJAPerf";
  std::string json_file;
  std::string tuning_cache;
  std::string progress_file;
  bool resume = false;
  std::string counters_file;
  InputMode input_mode = ReferenceInputs;
  std::uint64_t seed = 42;
  std::int64_t tsteps = 0;
  std::int64_t nk = 0;
  bool has_alpha = false, has_beta = false;
  double gemm_alpha = 0.0, gemm_beta = 0.0;
  int d1d = 0, q1d = 0;
  int gpu_block_size = 0;
  int items_per_thread = 0;

  std::vector<std::string> kernel_input, exclude_kernel_input;
  std::vector<std::string> variant_input, exclude_variant_input;
  std::vector<std::string> tuning_input, exclude_tuning_input;
  std::vector<std::string> invalid_input;
  std::string reference_variant;

  std::set<KernelID> run_kernels;
  std::set<VariantID> run_variants;
  VariantID reference_vid = NumVariants;
};

}  // namespace rajaperf
