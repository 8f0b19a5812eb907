// Executor: suite assembly, run loops and reports (src/common/Executor.hpp,
// Executor.cpp:246-357 setupSuite, :645-726 runSuite/runKernel,
// :825-1659 outputRunData).
#pragma once

#include <memory>
#include <set>
#include <string>
#include <vector>

#include "KernelBase.hpp"
#include "RunParams.hpp"

namespace rajaperf {

class Executor {
 public:
  Executor(int argc, const char** argv);
  ~Executor();

  void setupSuite();
  void reportRunSummary(std::ostream& str) const;
  void runSuite();
  void outputRunData();

  const RunParams& getRunParams() const { return run_params; }
  const std::vector<KernelBase*>& getKernels() const { return kernels; }
  const std::vector<VariantID>& getVariantIDs() const { return variant_ids; }
  // 0 after a successful run, 1 for bad input, 2 when a kernel failed.
  int getStatus() const { return status; }

 private:
  void runKernel(KernelBase* kern);
  void writeTimingCSV(const std::string& path, const std::string& combiner) const;
  void writeChecksumReport(const std::string& path) const;
  void writeKernelInfoCSV(const std::string& path) const;
  void writeJson(const std::string& path) const;
  void writeRooflineCSV(const std::string& path) const;
  void loadProgress();
  void appendProgress(const KernelBase* k, VariantID vid, size_t t) const;
  bool tuningSelected(const std::string& name) const;

  RunParams run_params;
  std::vector<KernelBase*> kernels;
  std::vector<VariantID> variant_ids;
  std::vector<std::string> failures;
  std::set<std::string> restored;  // "kernel|variant|tuning" restored from the progress file
  int status = 0;
};

}  // namespace rajaperf
