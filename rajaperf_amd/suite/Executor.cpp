#include "Executor.hpp"

#include "Autotune.hpp"

#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace rajaperf {

namespace {
void make_dir(const std::string& dir) {  // OutputUtils.cpp:36 recursiveMkdir
  std::string cur;
  std::stringstream ss(dir);
  std::string part;
  if (!dir.empty() && dir[0] == '/') cur = "/";
  while (std::getline(ss, part, '/')) {
    if (part.empty()) continue;
    cur += part + "/";
    mkdir(cur.c_str(), 0755);
  }
}

// Long double written exactly as a two-double split (hi + lo).
void split_ld(long double v, double& hi, double& lo) {
  hi = static_cast<double>(v);
  lo = static_cast<double>(v - static_cast<long double>(hi));
}

std::string json_escape(const std::string& s) {
  std::string o;
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    o += c;
  }
  return o;
}
}  // namespace

Executor::Executor(int argc, const char** argv) : run_params(argc, argv) {}

Executor::~Executor() {
  for (KernelBase* k : kernels) delete k;
}

void Executor::setupSuite() {  // Executor.cpp:246-357
  const RunParams::InputOpt in = run_params.getInputState();
  if (in == RunParams::InfoRequest || in == RunParams::BadInput) return;
  for (KernelID kid : run_params.getKernelIDsToRun()) kernels.push_back(getKernelObject(kid, run_params));
  for (VariantID vid : run_params.getVariantIDsToRun()) variant_ids.push_back(vid);
  Autotuner::instance().reset();
  if (!run_params.getTuningCache().empty()) Autotuner::instance().setCacheFile(run_params.getTuningCache());
}

bool Executor::tuningSelected(const std::string& name) const {
  const auto& in = run_params.getTuningInput();
  const auto& ex = run_params.getExcludeTuningInput();
  bool take = in.empty();
  for (const auto& t : in) take = take || (t == name);
  for (const auto& t : ex)
    if (t == name) take = false;
  return take;
}

void Executor::reportRunSummary(std::ostream& str) const {  // Executor.cpp:360-505
  const RunParams::InputOpt in = run_params.getInputState();
  if (in == RunParams::BadInput) {
    str << "\nRunParams state:\n";
    str << "  Bad input:";
    for (const auto& s : run_params.getInvalidInput()) str << " " << s;
    str << "\n  No kernels were run.\n";
    return;
  }
  if (in == RunParams::InfoRequest) return;
  str << "\nRAJA performance suite run summary (MI355X_Native harness)\n"
      << "--------------------------------------\n";
  run_params.print(str);
  str << "\nVariants and tunings:\n";
  for (VariantID vid : variant_ids) str << "  " << getVariantName(vid) << "\n";
  str << "\nKernels (" << kernels.size() << "):\n";
  for (const KernelBase* k : kernels) {
    str << "  " << std::left << std::setw(24) << k->getName() << " size " << k->getActualProblemSize() << " reps "
        << k->getRunReps() << " bytes/rep " << k->getBytesPerRep() << " flops/rep " << k->getFLOPsPerRep() << "\n";
  }
}

void Executor::runSuite() {  // Executor.cpp:645-670
  const RunParams::InputOpt in = run_params.getInputState();
  if (in != RunParams::PerfRun && in != RunParams::CheckRun) {
    if (in == RunParams::BadInput) status = 1;
    return;
  }
  for (int ip = 0; ip < run_params.getNumPasses(); ++ip) {
    if (run_params.showProgress()) getCout() << "\nPass through suite # " << ip << "\n";
    for (KernelBase* k : kernels) runKernel(k);
  }
}

void Executor::runKernel(KernelBase* kern) {  // Executor.cpp:679-726
  if (run_params.showProgress()) getCout() << "\nRun kernel -- " << kern->getName() << "\n";
  for (VariantID vid : variant_ids) {
    if (!kern->hasVariantDefined(vid)) continue;
    for (size_t t = 0; t < kern->getNumVariantTunings(vid); ++t) {
      const std::string& tname = kern->getVariantTuningName(vid, t);
      if (!tuningSelected(tname)) continue;
      try {
        kern->execute(vid, t);
      } catch (const std::exception& e) {
        failures.push_back(kern->getName() + " " + getVariantName(vid) + "-" + tname + ": " + e.what());
        getCout() << "\n  " << failures.back() << "\n";
        status = 2;
        continue;
      }
      if (run_params.showProgress())
        getCout() << "   Running " << getVariantName(vid) << " variant, " << tname << " tuning -- " << std::scientific
                  << std::setprecision(6) << kern->getLastTime() << " sec.\n"
                  << std::defaultfloat;
    }
  }
}

void Executor::outputRunData() {  // Executor.cpp:825-883
  const RunParams::InputOpt in = run_params.getInputState();
  if (in != RunParams::PerfRun && in != RunParams::CheckRun) return;
  const std::string dir = run_params.getOutputDirName();
  make_dir(dir);
  const std::string pre = dir + "/" + run_params.getOutputFilePrefix();
  writeTimingCSV(pre + "-timing-Average.csv");
  writeChecksumReport(pre + "-checksum.txt");
  writeKernelInfoCSV(pre + "-kernels.csv");
  if (!run_params.getJsonFile().empty()) writeJson(run_params.getJsonFile());
}

void Executor::writeTimingCSV(const std::string& path) const {
  std::ofstream f(path);
  f << "Mean Runtime Report (sec.)\nKernel,Variant,Tuning,Time/rep,GB/s,GFLOP/s,Frac HBM (8 TB/s)\n";
  for (const KernelBase* k : kernels)
    for (VariantID vid : variant_ids)
      for (size_t t = 0; t < k->getNumVariantTunings(vid); ++t) {
        f << k->getName() << "," << getVariantName(vid) << "," << k->getVariantTuningName(vid, t) << ",";
        if (!k->wasVariantTuningRun(vid, t)) {
          f << "Not run,,,\n";  // unexecuted tunings are marked, not reported as 0 s
          continue;
        }
        const double per_rep = k->getTotTime(vid, t) / k->getNumExec(vid, t) / double(std::max<Index_type>(1, k->getRunReps()));
        const double gbs = k->getBytesPerRep() / per_rep / 1e9;
        f << std::setprecision(8) << per_rep << "," << gbs << "," << k->getFLOPsPerRep() / per_rep / 1e9 << ","
          << gbs / 8000.0 << "\n";
      }
}

void Executor::writeChecksumReport(const std::string& path) const {  // Executor.cpp:1281-1512
  std::ofstream f(path);
  f << "Checksum Report\n";
  for (const KernelBase* k : kernels) {
    f << k->getName() << "\n";
    bool have_ref = false;
    long double ref = 0.0L;
    for (VariantID vid : variant_ids)
      for (size_t t = 0; t < k->getNumVariantTunings(vid); ++t) {
        f << "  " << std::left << std::setw(28) << (getVariantName(vid) + "-" + k->getVariantTuningName(vid, t));
        if (!k->wasVariantTuningRun(vid, t)) {
          f << "Not Run\n";
          continue;
        }
        const long double cs = k->getChecksum(vid, t);
        if (!have_ref) {
          ref = cs;
          have_ref = true;
        }
        f << std::setprecision(20) << cs << "  diff " << std::setprecision(6) << (cs - ref) << "\n";
      }
  }
}

void Executor::writeKernelInfoCSV(const std::string& path) const {  // Executor.cpp:508-642
  std::ofstream f(path);
  f << "Kernel,Feature,Problem size,Reps,Iterations/rep,Kernels/rep,Bytes/rep,FLOPS/rep\n";
  for (const KernelBase* k : kernels)
    f << k->getName() << ",," << k->getActualProblemSize() << "," << k->getRunReps() << "," << k->getItsPerRep()
      << "," << k->getKernelsPerRep() << "," << k->getBytesPerRep() << "," << k->getFLOPsPerRep() << "\n";
}

void Executor::writeJson(const std::string& path) const {
  std::ofstream f(path);
  f << "{\"status\": " << status << ", \"failures\": [";
  for (size_t i = 0; i < failures.size(); ++i) f << (i ? ", " : "") << "\"" << json_escape(failures[i]) << "\"";
  f << "], \"runs\": [";
  bool first = true;
  for (const KernelBase* k : kernels)
    for (VariantID vid : variant_ids)
      for (size_t t = 0; t < k->getNumVariantTunings(vid); ++t) {
        if (!k->wasVariantTuningRun(vid, t)) continue;
        double hi, lo;
        split_ld(k->getChecksum(vid, t), hi, lo);
        f << (first ? "" : ", ") << std::setprecision(17) << "{\"kernel\": \"" << k->getName()
          << "\", \"variant\": \"" << getVariantName(vid) << "\", \"tuning\": \"" << k->getVariantTuningName(vid, t)
          << "\", \"checksum_hi\": " << hi << ", \"checksum_lo\": " << lo << ", \"time_total_s\": "
          << k->getTotTime(vid, t) << ", \"time_min_s\": " << k->getMinTime(vid, t) << ", \"passes\": "
          << k->getNumExec(vid, t) << ", \"reps\": " << k->getRunReps() << ", \"problem_size\": "
          << k->getActualProblemSize() << ", \"bytes_per_rep\": " << k->getBytesPerRep()
          << ", \"flops_per_rep\": " << k->getFLOPsPerRep() << "}";
        first = false;
      }
  f << "], \"autotune\": [";
  const auto& ds = Autotuner::instance().decisions();
  for (size_t i = 0; i < ds.size(); ++i)
    f << (i ? ", " : "") << std::setprecision(6) << "{\"kernel\": \"" << json_escape(ds[i].kernel)
      << "\", \"shape\": \"" << json_escape(ds[i].shape) << "\", \"arch\": \"" << json_escape(ds[i].arch)
      << "\", \"choice\": \"" << json_escape(ds[i].choice) << "\", \"trial_ms\": " << ds[i].ms_per_rep
      << ", \"from_cache\": " << (ds[i].from_cache ? "true" : "false") << "}";
  f << "]}\n";
}

}  // namespace rajaperf

// In-process entry point for tests and tools: runs the suite with the given
// command line and returns the Executor status (the counterpart of the
// suite's in-process test, test/test-raja-perf-suite.cpp:50-178).
extern "C" int rajaperf_run(int argc, const char** argv) {
  rajaperf::Executor exec(argc, argv);
  exec.setupSuite();
  exec.reportRunSummary(rajaperf::getCout());
  exec.runSuite();
  exec.outputRunData();
  return exec.getStatus();
}
