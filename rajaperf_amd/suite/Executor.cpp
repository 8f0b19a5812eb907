#include "Executor.hpp"

#include "Autotune.hpp"

#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
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

std::string run_key(const std::string& k, const std::string& v, const std::string& t) { return k + "|" + v + "|" + t; }

// Field `key` of one flat JSON object line as written by appendProgress:
// a string (escapes as json_escape writes them) or a number.
bool json_string(const std::string& line, const std::string& key, std::string& out) {
  const std::string pat = "\"" + key + "\": \"";
  size_t p = line.find(pat);
  if (p == std::string::npos) return false;
  out.clear();
  for (p += pat.size(); p < line.size(); ++p) {
    if (line[p] == '\\' && p + 1 < line.size()) {
      out += line[++p];
    } else if (line[p] == '"') {
      return true;
    } else {
      out += line[p];
    }
  }
  return false;
}
bool json_number(const std::string& line, const std::string& key, double& out) {
  const std::string pat = "\"" + key + "\": ";
  const size_t p = line.find(pat);
  if (p == std::string::npos) return false;
  const char* b = line.c_str() + p + pat.size();
  char* e = nullptr;
  out = std::strtod(b, &e);
  return e != b;
}

// One CSV record with quoted fields ("" inside quotes is a quote).
std::vector<std::string> csv_fields(const std::string& line) {
  std::vector<std::string> f(1);
  bool q = false;
  for (size_t i = 0; i < line.size(); ++i) {
    const char c = line[i];
    if (q) {
      if (c == '"' && i + 1 < line.size() && line[i + 1] == '"') {
        f.back() += '"';
        ++i;
      } else if (c == '"') {
        q = false;
      } else {
        f.back() += c;
      }
    } else if (c == '"') {
      q = true;
    } else if (c == ',') {
      f.emplace_back();
    } else if (c != '\r') {
      f.back() += c;
    }
  }
  return f;
}

// Device kernels behind each suite kernel of this variant (substrings of the
// demangled names rocprofv3 reports), for the counter join.
const std::map<std::string, std::vector<std::string>>& device_kernels() {
  static const std::map<std::string, std::vector<std::string>> m = {
      {"Stream_COPY", {"stream_vec_kernel<0,"}},
      {"Stream_MUL", {"stream_vec_kernel<1,"}},
      {"Stream_ADD", {"stream_vec_kernel<2,"}},
      {"Stream_TRIAD", {"stream_vec_kernel<3,"}},
      {"Stream_DOT", {"reduce_kernel<true,"}},
      {"Algorithm_REDUCE_SUM", {"reduce_kernel<false,"}},
      {"Algorithm_SCAN", {"scan_kernel<"}},
      {"Algorithm_SORT", {"hist_kernel(", "offsets_kernel(", "scatter_kernel("}},
      {"Polybench_GEMM", {"gemm_mfma_f64_kernel(", "gemm_naive_kernel("}},
      {"Polybench_JACOBI_2D", {"jacobi_fused_kernel<", "jacobi_vec_kernel<", "jacobi_scalar_kernel("}},
      {"Apps_MASS3DPA", {"mass3dpa_kernel<", "mass3dpa_naive_kernel<"}},
      {"Apps_DIFFUSION3DPA", {"diffusion3dpa_kernel<"}},
  };
  return m;
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
  if (run_params.getResume()) loadProgress();
}

// Restores the runs recorded in the progress file whose kernel, variant and
// tuning are selected now, with the same problem size and rep count.
void Executor::loadProgress() {
  std::ifstream f(run_params.getProgressFile());
  std::string line;
  while (std::getline(f, line)) {
    std::string kname, vname, tname;
    double hi, lo, tot, mn, mx, passes, reps, size, iw = 0.0;
    if (!json_string(line, "kernel", kname) || !json_string(line, "variant", vname) ||
        !json_string(line, "tuning", tname) || !json_number(line, "checksum_hi", hi) ||
        !json_number(line, "checksum_lo", lo) || !json_number(line, "time_total_s", tot) ||
        !json_number(line, "time_min_s", mn) || !json_number(line, "time_max_s", mx) ||
        !json_number(line, "passes", passes) || !json_number(line, "reps", reps) ||
        !json_number(line, "problem_size", size))
      continue;  // not a complete record (e.g. a line cut short by a crash)
    json_number(line, "iweighted", iw);
    for (KernelBase* k : kernels) {
      if (k->getName() != kname) continue;
      for (VariantID vid : variant_ids) {
        if (getVariantName(vid) != vname || !k->hasVariantDefined(vid)) continue;
        const size_t t = k->getVariantTuningIndex(vid, tname);
        if (t == KernelBase::getUnknownTuningIdx() || !tuningSelected(tname)) continue;
        if (double(k->getActualProblemSize()) != size || double(k->getRunReps()) != reps) continue;
        if (int(passes) < run_params.getNumPasses()) continue;  // an incomplete record is rerun
        k->restoreRun(vid, t, static_cast<long double>(hi) + static_cast<long double>(lo), iw, tot, mn, mx,
                      int(passes));
        restored.insert(run_key(kname, vname, tname));
      }
    }
  }
}

void Executor::appendProgress(const KernelBase* k, VariantID vid, size_t t) const {
  if (run_params.getProgressFile().empty()) return;
  std::ofstream f(run_params.getProgressFile(), std::ios::app);
  double hi, lo;
  split_ld(k->getChecksum(vid, t), hi, lo);
  f << std::setprecision(17) << "{\"kernel\": \"" << json_escape(k->getName()) << "\", \"variant\": \""
    << json_escape(getVariantName(vid)) << "\", \"tuning\": \"" << json_escape(k->getVariantTuningName(vid, t))
    << "\", \"checksum_hi\": " << hi << ", \"checksum_lo\": " << lo << ", \"iweighted\": " << k->getIWeighted(vid, t)
    << ", \"time_total_s\": " << k->getTotTime(vid, t)
    << ", \"time_min_s\": " << k->getMinTime(vid, t) << ", \"time_max_s\": " << k->getMaxTime(vid, t)
    << ", \"passes\": " << k->getNumExec(vid, t) << ", \"reps\": " << k->getRunReps()
    << ", \"problem_size\": " << k->getActualProblemSize() << "}" << std::endl;
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
      if (restored.count(run_key(kern->getName(), getVariantName(vid), tname))) {
        if (run_params.showProgress())
          getCout() << "   " << getVariantName(vid) << " variant, " << tname << " tuning -- restored\n";
        continue;
      }
      try {
        kern->execute(vid, t);
      } catch (const std::exception& e) {
        failures.push_back(kern->getName() + " " + getVariantName(vid) + "-" + tname + ": " + e.what());
        getCout() << "\n  " << failures.back() << "\n";
        status = 2;
        continue;
      }
      appendProgress(kern, vid, t);
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
  for (const std::string& c : run_params.getNpassesCombiners()) writeTimingCSV(pre + "-timing-" + c + ".csv", c);
  writeChecksumReport(pre + "-checksum.txt");
  writeKernelInfoCSV(pre + "-kernels.csv");
  writeRooflineCSV(pre + "-roofline.csv");
  if (!run_params.getJsonFile().empty()) writeJson(run_params.getJsonFile());
}

// One timing report per --npasses-combiners entry (Executor.cpp:852-854,
// 1515-1600): the pass times of each run combined as their mean
// (total / passes), minimum or maximum, in seconds per pass, then per rep
// with the rates that per-rep time gives.
void Executor::writeTimingCSV(const std::string& path, const std::string& combiner) const {
  std::ofstream f(path);
  const char* title = combiner == "Minimum" ? "Min" : combiner == "Maximum" ? "Max" : "Mean";
  f << title << " Runtime Report (sec.)\nKernel,Variant,Tuning,Time/pass,Time/rep,GB/s,GFLOP/s,Frac HBM (8 TB/s)\n";
  for (const KernelBase* k : kernels)
    for (VariantID vid : variant_ids)
      for (size_t t = 0; t < k->getNumVariantTunings(vid); ++t) {
        f << k->getName() << "," << getVariantName(vid) << "," << k->getVariantTuningName(vid, t) << ",";
        if (!k->wasVariantTuningRun(vid, t)) {
          f << "Not run,,,,\n";  // unexecuted tunings are marked, not reported as 0 s
          continue;
        }
        const double per_pass = combiner == "Minimum"   ? k->getMinTime(vid, t)
                                : combiner == "Maximum" ? k->getMaxTime(vid, t)
                                                        : k->getTotTime(vid, t) / k->getNumExec(vid, t);
        const double per_rep = per_pass / double(std::max<Index_type>(1, k->getRunReps()));
        const double gbs = k->getBytesPerRep() / per_rep / 1e9;
        f << std::setprecision(8) << per_pass << "," << per_rep << "," << gbs << ","
          << k->getFLOPsPerRep() / per_rep / 1e9 << "," << gbs / 8000.0 << "\n";
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
        f << std::setprecision(20) << cs << "  diff " << std::setprecision(6) << (cs - ref)
          << "  i-weighted FP64 sum " << std::setprecision(17) << k->getIWeighted(vid, t) << "\n";
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

// Per run: algorithmic bytes and FLOPs (the kernel's own counts), the
// bound they imply on MI355X (HBM at 8 TB/s below the FP64 ridge of 78.6
// TF/s over 8 TB/s, FP64 above it), the fraction of that bound, and the
// checksum's difference from the kernel's first reported run (the reference
// variant's when --refvar names one).  With --counters, each rocprofv3
// counter of the file is joined as its mean value per dispatch over the
// device kernels behind the suite kernel (all of the profiled run's
// dispatches of those kernels).
void Executor::writeRooflineCSV(const std::string& path) const {
  constexpr double kHbmGBs = 8000.0, kFp64GFs = 78600.0;
  std::map<std::string, std::map<std::string, std::pair<double, long>>> counters;  // kernel -> counter -> (sum, n)
  std::vector<std::string> counter_names;
  if (!run_params.getCountersFile().empty()) {
    std::ifstream cf(run_params.getCountersFile());
    std::string line;
    int ik = -1, ic = -1, iv = -1;
    if (std::getline(cf, line)) {
      const auto h = csv_fields(line);
      for (size_t i = 0; i < h.size(); ++i) {
        if (h[i] == "Kernel_Name") ik = int(i);
        if (h[i] == "Counter_Name") ic = int(i);
        if (h[i] == "Counter_Value") iv = int(i);
      }
    }
    std::set<std::string> names;
    while (ik >= 0 && ic >= 0 && iv >= 0 && std::getline(cf, line)) {
      const auto r = csv_fields(line);
      if (int(r.size()) <= std::max(ik, std::max(ic, iv))) continue;
      for (const auto& kv : device_kernels())
        for (const auto& sub : kv.second)
          if (r[ik].find(sub) != std::string::npos) {
            auto& acc = counters[kv.first][r[ic]];
            acc.first += std::strtod(r[iv].c_str(), nullptr);
            acc.second += 1;
            names.insert(r[ic]);
            break;
          }
    }
    counter_names.assign(names.begin(), names.end());
  }
  std::ofstream f(path);
  f << "Kernel,Variant,Tuning,Problem size,Reps,Bytes/rep,FLOPs/rep,Time/rep (s),GB/s,GFLOP/s,FLOP/byte,Bound,"
       "Frac of bound,Checksum,Checksum diff,Rel checksum diff";
  for (const auto& c : counter_names) f << "," << c << " per dispatch";
  f << "\n";
  for (const KernelBase* k : kernels) {
    bool have_ref = false;
    long double ref = 0.0L;
    const VariantID rv = run_params.getReferenceVariantID();
    if (rv != NumVariants && k->hasVariantDefined(rv))
      for (size_t t = 0; t < k->getNumVariantTunings(rv) && !have_ref; ++t)
        if (k->wasVariantTuningRun(rv, t)) ref = k->getChecksum(rv, t), have_ref = true;
    for (VariantID vid : variant_ids)
      for (size_t t = 0; t < k->getNumVariantTunings(vid); ++t) {
        if (!k->wasVariantTuningRun(vid, t)) continue;
        const long double cs = k->getChecksum(vid, t);
        if (!have_ref) ref = cs, have_ref = true;
        const double per_rep =
            k->getTotTime(vid, t) / k->getNumExec(vid, t) / double(std::max<Index_type>(1, k->getRunReps()));
        const double bytes = double(k->getBytesPerRep()), flops = double(k->getFLOPsPerRep());
        const double gbs = bytes / per_rep / 1e9, gfs = flops / per_rep / 1e9;
        const double ai = bytes > 0 ? flops / bytes : 0.0;
        const bool hbm = ai < kFp64GFs / kHbmGBs;
        const long double diff = cs - ref;
        f << k->getName() << "," << getVariantName(vid) << "," << k->getVariantTuningName(vid, t) << ","
          << k->getActualProblemSize() << "," << k->getRunReps() << "," << k->getBytesPerRep() << ","
          << k->getFLOPsPerRep() << "," << std::setprecision(8) << per_rep << "," << gbs << "," << gfs << "," << ai
          << "," << (hbm ? "hbm" : "fp64") << "," << (hbm ? gbs / kHbmGBs : gfs / kFp64GFs) << ","
          << std::setprecision(20) << cs << "," << diff << "," << std::setprecision(6)
          << (ref != 0.0L ? double(diff / ref) : double(diff));
        for (const auto& c : counter_names) {
          f << ",";
          const auto ki = counters.find(k->getName());
          if (ki == counters.end()) continue;
          const auto ci = ki->second.find(c);
          if (ci != ki->second.end() && ci->second.second > 0)
            f << std::setprecision(10) << ci->second.first / double(ci->second.second);
        }
        f << "\n";
      }
  }
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
          << "\", \"checksum_hi\": " << hi << ", \"checksum_lo\": " << lo << ", \"iweighted\": "
          << k->getIWeighted(vid, t) << ", \"time_total_s\": "
          << k->getTotTime(vid, t) << ", \"time_min_s\": " << k->getMinTime(vid, t) << ", \"time_max_s\": "
          << k->getMaxTime(vid, t) << ", \"passes\": "
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
