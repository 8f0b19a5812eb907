#include "RunParams.hpp"

#include <cmath>
#include <cstdlib>
#include <iostream>
#include <sstream>

namespace rajaperf {

namespace {
bool is_option(const char* s) { return s && s[0] == '-'; }

// Collects the non-option arguments that follow position i.
std::vector<std::string> take_list(int argc, const char** argv, int& i) {
  std::vector<std::string> out;
  while (i + 1 < argc && !is_option(argv[i + 1])) out.emplace_back(argv[++i]);
  return out;
}

bool parse_int(const char* s, std::int64_t& v) {
  if (!s) return false;
  char* end = nullptr;
  const long long x = std::strtoll(s, &end, 10);
  if (end == s || *end != '\0') return false;
  v = x;
  return true;
}

bool parse_double(const char* s, double& v) {
  if (!s) return false;
  char* end = nullptr;
  v = std::strtod(s, &end);
  return end != s && *end == '\0';
}
}  // namespace

RunParams::RunParams(int argc, const char** argv) {
  parse(argc, argv);
  if (input_state == InfoRequest || input_state == BadInput) return;
  if (size_meaning == Unset) size_meaning = Factor;  // RunParams.cpp:1199-1203
  processKernelInput();
  processVariantInput();
  if (!invalid_input.empty()) input_state = BadInput;
}

void RunParams::parse(int argc, const char** argv) {
  for (int i = 1; i < argc; ++i) {
    const std::string opt(argv[i]);
    std::int64_t iv = 0;
    double dv = 0.0;
    auto need = [&](bool ok) {
      if (!ok) {
        invalid_input.push_back(opt);
        input_state = BadInput;
      }
      return ok;
    };
    if (opt == "-h" || opt == "--help") {
      input_state = InfoRequest;
      printHelpMessage(getCout());
    } else if (opt == "--print-kernels" || opt == "-pk") {
      print_kernels = true;
      input_state = InfoRequest;
    } else if (opt == "--print-variants" || opt == "-pv") {
      print_variants = true;
      input_state = InfoRequest;
    } else if (opt == "--show-progress" || opt == "-sp") {
      show_progress = true;
    } else if (opt == "--npasses") {
      if (need(i + 1 < argc && parse_int(argv[i + 1], iv) && iv > 0)) npasses = int(iv), ++i;
    } else if (opt == "--repfact") {
      if (need(i + 1 < argc && parse_double(argv[i + 1], dv) && dv >= 0.0)) rep_fact = dv, ++i;
    } else if (opt == "--sizefact") {
      if (need(size_meaning != Direct && i + 1 < argc && parse_double(argv[i + 1], dv) && dv >= 0.0))
        size_factor = dv, size_meaning = Factor, ++i;
    } else if (opt == "--size") {
      if (need(size_meaning != Factor && i + 1 < argc && parse_int(argv[i + 1], iv) && iv >= 0))
        size = iv, size_meaning = Direct, ++i;
    } else if (opt == "--checkrun") {
      input_state = CheckRun;
      if (i + 1 < argc && !is_option(argv[i + 1])) {
        if (need(parse_int(argv[i + 1], iv) && iv > 0)) checkrun_reps = int(iv), ++i;
      }
    } else if (opt == "--dryrun") {
      if (input_state != BadInput) input_state = DryRun;
    } else if (opt == "--disable-warmup") {
      disable_warmup = true;
    } else if (opt == "--gpu_block_size") {
      if (need(i + 1 < argc && parse_int(argv[i + 1], iv) && iv > 0)) gpu_block_size = int(iv), ++i;
    } else if (opt == "--items_per_thread") {
      if (need(i + 1 < argc && parse_int(argv[i + 1], iv) && iv > 0)) items_per_thread = int(iv), ++i;
    } else if (opt == "-od" || opt == "--outdir") {
      if (need(i + 1 < argc && !is_option(argv[i + 1]))) outdir = argv[++i];
    } else if (opt == "-of" || opt == "--outfile") {
      if (need(i + 1 < argc && !is_option(argv[i + 1]))) outfile_prefix = argv[++i];
    } else if (opt == "--json") {
      if (need(i + 1 < argc && !is_option(argv[i + 1]))) json_file = argv[++i];
    } else if (opt == "--tuning-cache") {
      if (need(i + 1 < argc && !is_option(argv[i + 1]))) tuning_cache = argv[++i];
    } else if (opt == "--progress") {
      if (need(i + 1 < argc && !is_option(argv[i + 1]))) progress_file = argv[++i];
    } else if (opt == "--resume") {
      resume = true;
    } else if (opt == "--counters") {
      if (need(i + 1 < argc && !is_option(argv[i + 1]))) counters_file = argv[++i];
    } else if (opt == "--refvar" || opt == "-rv") {
      if (need(i + 1 < argc && !is_option(argv[i + 1]))) reference_variant = argv[++i];
    } else if (opt == "-k" || opt == "--kernels") {
      auto l = take_list(argc, argv, i);
      need(!l.empty());
      kernel_input.insert(kernel_input.end(), l.begin(), l.end());
    } else if (opt == "-ek" || opt == "--exclude-kernels") {
      auto l = take_list(argc, argv, i);
      need(!l.empty());
      exclude_kernel_input.insert(exclude_kernel_input.end(), l.begin(), l.end());
    } else if (opt == "-v" || opt == "--variants") {
      auto l = take_list(argc, argv, i);
      need(!l.empty());
      variant_input.insert(variant_input.end(), l.begin(), l.end());
    } else if (opt == "-ev" || opt == "--exclude-variants") {
      auto l = take_list(argc, argv, i);
      need(!l.empty());
      exclude_variant_input.insert(exclude_variant_input.end(), l.begin(), l.end());
    } else if (opt == "-t" || opt == "--tunings") {
      auto l = take_list(argc, argv, i);
      need(!l.empty());
      tuning_input.insert(tuning_input.end(), l.begin(), l.end());
    } else if (opt == "-et" || opt == "--exclude-tunings") {
      auto l = take_list(argc, argv, i);
      need(!l.empty());
      exclude_tuning_input.insert(exclude_tuning_input.end(), l.begin(), l.end());
    } else if (opt == "--npasses-combiners") {  // RunParams.cpp:386-399
      auto l = take_list(argc, argv, i);
      need(!l.empty());
      npasses_combiners.insert(npasses_combiners.end(), l.begin(), l.end());
    } else if (opt == "--graph") {
      graph_mode = true;
    } else if (opt == "--input-mode") {
      if (need(i + 1 < argc)) {
        const std::string m(argv[++i]);
        if (m == "reference") input_mode = ReferenceInputs;
        else if (m == "seeded") input_mode = SeededInputs;
        else need(false);
      }
    } else if (opt == "--seed") {
      if (need(i + 1 < argc && parse_int(argv[i + 1], iv) && iv >= 0)) seed = std::uint64_t(iv), ++i;
    } else if (opt == "--tsteps") {
      if (need(i + 1 < argc && parse_int(argv[i + 1], iv) && iv >= 0)) tsteps = iv, ++i;
    } else if (opt == "--nk") {
      if (need(i + 1 < argc && parse_int(argv[i + 1], iv) && iv >= 0)) nk = iv, ++i;
    } else if (opt == "--alpha") {
      if (need(i + 1 < argc && parse_double(argv[i + 1], dv) && std::isfinite(dv)))
        gemm_alpha = dv, has_alpha = true, ++i;
    } else if (opt == "--beta") {
      if (need(i + 1 < argc && parse_double(argv[i + 1], dv) && std::isfinite(dv)))
        gemm_beta = dv, has_beta = true, ++i;
    } else if (opt == "--d1d") {
      if (need(i + 1 < argc && parse_int(argv[i + 1], iv) && iv >= 1 && iv <= 8)) d1d = int(iv), ++i;
    } else if (opt == "--q1d") {
      if (need(i + 1 < argc && parse_int(argv[i + 1], iv) && iv >= 1 && iv <= 9)) q1d = int(iv), ++i;
    } else {
      invalid_input.push_back(opt);
      input_state = BadInput;
    }
  }
  // Combiners of the per-pass times (RunParams.cpp:1858-1878): Average by default.
  if (npasses_combiners.empty()) npasses_combiners.push_back("Average");
  for (const auto& c : npasses_combiners)
    if (c != "Average" && c != "Minimum" && c != "Maximum") {
      invalid_input.push_back(c);
      input_state = BadInput;
    }
  if (input_state == Undefined) input_state = PerfRun;
  if (input_state == InfoRequest) {
    if (print_kernels)
      for (int k = 0; k < NumKernels; ++k) getCout() << getFullKernelName(KernelID(k)) << "\n";
    if (print_variants)
      for (int v = 0; v < NumVariants; ++v)
        if (isVariantAvailable(VariantID(v))) getCout() << getVariantName(VariantID(v)) << "\n";
  }
}

// Kernel selection by full name, base name or group name
// (RunParams.cpp:1893-2120).
void RunParams::processKernelInput() {
  auto matches = [](const std::string& in, KernelID k) {
    return in == getFullKernelName(k) || in == getKernelName(k) || in == getGroupName(getGroupOf(k));
  };
  for (int k = 0; k < NumKernels; ++k) {
    const KernelID kid = KernelID(k);
    bool take = kernel_input.empty();
    for (const auto& in : kernel_input) take = take || matches(in, kid);
    for (const auto& in : exclude_kernel_input)
      if (matches(in, kid)) take = false;
    if (take) run_kernels.insert(kid);
  }
  for (const auto& in : kernel_input) {
    bool known = false;
    for (int k = 0; k < NumKernels; ++k) known = known || matches(in, KernelID(k));
    if (!known) invalid_input.push_back(in);
  }
  for (const auto& in : exclude_kernel_input) {
    bool known = false;
    for (int k = 0; k < NumKernels; ++k) known = known || matches(in, KernelID(k));
    if (!known) invalid_input.push_back(in);
  }
}

// Variant selection among available variants (RunParams.cpp:2302-2425).
void RunParams::processVariantInput() {
  for (const auto& in : variant_input) {
    bool found = false;
    for (int v = 0; v < NumVariants; ++v)
      if (in == getVariantName(VariantID(v))) {
        found = true;
        if (isVariantAvailable(VariantID(v))) run_variants.insert(VariantID(v));
        else invalid_input.push_back(in + " (variant not built into this harness)");
      }
    if (!found) invalid_input.push_back(in);
  }
  if (variant_input.empty())
    for (int v = 0; v < NumVariants; ++v)
      if (isVariantAvailable(VariantID(v))) run_variants.insert(VariantID(v));
  for (const auto& in : exclude_variant_input) {
    bool found = false;
    for (int v = 0; v < NumVariants; ++v)
      if (in == getVariantName(VariantID(v))) {
        found = true;
        run_variants.erase(VariantID(v));
      }
    if (!found) invalid_input.push_back(in);
  }
  if (!reference_variant.empty()) {
    for (int v = 0; v < NumVariants; ++v)
      if (reference_variant == getVariantName(VariantID(v))) reference_vid = VariantID(v);
    if (reference_vid == NumVariants) invalid_input.push_back(reference_variant);
  } else if (!run_variants.empty()) {
    reference_vid = *run_variants.begin();
  }
}

void RunParams::printHelpMessage(std::ostream& str) const {
  str << "Usage: rajaperf-mi355x [options]\n"
      << "  -k/--kernels <names...>     kernels, groups or full names to run\n"
      << "  -ek/--exclude-kernels <..>  kernels to skip\n"
      << "  -v/--variants <names...>    variants to run (this build: MI355X_Native)\n"
      << "  -ev/--exclude-variants <..>\n"
      << "  -t/--tunings <names...>     tunings to run, -et/--exclude-tunings\n"
      << "  --size <n> | --sizefact <f> target problem size, direct or as a factor\n"
      << "  --repfact <f>               scale the default rep counts\n"
      << "  --npasses <n>               passes over the suite\n"
      << "  --npasses-combiners <Average|Minimum|Maximum...>  one timing CSV per combiner\n"
      << "  --graph                     also run every tuning as a HIP graph of its rep loop (<tuning>-graph)\n"
      << "  --checkrun [reps]           run each kernel for <reps> reps (default 1)\n"
      << "  --gpu_block_size <n>        workgroup size for the default tuning\n"
      << "  --items_per_thread <n>      items per work-item for the default tuning\n"
      << "  --input-mode reference|seeded, --seed <s>\n"
      << "  --tsteps <n> (JACOBI_2D), --nk <n>, --alpha <x>, --beta <x> (GEMM), --d1d/--q1d <n> (FEM)\n"
      << "  -od/--outdir <dir>, -of/--outfile <prefix>, --json <file>\n"
      << "  --tuning-cache <file>  (decisions of the \"auto\" tunings)\n"
      << "  --progress <file.jsonl> [--resume]  one line per finished run; resume skips recorded runs\n"
      << "  --counters <rocprofv3 counter_collection.csv>  joined into <prefix>-roofline.csv\n"
      << "  --refvar <variant>, --dryrun, --disable-warmup, --show-progress\n"
      << "  --print-kernels, --print-variants, -h/--help\n";
}

void RunParams::print(std::ostream& str) const {
  str << "\n npasses = " << npasses << "\n repfact = " << rep_fact
      << "\n size_meaning = " << (size_meaning == Direct ? "direct" : "factor")
      << "\n size = " << size << "\n sizefact = " << size_factor
      << "\n input_mode = " << (input_mode == SeededInputs ? "seeded" : "reference") << "\n seed = " << seed
      << "\n checkrun_reps = " << checkrun_reps << "\n";
}

}  // namespace rajaperf
