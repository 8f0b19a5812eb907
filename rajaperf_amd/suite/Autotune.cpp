#include "Autotune.hpp"

#include <fstream>
#include <sstream>
#include <stdexcept>

namespace rajaperf {

namespace {
void check(int st, const char* what) {
  if (st != RPX_SUCCESS) throw std::runtime_error(std::string("autotune ") + what + ": " + rpx_status_string(st));
}
constexpr int kTrialReps = 3;
}  // namespace

Autotuner& Autotuner::instance() {
  static Autotuner a;
  return a;
}

std::string Autotuner::arch() {
  if (arch_name.empty()) {
    char buf[64] = {0};
    check(rpx_init(0, buf, sizeof(buf)), "device");
    arch_name = buf;
  }
  return arch_name;
}

void Autotuner::reset() {
  cache_path.clear();
  known.clear();
  made.clear();
}

void Autotuner::setCacheFile(const std::string& path) {
  cache_path = path;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {  // kernel \t shape \t arch \t choice \t ms
    std::istringstream in(line);
    TuneDecision d;
    if (std::getline(in, d.kernel, '\t') && std::getline(in, d.shape, '\t') && std::getline(in, d.arch, '\t') &&
        std::getline(in, d.choice, '\t') && (in >> d.ms_per_rep)) {
      d.from_cache = true;
      known.push_back(d);
    }
  }
}

void Autotuner::save() const {
  if (cache_path.empty()) return;
  std::ofstream f(cache_path);
  for (const TuneDecision& d : known)
    f << d.kernel << '\t' << d.shape << '\t' << d.arch << '\t' << d.choice << '\t' << d.ms_per_rep << '\n';
}

const TuneCandidate& Autotuner::choose(const std::string& kernel, const std::string& shape,
                                       const std::vector<TuneCandidate>& cands,
                                       const std::function<void(const rpx_tuning&)>& trial) {
  if (cands.empty()) throw std::invalid_argument("autotune: no candidates for " + kernel);
  const std::string a = arch();
  for (const TuneDecision& d : known)
    if (d.kernel == kernel && d.shape == shape && d.arch == a)
      for (const TuneCandidate& c : cands)
        if (c.name == d.choice) {
          TuneDecision used = d;
          used.from_cache = true;
          made.push_back(used);
          return c;
        }
  void *e0 = nullptr, *e1 = nullptr;
  check(rpx_event_create(&e0), "event");
  check(rpx_event_create(&e1), "event");
  size_t best = 0;
  float best_ms = -1.0f;
  for (size_t i = 0; i < cands.size(); ++i) {
    trial(cands[i].tuning);  // warm-up
    check(rpx_device_synchronize(), "sync");
    check(rpx_event_record(e0, nullptr), "record");
    for (int r = 0; r < kTrialReps; ++r) trial(cands[i].tuning);
    check(rpx_event_record(e1, nullptr), "record");
    check(rpx_device_synchronize(), "sync");
    float ms = 0.0f;
    check(rpx_event_elapsed_ms(&ms, e0, e1), "elapsed");
    ms /= kTrialReps;
    if (best_ms < 0.0f || ms < best_ms) best_ms = ms, best = i;
  }
  rpx_event_destroy(e0);
  rpx_event_destroy(e1);
  const TuneDecision d{kernel, shape, a, cands[best].name, double(best_ms), false};
  known.push_back(d);
  made.push_back(d);
  save();
  return cands[best];
}

}  // namespace rajaperf
