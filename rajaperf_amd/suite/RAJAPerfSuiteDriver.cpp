// Suite driver (src/RAJAPerfSuiteDriver.cpp:22-61): Executor steps 1-5.
#include <iostream>

#include "Executor.hpp"
#include "rajaperf_mi355x.h"

int main(int argc, char** argv) {
  std::cout << "\n\nRunning RAJA performance suite (MI355X_Native)\n";
  char arch[64] = {0};
  const int st = rpx_init(0, arch, sizeof(arch));
  rajaperf::Executor executor(argc, const_cast<const char**>(argv));
  if (st != RPX_SUCCESS && executor.getRunParams().getInputState() != rajaperf::RunParams::InfoRequest &&
      executor.getRunParams().getInputState() != rajaperf::RunParams::BadInput &&
      executor.getRunParams().getInputState() != rajaperf::RunParams::DryRun) {
    std::cerr << "rajaperf-mi355x: no gfx950 device (" << rpx_status_string(st) << ")\n";
    return 3;
  }
  executor.setupSuite();
  executor.reportRunSummary(std::cout);
  executor.runSuite();
  executor.outputRunData();
  std::cout << "\n\nDONE!!!....\n";
  return executor.getStatus();
}
