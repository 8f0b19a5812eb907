// Kernel / variant registry of the harness (src/common/RAJAPerfSuite.hpp).
//
// The VariantID enum keeps the suite's full list and order and appends the
// new MI355X_Native variant before NumVariants, as a maintainer would add it
// upstream (RAJAPerfSuite.hpp:200-228, RAJAPerfSuite.cpp:294-323).  This
// harness builds only the MI355X_Native variant; the others are provided by
// the suite itself.  The KernelID list holds the kernels this variant
// implements, under the suite's names.
#pragma once

#include <iosfwd>
#include <string>

namespace rajaperf {

class KernelBase;
class RunParams;

enum GroupID { Polybench = 0, Stream, Apps, Algorithm, NumGroups };

enum KernelID {
  Polybench_GEMM = 0,
  Polybench_JACOBI_2D,
  Stream_ADD,
  Stream_COPY,
  Stream_DOT,
  Stream_MUL,
  Stream_TRIAD,
  Apps_DIFFUSION3DPA,
  Apps_MASS3DPA,
  Algorithm_SCAN,
  Algorithm_SORT,
  Algorithm_REDUCE_SUM,
  NumKernels
};

enum VariantID {
  Base_Seq = 0,
  Lambda_Seq,
  RAJA_Seq,
  Base_OpenMP,
  Lambda_OpenMP,
  RAJA_OpenMP,
  Base_OpenMPTarget,
  RAJA_OpenMPTarget,
  Base_CUDA,
  Lambda_CUDA,
  RAJA_CUDA,
  Base_HIP,
  Lambda_HIP,
  RAJA_HIP,
  Kokkos_Lambda,
  Base_SYCL,
  RAJA_SYCL,
  MI355X_Native,  // the variant this build adds
  NumVariants
};

enum Complexity { N = 0, N_logN, N_to_the_three_halves, N_to_the_two_thirds, NumComplexities };

const std::string& getGroupName(GroupID gid);
const std::string& getKernelName(KernelID kid);
std::string getFullKernelName(KernelID kid);
const std::string& getVariantName(VariantID vid);
const std::string& getComplexityName(Complexity ac);
// True for the variants this harness can run (RAJAPerfSuite.cpp:503-562).
bool isVariantAvailable(VariantID vid);
bool isVariantGPU(VariantID vid);
GroupID getGroupOf(KernelID kid);

// Factory (RAJAPerfSuite.cpp:788-1135).  Caller owns the object.
KernelBase* getKernelObject(KernelID kid, const RunParams& run_params);

// Output stream of the suite (rank 0 stdout; RAJAPerfSuite.cpp:1163-1173).
std::ostream& getCout();

}  // namespace rajaperf
