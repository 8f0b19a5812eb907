#include "RAJAPerfSuite.hpp"

#include <iostream>

#include "algorithm/Algorithm.hpp"
#include "apps/Apps.hpp"
#include "polybench/Polybench.hpp"
#include "stream/Stream.hpp"

namespace rajaperf {

namespace {
const std::string kGroupNames[] = {"Polybench", "Stream", "Apps", "Algorithm", "Unknown Group"};
// Kernel base names as in RAJAPerfSuite.cpp:168-279.
const std::string kKernelNames[] = {"GEMM",          "JACOBI_2D", "ADD",  "COPY", "DOT",  "MUL",
                                    "TRIAD",         "DIFFUSION3DPA",    "MASS3DPA", "SCAN", "SORT",
                                    "REDUCE_SUM",    "Unknown Kernel"};
const GroupID kKernelGroups[] = {Polybench, Polybench, Stream, Stream, Stream, Stream,
                                 Stream,    Apps,      Apps,   Algorithm, Algorithm, Algorithm};
const std::string kVariantNames[] = {"Base_Seq",     "Lambda_Seq",  "RAJA_Seq",          "Base_OpenMP",
                                     "Lambda_OpenMP", "RAJA_OpenMP", "Base_OpenMPTarget", "RAJA_OpenMPTarget",
                                     "Base_CUDA",    "Lambda_CUDA", "RAJA_CUDA",         "Base_HIP",
                                     "Lambda_HIP",   "RAJA_HIP",    "Kokkos_Lambda",     "Base_SYCL",
                                     "RAJA_SYCL",    "MI355X_Native", "Unknown Variant"};
const std::string kComplexityNames[] = {"N", "N_logN", "N^(3/2)", "N^(2/3)", "Unknown Complexity"};
}  // namespace

const std::string& getGroupName(GroupID gid) { return kGroupNames[gid < NumGroups ? gid : NumGroups]; }
const std::string& getKernelName(KernelID kid) { return kKernelNames[kid < NumKernels ? kid : NumKernels]; }
GroupID getGroupOf(KernelID kid) { return kKernelGroups[kid]; }
std::string getFullKernelName(KernelID kid) { return getGroupName(getGroupOf(kid)) + "_" + getKernelName(kid); }
const std::string& getVariantName(VariantID vid) { return kVariantNames[vid < NumVariants ? vid : NumVariants]; }
const std::string& getComplexityName(Complexity ac) {
  return kComplexityNames[ac < NumComplexities ? ac : NumComplexities];
}

bool isVariantAvailable(VariantID vid) { return vid == MI355X_Native; }
bool isVariantGPU(VariantID vid) {
  return vid == Base_CUDA || vid == Lambda_CUDA || vid == RAJA_CUDA || vid == Base_HIP || vid == Lambda_HIP ||
         vid == RAJA_HIP || vid == Base_SYCL || vid == RAJA_SYCL || vid == MI355X_Native;
}

KernelBase* getKernelObject(KernelID kid, const RunParams& p) {
  switch (kid) {
    case Polybench_GEMM: return new polybench::POLYBENCH_GEMM(p);
    case Polybench_JACOBI_2D: return new polybench::POLYBENCH_JACOBI_2D(p);
    case Stream_ADD: return new stream::ADD(p);
    case Stream_COPY: return new stream::COPY(p);
    case Stream_DOT: return new stream::DOT(p);
    case Stream_MUL: return new stream::MUL(p);
    case Stream_TRIAD: return new stream::TRIAD(p);
    case Apps_DIFFUSION3DPA: return new apps::DIFFUSION3DPA(p);
    case Apps_MASS3DPA: return new apps::MASS3DPA(p);
    case Algorithm_SCAN: return new algorithm::SCAN(p);
    case Algorithm_SORT: return new algorithm::SORT(p);
    case Algorithm_REDUCE_SUM: return new algorithm::REDUCE_SUM(p);
    default: return nullptr;
  }
}

std::ostream& getCout() { return std::cout; }

}  // namespace rajaperf
