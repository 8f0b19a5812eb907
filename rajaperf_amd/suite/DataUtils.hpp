// Data utilities of the harness (src/common/DataUtils.hpp/.cpp).
//
// The MI355X variant's data space is HBM (hipMalloc through the C ABI).
// Reference-mode initialisers follow the suite's formulas and call order
// (DataUtils.cpp:131-147, 477-595): the formula is evaluated on the device
// (its multiply and divide are correctly rounded, so the bits are the
// suite's), the rand()-based ones on the host and copied in.  Seeded mode
// draws every array initialiser from SplitMix64 on the device, keyed by the
// initialiser's call ordinal.  Checksums copy the data to the host and take
// the suite's Kahan long-double weighted sum (DataUtils.cpp:601-621).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "RPTypes.hpp"

namespace rajaperf {
namespace detail {

void resetDataInitCount();
void incDataInitCount();
int getDataInitCount();

// Device memory.
void* deviceAlloc(Size_type bytes);
void deviceFree(void* p);
void copyToDevice(void* dst, const void* src_host, Size_type bytes);
void copyToHost(void* dst_host, const void* src, Size_type bytes);
void deviceSynchronize();

// HIP graph capture of rep loops (KernelBase::timedReps): one created stream
// per process, capture, replay and teardown; errors throw.
void* captureStream();
void beginCapture(void* stream);
void* endCapture(void* stream);
void abandonCapture(void* stream);
void launchGraph(void* graph, void* stream);
void destroyGraph(void* graph);
// roctx ranges around timed rep loops (no-ops when no profiler listens).
void rangePush(const std::string& name);
void rangePop();

// Initialisers writing device arrays; each advances the init counter once.
void initDataDevice(Real_ptr d, Size_type len);                   // DataUtils.cpp:504-513
void initDataConstDevice(Real_ptr d, Size_type len, Real_type v);  // :518-528
void initDataRandValueDevice(Real_ptr d, Size_type len);          // :560-569 (host rand())
void initDataSeededDevice(Real_ptr d, Size_type len, std::uint64_t seed, Real_type lo, Real_type hi);
void initDataSeededGridDevice(Real_ptr d, Index_type N, std::uint64_t seed);
void initDataSeededSymTensorDevice(Real_ptr d, Index_type ne, int q3, int sym, std::uint64_t seed);
void initDataHostToDevice(Real_ptr d, const std::vector<Real_type>& host);  // counts once
void initData(Real_type& d);                                        // :589-595 (scalar)

// Lagrange basis on Gauss-Lobatto nodes at Gauss-Legendre points, exactly
// symmetric under (q,d) -> (Q-1-q, D-1-d); B[q + Q*d], G likewise.
void femBasis(int d1d, int q1d, std::vector<Real_type>& B, std::vector<Real_type>& G);

// Checksums of device arrays (DataUtils.hpp:386-409, DataUtils.cpp:601-621).
// The suite checksum of device data; with `iweighted`, also adds the
// data's i-weighted sum to *iweighted.
long double calcChecksumDevice(const Real_type* d, Size_type len, Real_type scale_factor,
                               double* iweighted = nullptr);
// sum_i (i+1)*h[i] in FP64, index order, no fused multiply-add: the second
// per-kernel checksum (the "sum of i-weighted elements in FP64" of the
// stream workload); weights start at 1 so a scalar result is itself.
double iWeightedSumHost(const Real_type* h, Size_type len);
long double calcChecksumHost(const Real_type* h, Size_type len, Real_type scale_factor);

}  // namespace detail
}  // namespace rajaperf
