#include "DataUtils.hpp"

#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "rajaperf_mi355x.h"

namespace rajaperf {
namespace detail {

namespace {
int data_init_count = 0;  // DataUtils.cpp:131

void check(int status, const char* what) {
  if (status != RPX_SUCCESS)
    throw std::runtime_error(std::string(what) + ": " + rpx_status_string(status));
}
}  // namespace

void resetDataInitCount() { data_init_count = 0; }
void incDataInitCount() { data_init_count++; }
int getDataInitCount() { return data_init_count; }

void* deviceAlloc(Size_type bytes) {
  void* p = nullptr;
  check(rpx_device_malloc(&p, bytes), "rpx_device_malloc");
  return p;
}
void deviceFree(void* p) { check(rpx_device_free(p), "rpx_device_free"); }
void copyToDevice(void* dst, const void* src, Size_type bytes) { check(rpx_memcpy_h2d(dst, src, bytes), "h2d"); }
void copyToHost(void* dst, const void* src, Size_type bytes) { check(rpx_memcpy_d2h(dst, src, bytes), "d2h"); }
void deviceSynchronize() { check(rpx_device_synchronize(), "rpx_device_synchronize"); }

void* captureStream() {
  static void* s = nullptr;  // created once, kept until the process ends
  if (!s) check(rpx_stream_create(&s), "rpx_stream_create");
  return s;
}
void beginCapture(void* stream) { check(rpx_graph_begin_capture(stream), "rpx_graph_begin_capture"); }
void* endCapture(void* stream) {
  void* g = nullptr;
  check(rpx_graph_end_capture(stream, &g), "rpx_graph_end_capture");
  return g;
}
void abandonCapture(void* stream) {
  void* g = nullptr;
  if (rpx_graph_end_capture(stream, &g) == RPX_SUCCESS) rpx_graph_destroy(g);
}
void launchGraph(void* graph, void* stream) { check(rpx_graph_launch(graph, stream), "rpx_graph_launch"); }
void destroyGraph(void* graph) { check(rpx_graph_destroy(graph), "rpx_graph_destroy"); }
void rangePush(const std::string& name) { roctxRangePushA(name.c_str()); }
void rangePop() { roctxRangePop(); }

void initDataDevice(Real_ptr d, Size_type len) {
  const Real_type factor = (data_init_count % 2 ? 0.1 : 0.2);
  check(rpx_fill_suite_formula_f64(d, Index_type(len), factor, nullptr), "rpx_fill_suite_formula_f64");
  incDataInitCount();
}

void initDataConstDevice(Real_ptr d, Size_type len, Real_type v) {
  check(rpx_fill_const_f64(d, Index_type(len), v, nullptr), "rpx_fill_const_f64");
  incDataInitCount();
}

void initDataRandValueDevice(Real_ptr d, Size_type len) {
  std::vector<Real_type> h(len);
  srand(4793);
  for (Size_type i = 0; i < len; ++i) h[i] = Real_type(rand()) / RAND_MAX;
  copyToDevice(d, h.data(), len * sizeof(Real_type));
  incDataInitCount();
}

void initDataSeededDevice(Real_ptr d, Size_type len, std::uint64_t seed, Real_type lo, Real_type hi) {
  check(rpx_fill_uniform_f64(d, Index_type(len), seed, std::uint64_t(data_init_count), lo, hi, nullptr),
        "rpx_fill_uniform_f64");
  incDataInitCount();
}

void initDataSeededGridDevice(Real_ptr d, Index_type N, std::uint64_t seed) {
  check(rpx_fill_grid_f64(d, N, seed, std::uint64_t(data_init_count), 0.0, 1.0, 0.0, nullptr), "rpx_fill_grid_f64");
  incDataInitCount();
}

void initDataSeededSymTensorDevice(Real_ptr d, Index_type ne, int q3, int sym, std::uint64_t seed) {
  check(rpx_fill_sym_tensor_f64(d, ne, q3, sym, seed, std::uint64_t(data_init_count), nullptr),
        "rpx_fill_sym_tensor_f64");
  incDataInitCount();
}

void initDataHostToDevice(Real_ptr d, const std::vector<Real_type>& host) {
  copyToDevice(d, host.data(), host.size() * sizeof(Real_type));
  incDataInitCount();
}

void initData(Real_type& d) {  // DataUtils.cpp:589-595
  const Real_type factor = (data_init_count % 2 ? 0.1 : 0.2);
  d = factor * 1.1 / 1.12345;
  incDataInitCount();
}

namespace {
void legendre(int n, double x, double* p, double* dp) {
  double p0 = 1.0, p1 = x;
  if (n == 0) {
    *p = 1.0;
    *dp = 0.0;
    return;
  }
  for (int k = 2; k <= n; ++k) {
    const double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = pk;
  }
  *p = p1;
  *dp = n * (x * p1 - p0) / (x * x - 1.0);
}

// Points on [0,1], mirrored so that x[n-1-i] = 1 - x[i] holds exactly.
std::vector<double> gauss_legendre(int n) {
  std::vector<double> x(n);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      legendre(n, z, &p, &dp);
      const double dz = p / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[n - 1 - i] = 0.5 * (z + 1.0);
    x[i] = 1.0 - x[n - 1 - i];
  }
  if (n % 2 == 1) x[n / 2] = 0.5;
  return x;
}

std::vector<double> gauss_lobatto(int n) {
  std::vector<double> x(n);
  if (n == 1) {
    x[0] = 0.5;
    return x;
  }
  x[0] = 0.0;
  x[n - 1] = 1.0;
  const int m = n - 1;
  for (int i = 1; i <= (n - 1) / 2; ++i) {
    double z = std::cos(M_PI * i / m);
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      legendre(m, z, &p, &dp);
      const double d2p = (2.0 * z * dp - m * (m + 1.0) * p) / (1.0 - z * z);
      const double dz = dp / d2p;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[n - 1 - i] = 0.5 * (z + 1.0);
    x[i] = 1.0 - x[n - 1 - i];
  }
  if (n % 2 == 1) x[n / 2] = 0.5;
  return x;
}
}  // namespace

void femBasis(int d1d, int q1d, std::vector<Real_type>& B, std::vector<Real_type>& G) {
  B.assign(size_t(q1d) * d1d, 0.0);
  G.assign(size_t(q1d) * d1d, 0.0);
  const std::vector<double> xd = gauss_lobatto(d1d), xq = gauss_legendre(q1d);
  for (int q = 0; q < q1d; ++q) {
    const bool mirror = (2 * q > q1d - 1);
    for (int d = 0; d < d1d; ++d) {
      if (mirror) {
        B[q + q1d * d] = B[(q1d - 1 - q) + q1d * (d1d - 1 - d)];
        G[q + q1d * d] = -G[(q1d - 1 - q) + q1d * (d1d - 1 - d)];
        continue;
      }
      const double x = xq[q];
      double l = 1.0;
      for (int m = 0; m < d1d; ++m)
        if (m != d) l *= (x - xd[m]) / (xd[d] - xd[m]);
      double dl = 0.0;
      for (int k = 0; k < d1d; ++k) {
        if (k == d) continue;
        double t = 1.0 / (xd[d] - xd[k]);
        for (int m = 0; m < d1d; ++m)
          if (m != d && m != k) t *= (x - xd[m]) / (xd[d] - xd[m]);
        dl += t;
      }
      B[q + q1d * d] = l;
      G[q + q1d * d] = dl;
    }
  }
  if (q1d % 2 == 1) {
    const int q = q1d / 2;
    for (int d = 0; d < d1d / 2; ++d) {
      B[q + q1d * (d1d - 1 - d)] = B[q + q1d * d];
      G[q + q1d * (d1d - 1 - d)] = -G[q + q1d * d];
    }
    if (d1d % 2 == 1) G[q + q1d * (d1d / 2)] = 0.0;
  }
}

long double calcChecksumHost(const Real_type* h, Size_type len, Real_type scale_factor) {
  long double tchk = 0.0;
  long double ckahan = 0.0;
  for (Size_type j = 0; j < len; ++j) {
    long double x = (std::abs(std::sin(j + 1.0)) + 0.5) * static_cast<long double>(h[j]);
    long double y = x - ckahan;
    volatile long double t = tchk + y;
    volatile long double z = t - tchk;
    ckahan = z - y;
    tchk = t;
  }
  tchk *= scale_factor;
  return tchk;
}

long double calcChecksumDevice(const Real_type* d, Size_type len, Real_type scale_factor, double* iweighted) {
  std::vector<Real_type> h(len);
  if (len) copyToHost(h.data(), d, len * sizeof(Real_type));
  if (iweighted) *iweighted += iWeightedSumHost(h.data(), len);
  return calcChecksumHost(h.data(), len, scale_factor);
}

double iWeightedSumHost(const Real_type* h, Size_type len) {
  double s = 0.0;
  for (Size_type i = 0; i < len; ++i) s += static_cast<double>(i + 1) * h[i];
  return s;
}

}  // namespace detail
}  // namespace rajaperf
