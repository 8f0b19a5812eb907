// Oracle restatement of the suite's data utilities (TEST INFRASTRUCTURE).
//
// src/common/DataUtils.cpp:131-147 (init counter), :477-595 (initialisers),
// :601-653 (checksum).  Plus the seeded generator and FEM bases used by the
// seeded input mode, which have no counterpart in the suite.
#include <cmath>
#include <cstdlib>
#include <vector>

#include "oracle.h"

static int g_init_count = 0;  // DataUtils.cpp:131

extern "C" {

void orc_reset_init_count(void) { g_init_count = 0; }  // DataUtils.cpp:136-139
int orc_get_init_count(void) { return g_init_count; }

// DataUtils.cpp:477-498: +-1 by rand() sign, then two outliers.
void orc_init_data_int(int* p, int64_t n) {
  srand(4793);
  double signfact = 0.0;
  for (size_t i = 0; i < size_t(n); ++i) {
    signfact = double(rand()) / RAND_MAX;
    p[i] = (signfact < 0.5 ? -1 : 1);
  }
  signfact = double(rand()) / RAND_MAX;
  size_t ilo = size_t(n) * signfact;
  p[ilo] = -58;
  signfact = double(rand()) / RAND_MAX;
  size_t ihi = size_t(n) * signfact;
  p[ihi] = 19;
  g_init_count++;
}

// DataUtils.cpp:504-513: factor*(i + 1.1)/(i + 1.12345), factor by parity.
void orc_init_data(double* p, int64_t n) {
  const double factor = (g_init_count % 2 ? 0.1 : 0.2);
  for (size_t i = 0; i < size_t(n); ++i) p[i] = factor * (i + 1.1) / (i + 1.12345);
  g_init_count++;
}

// DataUtils.cpp:518-528.
void orc_init_data_const(double* p, int64_t n, double v) {
  for (size_t i = 0; i < size_t(n); ++i) p[i] = v;
  g_init_count++;
}

// DataUtils.cpp:542-556.
void orc_init_data_rand_sign(double* p, int64_t n) {
  const double factor = (g_init_count % 2 ? 0.1 : 0.2);
  srand(4793);
  for (size_t i = 0; i < size_t(n); ++i) {
    double signfact = double(rand()) / RAND_MAX;
    signfact = (signfact < 0.5 ? -1.0 : 1.0);
    p[i] = signfact * factor * (i + 1.1) / (i + 1.12345);
  }
  g_init_count++;
}

// DataUtils.cpp:560-569.
void orc_init_data_rand_value(double* p, int64_t n) {
  srand(4793);
  for (size_t i = 0; i < size_t(n); ++i) p[i] = double(rand()) / RAND_MAX;
  g_init_count++;
}

// DataUtils.cpp:589-595.
void orc_init_scalar(double* d) {
  const double factor = (g_init_count % 2 ? 0.1 : 0.2);
  *d = factor * 1.1 / 1.12345;
  g_init_count++;
}

// DataUtils.cpp:601-621: Kahan sum in long double of (|sin(j+1)|+0.5)*x[j];
// the weight is computed in double (std::sin of a double argument).
long double orc_checksum(const double* p, int64_t n, double scale) {
  long double tchk = 0.0;
  long double ckahan = 0.0;
  for (size_t j = 0; j < size_t(n); ++j) {
    long double x = (std::abs(std::sin(j + 1.0)) + 0.5) * static_cast<long double>(p[j]);
    long double y = x - ckahan;
    volatile long double t = tchk + y;
    volatile long double z = t - tchk;
    ckahan = z - y;
    tchk = t;
  }
  tchk *= scale;
  return tchk;
}

// A long double split exactly into two doubles (hi + lo), so callers without
// an 80-bit type see every bit.
void orc_split_ld(long double v, double* hi, double* lo) {
  *hi = static_cast<double>(v);
  *lo = static_cast<double>(v - static_cast<long double>(*hi));
}

// The i-weighted FP64 sum sum_i (i+1)*x[i], summed in index order with no
// fused multiply-add: a second per-kernel checksum, the "sum of i-weighted
// elements in FP64" of the stream workload's description.  Weights start at
// 1 (as the suite's |sin(j+1)| weights do) so that a scalar result (DOT,
// REDUCE_SUM) contributes itself.
double orc_iweighted_sum(const double* p, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += static_cast<double>(i + 1) * p[i];
  return s;
}

void orc_checksum_parts(const double* p, int64_t n, double scale, double* hi, double* lo) {
  orc_split_ld(orc_checksum(p, n, scale), hi, lo);
}

long double orc_checksum_int(const int* p, int64_t n, double scale) {
  long double tchk = 0.0;
  long double ckahan = 0.0;
  for (size_t j = 0; j < size_t(n); ++j) {
    long double x = (std::abs(std::sin(j + 1.0)) + 0.5) * static_cast<long double>(p[j]);
    long double y = x - ckahan;
    volatile long double t = tchk + y;
    volatile long double z = t - tchk;
    ckahan = z - y;
    tchk = t;
  }
  tchk *= scale;
  return tchk;
}

// ---- seeded mode ----------------------------------------------------------

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// SplitMix64 next(): state += golden gamma; return mix64(state).
uint64_t orc_splitmix64_next(uint64_t* state) {
  *state += 0x9E3779B97F4A7C15ull;
  return mix64(*state);
}

// x[i] = lo + (hi-lo)*u_i, u_i = (s_i >> 11) * 2^-53, s_i the SplitMix64
// sequence from state0 = mix64(seed ^ stream_id*0xD1B54A32D192ED03).
void orc_fill_uniform(double* p, int64_t n, uint64_t seed, uint64_t stream_id, double lo, double hi) {
  uint64_t state = mix64(seed ^ (stream_id * 0xD1B54A32D192ED03ull));
  const double w = hi - lo;
  for (int64_t i = 0; i < n; ++i) {
    const double u = double(orc_splitmix64_next(&state) >> 11) * 0x1.0p-53;
    const double t = w * u;
    p[i] = lo + t;
  }
}

// Symmetric tensor field [e][s][q3]: diagonal components (11,22,33) uniform
// in [0.5,2], off-diagonal (12,13,23) uniform in [-0.25,0.25].  Component
// order (11,12,13,22,23,33) for sym == 6; for sym == 9 the row-major 3x3.
void orc_fill_sym_tensor(double* p, int64_t ne, int q3, int sym, uint64_t seed, uint64_t stream_id) {
  uint64_t state = mix64(seed ^ (stream_id * 0xD1B54A32D192ED03ull));
  const int64_t n = ne * int64_t(sym) * q3;
  for (int64_t i = 0; i < n; ++i) {
    const double u = double(orc_splitmix64_next(&state) >> 11) * 0x1.0p-53;
    const int s = int((i / q3) % sym);
    const bool diag = (sym == 6) ? (s == 0 || s == 3 || s == 5) : (s == 0 || s == 4 || s == 8);
    const double lo = diag ? 0.5 : -0.25;
    const double w = diag ? 1.5 : 0.5;
    const double t = w * u;
    p[i] = lo + t;
  }
}

// Legendre P_n(x) and P'_n(x) by the three-term recurrence.
static void legendre(int n, double x, double* p, double* dp) {
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

// Gauss-Legendre points on [0,1], mirrored so x[n-1-i] = 1 - x[i] exactly.
static std::vector<double> gauss_legendre_01(int n) {
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
    // z descends from near +1; map to ascending [0,1].
    x[n - 1 - i] = 0.5 * (z + 1.0);
    x[i] = 1.0 - x[n - 1 - i];
  }
  if (n % 2 == 1) x[n / 2] = 0.5;
  return x;
}

// Gauss-Lobatto nodes on [0,1]: endpoints and the roots of P'_{n-1}.
static std::vector<double> gauss_lobatto_01(int n) {
  std::vector<double> x(n);
  if (n == 1) {
    x[0] = 0.5;
    return x;
  }
  x[0] = 0.0;
  x[n - 1] = 1.0;
  const int m = n - 1;  // roots of P'_m
  for (int i = 1; i <= (n - 1) / 2; ++i) {
    // Chebyshev-Gauss-Lobatto initial guess, descending from +1.
    double z = std::cos(M_PI * i / m);
    for (int it = 0; it < 100; ++it) {
      // Newton on P'_m using P''_m from the Legendre ODE.
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

void orc_fem_basis(int d1d, int q1d, double* B, double* G) {
  const std::vector<double> xd = gauss_lobatto_01(d1d);
  const std::vector<double> xq = gauss_legendre_01(q1d);
  // Compute the half q <= (q1d-1)/2 (all d) and mirror, so that
  // B(q,d) = B(Q-1-q, D-1-d) and G(q,d) = -G(Q-1-q, D-1-d) hold exactly.
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
  // Centre row (odd q1d): enforce the symmetry on the row itself.
  if (q1d % 2 == 1) {
    const int q = q1d / 2;
    for (int d = 0; d < d1d / 2; ++d) {
      B[q + q1d * (d1d - 1 - d)] = B[q + q1d * d];
      G[q + q1d * (d1d - 1 - d)] = -G[q + q1d * d];
    }
    if (d1d % 2 == 1) G[q + q1d * (d1d / 2)] = 0.0;
  }
}

}  // extern "C"
