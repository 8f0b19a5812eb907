// Oracle restatement of the Base_Seq / Base_OpenMP loop bodies
// (TEST INFRASTRUCTURE).  variant 0 = Base_Seq, variant 1 = Base_OpenMP.
#include <algorithm>
#include <vector>

#include <omp.h>

#include "oracle.h"

extern "C" {

// src/stream/COPY.hpp:24-25; COPY-Seq.cpp / COPY-OMP.cpp Base cases.
void orc_copy(int variant, double* c, const double* a, int64_t n) {
  if (variant == 1) {
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) c[i] = a[i];
  } else {
    for (int64_t i = 0; i < n; ++i) c[i] = a[i];
  }
}

// src/stream/MUL.hpp:25-26.
void orc_mul(int variant, double* b, const double* c, double alpha, int64_t n) {
  if (variant == 1) {
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) b[i] = alpha * c[i];
  } else {
    for (int64_t i = 0; i < n; ++i) b[i] = alpha * c[i];
  }
}

// src/stream/ADD.hpp:25-26.
void orc_add(int variant, double* c, const double* a, const double* b, int64_t n) {
  if (variant == 1) {
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) c[i] = a[i] + b[i];
  } else {
    for (int64_t i = 0; i < n; ++i) c[i] = a[i] + b[i];
  }
}

// src/stream/TRIAD.hpp:26-27; TRIAD-Seq.cpp:37-50, TRIAD-OMP.cpp:37-51.
void orc_triad(int variant, double* a, const double* b, const double* c, double alpha, int64_t n) {
  if (variant == 1) {
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) a[i] = b[i] + alpha * c[i];
  } else {
    for (int64_t i = 0; i < n; ++i) a[i] = b[i] + alpha * c[i];
  }
}

// src/stream/DOT.hpp:24-25; DOT-Seq.cpp:38-43 (one rep), DOT-OMP.cpp:38-46.
double orc_dot(int variant, const double* a, const double* b, int64_t n) {
  double dot = 0.0;
  if (variant == 1) {
#pragma omp parallel for reduction(+ : dot)
    for (int64_t i = 0; i < n; ++i) dot += a[i] * b[i];
  } else {
    for (int64_t i = 0; i < n; ++i) dot += a[i] * b[i];
  }
  return dot;
}

// src/algorithm/REDUCE_SUM.hpp:31-32; REDUCE_SUM-Seq.cpp:38-45, -OMP.cpp:33-51.
double orc_reduce_sum(int variant, const double* x, int64_t n) {
  double sum = 0.0;
  if (variant == 1) {
#pragma omp parallel for reduction(+ : sum)
    for (int64_t i = 0; i < n; ++i) sum += x[i];
  } else {
    for (int64_t i = 0; i < n; ++i) sum += x[i];
  }
  return sum;
}

// src/algorithm/SCAN.hpp:27-32; SCAN-Seq.cpp:31-45 and the two-pass
// per-thread Base_OpenMP path SCAN-OMP.cpp:33-90 (non-OpenMP-5 branch).
void orc_scan(int variant, const double* x, double* y, int64_t n) {
  if (variant != 1) {
    double scan_var = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      y[i] = scan_var;
      scan_var += x[i];
    }
    return;
  }
  const int64_t ibegin = 0, iend = n;
  const int p0 = int(std::min<int64_t>(n, int64_t(omp_get_max_threads())));
  if (p0 <= 0) return;
  std::vector<double> thread_sums(p0);
  double scan_var = 0.0;
#pragma omp parallel num_threads(p0)
  {
    const int p = omp_get_num_threads();
    const int pid = omp_get_thread_num();
    const int64_t step = n / p;
    const int64_t local_begin = pid * step + ibegin;
    const int64_t local_end = (pid == p - 1) ? iend : (pid + 1) * step + ibegin;
    double local_scan_var = (pid == 0) ? scan_var : 0;
    for (int64_t i = local_begin; i < local_end; ++i) {
      y[i] = local_scan_var;
      local_scan_var += x[i];
    }
    thread_sums[pid] = local_scan_var;
#pragma omp barrier
    if (pid != 0) {
      double prev_sum = 0;
      for (int ip = 0; ip < pid; ++ip) prev_sum += thread_sums[ip];
      for (int64_t i = local_begin; i < local_end; ++i) y[i] += prev_sum;
    }
  }
}

// src/algorithm/SORT-Seq.cpp:31-42: std::sort of one segment.
void orc_sort(double* x, int64_t n) { std::sort(x, x + n); }

// src/polybench/POLYBENCH_JACOBI_2D.hpp:41-45, -Seq.cpp:30-54, -OMP.cpp:32-58.
void orc_jacobi2d(int variant, double* A, double* B, int64_t N, int64_t tsteps) {
  for (int64_t t = 0; t < tsteps; ++t) {
    if (variant == 1) {
#pragma omp parallel for
      for (int64_t i = 1; i < N - 1; ++i)
        for (int64_t j = 1; j < N - 1; ++j)
          B[j + i * N] = 0.2 * (A[j + i * N] + A[j - 1 + i * N] + A[j + 1 + i * N] + A[j + (i + 1) * N] +
                                A[j + (i - 1) * N]);
#pragma omp parallel for
      for (int64_t i = 1; i < N - 1; ++i)
        for (int64_t j = 1; j < N - 1; ++j)
          A[j + i * N] = 0.2 * (B[j + i * N] + B[j - 1 + i * N] + B[j + 1 + i * N] + B[j + (i + 1) * N] +
                                B[j + (i - 1) * N]);
    } else {
      for (int64_t i = 1; i < N - 1; ++i)
        for (int64_t j = 1; j < N - 1; ++j)
          B[j + i * N] = 0.2 * (A[j + i * N] + A[j - 1 + i * N] + A[j + 1 + i * N] + A[j + (i + 1) * N] +
                                A[j + (i - 1) * N]);
      for (int64_t i = 1; i < N - 1; ++i)
        for (int64_t j = 1; j < N - 1; ++j)
          A[j + i * N] = 0.2 * (B[j + i * N] + B[j - 1 + i * N] + B[j + 1 + i * N] + B[j + (i + 1) * N] +
                                B[j + (i - 1) * N]);
    }
  }
}

// src/polybench/POLYBENCH_GEMM.hpp:27-50; -Seq.cpp:30-50 (i-j-k, C *= beta
// then overwritten by the dot), -OMP.cpp:32-52 (collapse(2) over i,j).
static inline void gemm_ij(double* C, const double* A, const double* B, double alpha, double beta, int64_t i,
                           int64_t j, int64_t nj, int64_t nk) {
  double dot = 0.0;
  C[j + i * nj] *= beta;
  for (int64_t k = 0; k < nk; ++k) dot += alpha * A[k + i * nk] * B[j + k * nj];
  C[j + i * nj] = dot;
}

void orc_gemm(int variant, double* C, const double* A, const double* B, double alpha, double beta, int64_t ni,
              int64_t nj, int64_t nk) {
  if (variant == 1) {
#pragma omp parallel for collapse(2)
    for (int64_t i = 0; i < ni; ++i)
      for (int64_t j = 0; j < nj; ++j) gemm_ij(C, A, B, alpha, beta, i, j, nj, nk);
  } else {
    for (int64_t i = 0; i < ni; ++i)
      for (int64_t j = 0; j < nj; ++j) gemm_ij(C, A, B, alpha, beta, i, j, nj, nk);
  }
}

// ---- MASS3DPA: src/apps/MASS3DPA.hpp:158-339, MASS3DPA-Seq.cpp:29-94 ------
// One element, with the phase structure and loop orders of the Seq variant
// generalised to runtime (d1d, q1d); Bt is read as Bt[q + d1d*d]
// (MASS3DPA.hpp:175, MASS3DPA_6).
static void mass3dpa_element(const double* B, const double* Bt, const double* D, const double* X, double* Y,
                             int64_t e, int D1, int Q1) {
  const int MDQ = Q1 > D1 ? Q1 : D1;
  std::vector<double> sDQ(size_t(Q1) * D1), sm0(size_t(MDQ) * MDQ * MDQ), sm1(size_t(MDQ) * MDQ * MDQ);
  auto Bsmem = [&](int q, int d) -> double& { return sDQ[size_t(q) * D1 + d]; };   // [MQ1][MD1]
  auto Btsmem = [&](int d, int q) -> double& { return sDQ[size_t(d) * Q1 + q]; };  // [MD1][MQ1]
  auto Xsmem = [&](int dz, int dy, int dx) -> double& { return sm0[(size_t(dz) * D1 + dy) * D1 + dx]; };
  auto DDQ = [&](int dz, int dy, int qx) -> double& { return sm1[(size_t(dz) * D1 + dy) * Q1 + qx]; };
  auto DQQ = [&](int dz, int qy, int qx) -> double& { return sm0[(size_t(dz) * Q1 + qy) * Q1 + qx]; };
  auto QQQ = [&](int qz, int qy, int qx) -> double& { return sm1[(size_t(qz) * Q1 + qy) * Q1 + qx]; };
  auto QQD = [&](int qz, int qy, int dx) -> double& { return sm0[(size_t(qz) * Q1 + qy) * D1 + dx]; };
  auto QDD = [&](int qz, int dy, int dx) -> double& { return sm1[(size_t(qz) * D1 + dy) * D1 + dx]; };
  const size_t D3 = size_t(D1) * D1 * D1, Q3 = size_t(Q1) * Q1 * Q1;
  auto Xg = [&](int dx, int dy, int dz) { return X[dx + D1 * dy + D1 * D1 * dz + D3 * e]; };
  auto Yg = [&](int dx, int dy, int dz) -> double& { return Y[dx + D1 * dy + D1 * D1 * dz + D3 * e]; };
  auto Dg = [&](int qx, int qy, int qz) { return D[qx + Q1 * qy + Q1 * Q1 * qz + Q3 * e]; };
  std::vector<double> u(MDQ);

  for (int dy = 0; dy < D1; ++dy) {  // MASS3DPA_1, MASS3DPA_2
    for (int dx = 0; dx < D1; ++dx)
      for (int dz = 0; dz < D1; ++dz) Xsmem(dz, dy, dx) = Xg(dx, dy, dz);
    for (int dx = 0; dx < Q1; ++dx) Bsmem(dx, dy) = B[dx + Q1 * dy];
  }
  for (int dy = 0; dy < D1; ++dy)  // MASS3DPA_3
    for (int qx = 0; qx < Q1; ++qx) {
      for (int dz = 0; dz < D1; dz++) u[dz] = 0;
      for (int dx = 0; dx < D1; ++dx)
        for (int dz = 0; dz < D1; ++dz) u[dz] += Xsmem(dz, dy, dx) * Bsmem(qx, dx);
      for (int dz = 0; dz < D1; ++dz) DDQ(dz, dy, qx) = u[dz];
    }
  for (int qy = 0; qy < Q1; ++qy)  // MASS3DPA_4
    for (int qx = 0; qx < Q1; ++qx) {
      for (int dz = 0; dz < D1; dz++) u[dz] = 0;
      for (int dy = 0; dy < D1; ++dy)
        for (int dz = 0; dz < D1; dz++) u[dz] += DDQ(dz, dy, qx) * Bsmem(qy, dy);
      for (int dz = 0; dz < D1; dz++) DQQ(dz, qy, qx) = u[dz];
    }
  for (int qy = 0; qy < Q1; ++qy)  // MASS3DPA_5
    for (int qx = 0; qx < Q1; ++qx) {
      for (int qz = 0; qz < Q1; qz++) u[qz] = 0;
      for (int dz = 0; dz < D1; ++dz)
        for (int qz = 0; qz < Q1; qz++) u[qz] += DQQ(dz, qy, qx) * Bsmem(qz, dz);
      for (int qz = 0; qz < Q1; qz++) QQQ(qz, qy, qx) = u[qz] * Dg(qx, qy, qz);
    }
  for (int d = 0; d < D1; ++d)  // MASS3DPA_6
    for (int q = 0; q < Q1; ++q) Btsmem(d, q) = Bt[q + D1 * d];
  for (int qy = 0; qy < Q1; ++qy)  // MASS3DPA_7
    for (int dx = 0; dx < D1; ++dx) {
      for (int qz = 0; qz < Q1; ++qz) u[qz] = 0;
      for (int qx = 0; qx < Q1; ++qx)
        for (int qz = 0; qz < Q1; ++qz) u[qz] += QQQ(qz, qy, qx) * Btsmem(dx, qx);
      for (int qz = 0; qz < Q1; ++qz) QQD(qz, qy, dx) = u[qz];
    }
  for (int dy = 0; dy < D1; ++dy)  // MASS3DPA_8
    for (int dx = 0; dx < D1; ++dx) {
      for (int qz = 0; qz < Q1; ++qz) u[qz] = 0;
      for (int qy = 0; qy < Q1; ++qy)
        for (int qz = 0; qz < Q1; ++qz) u[qz] += QQD(qz, qy, dx) * Btsmem(dy, qy);
      for (int qz = 0; qz < Q1; ++qz) QDD(qz, dy, dx) = u[qz];
    }
  for (int dy = 0; dy < D1; ++dy)  // MASS3DPA_9
    for (int dx = 0; dx < D1; ++dx) {
      for (int dz = 0; dz < D1; ++dz) u[dz] = 0;
      for (int qz = 0; qz < Q1; ++qz)
        for (int dz = 0; dz < D1; ++dz) u[dz] += QDD(qz, dy, dx) * Btsmem(dz, qz);
      for (int dz = 0; dz < D1; ++dz) Yg(dx, dy, dz) += u[dz];
    }
}

void orc_mass3dpa(int variant, const double* B, const double* Bt, const double* D, const double* X, double* Y,
                  int64_t NE, int d1d, int q1d) {
  if (variant == 1) {  // MASS3DPA-OMP.cpp:36, parallel over elements
#pragma omp parallel for
    for (int64_t e = 0; e < NE; ++e) mass3dpa_element(B, Bt, D, X, Y, e, d1d, q1d);
  } else {
    for (int64_t e = 0; e < NE; ++e) mass3dpa_element(B, Bt, D, X, Y, e, d1d, q1d);
  }
}

// ---- DIFFUSION3DPA: src/apps/DIFFUSION3DPA.hpp:216-458, -Seq.cpp:28-110 --
// Half-storage index maps DIFFUSION3DPA.hpp:243-268.
static inline int qi(int q, int d, int Q) { return (q <= d) ? q : Q - 1 - q; }
static inline int dj(int q, int d, int D) { return (q <= d) ? d : D - 1 - d; }
static inline int qk(int q, int d, int Q) { return (q <= d) ? Q - 1 - q : q; }
static inline int dl(int q, int d, int D) { return (q <= d) ? D - 1 - d : d; }
static inline double sgn(int q, int d) { return (q <= d) ? -1.0 : 1.0; }

static void diffusion3dpa_element(const double* Basis, const double* dBasis, const double* Dd, const double* X,
                                  double* Y, int64_t e, int D1, int Q1, bool symmetric) {
  const int MDQ = Q1 > D1 ? Q1 : D1;
  const int SYMC = symmetric ? 6 : 9;
  const size_t S = size_t(MDQ) * MDQ * MDQ;
  std::vector<double> sBG(size_t(Q1) * D1), sm0(3 * S), sm1(3 * S);
  auto Bm = [&](int i, int j) -> double& { return sBG[size_t(i) * D1 + j]; };   // B[MQ1][MD1]
  auto Gm = [&](int k, int l) -> double& { return sBG[size_t(k) * D1 + l]; };   // G[MQ1][MD1]
  auto Btm = [&](int j, int i) -> double& { return sBG[size_t(j) * Q1 + i]; };  // Bt[MD1][MQ1]
  auto Gtm = [&](int l, int k) -> double& { return sBG[size_t(l) * Q1 + k]; };  // Gt[MD1][MQ1]
  auto b = [&](int x, int y) { return Basis[x + Q1 * y]; };
  auto g = [&](int x, int y) { return dBasis[x + Q1 * y]; };
  auto s_X = [&](int dz, int dy, int dx) -> double& { return sm0[2 * S + (size_t(dz) * D1 + dy) * D1 + dx]; };
  auto DDQ = [&](int c, int dz, int dy, int qx) -> double& { return sm0[c * S + (size_t(dz) * D1 + dy) * Q1 + qx]; };
  auto DQQ = [&](int c, int dz, int qy, int qx) -> double& { return sm1[c * S + (size_t(dz) * Q1 + qy) * Q1 + qx]; };
  auto QQQ = [&](int c, int qz, int qy, int qx) -> double& { return sm0[c * S + (size_t(qz) * Q1 + qy) * Q1 + qx]; };
  auto QQD = [&](int c, int qz, int qy, int dx) -> double& { return sm1[c * S + (size_t(qz) * Q1 + qy) * D1 + dx]; };
  auto QDD = [&](int c, int qz, int dy, int dx) -> double& { return sm0[c * S + (size_t(qz) * D1 + dy) * D1 + dx]; };
  const size_t D3 = size_t(D1) * D1 * D1, Q3 = size_t(Q1) * Q1 * Q1;
  auto Xg = [&](int dx, int dy, int dz) { return X[dx + D1 * dy + D1 * D1 * dz + D3 * e]; };
  auto Yg = [&](int dx, int dy, int dz) -> double& { return Y[dx + D1 * dy + D1 * D1 * dz + D3 * e]; };
  auto dg = [&](int qx, int qy, int qz, int s) {
    return Dd[qx + Q1 * qy + Q1 * Q1 * qz + Q3 * s + Q3 * SYMC * e];
  };

  for (int dz = 0; dz < D1; ++dz)  // DIFFUSION3DPA_1
    for (int dy = 0; dy < D1; ++dy)
      for (int dx = 0; dx < D1; ++dx) s_X(dz, dy, dx) = Xg(dx, dy, dz);
  for (int dy = 0; dy < D1; ++dy)  // DIFFUSION3DPA_2
    for (int qx = 0; qx < Q1; ++qx) {
      const int i = qi(qx, dy, Q1), j = dj(qx, dy, D1), k = qk(qx, dy, Q1), l = dl(qx, dy, D1);
      Bm(i, j) = b(qx, dy);
      Gm(k, l) = g(qx, dy) * sgn(qx, dy);
    }
  for (int dz = 0; dz < D1; ++dz)  // DIFFUSION3DPA_3
    for (int dy = 0; dy < D1; ++dy)
      for (int qx = 0; qx < Q1; ++qx) {
        double u = 0.0, v = 0.0;
        for (int dx = 0; dx < D1; ++dx) {
          const int i = qi(qx, dx, Q1), j = dj(qx, dx, D1), k = qk(qx, dx, Q1), l = dl(qx, dx, D1);
          const double s = sgn(qx, dx);
          const double coords = s_X(dz, dy, dx);
          u += coords * Bm(i, j);
          v += coords * Gm(k, l) * s;
        }
        DDQ(0, dz, dy, qx) = u;
        DDQ(1, dz, dy, qx) = v;
      }
  for (int dz = 0; dz < D1; ++dz)  // DIFFUSION3DPA_4
    for (int qy = 0; qy < Q1; ++qy)
      for (int qx = 0; qx < Q1; ++qx) {
        double u = 0.0, v = 0.0, w = 0.0;
        for (int dy = 0; dy < D1; ++dy) {
          const int i = qi(qy, dy, Q1), j = dj(qy, dy, D1), k = qk(qy, dy, Q1), l = dl(qy, dy, D1);
          const double s = sgn(qy, dy);
          u += DDQ(1, dz, dy, qx) * Bm(i, j);
          v += DDQ(0, dz, dy, qx) * Gm(k, l) * s;
          w += DDQ(0, dz, dy, qx) * Bm(i, j);
        }
        DQQ(0, dz, qy, qx) = u;
        DQQ(1, dz, qy, qx) = v;
        DQQ(2, dz, qy, qx) = w;
      }
  for (int qz = 0; qz < Q1; ++qz)  // DIFFUSION3DPA_5
    for (int qy = 0; qy < Q1; ++qy)
      for (int qx = 0; qx < Q1; ++qx) {
        double u = 0.0, v = 0.0, w = 0.0;
        for (int dz = 0; dz < D1; ++dz) {
          const int i = qi(qz, dz, Q1), j = dj(qz, dz, D1), k = qk(qz, dz, Q1), l = dl(qz, dz, D1);
          const double s = sgn(qz, dz);
          u += DQQ(0, dz, qy, qx) * Bm(i, j);
          v += DQQ(1, dz, qy, qx) * Bm(i, j);
          w += DQQ(2, dz, qy, qx) * Gm(k, l) * s;
        }
        const double O11 = dg(qx, qy, qz, 0);
        const double O12 = dg(qx, qy, qz, 1);
        const double O13 = dg(qx, qy, qz, 2);
        const double O21 = symmetric ? O12 : dg(qx, qy, qz, 3);
        const double O22 = symmetric ? dg(qx, qy, qz, 3) : dg(qx, qy, qz, 4);
        const double O23 = symmetric ? dg(qx, qy, qz, 4) : dg(qx, qy, qz, 5);
        const double O31 = symmetric ? O13 : dg(qx, qy, qz, 6);
        const double O32 = symmetric ? O23 : dg(qx, qy, qz, 7);
        const double O33 = symmetric ? dg(qx, qy, qz, 5) : dg(qx, qy, qz, 8);
        const double gX = u, gY = v, gZ = w;
        QQQ(0, qz, qy, qx) = (O11 * gX) + (O12 * gY) + (O13 * gZ);
        QQQ(1, qz, qy, qx) = (O21 * gX) + (O22 * gY) + (O23 * gZ);
        QQQ(2, qz, qy, qx) = (O31 * gX) + (O32 * gY) + (O33 * gZ);
      }
  for (int d = 0; d < D1; ++d)  // DIFFUSION3DPA_6
    for (int q = 0; q < Q1; ++q) {
      const int i = qi(q, d, Q1), j = dj(q, d, D1), k = qk(q, d, Q1), l = dl(q, d, D1);
      Btm(j, i) = b(q, d);
      Gtm(l, k) = g(q, d) * sgn(q, d);
    }
  for (int qz = 0; qz < Q1; ++qz)  // DIFFUSION3DPA_7
    for (int qy = 0; qy < Q1; ++qy)
      for (int dx = 0; dx < D1; ++dx) {
        double u = 0.0, v = 0.0, w = 0.0;
        for (int qx = 0; qx < Q1; ++qx) {
          const int i = qi(qx, dx, Q1), j = dj(qx, dx, D1), k = qk(qx, dx, Q1), l = dl(qx, dx, D1);
          const double s = sgn(qx, dx);
          u += QQQ(0, qz, qy, qx) * Gtm(l, k) * s;
          v += QQQ(1, qz, qy, qx) * Btm(j, i);
          w += QQQ(2, qz, qy, qx) * Btm(j, i);
        }
        QQD(0, qz, qy, dx) = u;
        QQD(1, qz, qy, dx) = v;
        QQD(2, qz, qy, dx) = w;
      }
  for (int qz = 0; qz < Q1; ++qz)  // DIFFUSION3DPA_8
    for (int dy = 0; dy < D1; ++dy)
      for (int dx = 0; dx < D1; ++dx) {
        double u = 0.0, v = 0.0, w = 0.0;
        for (int qy = 0; qy < Q1; ++qy) {
          const int i = qi(qy, dy, Q1), j = dj(qy, dy, D1), k = qk(qy, dy, Q1), l = dl(qy, dy, D1);
          const double s = sgn(qy, dy);
          u += QQD(0, qz, qy, dx) * Btm(j, i);
          v += QQD(1, qz, qy, dx) * Gtm(l, k) * s;
          w += QQD(2, qz, qy, dx) * Btm(j, i);
        }
        QDD(0, qz, dy, dx) = u;
        QDD(1, qz, dy, dx) = v;
        QDD(2, qz, dy, dx) = w;
      }
  for (int dz = 0; dz < D1; ++dz)  // DIFFUSION3DPA_9
    for (int dy = 0; dy < D1; ++dy)
      for (int dx = 0; dx < D1; ++dx) {
        double u = 0.0, v = 0.0, w = 0.0;
        for (int qz = 0; qz < Q1; ++qz) {
          const int i = qi(qz, dz, Q1), j = dj(qz, dz, D1), k = qk(qz, dz, Q1), l = dl(qz, dz, D1);
          const double s = sgn(qz, dz);
          u += QDD(0, qz, dy, dx) * Btm(j, i);
          v += QDD(1, qz, dy, dx) * Btm(j, i);
          w += QDD(2, qz, dy, dx) * Gtm(l, k) * s;
        }
        Yg(dx, dy, dz) += (u + v + w);
      }
}

void orc_diffusion3dpa(int variant, const double* Basis, const double* dBasis, const double* D, const double* X,
                       double* Y, int64_t NE, int d1d, int q1d, int symmetric) {
  if (variant == 1) {  // DIFFUSION3DPA-OMP.cpp:35
#pragma omp parallel for
    for (int64_t e = 0; e < NE; ++e) diffusion3dpa_element(Basis, dBasis, D, X, Y, e, d1d, q1d, symmetric != 0);
  } else {
    for (int64_t e = 0; e < NE; ++e) diffusion3dpa_element(Basis, dBasis, D, X, Y, e, d1d, q1d, symmetric != 0);
  }
}

}  // extern "C"
