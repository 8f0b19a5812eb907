// Oracle restatement of KernelBase::execute for the Base_Seq / Base_OpenMP
// variants (TEST INFRASTRUCTURE).
//
// For each kernel: problem shape from the target size and rep count
// (src/common/KernelBase.cpp:115-136), setUp in the reference's
// initialisation order (each kernel's .cpp setUp), the timed rep loop, and
// updateChecksum; all per pass, with checksums accumulated over passes
// (KernelBase.cpp:359-377, Executor.cpp:645-726).
//
// Seeded input mode replaces every formula/random array initialiser by
// orc_fill_uniform(seed, stream_id = the initialiser's call ordinal) in the
// kernel's range, keeps the constant initialisers, and uses fixed scalars.
#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "oracle.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Ctx {
  const orc_params* p;
  bool seeded() const { return p->input_mode == 1; }
  int64_t target(int64_t default_size) const {  // KernelBase.cpp:115-125
    if (p->size > 0) return p->size;
    return static_cast<int64_t>(default_size * (p->sizefact > 0 ? p->sizefact : 1.0));
  }
  int64_t reps(int64_t default_reps) const {  // KernelBase.cpp:127-136
    if (p->checkrun_reps > 0) return p->checkrun_reps;
    return static_cast<int64_t>(default_reps * (p->repfact > 0 ? p->repfact : 1.0));
  }
  // allocAndInitData / allocAndInitDataRandValue in seeded mode.
  void fill(std::vector<double>& v, double lo, double hi) const {
    const int sid = orc_get_init_count();
    orc_fill_uniform(v.data(), int64_t(v.size()), p->seed, uint64_t(sid), lo, hi);
    orc_init_data_const(nullptr, 0, 0.0);  // advances the init counter
  }
  void init_data(std::vector<double>& v, double lo, double hi) const {
    if (seeded()) fill(v, lo, hi);
    else orc_init_data(v.data(), int64_t(v.size()));
  }
  void init_rand_value(std::vector<double>& v, double lo, double hi) const {
    if (seeded()) fill(v, lo, hi);
    else orc_init_data_rand_value(v.data(), int64_t(v.size()));
  }
  void init_const(std::vector<double>& v, double val) const { orc_init_data_const(v.data(), int64_t(v.size()), val); }
  double init_scalar(double seeded_value) const {
    double d;
    orc_init_scalar(&d);
    return seeded() ? seeded_value : d;
  }
};

struct Timer {
  Clock::time_point t0;
  double total = 0.0;
  void start() { t0 = Clock::now(); }
  void stop() { total += std::chrono::duration<double>(Clock::now() - t0).count(); }
};

// The suite checksum of an array, with the array's i-weighted sum added to
// the pass's running total.
double g_iweighted = 0.0;
long double checksum(const double* p, int64_t n, double scale) {
  g_iweighted += orc_iweighted_sum(p, n);
  return orc_checksum(p, n, scale);
}

int variant_id(const char* v) {
  if (std::strcmp(v, "Base_Seq") == 0) return 0;
  if (std::strcmp(v, "Base_OpenMP") == 0) return 1;
  return -1;
}

// ---- stream (src/stream/*.cpp constructors and setUp) -----------------------

long double run_triad(const Ctx& c, int vid, orc_result* r, Timer& t) {
  const int64_t n = c.target(1000000), reps = c.reps(1000);  // TRIAD.cpp:24-34
  r->problem_size = n;
  r->reps = reps;
  r->bytes_per_rep = 3 * 8 * n;
  r->flops_per_rep = 2 * n;
  std::vector<double> a(n), b(n), cc(n);
  c.init_const(a, 0.0);  // TRIAD.cpp:73-79
  c.init_data(b, 0.0, 1.0);
  c.init_data(cc, 0.0, 1.0);
  const double alpha = c.init_scalar(0.75);
  t.start();
  for (int64_t irep = 0; irep < reps; ++irep) orc_triad(vid, a.data(), b.data(), cc.data(), alpha, n);
  t.stop();
  const long double scale = 0.001 * (static_cast<long double>(1000000) / n);  // TRIAD.cpp:36-38
  return checksum(a.data(), n, static_cast<double>(scale));
}

long double run_copy(const Ctx& c, int vid, orc_result* r, Timer& t) {
  const int64_t n = c.target(1000000), reps = c.reps(1800);  // COPY.cpp:24-34
  r->problem_size = n;
  r->reps = reps;
  r->bytes_per_rep = 2 * 8 * n;
  r->flops_per_rep = 0;
  std::vector<double> a(n), cc(n);
  c.init_data(a, 0.0, 1.0);  // COPY.cpp:71-75
  c.init_const(cc, 0.0);
  t.start();
  for (int64_t irep = 0; irep < reps; ++irep) orc_copy(vid, cc.data(), a.data(), n);
  t.stop();
  return checksum(cc.data(), n, 1.0);
}

long double run_mul(const Ctx& c, int vid, orc_result* r, Timer& t) {
  const int64_t n = c.target(1000000), reps = c.reps(1800);  // MUL.cpp:24-34
  r->problem_size = n;
  r->reps = reps;
  r->bytes_per_rep = 2 * 8 * n;
  r->flops_per_rep = n;
  std::vector<double> b(n), cc(n);
  c.init_const(b, 0.0);  // MUL.cpp:71-76
  c.init_data(cc, 0.0, 1.0);
  const double alpha = c.init_scalar(0.75);
  t.start();
  for (int64_t irep = 0; irep < reps; ++irep) orc_mul(vid, b.data(), cc.data(), alpha, n);
  t.stop();
  return checksum(b.data(), n, 1.0);
}

long double run_add(const Ctx& c, int vid, orc_result* r, Timer& t) {
  const int64_t n = c.target(1000000), reps = c.reps(1000);  // ADD.cpp:24-34
  r->problem_size = n;
  r->reps = reps;
  r->bytes_per_rep = 3 * 8 * n;
  r->flops_per_rep = n;
  std::vector<double> a(n), b(n), cc(n);
  c.init_data(a, 0.0, 1.0);  // ADD.cpp:71-76
  c.init_data(b, 0.0, 1.0);
  c.init_const(cc, 0.0);
  t.start();
  for (int64_t irep = 0; irep < reps; ++irep) orc_add(vid, cc.data(), a.data(), b.data(), n);
  t.stop();
  return checksum(cc.data(), n, 1.0);
}

long double run_dot(const Ctx& c, int vid, orc_result* r, Timer& t) {
  const int64_t n = c.target(1000000), reps = c.reps(2000);  // DOT.cpp:24-35
  r->problem_size = n;
  r->reps = reps;
  r->bytes_per_rep = 8 + 2 * 8 * n + 8;
  r->flops_per_rep = 2 * n;
  std::vector<double> a(n), b(n);
  c.init_data(a, 0.0, 1.0);  // DOT.cpp:68-74
  c.init_data(b, 0.0, 1.0);
  double m_dot = 0.0;
  t.start();
  for (int64_t irep = 0; irep < reps; ++irep) m_dot += orc_dot(vid, a.data(), b.data(), n);  // DOT-Seq.cpp:39-45
  t.stop();
  g_iweighted += m_dot;  // a scalar's i-weighted sum is itself
  return m_dot;         // DOT.cpp:78-81
}

// ---- algorithm (src/algorithm/*.cpp) ----------------------------------------

long double run_reduce_sum(const Ctx& c, int vid, orc_result* r, Timer& t) {
  const int64_t n = c.target(1000000), reps = c.reps(50);  // REDUCE_SUM.cpp:24-36
  r->problem_size = n;
  r->reps = reps;
  r->bytes_per_rep = 8 * (1 + n) + 8;
  r->flops_per_rep = n;
  std::vector<double> x(n);
  c.init_data(x, -1.0, 1.0);  // REDUCE_SUM.cpp:66-71
  double m_sum = 0.0;
  t.start();
  for (int64_t irep = 0; irep < reps; ++irep) m_sum = orc_reduce_sum(vid, x.data(), n);  // -Seq.cpp:38-45
  t.stop();
  return checksum(&m_sum, 1, 1.0);  // REDUCE_SUM.cpp:73-76
}

long double run_scan(const Ctx& c, int vid, orc_result* r, Timer& t) {
  const int64_t n = c.target(1000000), reps = c.reps(100);  // SCAN.cpp:24-39
  r->problem_size = n;
  r->reps = reps;
  r->bytes_per_rep = 2 * 8 * n;
  r->flops_per_rep = n;
  std::vector<double> x(n), y(n);
  c.init_rand_value(x, -1.0, 1.0);  // SCAN.cpp:68-72
  c.init_const(y, 0.0);
  t.start();
  for (int64_t irep = 0; irep < reps; ++irep) orc_scan(vid, x.data(), y.data(), n);
  t.stop();
  const long double scale =
      1e-2 * (static_cast<long double>(1000000) / n) / n;  // SCAN.cpp:36-39
  return checksum(y.data(), n, static_cast<double>(scale));
}

long double run_sort(const Ctx& c, int vid, orc_result* r, Timer& t) {
  if (vid != 0) return NAN;  // SORT.cpp:40-47: no Base_OpenMP
  const int64_t n = c.target(1000000), reps = c.reps(20);  // SORT.cpp:24-36
  r->problem_size = n;
  r->reps = reps;
  r->bytes_per_rep = 2 * 8 * n;
  r->flops_per_rep = 0;
  std::vector<double> x(size_t(n) * reps);
  c.init_rand_value(x, -1.0, 1.0);  // SORT.cpp:54-57
  t.start();
  for (int64_t irep = 0; irep < reps; ++irep) orc_sort(x.data() + n * irep, n);  // SORT-Seq.cpp:31-42
  t.stop();
  return checksum(x.data(), n * reps, 1.0);
}

// ---- polybench ----------------------------------------------------------------

long double run_gemm(const Ctx& c, int vid, orc_result* r, Timer& t) {
  const int64_t target = c.target(1000 * 1000);  // POLYBENCH_GEMM.cpp:24-48
  const int64_t ni = static_cast<int64_t>(std::sqrt(double(target)) + std::sqrt(2.0) - 1);
  const int64_t nj = ni;
  const int64_t nk = c.p->nk > 0 ? c.p->nk : static_cast<int64_t>(double(1200) / 1000 * ni);
  const int64_t reps = c.reps(4);
  const double alpha = c.p->has_alpha ? c.p->alpha : (c.seeded() ? 1.5 : 0.62);  // or --alpha / --beta
  const double beta = c.p->has_beta ? c.p->beta : (c.seeded() ? 1.2 : 1.002);
  r->problem_size = ni * nj;
  r->reps = reps;
  r->bytes_per_rep = 8 * ni * nk + 8 * nj * nk + 8 * ni * nj;
  r->flops_per_rep = (1 + 3 * nk) * ni * nj;
  std::vector<double> A(ni * nk), B(nk * nj), C(ni * nj);
  c.init_data(A, -1.0, 1.0);  // POLYBENCH_GEMM.cpp:85-91
  c.init_data(B, -1.0, 1.0);
  c.init_const(C, 0.0);
  t.start();
  for (int64_t irep = 0; irep < reps; ++irep) orc_gemm(vid, C.data(), A.data(), B.data(), alpha, beta, ni, nj, nk);
  t.stop();
  const long double scale = 0.001 * (static_cast<long double>(1000 * 1000) / (ni * nj));  // :50-52
  return checksum(C.data(), ni * nj, static_cast<double>(scale));
}

void seeded_grid(std::vector<double>& g, int64_t N, uint64_t seed, int sid) {
  orc_fill_uniform(g.data(), N * N, seed, uint64_t(sid), 0.0, 1.0);
  for (int64_t i = 0; i < N; ++i)
    for (int64_t j = 0; j < N; ++j)
      if (i == 0 || j == 0 || i == N - 1 || j == N - 1) g[j + i * N] = 0.0;
}

long double run_jacobi2d(const Ctx& c, int vid, orc_result* r, Timer& t) {
  const int64_t default_size = (1002 - 2) * (1002 - 2);  // POLYBENCH_JACOBI_2D.cpp:24-45
  const int64_t target = c.target(default_size);
  const int64_t N = static_cast<int64_t>(std::sqrt(double(target)) + 2 + std::sqrt(2.0) - 1);
  const int64_t tsteps = c.p->tsteps > 0 ? c.p->tsteps : 40;
  const int64_t reps = c.reps(50);
  r->problem_size = (N - 2) * (N - 2);
  r->reps = reps;
  r->bytes_per_rep = (8 * (N * N - 4) * 2 + 8 * (N - 2) * (N - 2) * 2) * tsteps;
  r->flops_per_rep = tsteps * 10 * (N - 2) * (N - 2);
  std::vector<double> Ainit(N * N), Binit(N * N), A(N * N), B(N * N);
  if (c.seeded()) {  // interior uniform [0,1), zero boundary
    seeded_grid(Ainit, N, c.p->seed, orc_get_init_count());
    orc_init_data_const(nullptr, 0, 0.0);
    seeded_grid(Binit, N, c.p->seed, orc_get_init_count());
    orc_init_data_const(nullptr, 0, 0.0);
  } else {  // POLYBENCH_JACOBI_2D.cpp:82-89
    orc_init_data(Ainit.data(), N * N);
    orc_init_data(Binit.data(), N * N);
  }
  A = Ainit;  // POLYBENCH_JACOBI_2D.hpp:30-38 (DATA_SETUP, before the timer)
  B = Binit;
  t.start();
  for (int64_t irep = 0; irep < reps; ++irep) orc_jacobi2d(vid, A.data(), B.data(), N, tsteps);
  t.stop();
  const long double scale = 0.0001 * (static_cast<long double>(default_size) / ((N - 2) * (N - 2)));
  long double cs = checksum(A.data(), N * N, static_cast<double>(scale));
  cs += checksum(B.data(), N * N, static_cast<double>(scale));  // :91-95
  return cs;
}

// ---- apps -----------------------------------------------------------------------

long double run_mass3dpa(const Ctx& c, int vid, orc_result* r, Timer& t) {
  const int D1 = c.p->d1d > 0 ? c.p->d1d : 4;  // MASS3DPA.hpp:172-173
  const int Q1 = c.p->q1d > 0 ? c.p->q1d : (c.p->d1d > 0 ? D1 + 1 : 5);
  const int64_t Q3 = int64_t(Q1) * Q1 * Q1, D3 = int64_t(D1) * D1 * D1;
  const int64_t target = c.target(8000 * Q3);  // MASS3DPA.cpp:26-49
  int64_t NE = (target + Q3 / 2) / Q3;
  if (NE < 1) NE = 1;
  const int64_t reps = c.reps(50);
  r->problem_size = NE * Q3;
  r->reps = reps;
  r->bytes_per_rep = 2 * 8 * Q1 * D1 + 2 * 8 * D3 * NE + 8 * Q3 * NE + 8 * D3 * NE;
  r->flops_per_rep = NE * (2 * D3 * Q1 + 2 * D1 * D1 * Q1 * Q1 + 2 * D1 * Q3 + Q3 + 2 * Q3 * D1 +
                           2 * Q1 * Q1 * D1 * D1 + 2 * Q1 * D3 + D3);
  std::vector<double> B(Q1 * D1), Bt(Q1 * D1), D(Q3 * NE), X(D3 * NE), Y(D3 * NE);
  if (c.seeded()) {
    std::vector<double> G(Q1 * D1);
    orc_fem_basis(D1, Q1, B.data(), G.data());
    orc_init_data_const(nullptr, 0, 0.0);
    for (int q = 0; q < Q1; ++q)  // Bt holds B^T as a d1d x q1d column-major array
      for (int d = 0; d < D1; ++d) Bt[d + D1 * q] = B[q + Q1 * d];
    orc_init_data_const(nullptr, 0, 0.0);
    c.fill(D, 0.5, 2.0);
    c.fill(X, -1.0, 1.0);
    c.init_const(Y, 0.0);
  } else {  // MASS3DPA.cpp:76-84
    c.init_const(B, 1.0);
    c.init_const(Bt, 1.0);
    c.init_const(D, 1.0);
    c.init_const(X, 1.0);
    c.init_const(Y, 0.0);
  }
  t.start();
  for (int64_t irep = 0; irep < reps; ++irep)
    orc_mass3dpa(vid, B.data(), Bt.data(), D.data(), X.data(), Y.data(), NE, D1, Q1);
  t.stop();
  return checksum(Y.data(), D3 * NE, 1.0);  // MASS3DPA.cpp:86-89
}

long double run_diffusion3dpa(const Ctx& c, int vid, orc_result* r, Timer& t) {
  const int D1 = c.p->d1d > 0 ? c.p->d1d : 3;  // DIFFUSION3DPA.hpp:231-233
  const int Q1 = c.p->q1d > 0 ? c.p->q1d : D1 + 1;
  const int SYM = 6;
  const int64_t Q3 = int64_t(Q1) * Q1 * Q1, D3 = int64_t(D1) * D1 * D1;
  const int64_t target = c.target(15625 * Q3);  // DIFFUSION3DPA.cpp:26-53
  int64_t NE = (target + Q3 / 2) / Q3;
  if (NE < 1) NE = 1;
  const int64_t reps = c.reps(50);
  const int64_t QD = Q1 * D1;
  r->problem_size = NE * Q3;
  r->reps = reps;
  r->bytes_per_rep = 2 * 8 * QD + 2 * 8 * D3 * NE + SYM * 8 * Q3 * NE + 8 * D3 * NE;
  r->flops_per_rep = NE * (QD + 5 * D1 * D1 * Q1 * D1 + 7 * D1 * D1 * Q1 * Q1 + 7 * Q1 * D1 * Q1 * Q1 +
                           15 * Q3 + QD + 7 * Q1 * Q1 * D1 * Q1 + 7 * Q1 * Q1 * D1 * D1 + 7 * D1 * Q1 * D1 * D1 +
                           3 * D3);
  std::vector<double> B(QD), G(QD), D(Q3 * SYM * NE), X(D3 * NE), Y(D3 * NE);
  if (c.seeded()) {
    orc_fem_basis(D1, Q1, B.data(), G.data());
    orc_init_data_const(nullptr, 0, 0.0);
    orc_init_data_const(nullptr, 0, 0.0);
    orc_fill_sym_tensor(D.data(), NE, int(Q3), SYM, c.p->seed, uint64_t(orc_get_init_count()));
    orc_init_data_const(nullptr, 0, 0.0);
    c.fill(X, -1.0, 1.0);
    c.init_const(Y, 0.0);
  } else {  // DIFFUSION3DPA.cpp:80-88
    c.init_const(B, 1.0);
    c.init_const(G, 1.0);
    c.init_const(D, 1.0);
    c.init_const(X, 1.0);
    c.init_const(Y, 0.0);
  }
  t.start();
  for (int64_t irep = 0; irep < reps; ++irep)
    orc_diffusion3dpa(vid, B.data(), G.data(), D.data(), X.data(), Y.data(), NE, D1, Q1, 1);
  t.stop();
  return checksum(Y.data(), D3 * NE, 1.0);  // DIFFUSION3DPA.cpp:90-93
}

using RunFn = long double (*)(const Ctx&, int, orc_result*, Timer&);
struct Entry {
  const char* name;
  RunFn fn;
};
const Entry kTable[] = {
    {"Stream_COPY", run_copy},           {"Stream_MUL", run_mul},
    {"Stream_ADD", run_add},             {"Stream_TRIAD", run_triad},
    {"Stream_DOT", run_dot},             {"Algorithm_REDUCE_SUM", run_reduce_sum},
    {"Algorithm_SCAN", run_scan},        {"Algorithm_SORT", run_sort},
    {"Polybench_GEMM", run_gemm},        {"Polybench_JACOBI_2D", run_jacobi2d},
    {"Apps_MASS3DPA", run_mass3dpa},     {"Apps_DIFFUSION3DPA", run_diffusion3dpa},
};

}  // namespace

extern "C" int orc_run(const char* kernel, const char* variant, const orc_params* p, orc_result* out) {
  if (!kernel || !variant || !p || !out) return -1;
  const int vid = variant_id(variant);
  if (vid < 0) return -1;
  if (vid == 1 && std::strcmp(kernel, "Algorithm_SORT") == 0) return -1;  // SORT.cpp:40-47
  RunFn fn = nullptr;
  for (const Entry& e : kTable)
    if (std::strcmp(e.name, kernel) == 0) fn = e.fn;
  if (!fn) return -1;
  if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
  Ctx c{p};
  Timer t;
  std::memset(out, 0, sizeof(*out));
  long double cs = 0.0;
  const int npasses = p->npasses > 0 ? p->npasses : 1;
  g_iweighted = 0.0;
  for (int pass = 0; pass < npasses; ++pass) {
    orc_reset_init_count();  // KernelBase.cpp:366
    cs += fn(c, vid, out, t);
  }
  out->checksum = cs;
  orc_split_ld(cs, &out->checksum_hi, &out->checksum_lo);
  out->time_total_s = t.total;
  out->iweighted = g_iweighted;
  return 0;
}
