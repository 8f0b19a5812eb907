/*
 * oracle.h -- CPU restatement of the RAJA Performance Suite's Base_Seq and
 * Base_OpenMP variants for the kernels the MI355X variant implements, plus
 * the suite's input initialisers and checksum.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the rajaperf_amd
 * package, its HIP library or its suite harness) links or calls this code;
 * only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
 * it, as the checker and as the timed CPU baseline.
 *
 * Every function cites the reference file:line it restates.  Loop order,
 * expression association, accumulate/overwrite semantics and the order of
 * the initialisation calls follow the reference literally; the library is
 * built with -ffp-contract=off so no multiply-add is fused.
 *
 * Pinning (tests/test_oracle_*.py): SplitMix64 against its published
 * vectors; initialisers against glibc rand() and closed forms; the checksum
 * against a brute-force long-double sum; stream kernels against closed
 * forms; DOT/REDUCE_SUM against exactly-summable integer inputs; SCAN
 * against difference invariants; SORT against sortedness + permutation;
 * JACOBI_2D against the constant-field fixed point and a brute-force 8x8;
 * GEMM against identity/alpha closed forms and numpy; MASS3DPA against the
 * all-ones closed form (D1D^3*Q1D^3 per entry) and an assembled element
 * matrix; DIFFUSION3DPA against an assembled G^T D G for symmetric bases.
 */
#ifndef RAJAPERF_ORACLE_H
#define RAJAPERF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- DataUtils (src/common/DataUtils.cpp:131-147,477-653) ------------- */
void orc_reset_init_count(void);
int  orc_get_init_count(void);
void orc_init_data(double* p, int64_t n);                 /* :504-513 */
void orc_init_data_int(int* p, int64_t n);                /* :477-498 */
void orc_init_data_const(double* p, int64_t n, double v); /* :518-528 */
void orc_init_data_rand_sign(double* p, int64_t n);       /* :542-556 */
void orc_init_data_rand_value(double* p, int64_t n);      /* :560-569 */
void orc_init_scalar(double* d);                          /* :589-595 */
long double orc_checksum(const double* p, int64_t n, double scale); /* :601-621 */
long double orc_checksum_int(const int* p, int64_t n, double scale);
/* sum_i (i+1)*x[i] in FP64, index order, no FMA (the i-weighted checksum). */
double orc_iweighted_sum(const double* p, int64_t n);
/* Exact two-double split (hi + lo) of a long double / of a checksum. */
void orc_split_ld(long double v, double* hi, double* lo);
void orc_checksum_parts(const double* p, int64_t n, double scale, double* hi, double* lo);

/* Seeded input mode (no reference counterpart; the suite only has the
 * formula initialisers): SplitMix64, see include/rajaperf_mi355x.h. */
uint64_t orc_splitmix64_next(uint64_t* state);
void orc_fill_uniform(double* p, int64_t n, uint64_t seed, uint64_t stream_id, double lo, double hi);
void orc_fill_sym_tensor(double* p, int64_t ne, int q3, int sym, uint64_t seed, uint64_t stream_id);

/* 1D bases for the seeded FEM inputs: Lagrange basis on d1d Gauss-Lobatto
 * nodes evaluated at q1d Gauss-Legendre points; B[q + q1d*d], G likewise. */
void orc_fem_basis(int d1d, int q1d, double* B, double* G);

/* ---- kernels: Base_Seq (variant 0) / Base_OpenMP (variant 1) ---------- */
void orc_copy(int variant, double* c, const double* a, int64_t n);
void orc_mul(int variant, double* b, const double* c, double alpha, int64_t n);
void orc_add(int variant, double* c, const double* a, const double* b, int64_t n);
void orc_triad(int variant, double* a, const double* b, const double* c, double alpha, int64_t n);
double orc_dot(int variant, const double* a, const double* b, int64_t n);
double orc_reduce_sum(int variant, const double* x, int64_t n);
void orc_scan(int variant, const double* x, double* y, int64_t n);
void orc_sort(double* x, int64_t n);
void orc_jacobi2d(int variant, double* A, double* B, int64_t N, int64_t tsteps);
void orc_gemm(int variant, double* C, const double* A, const double* B, double alpha, double beta,
              int64_t ni, int64_t nj, int64_t nk);
void orc_mass3dpa(int variant, const double* B, const double* Bt, const double* D, const double* X,
                  double* Y, int64_t NE, int d1d, int q1d);
void orc_diffusion3dpa(int variant, const double* Basis, const double* dBasis, const double* D,
                       const double* X, double* Y, int64_t NE, int d1d, int q1d, int symmetric);

/* ---- suite-level execute (KernelBase::execute, src/common/KernelBase.cpp:359-377) */
typedef struct orc_params {
  int32_t input_mode;   /* 0 = reference formulas, 1 = seeded               */
  int32_t npasses;      /* checksums accumulate over passes                 */
  uint64_t seed;
  int64_t size;         /* direct target problem size; <= 0 = use sizefact  */
  double sizefact;
  double repfact;
  int64_t checkrun_reps;/* > 0: run exactly this many reps (--checkrun)     */
  int64_t tsteps;       /* JACOBI_2D time steps; <= 0 = suite default 40    */
  int64_t nk;           /* GEMM nk; <= 0 = suite default 1.2*ni             */
  int32_t d1d, q1d;     /* FEM orders; <= 0 = suite defaults                */
  int32_t nthreads;     /* Base_OpenMP threads; <= 0 = OpenMP default       */
  int32_t pad_;
  double alpha, beta;   /* GEMM scalars when has_alpha / has_beta are set   */
  int32_t has_alpha, has_beta;
} orc_params;

typedef struct orc_result {
  long double checksum;
  double checksum_hi;    /* checksum == checksum_hi + checksum_lo exactly */
  double checksum_lo;
  double time_total_s;  /* sum over passes of the timed rep loops           */
  int64_t reps;
  int64_t problem_size;
  int64_t bytes_per_rep;
  int64_t flops_per_rep;
  double iweighted;     /* i-weighted sums of the checksummed arrays, summed
                           over arrays and passes                          */
} orc_result;

/* kernel: "Stream_TRIAD", ...; variant: "Base_Seq" or "Base_OpenMP".
 * Returns 0, or -1 for an unknown kernel/variant pair. */
int orc_run(const char* kernel, const char* variant, const orc_params* p, orc_result* out);

#ifdef __cplusplus
}
#endif
#endif
