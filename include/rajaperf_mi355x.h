/*
 * rajaperf_mi355x.h -- C ABI of the MI355X-native execution variant.
 *
 * One entry point per kernel of the RAJA Performance Suite subset this
 * library accelerates.  Each one replaces the device work done inside the
 * timed rep loop of a `run<Backend>Variant` method of the suite's
 * KernelBase plugin API (src/common/KernelBase.hpp:451-487): the suite's
 * setUp/updateChecksum/tearDown stay on the caller's side.
 *
 * Conventions (all entry points):
 *   - every array pointer is a DEVICE pointer owned by the caller, except
 *     where a parameter name ends in `_host`;
 *   - no allocation and no host synchronisation inside a compute entry
 *     point; work is queued on `stream` (an hipStream_t passed as void*,
 *     NULL = the null stream) and completes asynchronously;
 *   - the return value is RPX_SUCCESS (0), a negative RPX_ERR_* code for
 *     argument errors, or -(hipError_t) for a HIP runtime failure;
 *   - calls on different streams with disjoint buffers may run concurrently;
 *     scratch buffers must not be shared by calls in flight at the same time;
 *   - `tune` may be NULL: the library then uses its own tuned defaults.
 *
 * Floating-point policy: element-wise kernels (COPY/MUL/ADD/TRIAD),
 * JACOBI_2D and the FP64 FEM operators evaluate every expression in the
 * association of the suite's Base_Seq loop bodies with no fused
 * multiply-add contraction, so their results equal the Base_Seq result bit
 * for bit.  Reductions, scans and GEMM change the summation order and are
 * compared within a stated tolerance.  SORT is value-exact.
 */
#ifndef RAJAPERF_MI355X_H
#define RAJAPERF_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RPX_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
#define RPX_SUCCESS                  0
#define RPX_ERR_INVALID_ARGUMENT  -1001 /* null pointer, negative size, ...  */
#define RPX_ERR_SCRATCH_TOO_SMALL -1002 /* scratch_bytes below the query     */
#define RPX_ERR_UNSUPPORTED       -1003 /* shape outside the compiled range  */
#define RPX_ERR_NO_DEVICE         -1004 /* no gfx950 device visible          */
#define RPX_ERR_WAIT_EXPIRED      -1005 /* a bounded device wait gave up     */

/* ---- tuning ------------------------------------------------------------ */
/* Replaces the suite's compile-time block-size / items-per-thread lists
 * (src/common/GPUUtils.hpp:146-179, rajaperf_config.hpp.in:121-128).
 * Zero fields mean "library default". */
typedef struct rpx_tuning {
  int32_t block_size;        /* work-items per workgroup                    */
  int32_t items_per_thread;  /* elements per work-item per grid step        */
  int32_t grid_size;         /* workgroups; 0 = derived from the size       */
  int32_t flags;             /* RPX_TUNE_* bits                             */
  int32_t depth;             /* temporal blocking: JACOBI_2D time steps per
                                fused launch (1..4)                          */
} rpx_tuning;

#define RPX_TUNE_CACHED            0x1 /* stream/DOT/REDUCE_SUM/SCAN: regular cached
                                           loads and stores instead of the default
                                           nontemporal (streaming) ones */
#define RPX_TUNE_NAIVE             0x2 /* plain FP64 comparators in the suite's Base_HIP
                                           shape: GEMM one work-item per C element,
                                           MASS3DPA one element per Q x Q workgroup */
#define RPX_TUNE_JACOBI_SKIP_B     0x4 /* rpx_jacobi2d_fused_launch_f64: leave B's
                                           interior unwritten, for a launch whose B a
                                           later launch rewrites before it is read */
#define RPX_TUNE_MASS_SINGLE_LAUNCH 0x8 /* rpx_mass3dpa_batched_run: one persistent
                                           launch for all orders instead of one per
                                           order present */
#define RPX_TUNE_GEMM_BETA         0x10 /* rpx_gemm_f64, rpx_gemm_mixed_tuned_f64: the
                                           textbook C = alpha*A*B + beta*C (each C
                                           read once before it is written) instead of
                                           the suite's C = alpha*A*B */

/* ---- library / runtime support ---------------------------------------- *
 * Counterparts of the suite's HIP data utilities
 * (src/common/HipDataUtils.hpp:141-255) and of the device synchronisation
 * in KernelBase::synchronize (src/common/KernelBase.hpp:271-294). */
int         rpx_abi_version(void);
const char* rpx_status_string(int status);
/* Selects device `device` and checks that it is a gfx950 part.
 * `arch_out` (optional, host) receives the gcnArchName. */
int rpx_init(int device, char* arch_out, size_t arch_len);
int rpx_device_count(int* count_host);
int rpx_device_malloc(void** ptr_host, size_t bytes);
int rpx_device_free(void* ptr);
int rpx_host_malloc_pinned(void** ptr_host, size_t bytes);
int rpx_host_free_pinned(void* ptr);
int rpx_memcpy_h2d(void* dst, const void* src_host, size_t bytes);
int rpx_memcpy_d2h(void* dst_host, const void* src, size_t bytes);
int rpx_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int rpx_stream_create(void** stream_host);
int rpx_stream_destroy(void* stream);
int rpx_stream_synchronize(void* stream);
int rpx_device_synchronize(void);
/* Graph capture of a sequence of calls (the suite has none; the rep loops of
 * Executor::runSuite, Executor.cpp:645-670, launch kernel by kernel).
 * Between begin and end, calls on `stream` (a created stream, not the null
 * stream) are recorded instead of run; rpx_graph_launch replays them.  Entry
 * points that upload host tables (the segmented SCAN, the ragged SORT and
 * rpx_mass3dpa_batched_plan) or read results back (the scratch status
 * queries) cannot be captured and return an error under capture; every
 * other kernel entry point can. */
int rpx_graph_begin_capture(void* stream);
int rpx_graph_end_capture(void* stream, void** graph_exec_host);
int rpx_graph_launch(void* graph_exec, void* stream);
int rpx_graph_destroy(void* graph_exec);
int rpx_event_create(void** event_host);
int rpx_event_destroy(void* event);
int rpx_event_record(void* event, void* stream);
/* Milliseconds between two recorded events (synchronises on `stop`). */
int rpx_event_elapsed_ms(float* ms_host, void* start, void* stop);

/* ---- input generation -------------------------------------------------- *
 * Seeded input mode: x[i] = lo + (hi-lo)*u_i with u_i = (s_i >> 11)*2^-53
 * and s_i the i-th output of SplitMix64 started from
 * state0 = mix64(seed ^ (stream_id * 0xD1B54A32D192ED03)).  Evaluated with
 * the same rounding on host and device.  Replaces the host-side formula
 * initialisers (src/common/DataUtils.cpp:477-595) for large inputs so they
 * need no host-to-device copy. */
int rpx_fill_uniform_f64(double* x, int64_t n, uint64_t seed,
                         uint64_t stream_id, double lo, double hi,
                         void* stream);
/* N x N grid: interior points seeded as rpx_fill_uniform_f64 over the flat
 * index j + i*N, boundary rows/columns set to `boundary`. */
int rpx_fill_grid_f64(double* x, int64_t N, uint64_t seed, uint64_t stream_id,
                      double lo, double hi, double boundary, void* stream);
/* Symmetric tensor field [e][s][q3] (s < sym, sym = 6 or 9): diagonal
 * components uniform in [0.5,2], off-diagonal in [-0.25,0.25], from the same
 * SplitMix64 sequence over the flat index. */
int rpx_fill_sym_tensor_f64(double* x, int64_t ne, int q3, int sym,
                            uint64_t seed, uint64_t stream_id, void* stream);
/* x[i] = val (src/common/DataUtils.cpp:518-528 initDataConst). */
int rpx_fill_const_f64(double* x, int64_t n, double val, void* stream);
/* Suite formula x[i] = factor*(i+1.1)/(i+1.12345)
 * (src/common/DataUtils.cpp:504-513 initData(Real_ptr)). */
int rpx_fill_suite_formula_f64(double* x, int64_t n, double factor,
                               void* stream);

/* ---- stream group ------------------------------------------------------ */
/* c[i] = a[i]             src/stream/COPY.hpp:24-25, COPY-Hip.cpp          */
int rpx_copy_f64(double* c, const double* a, int64_t n,
                 const rpx_tuning* tune, void* stream);
/* b[i] = alpha*c[i]       src/stream/MUL.hpp:25-26, MUL-Hip.cpp            */
int rpx_mul_f64(double* b, const double* c, double alpha, int64_t n,
                const rpx_tuning* tune, void* stream);
/* c[i] = a[i] + b[i]      src/stream/ADD.hpp:25-26, ADD-Hip.cpp            */
int rpx_add_f64(double* c, const double* a, const double* b, int64_t n,
                const rpx_tuning* tune, void* stream);
/* a[i] = b[i] + alpha*c[i]  src/stream/TRIAD.hpp:26-27, TRIAD-Hip.cpp:24-61 */
int rpx_triad_f64(double* a, const double* b, const double* c, double alpha,
                  int64_t n, const rpx_tuning* tune, void* stream);

/* ---- reductions -------------------------------------------------------- *
 * Deterministic: each workgroup reduces one fixed chunk of the input
 * (tune->items_per_thread doubles per work-item, a power of two >= 2;
 * tune->grid_size, when set, caps the workgroup count by lengthening the
 * chunks) and writes one FP64 partial into `scratch`; partials are combined
 * in index order through two ticket levels (groups of 64 workgroups, then
 * the groups).  Results are identical from run to run for a fixed
 * (n, tune).  x (and b for DOT) must be 16-byte aligned.  `scratch` must be
 * zeroed once before its first use (rpx_reduce_scratch_init); every
 * complete call leaves it zero again, so one scratch serves later calls of
 * any n and tuning that fit it. */
size_t rpx_reduce_scratch_bytes(int64_t n, const rpx_tuning* tune);
int rpx_reduce_scratch_init(void* scratch, size_t scratch_bytes,
                            void* stream);
/* dot = sum a[i]*b[i]; *result = dot, or *result += dot when `accumulate`
 * (src/stream/DOT.hpp:24-25; the suite's per-rep `m_dot += dot`,
 * DOT-Seq.cpp:39-45, DOT-Hip.cpp:28-94).  `result` is a device scalar. */
int rpx_dot_f64(const double* a, const double* b, int64_t n, double* result,
                int accumulate, void* scratch, size_t scratch_bytes,
                const rpx_tuning* tune, void* stream);
/* Allocation-free form of rpx_dot_f64, with the same kernels and results
 * (bit-identical for a given (n, tune)).  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The scratch is a library-owned device array of
 * rpx_dot_noalloc_max_scratch_bytes() bytes (1 MiB) per GPU, held until the
 * process ends; shapes whose rpx_reduce_scratch_bytes exceeds it (beyond
 * about 1.2e9 doubles at the default tuning) return
 * RPX_ERR_SCRATCH_TOO_SMALL.  Every call on a GPU shares that GPU's array,
 * so calls that overlap on one GPU (other streams, other host threads)
 * race; calls ordered on one stream do not.  The first call on a device
 * looks the array up and zeroes it on `stream`. */
size_t rpx_dot_noalloc_max_scratch_bytes(void);
int rpx_dot_f64_noalloc(const double* a, const double* b, int64_t n,
                        double* result, int accumulate,
                        const rpx_tuning* tune, void* stream);
/* *result = sum x[i]  (src/algorithm/REDUCE_SUM.hpp:31-32,
 * REDUCE_SUM-Seq.cpp:34-51, REDUCE_SUM-Hip.cpp:36-198). */
int rpx_reduce_sum_f64(const double* x, int64_t n, double* result,
                       void* scratch, size_t scratch_bytes,
                       const rpx_tuning* tune, void* stream);

/* ---- scan -------------------------------------------------------------- *
 * Exclusive prefix sum y[i] = x[0]+...+x[i-1], y[0] = 0
 * (src/algorithm/SCAN.hpp:27-32, SCAN-Seq.cpp:31-45, SCAN-Hip.cpp:39-326).
 * Single pass with decoupled look-back over tiles of 256 x
 * tune->items_per_thread doubles (16, 32 or 64; default 64); a tile only
 * waits on tiles with lower workgroup ids, which are dispatched before it.
 * The scratch size does not depend on the tuning.  x and y may alias. */
size_t rpx_scan_scratch_bytes(int64_t n, const rpx_tuning* tune);
/* Zeroes a scan scratch buffer; required once before its first use.  Each
 * call then leaves the buffer ready for the next (epoch-tagged flags).  A
 * scratch may be shared with the segmented form below (sized for the larger
 * of the two): the segmented form's flags carry a tag the single-array form
 * never accepts. */
int rpx_scan_scratch_init(void* scratch, size_t scratch_bytes, void* stream);
/* Synchronously reads the scratch's error word: RPX_ERR_WAIT_EXPIRED if a
 * look-back wait of an earlier call gave up, else RPX_SUCCESS. */
int rpx_scan_scratch_status(const void* scratch);
int rpx_scan_exclusive_f64(const double* x, double* y, int64_t n,
                           void* scratch, size_t scratch_bytes,
                           const rpx_tuning* tune, void* stream);
/* Inclusive form, y[i] = x[0] + ... + x[i] (within a wavefront, the
 * exclusive prefix of element i+1), an extra output beside the suite's exclusive SCAN
 * (src/algorithm/SCAN.hpp:27-32 defines only the exclusive one).  Same
 * scratch, tunings and conventions as rpx_scan_exclusive_f64. */
int rpx_scan_inclusive_f64(const double* x, double* y, int64_t n,
                           void* scratch, size_t scratch_bytes,
                           const rpx_tuning* tune, void* stream);
/* Segmented form: segment s covers [offsets[s], offsets[s+1]) and is scanned
 * independently; `offsets_host` is a host array of nseg+1 ascending values.
 * All segments run in one launch (tiles of 16384 doubles counted from the
 * segment's start rounded down to an even element, so a segment's first
 * tile may be one element short; the look-back of a tile stops at its
 * segment's first tile).  The call copies the offsets and each segment's
 * first tile into `scratch` on `stream` in one copy from the library's
 * pinned staging memory, without a host synchronisation (a staging slot is
 * reused once its copy has completed); the scratch needs no
 * initialisation. */
size_t rpx_scan_segmented_scratch_bytes(const int64_t* offsets_host,
                                        int64_t nseg);
int rpx_scan_exclusive_segmented_f64(const double* x, double* y,
                                     const int64_t* offsets_host,
                                     int64_t nseg, void* scratch,
                                     size_t scratch_bytes, void* stream);

/* ---- sort -------------------------------------------------------------- *
 * Ascending sort of FP64 keys, value-identical to std::sort
 * (src/algorithm/SORT.hpp:18-25, SORT-Seq.cpp:31-42, SORT-Hip.cpp:35-43).
 * -0.0 orders before +0.0; NaNs order after +inf (sign bit clear) or before
 * -inf (sign bit set).  Segment s covers [offsets[s], offsets[s+1]) with
 * `offsets_host` a host array; each segment is sorted independently, in
 * place.  The suite's layout (reps segments of length n) is
 * offsets[s] = s*n.  Other layouts copy a tile table into `scratch` on
 * `stream` from pinned staging memory, without a host synchronisation. */
size_t rpx_sort_segmented_scratch_bytes(const int64_t* offsets_host,
                                        int64_t nseg);
int rpx_sort_segmented_f64(double* x, const int64_t* offsets_host,
                           int64_t nseg, void* scratch, size_t scratch_bytes,
                           void* stream);

/* ---- polybench --------------------------------------------------------- */
/* One Jacobi sweep dst[i][j] = 0.2*(src[c]+src[w]+src[e]+src[s]+src[n]) on
 * the interior 1..N-2 of an N x N row-major grid; the boundary of dst is not
 * written (src/polybench/POLYBENCH_JACOBI_2D.hpp:41-45). */
int rpx_jacobi2d_sweep_f64(double* dst, const double* src, int64_t N,
                           const rpx_tuning* tune, void* stream);
/* `tsteps` time steps of the two-sweep A->B->A update
 * (POLYBENCH_JACOBI_2D-Seq.cpp:30-54, -Hip.cpp:27-114). */
int rpx_jacobi2d_f64(double* A, double* B, int64_t N, int64_t tsteps,
                     const rpx_tuning* tune, void* stream);
/* Same result with temporal blocking: each launch runs `depth` whole time
 * steps (2*depth sweeps, tune->depth: 1..4, default 3) on register-resident
 * tiles or chunks,
 * reading A once and writing only the final B and the new A, the new A going
 * to `scratch` (an N x N grid) and back, so no launch overwrites values a
 * neighbouring tile still reads.  Ends with the result in A and B, bit-equal
 * to rpx_jacobi2d_f64.  Falls back to the two-sweep path for odd N,
 * N > 2^22 or pointers that are not 16-byte aligned. */
size_t rpx_jacobi2d_fused_scratch_bytes(int64_t N);
/* The building block of rpx_jacobi2d_fused_f64: one launch advancing `depth`
 * time steps (1..4) from Ain into Aout, writing the final B (unless
 * tune->flags has RPX_TUNE_JACOBI_SKIP_B: in a run of launches only the last
 * one needs to store B).  Requires even N, 16-byte aligned grids,
 * Aout != Ain, and Aout's boundary equal to Ain's (copy it once with
 * rpx_jacobi2d_copy_boundary_f64; the launch may rewrite boundary columns
 * with those same values); callers ping-pong Ain and Aout between launches.
 * tune->items_per_thread: output rows per wavefront, as a tile of 4, 8,
 * 12 or 16 rows (32 at depth 1), or as a chunk of 48..4096 rows that the
 * wavefront walks down in a loop (streaming).  `depth` 4 runs streaming
 * chunks only.  Default rows: at depth 4, chunks of 144; otherwise chunks of
 * 99 rows for a launch that skips B on a grid of N >= 3072, tiles of 16
 * rows for one that stores B and on smaller grids.
 * Returns RPX_ERR_UNSUPPORTED for odd N, N > 2^22 or an unsupported
 * rows/depth pair. */
int rpx_jacobi2d_fused_launch_f64(double* Aout, double* B, const double* Ain,
                                  int64_t N, int depth,
                                  const rpx_tuning* tune, void* stream);
/* dst's boundary rows and columns := src's. */
int rpx_jacobi2d_copy_boundary_f64(double* dst, const double* src,
                                   int64_t N, void* stream);
int rpx_jacobi2d_fused_f64(double* A, double* B, int64_t N, int64_t tsteps,
                           void* scratch, size_t scratch_bytes,
                           const rpx_tuning* tune, void* stream);

/* GEMM with the suite's semantics (src/polybench/POLYBENCH_GEMM.hpp:12-50):
 *   C[i][j] = sum_k (alpha*A[i][k])*B[k][j]
 * `beta` is accepted for interface fidelity: the suite scales C by beta and
 * then overwrites it, so beta never reaches the result.  With
 * RPX_TUNE_GEMM_BETA in tune->flags the result is instead the textbook
 *   C[i][j] = sum_k (alpha*A[i][k])*B[k][j] + beta*C[i][j]
 * (the sum as above, then beta*C added in FP64).  A is ni x nk, B is
 * nk x nj, C is ni x nj, all row-major.  FP64 on the MFMA f64 units;
 * RPX_TUNE_NAIVE selects the one-thread-per-element comparator. */
int rpx_gemm_f64(double* C, const double* A, const double* B, double alpha,
                 double beta, int64_t ni, int64_t nj, int64_t nk,
                 const rpx_tuning* tune, void* stream);

/* Mixed-precision GEMM, same semantics, emulated on int8 matrix cores with
 * the modular (Chinese remainder) scheme: rows of A and columns of B are
 * scaled by powers of two 2^e_i > max_k |A[i][k]| >= 2^(e_i-1) (likewise
 * 2^f_j for B) and rounded to b-bit integers; their exact integer product
 * is assembled from one int8 GEMM per modulus (`nmoduli` of the 16 largest
 * pairwise coprime integers <= 256) and converted to FP64.  b is the
 * largest value with 2 nk 4^b below the moduli's product (at most 46;
 * rpx_gemm_mixed_bits returns it), and the only error besides the FP64
 * rounding of C is that rounding of A and B:
 *   |C[i][j] - alpha (AB)[i][j]| <= |alpha| 2^(e_i+f_j-b) nk (1 + 2^-b),
 * at most 4 2^-b relative to |alpha| max_k|A[i][k]| max_k|B[k][j]| nk, so
 * it grows with the dynamic range within row i and column j.
 * nmoduli in [1, 16]; 0 picks the smallest count with b >= 42 (13 moduli,
 * b = 44 at nk = 4096).  nk is limited to 131071 (RPX_ERR_UNSUPPORTED
 * above), and to sizes where b >= 1.  A's conversion runs on a
 * library-owned auxiliary stream (created with two events on the first call
 * on a device) forked from and joined back to `stream` with events, so the
 * call stays ordered on `stream` and can be captured in a HIP graph. */
size_t rpx_gemm_mixed_scratch_bytes(int64_t ni, int64_t nj, int64_t nk,
                                    int nmoduli);
int rpx_gemm_mixed_bits(int64_t nk, int nmoduli);
int rpx_gemm_mixed_f64(double* C, const double* A, const double* B,
                       double alpha, double beta, int64_t ni, int64_t nj,
                       int64_t nk, int nmoduli, void* scratch,
                       size_t scratch_bytes, void* stream);
/* rpx_gemm_mixed_f64 with a tuning (src/polybench/POLYBENCH_GEMM.hpp:12-50
 * as above): RPX_TUNE_GEMM_BETA selects the textbook C = alpha*A*B + beta*C,
 * beta*C added in FP64 to the converted product; tune may be NULL. */
int rpx_gemm_mixed_tuned_f64(double* C, const double* A, const double* B,
                             double alpha, double beta, int64_t ni,
                             int64_t nj, int64_t nk, int nmoduli,
                             const rpx_tuning* tune, void* scratch,
                             size_t scratch_bytes, void* stream);

/* ---- apps (partial-assembly finite elements) -------------------------- */
/* Mass operator action Y_e += Bt'^3 (D_e .* (B^3 X_e)) for NE hex elements
 * (src/apps/MASS3DPA.hpp:158-339, MASS3DPA-Seq.cpp:29-94).
 * B is read as B[q + q1d*d]; Bt is read LITERALLY as Bt[q + d1d*d], the
 * suite's indexing (MASS3DPA.hpp:175).  X, Y: [e][dz][dy][dx];
 * D: [e][qz][qy][qx].  Supported: 1 <= d1d <= 8, q1d in {d1d, d1d+1,
 * d1d+2}, q1d <= 9.  Bit-identical to Base_Seq.  RPX_TUNE_NAIVE selects the
 * comparator in the suite's Base_HIP shape (MASS3DPA-Hip.cpp:25-117): one
 * element per q1d x q1d workgroup. */
int rpx_mass3dpa_f64(const double* B, const double* Bt, const double* D,
                     const double* X, double* Y, int64_t NE, int d1d,
                     int q1d, const rpx_tuning* tune, void* stream);

/* Ragged batches from one chunk table (the suite runs one fixed order and
 * size per kernel object; MASS3DPA.cpp:26-31).  Batch i is one operator
 * application, as rpx_mass3dpa_f64 with its own bases, data, element count
 * and order; orders may differ between batches (1 <= d1d <= 8, q1d =
 * d1d + 1).  Results are bit-identical to calling rpx_mass3dpa_f64 per batch.
 * rpx_mass3dpa_batched_plan uploads the batch and chunk tables into
 * `scratch` and waits for that copy (a host synchronisation of `stream`);
 * rpx_mass3dpa_batched_run then runs a planned scratch any number of times
 * (no host synchronisation; it may be captured in a graph): one launch per
 * order present, with one workgroup per chunk of that order, then one short
 * persistent launch for any chunks those grids missed.  The grids come from
 * chunk counts that the plan call records on the host, keyed by the scratch
 * address (the 64 most recent plans).  The closing launch keeps results
 * right when a record no longer matches the scratch, e.g. a larger plan
 * copied into a reused address.  Without a record (a plan copied from
 * another buffer), or with tune->grid_size, the run is one persistent launch
 * per order (eight launches; an order without batches exits at once).  With
 * RPX_TUNE_MASS_SINGLE_LAUNCH it is one persistent launch for all orders.
 * rpx_mass3dpa_batched_f64 does both calls.  tune->grid_size sets the
 * number of persistent workgroups per launch (default: as many as are
 * resident at once; 3 per CU for the single launch). */
typedef struct rpx_mass_batch {
  const double* B;
  const double* Bt;
  const double* D;
  const double* X;
  double* Y;
  int64_t ne;
  int32_t d1d;
  int32_t q1d;
} rpx_mass_batch;
size_t rpx_mass3dpa_batched_scratch_bytes(const rpx_mass_batch* batches_host,
                                          int nbatch);
int rpx_mass3dpa_batched_plan(const rpx_mass_batch* batches_host, int nbatch,
                              void* scratch, size_t scratch_bytes,
                              void* stream);
int rpx_mass3dpa_batched_run(const void* scratch, const rpx_tuning* tune,
                             void* stream);
int rpx_mass3dpa_batched_f64(const rpx_mass_batch* batches_host, int nbatch,
                             void* scratch, size_t scratch_bytes,
                             const rpx_tuning* tune, void* stream);

/* The same operator with the basis contractions on int8 matrix cores: the
 * forward and backward Kronecker products of the bases (q1d^3 x d1d^3 and
 * d1d^3 x q1d^3) applied as error-free int8 slice products (`nslices`
 * digits of 7 bits, 0 = 7) with int32 accumulation, FP64 recombination, FP64
 * quadrature scaling and FP64 Y +=.  Each output is within about
 * (nslices+3) 2^(-7 nslices) of the FP64 result relative to
 * |Bt'|^3 (|D| .* (|B|^3 |X|)).  No counterpart in the suite (the mixed
 * precision path of MASS3DPA-Hip.cpp's contractions).  Supported:
 * 1 <= d1d <= 4, d1d <= q1d <= min(5, d1d+2).  `scratch` (query
 * rpx_mass3dpa_mixed_scratch_bytes) receives the sliced Kronecker matrices;
 * it needs no initialisation. */
size_t rpx_mass3dpa_mixed_scratch_bytes(int d1d, int q1d, int nslices);
int rpx_mass3dpa_mixed_f64(const double* B, const double* Bt, const double* D,
                           const double* X, double* Y, int64_t NE, int d1d,
                           int q1d, int nslices, void* scratch,
                           size_t scratch_bytes, void* stream);

/* Diffusion operator action with the suite's half-storage basis tables
 * (src/apps/DIFFUSION3DPA.hpp:216-458, DIFFUSION3DPA-Seq.cpp:28-110).
 * Basis/dBasis read as [q + q1d*d]; D: [e][s][qz][qy][qx] with 6 symmetric
 * components (11,12,13,22,23,33) when `symmetric`, 9 otherwise.  The shared
 * B/G table is filled in the Base_Seq write order, so bases that break the
 * half-storage symmetry give the Base_Seq result.  Requires q1d = d1d+1,
 * 1 <= d1d <= 8. */
int rpx_diffusion3dpa_f64(const double* Basis, const double* dBasis,
                          const double* D, const double* X, double* Y,
                          int64_t NE, int d1d, int q1d, int symmetric,
                          const rpx_tuning* tune, void* stream);

/* The same operator with its basis contractions on int8 matrix cores: the
 * forward and backward products with the three Kronecker matrices
 * K_0 = B(x)B(x)G, K_1 = B(x)G(x)B, K_2 = G(x)B(x)B (built from the same
 * effective half-storage tables, so the result is the Base_Seq operator for
 * any bases) run as error-free int8 slice products (`nslices` digits of 7
 * bits, 5..8, 0 = 7) with int32 accumulation and FP64 recombination; the
 * flux O.grad at each point and Y += are FP64.  Rows of K_c and K_c^T, and
 * X and the flux per element, are scaled by powers of two; each product's
 * dropped digits are below (nslices+3) 2^(-7 nslices) times its operands'
 * scales times its length, and at nslices = 7 every output is within 1e-12
 * of the FP64 result relative to the magnitude operator
 * sum_c |K_c|^T (|O| (|K| |X|))_c on the test data (tests/test_gpu_gemm_mixed.py).
 * No counterpart in the suite (the mixed-precision path of
 * DIFFUSION3DPA-Hip.cpp's contractions).  Supported: 1 <= d1d <= 4,
 * q1d = d1d+1.  `scratch` (query rpx_diffusion3dpa_mixed_scratch_bytes)
 * receives the sliced Kronecker matrices; it needs no initialisation. */
size_t rpx_diffusion3dpa_mixed_scratch_bytes(int d1d, int q1d, int nslices);
int rpx_diffusion3dpa_mixed_f64(const double* Basis, const double* dBasis,
                                const double* D, const double* X, double* Y,
                                int64_t NE, int d1d, int q1d, int symmetric,
                                int nslices, void* scratch,
                                size_t scratch_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RAJAPERF_MI355X_H */
