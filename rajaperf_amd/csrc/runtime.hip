// Runtime support entry points: device selection, memory, streams, events
// and the device-side input generators.
//
// Counterparts of the suite's HIP allocation/copy helpers
// (src/common/HipDataUtils.hpp:141-255) and of DataUtils' initialisers
// (src/common/DataUtils.cpp:477-595).
#include <cstring>
#include <mutex>

#include "rpx_common.hpp"

using namespace rpx;

// ---- asynchronous table uploads ---------------------------------------------

namespace {

struct StageSlot {
  void* host = nullptr;  // pinned staging memory
  size_t cap = 0;
  hipEvent_t done = nullptr;  // recorded after the copies that read `host`
  int device = -1;            // device of `done`
  bool pending = false;
};
constexpr int kStageSlots = 4;
std::mutex g_stage_mu;
StageSlot g_stage[kStageSlots];
int g_stage_next = 0;

inline size_t align16(size_t b) { return (b + 15) & ~size_t(15); }

}  // namespace

namespace rpx {

int upload_async(const HostChunk* chunks, int n, hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  RPX_TRY(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone) return RPX_ERR_UNSUPPORTED;
  size_t total = 0;
  for (int i = 0; i < n; ++i) total += align16(chunks[i].bytes);
  if (total == 0) return RPX_SUCCESS;
  int dev = 0;
  RPX_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_stage_mu);
  StageSlot& slot = g_stage[g_stage_next];
  g_stage_next = (g_stage_next + 1) % kStageSlots;
  if (slot.pending) {  // the slot's previous copies, long done in practice
    RPX_TRY(hipEventSynchronize(slot.done));
    slot.pending = false;
  }
  if (slot.done && slot.device != dev) {
    RPX_TRY(hipEventDestroy(slot.done));
    slot.done = nullptr;
  }
  if (!slot.done) {
    RPX_TRY(hipEventCreateWithFlags(&slot.done, hipEventDisableTiming));
    slot.device = dev;
  }
  if (slot.cap < total) {
    // At least 64 KiB and at least twice the previous size, so that a run of
    // calls with growing tables pins new memory only a few times.
    size_t want = slot.cap * 2 > total ? slot.cap * 2 : total;
    if (want < (size_t(64) << 10)) want = size_t(64) << 10;
    if (slot.host) RPX_TRY(hipHostFree(slot.host));
    slot.host = nullptr;
    slot.cap = 0;
    RPX_TRY(hipHostMalloc(&slot.host, want, hipHostMallocDefault));
    slot.cap = want;
  }
  char* p = static_cast<char*>(slot.host);
  size_t o = 0;
  hipError_t e = hipSuccess;
  for (int i = 0; i < n && e == hipSuccess; ++i) {
    if (chunks[i].bytes == 0) continue;
    std::memcpy(p + o, chunks[i].src, chunks[i].bytes);
    e = hipMemcpyAsync(chunks[i].dst, p + o, chunks[i].bytes, hipMemcpyHostToDevice, s);
    o += align16(chunks[i].bytes);
  }
  if (e == hipSuccess) e = hipEventRecord(slot.done, s);
  if (e != hipSuccess) {
    // Copies already queued may still read the slot: let them finish
    // before it can be reused.
    (void)hipStreamSynchronize(s);
    return hip_status(e);
  }
  slot.pending = true;
  return RPX_SUCCESS;
}

}  // namespace rpx

namespace {

// x[i] = lo + (hi-lo)*u_i, evaluated as two separately rounded operations
// (no contraction) so the host restatement reproduces it exactly.
__global__ __launch_bounds__(256) void fill_uniform_kernel(double* __restrict__ x, int64_t n,
                                                           uint64_t state0, double lo, double w) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double u = seeded_unit(state0, uint64_t(i));
    const double t = w * u;
    x[i] = lo + t;
  }
}

__global__ __launch_bounds__(256) void fill_const_kernel(double* __restrict__ x, int64_t n, double v) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = v;
}

// factor*(i + 1.1)/(i + 1.12345): IEEE double multiply and correctly rounded
// divide, in the suite's association ((factor*(i+1.1)) / (i+1.12345)).
__global__ __launch_bounds__(256) void fill_formula_kernel(double* __restrict__ x, int64_t n,
                                                           double factor) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double di = static_cast<double>(i);
    const double num = factor * (di + 1.1);
    x[i] = num / (di + 1.12345);
  }
}

__global__ __launch_bounds__(256) void fill_grid_kernel(double* __restrict__ x, int64_t N, uint64_t state0,
                                                        double lo, double w, double boundary) {
  const int64_t n = N * N;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += stride) {
    const int64_t i = k / N, j = k - i * N;
    if (i == 0 || j == 0 || i == N - 1 || j == N - 1) {
      x[k] = boundary;
    } else {
      const double u = seeded_unit(state0, uint64_t(k));
      const double t = w * u;
      x[k] = lo + t;
    }
  }
}

// [e][s][q3] tensor field; component s of 6 (11,12,13,22,23,33) or 9
// (row-major 3x3): diagonal in [0.5,2], off-diagonal in [-0.25,0.25].
__global__ __launch_bounds__(256) void fill_sym_tensor_kernel(double* __restrict__ x, int64_t n, int q3, int sym,
                                                              uint64_t state0) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double u = seeded_unit(state0, uint64_t(i));
    const int s = int((i / q3) % sym);
    const bool diag = (sym == 6) ? (s == 0 || s == 3 || s == 5) : (s == 0 || s == 4 || s == 8);
    const double lo = diag ? 0.5 : -0.25;
    const double w = diag ? 1.5 : 0.5;
    const double t = w * u;
    x[i] = lo + t;
  }
}

unsigned fill_grid(int64_t n) {
  int64_t g = ceil_div(n, 256);
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

}  // namespace

extern "C" {

int rpx_abi_version(void) { return RPX_ABI_VERSION; }

const char* rpx_status_string(int status) {
  switch (status) {
    case RPX_SUCCESS: return "success";
    case RPX_ERR_INVALID_ARGUMENT: return "invalid argument";
    case RPX_ERR_SCRATCH_TOO_SMALL: return "scratch buffer too small";
    case RPX_ERR_UNSUPPORTED: return "unsupported shape";
    case RPX_ERR_NO_DEVICE: return "no gfx950 device";
    case RPX_ERR_WAIT_EXPIRED: return "bounded device wait expired";
    default: break;
  }
  if (status < 0 && status > -10000) return hipGetErrorString(static_cast<hipError_t>(-status));
  return "unknown status";
}

int rpx_device_count(int* count_host) {
  if (!count_host) return RPX_ERR_INVALID_ARGUMENT;
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *count_host = 0;
    return RPX_ERR_NO_DEVICE;
  }
  *count_host = c;
  return RPX_SUCCESS;
}

int rpx_init(int device, char* arch_out, size_t arch_len) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return RPX_ERR_NO_DEVICE;
  RPX_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  RPX_TRY(hipGetDeviceProperties(&prop, device));
  if (arch_out && arch_len > 0) {
    std::strncpy(arch_out, prop.gcnArchName, arch_len - 1);
    arch_out[arch_len - 1] = '\0';
  }
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return RPX_ERR_NO_DEVICE;
  return RPX_SUCCESS;
}

int rpx_device_malloc(void** ptr_host, size_t bytes) {
  if (!ptr_host) return RPX_ERR_INVALID_ARGUMENT;
  *ptr_host = nullptr;
  if (bytes == 0) return RPX_SUCCESS;
  return hip_status(hipMalloc(ptr_host, bytes));
}
int rpx_device_free(void* ptr) { return ptr ? hip_status(hipFree(ptr)) : RPX_SUCCESS; }
int rpx_host_malloc_pinned(void** ptr_host, size_t bytes) {
  if (!ptr_host) return RPX_ERR_INVALID_ARGUMENT;
  *ptr_host = nullptr;
  if (bytes == 0) return RPX_SUCCESS;
  return hip_status(hipHostMalloc(ptr_host, bytes, hipHostMallocDefault));
}
int rpx_host_free_pinned(void* ptr) { return ptr ? hip_status(hipHostFree(ptr)) : RPX_SUCCESS; }
int rpx_memcpy_h2d(void* dst, const void* src_host, size_t bytes) {
  if (bytes == 0) return RPX_SUCCESS;
  if (!dst || !src_host) return RPX_ERR_INVALID_ARGUMENT;
  return hip_status(hipMemcpy(dst, src_host, bytes, hipMemcpyHostToDevice));
}
int rpx_memcpy_d2h(void* dst_host, const void* src, size_t bytes) {
  if (bytes == 0) return RPX_SUCCESS;
  if (!dst_host || !src) return RPX_ERR_INVALID_ARGUMENT;
  return hip_status(hipMemcpy(dst_host, src, bytes, hipMemcpyDeviceToHost));
}
int rpx_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return RPX_SUCCESS;
  if (!dst || !src) return RPX_ERR_INVALID_ARGUMENT;
  return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
}
int rpx_stream_create(void** stream_host) {
  if (!stream_host) return RPX_ERR_INVALID_ARGUMENT;
  hipStream_t s = nullptr;
  RPX_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream_host = s;
  return RPX_SUCCESS;
}
int rpx_stream_destroy(void* stream) { return stream ? hip_status(hipStreamDestroy(as_stream(stream))) : RPX_SUCCESS; }
int rpx_stream_synchronize(void* stream) { return hip_status(hipStreamSynchronize(as_stream(stream))); }
int rpx_device_synchronize(void) { return hip_status(hipDeviceSynchronize()); }
// Stream capture into a graph.  The capture mode is thread-local, so other
// host threads may keep issuing work while one thread captures.
int rpx_graph_begin_capture(void* stream) {
  if (!stream) return RPX_ERR_INVALID_ARGUMENT;  // the null stream cannot be captured
  return hip_status(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
}
int rpx_graph_end_capture(void* stream, void** graph_exec_host) {
  if (!stream || !graph_exec_host) return RPX_ERR_INVALID_ARGUMENT;
  hipGraph_t g = nullptr;
  RPX_TRY(hipStreamEndCapture(as_stream(stream), &g));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);  // the executable graph keeps what it needs
  if (e != hipSuccess) return hip_status(e);
  *graph_exec_host = exec;
  return RPX_SUCCESS;
}
int rpx_graph_launch(void* graph_exec, void* stream) {
  if (!graph_exec) return RPX_ERR_INVALID_ARGUMENT;
  return hip_status(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph_exec), as_stream(stream)));
}
int rpx_graph_destroy(void* graph_exec) {
  return graph_exec ? hip_status(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph_exec))) : RPX_SUCCESS;
}

int rpx_event_create(void** event_host) {
  if (!event_host) return RPX_ERR_INVALID_ARGUMENT;
  hipEvent_t e = nullptr;
  RPX_TRY(hipEventCreate(&e));
  *event_host = e;
  return RPX_SUCCESS;
}
int rpx_event_destroy(void* event) {
  return event ? hip_status(hipEventDestroy(reinterpret_cast<hipEvent_t>(event))) : RPX_SUCCESS;
}
int rpx_event_record(void* event, void* stream) {
  if (!event) return RPX_ERR_INVALID_ARGUMENT;
  return hip_status(hipEventRecord(reinterpret_cast<hipEvent_t>(event), as_stream(stream)));
}
int rpx_event_elapsed_ms(float* ms_host, void* start, void* stop) {
  if (!ms_host || !start || !stop) return RPX_ERR_INVALID_ARGUMENT;
  RPX_TRY(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)));
  return hip_status(hipEventElapsedTime(ms_host, reinterpret_cast<hipEvent_t>(start),
                                        reinterpret_cast<hipEvent_t>(stop)));
}

int rpx_fill_uniform_f64(double* x, int64_t n, uint64_t seed, uint64_t stream_id, double lo, double hi,
                         void* stream) {
  if (n < 0 || (n > 0 && !x)) return RPX_ERR_INVALID_ARGUMENT;
  if (n == 0) return RPX_SUCCESS;
  const double w = hi - lo;
  fill_uniform_kernel<<<fill_grid(n), 256, 0, as_stream(stream)>>>(x, n, seeded_state0(seed, stream_id), lo, w);
  return launch_status();
}

int rpx_fill_grid_f64(double* x, int64_t N, uint64_t seed, uint64_t stream_id, double lo, double hi,
                      double boundary, void* stream) {
  if (N < 0 || (N > 0 && !x)) return RPX_ERR_INVALID_ARGUMENT;
  if (N == 0) return RPX_SUCCESS;
  fill_grid_kernel<<<fill_grid(N * N), 256, 0, as_stream(stream)>>>(x, N, seeded_state0(seed, stream_id), lo,
                                                                     hi - lo, boundary);
  return launch_status();
}

int rpx_fill_sym_tensor_f64(double* x, int64_t ne, int q3, int sym, uint64_t seed, uint64_t stream_id,
                            void* stream) {
  if (ne < 0 || q3 <= 0 || (sym != 6 && sym != 9) || (ne > 0 && !x)) return RPX_ERR_INVALID_ARGUMENT;
  const int64_t n = ne * int64_t(sym) * q3;
  if (n == 0) return RPX_SUCCESS;
  fill_sym_tensor_kernel<<<fill_grid(n), 256, 0, as_stream(stream)>>>(x, n, q3, sym,
                                                                      seeded_state0(seed, stream_id));
  return launch_status();
}

int rpx_fill_const_f64(double* x, int64_t n, double val, void* stream) {
  if (n < 0 || (n > 0 && !x)) return RPX_ERR_INVALID_ARGUMENT;
  if (n == 0) return RPX_SUCCESS;
  fill_const_kernel<<<fill_grid(n), 256, 0, as_stream(stream)>>>(x, n, val);
  return launch_status();
}

int rpx_fill_suite_formula_f64(double* x, int64_t n, double factor, void* stream) {
  if (n < 0 || (n > 0 && !x)) return RPX_ERR_INVALID_ARGUMENT;
  if (n == 0) return RPX_SUCCESS;
  fill_formula_kernel<<<fill_grid(n), 256, 0, as_stream(stream)>>>(x, n, factor);
  return launch_status();
}

}  // extern "C"
