// Library runtime: device selection, error strings, per-stream workspaces,
// memory/stream/event utilities and the seeded synthetic-data generators.
#include "coral_common.hpp"

#include <map>
#include <mutex>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>

namespace coral {

int autotune_load_default();  // autotune.hip

static thread_local char g_last_error[512] = "";
static std::mutex g_mu;
static int g_device = -1;
static int g_num_cus = 0;

void set_last_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
  set_last_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  if (e == hipErrorOutOfMemory) return CORAL_ERR_ALLOC;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return CORAL_ERR_NO_DEVICE;
  return CORAL_ERR_HIP;
}

int ensure_init() {
  if (g_device >= 0) return CORAL_OK;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_device >= 0) return CORAL_OK;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    set_last_error("no HIP device visible (%s)", hipGetErrorString(e));
    return CORAL_ERR_NO_DEVICE;
  }
  int dev = 0;
  CORAL_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  CORAL_HIP(hipGetDeviceProperties(&prop, dev));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_last_error("device %d is %s; libcoralhip is built for gfx950 only", dev, prop.gcnArchName);
    return CORAL_ERR_NO_DEVICE;
  }
  g_num_cus = prop.multiProcessorCount;
  // a shipped tile table that cannot be read leaves the heuristics in place
  (void)autotune_load_default();
  g_device = dev;
  return CORAL_OK;
}

int num_cus() { return g_num_cus > 0 ? g_num_cus : 256; }

hipStream_t resolve(coral_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// A workspace buffer is reference counted: its stream slot holds one reference and
// every instantiated graph whose captured work uses it holds another, so a buffer a
// graph still reads is never freed when an eager call grows the slot or the stream
// is destroyed; the last reference frees it.
struct WsBuf { void *ptr = nullptr; size_t size = 0; int refs = 0; };
struct WsEntry { WsBuf *buf[WS_NSLOTS] = {}; };
static std::map<hipStream_t, WsEntry> g_ws;
static std::map<hipStream_t, std::vector<WsBuf *>> g_capture_refs;  // buffers used by work being captured
static std::map<hipGraphExec_t, std::vector<WsBuf *>> g_graph_refs;  // buffers an instantiated graph reads

static void ws_release(WsBuf *b) {  // caller holds g_mu
  if (--b->refs == 0) {
    (void)hipFree(b->ptr);
    delete b;
  }
}

static bool capturing(hipStream_t stream) {
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(stream, &cap) == hipSuccess && cap != hipStreamCaptureStatusNone;
}

int workspace(hipStream_t stream, int slot, size_t bytes, void **ptr) {
  const bool cap = capturing(stream);
  std::lock_guard<std::mutex> lk(g_mu);
  WsEntry &w = g_ws[stream];
  if (bytes == 0) bytes = 256;
  WsBuf *b = w.buf[slot];
  if (!b || b->size < bytes) {
    if (cap) {
      set_last_error("a call under stream capture needs a larger workspace than the stream holds: make the "
                     "same call once eagerly on this stream before capturing it");
      return CORAL_ERR_UNSUPPORTED;
    }
    if (b) {
      // queued work on this stream may still read the old buffer
      CORAL_HIP(hipStreamSynchronize(stream));
      w.buf[slot] = nullptr;
      ws_release(b);
    }
    size_t grow = bytes + bytes / 8;
    grow = (grow + 4095) & ~size_t(4095);
    void *p = nullptr;
    CORAL_HIP(hipMalloc(&p, grow));
    if (slot == WS_ERROR) {
      // ordered before the stream's kernels that set flags in it (a plain hipMemset
      // goes to the null stream, which a non-blocking stream does not wait for); never
      // under capture: growth is refused there
      hipError_t e = hipMemsetAsync(p, 0, grow, stream);
      if (e != hipSuccess) {
        (void)hipFree(p);
        return hip_fail(e, "hipMemsetAsync", __FILE__, __LINE__);
      }
    }
    b = new WsBuf;
    b->ptr = p;
    b->size = grow;
    b->refs = 1;
    w.buf[slot] = b;
  }
  if (cap) {
    std::vector<WsBuf *> &refs = g_capture_refs[stream];
    bool seen = false;
    for (WsBuf *r : refs) seen = seen || r == b;
    if (!seen) {
      ++b->refs;
      refs.push_back(b);
    }
  }
  *ptr = b->ptr;
  return CORAL_OK;
}

int error_word(hipStream_t stream, int **ptr) {
  void *p = nullptr;
  CORAL_TRY(workspace(stream, WS_ERROR, 256, &p));
  *ptr = (int *)p;
  return CORAL_OK;
}

static uint32_t g_wait_limit = 1u << 22;
uint32_t wait_limit() { return g_wait_limit; }

// ---------------------------------------------------------------------------
// kernel timing: event pairs recorded on the launch stream
// ---------------------------------------------------------------------------
struct TimedLaunch { int kind; hipEvent_t start, stop; };
static bool g_timing_on = false;
static std::vector<TimedLaunch> g_timed;
static std::vector<hipEvent_t> g_event_pool;

static hipEvent_t pool_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int timing_begin(int kind, hipStream_t s) {
  if (!g_timing_on) return -1;
  std::lock_guard<std::mutex> lk(g_mu);
  TimedLaunch t{kind, pool_event(), pool_event()};
  if (!t.start || !t.stop) return -1;
  if (hipEventRecord(t.start, s) != hipSuccess) return -1;
  g_timed.push_back(t);
  return (int)g_timed.size() - 1;
}

void timing_end(int slot, hipStream_t s) {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lk(g_mu);
  (void)hipEventRecord(g_timed[slot].stop, s);
}

// ---------------------------------------------------------------------------
// counter-based generator: h(seed, i) = mix(mix(seed + G) + (i + 1) * G)
// ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t rng_at(uint64_t seed, uint64_t i) {
  const uint64_t G = 0x9E3779B97F4A7C15ULL;
  return mix64(mix64(seed + G) + (i + 1) * G);
}

template <typename T>
__global__ void fill_random_kernel(T *dst, int64_t count, uint64_t seed, int kind) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) {
    if (kind == 0) {
      uint64_t h = rng_at(seed, (uint64_t)i);
      if (sizeof(T) == 4) {
        float u = (float)(h >> 40) * 0x1.0p-24f;
        dst[i] = (T)(2.0f * u - 1.0f);
      } else {
        double u = (double)(h >> 11) * 0x1.0p-53;
        dst[i] = (T)(2.0 * u - 1.0);
      }
    } else {
      uint64_t h1 = rng_at(seed, 2 * (uint64_t)i), h2 = rng_at(seed, 2 * (uint64_t)i + 1);
      double u1 = (double)((h1 >> 11) + 1) * 0x1.0p-53;
      double u2 = (double)(h2 >> 11) * 0x1.0p-53;
      double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
      dst[i] = (T)z;
    }
  }
}

__global__ void scale_pow2_kernel(float *m, int64_t rows, int64_t cols, int64_t ld,
                                  const int32_t *rexp, const int32_t *cexp) {
  int64_t total = rows * cols;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < total; i += stride) {
    int64_t r = i % rows, c = i / rows;
    int e = (rexp ? rexp[r] : 0) + (cexp ? cexp[c] : 0);
    float *p = m + r + c * ld;
    *p = ldexpf(*p, e);
  }
}

}  // namespace coral

using namespace coral;

extern "C" {

int coral_init(void) { return ensure_init(); }

const char *coral_status_string(int s) {
  if (s < 0) return "invalid argument";
  switch (s) {
    case CORAL_OK: return "ok";
    case CORAL_ERR_HIP: return "HIP runtime error";
    case CORAL_ERR_NO_DEVICE: return "no usable gfx950 device";
    case CORAL_ERR_ALLOC: return "device allocation failed";
    case CORAL_ERR_UNSUPPORTED: return "unsupported argument combination";
    case CORAL_ERR_IO: return "autotuner file i/o failed";
    case CORAL_ERR_WAIT_BOUND: return "a device-side wait gave up; the solve's result is invalid";
    case CORAL_ERR_PROBLEM_ARG: return "a batched call skipped problems with invalid per-problem arguments";
    default: return "unknown status";
  }
}

const char *coral_last_error(void) { return g_last_error; }

int coral_device_name(char *buf, size_t len) {
  CORAL_TRY(ensure_init());
  hipDeviceProp_t prop;
  CORAL_HIP(hipGetDeviceProperties(&prop, g_device));
  snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return CORAL_OK;
}

int coral_malloc(void **ptr, size_t bytes) {
  CORAL_TRY(ensure_init());
  CORAL_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return CORAL_OK;
}
int coral_free(void *ptr) {
  if (!ptr) return CORAL_OK;
  CORAL_HIP(hipFree(ptr));
  return CORAL_OK;
}
int coral_memcpy_htod(void *dst, const void *src, size_t bytes) {
  CORAL_TRY(ensure_init());
  CORAL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  // from pageable memory the call can return once the data is staged for the DMA
  CORAL_HIP(hipStreamSynchronize(nullptr));
  return CORAL_OK;
}
int coral_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  CORAL_TRY(ensure_init());
  CORAL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return CORAL_OK;
}
// Device-to-device copies and fills are queued on the null stream and may return
// before they are done; both wait for completion, so work a caller then queues on a
// non-blocking stream (which does not wait for the null stream) sees the result.
int coral_memcpy_dtod(void *dst, const void *src, size_t bytes) {
  CORAL_TRY(ensure_init());
  CORAL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  CORAL_HIP(hipStreamSynchronize(nullptr));
  return CORAL_OK;
}
int coral_memset(void *dst, int value, size_t bytes) {
  CORAL_TRY(ensure_init());
  CORAL_HIP(hipMemset(dst, value, bytes));
  CORAL_HIP(hipStreamSynchronize(nullptr));
  return CORAL_OK;
}
int coral_synchronize(void) {
  CORAL_TRY(ensure_init());
  CORAL_HIP(hipDeviceSynchronize());
  return CORAL_OK;
}
int coral_stream_create(coral_stream_t *stream) {
  CORAL_TRY(ensure_init());
  hipStream_t s;
  CORAL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (coral_stream_t)s;
  return CORAL_OK;
}
int coral_stream_destroy(coral_stream_t stream) {
  if (!stream) return CORAL_OK;
  CORAL_HIP(hipStreamSynchronize(resolve(stream)));
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_ws.find(resolve(stream));
    if (it != g_ws.end()) {
      for (int s = 0; s < WS_NSLOTS; ++s)
        if (it->second.buf[s]) ws_release(it->second.buf[s]);  // graphs keep what they read
      g_ws.erase(it);
    }
  }
  CORAL_HIP(hipStreamDestroy(resolve(stream)));
  return CORAL_OK;
}
int coral_stream_check(coral_stream_t stream) {
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  if (capturing(s)) {
    set_last_error("coral_stream_check synchronises the stream and cannot run under capture");
    return CORAL_ERR_UNSUPPORTED;
  }
  CORAL_HIP(hipStreamSynchronize(s));
  int *err = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_ws.find(s);
    if (it == g_ws.end() || !it->second.buf[WS_ERROR]) return CORAL_OK;  // no checked routine ran
    err = (int *)it->second.buf[WS_ERROR]->ptr;
  }
  int flags = 0;
  CORAL_HIP(hipMemcpy(&flags, err, sizeof(int), hipMemcpyDeviceToHost));
  if (flags == 0) return CORAL_OK;
  // cleared on the stream itself, so flags set by work queued on it later are kept
  CORAL_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
  set_last_error("%s%s", (flags & ERRF_WAIT_BOUND) ? "a triangular solve's bounded wait gave up (its x is invalid); " : "",
                 (flags & ERRF_PROBLEM_ARG) ? "a batched call skipped problems with invalid per-problem "
                                              "dimensions or leading dimensions" : "");
  return (flags & ERRF_WAIT_BOUND) ? CORAL_ERR_WAIT_BOUND : CORAL_ERR_PROBLEM_ARG;
}
int coral_testing_set_wait_limit(int64_t spins) {
  if (spins < -1 || spins > (int64_t)UINT32_MAX) return -1;
  g_wait_limit = spins == 0 ? (1u << 22) : (spins < 0 ? 0u : (uint32_t)spins);
  return CORAL_OK;
}
int coral_stream_synchronize(coral_stream_t stream) {
  CORAL_TRY(ensure_init());
  CORAL_HIP(hipStreamSynchronize(resolve(stream)));
  return CORAL_OK;
}
int coral_event_create(void **event) {
  CORAL_TRY(ensure_init());
  hipEvent_t e;
  CORAL_HIP(hipEventCreate(&e));
  *event = (void *)e;
  return CORAL_OK;
}
int coral_event_destroy(void *event) {
  if (event) CORAL_HIP(hipEventDestroy((hipEvent_t)event));
  return CORAL_OK;
}
int coral_event_record(void *event, coral_stream_t stream) {
  CORAL_HIP(hipEventRecord((hipEvent_t)event, resolve(stream)));
  return CORAL_OK;
}
int coral_event_elapsed_ms(float *ms, void *start, void *stop) {
  CORAL_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return CORAL_OK;
}
int coral_event_synchronize(void *event) {
  CORAL_HIP(hipEventSynchronize((hipEvent_t)event));
  return CORAL_OK;
}

int coral_graph_begin_capture(coral_stream_t stream) {
  CORAL_TRY(ensure_init());
  if (!stream) {
    set_last_error("graph capture needs a created stream (not the null stream)");
    return -1;
  }
  CORAL_HIP(hipStreamBeginCapture(resolve(stream), hipStreamCaptureModeThreadLocal));
  return CORAL_OK;
}
int coral_graph_end_capture(coral_stream_t stream, void **graph_exec) {
  if (!graph_exec) return -2;
  hipStream_t s = resolve(stream);
  std::vector<WsBuf *> refs;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_capture_refs.find(s);
    if (it != g_capture_refs.end()) {
      refs.swap(it->second);
      g_capture_refs.erase(it);
    }
  }
  auto drop = [&]() {
    std::lock_guard<std::mutex> lk(g_mu);
    for (WsBuf *b : refs) ws_release(b);
  };
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(s, &graph);
  if (e != hipSuccess) {
    drop();
    return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    drop();
    return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
  }
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_graph_refs[exec] = std::move(refs);
  }
  *graph_exec = (void *)exec;
  return CORAL_OK;
}
int coral_graph_launch(void *graph_exec, coral_stream_t stream) {
  if (!graph_exec) return -1;
  CORAL_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, resolve(stream)));
  return CORAL_OK;
}
int coral_graph_destroy(void *graph_exec) {
  if (!graph_exec) return CORAL_OK;
  hipGraphExec_t exec = (hipGraphExec_t)graph_exec;
  CORAL_HIP(hipGraphExecDestroy(exec));
  // the buffers are freed only after replays already queued have finished
  CORAL_HIP(hipDeviceSynchronize());
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_graph_refs.find(exec);
  if (it != g_graph_refs.end()) {
    for (WsBuf *b : it->second) ws_release(b);
    g_graph_refs.erase(it);
  }
  return CORAL_OK;
}

int coral_timing_enable(int32_t on) {
  CORAL_TRY(ensure_init());
  g_timing_on = on != 0;
  return CORAL_OK;
}

int coral_timing_reset(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &t : g_timed) {
    (void)hipEventSynchronize(t.stop);
    g_event_pool.push_back(t.start);
    g_event_pool.push_back(t.stop);
  }
  g_timed.clear();
  return CORAL_OK;
}

int coral_timing_read(int32_t kind, double *total_ms, int64_t *launches) {
  if (kind < 0 || kind >= TK_NKINDS) return -1;
  if (!total_ms) return -2;
  if (!launches) return -3;
  std::lock_guard<std::mutex> lk(g_mu);
  double sum = 0.0;
  int64_t cnt = 0;
  for (auto &t : g_timed) {
    if (t.kind != kind) continue;
    CORAL_HIP(hipEventSynchronize(t.stop));
    float ms = 0.f;
    CORAL_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
    sum += ms;
    ++cnt;
  }
  *total_ms = sum;
  *launches = cnt;
  return CORAL_OK;
}

int coral_fill_random_f32(float *dst, int64_t count, uint64_t seed, int32_t kind, coral_stream_t stream) {
  if (!dst && count > 0) return -1;
  if (count < 0) return -2;
  if (kind != 0 && kind != 1) return -4;
  CORAL_TRY(ensure_init());
  if (count == 0) return CORAL_OK;
  int64_t blocks = cdiv(count, 256);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(fill_random_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, resolve(stream), dst, count, seed, kind);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}
int coral_fill_random_f64(double *dst, int64_t count, uint64_t seed, int32_t kind, coral_stream_t stream) {
  if (!dst && count > 0) return -1;
  if (count < 0) return -2;
  if (kind != 0 && kind != 1) return -4;
  CORAL_TRY(ensure_init());
  if (count == 0) return CORAL_OK;
  int64_t blocks = cdiv(count, 256);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(fill_random_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, resolve(stream), dst, count, seed, kind);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}
int coral_scale_pow2_f32(float *m, int64_t rows, int64_t cols, int64_t ld, const int32_t *row_exp,
                         const int32_t *col_exp, coral_stream_t stream) {
  if (rows < 0) return -2;
  if (cols < 0) return -3;
  if (ld < imax64(1, rows)) return -4;
  CORAL_TRY(ensure_init());
  if (rows == 0 || cols == 0) return CORAL_OK;
  int64_t blocks = cdiv(rows * cols, 256);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(scale_pow2_kernel, dim3((unsigned)blocks), dim3(256), 0, resolve(stream), m, rows, cols, ld, row_exp, col_exp);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

}  // extern "C"
