// Device context, owning device buffers and small elementwise helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "../../include/afqmc_mi355x.h"
#include "common.hpp"
#include "cp_kernels.hpp"
#include "group.hpp"

// Optional per-region kernel timing: when enabled, tagged launches are bracketed
// by HIP events on the context stream; totals are resolved on query.
enum TimedRegion {
  REG_PROPAGATE = 0, REG_OVERLAP, REG_GJ, REG_GREEN, REG_HS_SAMPLE, REG_QR, REG_COMB, REG_MEASURE, REG_PH_FORCE,
  REG_PH_SAMPLE, REG_PH_TAYLOR, REG_HS_APPLY, REG_COUNT
};
struct afqmc_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  int timing = 0;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> events;
  double total_ms[REG_COUNT] = {0};
  long long count[REG_COUNT] = {0};
};

template <typename T> struct DevBuf;

// Stream capture into a HIP graph (afqmc_capture_begin / afqmc_capture_end).  While a capture is
// open on this thread, entry points that only enqueue work skip their host synchronisation and
// timing events, and walker sets record the buffer pairs whose roles the captured work swaps:
// the graph replays those swaps on the host at every launch.
struct CaptureSession {
  afqmc_ctx origin = nullptr;
  std::vector<afqmc_ctx> members;
  std::vector<std::pair<DevBuf<char>*, DevBuf<char>*>> pairs;
  std::vector<void*> first;  // pairs[i].first->p when the pair was noted
};
extern thread_local CaptureSession* g_capture;
inline bool capturing() { return g_capture != nullptr; }
inline hipError_t sync_unless_capturing(hipStream_t s) { return g_capture ? hipSuccess : hipStreamSynchronize(s); }
void capture_note_pair(DevBuf<char>* a, DevBuf<char>* b);
#define NOT_CAPTURING() REQUIRE(!capturing(), "%s cannot be recorded into a graph (it synchronises)", __func__)

struct GraphPair {
  DevBuf<char>* a;
  DevBuf<char>* b;
  void* start;  // a->p when the capture began: the buffer the graph reads as `a`
  int swapped;  // the captured work leaves the pair with roles exchanged
};
struct afqmc_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  afqmc_ctx origin = nullptr;
  std::vector<afqmc_ctx> members;
  std::vector<GraphPair> pairs;
  hipEvent_t ev = nullptr;
};

// RAII bracket of one tagged region on the context stream
struct RegionTimer {
  afqmc_ctx ctx;
  int reg;
  hipEvent_t a = nullptr, b = nullptr;
  RegionTimer(afqmc_ctx c, int r) : ctx(c), reg(r) {
    if (ctx && ctx->timing && !capturing() && !grouping()) {
      // a region whose events cannot be made or recorded is left out of the timing
      if (hipEventCreate(&a) != hipSuccess) a = nullptr;
      if (a && hipEventCreate(&b) != hipSuccess) b = nullptr;
      if (a && (!b || hipEventRecord(a, ctx->stream) != hipSuccess)) {
        (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        a = b = nullptr;
      }
    }
  }
  ~RegionTimer() {
    if (!a) return;
    if (hipEventRecord(b, ctx->stream) == hipSuccess) {
      ctx->events.push_back({reg, {a, b}});
    } else {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  }
};

// Selects the context's device for the lifetime of a C entry point.
struct ScopedDevice {
  int prev = 0;
  // device selection errors surface at the entry point's first HIP call on that device
  explicit ScopedDevice(afqmc_ctx ctx) {
    (void)hipGetDevice(&prev);
    if (ctx && ctx->device != prev) (void)hipSetDevice(ctx->device);
  }
  explicit ScopedDevice(int device) {
    (void)hipGetDevice(&prev);
    if (device != prev) (void)hipSetDevice(device);
  }
  ~ScopedDevice() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != prev) (void)hipSetDevice(prev);
  }
};

// Owning device allocation (hipMalloc / hipFree), movable.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
  }
  int alloc(size_t count) { return alloc_bytes(count * sizeof(T)); }
  int alloc_bytes(size_t bytes) {
    release();
    if (bytes == 0) bytes = 16;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
      afqmc_set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
      return -3;
    }
    p = reinterpret_cast<T*>(q);
    n = bytes / sizeof(T);
    return 0;
  }
};
template <typename T>
void swap(DevBuf<T>& a, DevBuf<T>& b) noexcept {
  std::swap(a.p, b.p);
  std::swap(a.n, b.n);
}

// looks up the device's copy of the comb scratch array used by afqmc_comb_noalloc (device current)
int afqmc_noalloc_scratch_init(int device);

int launch_convert_f32_f64(const void* in, void* out, long long ncomplex, hipStream_t s);
int launch_convert_f64_f32(const void* in, void* out, long long ncomplex, hipStream_t s);
// the per_walker elements of each walker w with mask[w] set
int launch_convert_f64_f32_masked(const void* in, void* out, long long per_walker, int nw, const int* mask,
                                  hipStream_t s);
// logw_out[w] = logw[w] + c
int launch_logw_shift(int nw, const zd* logw, zd* logw_out, zd c, hipStream_t s);
// logw[w] += c for alive walkers
int launch_logw_const_add(int nw, zd* logw, const int* alive, zd c, hipStream_t s);
// logw[w] += v[w] for alive walkers
int launch_logw_vec_add(int nw, zd* logw, const int* alive, const zd* v, hipStream_t s);
// dst[w] = src[w] (per-walker blocks of n elements of size esz) for alive walkers
int launch_select_alive(int nw, long long n, size_t esz, const int* alive, const void* src, void* dst, hipStream_t s);
int launch_reset_measure(DevScalars* sc, zd* site_acc, int L, hipStream_t s);
