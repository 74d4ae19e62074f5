// Context management, error reporting and elementwise helper kernels.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <functional>

#include "runtime.hpp"

#include "group.hpp"

static thread_local char g_err[1024] = "";

void afqmc_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

thread_local GroupCollector* g_group = nullptr;

GroupTables::~GroupTables() {
  for (auto& kv : map_) (void)hipFree(kv.second);
  if (side_) (void)hipStreamDestroy(side_);
  for (hipStream_t& f : fork_)
    if (f) (void)hipStreamDestroy(f);
  for (hipEvent_t& e : ev_)
    if (e) (void)hipEventDestroy(e);
}

bool GroupTables::fork_resources(hipStream_t* streams, hipEvent_t* events) {
  for (int i = 0; i < FORKS; ++i) {
    if (!fork_[i] && hipStreamCreateWithFlags(&fork_[i], hipStreamNonBlocking) != hipSuccess) return false;
    streams[i] = fork_[i];
  }
  for (int i = 0; i <= FORKS; ++i) {
    if (!ev_[i] && hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming) != hipSuccess) return false;
    events[i] = ev_[i];
  }
  return true;
}

const void* GroupTables::get(const void* p, size_t n) {
  std::string key(reinterpret_cast<const char*>(p), n);
  auto it = map_.find(key);
  if (it != map_.end()) return it->second;
  // a stream of its own (not the capturing one) for the upload, so the call is legal while a
  // graph capture in relaxed mode is open
  if (!side_ && hipStreamCreateWithFlags(&side_, hipStreamNonBlocking) != hipSuccess) return nullptr;
  void* d = nullptr;
  if (hipMalloc(&d, n) != hipSuccess) return nullptr;
  if (hipMemcpyAsync(d, p, n, hipMemcpyHostToDevice, side_) != hipSuccess || hipStreamSynchronize(side_) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  map_.emplace(std::move(key), d);
  return d;
}

int group_flush_all(GroupCollector& c, hipStream_t s) {
  std::vector<GroupRecord> recs = std::move(c.recs);
  c.recs.clear();
  // by position, then kernel and block shape, then member order
  std::vector<const GroupRecord*> order;
  order.reserve(recs.size());
  for (const GroupRecord& r : recs) order.push_back(&r);
  auto shape = [](const GroupRecord* r) {
    return ((unsigned long long)r->block.x << 40) ^ ((unsigned long long)r->block.y << 20) ^ r->block.z;
  };
  std::stable_sort(order.begin(), order.end(), [&](const GroupRecord* a, const GroupRecord* b) {
    if (a->pos != b->pos) return a->pos < b->pos;
    if (a->key != b->key) return std::less<const void*>()(a->key, b->key);
    if (shape(a) != shape(b)) return shape(a) < shape(b);
    return a->seq < b->seq;
  });
  hipStream_t fork[GroupTables::FORKS];
  hipEvent_t ev[GroupTables::FORKS + 1];
  bool have_forks = false;
  for (size_t i = 0; i < order.size();) {
    // the runs of one position: [i, e), each run one kernel and block shape
    size_t e = i + 1;
    while (e < order.size() && order[e]->pos == order[i]->pos) ++e;
    std::vector<size_t> runs;
    for (size_t k = i; k < e; ++k)
      if (k == i || order[k]->key != order[k - 1]->key || shape(order[k]) != shape(order[k - 1])) runs.push_back(k);
    runs.push_back(e);
    const int nrun = (int)runs.size() - 1;
    if (nrun > 1 && !have_forks) {
      REQUIRE(c.tables->fork_resources(fork, ev), "grouped launch: no side streams");
      have_forks = true;
    }
    if (nrun > 1) HIP_TRY(hipEventRecord(ev[0], s));
    const int used = nrun > 1 ? std::min(nrun - 1, GroupTables::FORKS) : 0;
    for (int f = 0; f < used; ++f) HIP_TRY(hipStreamWaitEvent(fork[f], ev[0], 0));
    for (int r = 0; r < nrun; ++r) {
      hipStream_t rs = r == 0 ? s : fork[(r - 1) % GroupTables::FORKS];
      int rc = order[runs[r]]->flush(order.data() + runs[r], (int)(runs[r + 1] - runs[r]), rs, c.tables);
      if (rc) return rc;
    }
    for (int f = 0; f < used; ++f) {
      HIP_TRY(hipEventRecord(ev[1 + f], fork[f]));
      HIP_TRY(hipStreamWaitEvent(s, ev[1 + f], 0));
    }
    i = e;
  }
  return 0;
}

namespace {

template <typename A, typename B>
struct ConvertArgs {
  const cplx<A>* in;
  cplx<B>* out;
  long long n;
};
template <typename A, typename B>
struct ConvertK {
  static constexpr int NT = 256, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const ConvertArgs<A, B>& g);
};
template <typename A, typename B>
__device__ __forceinline__ void ConvertK<A, B>::run(const BlockIdx3& bk, const ConvertArgs<A, B>& g) {
  const cplx<A>* in = g.in;
  cplx<B>* out = g.out;
  const long long n = g.n;
  for (long long i = (long long)bk.x * blockDim.x + threadIdx.x; i < n; i += (long long)bk.gx * blockDim.x)
    out[i] = cmk<B>((B)in[i].re, (B)in[i].im);
}

// per walker: convert the n elements of walker w where mask[w] is set
template <typename A, typename B>
struct ConvertMaskedArgs {
  const cplx<A>* in;
  cplx<B>* out;
  const int* mask;
  long long n;
};
template <typename A, typename B>
struct ConvertMaskedK {
  static constexpr int NT = 256, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const ConvertMaskedArgs<A, B>& g);
};
template <typename A, typename B>
__device__ __forceinline__ void ConvertMaskedK<A, B>::run(const BlockIdx3& bk, const ConvertMaskedArgs<A, B>& g) {
  const int w = bk.y;
  if (!g.mask[w]) return;
  const long long n = g.n;
  const cplx<A>* in = g.in + (long long)w * n;
  cplx<B>* out = g.out + (long long)w * n;
  for (long long i = (long long)bk.x * blockDim.x + threadIdx.x; i < n; i += (long long)bk.gx * blockDim.x)
    out[i] = cmk<B>((B)in[i].re, (B)in[i].im);
}

__global__ void logw_shift_kernel(int nw, const zd* logw, zd* out, zd c) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nw) out[w] = cadd(logw[w], c);
}
__global__ void logw_const_add_kernel(int nw, zd* logw, const int* alive, zd c) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nw && alive[w]) logw[w] = cadd(logw[w], c);
}
__global__ void logw_vec_add_kernel(int nw, zd* logw, const int* alive, const zd* v) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nw && alive[w]) logw[w] = cadd(logw[w], v[w]);
}
template <typename E>
struct SelectArgs {
  const int* alive;
  const E* src;
  E* dst;
  long long n;
};
template <typename E>
struct SelectK {
  static constexpr int NT = 256, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const SelectArgs<E>& g);
};
template <typename E>
__device__ __forceinline__ void SelectK<E>::run(const BlockIdx3& bk, const SelectArgs<E>& g) {
  const long long n = g.n;
  const int w = bk.y;
  if (!g.alive[w]) return;
  for (long long i = (long long)bk.x * blockDim.x + threadIdx.x; i < n; i += (long long)bk.gx * blockDim.x)
    g.dst[(long long)w * n + i] = g.src[(long long)w * n + i];
}
__global__ void reset_measure_kernel(DevScalars* sc, zd* site_acc, int L) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    zd z = cmk<double>(0, 0);
    sc->den = z; sc->HNum = z; sc->KNum = z; sc->VNum = z; sc->RNum = z;
    sc->NupTot = z; sc->NdnTot = z; sc->SplusTot = z; sc->SminusTot = z;
  }
  if (i < 4 * L) site_acc[i] = cmk<double>(0, 0);
}

}  // namespace

int launch_convert_f32_f64(const void* in, void* out, long long n, hipStream_t s) {
  unsigned g = (unsigned)std::min<long long>((n + 255) / 256, 65536);
  if (n <= 0) return 0;
  ConvertArgs<float, double> a{(const zf*)in, (zd*)out, n};
  return glaunch<ConvertK<float, double>>(a, dim3(g), dim3(256), 0, s);
}
int launch_convert_f64_f32(const void* in, void* out, long long n, hipStream_t s) {
  unsigned g = (unsigned)std::min<long long>((n + 255) / 256, 65536);
  if (n <= 0) return 0;
  ConvertArgs<double, float> a{(const zd*)in, (zf*)out, n};
  return glaunch<ConvertK<double, float>>(a, dim3(g), dim3(256), 0, s);
}
int launch_convert_f64_f32_masked(const void* in, void* out, long long per_walker, int nw, const int* mask,
                                  hipStream_t s) {
  if (nw <= 0 || per_walker <= 0) return 0;
  ConvertMaskedArgs<double, float> a{(const zd*)in, (zf*)out, mask, per_walker};
  const unsigned gx = (unsigned)std::min<long long>((per_walker + 255) / 256, 64);
  return glaunch<ConvertMaskedK<double, float>>(a, dim3(gx, nw), dim3(256), 0, s);
}
int launch_logw_shift(int nw, const zd* logw, zd* out, zd c, hipStream_t s) {
  hipLaunchKernelGGL(logw_shift_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, nw, logw, out, c);
  HIP_LAUNCH_CHECK();
  return 0;
}
int launch_logw_const_add(int nw, zd* logw, const int* alive, zd c, hipStream_t s) {
  hipLaunchKernelGGL(logw_const_add_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, nw, logw, alive, c);
  HIP_LAUNCH_CHECK();
  return 0;
}
int launch_logw_vec_add(int nw, zd* logw, const int* alive, const zd* v, hipStream_t s) {
  hipLaunchKernelGGL(logw_vec_add_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, nw, logw, alive, v);
  HIP_LAUNCH_CHECK();
  return 0;
}
int launch_select_alive(int nw, long long n, size_t esz, const int* alive, const void* src, void* dst, hipStream_t s) {
  if (nw <= 0 || n <= 0) return 0;
  dim3 grid((unsigned)std::min<long long>((n + 255) / 256, 1024), nw);
  if (esz == 16) {
    SelectArgs<uint4> g{alive, (const uint4*)src, (uint4*)dst, n};
    return glaunch<SelectK<uint4>>(g, grid, dim3(256), 0, s);
  }
  SelectArgs<uint2> g{alive, (const uint2*)src, (uint2*)dst, n};
  return glaunch<SelectK<uint2>>(g, grid, dim3(256), 0, s);
}
int launch_reset_measure(DevScalars* sc, zd* site_acc, int L, hipStream_t s) {
  hipLaunchKernelGGL(reset_measure_kernel, dim3((4 * L + 255) / 256 + (L == 0)), dim3(256), 0, s, sc, site_acc, L);
  HIP_LAUNCH_CHECK();
  return 0;
}

thread_local CaptureSession* g_capture = nullptr;

void capture_note_pair(DevBuf<char>* a, DevBuf<char>* b) {
  if (!g_capture) return;
  for (auto& q : g_capture->pairs)
    if (q.first == a || q.first == b) return;
  g_capture->pairs.push_back({a, b});
  g_capture->first.push_back(a->p);
}

extern "C" {

const char* afqmc_last_error(void) { return g_err; }

int afqmc_version(void) { return 1; }

int afqmc_ctx_create(int device, afqmc_ctx* out) {
  REQUIRE(out, "null argument");
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  REQUIRE(device >= 0 && device < n, "device %d not available (%d devices)", device, n);
  afqmc_ctx c = new afqmc_ctx_s();
  c->device = device;
  int prev = 0;
  hipGetDevice(&prev);
  hipSetDevice(device);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  // the comb scratch array's address on this device, looked up once with the device current
  const int src = e == hipSuccess ? afqmc_noalloc_scratch_init(device) : 0;
  hipSetDevice(prev);
  if (e != hipSuccess) {
    delete c;
    afqmc_set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
    return -2;
  }
  if (src) {
    hipStreamDestroy(c->stream);
    delete c;
    return src;
  }
  *out = c;
  return 0;
}

int afqmc_ctx_destroy(afqmc_ctx ctx) {
  if (!ctx) return 0;
  ScopedDevice sd(ctx);
  hipStreamSynchronize(ctx->stream);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

int afqmc_ctx_synchronize(afqmc_ctx ctx) {
  REQUIRE(ctx, "null argument");
  ScopedDevice sd(ctx);
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return 0;
}

static int resolve_events(afqmc_ctx ctx) {
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  // every pending pair is read and released; the first failure is reported after all of them
  hipError_t first = hipSuccess;
  for (auto& e : ctx->events) {
    float ms = 0.f;
    const hipError_t err = hipEventElapsedTime(&ms, e.second.first, e.second.second);
    if (err == hipSuccess) {
      ctx->total_ms[e.first] += ms;
      ctx->count[e.first] += 1;
    } else if (first == hipSuccess) {
      first = err;
    }
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  ctx->events.clear();
  HIP_TRY(first);
  return 0;
}

int afqmc_ctx_set_timing(afqmc_ctx ctx, int enable) {
  REQUIRE(ctx, "null argument");
  ScopedDevice sd(ctx);
  int rc = resolve_events(ctx);
  if (rc) return rc;
  ctx->timing = enable ? 1 : 0;
  for (int r = 0; r < REG_COUNT; ++r) { ctx->total_ms[r] = 0.0; ctx->count[r] = 0; }
  return 0;
}

int afqmc_ctx_get_timing(afqmc_ctx ctx, double* total_ms, long long* count, int n) {
  REQUIRE(ctx && total_ms && count, "null argument");
  ScopedDevice sd(ctx);
  int rc = resolve_events(ctx);
  if (rc) return rc;
  for (int r = 0; r < n && r < REG_COUNT; ++r) { total_ms[r] = ctx->total_ms[r]; count[r] = ctx->count[r]; }
  return 0;
}

int afqmc_capture_begin(afqmc_ctx ctx) { return afqmc_capture_begin_group(ctx, 0, nullptr); }

int afqmc_capture_begin_group(afqmc_ctx origin, int n, const afqmc_ctx* members) {
  REQUIRE(origin && n >= 0 && (n == 0 || members), "invalid argument");
  REQUIRE(!g_capture, "a graph capture is already open on this thread");
  for (int i = 0; i < n; ++i) {
    REQUIRE(members[i] && members[i]->device == origin->device, "group members must be contexts on the origin's device");
    for (int j = 0; j < i; ++j) REQUIRE(members[j] != members[i], "group members must be distinct contexts");
    REQUIRE(members[i] != origin, "the origin context is not a member");
  }
  ScopedDevice sd(origin);
  HIP_TRY(hipStreamSynchronize(origin->stream));
  for (int i = 0; i < n; ++i) HIP_TRY(hipStreamSynchronize(members[i]->stream));
  HIP_TRY(hipStreamBeginCapture(origin->stream, hipStreamCaptureModeRelaxed));
  auto* cs = new CaptureSession();
  cs->origin = origin;
  if (n > 0) {
    // members join the capture by waiting on an event recorded in it
    hipEvent_t fork = nullptr;
    hipError_t e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(fork, origin->stream);
    for (int i = 0; i < n && e == hipSuccess; ++i) {
      e = hipStreamWaitEvent(members[i]->stream, fork, 0);
      if (e == hipSuccess) cs->members.push_back(members[i]);
    }
    if (fork) hipEventDestroy(fork);
    if (e != hipSuccess) {
      hipGraph_t g = nullptr;
      hipStreamEndCapture(origin->stream, &g);
      if (g) hipGraphDestroy(g);
      delete cs;
      afqmc_set_error("afqmc_capture_begin_group: %s", hipGetErrorString(e));
      return -2;
    }
  }
  g_capture = cs;
  return 0;
}

int afqmc_capture_end(afqmc_ctx origin, afqmc_graph* out) {
  REQUIRE(origin && out, "null argument");
  REQUIRE(g_capture && g_capture->origin == origin, "no graph capture is open for this context");
  CaptureSession* cs = g_capture;
  g_capture = nullptr;
  ScopedDevice sd(origin);
  hipError_t e = hipSuccess;
  std::vector<hipEvent_t> joins;
  for (afqmc_ctx m : cs->members) {
    hipEvent_t j = nullptr;
    if (e == hipSuccess) e = hipEventCreateWithFlags(&j, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(j, m->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(origin->stream, j, 0);
    if (j) joins.push_back(j);
  }
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(origin->stream, &g);
  for (hipEvent_t j : joins) hipEventDestroy(j);
  // the captured work has not run: give the swapped buffer pairs back their roles
  std::vector<GraphPair> pairs;
  for (size_t i = 0; i < cs->pairs.size(); ++i) {
    auto& q = cs->pairs[i];
    const int swapped = q.first->p != cs->first[i];
    if (swapped) swap(*q.first, *q.second);
    pairs.push_back({q.first, q.second, cs->first[i], swapped});
  }
  std::vector<afqmc_ctx> members = cs->members;
  delete cs;
  if (e != hipSuccess || e2 != hipSuccess || !g) {
    if (g) hipGraphDestroy(g);
    afqmc_set_error("afqmc_capture_end: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    return -2;
  }
  auto* gr = new afqmc_graph_s();
  gr->graph = g;
  gr->origin = origin;
  gr->members = std::move(members);
  gr->pairs = std::move(pairs);
  e = hipGraphInstantiate(&gr->exec, g, nullptr, nullptr, 0);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&gr->ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    if (gr->exec) hipGraphExecDestroy(gr->exec);
    hipGraphDestroy(g);
    delete gr;
    afqmc_set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
    return -2;
  }
  *out = gr;
  return 0;
}

int afqmc_graph_launch(afqmc_graph g) {
  REQUIRE(g && g->exec, "null graph");
  REQUIRE(!capturing(), "afqmc_graph_launch cannot be recorded into a graph");
  ScopedDevice sd(g->origin);
  hipStream_t s = g->origin->stream;
  // order after work already queued on the member streams
  for (afqmc_ctx m : g->members) {
    HIP_TRY(hipEventRecord(g->ev, m->stream));
    HIP_TRY(hipStreamWaitEvent(s, g->ev, 0));
  }
  // the graph reads each pair's data from the buffer it had at capture time: if eager calls
  // since then left the data in the other buffer, move it back first
  for (auto& q : g->pairs) {
    if (q.a->p != q.start) {
      HIP_TRY(hipMemcpyAsync(q.start, q.a->p, q.a->n, hipMemcpyDeviceToDevice, s));
      swap(*q.a, *q.b);
    }
  }
  HIP_TRY(hipGraphLaunch(g->exec, s));
  for (auto& q : g->pairs)
    if (q.swapped) swap(*q.a, *q.b);
  // later work on the member streams follows the graph
  if (!g->members.empty()) {
    HIP_TRY(hipEventRecord(g->ev, s));
    for (afqmc_ctx m : g->members) HIP_TRY(hipStreamWaitEvent(m->stream, g->ev, 0));
  }
  return 0;
}

int afqmc_graph_node_count(afqmc_graph g) {
  if (!g || !g->graph) return -1;
  size_t n = 0;
  if (hipGraphGetNodes(g->graph, nullptr, &n) != hipSuccess) return -1;
  return (int)n;
}

int afqmc_graph_destroy(afqmc_graph g) {
  if (!g) return 0;
  ScopedDevice sd(g->origin);
  hipStreamSynchronize(g->origin->stream);
  for (afqmc_ctx m : g->members) hipStreamSynchronize(m->stream);
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  if (g->ev) hipEventDestroy(g->ev);
  delete g;
  return 0;
}

}  // extern "C"
