// ctx.hip: context lifetime, staging channels, solution state and the graph entry points of
// the C ABI (include/scout_hip.h).
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.cuh"

namespace scout {

namespace {
thread_local std::string g_err;
}

void set_error(const std::string& msg) { g_err = msg; }

int box_read(Ctx& c, int count) {
  SCOUT_TRY(hipMemcpyAsync(c.h_box, c.d_box, (size_t)count * sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  c.readbacks++;
  return SCOUT_OK;
}

int box_read_async(Ctx& c, int first, int count) {
  if (!c.box_ev) SCOUT_TRY(hipEventCreateWithFlags(&c.box_ev, hipEventDisableTiming));
  SCOUT_TRY(hipMemcpyAsync(c.h_box + first, c.d_box + first, (size_t)count * sizeof(int64_t), hipMemcpyDeviceToHost,
                           c.s));
  SCOUT_TRY(hipEventRecord(c.box_ev, c.s));
  return SCOUT_OK;
}

int box_wait(Ctx& c) {
  SCOUT_TRY(hipEventSynchronize(c.box_ev));
  c.readbacks++;
  return SCOUT_OK;
}

int next_epoch(Ctx& c) {
  if (c.epoch >= 0x7ffffff0) {
    // every array compared against epochs starts over (epochs are > 0)
    size_t bytes = ((size_t)c.g.n + 1) * sizeof(int32_t);
    SCOUT_TRY(hipMemsetAsync(c.stamp, 0, bytes, c.s));
    if (c.aux) SCOUT_TRY(hipMemsetAsync(c.aux, 0xff, bytes, c.s));
    c.epoch = 0;
  }
  return ++c.epoch;
}

namespace {

int alloc_n(Ctx& c, void** p, size_t elem) {
  if (*p) return SCOUT_OK;
  size_t bytes = ((size_t)c.g.n + 1) * elem;
  if (c.budget.hbm_bytes > 0 && c.hbm_used + (int64_t)bytes > c.budget.hbm_bytes) {
    set_error("HBM budget exceeded by solution state");
    return SCOUT_ERR_BUDGET;
  }
  SCOUT_TRY(hipMalloc(p, bytes));
  c.hbm_used += (int64_t)bytes;
  return SCOUT_OK;
}

void free_state(Ctx& c) {
  void* ps[] = {c.level, c.parent, c.label, c.color, c.mis, c.stamp, c.aux, c.aux2,
                c.cc_owner, c.col_marker, c.mis_minw, c.cc_level, c.cc_parent};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  c.cc_level = c.cc_parent = c.parent = nullptr;
  c.level_seq = c.cc_level_seq = c.label_seq = -1;
  c.cc_components = -1;
  c.level = c.label = c.color = nullptr;
  c.mis = nullptr;
  c.stamp = c.aux = c.aux2 = nullptr;
  c.cc_owner = c.col_marker = c.mis_minw = nullptr;
}

int check_ctx(scout_ctx* x) {
  if (!x) {
    set_error("null context");
    return SCOUT_ERR_ARG;
  }
  if (hipSetDevice(x->c.dev) != hipSuccess) {
    set_error("hipSetDevice failed");
    return SCOUT_ERR_HIP;
  }
  return SCOUT_OK;
}

int check_graph(scout_ctx* x) {
  SCOUT_CHECK(check_ctx(x));
  if (!x->c.g.boff) {
    set_error("no graph loaded");
    return SCOUT_ERR_STATE;
  }
  return SCOUT_OK;
}

}  // namespace

int ensure_state(Ctx& c, int kind) {
  switch (kind) {
    case SCOUT_STATE_MIS: return alloc_n(c, (void**)&c.mis, 1);
    case SCOUT_STATE_COLOR: return alloc_n(c, (void**)&c.color, 4);
    case SCOUT_STATE_LEVEL:
      SCOUT_CHECK(alloc_n(c, (void**)&c.level, 4));
      return alloc_n(c, (void**)&c.parent, 4);
    case SCOUT_STATE_LABEL: return alloc_n(c, (void**)&c.label, 4);
  }
  set_error("unknown state kind");
  return SCOUT_ERR_ARG;
}

int ensure_scratch(Ctx& c, int32_t** p, int32_t fill) {
  if (*p) return SCOUT_OK;
  SCOUT_CHECK(alloc_n(c, (void**)p, 4));
  // fill with a repeated byte when possible, else a kernel-free host upload
  uint8_t b = (uint8_t)(fill & 0xff);
  if (fill == (int32_t)(b * 0x01010101u)) {
    SCOUT_TRY(hipMemsetAsync(*p, b, ((size_t)c.g.n + 1) * 4, c.s));
  } else {
    std::vector<int32_t> h((size_t)c.g.n + 1, fill);
    SCOUT_TRY(hipMemcpyAsync(*p, h.data(), h.size() * 4, hipMemcpyHostToDevice, c.s));
    SCOUT_TRY(hipStreamSynchronize(c.s));
  }
  return SCOUT_OK;
}

int ensure_list(DevBuf<int32_t>& l, int64_t count) { return l.reserve(count); }

int Span::begin(hipStream_t s) {
  (void)s;
  t0 = std::chrono::steady_clock::now();
  return SCOUT_OK;
}
int Span::end(hipStream_t s) {
  (void)s;
  return SCOUT_OK;
}
double Span::ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

struct ProfState {
  bool on = false;
  struct Entry {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> spans;
    int64_t bytes = 0;
    hipEvent_t open = nullptr;
  } e[PROF_COUNT];
  std::vector<hipEvent_t> pool;
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t ev = pool.back();
      pool.pop_back();
      return ev;
    }
    hipEvent_t ev = nullptr;
    (void)hipEventCreate(&ev);
    return ev;
  }
  void clear() {
    for (auto& en : e) {
      for (auto& sp : en.spans) {
        pool.push_back(sp.first);
        pool.push_back(sp.second);
      }
      en.spans.clear();
      en.bytes = 0;
      if (en.open) pool.push_back(en.open);
      en.open = nullptr;
    }
  }
};

int prof_begin(Ctx& c, int id) {
  if (!c.prof || !c.prof->on) return SCOUT_OK;
  ProfState::Entry& en = c.prof->e[id];
  en.open = c.prof->get();
  SCOUT_TRY(hipEventRecord(en.open, c.s));
  return SCOUT_OK;
}

int prof_end(Ctx& c, int id, int64_t bytes) {
  if (!c.prof || !c.prof->on) return SCOUT_OK;
  ProfState::Entry& en = c.prof->e[id];
  if (!en.open) return SCOUT_OK;
  hipEvent_t b = c.prof->get();
  SCOUT_TRY(hipEventRecord(b, c.s));
  en.spans.emplace_back(en.open, b);
  en.open = nullptr;
  en.bytes += bytes;
  return SCOUT_OK;
}

bool prof_kernel_events(Ctx& c, int id, hipEvent_t* start, hipEvent_t* stop) {
  if (!c.prof || !c.prof->on) return false;
  hipEvent_t a = c.prof->get(), b = c.prof->get();
  if (!a || !b) return false;
  c.prof->e[id].spans.emplace_back(a, b);
  *start = a;
  *stop = b;
  return true;
}

// Adds algorithmic bytes to a kernel family after the counts were read back.
void prof_bytes(Ctx& c, int id, int64_t bytes) {
  if (c.prof && c.prof->on) c.prof->e[id].bytes += bytes;
}

void prof_release(Ctx& c) {
  if (!c.prof) return;
  c.prof->clear();
  for (hipEvent_t ev : c.prof->pool) (void)hipEventDestroy(ev);
  delete c.prof;
  c.prof = nullptr;
}

Channel* channel_of(Ctx& c, int ch) {
  if (ch < 0 || ch > 1) {
    set_error("channel must be 0 or 1");
    return nullptr;
  }
  return &c.ch[ch];
}

}  // namespace scout

using namespace scout;

extern "C" {

const char* scout_last_error(void) { return g_err.c_str(); }

int scout_ctx_create(int gpu_id, const scout_budget* budget, scout_ctx** out) {
  if (!out) {
    set_error("null output pointer");
    return SCOUT_ERR_ARG;
  }
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    set_error("no HIP device available");
    return SCOUT_ERR_HIP;
  }
  if (gpu_id < 0 || gpu_id >= count) {
    set_error("gpu id out of range");
    return SCOUT_ERR_ARG;
  }
  SCOUT_TRY(hipSetDevice(gpu_id));
  SCOUT_TRY(hipFree(nullptr));
  SCOUT_CHECK(noalloc_init(gpu_id));
  scout_ctx* x = new scout_ctx();
  Ctx& c = x->c;
  c.dev = gpu_id;
  if (budget) c.budget = *budget;
  if (hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c.cs, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void**)&c.h_box, (kBoxSlots + kPathCount) * sizeof(int64_t), hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc((void**)&c.h_apply, kBoxSlots * sizeof(int64_t), hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void**)&c.d_box, (kBoxSlots + kPathCount) * sizeof(int64_t)) != hipSuccess ||
      hipMemset(c.d_box, 0, (kBoxSlots + kPathCount) * sizeof(int64_t)) != hipSuccess) {
    set_error("context set-up failed");
    delete x;
    return SCOUT_ERR_HIP;
  }
  c.d_paths = c.d_box + kBoxSlots;
  for (int k = 0; k < 2; k++) {
    Channel& ch = c.ch[k];
    if (hipEventCreateWithFlags(&ch.staged, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ch.consumed, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&ch.h2d0) != hipSuccess || hipEventCreate(&ch.h2d1) != hipSuccess) {
      set_error("event creation failed");
      delete x;
      return SCOUT_ERR_HIP;
    }
  }
  *out = x;
  return SCOUT_OK;
}

int scout_ctx_destroy(scout_ctx* x) {
  if (!x) return SCOUT_OK;
  Ctx& c = x->c;
  (void)hipSetDevice(c.dev);
  (void)hipDeviceSynchronize();
  graph_free(c);
  free_state(c);
  cc_release(c);
  mlp_release(c);
  mis_release(c);
  color_release(c);
  stream_release(c);
  snapshot_release(c);
  prof_release(c);
  for (int k = 0; k < 2; k++) {
    c.ch[k].h_uv.release();
    c.ch[k].d_uv.release();
    c.ch[k].d_ins.release();
    c.ch[k].d_done.release();
    for (hipEvent_t ev : {c.ch[k].staged, c.ch[k].consumed, c.ch[k].h2d0, c.ch[k].h2d1})
      if (ev) (void)hipEventDestroy(ev);
  }
  DevBuf<int32_t>* ls[] = {&c.la, &c.lb, &c.lc, &c.ld, &c.le, &c.lf, &c.lg, &c.lh};
  for (auto* l : ls) l->release();
  c.wa.release();
  c.wb.release();
  c.wc.release();
  c.lvlb.release();
  c.cc_bits.release();
  c.tmp.release();
  if (c.box_ev) (void)hipEventDestroy(c.box_ev);
  if (c.apply_ev) (void)hipEventDestroy(c.apply_ev);
  if (c.h_apply) (void)hipHostFree(c.h_apply);
  if (c.h_box) (void)hipHostFree(c.h_box);
  if (c.d_box) (void)hipFree(c.d_box);
  if (c.s) (void)hipStreamDestroy(c.s);
  if (c.cs) (void)hipStreamDestroy(c.cs);
  delete x;
  return SCOUT_OK;
}

int scout_device_name(scout_ctx* x, char* buf, int len) {
  SCOUT_CHECK(check_ctx(x));
  hipDeviceProp_t prop;
  SCOUT_TRY(hipGetDeviceProperties(&prop, x->c.dev));
  if (len > 0) {
    strncpy(buf, prop.name, (size_t)len - 1);
    buf[len - 1] = 0;
  }
  return SCOUT_OK;
}

int scout_mem_info(scout_ctx* x, int64_t* free_b, int64_t* total_b) {
  SCOUT_CHECK(check_ctx(x));
  size_t f = 0, t = 0;
  SCOUT_TRY(hipMemGetInfo(&f, &t));
  if (free_b) *free_b = (int64_t)f;
  if (total_b) *total_b = (int64_t)t;
  return SCOUT_OK;
}

int scout_graph_load(scout_ctx* x, int32_t n, int64_t entries, const void* nindex, int wide, const int32_t* nlist) {
  SCOUT_CHECK(check_ctx(x));
  Ctx& c = x->c;
  if (!nindex || (!nlist && entries > 0)) {
    set_error("graph_load: null array");
    return SCOUT_ERR_ARG;
  }
  // offsets, inserted-row heads and the n-sized scratch stay in HBM; the base entries too when
  // the budget holds them, else they stay in pinned host memory (storage-resident mode)
  int64_t need_small = ((int64_t)n + 1) * 16 + (int64_t)n * 12;
  int64_t need = need_small + entries * 4;
  bool host_resident = c.budget.hbm_bytes > 0 && need > c.budget.hbm_bytes;
  if (host_resident && need_small > c.budget.hbm_bytes) {
    set_error("graph_load: the offsets and per-vertex state exceed the HBM budget even with the base entries "
              "kept on the host");
    return SCOUT_ERR_BUDGET;
  }
  if (host_resident && c.budget.pinned_bytes > 0 && entries * 4 > c.budget.pinned_bytes) {
    set_error("graph_load: the base entries exceed both the HBM and the pinned-memory budget");
    return SCOUT_ERR_BUDGET;
  }
  free_state(c);
  c.hbm_used = 0;
  SCOUT_CHECK(graph_upload_egr(c, n, entries, nindex, wide, nlist, host_resident));
  // the n-sized scratch below counts itself as it is allocated
  c.hbm_used = ((int64_t)n + 1) * 16 + (host_resident ? 0 : entries * 4);
  SCOUT_CHECK(alloc_n(c, (void**)&c.stamp, 4));
  SCOUT_CHECK(alloc_n(c, (void**)&c.aux, 4));
  SCOUT_CHECK(alloc_n(c, (void**)&c.aux2, 4));
  SCOUT_TRY(hipMemsetAsync(c.stamp, 0, ((size_t)n + 1) * 4, c.s));
  SCOUT_TRY(hipMemsetAsync(c.aux, 0xff, ((size_t)n + 1) * 4, c.s));
  SCOUT_TRY(hipMemsetAsync(c.aux2, 0xff, ((size_t)n + 1) * 4, c.s));
  c.epoch = 0;
  SCOUT_TRY(hipStreamSynchronize(c.s));
  return SCOUT_OK;
}

int scout_graph_info(scout_ctx* x, int64_t* out) {
  SCOUT_CHECK(check_graph(x));
  Ctx& c = x->c;
  out[0] = c.g.n;
  out[1] = c.g.E;
  out[2] = c.g.live;
  DevBuf<int64_t> deg, sum;
  SCOUT_CHECK(deg.reserve((int64_t)c.g.n + 1));
  SCOUT_CHECK(sum.reserve(1));
  SCOUT_CHECK(graph_logical_degrees(c, deg.p));
  SCOUT_CHECK(reduce_sum_i64(c, deg.p, c.g.n, sum.p));
  int64_t total = 0;
  SCOUT_TRY(hipMemcpyAsync(&total, sum.p, sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  deg.release();
  sum.release();
  out[3] = total;
  out[4] = c.g.sorted ? 1 : 0;
  out[5] = c.g.host_bcol ? 1 : 0;
  out[6] = c.fetch_bytes;
  out[7] = (int64_t)(c.fetch_ms * 1e3);  // microseconds
  return SCOUT_OK;
}

int scout_graph_download(scout_ctx* x, int64_t* off, int32_t* col) {
  SCOUT_CHECK(check_graph(x));
  return graph_download(x->c, off, col);
}

int scout_graph_compact(scout_ctx* x) {
  SCOUT_CHECK(check_graph(x));
  return graph_compact(x->c);
}

int scout_graph_degrees(scout_ctx* x, int64_t* deg) {
  SCOUT_CHECK(check_graph(x));
  Ctx& c = x->c;
  DevBuf<int64_t> d;
  SCOUT_CHECK(d.reserve((int64_t)c.g.n + 1));
  SCOUT_CHECK(graph_logical_degrees(c, d.p));
  SCOUT_TRY(hipMemcpyAsync(deg, d.p, (size_t)c.g.n * sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  d.release();
  return SCOUT_OK;
}

int scout_state_upload(scout_ctx* x, int kind, const void* host, int64_t n) {
  SCOUT_CHECK(check_graph(x));
  Ctx& c = x->c;
  if (n != c.g.n || !host) {
    set_error("state_upload: size must equal the vertex count");
    return SCOUT_ERR_ARG;
  }
  SCOUT_CHECK(ensure_state(c, kind));
  void* dst = kind == SCOUT_STATE_MIS ? (void*)c.mis
              : kind == SCOUT_STATE_COLOR ? (void*)c.color
              : kind == SCOUT_STATE_LEVEL ? (void*)c.level
                                          : (void*)c.label;
  size_t elem = kind == SCOUT_STATE_MIS ? 1 : 4;
  if (kind == SCOUT_STATE_MIS) c.mis_version++;
  SCOUT_TRY(hipMemcpyAsync(dst, host, (size_t)n * elem, hipMemcpyHostToDevice, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  // uploaded levels are taken as exact for the current graph; their parents are derived
  if (kind == SCOUT_STATE_LEVEL) {
    SCOUT_CHECK(compute_parents(c, c.level, c.parent));
    c.level_seq = c.apply_seq;
  }
  if (kind == SCOUT_STATE_LABEL) {
    c.cc_components = -1;
    c.label_seq = c.apply_seq;
  }
  return SCOUT_OK;
}

int scout_state_download(scout_ctx* x, int kind, void* host, int64_t n) {
  SCOUT_CHECK(check_graph(x));
  Ctx& c = x->c;
  if (n != c.g.n || !host) {
    set_error("state_download: size must equal the vertex count");
    return SCOUT_ERR_ARG;
  }
  const void* src = kind == SCOUT_STATE_MIS ? (const void*)c.mis
                    : kind == SCOUT_STATE_COLOR ? (const void*)c.color
                    : kind == SCOUT_STATE_LEVEL ? (const void*)c.level
                                                : (const void*)c.label;
  if (!src) {
    set_error("state_download: state not initialised");
    return SCOUT_ERR_STATE;
  }
  size_t elem = kind == SCOUT_STATE_MIS ? 1 : 4;
  SCOUT_TRY(hipMemcpyAsync(host, src, (size_t)n * elem, hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  return SCOUT_OK;
}

int scout_set_source(scout_ctx* x, int32_t source) {
  SCOUT_CHECK(check_graph(x));
  if (source < 0 || source >= x->c.g.n) {
    set_error("source out of range");
    return SCOUT_ERR_ARG;
  }
  x->c.source = source;
  return SCOUT_OK;
}

int scout_stage_batch(scout_ctx* x, int channel, const int32_t* uv, int64_t m) {
  SCOUT_CHECK(check_graph(x));
  Ctx& c = x->c;
  Channel* ch = channel_of(c, channel);
  if (!ch || m < 0 || (m > 0 && !uv)) {
    if (ch) set_error("stage_batch: bad arguments");
    return SCOUT_ERR_ARG;
  }
  if (m > 0x7fffffffLL) {
    set_error("stage_batch: more than 2^31-1 lines");
    return SCOUT_ERR_ARG;
  }
  // The previous transfer from this channel's pinned buffer must be complete before refill.
  SCOUT_TRY(hipEventSynchronize(ch->staged));
  if (c.budget.pinned_bytes > 0 && 8 * m > c.budget.pinned_bytes) {
    set_error("stage_batch: batch exceeds the pinned-memory budget");
    return SCOUT_ERR_BUDGET;
  }
  SCOUT_CHECK(ch->h_uv.reserve(2 * m + 2));
  SCOUT_CHECK(ch->d_uv.reserve(2 * m + 2));
  SCOUT_CHECK(ch->d_ins.reserve(m + 1));
  SCOUT_CHECK(ch->d_done.reserve(m + 1));
  auto t0 = std::chrono::steady_clock::now();
  if (m > 0) memcpy(ch->h_uv.p, uv, (size_t)(2 * m) * sizeof(int32_t));
  ch->pack_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  // the compute stream may still read this channel's previous lines
  SCOUT_TRY(hipStreamWaitEvent(c.cs, ch->consumed, 0));
  SCOUT_TRY(hipEventRecord(ch->h2d0, c.cs));
  if (m > 0)
    SCOUT_TRY(hipMemcpyAsync(ch->d_uv.p, ch->h_uv.p, (size_t)(2 * m) * sizeof(int32_t), hipMemcpyHostToDevice, c.cs));
  SCOUT_TRY(hipEventRecord(ch->h2d1, c.cs));
  SCOUT_TRY(hipEventRecord(ch->staged, c.cs));
  ch->h2d_bytes = 8 * m;
  channel_staged(*ch, m);
  return SCOUT_OK;
}

int scout_stage_info(scout_ctx* x, int channel, double* pack_ms, double* h2d_ms, int64_t* bytes) {
  SCOUT_CHECK(check_ctx(x));
  Ctx& c = x->c;
  Channel* ch = channel_of(c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  SCOUT_TRY(hipEventSynchronize(ch->h2d1));
  float t = 0.f;
  SCOUT_TRY(hipEventElapsedTime(&t, ch->h2d0, ch->h2d1));
  if (pack_ms) *pack_ms = ch->pack_ms;
  if (h2d_ms) *h2d_ms = (double)t;
  if (bytes) *bytes = ch->h2d_bytes;
  return SCOUT_OK;
}

int scout_stage_batch_device(scout_ctx* x, int channel, const int32_t* d_uv, int64_t m) {
  SCOUT_CHECK(check_graph(x));
  Ctx& c = x->c;
  Channel* ch = channel_of(c, channel);
  if (!ch || m < 0) return SCOUT_ERR_ARG;
  SCOUT_TRY(hipEventSynchronize(ch->staged));
  SCOUT_CHECK(ch->d_uv.reserve(2 * m + 2));
  SCOUT_CHECK(ch->d_ins.reserve(m + 1));
  SCOUT_CHECK(ch->d_done.reserve(m + 1));
  SCOUT_TRY(hipStreamWaitEvent(c.cs, ch->consumed, 0));
  SCOUT_TRY(hipEventRecord(ch->h2d0, c.cs));
  if (m > 0)
    SCOUT_TRY(hipMemcpyAsync(ch->d_uv.p, d_uv, (size_t)(2 * m) * sizeof(int32_t), hipMemcpyDeviceToDevice, c.cs));
  SCOUT_TRY(hipEventRecord(ch->h2d1, c.cs));
  SCOUT_TRY(hipEventRecord(ch->staged, c.cs));
  ch->pack_ms = 0.0;
  ch->h2d_bytes = 0;
  channel_staged(*ch, m);
  return SCOUT_OK;
}

int scout_classify(scout_ctx* x, int channel, int rule, uint8_t* ins) {
  SCOUT_CHECK(check_graph(x));
  Ctx& c = x->c;
  Channel* ch = channel_of(c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  if (rule != SCOUT_CLASSIFY_CURRENT && rule != SCOUT_CLASSIFY_BASE) {
    set_error("classify: unknown rule");
    return SCOUT_ERR_ARG;
  }
  SCOUT_CHECK(channel_wait(c.s, *ch));
  SCOUT_CHECK(graph_classify(c, *ch, rule));
  SCOUT_CHECK(channel_used(c.s, *ch));
  if (ins && ch->m > 0) {
    SCOUT_TRY(hipMemcpyAsync(ins, ch->d_ins.p, (size_t)ch->m, hipMemcpyDeviceToHost, c.s));
    SCOUT_TRY(hipStreamSynchronize(c.s));
  }
  return SCOUT_OK;
}

int scout_apply(scout_ctx* x, int channel, int conditional) {
  SCOUT_CHECK(check_graph(x));
  Ctx& c = x->c;
  Channel* ch = channel_of(c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  SCOUT_CHECK(channel_wait(c.s, *ch));
  SCOUT_CHECK(graph_apply(c, *ch, conditional));
  return channel_used(c.s, *ch);
}

int scout_commit(scout_ctx* x, int channel, int strategy, scout_stats* bfs_st, scout_stats* cc_st) {
  SCOUT_CHECK(check_graph(x));
  Ctx& c = x->c;
  Channel* ch = channel_of(c, channel);
  if (!ch || !bfs_st || !cc_st) return SCOUT_ERR_ARG;
  if (!c.level || !c.label) {
    set_error("commit: levels and labels must be loaded first");
    return SCOUT_ERR_STATE;
  }
  SCOUT_CHECK(channel_wait(c.s, *ch));
  SCOUT_CHECK(graph_classify(c, *ch, SCOUT_CLASSIFY_CURRENT));
  // the update's counters are read with the BFS repair's first read-back, which settles it
  SCOUT_CHECK(graph_apply_deferred(c, *ch));
  // The CC scan reads only the final levels: it is queued behind each read-back of the BFS
  // relax, shut until the relax has ended, so the device goes from one repair to the next
  // without waiting for the host.
  Channel* chp = ch;
  c.after_relax = [&c, chp, strategy](const int64_t* gate, bool* queued) -> int {
    *queued = cc_scan_ahead_ok(c, *chp, strategy);
    return *queued ? cc_scan_queue(c, *chp, c.level, gate) : SCOUT_OK;
  };
  int rc = repair_bfs(c, *ch, strategy, bfs_st);
  c.after_relax = nullptr;
  bool complete;
  int rs = apply_settle(c, &complete);  // settled by the repair on every path; kept for errors
  SCOUT_CHECK(rc);
  SCOUT_CHECK(rs);
  SCOUT_CHECK(repair_cc(c, *ch, strategy, cc_st));
  return channel_used(c.s, *ch);
}

int scout_apply_segments(scout_ctx* x, int channel, int32_t nseg, const int64_t* seg_off, int rule,
                         int conditional, uint8_t* ins) {
  SCOUT_CHECK(check_graph(x));
  Ctx& c = x->c;
  Channel* ch = channel_of(c, channel);
  if (!ch || nseg < 0 || (nseg > 0 && !seg_off)) return SCOUT_ERR_ARG;
  if (rule != SCOUT_CLASSIFY_CURRENT && rule != SCOUT_CLASSIFY_BASE) {
    set_error("apply_segments: unknown rule");
    return SCOUT_ERR_ARG;
  }
  if (nseg > 0 && (seg_off[0] != 0 || seg_off[nseg] != ch->m)) {
    set_error("apply_segments: segments must cover the staged lines from 0 to the line count");
    return SCOUT_ERR_ARG;
  }
  for (int32_t k = 0; k < nseg; k++)
    if (seg_off[k] > seg_off[k + 1]) {
      set_error("apply_segments: segment offsets must not decrease");
      return SCOUT_ERR_ARG;
    }
  SCOUT_CHECK(channel_wait(c.s, *ch));
  c.apply_seq++;  // one commit: the repairs that follow see one batch of all segments
  // When no pair occurs twice in the whole range, no segment touches a pair of an earlier one:
  // each line's pair stands in the graph as it stood before the segments, so one
  // classification and one application of all lines (inserted entries in line order) equal
  // the segments applied one by one. Otherwise the segments go one by one.
  int64_t repeated = 1;
  if (nseg > 1) SCOUT_CHECK(graph_repeated_pairs(c, *ch, 0, ch->m, &repeated));
  if (nseg > 1 && repeated == 0) {
    SCOUT_CHECK(graph_classify_lines(c, *ch, 0, ch->m, rule));
    SCOUT_CHECK(graph_apply_lines(c, *ch, 0, ch->m, conditional));
  } else {
    for (int32_t k = 0; k < nseg; k++) {
      SCOUT_CHECK(graph_classify_lines(c, *ch, seg_off[k], seg_off[k + 1], rule));
      SCOUT_CHECK(graph_apply_lines(c, *ch, seg_off[k], seg_off[k + 1], conditional));
    }
  }
  SCOUT_CHECK(channel_used(c.s, *ch));
  if (ins && ch->m > 0) {
    SCOUT_TRY(hipMemcpyAsync(ins, ch->d_ins.p, (size_t)ch->m, hipMemcpyDeviceToHost, c.s));
    SCOUT_TRY(hipStreamSynchronize(c.s));
  }
  return SCOUT_OK;
}

int scout_prof_enable(scout_ctx* x, int on) {
  SCOUT_CHECK(check_ctx(x));
  Ctx& c = x->c;
  SCOUT_TRY(hipStreamSynchronize(c.s));
  if (!c.prof) c.prof = new ProfState();
  c.prof->clear();
  c.prof->on = on != 0;
  return SCOUT_OK;
}

int scout_prof_read(scout_ctx* x, int id, double* total_ms, int64_t* launches, int64_t* bytes) {
  SCOUT_CHECK(check_ctx(x));
  Ctx& c = x->c;
  if (id < 0 || id >= PROF_COUNT) {
    set_error("prof_read: unknown kernel id");
    return SCOUT_ERR_ARG;
  }
  double ms = 0.0;
  int64_t cnt = 0, by = 0;
  if (c.prof) {
    SCOUT_TRY(hipStreamSynchronize(c.s));
    for (auto& sp : c.prof->e[id].spans) {
      float t = 0.f;
      SCOUT_TRY(hipEventElapsedTime(&t, sp.first, sp.second));
      ms += t;
    }
    cnt = (int64_t)c.prof->e[id].spans.size();
    by = c.prof->e[id].bytes;
  }
  if (total_ms) *total_ms = ms;
  if (launches) *launches = cnt;
  if (bytes) *bytes = by;
  return SCOUT_OK;
}

int scout_touched_enable(scout_ctx* x, int on) {
  SCOUT_CHECK(check_graph(x));
  Ctx& c = x->c;
  if (on) {
    int64_t words = ((int64_t)c.g.n + 31) / 32;
    SCOUT_CHECK(c.touched_bits.reserve(words));
    SCOUT_TRY(hipMemsetAsync(c.touched_bits.p, 0, (size_t)words * sizeof(uint32_t), c.s));
    SCOUT_TRY(hipStreamSynchronize(c.s));
  }
  c.touched_on = on != 0;
  return SCOUT_OK;
}

int scout_touched_download(scout_ctx* x, int32_t* ids, int64_t cap, int64_t* count) {
  SCOUT_CHECK(check_graph(x));
  Ctx& c = x->c;
  if (!count || (cap > 0 && !ids)) {
    set_error("touched_download: null output");
    return SCOUT_ERR_ARG;
  }
  *count = 0;
  if (!c.touched_bits.p) return SCOUT_OK;
  int64_t words = ((int64_t)c.g.n + 31) / 32;
  std::vector<uint32_t> bits((size_t)words);
  SCOUT_TRY(hipStreamSynchronize(c.s));
  SCOUT_TRY(hipMemcpy(bits.data(), c.touched_bits.p, (size_t)words * sizeof(uint32_t), hipMemcpyDeviceToHost));
  int64_t k = 0;
  for (int64_t w = 0; w < words; w++)
    for (uint32_t b = bits[(size_t)w]; b; b &= b - 1) {
      int64_t v = 32 * w + __builtin_ctz(b);
      if (k < cap) ids[k] = (int32_t)v;
      k++;
    }
  *count = k;
  return SCOUT_OK;
}

int scout_path_counts(scout_ctx* x, int64_t* out, int count) {
  SCOUT_CHECK(check_ctx(x));
  Ctx& c = x->c;
  if (!out || count < 0 || count > kPathCount) {
    set_error("path_counts: count must be 0.." + std::to_string(kPathCount));
    return SCOUT_ERR_ARG;
  }
  SCOUT_TRY(hipStreamSynchronize(c.s));
  if (count > 0) SCOUT_TRY(hipMemcpy(out, c.d_paths, (size_t)count * sizeof(int64_t), hipMemcpyDeviceToHost));
  return SCOUT_OK;
}

int scout_sync(scout_ctx* x) {
  SCOUT_CHECK(check_ctx(x));
  SCOUT_TRY(hipStreamSynchronize(x->c.s));
  SCOUT_TRY(hipStreamSynchronize(x->c.cs));
  return SCOUT_OK;
}

}  // extern "C"
