// noalloc.hip: the full CC recompute into a caller's device array with no allocation, for
// callers that cannot call hipMalloc (code under stream capture, tight serving loops,
// harnesses that forbid it).
//
// scout_cc_full (cc.hip) takes its chunk-task list from a context buffer that grows with the
// graph and its counters from the context mailbox. This variant runs the same kernels with
// both in `scratch`, a file-scope device array of fixed size: one copy per GPU, held until the
// process ends. Host code never passes the array's name to a kernel (on the host it is only a
// stand-in for the symbol); it uses the device address hipGetSymbolAddress returned for the
// current device, kept per device in g_scratch.
#include <mutex>

#include "common.cuh"

// Chunk-task slots of the scratch array (int64 each) and the counters after them.
constexpr int64_t kScratchTasks = int64_t(1) << 21;  // 2M tasks: 16 MiB
constexpr int64_t kScratchCounters = 32;             // 256 B
constexpr int64_t kScratchWords = kScratchTasks + kScratchCounters;

static __device__ __align__(256) int64_t scratch[kScratchWords];

namespace scout {

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int64_t* g_scratch[kMaxDevices] = {};

}  // namespace

int noalloc_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("noalloc: device id beyond the supported range");
    return SCOUT_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev]) return SCOUT_OK;
  void* p = nullptr;
  SCOUT_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch[dev] = static_cast<int64_t*>(p);
  return SCOUT_OK;
}

}  // namespace scout

using namespace scout;

extern "C" {

int scout_cc_full_noalloc(scout_ctx* x, int32_t* d_label) {
  if (!x) {
    set_error("null context");
    return SCOUT_ERR_ARG;
  }
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  if (!c.g.boff) {
    set_error("no graph loaded");
    return SCOUT_ERR_STATE;
  }
  if (!d_label) {
    set_error("cc_full_noalloc: null label array");
    return SCOUT_ERR_ARG;
  }
  int64_t bound = cc_full_task_bound(c);
  if (bound > kScratchTasks) {
    set_error("cc_full_noalloc: the graph needs up to " + std::to_string(bound) + " chunk tasks, more than the " +
              std::to_string(kScratchTasks) + " of the fixed scratch (about 2^32 adjacency entries at most)");
    return SCOUT_ERR_ARG;
  }
  int64_t* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = c.dev < kMaxDevices ? g_scratch[c.dev] : nullptr;
  }
  if (!base) {
    set_error("cc_full_noalloc: scratch not resolved for this device");
    return SCOUT_ERR_STATE;
  }
  if (c.g.n == 0) return SCOUT_OK;
  int64_t* tasks = base;
  int64_t* counters = base + kScratchTasks;
  SCOUT_CHECK(cc_full_into(c, d_label, tasks, kScratchTasks, counters));
  // the bound above keeps the task count within the scratch; checked again from the counter
  SCOUT_TRY(hipMemcpyAsync(c.h_box, counters, sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  if (c.h_box[0] > kScratchTasks) {
    set_error("cc_full_noalloc: chunk task overflow");
    return SCOUT_ERR_STATE;
  }
  return SCOUT_OK;
}

}  // extern "C"
