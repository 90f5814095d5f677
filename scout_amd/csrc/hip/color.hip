// color.hip: colouring repair (LCL, radius 1), restating src/coloring/scout_coloring.cu:46-100
// and :243-277 without the per-round truncation at 200,000 vertices / 2e7 entries (:181-183).
//
// Marks: the larger endpoint of every insertion line whose endpoints had equal colours before
// the batch (a set). Rounds until no loser remains:
//   propose  smallest colour >= 0 absent from the neighbours' snapshot colours, published in
//            marker[v] (propose_color_kernel, :46-64)
//   detect   v loses iff a lower-id round neighbour proposed the same colour
//            (detect_conflict_kernel, :69-87)
//   apply    every proposal is written (apply_color_kernel, :89-93); markers are cleared
// The result depends only on the round sets, not on edit or row order, nor on how rows are
// split between threads, waves and blocks.
//
// Rows are tiered by length: up to `thread_max` entries one thread, up to `block_min` one wave,
// longer rows one block. First fit over a wave or block ORs the colours of each lane's slice
// that fall in the window [base, base + 64) into a 64-bit mask, reduces the masks and takes
// the first zero bit; a full window moves up by 64. The thresholds and the grid size are
// runtime settings (scout_color_config), chosen by timing in scout_color_autotune.
//
// Rounds are device-driven: every kernel reads its list length from device memory and a
// one-thread kernel does the round bookkeeping, so a group of rounds (list A -> B -> A ...)
// can be captured once in a hipGraph and replayed until the list is empty, with one read-back
// per replay. With the graph off, the same kernels are launched one by one.
#include <vector>

#include "common.cuh"

namespace scout {

struct ColorState {
  int thread_max = 8;          // rows up to this length: one thread
  int block_min = 4096;        // rows longer than this: one block
  unsigned grid = kListGrid;   // blocks of the list kernels
  bool use_graph = true;
  int group_rounds = 2;        // rounds per graph replay (even)
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> key;  // what the captured graph was built with
  DevBuf<int32_t> wave_list, block_list;
};

namespace {

constexpr int32_t kUnmarked = -2;

// box layout of the colouring repair: [0] round count A, [1] conflicts, [2] insertion lines,
// [3] round count B, [4] wave-list count, [5] block-list count, [6] round vertices over all
// rounds, [7] non-empty rounds, [8] scanned entries
constexpr int kNrA = 0, kNrB = 3, kWc = 4, kBc = 5, kTotal = 6, kRounds = 7, kScan = 8;

__device__ __forceinline__ unsigned long long wave_or(unsigned long long x) {
  for (int off = kWave / 2; off > 0; off >>= 1) x |= __shfl_xor(x, off);
  return x;
}

// Insertion lines with equal pre-batch colours mark max(u, v) once (stamp dedup).
// box[0] = round size, box[1] = conflicts, box[2] = insertion lines.
__global__ void k_color_marks(const int32_t* uv, const uint8_t* ins, int64_t m, const int32_t* color, int32_t* round,
                              int64_t cap, int32_t* stamp, int32_t epoch, int64_t* box) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool is_ins = i < m && ins[i];
  bool conflict = false, take = false;
  int32_t loser = 0;
  if (is_ins) {
    int32_t u = uv[2 * i], v = uv[2 * i + 1];
    if (color[u] == color[v]) {
      conflict = true;
      loser = u > v ? u : v;
      take = atomicExch(&stamp[loser], epoch) != epoch;
    }
  }
  block_append(take, loser, round, &box[0], cap);
  block_add(conflict ? 1 : 0, &box[1]);
  block_add(is_ins ? 1 : 0, &box[2]);
}

// Round bookkeeping: adds the round size to the totals, zeroes the next round's count and the
// tier lists.
__global__ void k_color_round_begin(const int64_t* nr, int64_t* next_count, int64_t* box) {
  int64_t r = *nr;
  box[kTotal] += r;
  if (r > 0) box[kRounds] += 1;
  *next_count = 0;
  box[kWc] = 0;
  box[kBc] = 0;
}

// Colour bit of w in the window [base, base + 64), or 0.
__device__ __forceinline__ unsigned long long color_bit(const int32_t* color, int32_t w, int32_t base) {
  if (w < 0) return 0;
  int32_t cw = color[w] - base;
  return (cw >= 0 && cw < 64) ? 1ULL << cw : 0ULL;
}

// Thread per round vertex: short rows propose here, longer rows are listed for waves/blocks.
__global__ void k_color_propose_t(GView g, const int32_t* round, const int64_t* nr_p, const int32_t* color,
                                  int32_t* marker, int32_t* prop, int thread_max, int block_min, int32_t* wlist,
                                  int32_t* blist, int64_t* box, int64_t cap) {
  int64_t nr = *nr_p;
  int64_t work = 0;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nr; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = base + threadIdx.x;
    bool to_wave = false, to_block = false;
    if (i < nr) {
      int32_t v = round[i];
      Row r = row_of(g, v);
      int64_t len = row_len(r);
      if (len <= thread_max) {
        int32_t cbase = 0;
        while (true) {
          unsigned long long mask = 0;
          for (int64_t k = 0; k < len; k++) mask |= color_bit(color, row_at(g, r, k), cbase);
          work += len;
          if (~mask != 0ULL) {
            int32_t c = cbase + (__ffsll((long long)~mask) - 1);
            prop[i] = c;
            marker[v] = c;
            break;
          }
          cbase += 64;
        }
      } else if (len <= block_min) {
        to_wave = true;
      } else {
        to_block = true;
      }
    }
    int32_t idx = (int32_t)(i < nr ? i : 0);
    wave_append(to_wave, idx, wlist, &box[kWc], cap);
    wave_append(to_block, idx, blist, &box[kBc], cap);
  }
  work = wave_sum(work);
  if (lane_id() == 0 && work) atomicAdd((unsigned long long*)&box[kScan], (unsigned long long)work);
}

// Wave per listed round index.
__global__ void k_color_propose_w(GView g, const int32_t* round, const int32_t* wlist, const int64_t* wcount,
                                  const int32_t* color, int32_t* marker, int32_t* prop, int64_t* scanned) {
  int64_t nw = *wcount;
  int64_t work = 0;
  for (int64_t t = global_wave(); t < nw; t += total_waves()) {
    int64_t i = wlist[t];
    int32_t v = round[i];
    Row r = row_of(g, v);
    int64_t len = row_len(r);
    int32_t cbase = 0;
    while (true) {
      unsigned long long mask = 0;
      for (int64_t k = lane_id(); k < len; k += kWave) mask |= color_bit(color, row_at(g, r, k), cbase);
      work += len;
      mask = wave_or(mask);
      if (~mask != 0ULL) {
        int32_t c = cbase + (__ffsll((long long)~mask) - 1);
        if (lane_id() == 0) {
          prop[i] = c;
          marker[v] = c;
        }
        break;
      }
      cbase += 64;
    }
  }
  if (lane_id() == 0 && work) atomicAdd((unsigned long long*)scanned, (unsigned long long)work);
}

// Block per listed round index (the longest rows).
__global__ void k_color_propose_b(GView g, const int32_t* round, const int32_t* blist, const int64_t* bcount,
                                  const int32_t* color, int32_t* marker, int32_t* prop, int64_t* scanned) {
  __shared__ unsigned long long part[kWavesPerBlock];
  int64_t nb = *bcount;
  int64_t work = 0;
  int w = threadIdx.x / kWave;
  for (int64_t t = blockIdx.x; t < nb; t += gridDim.x) {
    int64_t i = blist[t];
    int32_t v = round[i];
    Row r = row_of(g, v);
    int64_t len = row_len(r);
    int32_t cbase = 0;
    while (true) {
      unsigned long long mask = 0;
      for (int64_t k = threadIdx.x; k < len; k += blockDim.x) mask |= color_bit(color, row_at(g, r, k), cbase);
      work += len;
      mask = wave_or(mask);
      if (lane_id() == 0) part[w] = mask;
      __syncthreads();
      unsigned long long all = 0;
      for (int k = 0; k < kWavesPerBlock; k++) all |= part[k];
      __syncthreads();
      if (~all != 0ULL) {
        int32_t c = cbase + (__ffsll((long long)~all) - 1);
        if (threadIdx.x == 0) {
          prop[i] = c;
          marker[v] = c;
        }
        break;
      }
      cbase += 64;
    }
  }
  if (threadIdx.x == 0 && work) atomicAdd((unsigned long long*)scanned, (unsigned long long)work);
}

// Thread per round vertex: v loses iff a lower-id neighbour proposed the same colour. Only
// rows of at most thread_max entries are checked here.
__global__ void k_color_detect_t(GView g, const int32_t* round, const int64_t* nr_p, const int32_t* marker,
                                 const int32_t* prop, int thread_max, int32_t* retry, int64_t* retry_count,
                                 int64_t cap) {
  int64_t nr = *nr_p;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nr; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = base + threadIdx.x;
    bool lose = false;
    int32_t v = 0;
    if (i < nr) {
      v = round[i];
      Row r = row_of(g, v);
      int64_t len = row_len(r);
      if (len <= thread_max) {
        int32_t mine = prop[i];
        for (int64_t k = 0; k < len && !lose; k++) {
          int32_t w = row_at(g, r, k);
          lose = w >= 0 && w < v && marker[w] == mine;
        }
      }
    }
    wave_append(lose, v, retry, retry_count, cap);
  }
}

__global__ void k_color_detect_w(GView g, const int32_t* round, const int32_t* wlist, const int64_t* wcount,
                                 const int32_t* marker, const int32_t* prop, int32_t* retry, int64_t* retry_count,
                                 int64_t cap) {
  int64_t nw = *wcount;
  for (int64_t t = global_wave(); t < nw; t += total_waves()) {
    int64_t i = wlist[t];
    int32_t v = round[i];
    int32_t mine = prop[i];
    Row r = row_of(g, v);
    int64_t hit = wave_find_first(g, r, 0, row_len(r), [&](int32_t w) { return w < v && marker[w] == mine; });
    if (hit >= 0 && lane_id() == 0) thread_append(v, retry, retry_count, cap);
  }
}

__global__ void k_color_detect_b(GView g, const int32_t* round, const int32_t* blist, const int64_t* bcount,
                                 const int32_t* marker, const int32_t* prop, int32_t* retry, int64_t* retry_count,
                                 int64_t cap) {
  __shared__ int hit_s;
  int64_t nb = *bcount;
  for (int64_t t = blockIdx.x; t < nb; t += gridDim.x) {
    int64_t i = blist[t];
    int32_t v = round[i];
    int32_t mine = prop[i];
    Row r = row_of(g, v);
    int64_t len = row_len(r);
    if (threadIdx.x == 0) hit_s = 0;
    __syncthreads();
    for (int64_t k0 = 0; k0 < len; k0 += blockDim.x) {
      int64_t k = k0 + threadIdx.x;
      if (k < len) {
        int32_t w = row_at(g, r, k);
        if (w >= 0 && w < v && marker[w] == mine) hit_s = 1;
      }
      __syncthreads();
      int h = hit_s;
      __syncthreads();
      if (h) break;
    }
    if (threadIdx.x == 0 && hit_s) thread_append(v, retry, retry_count, cap);
    __syncthreads();
  }
}

__global__ void k_color_apply(const int32_t* round, const int64_t* nr_p, const int32_t* prop, int32_t* color,
                              int32_t* marker) {
  int64_t nr = *nr_p;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t v = round[i];
    color[v] = prop[i];
    marker[v] = kUnmarked;
  }
}

struct RoundBufs {
  int32_t* la;
  int32_t* lb;
  int32_t* prop;
  int64_t cap;
};

// One round: list `round` (count at box[nr]) -> list `retry` (count at box[nr_next]).
void launch_round(Ctx& c, ColorState& s, const RoundBufs& b, bool a_to_b) {
  int64_t* box = c.d_box;
  GView g = c.g.view();
  const int32_t* round = a_to_b ? b.la : b.lb;
  int32_t* retry = a_to_b ? b.lb : b.la;
  int64_t* nr = box + (a_to_b ? kNrA : kNrB);
  int64_t* next = box + (a_to_b ? kNrB : kNrA);
  unsigned gl = s.grid;
  k_color_round_begin<<<1, 1, 0, c.s>>>(nr, next, box);
  k_color_propose_t<<<gl, kBlock, 0, c.s>>>(g, round, nr, c.color, c.col_marker, b.prop, s.thread_max, s.block_min,
                                            s.wave_list.p, s.block_list.p, box, b.cap);
  k_color_propose_w<<<gl, kBlock, 0, c.s>>>(g, round, s.wave_list.p, box + kWc, c.color, c.col_marker, b.prop,
                                            box + kScan);
  k_color_propose_b<<<gl, kBlock, 0, c.s>>>(g, round, s.block_list.p, box + kBc, c.color, c.col_marker, b.prop,
                                            box + kScan);
  k_color_detect_t<<<gl, kBlock, 0, c.s>>>(g, round, nr, c.col_marker, b.prop, s.thread_max, retry, next, b.cap);
  k_color_detect_w<<<gl, kBlock, 0, c.s>>>(g, round, s.wave_list.p, box + kWc, c.col_marker, b.prop, retry, next,
                                           b.cap);
  k_color_detect_b<<<gl, kBlock, 0, c.s>>>(g, round, s.block_list.p, box + kBc, c.col_marker, b.prop, retry, next,
                                           b.cap);
  k_color_apply<<<gl, kBlock, 0, c.s>>>(round, nr, b.prop, c.color, c.col_marker);
}

void launch_group(Ctx& c, ColorState& s, const RoundBufs& b) {
  for (int k = 0; k < s.group_rounds; k++) launch_round(c, s, b, k % 2 == 0);
}

// Everything the captured kernels hold by value: buffers, the row arrays and the settings.
std::vector<const void*> graph_key(Ctx& c, ColorState& s, const RoundBufs& b) {
  GView g = c.g.view();
  return {b.la, b.lb, b.prop, (const void*)(intptr_t)b.cap, c.color, c.col_marker, c.d_box, s.wave_list.p,
          s.block_list.p, g.boff, g.bcol, g.ahead, g.acol, g.bsamp, g.touched, (const void*)(intptr_t)g.n,
          (const void*)(intptr_t)s.thread_max, (const void*)(intptr_t)s.block_min, (const void*)(intptr_t)s.grid,
          (const void*)(intptr_t)s.group_rounds};
}

void drop_graph(ColorState& s) {
  if (s.exec) (void)hipGraphExecDestroy(s.exec);
  s.exec = nullptr;
  s.key.clear();
}

// Captures one group of rounds, unless the graph in hand was built from the same key.
int color_graph(Ctx& c, ColorState& s, const RoundBufs& b) {
  std::vector<const void*> key = graph_key(c, s, b);
  if (s.exec && key == s.key) return SCOUT_OK;
  drop_graph(s);
  hipGraph_t graph = nullptr;
  SCOUT_TRY(hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal));
  launch_group(c, s, b);
  hipError_t e = hipStreamEndCapture(c.s, &graph);
  if (e != hipSuccess) {
    set_error(std::string("color: graph capture failed: ") + hipGetErrorString(e));
    return SCOUT_ERR_HIP;
  }
  e = hipGraphInstantiate(&s.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    s.exec = nullptr;
    set_error(std::string("color: graph instantiation failed: ") + hipGetErrorString(e));
    return SCOUT_ERR_HIP;
  }
  s.key = key;
  return SCOUT_OK;
}

}  // namespace

void color_release(Ctx& c) {
  if (!c.color_state) return;
  drop_graph(*c.color_state);
  c.color_state->wave_list.release();
  c.color_state->block_list.release();
  delete c.color_state;
  c.color_state = nullptr;
}

int color_repair(Ctx& c, Channel& ch, scout_stats* st) {
  if (!c.color_state) c.color_state = new ColorState();
  ColorState& s = *c.color_state;
  int32_t n = c.g.n;
  int64_t m = ch.m;
  SCOUT_CHECK(ensure_state(c, SCOUT_STATE_COLOR));
  SCOUT_CHECK(ensure_scratch(c, &c.col_marker, kUnmarked));
  int64_t cap = (int64_t)n + 1;
  SCOUT_CHECK(c.la.reserve(cap));
  SCOUT_CHECK(c.lb.reserve(cap));
  SCOUT_CHECK(c.lc.reserve(cap));
  SCOUT_CHECK(s.wave_list.reserve(cap));
  SCOUT_CHECK(s.block_list.reserve(cap));
  int64_t* box = c.d_box;
  SCOUT_TRY(hipMemsetAsync(box, 0, 16 * sizeof(int64_t), c.s));
  int32_t epoch = next_epoch(c);
  if (m > 0)
    k_color_marks<<<grid_for(m), kBlock, 0, c.s>>>(ch.d_uv.p, ch.d_ins.p, m, c.color, c.la.p, cap, c.stamp, epoch, box);
  SCOUT_CHECK(box_read(c, 3));
  int64_t nr = c.h_box[0];
  st->conflicts += c.h_box[1];
  st->insertions += c.h_box[2];
  st->deletions += m - c.h_box[2];
  st->seeds += nr;
  if (nr == 0) return SCOUT_OK;
  SCOUT_CHECK(prof_begin(c, PROF_COLOR));
  RoundBufs b{c.la.p, c.lb.p, c.lc.p, cap};
  if (s.use_graph) SCOUT_CHECK(color_graph(c, s, b));
  while (true) {
    if (s.use_graph) SCOUT_TRY(hipGraphLaunch(s.exec, c.s));
    else launch_group(c, s, b);
    // a group has an even number of rounds, so the live list is A again
    SCOUT_TRY(hipMemcpyAsync(c.h_box, box, 9 * sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
    SCOUT_TRY(hipStreamSynchronize(c.s));
    c.readbacks++;
    if (c.h_box[kNrA] == 0) break;
  }
  // per round vertex its id, row offsets, proposal, marker and colour (~36 B); per scanned
  // entry the neighbour id and its colour or marker (8 B)
  SCOUT_CHECK(prof_end(c, PROF_COLOR, 36 * c.h_box[kTotal] + 8 * c.h_box[kScan]));
  st->touched += c.h_box[kTotal];
  st->rounds += c.h_box[kRounds];
  st->edges_scanned += c.h_box[kScan];
  return SCOUT_OK;
}

}  // namespace scout

using namespace scout;

extern "C" int scout_repair_color(scout_ctx* x, int channel, scout_stats* st) {
  if (!x || !st) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  Channel* ch = channel_of(c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  if (!c.g.boff || !c.color) {
    set_error("repair_color: graph and colours must be loaded first");
    return SCOUT_ERR_STATE;
  }
  TouchScope touch(c);
  Span span(c);
  SCOUT_CHECK(span.begin(c.s));
  st->lines += ch->m;
  int64_t rb0 = c.readbacks;
  SCOUT_CHECK(color_repair(c, *ch, st));
  SCOUT_CHECK(span.end(c.s));
  st->repair_ms += span.ms();
  st->readbacks += c.readbacks - rb0;
  return channel_used(c.s, *ch);
}

extern "C" int scout_color_config(scout_ctx* x, int thread_max, int block_min, int grid, int use_graph) {
  if (!x || thread_max < 0 || block_min < thread_max || grid < 1 || grid > 65535) {
    set_error("color_config: need 0 <= thread_max <= block_min and 1 <= grid <= 65535");
    return SCOUT_ERR_ARG;
  }
  Ctx& c = x->c;
  if (!c.color_state) c.color_state = new ColorState();
  c.color_state->thread_max = thread_max;
  c.color_state->block_min = block_min;
  c.color_state->grid = (unsigned)grid;
  c.color_state->use_graph = use_graph != 0;
  return SCOUT_OK;
}

extern "C" int scout_color_get_config(scout_ctx* x, int* out4) {
  if (!x || !out4) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  if (!c.color_state) c.color_state = new ColorState();
  out4[0] = c.color_state->thread_max;
  out4[1] = c.color_state->block_min;
  out4[2] = (int)c.color_state->grid;
  out4[3] = c.color_state->use_graph ? 1 : 0;
  return SCOUT_OK;
}

// Times the colouring repair of the staged, already applied batch under each candidate
// setting (colours restored before every run, one untimed run per setting to build its graph)
// and keeps the fastest. Leaves the colours as they were.
extern "C" int scout_color_autotune(scout_ctx* x, int channel, double* best_ms) {
  if (!x) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  Channel* ch = channel_of(c, channel);
  if (!ch || !c.g.boff || !c.color) {
    set_error("color_autotune: graph, colours and a staged batch are needed");
    return SCOUT_ERR_STATE;
  }
  if (!c.color_state) c.color_state = new ColorState();
  ColorState& s = *c.color_state;
  int32_t n = c.g.n;
  DevBuf<int32_t> saved;
  SCOUT_CHECK(saved.reserve(n));
  SCOUT_TRY(hipMemcpyAsync(saved.p, c.color, (size_t)n * 4, hipMemcpyDeviceToDevice, c.s));
  const int tms[] = {0, 4, 8, 16, 32};
  const int bms[] = {1024, 8192};
  const unsigned grids[] = {1024, 2048, 4096};
  double best = 1e30;
  int keep_tm = s.thread_max, keep_bm = s.block_min;
  unsigned keep_grid = s.grid;
  // device time of each trial (events on the stream around the repair alone)
  hipEvent_t ev[2];
  SCOUT_TRY(hipEventCreate(&ev[0]));
  SCOUT_TRY(hipEventCreate(&ev[1]));
  for (int tm : tms)
    for (int bm : bms)
      for (unsigned gr : grids) {
        s.thread_max = tm;
        s.block_min = bm;
        s.grid = gr;
        for (int rep = 0; rep < 2; rep++) {
          SCOUT_TRY(hipMemcpyAsync(c.color, saved.p, (size_t)n * 4, hipMemcpyDeviceToDevice, c.s));
          scout_stats st{};
          SCOUT_TRY(hipEventRecord(ev[0], c.s));
          SCOUT_CHECK(color_repair(c, *ch, &st));
          SCOUT_TRY(hipEventRecord(ev[1], c.s));
          SCOUT_TRY(hipEventSynchronize(ev[1]));
          float t = 0.f;
          SCOUT_TRY(hipEventElapsedTime(&t, ev[0], ev[1]));
          double ms = (double)t;
          if (rep == 1 && ms < best) {
            best = ms;
            keep_tm = tm;
            keep_bm = bm;
            keep_grid = gr;
          }
        }
      }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  s.thread_max = keep_tm;
  s.block_min = keep_bm;
  s.grid = keep_grid;
  SCOUT_TRY(hipMemcpyAsync(c.color, saved.p, (size_t)n * 4, hipMemcpyDeviceToDevice, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  saved.release();
  if (best_ms) *best_ms = best;
  return SCOUT_OK;
}
