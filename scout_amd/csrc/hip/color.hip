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
// The result depends only on the round sets, not on edit or row order.
//
// First-fit is a wave-wide bitmap: each lane ORs the colours of its slice of the row that fall
// in the window [base, base + 64) into a 64-bit mask, the wave OR-reduces it, and the first
// zero bit is the answer; a full window moves the window up by 64 (one extra row pass per 64
// colours in use).
#include "common.cuh"

namespace scout {

namespace {

constexpr int32_t kUnmarked = -2;

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

// Wave per round vertex: first-fit against the snapshot colours.
__global__ void k_color_propose(GView g, const int32_t* round, int64_t nr, const int32_t* color, int32_t* marker,
                                int32_t* prop, int64_t* scanned) {
  int64_t work = 0;
  for (int64_t i = global_wave(); i < nr; i += total_waves()) {
    int32_t v = round[i];
    Row r = row_of(g, v);
    int64_t len = row_len(r);
    int32_t base = 0;
    while (true) {
      unsigned long long mask = 0;
      for (int64_t k = lane_id(); k < len; k += kWave) {
        int32_t w = row_at(g, r, k);
        if (w >= 0) {
          int32_t cw = color[w] - base;
          if (cw >= 0 && cw < 64) mask |= 1ULL << cw;
        }
      }
      work += len;
      mask = wave_or(mask);
      if (~mask != 0ULL) {
        int32_t c = base + (__ffsll((long long)~mask) - 1);
        if (lane_id() == 0) {
          prop[i] = c;
          marker[v] = c;
        }
        break;
      }
      base += 64;
    }
  }
  if (lane_id() == 0 && work) atomicAdd((unsigned long long*)scanned, (unsigned long long)work);
}

// Wave per round vertex: lose iff a lower-id neighbour proposed the same colour.
__global__ void k_color_detect(GView g, const int32_t* round, int64_t nr, const int32_t* marker, const int32_t* prop,
                               int32_t* retry, int64_t* retry_count, int64_t cap, int64_t* scanned) {
  int64_t work = 0;
  for (int64_t i = global_wave(); i < nr; i += total_waves()) {
    int32_t v = round[i];
    int32_t mine = prop[i];
    Row r = row_of(g, v);
    int64_t len = row_len(r);
    work += len;
    int64_t hit = wave_find_first(g, r, 0, len, [&](int32_t w) { return w < v && marker[w] == mine; });
    if (hit >= 0 && lane_id() == 0) thread_append(v, retry, retry_count, cap);
  }
  if (lane_id() == 0 && work) atomicAdd((unsigned long long*)scanned, (unsigned long long)work);
}

__global__ void k_color_apply(const int32_t* round, int64_t nr, const int32_t* prop, int32_t* color, int32_t* marker) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  int32_t v = round[i];
  color[v] = prop[i];
  marker[v] = kUnmarked;
}

}  // namespace

int color_repair(Ctx& c, Channel& ch, scout_stats* st) {
  int32_t n = c.g.n;
  int64_t m = ch.m;
  SCOUT_CHECK(ensure_state(c, SCOUT_STATE_COLOR));
  SCOUT_CHECK(ensure_scratch(c, &c.col_marker, kUnmarked));
  SCOUT_CHECK(c.la.reserve((int64_t)n + 1));
  SCOUT_CHECK(c.lb.reserve((int64_t)n + 1));
  SCOUT_CHECK(c.lc.reserve((int64_t)n + 1));
  int64_t cap = (int64_t)n + 1;
  int64_t* box = c.d_box;
  SCOUT_TRY(hipMemsetAsync(box, 0, 8 * sizeof(int64_t), c.s));
  int32_t epoch = next_epoch(c);
  if (m > 0)
    k_color_marks<<<grid_for(m), kBlock, 0, c.s>>>(ch.d_uv.p, ch.d_ins.p, m, c.color, c.la.p, cap, c.stamp, epoch, box);
  SCOUT_CHECK(box_read(c, 3));
  int64_t nr = c.h_box[0];
  st->conflicts += c.h_box[1];
  st->insertions += c.h_box[2];
  st->deletions += m - c.h_box[2];
  st->seeds += nr;
  int32_t* round = c.la.p;
  int32_t* retry = c.lc.p;
  int32_t* prop = c.lb.p;
  int64_t rounds = 0;
  while (nr > 0) {
    rounds++;
    st->touched += nr;
    SCOUT_TRY(hipMemsetAsync(box + 4, 0, sizeof(int64_t), c.s));
    k_color_propose<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), round, nr, c.color, c.col_marker, prop, box + 5);
    k_color_detect<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), round, nr, c.col_marker, prop, retry, box + 4, cap,
                                                  box + 5);
    k_color_apply<<<grid_for(nr), kBlock, 0, c.s>>>(round, nr, prop, c.color, c.col_marker);
    SCOUT_TRY(hipMemcpyAsync(c.h_box + 4, box + 4, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
    SCOUT_TRY(hipStreamSynchronize(c.s));
    nr = c.h_box[4];
    std::swap(round, retry);
  }
  st->rounds += rounds;
  st->edges_scanned += rounds > 0 ? c.h_box[5] : 0;
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
  Span span;
  SCOUT_CHECK(span.begin(c.s));
  st->lines += ch->m;
  SCOUT_CHECK(color_repair(c, *ch, st));
  SCOUT_CHECK(span.end(c.s));
  st->repair_ms += span.ms();
  return SCOUT_OK;
}
