// bfs.hip: BFS levels from the source, kept exact after every batch.
//
// The reference repairs one edit at a time: an improving insertion starts CAS-min relax
// rounds, a deletion between consecutive levels starts Jacobi recompute rounds over the
// child's dependants, and a cascade past b = 200 vertices is deferred to one full BFS after
// the stream (src/bfs/scout_bfs.cu:436-515). Here one batch is one device job whose result
// equals textbook BFS on the updated graph:
//   1. candidates: the deeper endpoint of every deleted consecutive-level pair;
//   2. invalidation, level by level: a candidate at level l with no valid neighbour at
//      level l-1 is invalid, and its neighbours at level l+1 become candidates;
//   3. every invalid vertex restarts at 1 + the smallest valid neighbour level (or
//      unreached);
//   4. insertions that now join u and v lower the far endpoint by CAS-min (gpu_relax,
//      src/bfs/scout_bfs.cu:46-54);
//   5. level-synchronous relax from all lowered vertices, processed in increasing level so
//      each vertex is expanded once, at its final level.
// Valid vertices keep a parent chain of valid vertices, so their old level is still an
// upper bound; relaxing every vertex whose value dropped, in level order, is Dijkstra with
// unit weights, hence exact.
#include "common.cuh"

namespace scout {

namespace {

// Wave per frontier vertex: items with row length <= kChunk are walked whole; longer rows
// are split into (vertex, chunk) tasks for k_relax_tasks. Frontier entries whose level is no
// longer `lvl` were lowered after being queued and are skipped.
__global__ void k_relax(GView g, int32_t* level, const int32_t* front, const int64_t* count, int32_t lvl,
                        int32_t* next, int64_t* next_count, int64_t cap, int32_t* stamp, int32_t epoch,
                        int64_t* tasks, int64_t* task_count, int64_t task_cap, int64_t* scanned) {
  int64_t nf = *count;
  int64_t work = 0;
  for (int64_t i = global_wave(); i < nf; i += total_waves()) {
    int32_t f = front[i];
    if (level[f] != lvl) continue;
    Row r = row_of(g, f);
    int64_t len = row_len(r);
    if (len > kChunk) {
      if (lane_id() == 0) {
        int64_t chunks = (len + kChunk - 1) / kChunk;
        int64_t pos = (int64_t)atomicAdd((unsigned long long*)task_count, (unsigned long long)chunks);
        for (int64_t k = 0; k < chunks && pos + k < task_cap; k++) tasks[pos + k] = ((int64_t)f << 32) | k;
      }
      continue;
    }
    work += len;
    for (int64_t base = 0; base < len; base += kWave) {
      int64_t k = base + lane_id();
      bool take = false;
      int32_t w = -1;
      if (k < len) {
        w = row_at(g, r, k);
        if (w >= 0 && cas_min_level(&level[w], lvl + 1)) take = atomicExch(&stamp[w], epoch) != epoch;
      }
      wave_append(take, w, next, next_count, cap);
    }
  }
  if (lane_id() == 0 && work) atomicAdd((unsigned long long*)scanned, (unsigned long long)work);
}

__global__ void k_relax_tasks(GView g, int32_t* level, const int64_t* tasks, const int64_t* task_count,
                              int64_t task_cap, int32_t lvl, int32_t* next, int64_t* next_count, int64_t cap,
                              int32_t* stamp, int32_t epoch, int64_t* scanned) {
  int64_t nt = min(*task_count, task_cap);
  int64_t work = 0;
  for (int64_t i = global_wave(); i < nt; i += total_waves()) {
    int64_t t = tasks[i];
    int32_t f = (int32_t)(t >> 32);
    int64_t chunk = t & 0xffffffffLL;
    Row r = row_of(g, f);
    int64_t lo = chunk * kChunk, hi = lo + kChunk;
    int64_t len = row_len(r);
    if (hi > len) hi = len;
    work += hi - lo;
    for (int64_t base = lo; base < hi; base += kWave) {
      int64_t k = base + lane_id();
      bool take = false;
      int32_t w = -1;
      if (k < hi) {
        w = row_at(g, r, k);
        if (w >= 0 && cas_min_level(&level[w], lvl + 1)) take = atomicExch(&stamp[w], epoch) != epoch;
      }
      wave_append(take, w, next, next_count, cap);
    }
  }
  if (lane_id() == 0 && work) atomicAdd((unsigned long long*)scanned, (unsigned long long)work);
}

// Deeper endpoint of each deleted pair whose levels were consecutive.
// box[0] = candidates, box[1] = min level, box[2] = max level, box[3] = deletion lines.
__global__ void k_del_candidates(const int32_t* uv, const uint8_t* ins, int64_t m, const int32_t* level,
                                 int32_t* cand, int64_t cap, int64_t* box) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || ins[i]) return;
  atomicAdd((unsigned long long*)&box[3], 1ULL);
  int32_t u = uv[2 * i], v = uv[2 * i + 1];
  int32_t lu = level[u], lv = level[v];
  int32_t c = -1, lc = 0;
  if (lu >= 0 && lv == lu + 1) c = v, lc = lv;
  else if (lv >= 0 && lu == lv + 1) c = u, lc = lu;
  if (c < 0) return;
  thread_append(c, cand, &box[0], cap);
  atomicMin((long long*)&box[1], (long long)lc);
  atomicMax((long long*)&box[2], (long long)lc);
}

// Candidates whose level is `lvl`, appended to the current list once (stamp dedup).
__global__ void k_cand_at_level(const int32_t* cand, int64_t nc, const int32_t* level, int32_t lvl, int32_t* cur,
                                int64_t* cur_count, int64_t cap, int32_t* stamp, int32_t epoch) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  int32_t c = cand[i];
  if (level[c] == lvl && atomicExch(&stamp[c], epoch) != epoch) thread_append(c, cur, cur_count, cap);
}

// Wave per candidate at level lvl: keep it if some neighbour at lvl-1 is still valid,
// otherwise mark it invalid and queue its neighbours at lvl+1.
__global__ void k_invalidate(GView g, const int32_t* level, int32_t* mark, int32_t mepoch, const int32_t* cur,
                             const int64_t* cur_count, int32_t lvl, int32_t* inval, int64_t* inval_count,
                             int64_t inval_cap, int32_t* next, int64_t* next_count, int64_t cap, int32_t* stamp,
                             int32_t epoch, int64_t* scanned) {
  int64_t nc = *cur_count;
  int64_t work = 0;
  for (int64_t i = global_wave(); i < nc; i += total_waves()) {
    int32_t c = cur[i];
    Row r = row_of(g, c);
    int64_t len = row_len(r);
    int32_t parent = wave_find_first(g, r, 0, len, [&](int32_t x) { return level[x] == lvl - 1 && mark[x] != mepoch; });
    if (parent >= 0) {
      work += 1;
      continue;
    }
    work += len;
    if (lane_id() == 0) {
      mark[c] = mepoch;
      thread_append(c, inval, inval_count, inval_cap);
    }
    for (int64_t base = 0; base < len; base += kWave) {
      int64_t k = base + lane_id();
      bool take = false;
      int32_t w = -1;
      if (k < len) {
        w = row_at(g, r, k);
        take = w >= 0 && level[w] == lvl + 1 && atomicExch(&stamp[w], epoch) != epoch;
      }
      wave_append(take, w, next, next_count, cap);
    }
  }
  if (lane_id() == 0 && work) atomicAdd((unsigned long long*)scanned, (unsigned long long)work);
}

__global__ void k_reset_levels(const int32_t* xs, int64_t nx, int32_t* level) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nx) level[xs[i]] = -1;
}

// Wave per invalid vertex: 1 + smallest reached neighbour level, or -1.
__global__ void k_tentative(GView g, const int32_t* level, const int32_t* xs, int64_t nx, int32_t* tent,
                            int64_t* scanned) {
  int64_t work = 0;
  for (int64_t i = global_wave(); i < nx; i += total_waves()) {
    int32_t x = xs[i];
    Row r = row_of(g, x);
    int64_t len = row_len(r);
    work += len;
    int32_t best = INT32_MAX;
    for (int64_t k = lane_id(); k < len; k += kWave) {
      int32_t w = row_at(g, r, k);
      if (w >= 0) {
        int32_t lw = level[w];
        if (lw >= 0 && lw < best) best = lw;
      }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) best = min(best, __shfl_xor(best, off));
    if (lane_id() == 0) tent[i] = best == INT32_MAX ? -1 : best + 1;
  }
  if (lane_id() == 0 && work) atomicAdd((unsigned long long*)scanned, (unsigned long long)work);
}

// Seeds are packed level << 32 | vertex. box[4] = seeds, box[5] = min level, box[6] = max.
__device__ __forceinline__ void push_seed(int32_t v, int32_t lv, int64_t* seeds, int64_t* box, int64_t cap) {
  int64_t pos = (int64_t)atomicAdd((unsigned long long*)&box[4], 1ULL);
  if (pos < cap) seeds[pos] = ((int64_t)lv << 32) | (int64_t)(uint32_t)v;
  atomicMin((long long*)&box[5], (long long)lv);
  atomicMax((long long*)&box[6], (long long)lv);
}

__global__ void k_apply_tentative(const int32_t* xs, const int32_t* tent, int64_t nx, int32_t* level, int64_t* seeds,
                                  int64_t cap, int64_t* box) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  int32_t t = tent[i];
  level[xs[i]] = t;
  if (t >= 0) push_seed(xs[i], t, seeds, box, cap);
}

// Insertion lines whose endpoints are adjacent after the batch: lower the far endpoint.
__global__ void k_insert_seeds(GView g, bool sorted, const int32_t* uv, const uint8_t* ins, int64_t m,
                               int32_t* level, int64_t* seeds, int64_t cap, int64_t* box) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !ins[i]) return;
  atomicAdd((unsigned long long*)&box[7], 1ULL);
  int32_t u = uv[2 * i], v = uv[2 * i + 1];
  int32_t lu = level[u], lv = level[v];
  bool maybe = (lu >= 0 && (lv < 0 || lu + 1 < lv)) || (lv >= 0 && (lu < 0 || lv + 1 < lu));
  if (!maybe) return;
  if (!adjacent_now(pair_state(g, sorted, u, v))) return;
  if (lu >= 0 && cas_min_level(&level[v], lu + 1)) push_seed(v, lu + 1, seeds, box, cap);
  lv = level[v];
  if (lv >= 0 && cas_min_level(&level[u], lv + 1)) push_seed(u, lv + 1, seeds, box, cap);
}

// Seeds at `lvl` that still hold that level join the current frontier once.
__global__ void k_seed_at_level(const int64_t* seeds, int64_t ns, const int32_t* level, int32_t lvl, int32_t* cur,
                                int64_t* cur_count, int64_t cap, int32_t* stamp, int32_t epoch) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  int64_t s = seeds[i];
  int32_t lv = (int32_t)(s >> 32), v = (int32_t)(s & 0xffffffff);
  if (lv == lvl && level[v] == lvl && atomicExch(&stamp[v], epoch) != epoch) thread_append(v, cur, cur_count, cap);
}

__global__ void k_fill_level(int32_t* level, int32_t n, int32_t src) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) level[i] = (i == src) ? 0 : -1;
}

__global__ void k_set_one(int32_t* list, int32_t v, int64_t* count) {
  list[0] = v;
  *count = 1;
}

// Level-synchronous relax starting at level `lvl0` with the current list already in `cur`
// (count at box slot `cur_slot`) and optional seeds; runs until no frontier and no seed
// above the current level remains. Returns the rounds run via *rounds.
int relax_levels(Ctx& c, int32_t* L, int32_t lvl0, int32_t max_seed, const int64_t* seeds, int64_t nseeds, int32_t epoch,
                 int64_t* rounds, int64_t* scanned_total) {
  int32_t n = c.g.n;
  int64_t cap = (int64_t)n + 1;
  int64_t task_cap = c.wb.cap;
  int32_t* cur = c.lb.p;
  int32_t* nxt = c.lc.p;
  // box slots: [8] cur count, [9] next count, [10] task count, [11] scanned
  int64_t* bcur = c.d_box + 8;
  int64_t* bnext = c.d_box + 9;
  int64_t* btask = c.d_box + 10;
  int64_t* bscan = c.d_box + 11;
  SCOUT_TRY(hipMemsetAsync(bscan, 0, sizeof(int64_t), c.s));
  int32_t lvl = lvl0;
  int64_t r = 0;
  int64_t scanned_before = 0;
  while (true) {
    if (nseeds > 0 && lvl <= max_seed)
      k_seed_at_level<<<grid_for(nseeds), kBlock, 0, c.s>>>(seeds, nseeds, L, lvl, cur, bcur, cap, c.stamp, epoch);
    SCOUT_TRY(hipMemsetAsync(bnext, 0, 2 * sizeof(int64_t), c.s));
    SCOUT_CHECK(prof_begin(c, PROF_BFS_RELAX));
    k_relax<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), L, cur, bcur, lvl, nxt, bnext, cap, c.stamp, epoch,
                                            c.wb.p, btask, task_cap, bscan);
    k_relax_tasks<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), L, c.wb.p, btask, task_cap, lvl, nxt, bnext, cap,
                                                  c.stamp, epoch, bscan);
    SCOUT_CHECK(prof_end(c, PROF_BFS_RELAX, 0));
    r++;
    SCOUT_TRY(hipMemcpyAsync(c.h_box + 8, c.d_box + 8, 4 * sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
    SCOUT_TRY(hipStreamSynchronize(c.s));
    int64_t nn = c.h_box[9];
    // algorithmic bytes: per frontier vertex its id, level and the four row offsets (36 B);
    // per scanned entry the neighbour id and its level (8 B); per lowered vertex the level
    // write and its queue entry (8 B)
    prof_bytes(c, PROF_BFS_RELAX, 36 * c.h_box[8] + 8 * (c.h_box[11] - scanned_before) + 8 * nn);
    scanned_before = c.h_box[11];
    if (c.h_box[10] > task_cap) {
      set_error("bfs relax: chunk task list overflow");
      return SCOUT_ERR_STATE;
    }
    lvl++;
    // next becomes current
    std::swap(cur, nxt);
    SCOUT_TRY(hipMemcpyAsync(bcur, bnext, sizeof(int64_t), hipMemcpyDeviceToDevice, c.s));
    if (nn == 0 && (nseeds == 0 || lvl > max_seed)) break;
  }
  // keep list ownership consistent: cur/nxt are lb/lc in some order, both owned by c.
  *rounds += r;
  *scanned_total += c.h_box[11];
  return SCOUT_OK;
}

int ensure_bfs(Ctx& c, int64_t extra) {
  int64_t n = c.g.n;
  SCOUT_CHECK(ensure_scratch(c, &c.bfs_mark, 0));
  SCOUT_CHECK(c.la.reserve(n + 2 * extra + 1));
  SCOUT_CHECK(c.lb.reserve(n + 1));
  SCOUT_CHECK(c.lc.reserve(n + 1));
  SCOUT_CHECK(c.ld.reserve(n + 1));
  SCOUT_CHECK(c.le.reserve(n + 1));
  SCOUT_CHECK(c.wa.reserve(n + 2 * extra + 1));
  // chunk tasks: at most 2 * entries / kChunk (every split row has more than kChunk entries)
  SCOUT_CHECK(c.wb.reserve(2 * (c.g.E + c.g.A) / kChunk + 1024));
  return SCOUT_OK;
}

}  // namespace

int bfs_full_on(Ctx& c, int32_t* L, int32_t src, int64_t* rounds, int64_t* scanned) {
  SCOUT_CHECK(ensure_bfs(c, 0));
  int32_t n = c.g.n;
  if (n == 0) return SCOUT_OK;
  k_fill_level<<<grid_for(n), kBlock, 0, c.s>>>(L, n, src);
  int32_t epoch = next_epoch(c);
  k_set_one<<<1, 1, 0, c.s>>>(c.lb.p, src, c.d_box + 8);
  int64_t r = 0, sc = 0;
  SCOUT_CHECK(relax_levels(c, L, 0, -1, nullptr, 0, epoch, &r, &sc));
  if (rounds) *rounds += r;
  if (scanned) *scanned += sc;
  return SCOUT_OK;
}

int bfs_local_on(Ctx& c, Channel& ch, int32_t* L, scout_stats* st) {
  SCOUT_CHECK(ensure_bfs(c, ch.m));
  int64_t m = ch.m;
  int32_t n = c.g.n;
  int64_t cap = (int64_t)n + 1;
  int64_t* box = c.d_box;
  // box: [0] cand, [1] min, [2] max, [3] deletions, [4] seeds, [5] seed min, [6] seed max, [7] insertions
  int64_t init[8] = {0, INT32_MAX, -1, 0, 0, INT32_MAX, -1, 0};
  SCOUT_TRY(hipMemcpyAsync(box, init, sizeof(init), hipMemcpyHostToDevice, c.s));
  SCOUT_TRY(hipMemsetAsync(box + 8, 0, 8 * sizeof(int64_t), c.s));
  if (m > 0)
    k_del_candidates<<<grid_for(m), kBlock, 0, c.s>>>(ch.d_uv.p, ch.d_ins.p, m, L, c.la.p, c.la.cap, box);
  SCOUT_CHECK(box_read(c, 8));
  int64_t ncand = c.h_box[0];
  int32_t cmin = (int32_t)c.h_box[1], cmax = (int32_t)c.h_box[2];
  st->deletions += c.h_box[3];
  int64_t nx = 0, scanned = 0, rounds = 0;
  if (ncand > 0) {
    // Invalidation, level by level.
    int32_t mepoch = next_epoch(c);
    int32_t sepoch = next_epoch(c);
    int32_t* cur = c.lb.p;
    int32_t* nxt = c.lc.p;
    int64_t* bcur = box + 8;
    int64_t* bnext = box + 9;
    int64_t* binval = box + 10;
    int64_t* bscan = box + 11;
    SCOUT_TRY(hipMemsetAsync(box + 8, 0, 4 * sizeof(int64_t), c.s));
    for (int32_t lvl = cmin;; lvl++) {
      if (lvl <= cmax)
        k_cand_at_level<<<grid_for(ncand), kBlock, 0, c.s>>>(c.la.p, ncand, L, lvl, cur, bcur, cap, c.stamp,
                                                              sepoch);
      SCOUT_TRY(hipMemsetAsync(bnext, 0, sizeof(int64_t), c.s));
      k_invalidate<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), L, c.bfs_mark, mepoch, cur, bcur, lvl, c.ld.p,
                                                   binval, cap, nxt, bnext, cap, c.stamp, sepoch, bscan);
      rounds++;
      SCOUT_TRY(hipMemcpyAsync(c.h_box + 8, box + 8, 4 * sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
      SCOUT_TRY(hipStreamSynchronize(c.s));
      int64_t nn = c.h_box[9];
      std::swap(cur, nxt);
      SCOUT_TRY(hipMemcpyAsync(bcur, bnext, sizeof(int64_t), hipMemcpyDeviceToDevice, c.s));
      if (nn == 0 && lvl + 1 > cmax) break;
    }
    nx = c.h_box[10];
    scanned += c.h_box[11];
    if (nx > 0) {
      k_reset_levels<<<grid_for(nx), kBlock, 0, c.s>>>(c.ld.p, nx, L);
      SCOUT_TRY(hipMemsetAsync(box + 11, 0, sizeof(int64_t), c.s));
      k_tentative<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), L, c.ld.p, nx, c.le.p, box + 11);
      k_apply_tentative<<<grid_for(nx), kBlock, 0, c.s>>>(c.ld.p, c.le.p, nx, L, c.wa.p, c.wa.cap, box);
    }
  }
  if (m > 0)
    k_insert_seeds<<<grid_for(m), kBlock, 0, c.s>>>(c.g.view(), c.g.sorted, ch.d_uv.p, ch.d_ins.p, m, L, c.wa.p,
                                                    c.wa.cap, box);
  SCOUT_CHECK(box_read(c, 12));
  if (nx > 0) scanned += c.h_box[11];
  st->insertions += c.h_box[7];
  int64_t nseeds = c.h_box[4];
  if (nseeds > c.wa.cap) {
    set_error("bfs: seed list overflow");
    return SCOUT_ERR_STATE;
  }
  st->seeds += nseeds;
  st->touched += nx;
  if (nseeds > 0) {
    int32_t smin = (int32_t)c.h_box[5], smax = (int32_t)c.h_box[6];
    int32_t epoch = next_epoch(c);
    SCOUT_TRY(hipMemsetAsync(box + 8, 0, sizeof(int64_t), c.s));
    SCOUT_CHECK(relax_levels(c, L, smin, smax, c.wa.p, nseeds, epoch, &rounds, &scanned));
  }
  st->rounds += rounds;
  st->edges_scanned += scanned;
  st->noop += (ncand == 0 && nseeds == 0) ? m : 0;
  return SCOUT_OK;
}

int bfs_full(Ctx& c, int32_t src, int64_t* rounds, int64_t* scanned) {
  SCOUT_CHECK(ensure_state(c, SCOUT_STATE_LEVEL));
  SCOUT_CHECK(bfs_full_on(c, c.level, src, rounds, scanned));
  c.level_seq = c.apply_seq;
  return SCOUT_OK;
}

int bfs_local(Ctx& c, Channel& ch, scout_stats* st) {
  SCOUT_CHECK(bfs_local_on(c, ch, c.level, st));
  c.level_seq = c.apply_seq;
  return SCOUT_OK;
}

}  // namespace scout

using namespace scout;

extern "C" {

int scout_bfs_full(scout_ctx* x, int32_t source) {
  if (!x) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  if (!c.g.boff) {
    set_error("no graph loaded");
    return SCOUT_ERR_STATE;
  }
  if (source < 0 || source >= c.g.n) {
    set_error("source out of range");
    return SCOUT_ERR_ARG;
  }
  c.source = source;
  SCOUT_CHECK(bfs_full(c, source, nullptr, nullptr));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  return SCOUT_OK;
}

int scout_repair_bfs(scout_ctx* x, int channel, int strategy, scout_stats* st) {
  if (!x || !st) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  Channel* ch = channel_of(c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  if (!c.g.boff || !c.level) {
    set_error("repair_bfs: graph and levels must be loaded first");
    return SCOUT_ERR_STATE;
  }
  Span span;
  SCOUT_CHECK(span.begin(c.s));
  st->lines += ch->m;
  int rc;
  if (strategy == SCOUT_STRATEGY_RECOMPUTE) {
    st->fallback = 1;
    rc = bfs_full(c, c.source, &st->rounds, &st->edges_scanned);
  } else {
    rc = bfs_local(c, *ch, st);
  }
  SCOUT_CHECK(rc);
  SCOUT_CHECK(span.end(c.s));
  st->repair_ms += span.ms();
  return SCOUT_OK;
}

}  // extern "C"
