// bfs.hip: BFS levels from the source, kept exact after every batch.
//
// The reference repairs one edit at a time: an improving insertion starts CAS-min relax
// rounds, a deletion between consecutive levels starts Jacobi recompute rounds over the
// child's dependants, and a cascade past b = 200 vertices is deferred to one full BFS after
// the stream (src/bfs/scout_bfs.cu:436-515). Here one batch is one device job whose result
// equals textbook BFS on the updated graph.
//
// Besides the levels L, the engine keeps a parent P[v] for every reached vertex other than
// the source: a neighbour with L[P[v]] = L[v] - 1 (-1 for the source and unreached vertices).
// A batch then runs:
//   1. candidates: every vertex whose parent edge was deleted;
//   2. invalidation, level by level: a candidate at level l that has a neighbour at level
//      l-1 which is still valid takes it as its new parent; otherwise it is invalid, and its
//      children (neighbours whose parent it was) become candidates at level l+1;
//   3. every invalid vertex restarts at 1 + the smallest valid neighbour level (that
//      neighbour becomes its parent), or unreached;
//   4. insertions that now join u and v lower the far endpoint by CAS-min (gpu_relax,
//      src/bfs/scout_bfs.cu:46-54);
//   5. level-synchronous relax from all lowered vertices, processed in increasing level so
//      each vertex is expanded once, at its final level; each successful lowering records
//      the lowering vertex as the parent.
// Valid vertices keep a parent chain of valid vertices down to the source, so their old level
// is still the length of a path and, deletions only lengthening paths, still exact until the
// insertions are relaxed. Relaxing every vertex whose value dropped, in level order, is
// Dijkstra with unit weights, hence exact. Only parent edges start work: a deleted edge
// between consecutive levels that is not a parent edge changes nothing.
//
// Rows are walked by length: up to kShort entries by one thread (the entries and their
// levels are gathered first, so the loads of one row are in flight together), up to kSmall by
// a group of kGroup lanes (invalidation), up to kTask (relax, tentative levels) or kChunk
// (invalidation, which stops at the first support) by one wave, longer rows as kTask-entry
// tasks spread over many waves.
#include <algorithm>

#include "common.cuh"

namespace scout {

namespace {

constexpr int kShort = 16;
// Rows of kShort < length <= kSmall entries are walked by kGroup lanes of a wave each (four
// rows per wave at once, kU entries per lane in flight: the whole row in one step), so a wave
// does not take such rows one after another.
constexpr int kSmall = 64;
constexpr int kGroup = 16;
static_assert(kGroup * 4 == kSmall && kWave % kGroup == 0, "a group's four loads per lane cover a small row");
// Entries per chunk task of a split row, and the longest row the relax and tentative passes
// give one wave: a wave walks at most kTask entries (kTask / (kU * kWave) dependent steps),
// so one long row does not set a round's duration.
constexpr int64_t kTask = 512;
// Row steps per iteration of the wave walks: kU entries per lane loaded before any is used.
constexpr int kU = 4;
// LDS staging of a wave's appends to the next frontier (entries per wave): the list counter
// is then taken once per block, or once per wave when the segment fills, instead of once per
// row step of every wave.
constexpr int kNextSeg = 256;
// Level of a vertex found invalid in this batch until its tentative level is written. It is
// negative, so it never matches a level and reads as unreached.
constexpr int32_t kInvalid = -2;
// Asynchronous invalidation (SCOUT_BFS_ASYNC): while a round decides its members, each member
// at level l is stored as kPend - l, so that no vertex of the same round takes it as support
// before it is decided (a vertex seen pending waits for the next round instead). Levels
// kPend - l are all below kInvalid and read as unreached everywhere else.
constexpr int32_t kPend = -3;
__device__ __forceinline__ int32_t pend_of(int32_t l) { return kPend - l; }
// Rounds queued between two read-backs of the round counters: the first group is short (most
// repairs end within a few levels), later groups double up to kMaxGroup. A round whose list
// is empty costs a few launches and no work.
constexpr int kFirstGroup = 2;
// Batch lines per thread in the per-line kernels: fewer blocks, so fewer block-wide
// reductions and shared-counter atomics per batch.
constexpr int kLinesPerThread = 4;
constexpr int kMaxGroup = 16;
// The learned scheduler switches a batch to a full recompute when the predicted cascades
// cover more than this fraction of the vertices.
constexpr double kRecomputeFraction = 0.25;

// Gathers the row entries of a short row (len <= kShort) into registers; -1 past the end.
__device__ __forceinline__ void gather_short(const GView& g, const Row& r, int64_t len, bool active,
                                             int32_t (&xs)[kShort]) {
#pragma unroll
  for (int k = 0; k < kShort; k++) xs[k] = (active && k < len) ? row_at(g, r, k) : -1;
}

// Writes xs[k] for every take[k], in k order, from slot p on (slots at or past cap dropped).
__device__ __forceinline__ void write_taken(const bool (&take)[kShort], const int32_t (&xs)[kShort], int32_t* out,
                                            int64_t p, int64_t cap) {
#pragma unroll
  for (int k = 0; k < kShort; k++) {
    if (take[k]) {
      if (p < cap) out[p] = xs[k];
      p++;
    }
  }
}

// Round bookkeeping on the device, so the host can queue several rounds between read-backs.
// Each round's list counters come in two sets used by alternate rounds. The last kernel of a
// round calls round_book: one thread of its last block (whose share of the work list is the
// smallest) adds the round's list size to *sum, counts the round in *nonempty when its list
// was not empty or `side` says that side-list items of its level may exist, folds the other
// set's chunk-task count (the previous round's) into *max_tasks and zeroes that set for the
// next round, and zeroes the round's list count, which the next round appends to.
struct RoundBook {
  int64_t* fcur;         // this round's list count
  int64_t* sum;
  int64_t* nonempty;
  int64_t* max_tasks;
  int64_t* other_tasks;  // the other set: chunk tasks, then up to three more counters (or null)
  int64_t* other_a;
  int64_t* other_b;
  int side;
  int64_t* other_c = nullptr;
};

__device__ __forceinline__ void round_book(const RoundBook& b) {
  if (blockIdx.x != gridDim.x - 1 || threadIdx.x != 0) return;
  int64_t nc = *b.fcur, t = *b.other_tasks;
  *b.sum += nc;
  if (nc > 0 || b.side) *b.nonempty += 1;
  if (t > *b.max_tasks) *b.max_tasks = t;
  *b.other_tasks = 0;
  if (b.other_a) *b.other_a = 0;
  if (b.other_b) *b.other_b = 0;
  if (b.other_c) *b.other_c = 0;
  *b.fcur = 0;
}

// Bounds of chunk task `task` of row r.
__device__ __forceinline__ void task_range(const Row& r, int64_t task, int64_t& lo, int64_t& hi) {
  lo = (task & 0xffffffffLL) * kTask;
  hi = min(lo + kTask, row_len(r));
}

// Loads entries base + j * kWave + lane (j < kU) of row r, -1 at or past hi.
__device__ __forceinline__ void load_steps(const GView& g, const Row& r, int64_t base, int64_t hi,
                                           int32_t (&xs)[kU]) {
#pragma unroll
  for (int j = 0; j < kU; j++) {
    int64_t k = base + (int64_t)j * kWave + lane_id();
    xs[j] = k < hi ? row_at(g, r, k) : -1;
  }
}

// Group lanes: kGroup consecutive lanes walk one row of at most kSmall entries.
__device__ __forceinline__ int group_lane() { return lane_id() & (kGroup - 1); }

// Loads entries group_lane() + j * kGroup (j < kU) of row r, -1 at or past len (<= kSmall).
__device__ __forceinline__ void group_load(const GView& g, const Row& r, int64_t len, bool active, int32_t (&xs)[kU]) {
#pragma unroll
  for (int j = 0; j < kU; j++) {
    int64_t k = group_lane() + (int64_t)j * kGroup;
    xs[j] = active && k < len ? row_at(g, r, k) : -1;
  }
}

// Minimum and OR over the kGroup lanes of a group (the xor offsets stay inside the group).
__device__ __forceinline__ int32_t group_min(int32_t v) {
  for (int off = kGroup / 2; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ bool group_any(bool b) {
  int v = b ? 1 : 0;
  for (int off = kGroup / 2; off > 0; off >>= 1) v |= __shfl_xor(v, off);
  return v != 0;
}

// Cascade attribution while the learned scheduler logs training data: every invalid vertex
// is charged to the deletion candidate its cascade started from (root), for the first
// kMlpSample candidates. root == nullptr when not logging.
struct Attrib {
  int32_t* root;
  int32_t* subtree;
  int32_t* maxlvl;
};

__device__ __forceinline__ void attrib_invalid(const Attrib& a, int32_t c, int32_t lvl) {
  if (!a.root) return;
  int32_t rt = a.root[c];
  if (rt >= 0 && rt < kMlpSample) {
    atomicAdd(&a.subtree[rt], 1);
    atomicMax(&a.maxlvl[rt], lvl);
  }
}

__device__ __forceinline__ void attrib_child(const Attrib& a, int32_t c, int32_t w) {
  if (a.root) a.root[w] = a.root[c];
}

// Lowers w to lvl+1 on behalf of f; on success f becomes w's parent. True if w must be queued
// (lowered and not queued before under this epoch: the whole relax when it is
// level-synchronous, the next round when it is asynchronous).
__device__ __forceinline__ bool relax_to(int32_t* level, int32_t* parent, int32_t* stamp, int32_t epoch, int32_t f,
                                         int32_t w, int32_t lvl) {
  if (!cas_min_level(&level[w], lvl + 1)) return false;
  parent[w] = f;
  return atomicExch(&stamp[w], epoch) != epoch;
}

// ---------------------------------------------------------------------------------------
// Level buckets: the batch's deletion candidates and relax seeds grouped by level once, so
// each round reads only its own level's items instead of filtering the whole list.
// ---------------------------------------------------------------------------------------

// Levels spanned by one bucketing (LDS counters per block); wider spans keep the unbucketed
// per-round filter.
constexpr int kMaxBuckets = 2048;
// Items per thread of the bucketing kernels (a block handles kBlock * kBucketItems items).
constexpr int kBucketItems = 8;

// Level of item i: a vertex list keyed by its current level, or packed (level << 32 | vertex)
// seeds keyed by their level.
struct VertexKey {
  const int32_t* items;
  const int32_t* level;
  __device__ int32_t key(int64_t i) const { return level[items[i]]; }
  __device__ int64_t value(int64_t i) const { return items[i]; }
};
struct SeedKey {
  const int64_t* items;
  __device__ int32_t key(int64_t i) const { return (int32_t)(items[i] >> 32); }
  __device__ int64_t value(int64_t i) const { return items[i]; }
};

// Zeroes start[0..nb] (the per-level counts land at start[b + 1]).
__global__ void k_bucket_clear(int64_t* start, int nb) {
  for (int i = (int)threadIdx.x; i <= nb; i += (int)blockDim.x) start[i] = 0;
}

// Per-level counts: LDS counters per block, one global atomic per (block, level).
template <class K>
__global__ void k_bucket_count(K src, int64_t count, int32_t lmin, int nb, int64_t* start) {
  __shared__ int32_t h[kMaxBuckets];
  for (int i = (int)threadIdx.x; i < nb; i += (int)blockDim.x) h[i] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * blockDim.x * kBucketItems;
#pragma unroll
  for (int k = 0; k < kBucketItems; k++) {
    int64_t i = base + (int64_t)k * blockDim.x + threadIdx.x;
    if (i < count) {
      int32_t b = src.key(i) - lmin;
      if (b >= 0 && b < nb) atomicAdd(&h[b], 1);  // always in range: lmin..lmin+nb-1 span the keys
    }
  }
  __syncthreads();
  for (int i = (int)threadIdx.x; i < nb; i += (int)blockDim.x)
    if (h[i]) atomicAdd((unsigned long long*)&start[i + 1], (unsigned long long)h[i]);
}

// Counts -> exclusive offsets start[0..nb], and cursor[b] = start[b] (one block).
__global__ void k_bucket_scan(int64_t* start, int64_t* cursor, int nb) {
  __shared__ int64_t part[kBlock];
  // each thread sums a contiguous run of counts, the runs are scanned, then written back
  int per = (nb + kBlock - 1) / kBlock;
  int lo = (int)threadIdx.x * per, hi = min(lo + per, nb);
  int64_t s = 0;
  for (int b = lo; b < hi; b++) s += start[b + 1];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int i = 0; i < kBlock; i++) {
      int64_t v = part[i];
      part[i] = t;
      t += v;
    }
  }
  __syncthreads();
  // offsets into cursor first (the counts in start are still read), then copied over them
  int64_t run = part[threadIdx.x];
  for (int b = lo; b < hi; b++) {
    int64_t cnt = start[b + 1];
    cursor[b] = run;
    run += cnt;
  }
  __syncthreads();
  if (threadIdx.x == kBlock - 1) start[nb] = run;
  for (int b = lo; b < hi; b++) start[b] = cursor[b];
}

// Scatter into level order: ranks inside the block from LDS counters, one global atomic per
// (block, level) reserves the block's slots.
template <class K, class T>
__global__ void k_bucket_scatter(K src, int64_t count, int32_t lmin, int nb, int64_t* cursor, T* out) {
  __shared__ int32_t h[kMaxBuckets];
  __shared__ int64_t at[kMaxBuckets];
  for (int i = (int)threadIdx.x; i < nb; i += (int)blockDim.x) h[i] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * blockDim.x * kBucketItems;
  int32_t bk[kBucketItems], rank[kBucketItems];
#pragma unroll
  for (int k = 0; k < kBucketItems; k++) {
    int64_t i = base + (int64_t)k * blockDim.x + threadIdx.x;
    bk[k] = -1;
    rank[k] = 0;
    if (i < count) {
      int32_t b = src.key(i) - lmin;
      if (b >= 0 && b < nb) {
        bk[k] = b;
        rank[k] = atomicAdd(&h[b], 1);
      }
    }
  }
  __syncthreads();
  for (int i = (int)threadIdx.x; i < nb; i += (int)blockDim.x)
    if (h[i]) at[i] = (int64_t)atomicAdd((unsigned long long*)&cursor[i], (unsigned long long)h[i]);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kBucketItems; k++) {
    int64_t i = base + (int64_t)k * blockDim.x + threadIdx.x;
    if (bk[k] >= 0) out[at[bk[k]] + rank[k]] = (T)src.value(i);
  }
}

// Groups count items of src by level into out; start[0..nb] receives each level's range
// (lmin + b -> [start[b], start[b + 1])), cursor is nb entries of scratch. Every key must lie
// in [lmin, lmin + nb) and nb <= kMaxBuckets.
template <class K, class T>
void bucket_by_level(hipStream_t s, K src, int64_t count, int32_t lmin, int nb, int64_t* start, int64_t* cursor,
                     T* out) {
  unsigned grid = grid_for(count, kBlock * kBucketItems);
  k_bucket_clear<<<1, kBlock, 0, s>>>(start, nb);
  k_bucket_count<<<grid, kBlock, 0, s>>>(src, count, lmin, nb, start);
  k_bucket_scan<<<1, kBlock, 0, s>>>(start, cursor, nb);
  k_bucket_scatter<<<grid, kBlock, 0, s>>>(src, count, lmin, nb, cursor, out);
}

// ---------------------------------------------------------------------------------------
// Relax rounds
// ---------------------------------------------------------------------------------------

// Thread per frontier vertex still at level lvl, and per seed (level << 32 | vertex) of this
// level not queued before: seeds[0..ns) filtered by level, or, when the seeds are grouped by
// level, seeds[sb[0]..sb[1]) (ns = 0 in rounds past the deepest seed). Short rows are relaxed
// here; rows up to kSmall go to the group list, up to kTask to the wave list, longer rows to
// chunk tasks.
// *seeded += the seeds taken.
// Asynchronous (A): every frontier vertex relaxes from its current level, whatever the round;
// the first round takes every seed (seeds[0..ns), or [0..*ns_dev) when the count is on the
// device, at most seed_cap; once each under seed_epoch). A vertex may be
// lowered in several rounds, and each lowering queues it for the next round (epoch is that
// round's). A frontier vertex also takes a neighbour one level up as its parent: concurrent
// lowerings of one vertex in one round may leave the parent of a level that lost the race,
// and the round after its last lowering sets it right (its last lowerer is such a neighbour).
template <bool A>
__global__ void k_relax_short(GView g, int32_t* level, int32_t* parent, const int32_t* front, const int64_t* count,
                              const int64_t* seeds, int64_t ns, const int64_t* sb, int64_t* seeded, int32_t lvl,
                              int32_t* next, int64_t* next_count, int64_t cap, int32_t* stamp, int32_t epoch,
                              int32_t seed_epoch, int32_t* small, int64_t* small_count, int32_t* medium,
                              int64_t* medium_count, int64_t* tasks,
                              int64_t* task_count, int64_t task_cap, int64_t* scanned, const int64_t* ns_dev,
                              int64_t seed_cap, const int64_t* gate) {
  if (gate && *gate != 0) return;
  int64_t nf = *count;
  if (ns_dev) ns = min(*ns_dev, seed_cap);
  if (!A && sb && ns > 0) {
    seeds += sb[0];
    ns = sb[1] - sb[0];
  }
  int64_t total = nf + ns;
  int64_t work = 0, taken = 0;
  // the trip count is uniform across the block, so the block-wide appends see every thread
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < total; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = base + threadIdx.x;
    int32_t f = -1, lf = lvl;
    Row r{0, 0, 0, 0};
    int64_t len = 0;
    bool active = false;
    if (i < nf) {
      f = front[i];
      if constexpr (A) {
        lf = level[f];
        active = lf >= 0;
      } else {
        active = level[f] == lvl;
      }
    } else if (i < total) {
      int64_t sd = seeds[i - nf];
      f = (int32_t)(sd & 0xffffffff);
      if constexpr (A) {
        lf = level[f];
        active = lf >= 0 && atomicExch(&stamp[f], seed_epoch) != seed_epoch;
      } else {
        active = (int32_t)(sd >> 32) == lvl && level[f] == lvl && atomicExch(&stamp[f], epoch) != epoch;
      }
      taken += active;
    }
    if (active) {
      r = row_of(g, f);
      len = row_len(r);
    }
    bool is_short = active && len <= kShort;
    bool is_small = active && !is_short && len <= kSmall;
    bool is_medium = active && len > kSmall && len <= kTask;
    bool is_long = active && len > kTask;
    int32_t xs[kShort];
    gather_short(g, r, len, is_short, xs);
    int32_t lv[kShort];
#pragma unroll
    for (int k = 0; k < kShort; k++) lv[k] = xs[k] >= 0 ? level[xs[k]] : 0;
    if (is_short) work += len;
    if constexpr (A) {
      int32_t up = -1;
#pragma unroll
      for (int k = 0; k < kShort; k++)
        if (up < 0 && xs[k] >= 0 && lv[k] == lf - 1) up = xs[k];
      if (is_short && lf > 0 && up >= 0) parent[f] = up;
    }
    bool take[kShort];
    int32_t nnext = 0;
#pragma unroll
    for (int k = 0; k < kShort; k++) {
      take[k] = xs[k] >= 0 && (lv[k] == -1 || lv[k] > lf + 1) && relax_to(level, parent, stamp, epoch, f, xs[k], lf);
      nnext += take[k];
    }
    // one reservation for the block's group and wave rows, chunk tasks and lowered neighbours
    int32_t need[4] = {is_medium, is_long ? (int32_t)((len + kTask - 1) / kTask) : 0, nnext, is_small};
    int64_t pos[4];
    block_reserve_n(need, pos, medium_count, task_count, next_count, small_count);
    if (is_medium && pos[0] < cap) medium[pos[0]] = f;
    if (is_small && pos[3] < cap) small[pos[3]] = f;
    for (int64_t k = 0; k < need[1] && pos[1] + k < task_cap; k++) tasks[pos[1] + k] = ((int64_t)f << 32) | k;
    write_taken(take, xs, next, pos[2], cap);
  }
  block_add(work, scanned);
  block_add(taken, seeded);
}

// Group of kGroup lanes per small frontier vertex (four per wave), then wave per medium one
// (both listed by k_relax_short) or per chunk task of a long one.
// Asynchronous (A): relaxes from the vertex's current level and sets its parent to a
// neighbour one level up, as k_relax_short does.
template <bool A>
__global__ void k_relax_rows(GView g, int32_t* level, int32_t* parent, const int32_t* small,
                             const int64_t* small_count, const int32_t* medium, const int64_t* medium_count,
                             const int64_t* tasks, const int64_t* task_count, int64_t task_cap, int32_t lvl,
                             int32_t* next, int64_t* next_count, int64_t cap, int32_t* stamp, int32_t epoch,
                             int64_t* scanned, int64_t* paths, RoundBook book, const int64_t* gate) {
  if (gate && *gate != 0) return;
  __shared__ int32_t seg_all[kWavesPerBlock * kNextSeg];
  WaveStage<kNextSeg> st{seg_all + (threadIdx.x / kWave) * kNextSeg, 0};
  int64_t nm = *medium_count;
  int64_t nt = min(*task_count, task_cap);
  int64_t work = 0, ntask = 0, gwork = 0;
  constexpr int kPerWave = kWave / kGroup;
  int64_t nsm = *small_count;
  for (int64_t base = global_wave() * kPerWave; base < nsm; base += total_waves() * kPerWave) {
    int64_t idx = base + lane_id() / kGroup;
    bool act = idx < nsm;
    int32_t f = act ? small[idx] : -1;
    Row r{0, 0, 0, 0};
    int64_t len = 0;
    int32_t lf = lvl;
    if (act) {
      r = row_of(g, f);
      len = row_len(r);
      if constexpr (A) lf = level[f];
    }
    act = act && lf >= 0;
    int32_t ws[kU];
    group_load(g, r, len, act, ws);
    int32_t lw[kU];
#pragma unroll
    for (int j = 0; j < kU; j++) lw[j] = ws[j] >= 0 ? level[ws[j]] : 0;
    if constexpr (A) {
      // a neighbour one level up as the parent: the first in row order over the group
      int32_t mine = INT32_MAX;
#pragma unroll
      for (int j = 0; j < kU; j++)
        if (mine == INT32_MAX && ws[j] >= 0 && lw[j] == lf - 1) mine = group_lane() + j * kGroup;
      int32_t at = group_min(mine);
#pragma unroll
      for (int j = 0; j < kU; j++)
        if (act && lf > 0 && group_lane() + j * kGroup == at) parent[f] = ws[j];
    }
#pragma unroll
    for (int j = 0; j < kU; j++) {
      bool take = act && ws[j] >= 0 && (lw[j] == -1 || lw[j] > lf + 1) && relax_to(level, parent, stamp, epoch, f,
                                                                                     ws[j], lf);
      stage_push(st, take, ws[j], next, next_count, cap);
    }
    if (act && group_lane() == 0) gwork += len;
  }
  for (int64_t t = global_wave(); t < nm + nt; t += total_waves()) {
    int32_t f;
    int64_t lo, hi;
    Row r;
    if (t < nm) {
      f = medium[t];
      r = row_of(g, f);
      lo = 0;
      hi = row_len(r);
    } else {
      int64_t task = tasks[t - nm];
      f = (int32_t)(task >> 32);
      r = row_of(g, f);
      task_range(r, task, lo, hi);
      ntask++;
    }
    int32_t lf = lvl;
    if constexpr (A) lf = level[f];
    bool fixed = !A || lf <= 0;
    work += hi - lo;
    for (int64_t base = lo; base < hi; base += (int64_t)kU * kWave) {
      int32_t ws[kU];
      load_steps(g, r, base, hi, ws);
      int32_t lw[kU];
#pragma unroll
      for (int j = 0; j < kU; j++) lw[j] = ws[j] >= 0 ? level[ws[j]] : 0;
      if constexpr (A) {
#pragma unroll
        for (int j = 0; j < kU; j++) {
          unsigned long long up = __ballot(ws[j] >= 0 && lw[j] == lf - 1);
          if (!fixed && up) {
            if (lane_id() == __ffsll((long long)up) - 1) parent[f] = ws[j];
            fixed = true;
          }
        }
      }
      bool take[kU];
#pragma unroll
      for (int j = 0; j < kU; j++)
        take[j] = ws[j] >= 0 && (lw[j] == -1 || lw[j] > lf + 1) && relax_to(level, parent, stamp, epoch, f, ws[j], lf);
#pragma unroll
      for (int j = 0; j < kU; j++) stage_push(st, take[j], ws[j], next, next_count, cap);
    }
  }
  stage_finish(st, next, next_count, cap);
  count_path(ntask, &paths[PATH_RELAX_TASKS]);
  block_add((lane_id() == 0 ? work : 0) + gwork, scanned);
  round_book(book);
}

// ---------------------------------------------------------------------------------------
// Deletion candidates and invalidation
// ---------------------------------------------------------------------------------------

// Endpoint whose parent edge a deletion line removed.
// box[0] = candidates, box[1] = min level, box[2] = max level, box[3] = deletion lines.
__global__ void k_del_candidates(const int32_t* uv, const uint8_t* ins, int64_t m, const int32_t* level,
                                 const int32_t* parent, int32_t* cand, int64_t cap, int64_t* box, int32_t* root) {
  // kLinesPerThread lines per thread (each pass coalesced over the block), one append and one
  // reduction per block
  constexpr int K = kLinesPerThread;
  int64_t first = (int64_t)blockIdx.x * blockDim.x * K + threadIdx.x;
  bool take[K];
  int32_t cs[K];
  int64_t dels = 0, lmin = INT64_MAX, lmax = -1;
#pragma unroll
  for (int k = 0; k < K; k++) {
    int64_t i = first + (int64_t)k * blockDim.x;
    bool del = i < m && !ins[i];
    int32_t c = -1, lc = 0;
    if (del) {
      int32_t u = uv[2 * i], v = uv[2 * i + 1];
      int32_t lu = level[u], lv = level[v];
      if (lu >= 0 && lv == lu + 1 && parent[v] == u) c = v, lc = lv;
      else if (lv >= 0 && lu == lv + 1 && parent[u] == v) c = u, lc = lu;
    }
    take[k] = c >= 0;
    cs[k] = c;
    dels += del;
    if (c >= 0) {
      lmin = min(lmin, (int64_t)lc);
      lmax = max(lmax, (int64_t)lc);
    }
  }
  int64_t pos[K];
  block_append_k(take, cs, cand, &box[0], cap, pos);
  if (root) {
#pragma unroll
    for (int k = 0; k < K; k++)
      if (take[k]) root[cs[k]] = pos[k] >= 0 && pos[k] < kMlpSample ? (int32_t)pos[k] : -1;
  }
  block_add(dels, &box[3]);
  block_min_max(lmin, lmax, &box[1], &box[2]);
}

// Asynchronous invalidation: the batch's deletion candidates (cand[0..*ncand)) become the
// first round's list, each marked pending (stored as pend_of(level)); a vertex listed twice
// is kept once. Runs after every reader of the plain candidate levels (the candidate pass and
// the scheduler's features).
// n_host: the candidate count when the host has it (-1: read *ncand, queued before the host has
// read it back).
__global__ void k_mark_pending(const int32_t* cand, const int64_t* ncand, int64_t n_host, int32_t* level,
                               int32_t* out, int64_t* out_count, int64_t cap, const int64_t* gate) {
  if (gate && *gate != 0) return;
  constexpr int K = kLinesPerThread;
  int64_t n = n_host >= 0 ? n_host : *ncand;
  int64_t first = (int64_t)blockIdx.x * blockDim.x * K + threadIdx.x;
  bool take[K];
  int32_t cs[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    int64_t i = first + (int64_t)k * blockDim.x;
    int32_t c = i < n ? cand[i] : -1;
    int32_t lc = c >= 0 ? level[c] : -1;
    take[k] = lc > 0 && atomicCAS(&level[c], lc, pend_of(lc)) == lc;
    cs[k] = c;
  }
  int64_t pos[K];
  block_append_k(take, cs, out, out_count, cap, pos);
}

// Invalidation of the candidates at level lvl: the children queued by the previous round
// (cur) and the batch's deletion candidates of this level (cand[0..ncand), or cand[cb[0]..cb[1])
// when the candidates are grouped by level; filtered by level and stamp; ncand = 0 past the
// deepest candidate; *listed += those taken). Thread per candidate; rows longer than kShort go
// to the group list (<= kSmall), the wave list (<= kChunk) or the long list, whose chunk tasks
// are pushed here.
// Asynchronous (A): the round's list (cur) holds pending vertices of any level, each decided
// at its own level l = kPend - level: a neighbour at l - 1 that is not pending supports it; if
// none does but one is pending, it waits for the next round (queued again, still pending);
// otherwise it is invalid and its children at l + 1 are marked pending and queued. An early
// "invalid" is final (during the invalidation support only disappears), and a supported
// vertex whose parent is later found invalid is queued again by that parent's child scan, so
// the result is the level-synchronous one in as many rounds as the longest cascade needs.
// A long row's level goes to long_lvl beside its entry in longs.
template <bool A>
__global__ void k_invalidate_short(GView g, int32_t* level, int32_t* parent, const int32_t* cur,
                                   const int64_t* cur_count, const int32_t* cand, int64_t ncand, const int64_t* cb,
                                   int64_t* listed, int32_t lvl, int32_t* inval, int64_t* inval_count, int32_t* next,
                                   int64_t* next_count, int64_t cap, int32_t* stamp, int32_t epoch, int32_t* small,
                                   int64_t* small_count, int32_t* medium, int64_t* medium_count, int32_t* longs,
                                   int32_t* long_lvl, int64_t* long_count, int64_t* tasks, int64_t* task_count,
                                   int64_t task_cap, int64_t* scanned, int64_t* paths, Attrib at,
                                   const int64_t* gate) {
  if (gate && *gate != 0) return;
  int64_t nc = *cur_count;
  if (A) ncand = 0;
  if (!A && cb && ncand > 0) {
    cand += cb[0];
    ncand = cb[1] - cb[0];
  }
  int64_t total = nc + ncand;
  int64_t work = 0, taken = 0, deferred = 0;
  // the trip count is uniform across the block, so the block-wide appends see every thread
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < total; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = base + threadIdx.x;
    int32_t c = -1, lm = lvl;
    Row r{0, 0, 0, 0};
    int64_t len = 0;
    bool active = i < nc;
    if (active) {
      c = cur[i];
      if constexpr (A) {
        int32_t e = level[c];
        active = e <= kPend;
        lm = kPend - e;
      }
    } else if (i < total) {
      // a deletion candidate whose level is this round's, listed once (stamp)
      int32_t x = cand[i - nc];
      active = level[x] == lvl && atomicExch(&stamp[x], epoch) != epoch;
      if (active) c = x;
      taken += active;
    }
    if (active) {
      r = row_of(g, c);
      len = row_len(r);
    }
    bool is_short = active && len <= kShort;
    int32_t xs[kShort];
    gather_short(g, r, len, is_short, xs);
    int32_t lv[kShort];
#pragma unroll
    for (int k = 0; k < kShort; k++) lv[k] = xs[k] >= 0 ? level[xs[k]] : kInvalid;
    int32_t found = -1;
    bool pending = false;
#pragma unroll
    for (int k = 0; k < kShort; k++) {
      if (found < 0 && lv[k] == lm - 1) found = xs[k];
      if (A) pending |= xs[k] >= 0 && lv[k] == pend_of(lm - 1);
    }
    bool defer = A && is_short && found < 0 && pending;
    bool invalid = is_short && found < 0 && !defer;
    deferred += defer;
    if (is_short) {
      work += len;
      if (invalid) {
        level[c] = kInvalid;
        attrib_invalid(at, c, lm);
      } else if (!defer) {
        parent[c] = found;
        if (A) level[c] = lm;
      }
    }
    bool is_small = active && !is_short && len <= kSmall;
    bool is_medium = active && len > kSmall && len <= kChunk;
    bool is_long = active && len > kChunk;
    // children of an invalid candidate: neighbours one level down (levels read above; no
    // vertex of that level changes in this kernel, except to pending) whose parent it is
    int32_t pw[kShort];
#pragma unroll
    for (int k = 0; k < kShort; k++) pw[k] = (invalid && xs[k] >= 0 && lv[k] == lm + 1) ? parent[xs[k]] : -1;
    bool take[kShort];
    int32_t nkids = 0;
#pragma unroll
    for (int k = 0; k < kShort; k++) {
      if constexpr (A)
        take[k] = pw[k] == c && invalid && atomicCAS(&level[xs[k]], lm + 1, pend_of(lm + 1)) == lm + 1;
      else
        take[k] = pw[k] == c && invalid && atomicExch(&stamp[xs[k]], epoch) != epoch;
      if (take[k]) attrib_child(at, c, xs[k]);
      nkids += take[k];
    }
    // one reservation for the block's invalid, medium and long rows, the long rows' chunk
    // tasks (index in longs << 32 | chunk) and the children (and a deferred vertex); a vertex
    // is listed once per round, so the lists (n + 1 slots) never fill
    int32_t need[6] = {invalid, is_medium, is_long, is_long ? (int32_t)((len + kTask - 1) / kTask) : 0,
                       nkids + (int32_t)defer, is_small};
    int64_t pos[6];
    block_reserve_n(need, pos, inval_count, medium_count, long_count, task_count, next_count, small_count);
    if (invalid && pos[0] < cap) inval[pos[0]] = c;
    if (is_medium && pos[1] < cap) medium[pos[1]] = c;
    if (is_small && pos[5] < cap) small[pos[5]] = c;
    if (is_long && pos[2] < cap) {
      longs[pos[2]] = c;
      if (A) long_lvl[pos[2]] = lm;
      for (int64_t k = 0; k < need[3] && pos[3] + k < task_cap; k++) tasks[pos[3] + k] = (pos[2] << 32) | k;
    }
    write_taken(take, xs, next, pos[4], cap);
    if (defer && pos[4] < cap) next[pos[4]] = c;
  }
  block_add(work, scanned);
  block_add(taken, listed);
  if (A) block_add(deferred, &paths[PATH_INVAL_DEFERRED]);
}

// Wave per medium candidate: a neighbour one level up supports it (and becomes its parent);
// otherwise it is invalid and its children (neighbours one level down whose parent it is)
// join the next level. The support search keeps the neighbours one level down in LDS, so an
// invalid candidate then reads only their parents.
// Then wave per chunk task of a long candidate: a neighbour one level up anywhere in the
// chunk supports the candidate and becomes its parent (any of them will do); supported[c] =
// supepoch records it for k_inval_push.
// Asynchronous (A): levels are the candidates' own (as in k_invalidate_short); a medium row
// with no support but a pending neighbour one level up waits for the next round, and a chunk
// that sees such a neighbour raises supported[c] to supepoch - 1 (atomicMax: support wins).
// Rows of at most kSmall entries (the group list) are walked first, kGroup lanes per row: the
// row, its levels and (for an invalid row) its children's parents are read in one step each.
// A candidate at level 1 can only be supported by the source src: it is supported exactly when
// still adjacent to it (one pair lookup instead of a search of its whole row, usually a hub's
// that lost its edge to the source); an unsupported one then reads its row once, for children.
// own[0] += entries and parents read, own[1] += rows and tasks.
template <bool A>
__global__ void k_invalidate(GView g, int32_t* level, int32_t* parent, const int32_t* small,
                             const int64_t* small_count, const int32_t* cur, const int64_t* cur_count,
                             const int32_t* longs, const int32_t* long_lvl, const int64_t* tasks,
                             const int64_t* task_count, int64_t task_cap, int32_t* supported, int32_t supepoch,
                             int32_t lvl, int32_t* inval, int64_t* inval_count, int64_t inval_cap, int32_t* next,
                             int64_t* next_count, int64_t cap, int32_t* stamp, int32_t epoch, int64_t* scanned,
                             int64_t* own, int64_t* paths, bool sorted, int32_t src, Attrib ab,
                             const int64_t* gate) {
  if (gate && *gate != 0) return;
  // neighbours one level down met while searching for support: the only possible children
  constexpr int kKids = 1024;
  constexpr int kInvSeg = 64;
  __shared__ int32_t kids_all[kWavesPerBlock * kKids];
  __shared__ int32_t seg_all[kWavesPerBlock * kNextSeg];
  __shared__ int32_t inv_all[kWavesPerBlock * kInvSeg];
  int32_t* kids = kids_all + (threadIdx.x / kWave) * kKids;
  WaveStage<kNextSeg> st{seg_all + (threadIdx.x / kWave) * kNextSeg, 0};
  WaveStage<kInvSeg> sinv{inv_all + (threadIdx.x / kWave) * kInvSeg, 0};
  int64_t nc = *cur_count;
  int64_t nt = min(*task_count, task_cap);
  int64_t work = 0, reads = 0, rows = 0, ntask = 0, overflow = 0, deferred = 0;
  // per-lane counts of the group walks (a group's leader counts its row)
  int64_t gwork = 0, greads = 0, grows = 0, gdeferred = 0;
  constexpr int kPerWave = kWave / kGroup;
  int64_t nsm = *small_count;
  for (int64_t base = global_wave() * kPerWave; base < nsm; base += total_waves() * kPerWave) {
    int64_t idx = base + lane_id() / kGroup;
    bool act = idx < nsm;
    bool lead = act && group_lane() == 0;
    int32_t c = act ? small[idx] : -1;
    Row r{0, 0, 0, 0};
    int64_t len = 0;
    int32_t lm = lvl;
    if (act) {
      r = row_of(g, c);
      len = row_len(r);
      if constexpr (A) lm = kPend - level[c];
    }
    int32_t xs[kU];
    group_load(g, r, len, act, xs);
    int32_t lx[kU];
#pragma unroll
    for (int j = 0; j < kU; j++) lx[j] = xs[j] >= 0 ? level[xs[j]] : INT32_MIN;
    // the first support in row order (entry group_lane() + j * kGroup), then the group's minimum
    int32_t mine = INT32_MAX;
    bool pend = false;
#pragma unroll
    for (int j = 0; j < kU; j++) {
      if (mine == INT32_MAX && xs[j] >= 0 && lx[j] == lm - 1) mine = group_lane() + j * kGroup;
      if (A) pend |= xs[j] >= 0 && lx[j] == pend_of(lm - 1);
    }
    int32_t at = group_min(mine);
    bool any_pend = A && group_any(pend);
    bool supported = act && at != INT32_MAX;
    bool defer = act && !supported && any_pend;
    bool invalid = act && !supported && !defer;
    if (supported) {
#pragma unroll
      for (int j = 0; j < kU; j++) {
        if (group_lane() + j * kGroup == at) {
          parent[c] = xs[j];
          if (A) level[c] = lm;
        }
      }
    }
    if (invalid && lead) {
      level[c] = kInvalid;
      attrib_invalid(ab, c, lm);
    }
    stage_push(sinv, invalid && lead, c, inval, inval_count, inval_cap);
    stage_push(st, defer && lead, c, next, next_count, cap);
    // children of an invalid row: its neighbours one level down whose parent it is
    int32_t pw[kU];
#pragma unroll
    for (int j = 0; j < kU; j++) pw[j] = (invalid && xs[j] >= 0 && lx[j] == lm + 1) ? parent[xs[j]] : -1;
#pragma unroll
    for (int j = 0; j < kU; j++) {
      // only an invalid row's lanes take children (an idle lane's c and pw[j] are both -1)
      bool take;
      if constexpr (A)
        take = invalid && pw[j] == c && atomicCAS(&level[xs[j]], lm + 1, pend_of(lm + 1)) == lm + 1;
      else
        take = invalid && pw[j] == c && atomicExch(&stamp[xs[j]], epoch) != epoch;
      if (take) attrib_child(ab, c, xs[j]);
      greads += invalid && xs[j] >= 0 && lx[j] == lm + 1;
      stage_push(st, take, xs[j], next, next_count, cap);
    }
    if (lead) {
      int64_t nread = supported ? at + 1 : len;
      gwork += nread;
      greads += nread;
      grows++;
      gdeferred += defer;
    }
  }
  for (int64_t i = global_wave(); i < nc + nt; i += total_waves()) {
    if (i >= nc) {
      int64_t task = tasks[i - nc];
      int32_t c = longs[task >> 32];
      int32_t lm = lvl;
      if constexpr (A) lm = long_lvl[task >> 32];
      if (supported[c] == supepoch) continue;
      ntask++;
      Row r = row_of(g, c);
      if (lm == 1) {
        // the source alone can support it: chunk 0 asks, the other chunks have nothing to do
        rows++;
        if ((task & 0xffffffffLL) == 0 && wave_adjacent(g, sorted, c, src) && lane_id() == 0) {
          parent[c] = src;
          if constexpr (A) {
            level[c] = lm;
            atomicMax(&supported[c], supepoch);
          } else {
            supported[c] = supepoch;
          }
        }
        continue;
      }
      int64_t lo, hi;
      task_range(r, task, lo, hi);
      int64_t at = -1;
      bool pending = false;
      // one row step first (support is usually among the first entries), then kU at a time
      for (int64_t base = lo, step = kWave; base < hi && at < 0; base += step, step = (int64_t)kU * kWave) {
        int32_t xs[kU];
        load_steps(g, r, base, min(hi, base + step), xs);
        int32_t lx[kU];
#pragma unroll
        for (int j = 0; j < kU; j++) lx[j] = xs[j] >= 0 ? level[xs[j]] : INT32_MIN;
#pragma unroll
        for (int j = 0; j < kU; j++) {
          unsigned long long sup = __ballot(xs[j] >= 0 && lx[j] == lm - 1);
          if (sup && at < 0) at = base + (int64_t)j * kWave + (__ffsll((long long)sup) - 1);
          if (A) pending |= __ballot(xs[j] >= 0 && lx[j] == pend_of(lm - 1)) != 0;
        }
      }
      int64_t nread = (at >= 0 ? at + 1 : hi) - lo;
      work += nread;
      reads += nread;
      rows++;
      if (at >= 0 && lane_id() == 0) {
        parent[c] = row_at(g, r, at);
        if constexpr (A) {
          level[c] = lm;
          atomicMax(&supported[c], supepoch);
        } else {
          supported[c] = supepoch;
        }
      } else if (A && at < 0 && pending && lane_id() == 0) {
        atomicMax(&supported[c], supepoch - 1);
      }
      continue;
    }
    int32_t c = cur[i];
    Row r = row_of(g, c);
    int32_t lm = lvl;
    if constexpr (A) lm = kPend - level[c];
    int64_t len = row_len(r);
    int64_t at = -1, nk = 0;
    bool pending = false;
    bool source_only = lm == 1;
    if (source_only) {
      // the source alone can support it (no search; children by one pass over the row below)
      rows++;
      if (wave_adjacent(g, sorted, c, src)) {
        if (lane_id() == 0) {
          parent[c] = src;
          if (A) level[c] = lm;
        }
        continue;
      }
      nk = kKids + 1;
    }
    // one row step first (support is usually among the first entries of a row, whose hubs
    // come first), then kU steps in flight per iteration; `at` is uniform across the wave
    for (int64_t base = 0, step = kWave; !source_only && base < len && at < 0;
         base += step, step = (int64_t)kU * kWave) {
      int32_t xs[kU];
      load_steps(g, r, base, min(len, base + step), xs);
      int32_t lx[kU];
#pragma unroll
      for (int j = 0; j < kU; j++) lx[j] = xs[j] >= 0 ? level[xs[j]] : INT32_MIN;
#pragma unroll
      for (int j = 0; j < kU; j++) {
        if (at >= 0) break;
        unsigned long long sup = __ballot(xs[j] >= 0 && lx[j] == lm - 1);
        if (sup) {
          at = base + (int64_t)j * kWave + (__ffsll((long long)sup) - 1);
          break;
        }
        if (A) pending |= __ballot(xs[j] >= 0 && lx[j] == pend_of(lm - 1)) != 0;
        bool kid = xs[j] >= 0 && lx[j] == lm + 1;
        unsigned long long km = __ballot(kid);
        int64_t slot = nk + __popcll(km & ((1ULL << lane_id()) - 1ULL));
        if (kid && slot < kKids) kids[slot] = xs[j];
        nk += __popcll(km);
      }
    }
    if (!source_only) rows++;
    if (at >= 0) {
      work += at + 1;
      reads += at + 1;
      if (lane_id() == 0) {
        parent[c] = row_at(g, r, at);
        if (A) level[c] = lm;
      }
      continue;
    }
    work += len;
    if (A && pending) {
      // a pending neighbour one level up may still support it: decided next round
      reads += len;
      deferred++;
      stage_push(st, lane_id() == 0, c, next, next_count, cap);
      continue;
    }
    if (lane_id() == 0) {
      level[c] = kInvalid;
      attrib_invalid(ab, c, lm);
    }
    stage_push(sinv, lane_id() == 0, c, inval, inval_count, inval_cap);
    if (nk <= kKids) {
      // children among the buffered neighbours (their levels were read above)
      __threadfence_block();
      reads += len + nk;
      for (int64_t base = 0; base < nk; base += (int64_t)kU * kWave) {
        int32_t ws[kU];
#pragma unroll
        for (int j = 0; j < kU; j++) {
          int64_t q = base + (int64_t)j * kWave + lane_id();
          ws[j] = q < nk ? kids[q] : -1;
        }
        int32_t pw[kU];
#pragma unroll
        for (int j = 0; j < kU; j++) pw[j] = ws[j] >= 0 ? parent[ws[j]] : -1;
        bool take[kU];
#pragma unroll
        for (int j = 0; j < kU; j++) {
          if constexpr (A)
            take[j] = pw[j] == c && atomicCAS(&level[ws[j]], lm + 1, pend_of(lm + 1)) == lm + 1;
          else
            take[j] = pw[j] == c && atomicExch(&stamp[ws[j]], epoch) != epoch;
          if (take[j]) attrib_child(ab, c, ws[j]);
        }
#pragma unroll
        for (int j = 0; j < kU; j++) stage_push(st, take[j], ws[j], next, next_count, cap);
      }
    } else {
      // the row again (the support search read it too, except for a level-1 candidate)
      reads += source_only ? len : 2 * len;
      overflow++;
      for (int64_t base = 0; base < len; base += (int64_t)kU * kWave) {
        int32_t ws[kU];
        load_steps(g, r, base, len, ws);
        int32_t pw[kU];
#pragma unroll
        for (int j = 0; j < kU; j++) pw[j] = ws[j] >= 0 ? parent[ws[j]] : -1;
        bool take[kU];
#pragma unroll
        for (int j = 0; j < kU; j++) {
          if constexpr (A)
            take[j] = pw[j] == c && atomicCAS(&level[ws[j]], lm + 1, pend_of(lm + 1)) == lm + 1;
          else
            take[j] = pw[j] == c && level[ws[j]] == lvl + 1 && atomicExch(&stamp[ws[j]], epoch) != epoch;
          if (take[j]) attrib_child(ab, c, ws[j]);
        }
#pragma unroll
        for (int j = 0; j < kU; j++) stage_push(st, take[j], ws[j], next, next_count, cap);
      }
    }
  }
  stage_finish(st, next, next_count, cap);
  stage_finish(sinv, inval, inval_count, inval_cap);
  count_path(ntask, &paths[PATH_INVAL_TASKS]);
  count_path(overflow, &paths[PATH_INVAL_KIDS_OVERFLOW]);
  count_path(deferred, &paths[PATH_INVAL_DEFERRED]);
  if (A) block_add(gdeferred, &paths[PATH_INVAL_DEFERRED]);
  // work, reads and rows of the wave walks are uniform across the wave: counted once per wave;
  // the group walks' by their leaders
  bool l0 = lane_id() == 0;
  block_add((l0 ? work : 0) + gwork, scanned);
  block_add((l0 ? reads : 0) + greads, &own[0]);
  block_add((l0 ? rows : 0) + grows, &own[1]);
}

// Wave per chunk of a long candidate that no chunk supported: it is invalid (recorded by its
// chunk 0), and every chunk queues its children.
// Asynchronous (A): a long row whose chunks saw a pending neighbour one level up (supported
// = supepoch - 1) waits for the next round instead (queued by its chunk 0, still pending);
// its level comes from long_lvl.
template <bool A>
__global__ void k_inval_push(GView g, int32_t* level, const int32_t* parent, int32_t lvl, const int32_t* longs,
                             const int32_t* long_lvl, const int64_t* tasks, const int64_t* task_count,
                             int64_t task_cap, const int32_t* supported, int32_t supepoch, int32_t* inval,
                             int64_t* inval_count, int64_t inval_cap, int32_t* next, int64_t* next_count, int64_t cap,
                             int32_t* stamp, int32_t epoch, int64_t* scanned, int64_t* paths, Attrib ab,
                             RoundBook book, const int64_t* gate) {
  if (gate && *gate != 0) return;
  constexpr int kInvSeg = 64;
  __shared__ int32_t seg_all[kWavesPerBlock * kNextSeg];
  __shared__ int32_t inv_all[kWavesPerBlock * kInvSeg];
  WaveStage<kNextSeg> st{seg_all + (threadIdx.x / kWave) * kNextSeg, 0};
  WaveStage<kInvSeg> sinv{inv_all + (threadIdx.x / kWave) * kInvSeg, 0};
  int64_t nt = min(*task_count, task_cap);
  int64_t work = 0, ntask = 0, deferred = 0;
  for (int64_t t = global_wave(); t < nt; t += total_waves()) {
    int64_t task = tasks[t];
    int32_t c = longs[task >> 32];
    int32_t sup = supported[c];
    if (sup == supepoch) continue;
    bool first = (task & 0xffffffffLL) == 0;
    if (A && sup == supepoch - 1) {
      deferred += first;
      stage_push(st, first && lane_id() == 0, c, next, next_count, cap);
      continue;
    }
    int32_t lm = lvl;
    if constexpr (A) lm = long_lvl[task >> 32];
    ntask++;
    if (first && lane_id() == 0) {
      level[c] = kInvalid;
      attrib_invalid(ab, c, lm);
    }
    stage_push(sinv, first && lane_id() == 0, c, inval, inval_count, inval_cap);
    Row r = row_of(g, c);
    int64_t lo, hi;
    task_range(r, task, lo, hi);
    work += hi - lo;
    for (int64_t base = lo; base < hi; base += (int64_t)kU * kWave) {
      int32_t ws[kU];
      load_steps(g, r, base, hi, ws);
      int32_t pw[kU];
#pragma unroll
      for (int j = 0; j < kU; j++) pw[j] = ws[j] >= 0 ? parent[ws[j]] : -1;
      bool take[kU];
#pragma unroll
      for (int j = 0; j < kU; j++) {
        if constexpr (A)
          take[j] = pw[j] == c && atomicCAS(&level[ws[j]], lm + 1, pend_of(lm + 1)) == lm + 1;
        else
          take[j] = pw[j] == c && level[ws[j]] == lvl + 1 && atomicExch(&stamp[ws[j]], epoch) != epoch;
        if (take[j]) attrib_child(ab, c, ws[j]);
      }
#pragma unroll
      for (int j = 0; j < kU; j++) stage_push(st, take[j], ws[j], next, next_count, cap);
    }
  }
  stage_finish(sinv, inval, inval_count, inval_cap);
  stage_finish(st, next, next_count, cap);
  count_path(ntask, &paths[PATH_INVAL_PUSH_TASKS]);
  count_path(deferred, &paths[PATH_INVAL_DEFERRED]);
  block_add(lane_id() == 0 ? work : 0, scanned);
  round_book(book);
}

// ---------------------------------------------------------------------------------------
// Tentative levels of invalid vertices and insertion seeds
// ---------------------------------------------------------------------------------------

// Smallest (level << 32 | neighbour) over the reached neighbours of each invalidated vertex
// (UINT64_MAX if none). Thread per vertex for rows of at most kShort entries; longer rows go
// to the wave list (<= kTask) or to chunk tasks.
__global__ void k_tentative_short(GView g, const int32_t* level, const int32_t* xs, const int64_t* nx_dev,
                                  uint64_t* tent, int32_t* medium, int64_t* medium_count, int64_t medium_cap,
                                  int64_t* tasks, int64_t* task_count, int64_t task_cap, int64_t* scanned,
                                  const int64_t* gate) {
  if (gate && *gate != 0) return;
  int64_t nx = *nx_dev;
  int64_t work = 0;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nx; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = base + threadIdx.x;
    Row r{0, 0, 0, 0};
    int64_t len = 0;
    bool active = i < nx;
    if (active) {
      r = row_of(g, xs[i]);
      len = row_len(r);
    }
    bool is_short = active && len <= kShort;
    int32_t ys[kShort];
    gather_short(g, r, len, is_short, ys);
    uint64_t best = UINT64_MAX;
#pragma unroll
    for (int k = 0; k < kShort; k++) {
      int32_t lw = ys[k] >= 0 ? level[ys[k]] : -1;
      if (lw >= 0) best = min(best, ((uint64_t)(uint32_t)lw << 32) | (uint32_t)ys[k]);
    }
    if (is_short) work += len;
    if (active) {
      tent[i] = best;
    }
    bool is_medium = active && !is_short && len <= kTask;
    bool is_long = active && len > kTask;
    int32_t need[2] = {is_medium, is_long ? (int32_t)((len + kTask - 1) / kTask) : 0};
    int64_t pos[2];
    block_reserve_n(need, pos, medium_count, task_count);
    if (is_medium && pos[0] < medium_cap) medium[pos[0]] = (int32_t)i;
    for (int64_t k = 0; k < need[1] && pos[1] + k < task_cap; k++) tasks[pos[1] + k] = (i << 32) | k;
  }
  block_add(work, scanned);
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
  for (int off = kWave / 2; off > 0; off >>= 1) {
    uint64_t o = ((uint64_t)(uint32_t)__shfl_xor((int)(v >> 32), off) << 32) | (uint32_t)__shfl_xor((int)v, off);
    v = min(v, o);
  }
  return v;
}

// Wave per medium row (index list) or per chunk of a long row; atomicMin into tent.
__global__ void k_tentative_rows(GView g, const int32_t* level, const int32_t* xs, const int32_t* medium,
                                 const int64_t* medium_count, const int64_t* tasks, const int64_t* task_count,
                                 int64_t task_cap, uint64_t* tent, int64_t* scanned, int64_t* paths,
                                 const int64_t* gate) {
  if (gate && *gate != 0) return;
  int64_t nm = *medium_count;
  int64_t nt = min(*task_count, task_cap);
  int64_t work = 0, ntask = 0;
  for (int64_t t = global_wave(); t < nm + nt; t += total_waves()) {
    int64_t i, lo, hi;
    Row r;
    if (t < nm) {
      i = medium[t];
      r = row_of(g, xs[i]);
      lo = 0;
      hi = row_len(r);
    } else {
      int64_t task = tasks[t - nm];
      i = task >> 32;
      r = row_of(g, xs[i]);
      task_range(r, task, lo, hi);
      ntask++;
    }
    work += hi - lo;
    uint64_t best = UINT64_MAX;
    for (int64_t base = lo; base < hi; base += (int64_t)kU * kWave) {
      int32_t ws[kU];
      load_steps(g, r, base, hi, ws);
      int32_t lw[kU];
#pragma unroll
      for (int j = 0; j < kU; j++) lw[j] = ws[j] >= 0 ? level[ws[j]] : -1;
#pragma unroll
      for (int j = 0; j < kU; j++)
        if (lw[j] >= 0) best = min(best, ((uint64_t)(uint32_t)lw[j] << 32) | (uint32_t)ws[j]);
    }
    best = wave_min_u64(best);
    if (lane_id() == 0 && best != UINT64_MAX) atomicMin((unsigned long long*)&tent[i], (unsigned long long)best);
  }
  count_path(ntask, &paths[PATH_TENTATIVE_TASKS]);
  block_add(lane_id() == 0 ? work : 0, scanned);
}

// Grid-stride over the nx_dev invalidated vertices, kLinesPerThread per thread and pass (the
// trip count is uniform across the block, so the block-wide appends see every thread).
__global__ void k_apply_tentative(const int32_t* xs, const uint64_t* tent, const int64_t* nx_dev, int32_t* level,
                                  int32_t* parent, int64_t* seeds, int32_t* seed_par, int64_t cap, int64_t* box,
                                  const int64_t* gate) {
  if (gate && *gate != 0) return;
  constexpr int K = kLinesPerThread;
  int64_t nx = *nx_dev;
  int64_t lmin = INT64_MAX, lmax = -1;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x * K; base < nx; base += (int64_t)gridDim.x * blockDim.x * K) {
    bool take[K];
    int64_t packed[K];
    int32_t par[K];
#pragma unroll
    for (int q = 0; q < K; q++) {
      int64_t i = base + threadIdx.x + (int64_t)q * blockDim.x;
      int32_t t = -1, x = 0, p = -1;
      if (i < nx) {
        uint64_t best = tent[i];
        if (best != UINT64_MAX) {
          t = (int32_t)(best >> 32) + 1;
          p = (int32_t)(best & 0xffffffffu);
        }
        x = xs[i];
        level[x] = t;
        parent[x] = p;
      }
      take[q] = t >= 0;
      packed[q] = ((int64_t)t << 32) | (int64_t)(uint32_t)x;
      par[q] = p;
      if (t >= 0) lmin = min(lmin, (int64_t)t), lmax = max(lmax, (int64_t)t);
    }
    int64_t at[K];
    block_append_k(take, packed, seeds, &box[4], cap, at);
#pragma unroll
    for (int q = 0; q < K; q++)
      if (at[q] >= 0) seed_par[at[q]] = par[q];
  }
  block_min_max(lmin, lmax, &box[5], &box[6]);
}

// Insertion lines whose endpoints are adjacent after the batch: lower the far endpoint. With
// distinct pairs every insertion line's pair is adjacent after the batch (applied, or skipped
// by a conditional apply because it already was), so the row search is needed only when a
// pair repeats.
__global__ void k_insert_seeds(GView g, bool sorted, bool distinct, const int32_t* uv, const uint8_t* ins, int64_t m,
                               int32_t* level, int64_t* seeds, int32_t* seed_par, int64_t cap, int64_t* box,
                               const int64_t* gate) {
  if (gate && *gate != 0) return;
  // kLinesPerThread lines per thread (each pass coalesced over the block): one append of the
  // seeds and one reduction of their levels per block. Seeds are packed level << 32 | vertex,
  // with the vertex that lowered them in seed_par; box[4] = seeds, box[5] / box[6] = their
  // smallest / largest level, box[7] = insertion lines.
  constexpr int K = kLinesPerThread;
  int64_t first = (int64_t)blockIdx.x * blockDim.x * K + threadIdx.x;
  bool take[2 * K];
  int64_t packed[2 * K];
  int32_t par[2 * K];
  int64_t nins = 0, lmin = INT64_MAX, lmax = -1;
#pragma unroll
  for (int q = 0; q < K; q++) {
    int64_t i = first + (int64_t)q * blockDim.x;
    bool is_ins = i < m && ins[i];
    bool sv = false, su = false;
    int32_t u = 0, v = 0, lu = -1, lv = -1;
    if (is_ins) {
      u = uv[2 * i];
      v = uv[2 * i + 1];
      lu = level[u];
      lv = level[v];
      bool maybe = (lu >= 0 && (lv < 0 || lu + 1 < lv)) || (lv >= 0 && (lu < 0 || lv + 1 < lu));
      if (maybe && (distinct || adjacent_now(pair_state(g, sorted, u, v)))) {
        if (lu >= 0 && cas_min_level(&level[v], lu + 1)) sv = true;
        lv = level[v];
        if (lv >= 0 && cas_min_level(&level[u], lv + 1)) su = true;
      }
    }
    nins += is_ins;
    // seed = (level << 32 | vertex) with the vertex that lowered it as its parent candidate
    take[2 * q] = sv;
    packed[2 * q] = ((int64_t)(lu + 1) << 32) | (int64_t)(uint32_t)v;
    par[2 * q] = u;
    take[2 * q + 1] = su;
    packed[2 * q + 1] = ((int64_t)(lv + 1) << 32) | (int64_t)(uint32_t)u;
    par[2 * q + 1] = v;
    if (sv) lmin = min(lmin, (int64_t)(lu + 1)), lmax = max(lmax, (int64_t)(lu + 1));
    if (su) lmin = min(lmin, (int64_t)(lv + 1)), lmax = max(lmax, (int64_t)(lv + 1));
  }
  int64_t at[2 * K];
  block_append_k(take, packed, seeds, &box[4], cap, at);
#pragma unroll
  for (int q = 0; q < 2 * K; q++)
    if (at[q] >= 0) seed_par[at[q]] = par[q];
  block_add(nins, &box[7]);
  block_min_max(lmin, lmax, &box[5], &box[6]);
}

// A seed that still holds the level it was lowered to takes the vertex that lowered it as
// its parent (several seeds of one vertex at one level are all valid parents).
// ns_dev (optional): the count on the device, at most cap (grid-stride over it).
__global__ void k_seed_parents(const int64_t* seeds, const int32_t* seed_par, int64_t ns, const int64_t* ns_dev,
                               int64_t cap, const int32_t* level, int32_t* parent, const int64_t* gate) {
  if (gate && *gate != 0) return;
  if (ns_dev) ns = min(*ns_dev, cap);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t s = seeds[i];
    int32_t lv = (int32_t)(s >> 32), v = (int32_t)(s & 0xffffffff);
    if (level[v] == lv) parent[v] = seed_par[i];
  }
}

// ---------------------------------------------------------------------------------------
// Bookkeeping
// ---------------------------------------------------------------------------------------

// Zeroes `count` consecutive counters (round bookkeeping without a copy or fill kernel).
__global__ void k_zero(int64_t* p, int count) {
  if ((int)threadIdx.x < count) p[threadIdx.x] = 0;
}

// Box slots of the work queued ahead of the read-back that ends an invalidation group.
constexpr int kGateSlot = 39;  // the gate: zero once the invalidation has ended
constexpr int kNxSlot = 33;    // the invalidated vertices' count, kept for the tentative levels

// Opens (gate = 0) or closes that work: the gate takes the next round's list count. When the
// invalidation has ended, the invalidated count moves to kNxSlot and the tentative-level and
// relax counters ([8..21], [36..38]) are zeroed.
__global__ void k_spec_gate(const int64_t* front, int64_t* box) {
  if (threadIdx.x != 0) return;
  int64_t f = *front;
  box[kGateSlot] = f;
  if (f != 0) return;
  box[kNxSlot] = box[10];
  for (int k = 8; k <= 21; k++) box[k] = 0;
  for (int k = 36; k <= 38; k++) box[k] = 0;
}

// The gate of the work queued behind a relax read-back (the CC scan under scout_commit):
// the next round's frontier count, zero once the relax has ended.
constexpr int kRelaxGateSlot = 35;
__global__ void k_copy_gate(const int64_t* front, int64_t* gate) {
  if (threadIdx.x == 0) *gate = *front;
}

// Zeroes two runs of counters.
__global__ void k_zero2(int64_t* a, int na, int64_t* b, int nb, const int64_t* gate = nullptr) {
  if (gate && *gate != 0) return;
  int t = (int)threadIdx.x;
  if (t < na) a[t] = 0;
  if (t < nb) b[t] = 0;
}

// box[0..7] = {0, INT32_MAX, -1, 0, 0, INT32_MAX, -1, 0}, box[8..15] = 0.
__global__ void k_init_box(int64_t* box) {
  int t = (int)threadIdx.x;
  if (t < 16) box[t] = (t == 1 || t == 5) ? (int64_t)INT32_MAX : (t == 2 || t == 6) ? -1 : 0;
}

__global__ void k_fill_level(int32_t* level, int32_t* parent, int32_t n, int32_t src) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    level[i] = (i == src) ? 0 : -1;
    parent[i] = -1;
  }
}

__global__ void k_set_one(int32_t* list, int32_t v, int64_t* count) {
  list[0] = v;
  *count = 1;
}

// Wave per vertex: a parent for every reached vertex of a level array that came from outside
// (an init file or a download); *missing counts reached vertices with no neighbour one level up.
__global__ void k_find_parents(GView g, const int32_t* level, int32_t* parent, int32_t n, int64_t* missing) {
  for (int64_t v = global_wave(); v < n; v += total_waves()) {
    int32_t lv = level[v];
    int32_t p = -1;
    if (lv > 0) {
      Row r = row_of(g, (int32_t)v);
      int64_t at = wave_find_first(g, r, 0, row_len(r), [&](int32_t x) { return level[x] == lv - 1; });
      if (at >= 0) p = row_at(g, r, at);
      else if (lane_id() == 0) atomicAdd((unsigned long long*)missing, 1ULL);
    }
    if (lane_id() == 0) parent[v] = p;
  }
}

// Launch shape of the asynchronous rounds. A phase's first rounds carry nearly all of its work
// (every candidate or seed and its first cascade step) and take the full list grid; the rounds
// after them hold a few hundred vertices or fewer, where dispatching kListGrid mostly idle
// blocks costs more than the work: they take one block per CU (R-MAT scale 26, 1M-line
// batches: 5.32e8 -> 5.44e8 updates/s; 128 and 512 blocks measured no better). A read-back
// that finds a frontier larger than one pass of the small grid restores the full grid for the
// next group's first rounds. The level-synchronous order, whose later levels can be the widest,
// always takes the full grid.
constexpr int kFullRounds = 2;
// Grid of k_apply_tentative, whose count is on the device: one pass over 512k vertices.
constexpr unsigned kTentGrid = 512;
constexpr unsigned kTailGrid = 256;
constexpr int64_t kTailFront = (int64_t)kTailGrid * kBlock;

// Grid of the next round; counts down the full-grid rounds left.
unsigned round_grid(bool async, int& full_rounds) {
  if (!async) return kListGrid;
  if (full_rounds > 0) {
    full_rounds--;
    return kListGrid;
  }
  return kTailGrid;
}

// One relax: rounds queued in groups, each group followed by one read-back of the counters.
// Box slots: [8]/[9] frontier counts (current and next alternate), [10]/[11] and [18]/[19]
// chunk tasks and medium rows, [20]/[21] small rows (the counter sets of alternate rounds),
// [12] scanned entries, [13] entries of short rows, [15] frontier vertices over all rounds,
// [16] largest task count, [17] non-empty rounds.
// start() zeroes the counters (unless the caller has) and queues the first group; finish()
// reads back and queues further groups until a round lowers nothing. Under a gate, start()
// runs ahead of the read-back that decides whether the relax may start: every kernel it queues
// returns at once while *gate is non-zero.
struct Relax {
  Ctx& c;
  int32_t* L;
  int32_t* P;
  int32_t lvl0, max_seed;
  const int64_t* seeds;
  int64_t nseeds;
  int32_t epoch;
  bool async;
  const int64_t* nseeds_dev;
  int64_t seed_cap;
  // progress
  int32_t* cur = nullptr;
  int32_t* nxt = nullptr;
  const int64_t* sbuckets = nullptr;
  int32_t lvl = 0;
  int p = 0;
  int group = kFirstGroup, first_group = kFirstGroup;
  int64_t done = 0;  // rounds queued so far
  int full_rounds = kFullRounds;
  // work to queue behind each read-back, shut until the relax has ended (Ctx::after_relax)
  const std::function<int(const int64_t*, bool*)>* ahead = nullptr;

  int queue_group(const int64_t* gate) {
    int32_t n = c.g.n;
    int64_t cap = (int64_t)n + 1;
    int64_t task_cap = c.wb.cap;
    int64_t* box = c.d_box;
    int64_t* cnt[2] = {box + 8, box + 9};
    int64_t* tasks_of[2] = {box + 10, box + 18};
    int64_t* medium_of[2] = {box + 11, box + 19};
    int64_t* small_of[2] = {box + 20, box + 21};
    for (int k = 0; k < group; k++, done++) {
      int64_t* bcur = cnt[p];
      int64_t* bnext = cnt[p ^ 1];
      // level-synchronous: the seeds of this level are taken by the round's first kernel;
      // asynchronous: all of them by the first round, each round queueing under its own epoch
      int64_t ns = async ? (done == 0 ? nseeds : 0) : (nseeds > 0 && lvl <= max_seed) ? nseeds : 0;
      const int64_t* ns_dev = async && done == 0 ? nseeds_dev : nullptr;
      const int64_t* sb = (!async && sbuckets && ns > 0) ? sbuckets + (lvl - lvl0) : nullptr;
      int32_t ep = async ? next_epoch(c) : epoch;
      unsigned rg = round_grid(async, full_rounds);
      RoundBook book{bcur, box + 15, box + 17, box + 16, tasks_of[p ^ 1], medium_of[p ^ 1], small_of[p ^ 1],
                     ns > 0 || ns_dev != nullptr};
      if (async) {
        SCOUT_CHECK(prof_launch(c, PROF_K_RELAX_SHORT, k_relax_short<true>, dim3(rg), dim3(kBlock), c.g.view(), L,
                                P, cur, bcur, seeds, ns, sb, box + 15, lvl, nxt, bnext, cap, c.stamp, ep, epoch,
                                c.lh.p, small_of[p], c.le.p, medium_of[p], c.wb.p, tasks_of[p], task_cap, box + 13,
                                ns_dev, seed_cap, gate));
        SCOUT_CHECK(prof_launch(c, PROF_K_RELAX_ROWS, k_relax_rows<true>, dim3(rg), dim3(kBlock), c.g.view(), L, P,
                                c.lh.p, small_of[p], c.le.p, medium_of[p], c.wb.p, tasks_of[p], task_cap, lvl, nxt,
                                bnext, cap, c.stamp, ep, box + 12, c.d_paths, book, gate));
      } else {
        SCOUT_CHECK(prof_launch(c, PROF_K_RELAX_SHORT, k_relax_short<false>, dim3(rg), dim3(kBlock), c.g.view(), L,
                                P, cur, bcur, seeds, ns, sb, box + 15, lvl, nxt, bnext, cap, c.stamp, ep, epoch,
                                c.lh.p, small_of[p], c.le.p, medium_of[p], c.wb.p, tasks_of[p], task_cap, box + 13,
                                nullptr, 0, gate));
        SCOUT_CHECK(prof_launch(c, PROF_K_RELAX_ROWS, k_relax_rows<false>, dim3(rg), dim3(kBlock), c.g.view(), L,
                                P, c.lh.p, small_of[p], c.le.p, medium_of[p], c.wb.p, tasks_of[p], task_cap, lvl,
                                nxt, bnext, cap, c.stamp, ep, box + 12, c.d_paths, book, gate));
      }
      lvl++;
      std::swap(cur, nxt);
      p ^= 1;
    }
    return SCOUT_OK;
  }

  int start(int first, const int64_t* gate, bool zeroed) {
    int64_t* box = c.d_box;
    cur = c.lb.p;
    nxt = c.lc.p;
    lvl = lvl0;
    p = 0;
    done = 0;
    full_rounds = kFullRounds;
    group = first_group = first;
    if (!zeroed) k_zero2<<<1, 64, 0, c.s>>>(box + 9, 5, box + 15, 7);
    // seeds grouped by level once (each round then reads only its own), when the levels they
    // span fit the bucket counters and the list fits the scratch
    int64_t span = (int64_t)max_seed - lvl0 + 1;
    if (!async && nseeds > 0 && span >= 1 && span <= kMaxBuckets && c.wc.cap >= nseeds) {
      bucket_by_level(c.s, SeedKey{seeds}, nseeds, lvl0, (int)span, c.lvlb.p, c.lvlb.p + kMaxBuckets + 1,
                      reinterpret_cast<int64_t*>(c.wc.p));
      seeds = reinterpret_cast<const int64_t*>(c.wc.p);
      sbuckets = c.lvlb.p;
    }
    SCOUT_CHECK(prof_begin(c, PROF_BFS_RELAX));
    return queue_group(gate);
  }

  // Adds the non-empty rounds to *rounds.
  int finish(int64_t* rounds, int64_t* scanned_total) {
    int64_t* box = c.d_box;
    while (true) {
      SCOUT_CHECK(prof_end(c, PROF_BFS_RELAX, 0));
      // [8..21] this relax's counters; [4..7] and [36..38] come back too (the seed and
      // tentative-level counters of a local repair, which read them after the relax)
      bool queued = false;
      if (ahead && *ahead) {
        SCOUT_CHECK(box_read_async(c, 4, 35));
        k_copy_gate<<<1, 64, 0, c.s>>>(box + 8 + p, box + kRelaxGateSlot);
        SCOUT_CHECK((*ahead)(box + kRelaxGateSlot, &queued));
        SCOUT_CHECK(box_wait(c));
      } else {
        SCOUT_TRY(hipMemcpyAsync(c.h_box + 4, box + 4, 35 * sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
        SCOUT_TRY(hipStreamSynchronize(c.s));
        c.readbacks++;
      }
      if (std::max({c.h_box[10], c.h_box[16], c.h_box[18]}) > c.wb.cap) {
        set_error("bfs relax: chunk task list overflow");
        return SCOUT_ERR_STATE;
      }
      if (c.h_box[8 + p] == 0 && (async || nseeds == 0 || lvl > max_seed)) {
        if (queued) c.cc_ahead_seq = c.apply_seq;  // the gate opened: what was queued ran
        break;
      }
      full_rounds = c.h_box[8 + p] > kTailFront ? kFullRounds : 0;
      // after a predicted first group the repair is near its end: start again from a short group
      group = group == first_group && first_group != kFirstGroup ? kFirstGroup : std::min(2 * group, kMaxGroup);
      first_group = kFirstGroup;
      SCOUT_CHECK(prof_begin(c, PROF_BFS_RELAX));
      SCOUT_CHECK(queue_group(nullptr));
    }
    // algorithmic bytes: per frontier vertex its id, level, the four row offsets and its queue
    // entry (40 B); per scanned entry the neighbour id and its level (8 B). k_relax_short reads
    // every frontier vertex and the short rows ([13]), k_relax_rows the longer rows ([12]).
    prof_bytes(c, PROF_BFS_RELAX, 40 * c.h_box[15] + 8 * (c.h_box[12] + c.h_box[13]));
    prof_bytes(c, PROF_K_RELAX_SHORT, 40 * c.h_box[15] + 8 * c.h_box[13]);
    prof_bytes(c, PROF_K_RELAX_ROWS, 8 * c.h_box[12]);
    *rounds += c.h_box[17];
    *scanned_total += c.h_box[12] + c.h_box[13];
    return SCOUT_OK;
  }
};

// Relax with the current list already in `cur` (count in box slot 8) and optional seeds.
// Level-synchronous (async = false): starts at level `lvl0` and runs until no frontier and no
// seed above the current level remains. Asynchronous: the first round takes every seed and
// later rounds their lowered neighbours, whatever their levels, until a round lowers nothing.
// Adds the non-empty rounds to *rounds.
int relax_levels(Ctx& c, int32_t* L, int32_t* P, int32_t lvl0, int32_t max_seed, const int64_t* seeds, int64_t nseeds,
                 int32_t epoch, int64_t* rounds, int64_t* scanned_total, int first_group = kFirstGroup,
                 bool async = false, const int64_t* nseeds_dev = nullptr, int64_t seed_cap = 0) {
  Relax r{c, L, P, lvl0, max_seed, seeds, nseeds, epoch, async, nseeds_dev, seed_cap};
  SCOUT_CHECK(r.start(first_group, nullptr, false));
  return r.finish(rounds, scanned_total);
}

int ensure_bfs(Ctx& c, int64_t extra) {
  int64_t n = c.g.n;
  SCOUT_CHECK(c.la.reserve(n + 2 * extra + 1));
  SCOUT_CHECK(c.lb.reserve(n + 1));
  SCOUT_CHECK(c.lc.reserve(n + 1));
  SCOUT_CHECK(c.ld.reserve(n + 1));
  SCOUT_CHECK(c.le.reserve(n + 1));
  SCOUT_CHECK(c.lf.reserve(n + 1));
  SCOUT_CHECK(c.lg.reserve(n + 2 * extra + 1));
  SCOUT_CHECK(c.lh.reserve(n + 1));
  SCOUT_CHECK(c.wa.reserve(n + 2 * extra + 1));
  SCOUT_CHECK(c.wc.reserve(n + 2 * extra + 1));
  SCOUT_CHECK(c.lvlb.reserve(2 * kMaxBuckets + 2));
  // chunk tasks: at most 2 * entries / kTask (every split row has more than kTask entries)
  SCOUT_CHECK(c.wb.reserve(2 * (c.g.E + c.g.A) / kTask + 1024));
  return SCOUT_OK;
}

}  // namespace

int compute_parents(Ctx& c, const int32_t* L, int32_t* P) {
  int32_t n = c.g.n;
  if (n == 0) return SCOUT_OK;
  k_zero<<<1, 64, 0, c.s>>>(c.d_box, 1);
  k_find_parents<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), L, P, n, c.d_box);
  SCOUT_CHECK(box_read(c, 1));
  if (c.h_box[0] != 0) {
    set_error("levels are not BFS levels of the graph: " + std::to_string(c.h_box[0]) +
              " reached vertices have no neighbour one level up");
    return SCOUT_ERR_ARG;
  }
  return SCOUT_OK;
}

int bfs_full_on(Ctx& c, int32_t* L, int32_t* P, int32_t src, int64_t* rounds, int64_t* scanned) {
  SCOUT_CHECK(ensure_bfs(c, 0));
  int32_t n = c.g.n;
  if (n == 0) return SCOUT_OK;
  k_fill_level<<<grid_for(n), kBlock, 0, c.s>>>(L, P, n, src);
  int32_t epoch = next_epoch(c);
  k_set_one<<<1, 1, 0, c.s>>>(c.lb.p, src, c.d_box + 8);
  int64_t r = 0, sc = 0;
  SCOUT_CHECK(relax_levels(c, L, P, 0, -1, nullptr, 0, epoch, &r, &sc));
  if (rounds) *rounds += r;
  if (scanned) *scanned += sc;
  return SCOUT_OK;
}

int bfs_local_on(Ctx& c, Channel& ch, int32_t* L, int32_t* P, int strategy, scout_stats* st) {
  SCOUT_CHECK(ensure_bfs(c, ch.m));
  int64_t m = ch.m;
  int32_t n = c.g.n;
  int64_t cap = (int64_t)n + 1;
  int64_t seed_cap = std::min(c.wa.cap, c.lg.cap);  // seeds and their parents
  int64_t* box = c.d_box;
  // the maintained levels grow from the BFS source; CC's own levels from vertex 0
  int32_t src = L == c.level ? c.source : 0;
  // box: [0] cand, [1] min, [2] max, [3] deletions, [4] seeds, [5] seed min, [6] seed max, [7] insertions
  k_init_box<<<1, 64, 0, c.s>>>(box);
  // the learned scheduler logs from, and predicts for, the maintained levels only
  bool logging = L == c.level && mlp_logging(c);
  bool learned = L == c.level && strategy == SCOUT_STRATEGY_LEARNED && mlp_ready(c);
  Attrib ab{nullptr, nullptr, nullptr};
  if (logging) ab = Attrib{mlp_root(c), mlp_subtree(c), mlp_maxlvl(c)};
  if (m > 0)
    k_del_candidates<<<grid_for(m, kBlock * kLinesPerThread), kBlock, 0, c.s>>>(ch.d_uv.p, ch.d_ins.p, m, L, P,
                                                                              c.la.p, c.la.cap, box, ab.root);
  // features and predictions are computed from the device-side counts and come back with the
  // same read-back as the candidate counts ([24] predicted cascade sizes, [25] last level,
  // [27] expected vertices cut off vertex 0's component, in sixteenths)
  int64_t sample = logging ? kMlpSample : kMlpInfer;
  if (logging || learned) SCOUT_CHECK(mlp_begin_batch(c, c.la.p, box, L, P, sample));
  if (learned) SCOUT_CHECK(mlp_predict_async(c, box, sample, box + 24));
  // Invalidation, queued in groups of rounds. Counts: [8]/[9] the round's list / the next
  // (alternating), [10] invalid vertices, [11] scanned entries, [12..14] / [28..30] medium
  // rows, long rows and chunk tasks (alternating), [31]/[32] rows walked by lane groups, [18]
  // candidates over all rounds, [19] largest task count, [20] non-empty rounds, [15] entries
  // read by k_inval_push, [23] entries of short rows read by k_invalidate_short.
  // Level-synchronous: one level per round. Asynchronous (c.bfs_async): the first round holds
  // every candidate, marked pending, and the rounds run until one queues nothing.
  bool async = c.bfs_async;
  int32_t sepoch = 0;
  int32_t* cur = c.lb.p;
  int32_t* nxt = c.lc.p;
  const int32_t* cand = c.la.p;
  const int64_t* cbuckets = nullptr;  // candidates grouped by level (level-synchronous)
  int64_t* cnt[2] = {box + 8, box + 9};
  int64_t* binval = box + 10;
  int64_t* bscan = box + 11;
  int64_t* medium_of[2] = {box + 12, box + 28};
  int64_t* longs_of[2] = {box + 13, box + 29};
  int64_t* tasks_of[2] = {box + 14, box + 30};
  int64_t* small_of[2] = {box + 31, box + 32};
  int p = 0;
  int32_t lvl = 1;
  int full_rounds = kFullRounds;
  int64_t ncand = 0;
  int32_t cmin = 1, cmax = 0;
  // zeroed counters; in the asynchronous order every candidate marked pending into the first
  // round's list (the count comes from the device; there are at most m candidates)
  // rgate: the gate of rounds queued before a deferred graph update is settled; ncand_host: the
  // candidate count once read back (the device slot may then hold a host-side apply's counter)
  const int64_t* rgate = nullptr;
  int64_t ncand_host = -1;
  auto inval_begin = [&]() {
    sepoch = next_epoch(c);
    k_zero2<<<1, 64, 0, c.s>>>(box + 8, 16, box + 28, 5, rgate);
    if (async && m > 0)
      k_mark_pending<<<grid_for(m, kBlock * kLinesPerThread), kBlock, 0, c.s>>>(c.la.p, box, ncand_host, L, cur,
                                                                                box + 8, cap, rgate);
  };
  // queues one round
  auto inval_round = [&]() -> int {
    unsigned ig = round_grid(async, full_rounds);
    int64_t* bcur = cnt[p];
    int64_t* bnext = cnt[p ^ 1];
    // candidates of this level are taken by the round's first kernel, which also lists the
    // long rows and their chunk tasks; k_invalidate searches the medium rows and the
    // tasks; k_inval_push settles the long rows no task supported
    int64_t nlevel = !async && lvl <= cmax ? ncand : 0;
    RoundBook book{bcur,           box + 18,          box + 20,        box + 19,
                   tasks_of[p ^ 1], medium_of[p ^ 1], longs_of[p ^ 1], nlevel > 0};
    book.other_c = small_of[p ^ 1];
    const int64_t* cb = (cbuckets && nlevel > 0) ? cbuckets + (lvl - cmin) : nullptr;
    // two fresh consecutive epochs: supepoch marks support, supepoch - 1 a pending neighbour
    int32_t supepoch = next_epoch(c);
    if (async) {
      supepoch = next_epoch(c);
      if (supepoch == 1) supepoch = next_epoch(c);  // the counter started over
      SCOUT_CHECK(prof_launch(c, PROF_K_INVALIDATE_SHORT, k_invalidate_short<true>, dim3(ig),
                              dim3(kBlock), c.g.view(), L, P, cur, bcur, cand, nlevel, cb, box + 18, lvl, c.ld.p,
                              binval, nxt, bnext, cap, c.stamp, sepoch, c.lh.p, small_of[p], c.le.p, medium_of[p],
                              c.lf.p, c.lg.p, longs_of[p], c.wb.p, tasks_of[p], c.wb.cap, box + 23, c.d_paths,
                              ab, rgate));
      SCOUT_CHECK(prof_launch(c, PROF_BFS_INVALIDATE_KERNEL, k_invalidate<true>, dim3(ig), dim3(kBlock),
                              c.g.view(), L, P, c.lh.p, small_of[p], c.le.p, medium_of[p], c.lf.p, c.lg.p, c.wb.p, tasks_of[p],
                              c.wb.cap, c.aux, supepoch, lvl, c.ld.p, binval, cap, nxt, bnext, cap, c.stamp,
                              sepoch, bscan, box + 21, c.d_paths, c.g.sorted, src, ab, rgate));
      SCOUT_CHECK(prof_launch(c, PROF_K_INVAL_PUSH, k_inval_push<true>, dim3(ig), dim3(kBlock),
                              c.g.view(), L, P, lvl, c.lf.p, c.lg.p, c.wb.p, tasks_of[p], c.wb.cap, c.aux,
                              supepoch, c.ld.p, binval, cap, nxt, bnext, cap, c.stamp, sepoch, box + 15,
                              c.d_paths, ab, book, rgate));
    } else {
      SCOUT_CHECK(prof_launch(c, PROF_K_INVALIDATE_SHORT, k_invalidate_short<false>, dim3(ig),
                              dim3(kBlock), c.g.view(), L, P, cur, bcur, cand, nlevel, cb, box + 18, lvl, c.ld.p,
                              binval, nxt, bnext, cap, c.stamp, sepoch, c.lh.p, small_of[p], c.le.p, medium_of[p],
                              c.lf.p, c.lg.p, longs_of[p], c.wb.p, tasks_of[p], c.wb.cap, box + 23, c.d_paths,
                              ab, rgate));
      SCOUT_CHECK(prof_launch(c, PROF_BFS_INVALIDATE_KERNEL, k_invalidate<false>, dim3(ig),
                              dim3(kBlock), c.g.view(), L, P, c.lh.p, small_of[p], c.le.p, medium_of[p], c.lf.p, c.lg.p, c.wb.p,
                              tasks_of[p], c.wb.cap, c.aux, supepoch, lvl, c.ld.p, binval, cap, nxt, bnext, cap,
                              c.stamp, sepoch, bscan, box + 21, c.d_paths, c.g.sorted, src, ab, rgate));
      SCOUT_CHECK(prof_launch(c, PROF_K_INVAL_PUSH, k_inval_push<false>, dim3(ig), dim3(kBlock),
                              c.g.view(), L, P, lvl, c.lf.p, c.lg.p, c.wb.p, tasks_of[p], c.wb.cap, c.aux,
                              supepoch, c.ld.p, binval, cap, nxt, bnext, cap, c.stamp, sepoch, box + 15,
                              c.d_paths, ab, book, rgate));
    }
    lvl++;
    std::swap(cur, nxt);
    p ^= 1;
    return SCOUT_OK;
  };
  // The asynchronous order takes nothing from the host in its first rounds: they are queued
  // behind the read-back of the candidate counts and predictions, and run while the host
  // reads those.
  // A graph update still to settle (scout_commit) gates the early rounds, which read rows; it
  // settles after the counts below are taken, and if it was not complete on the device the
  // early rounds did not run and are queued again.
  bool early = async && m > 0;
  int queued = 0;  // rounds of the first group already queued
  if (early) {
    SCOUT_CHECK(box_read_async(c, 0, learned ? 28 : 8));
    SCOUT_CHECK(prof_begin(c, PROF_BFS_INVALIDATE));
    if (c.apply_pend.on) rgate = box + kApplyGateSlot;
    inval_begin();
    for (; queued < kFirstGroup; queued++) SCOUT_CHECK(inval_round());
    rgate = nullptr;
    SCOUT_CHECK(box_wait(c));
  } else {
    SCOUT_CHECK(box_read(c, learned ? 28 : 8));
  }
  ncand = c.h_box[0];
  ncand_host = ncand;
  cmin = (int32_t)c.h_box[1];
  cmax = (int32_t)c.h_box[2];
  st->deletions += c.h_box[3];
  int64_t pred_box[4] = {c.h_box[24], c.h_box[25], c.h_box[26], c.h_box[27]};
  if (c.apply_pend.on) {
    bool complete = true;
    SCOUT_CHECK(apply_settle(c, &complete));
    if (!complete && early) {
      early = false;
      queued = 0;
    }
  }
  int64_t nx = 0, scanned = 0, rounds = 0;
  int64_t nsample = std::min<int64_t>(ncand, sample);
  int first_group = kFirstGroup;
  int64_t predicted_last = -1, predicted_depth = -1;
  if (learned && ncand > 0) {
    // predictions size and order the work; the plain search below decides every level
    double predicted = (double)pred_box[0] * (double)ncand / (double)nsample;
    // the split head's expected cut-off, kept for this batch's CC repair
    c.pred_cut = (int64_t)((double)pred_box[3] / 16.0 * (double)ncand / (double)nsample);
    c.pred_cut_seq = c.apply_seq;
    if (predicted > kRecomputeFraction * (double)n) {
      // cascades over a large part of the graph: one full BFS costs less than local repair
      st->fallback = 1;
      return bfs_full_on(c, L, P, src, &st->rounds, &st->edges_scanned);
    }
    // rounds cmin .. max(predicted last level, deepest candidate): one read-back when the
    // prediction holds; an empty extra round costs as much as a read-back, so no slack. The
    // prediction is a maximum over sampled candidates, which overshoots: it is lowered by the
    // running mean of its overshoot on earlier batches.
    if (c.bfs_async) {
      // asynchronous rounds: one per level of the deepest cascade, plus the round that finds
      // its last children supported. [26] holds minus the largest predicted depth over the
      // sampled candidates; the deepest cascade of the whole batch is rarely among them, so
      // the prediction is corrected by the running mean of its error on earlier batches
      // (rounded up: an empty extra round costs less than a second read-back).
      predicted_depth = std::max<int64_t>(-pred_box[2], 1) + 1;
      int64_t want = predicted_depth + (int64_t)std::ceil(c.pred_bias);
      first_group = (int)std::min<int64_t>(std::max<int64_t>(want, kFirstGroup), kMaxGroup);
    } else {
      predicted_last = std::max<int64_t>(pred_box[1] - (int64_t)std::floor(c.pred_over), cmax);
      first_group = (int)std::min<int64_t>(std::max<int64_t>(predicted_last - cmin + 1, kFirstGroup), kMaxGroup);
    }
  }
  // a schedule handed in for this repair (the rounds a plain repair of the batch needed)
  int hint_inval = c.hint_inval, hint_relax = c.hint_relax;
  c.hint_inval = c.hint_relax = 0;
  if (L == c.level && hint_inval > 0) first_group = hint_inval;
  // The asynchronous relax re-attaches the invalidated cascades and runs out from the
  // insertions; under the learned strategy its rounds follow the predicted depth, corrected by
  // the running mean of their difference on earlier batches (rounded up, as for the
  // invalidation).
  int relax_group = kFirstGroup;
  if (learned && predicted_depth >= 0)
    relax_group = (int)std::min<int64_t>(
        std::max<int64_t>(predicted_depth + (int64_t)std::ceil(c.relax_bias), kFirstGroup), kMaxGroup);
  if (L == c.level && hint_relax > 0) relax_group = hint_relax;
  Relax relax{c, L, P, 0, 0, c.wa.p, 0, 0, true, box + 4, seed_cap};
  // work that follows the relax (scout_commit's CC scan) may run behind its read-backs, on the
  // maintained levels and when no training data is logged (the log reads the candidate list)
  if (L == c.level && !logging && c.after_relax) relax.ahead = &c.after_relax;
  bool relax_queued = false;
  // Queues what follows the invalidation in the asynchronous relax: the tentative levels of the
  // invalidated vertices (their count from the device), the insertion seeds and the relax's
  // first group. No read-back: the seeds' count stays on the device for the relax's first
  // round, and the seed and tentative-level counters come back with the relax's read-back.
  // gate: queued ahead of the read-back that ends an invalidation group (k_spec_gate has
  // zeroed the counters and kept the invalidated count).
  auto queue_relax = [&](const int64_t* gate, bool tentative) -> int {
    if (tentative) {
      // [36] scanned, [37] medium rows, [38] chunk tasks (slots the relax leaves alone)
      if (!gate) k_zero<<<1, 64, 0, c.s>>>(box + 36, 3);
      const int64_t* nx_dev = gate ? box + kNxSlot : box + 10;
      uint64_t* tent = reinterpret_cast<uint64_t*>(c.wc.p);
      k_tentative_short<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), L, c.ld.p, nx_dev, tent, c.lf.p, box + 37,
                                                        c.lf.cap, c.wb.p, box + 38, c.wb.cap, box + 36, gate);
      k_tentative_rows<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), L, c.ld.p, c.lf.p, box + 37, c.wb.p, box + 38,
                                                       c.wb.cap, tent, box + 36, c.d_paths, gate);
      k_apply_tentative<<<kTentGrid, kBlock, 0, c.s>>>(c.ld.p, tent, nx_dev, L, P, c.wa.p, c.lg.p, seed_cap, box,
                                                        gate);
    }
    if (m > 0)
      k_insert_seeds<<<grid_for(m, kBlock * kLinesPerThread), kBlock, 0, c.s>>>(
          c.g.view(), c.g.sorted, ch.pairs_distinct, ch.d_uv.p, ch.d_ins.p, m, L, c.wa.p, c.lg.p, seed_cap, box, gate);
    k_seed_parents<<<kListGrid, kBlock, 0, c.s>>>(c.wa.p, c.lg.p, 0, box + 4, seed_cap, L, P, gate);
    relax.epoch = next_epoch(c);
    if (!gate) k_zero<<<1, 64, 0, c.s>>>(box + 8, 1);
    relax_queued = true;
    return relax.start(relax_group, gate, gate != nullptr);
  };
  // A first invalidation group sized by a prediction (or a schedule handed in) is expected to
  // end the invalidation: the relax is queued behind its read-back under the gate, so the
  // device does not wait for the host between the two phases.
  bool speculate = async && c.relax_async && ((learned && predicted_depth >= 0) || (L == c.level && hint_inval > 0));
  if (ncand == 0 && queued > 0) SCOUT_CHECK(prof_end(c, PROF_BFS_INVALIDATE, 0));
  if (ncand > 0) {
    if (!early) {
      lvl = cmin;
      int64_t cspan = (int64_t)cmax - cmin + 1;
      if (!async && cspan >= 1 && cspan <= kMaxBuckets) {
        // each round then reads only its own level's candidates
        bucket_by_level(c.s, VertexKey{c.la.p, L}, ncand, cmin, (int)cspan, c.lvlb.p, c.lvlb.p + kMaxBuckets + 1,
                        c.lg.p);
        cand = c.lg.p;
        cbuckets = c.lvlb.p;
      }
      inval_begin();
    }
    int group = first_group;
    while (true) {
      if (queued == 0) SCOUT_CHECK(prof_begin(c, PROF_BFS_INVALIDATE));
      for (; queued < group; queued++) SCOUT_CHECK(inval_round());
      queued = 0;
      SCOUT_CHECK(prof_end(c, PROF_BFS_INVALIDATE, 0));
      // [8..30]: the prediction slots [24..27] come back unchanged
      if (speculate) {
        SCOUT_CHECK(box_read_async(c, 8, 23));
        k_spec_gate<<<1, 64, 0, c.s>>>(box + 8 + p, box);
        SCOUT_CHECK(queue_relax(box + kGateSlot, true));
        SCOUT_CHECK(box_wait(c));
      } else {
        SCOUT_TRY(hipMemcpyAsync(c.h_box + 8, box + 8, 23 * sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
        SCOUT_TRY(hipStreamSynchronize(c.s));
        c.readbacks++;
      }
      if (std::max({c.h_box[14], c.h_box[19], c.h_box[30]}) > c.wb.cap) {
        set_error("bfs invalidation: chunk task list overflow");
        return SCOUT_ERR_STATE;
      }
      if (c.h_box[8 + p] == 0 && (async || lvl > cmax)) break;
      // the gate stayed shut: nothing queued behind the read-back ran
      speculate = false;
      relax_queued = false;
      full_rounds = c.h_box[8 + p] > kTailFront ? kFullRounds : 0;
      // after a predicted first group the repair is near its end: start again from a short group
      group = group == first_group && first_group != kFirstGroup ? kFirstGroup : std::min(2 * group, kMaxGroup);
      first_group = kFirstGroup;
    }
    rounds += c.h_box[20];
    st->inval_rounds += c.h_box[20];
    if (predicted_depth >= 0) {
      c.pred_bias = 0.75 * c.pred_bias + 0.25 * (double)(c.h_box[20] - predicted_depth);
    } else if (predicted_last >= 0) {
      int64_t last = cmin + c.h_box[20] - 1;  // deepest invalidation level
      double over = std::max<double>(0.0, (double)(pred_box[1] - last));
      c.pred_over = 0.75 * c.pred_over + 0.25 * over;
    }
    // per candidate its id, level, row offsets and queue entry (40 B); per scanned entry the
    // neighbour and its level (8 B)
    int64_t inval_scanned = c.h_box[11] + c.h_box[15] + c.h_box[23];
    prof_bytes(c, PROF_BFS_INVALIDATE, 40 * c.h_box[18] + 8 * inval_scanned);
    // k_invalidate_short: every candidate (40 B) and the entries of the short rows (8 B)
    prof_bytes(c, PROF_K_INVALIDATE_SHORT, 40 * c.h_box[18] + 8 * c.h_box[23]);
    // k_invalidate alone: per row its id, offsets, level and parent write (36 B); per entry
    // read the neighbour and its level, per buffered child candidate its entry and parent (8 B)
    prof_bytes(c, PROF_BFS_INVALIDATE_KERNEL, 36 * c.h_box[22] + 8 * c.h_box[21]);
    // k_inval_push: per entry of an unsupported long row the neighbour and its parent (8 B)
    prof_bytes(c, PROF_K_INVAL_PUSH, 8 * c.h_box[15]);
    nx = c.h_box[10];
    scanned += inval_scanned;
    if (c.bfs_bound > 0 && nx > c.bfs_bound * ncand) {
      // the cascades passed the bound (src/bfs/scout_bfs.cu:475-488 per deletion, here over the
      // batch's candidates): the batch is settled by a full BFS before it commits
      st->fallback = 1;
      st->touched += nx;
      st->rounds += rounds;
      st->edges_scanned += scanned;
      SCOUT_CHECK(bfs_full_on(c, L, P, src, &st->rounds, &st->edges_scanned));
      if (logging) SCOUT_CHECK(mlp_end_batch(c, c.la.p, nsample, L));
      return SCOUT_OK;
    }
    if (nx > 0 && !c.relax_async) {
      // [36] scanned, [37] medium rows, [38] chunk tasks (slots the relax leaves alone)
      k_zero<<<1, 64, 0, c.s>>>(box + 36, 3);
      uint64_t* tent = reinterpret_cast<uint64_t*>(c.wc.p);
      k_tentative_short<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), L, c.ld.p, box + 10, tent, c.lf.p, box + 37,
                                                        c.lf.cap, c.wb.p, box + 38, c.wb.cap, box + 36, nullptr);
      k_tentative_rows<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), L, c.ld.p, c.lf.p, box + 37, c.wb.p, box + 38,
                                                       c.wb.cap, tent, box + 36, c.d_paths, nullptr);
      k_apply_tentative<<<kTentGrid, kBlock, 0, c.s>>>(c.ld.p, tent, box + 10, L, P, c.wa.p, c.lg.p, seed_cap, box,
                                                        nullptr);
    }
  }
  if (m > 0 && !c.relax_async)
    k_insert_seeds<<<grid_for(m, kBlock * kLinesPerThread), kBlock, 0, c.s>>>(
        c.g.view(), c.g.sorted, ch.pairs_distinct, ch.d_uv.p, ch.d_ins.p, m, L, c.wa.p, c.lg.p, seed_cap, box,
        nullptr);
  int64_t nseeds = 0;
  if (c.relax_async) {
    if (!relax_queued) SCOUT_CHECK(queue_relax(nullptr, nx > 0));
    int64_t r0 = rounds;
    SCOUT_CHECK(relax.finish(&rounds, &scanned));
    st->relax_rounds += rounds - r0;
    if (learned && predicted_depth >= 0)
      c.relax_bias = 0.75 * c.relax_bias + 0.25 * (double)(rounds - r0 - predicted_depth);
    nseeds = c.h_box[4];
  }
  if (!c.relax_async) SCOUT_CHECK(box_read(c, 39));
  if (nx > 0) scanned += c.h_box[36];
  if (nx > 0 && c.h_box[38] > c.wb.cap) {
    set_error("bfs tentative levels: chunk task list overflow");
    return SCOUT_ERR_STATE;
  }
  st->insertions += c.h_box[7];
  nseeds = c.h_box[4];
  if (nseeds > seed_cap) {
    set_error("bfs: seed list overflow");
    return SCOUT_ERR_STATE;
  }
  st->seeds += nseeds;
  st->touched += nx;
  if (!c.relax_async && nseeds > 0) {
    int32_t smin = (int32_t)c.h_box[5], smax = (int32_t)c.h_box[6];
    k_seed_parents<<<grid_for(nseeds), kBlock, 0, c.s>>>(c.wa.p, c.lg.p, nseeds, nullptr, seed_cap, L, P, nullptr);
    int32_t epoch = next_epoch(c);
    k_zero<<<1, 64, 0, c.s>>>(box + 8, 1);
    // the learned scheduler's predicted last level bounds the relax as well
    int relax_group = kFirstGroup;
    if (learned && predicted_last >= 0)
      relax_group = (int)std::min<int64_t>(
          std::max<int64_t>(std::max<int64_t>(predicted_last, smax) - smin + 1, kFirstGroup), kMaxGroup);
    if (L == c.level && hint_relax > 0) relax_group = hint_relax;
    int64_t r0 = rounds;
    SCOUT_CHECK(relax_levels(c, L, P, smin, smax, c.wa.p, nseeds, epoch, &rounds, &scanned, relax_group, false));
    st->relax_rounds += rounds - r0;
  }
  st->rounds += rounds;
  st->edges_scanned += scanned;
  st->noop += (ncand == 0 && nseeds == 0) ? m : 0;
  if (logging) SCOUT_CHECK(mlp_end_batch(c, c.la.p, nsample, L));
  return SCOUT_OK;
}

int bfs_full(Ctx& c, int32_t src, int64_t* rounds, int64_t* scanned) {
  SCOUT_CHECK(ensure_state(c, SCOUT_STATE_LEVEL));
  SCOUT_CHECK(bfs_full_on(c, c.level, c.parent, src, rounds, scanned));
  c.level_seq = c.apply_seq;
  return SCOUT_OK;
}

int bfs_local(Ctx& c, Channel& ch, int strategy, scout_stats* st) {
  SCOUT_CHECK(bfs_local_on(c, ch, c.level, c.parent, strategy, st));
  c.level_seq = c.apply_seq;
  return SCOUT_OK;
}

int repair_bfs(Ctx& c, Channel& ch, int strategy, scout_stats* st) {
  if (!c.g.boff || !c.level) {
    set_error("repair_bfs: graph and levels must be loaded first");
    return SCOUT_ERR_STATE;
  }
  TouchScope touch(c);
  Span span(c);
  SCOUT_CHECK(span.begin(c.s));
  st->lines += ch.m;
  int64_t rb0 = c.readbacks;
  // a local repair needs levels that were exact for the graph before this batch
  bool stale = c.level_seq != c.apply_seq - 1;
  if (strategy == SCOUT_STRATEGY_RECOMPUTE || stale) {
    bool complete;
    SCOUT_CHECK(apply_settle(c, &complete));  // a full BFS reads every row
    st->fallback = 1;
    SCOUT_CHECK(bfs_full(c, c.source, &st->rounds, &st->edges_scanned));
  } else {
    SCOUT_CHECK(bfs_local(c, ch, strategy, st));
  }
  SCOUT_CHECK(span.end(c.s));
  st->repair_ms += span.ms();
  st->readbacks += c.readbacks - rb0;
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

int scout_bfs_schedule(scout_ctx* x, int inval_rounds, int relax_rounds) {
  if (!x || inval_rounds < 0 || relax_rounds < 0 || inval_rounds > 64 || relax_rounds > 64) {
    set_error("bfs_schedule: rounds must be 0..64");
    return SCOUT_ERR_ARG;
  }
  x->c.hint_inval = inval_rounds;
  x->c.hint_relax = relax_rounds;
  return SCOUT_OK;
}

int scout_bfs_bound(scout_ctx* x, int64_t bound) {
  if (!x || bound < 0) {
    set_error("bfs_bound: bound must be >= 0");
    return SCOUT_ERR_ARG;
  }
  x->c.bfs_bound = bound;
  return SCOUT_OK;
}

int scout_bfs_mode(scout_ctx* x, int mode) {
  if (!x || (mode != SCOUT_BFS_ASYNC && mode != SCOUT_BFS_LEVEL_SYNC && mode != SCOUT_BFS_ASYNC_INVALIDATION)) {
    set_error("bfs_mode: mode must be SCOUT_BFS_ASYNC, SCOUT_BFS_LEVEL_SYNC or SCOUT_BFS_ASYNC_INVALIDATION");
    return SCOUT_ERR_ARG;
  }
  x->c.bfs_async = mode != SCOUT_BFS_LEVEL_SYNC;
  x->c.relax_async = mode == SCOUT_BFS_ASYNC;
  return SCOUT_OK;
}

int scout_repair_bfs(scout_ctx* x, int channel, int strategy, scout_stats* st) {
  if (!x || !st) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  Channel* ch = channel_of(c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  SCOUT_CHECK(repair_bfs(c, *ch, strategy, st));
  return channel_used(c.s, *ch);
}

}  // extern "C"
