// graph.hip: the HBM-resident dynamic graph -- upload of the base CSR, batch classification,
// the scratchpad update rules, insertion-ordered rebuild of the inserted rows, logical degrees,
// download and compaction.
//
// The update rules restate DeltaOverlay (core/FlashGraph.h:200-310) on the encoding described
// in internal.h:
//   "v in removed[u]"    <=>  base row u holds ~v, or inserted row u holds the tombstone ~v
//   "v in added[u]" (k)  <=>  inserted row u holds v (k copies, insertion order)
// Lines of one batch whose unordered pairs are distinct touch disjoint (row, value) slots, so
// one thread per line applies them in parallel with the sequential result; a batch that
// repeats a pair is applied as consecutive runs of distinct pairs.
#include <omp.h>

#include <algorithm>
#include <cstring>
#include <unordered_set>

#include "common.cuh"

namespace scout {

namespace {

__global__ void k_widen_offsets(const int32_t* in, int64_t* out, int64_t count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (int64_t)in[i];
}

// Rows non-decreasing by decoded id, ids in range: flags[0] = unsorted rows, flags[1] =
// out-of-range entries. Thread per vertex for rows up to 64 entries; longer rows are listed
// for k_check_long_rows (wave per row).
__global__ void k_check_rows(const int64_t* boff, const int32_t* bcol, int32_t n, int64_t* flags, int32_t* longs,
                             int64_t* nlong, int64_t cap) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool unsorted = false, bad = false, is_long = false;
  if (v < n) {
    int64_t b = boff[v], e = boff[v + 1];
    if (e - b > 64) {
      is_long = true;
    } else {
      int32_t prev = -1;
      for (int64_t k = b; k < e; k++) {
        int32_t x = decode(bcol[k]);
        if (x < 0 || x >= n) bad = true;
        if (x < prev) unsorted = true;
        prev = x;
      }
    }
  }
  block_append(is_long, (int32_t)v, longs, nlong, cap);
  block_add(unsorted ? 1 : 0, &flags[0]);
  block_add(bad ? 1 : 0, &flags[1]);
}

__global__ void k_check_long_rows(const int64_t* boff, const int32_t* bcol, int32_t n, const int32_t* longs,
                                  const int64_t* nlong, int64_t* flags) {
  int64_t nl = *nlong;
  for (int64_t i = global_wave(); i < nl; i += total_waves()) {
    int64_t b = boff[longs[i]], e = boff[longs[i] + 1];
    bool unsorted = false, bad = false;
    for (int64_t k = b + lane_id(); k < e; k += kWave) {
      int32_t x = decode(bcol[k]);
      if (x < 0 || x >= n) bad = true;
      if (k + 1 < e && decode(bcol[k + 1]) < x) unsorted = true;
    }
    if (__ballot(unsorted) && lane_id() == 0) atomicAdd((unsigned long long*)&flags[0], 1ULL);
    if (__ballot(bad) && lane_id() == 0) atomicAdd((unsigned long long*)&flags[1], 1ULL);
  }
}

__device__ __forceinline__ uint64_t pair_key(int32_t u, int32_t v) {
  uint32_t a = (uint32_t)(u < v ? u : v), b = (uint32_t)(u < v ? v : u);
  return ((uint64_t)a << 32) | (uint64_t)b;
}

__device__ __forceinline__ uint64_t hash64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

constexpr uint64_t kEmpty = ~0ULL;

// Inserts the unordered pair {u, v} into an open-addressing set (at most a quarter full);
// *count += 1 when another line claimed it first.
__device__ __forceinline__ void pair_set_insert(int32_t u, int32_t v, uint64_t* hkey, uint64_t mask,
                                                int64_t* count) {
  uint64_t key = pair_key(u, v);
  uint64_t h = hash64(key) & mask;
  while (true) {
    uint64_t prev = atomicCAS((unsigned long long*)&hkey[h], (unsigned long long)kEmpty, (unsigned long long)key);
    if (prev == kEmpty) return;
    if (prev == key) {
      atomicAdd((unsigned long long*)count, 1ULL);
      return;
    }
    h = (h + 1) & mask;
  }
}

// *count += the lines whose unordered pair an earlier-claiming line holds (m minus the
// distinct pairs).
__global__ void k_dup_count(const int32_t* uv, int64_t m, uint64_t* hkey, uint64_t mask, int64_t* count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) pair_set_insert(uv[2 * i], uv[2 * i + 1], hkey, mask, count);
}

// ins[i] = 1 iff the pair is absent from the rule's graph. rule 0: current logical graph;
// rule 1: the loaded base graph (decoded, i.e. before any update).
// Every kSampleStride-th base entry, decoded (deleted entries keep their position and id).
__global__ void k_base_samples(const int32_t* bcol, int64_t E, int32_t* samp) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k * kSampleStride < E) samp[k] = decode(bcol[k * kSampleStride]);
}

// The row of every base entry, into out (a wave per row).
__global__ void k_twin_rows(const int64_t* boff, int32_t n, int32_t* out) {
  for (int64_t u = global_wave(); u < n; u += total_waves()) {
    int64_t b = boff[u], e = boff[u + 1];
    for (int64_t k = b + lane_id(); k < e; k += kWave) out[k] = (int32_t)u;
  }
}

// twin[k] (holding the row u of entry k on entry) becomes the offset in row v = entry k's
// neighbour of its one entry for u, or -1; *missing counts the -1s.
__global__ void k_twin_search(GView g, int64_t E, int32_t* twin, int64_t* missing) {
  int64_t miss = 0;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < E; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t k = base + threadIdx.x;
    if (k < E) {
      int32_t u = twin[k];
      int32_t v = decode(g.bcol[k]);
      int64_t lo, hi;
      base_range_sampled(g, true, v, u, lo, hi);
      twin[k] = hi - lo == 1 ? (int32_t)(lo - g.boff[v]) : -1;
      miss += hi - lo != 1;
    }
  }
  block_add(miss, missing);
}

// pos (optional): the base range of row u found for each line, reused by k_apply.
// hkey (optional): the repeated-pair count of k_dup_count is taken here too, into *repeated.
// by_short_row: every base entry has exactly one mirror entry (g.twin complete), so the pair
// stands in row u as it stands in row v; the shorter base row of the two is searched, and row
// u's entry is found from the mirror offset.
__global__ void k_classify(GView g, bool sorted, const int32_t* uv, int64_t m, int rule, uint8_t* ins,
                           int64_t* pos, uint64_t* hkey, uint64_t mask, int64_t* repeated, bool by_short_row) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) {
    int32_t u = uv[2 * i], v = uv[2 * i + 1];
    PairState s;
    int64_t lo, hi;
    AHead hu = g.ahead[u];
    AHead hv;
    if (by_short_row) hv = g.ahead[v];
    if (by_short_row && hv.blen < hu.blen) {
      s = pair_state_h(g, sorted, v, hv, u);
      lo = hi = hu.b0;
      if (s.hi - s.lo == 1) {
        lo = hu.b0 + g.twin[s.lo];
        hi = lo + 1;
      }
    } else {
      s = pair_state_h(g, sorted, u, hu, v);
      lo = s.lo;
      hi = s.hi;
    }
    if (pos) {
      pos[2 * i] = lo;
      pos[2 * i + 1] = hi;
    }
    bool adj = rule == SCOUT_CLASSIFY_BASE ? s.base_any : adjacent_now(s);
    ins[i] = adj ? 0 : 1;
    if (hkey) pair_set_insert(u, v, hkey, mask, repeated);
  }
}

// Flips the base entries of the pair in the range pair_state found (v <-> ~v).
__device__ __forceinline__ void flip_base(const GView& g, const PairState& s, int32_t v, bool to_tomb) {
  int32_t* col = const_cast<int32_t*>(g.bcol);
  int32_t from = to_tomb ? v : ~v, to = to_tomb ? ~v : v;
  for (int64_t k = s.lo; k < s.hi; k++)
    if (col[k] == from) col[k] = to;
}

// First inserted-row entry of u equal to x set to HOLE. True if one was found.
__device__ __forceinline__ bool hole_first(const GView& g, int32_t u, int32_t x) {
  int32_t* col = const_cast<int32_t*>(g.acol);
  AHead h = g.ahead[u];
  for (int64_t k = h.off, e = k + h.len; k < e; k++) {
    if (col[k] == x) {
      col[k] = HOLE;
      return true;
    }
  }
  return false;
}

struct NewEntry {
  bool take;
  int32_t src, dst;
};

// Batch lines per thread in k_apply: fewer blocks, so fewer block-wide appends.
constexpr int kApplyLines = 4;

// The DeltaOverlay rules for the lines of [lo, hi). done[i] = 1 if the line was applied.
// New inserted-row entries (key src << ks | line - lo, value dst; ks bits hold every line
// offset, so the sort needs only the key's low ks + log2(n) bits) are appended with one atomic
// per block. box[0] = new inserted-row entries, box[1] = holes created.
// pos (optional): base ranges of row u located by k_classify (the base rows' entries do not
// move within a batch; only their deletion marks flip).
__global__ void k_apply(GView g, bool sorted, const int32_t* uv, const uint8_t* ins, const int64_t* pos, int64_t lo,
                        int64_t hi, int conditional, uint8_t* done, uint64_t* nkey, int32_t* ndst, int64_t ncap,
                        int ks, int64_t* box, const int64_t* skip, int32_t* ncnt) {
  // the range repeats a pair (*skip, counted by k_dup_count just before): nothing is applied
  // here, and the host applies the range as runs of distinct pairs instead
  if (skip && *skip != 0) return;
  // kApplyLines lines per thread (each pass coalesced over the block): one block-wide append
  // of the new entries per block
  constexpr int K = kApplyLines;
  int64_t first = lo + (int64_t)blockIdx.x * blockDim.x * K + threadIdx.x;
  bool take[2 * K];
  uint64_t keys[2 * K];
  int32_t dsts[2 * K];
  int64_t holes = 0;
#pragma unroll
  for (int q = 0; q < K; q++) {
    int64_t i = first + (int64_t)q * blockDim.x;
    NewEntry e1{false, 0, 0}, e2{false, 0, 0};
    if (i < hi) {
      int32_t u = uv[2 * i], v = uv[2 * i + 1];
      bool insertion = ins[i] != 0;
      PairState su = pos ? pair_state_in(g, u, v, pos[2 * i], pos[2 * i + 1]) : pair_state(g, sorted, u, v);
      bool apply = !conditional || insertion != adjacent_now(su);
      done[i] = apply ? 1 : 0;
      if (apply && insertion) {
        // un_tombstone (FlashGraph.h:284-291), else append to both inserted rows.
        if (su.base_tomb || su.added_tomb) {
          if (su.base_tomb) flip_base(g, su, v, false);
          if (su.added_tomb) holes += hole_first(g, u, ~v);
          PairState sv = pair_state_twin(g, sorted, v, u, mirror_offset(g, su));
          if (sv.base_tomb) flip_base(g, sv, u, false);
          if (sv.added_tomb) holes += hole_first(g, v, ~u);
        } else {
          e1 = NewEntry{true, u, v};
          e2 = NewEntry{true, v, u};
        }
      } else if (apply) {
        // forget_delta_edge (FlashGraph.h:295-310), else tombstone both directions.
        if (su.added_live) {
          holes += hole_first(g, u, v);
          holes += hole_first(g, v, u);
        } else {
          if (su.base_live) flip_base(g, su, v, true);
          else if (!su.base_any && !su.added_tomb) e1 = NewEntry{true, u, ~v};
          PairState sv = pair_state_twin(g, sorted, v, u, mirror_offset(g, su));
          if (sv.base_live) flip_base(g, sv, u, true);
          else if (!sv.base_any && !sv.added_tomb) e2 = NewEntry{true, v, ~u};
        }
      }
    }
    // each source's new entries are counted here for their placement (k_place_single)
    if (e1.take) atomicAdd(&ncnt[e1.src], 1);
    if (e2.take) atomicAdd(&ncnt[e2.src], 1);
    take[2 * q] = e1.take;
    keys[2 * q] = ((uint64_t)(uint32_t)e1.src << ks) | (uint64_t)(i - lo);
    dsts[2 * q] = e1.dst;
    take[2 * q + 1] = e2.take;
    keys[2 * q + 1] = ((uint64_t)(uint32_t)e2.src << ks) | (uint64_t)(i - lo);
    dsts[2 * q + 1] = e2.dst;
  }
  int64_t at[2 * K];
  block_append_k(take, keys, nkey, &box[0], ncap, at);
#pragma unroll
  for (int q = 0; q < 2 * K; q++)
    if (at[q] >= 0) ndst[at[q]] = dsts[q];
  block_add(holes, &box[1]);
}

// --- inserted rows: appends into the slack pool ---------------------------------------
//
// A batch's new entries (key src << ks | line offset, value dst) are placed after k_apply
// without a host round trip: k_apply counts each source's new entries (ncnt); k_place_single
// places every entry whose source has no other (a group of one, in line order trivially) and
// lists the others; k_group_small groups those by source in one block, each source's run in
// line order (so a row keeps the batch's line order); k_place_multi places each run; k_move_rows copies the long rows that
// outgrew their span. Counts of the lists live in the mailbox, so every launch sizes itself on
// the device.

__device__ __forceinline__ int32_t grown_capacity(int64_t len) {
  int64_t c = 2 * len;
  if (c < 4) c = 4;
  return (int32_t)((c + 3) & ~3LL);
}

// Rows moved by one thread at most this long; longer ones by k_move_rows (wave per row).
constexpr int64_t kMoveThread = 64;

// Rows moving to fresh pool space that one thread does not copy: row u, its new span at pos,
// and its k new entries at dst[s..s + k) of the single entries (src = -1 - s, k = 1) or of the
// grouped multi entries (src = s).
struct MoveList {
  int32_t* u;
  int64_t* pos;
  int64_t* src;
  int32_t* k;
  int64_t* count;
};

// Places k new entries of row u (dst[0..k)) after the h.len entries of its header h, or moves
// the row to the span of `need` slots reserved at pos (need = 0: the span has room).
__device__ __forceinline__ void place_group(AHead* ahead, int32_t* acap, int32_t* acol, int32_t* aown,
                                            const int32_t* dst, int64_t k,
                                            int32_t u, AHead h, int64_t need, int64_t pos, int64_t pool_cap,
                                            int64_t src, const MoveList& mv, int64_t* overflow) {
  int32_t len = h.len;
  if (need == 0) {
    int64_t at = h.off + len;
    for (int64_t j = 0; j < k; j++) {
      acol[at + j] = dst[j];
      aown[at + j] = u;
    }
    ahead[u].len = (int32_t)(len + k);
  } else if (pos + need > pool_cap) {
    atomicAdd((unsigned long long*)overflow, 1ULL);
  } else if (len <= kMoveThread) {
    for (int64_t j = 0; j < len; j++) {
      acol[pos + j] = acol[h.off + j];
      aown[pos + j] = u;
    }
    for (int64_t j = 0; j < k; j++) {
      acol[pos + len + j] = dst[j];
      aown[pos + len + j] = u;
    }
    set_inserted(ahead, acap, u, pos, (int32_t)(len + k), (int32_t)need);
  } else {
    int64_t slot = (int64_t)atomicAdd((unsigned long long*)mv.count, 1ULL);
    mv.u[slot] = u;
    mv.pos[slot] = pos;
    mv.src[slot] = src;
    mv.k[slot] = (int32_t)k;
  }
}

// Thread per new entry (nkey/ndst, *count of them): an entry whose source has no other new
// entry is placed here and its source's count goes back to zero; the others go to the multi
// list (mkey/mdst, *nmulti). The span of a row that must move is reserved on *top with one
// atomic per block; the pool has room for the worst case (checked on the host before the
// batch), and *overflow counts reservations past it all the same.
__global__ void k_place_single(AHead* ahead, int32_t* acap, int32_t* acol, int32_t* aown,
                               const uint64_t* nkey, const int32_t* ndst, const int64_t* count, int64_t cap, int ks,
                               int32_t* ncnt, uint64_t* mkey, int32_t* mdst, int64_t* nmulti, int64_t* top,
                               int64_t pool_cap, MoveList mv, int64_t* overflow) {
  int64_t cnt = min(*count, cap);
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool have = i < cnt;
  uint64_t key = have ? nkey[i] : 0;
  int32_t dst = have ? ndst[i] : 0;
  int32_t u = (int32_t)(key >> ks);
  bool single = have && ncnt[u] == 1;
  int64_t pm = block_append(have && !single, key, mkey, nmulti, cap);
  if (pm >= 0) mdst[pm] = dst;
  AHead h{0, 0, 0};
  int64_t need = 0;
  if (single) {
    h = ahead[u];
    if (h.len + 1 > acap[u]) need = grown_capacity(h.len + 1);
  }
  int64_t pos = block_reserve(need, top);
  if (!single) return;
  ncnt[u] = 0;
  place_group(ahead, acap, acol, aown, &ndst[i], 1, u, h, need, pos, pool_cap, -1 - i, mv, overflow);
}

// Multi entries one block groups (at most kSmallSort / 2 sources, each with two or more
// entries, so the LDS table of kGroupSlots sources is at most half full), and the largest run
// it ranks by counting (a run of k costs k reads per entry).
constexpr int kSmallSort = 8192;
constexpr int kGroupSlots = 8192;
static_assert(kGroupSlots == 1 << 13, "the hash keeps the top 13 bits");
constexpr int kGroupMaxRun = 256;
constexpr int kSortThreads = 1024;
constexpr int kSortPerThread = kSmallSort / kSortThreads;

// One block: groups the *count multi entries of kin/vin by source into kout/vout, each
// source's run in line order (the order a sort by key gives inside each run; runs come in any
// order). Sources go to an LDS hash table (slot, run size, run start); each entry is written
// into its run at a cursor, then moved to run start + the number of its run's keys below its
// own. More than kSmallSort entries, or a run longer than kGroupMaxRun: sets *too_many and
// writes nothing (the host sorts them then). tmp_key holds the unordered runs.
__global__ __launch_bounds__(kSortThreads) void k_group_small(const uint64_t* kin, const int32_t* vin,
                                                              const int64_t* count, int64_t cap, int ks,
                                                              uint64_t* tmp_key, uint64_t* kout,
                                                              int32_t* vout, int64_t* too_many) {
  __shared__ int32_t hsrc[kGroupSlots];
  __shared__ int32_t hcnt[kGroupSlots];
  __shared__ int32_t hoff[kGroupSlots];
  __shared__ int32_t wsum[kSortThreads / kWave];
  __shared__ int32_t longest;
  int64_t n = min(*count, cap);
  if (n > kSmallSort) {
    if (threadIdx.x == 0) *too_many = 1;
    return;
  }
  int t = (int)threadIdx.x;
  for (int h = t; h < kGroupSlots; h += kSortThreads) {
    hsrc[h] = -1;
    hcnt[h] = 0;
  }
  if (t == 0) longest = 0;
  __syncthreads();
  uint64_t key[kSortPerThread];
  int32_t dst[kSortPerThread], slot[kSortPerThread];
#pragma unroll
  for (int r = 0; r < kSortPerThread; r++) {
    int64_t i = (int64_t)r * kSortThreads + t;
    slot[r] = -1;
    if (i < n) {
      key[r] = kin[i];
      dst[r] = vin[i];
      int32_t src = (int32_t)(key[r] >> ks);
      uint32_t h = ((uint32_t)src * 2654435761u) >> (32 - 13);  // kGroupSlots = 2^13
      while (true) {
        int32_t prev = atomicCAS(&hsrc[h], -1, src);
        if (prev == -1 || prev == src) break;
        h = (h + 1) & (kGroupSlots - 1);
      }
      slot[r] = (int32_t)h;
      atomicAdd(&hcnt[h], 1);
    }
  }
  __syncthreads();
  // run starts: exclusive scan of the run sizes (contiguous slots per thread, then the waves)
  constexpr int kPer = kGroupSlots / kSortThreads;
  int run = 0, big = 0;
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    int c = hcnt[t * kPer + q];
    run += c;
    big = max(big, c);
  }
  if (big > 0) atomicMax(&longest, big);
  int x = run;
  for (int off = 1; off < kWave; off <<= 1) {
    int y = __shfl_up(x, off);
    if (lane_id() >= off) x += y;
  }
  if (lane_id() == kWave - 1) wsum[t / kWave] = x;
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int w = 0; w < kSortThreads / kWave; w++) {
      int c = wsum[w];
      wsum[w] = acc;
      acc += c;
    }
  }
  __syncthreads();
  int start = wsum[t / kWave] + x - run;
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    hoff[t * kPer + q] = start;
    start += hcnt[t * kPer + q];
  }
  __syncthreads();
  if (longest > kGroupMaxRun) {
    if (t == 0) *too_many = 1;
    return;
  }
  // each entry into its run at the run's cursor (hoff moves to the run's end)
  int32_t at[kSortPerThread];
#pragma unroll
  for (int r = 0; r < kSortPerThread; r++) {
    at[r] = -1;
    if (slot[r] >= 0) {
      at[r] = atomicAdd(&hoff[slot[r]], 1);
      tmp_key[at[r]] = key[r];
    }
  }
  __threadfence_block();
  __syncthreads();
  // rank inside the run: its keys below this one (keys are distinct)
#pragma unroll
  for (int r = 0; r < kSortPerThread; r++) {
    if (slot[r] < 0) continue;
    int32_t len = hcnt[slot[r]];
    int32_t lo = hoff[slot[r]] - len;
    int32_t rank = 0;
    for (int32_t j = lo; j < lo + len; j++) rank += tmp_key[j] < key[r];
    kout[lo + rank] = key[r];
    vout[lo + rank] = dst[r];
  }
}

// The grouped multi entries (skey/sdst, *count of them): the first entry of each source's run
// places the run (one thread per run, a block-wide reservation per pass) and zeroes the
// source's count. Nothing when *skip is set (the entries were not sorted).
__global__ void k_place_multi(AHead* ahead, int32_t* acap, int32_t* acol, int32_t* aown,
                              const uint64_t* skey, const int32_t* sdst, const int64_t* count, int64_t cap, int ks,
                              int32_t* ncnt, int64_t* top, int64_t pool_cap, MoveList mv, int64_t* overflow,
                              const int64_t* skip) {
  if (skip && *skip) return;
  int64_t cnt = min(*count, cap);
  // the trip count is uniform across the block, so the block-wide reservations see every thread
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < cnt; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = base + threadIdx.x;
    bool head = i < cnt && (i == 0 || (skey[i] >> ks) != (skey[i - 1] >> ks));
    int32_t u = -1;
    AHead h{0, 0, 0};
    int64_t k = 0, need = 0;
    if (head) {
      u = (int32_t)(skey[i] >> ks);
      int64_t e = i + 1;
      while (e < cnt && (int32_t)(skey[e] >> ks) == u) e++;
      k = e - i;
      h = ahead[u];
      if (h.len + k > acap[u]) need = grown_capacity(h.len + k);
    }
    int64_t pos = block_reserve(need, top);
    if (head) {
      ncnt[u] = 0;
      place_group(ahead, acap, acol, aown, &sdst[i], k, u, h, need, pos, pool_cap, i, mv, overflow);
    }
  }
}

// Wave per long row that moves: its entries, then the new ones, into the reserved span.
__global__ void k_move_rows(AHead* ahead, int32_t* acap, int32_t* acol, int32_t* aown, const int32_t* ndst,
                            const int32_t* sdst,
                            MoveList mv) {
  int64_t nm = *mv.count;
  for (int64_t i = global_wave(); i < nm; i += total_waves()) {
    int32_t u = mv.u[i];
    int64_t src = mv.src[i], k = mv.k[i], pos = mv.pos[i];
    const int32_t* d = src < 0 ? ndst + (-1 - src) : sdst + src;
    AHead h = ahead[u];
    int64_t len = h.len, old = h.off;
    for (int64_t j = lane_id(); j < len; j += kWave) {
      acol[pos + j] = acol[old + j];
      aown[pos + j] = u;
    }
    for (int64_t j = lane_id(); j < k; j += kWave) {
      acol[pos + len + j] = d[j];
      aown[pos + len + j] = u;
    }
    if (lane_id() == 0) set_inserted(ahead, acap, u, pos, (int32_t)(len + k), grown_capacity(len + k));
  }
}

// Mailbox slots of an apply: [0] new entries, [1] holes created, [3] rows listed for
// k_move_rows, [4] pool top, [5] reservations past the pool; the repeated-pair count k_apply
// checks, the multi-entry count and the flag of a multi list k_group_small leaves to the host.
constexpr int kRepeatSlot = 48, kMultiSlot = 51, kSortSlot = 52;

// Zeroes mailbox slots [4..7] (seed counters) after a host-side apply finish.
__global__ void k_apply_rezero(int64_t* box) {
  if (threadIdx.x >= 4 && threadIdx.x < 8) box[threadIdx.x] = 0;
}

// Non-zero when the apply is not complete on the device: the range repeats a pair (*skip) or
// holds more multi entries than one block groups (box[kSortSlot]).
__global__ void k_apply_gate(const int64_t* box, const int64_t* skip, int64_t* gate) {
  if (threadIdx.x == 0) *gate = (skip ? *skip : 0) | box[kSortSlot];
}

// First kernel of an apply: zeroes the counts and flags above and sets the pool top.
__global__ void k_apply_begin(int64_t* box, int64_t pool_top) {
  int t = (int)threadIdx.x;
  if (t < 6) box[t] = t == 4 ? pool_top : 0;
  if (t == 6) box[kMultiSlot] = 0;
  if (t == 7) box[kSortSlot] = 0;
}

// --- inserted rows: full re-pack (holes dropped, fresh slack) --------------------------

// Non-hole entries of every row; capacities with half again as slack. *live sums the counts.
__global__ void k_repack_count(GView g, int32_t* cnt, int32_t* cap2, int64_t* live) {
  int64_t c = 0;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base <= g.n; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t v = base + threadIdx.x;
    if (v < g.n) {
      AHead h = g.ahead[v];
      int64_t a0 = h.off, a1 = a0 + h.len;
      int32_t k = 0;
      for (int64_t j = a0; j < a1; j++) k += g.acol[j] != HOLE;
      cnt[v] = k;
      cap2[v] = k ? (int32_t)((k + k / 2 + 4) & ~3) : 0;
      c += k;
    } else if (v == g.n) {
      cap2[v] = 0;
    }
  }
  block_add(c, live);
}

__global__ void k_repack_copy(GView g, const int64_t* aoff2, int32_t* acol2, int32_t* aown2) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += (int64_t)gridDim.x * blockDim.x) {
    AHead h = g.ahead[v];
    int64_t a0 = h.off, a1 = a0 + h.len, o = aoff2[v];
    for (int64_t j = a0; j < a1; j++) {
      int32_t x = g.acol[j];
      if (x != HOLE) {
        acol2[o] = x;
        aown2[o] = (int32_t)v;
        o++;
      }
    }
  }
}

// The re-packed headers: start, live count and capacity of every row (and the n-th), with the
// base rows of boff.
__global__ void k_repack_heads(AHead* ahead2, int32_t* acap, const int64_t* off2, const int32_t* cnt,
                               const int32_t* cap2, const int64_t* boff, int64_t n) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > n) return;
  AHead h;
  h.off = off2[v];
  h.len = v < n ? cnt[v] : 0;
  h.b0 = boff[v];
  h.blen = v < n ? (int32_t)(boff[v + 1] - boff[v]) : 0;
  ahead2[v] = h;
  acap[v] = cap2[v];
}

// Empty inserted rows with the base rows of boff (and the n-th header).
__global__ void k_reset_heads(AHead* ahead, int32_t* acap, const int64_t* boff, int64_t n) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > n) return;
  AHead h;
  h.off = 0;
  h.len = 0;
  h.b0 = boff[v];
  h.blen = v < n ? (int32_t)(boff[v + 1] - boff[v]) : 0;
  ahead[v] = h;
  acap[v] = 0;
}

// Live entries per vertex of the logical graph: thread per vertex for rows up to 64 physical
// entries, longer rows listed for k_logical_degree_long (wave per row).
__global__ void k_logical_degree(GView g, int64_t* deg, int32_t* longs, int64_t* nlong, int64_t cap) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool is_long = false;
  if (v < g.n) {
    Row r = row_of(g, (int32_t)v);
    int64_t len = row_len(r);
    if (len > 64) {
      is_long = true;
    } else {
      int64_t c = 0;
      for (int64_t k = 0; k < len; k++) c += (row_at(g, r, k) >= 0);
      deg[v] = c;
    }
  }
  block_append(is_long, (int32_t)v, longs, nlong, cap);
}

__global__ void k_logical_degree_long(GView g, const int32_t* longs, const int64_t* nlong, int64_t* deg) {
  int64_t nl = *nlong;
  for (int64_t i = global_wave(); i < nl; i += total_waves()) {
    int32_t v = longs[i];
    Row r = row_of(g, v);
    int64_t c = 0;
    for (int64_t k = lane_id(); k < row_len(r); k += kWave) c += (row_at(g, r, k) >= 0);
    c = wave_sum(c);
    if (lane_id() == 0) deg[v] = c;
  }
}

// Live entries of each row copied to the compacted CSR (wave per vertex, order kept).
__global__ void k_gather_live(GView g, const int64_t* off2, int32_t* col2) {
  for (int64_t v = global_wave(); v < g.n; v += total_waves()) {
    Row r = row_of(g, (int32_t)v);
    int64_t len = row_len(r);
    int64_t out = off2[v];
    for (int64_t base = 0; base < len; base += kWave) {
      int64_t k = base + lane_id();
      int32_t x = k < len ? row_at(g, r, k) : -1;
      bool live = x >= 0;
      unsigned long long mask = __ballot(live);
      if (live) col2[out + __popcll(mask & ((1ULL << lane_id()) - 1ULL))] = x;
      out += __popcll(mask);
    }
  }
}

int upload_chunked(Ctx& c, void* dst, const void* src, size_t bytes) {
  // Stream a (possibly mmap'd, pageable) host array through two pinned buffers on the copy
  // stream; filling buffer k overlaps the transfer of buffer k^1.
  const size_t CH = (size_t)64 << 20;
  if (bytes == 0) return SCOUT_OK;
  char* pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2];
  SCOUT_TRY(hipHostMalloc((void**)&pin[0], CH, hipHostMallocDefault));
  SCOUT_TRY(hipHostMalloc((void**)&pin[1], CH, hipHostMallocDefault));
  SCOUT_TRY(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  SCOUT_TRY(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int rc = SCOUT_OK;
  size_t done = 0;
  int k = 0;
  bool used[2] = {false, false};
  while (done < bytes) {
    size_t len = std::min(CH, bytes - done);
    if (used[k]) {
      if (hipEventSynchronize(ev[k]) != hipSuccess) {
        rc = SCOUT_ERR_HIP;
        break;
      }
    }
    const char* s = (const char*)src + done;
    int nt = omp_get_max_threads();
    size_t per = (len + (size_t)nt - 1) / (size_t)nt;
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nt; t++) {
      size_t a = (size_t)t * per;
      if (a < len) memcpy(pin[k] + a, s + a, std::min(per, len - a));
    }
    if (hipMemcpyAsync((char*)dst + done, pin[k], len, hipMemcpyHostToDevice, c.cs) != hipSuccess ||
        hipEventRecord(ev[k], c.cs) != hipSuccess) {
      rc = SCOUT_ERR_HIP;
      break;
    }
    used[k] = true;
    done += len;
    k ^= 1;
  }
  (void)hipStreamSynchronize(c.cs);
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  (void)hipHostFree(pin[0]);
  (void)hipHostFree(pin[1]);
  if (rc != SCOUT_OK) set_error("upload_chunked: HIP copy failed");
  return rc;
}

// Re-packs the inserted rows into a fresh pool: holes dropped, each row with half again its
// size as slack, and room for `headroom` more slots. O(n + pool); runs when holes or moved-away
// spans waste the pool, or the pool is full.
int repack_inserted(Ctx& c, int64_t headroom) {
  DevGraph& g = c.g;
  int32_t n = g.n;
  SCOUT_CHECK(g.rcap.reserve((int64_t)n + 1));
  SCOUT_CHECK(g.roff.reserve((int64_t)n + 1));
  SCOUT_CHECK(g.ahead2.reserve((int64_t)n + 1));
  SCOUT_TRY(hipMemsetAsync(c.d_box + 32, 0, sizeof(int64_t), c.s));
  // live counts staged in moves, new capacities in rcap, new starts (their scan) in roff
  SCOUT_CHECK(g.moves.reserve((int64_t)n + 1));
  k_repack_count<<<kListGrid, kBlock, 0, c.s>>>(g.view(), g.moves.p, g.rcap.p, c.d_box + 32);
  SCOUT_CHECK(scan_exclusive_i32_to_i64(c, g.rcap.p, g.roff.p, (int64_t)n + 1));
  int64_t total = 0;
  SCOUT_TRY(hipMemcpyAsync(&total, g.roff.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipMemcpyAsync(c.h_box + 32, c.d_box + 32, sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  c.readbacks++;
  SCOUT_CHECK(g.acol2.reserve(total + headroom + 1024));
  SCOUT_CHECK(g.aown2.reserve(total + headroom + 1024));
  k_repack_copy<<<kListGrid, kBlock, 0, c.s>>>(g.view(), g.roff.p, g.acol2.p, g.aown2.p);
  k_repack_heads<<<grid_for((int64_t)n + 1), kBlock, 0, c.s>>>(g.ahead2.p, g.acap.p, g.roff.p, g.moves.p, g.rcap.p,
                                                                g.boff, n);
  std::swap(g.ahead, g.ahead2);
  std::swap(g.acol, g.acol2);
  std::swap(g.aown, g.aown2);
  g.A = total;
  g.live = c.h_box[32];
  g.holes = 0;
  return SCOUT_OK;
}

// Slots usable in the pool: both the entry and the owner arrays hold them.
int64_t pool_cap(const DevGraph& g) { return std::min(g.acol.cap, g.aown.cap); }

// Appends this batch's new entries (nkey/ndst, cnt_new of them) to their rows.
// Bits that hold every value in [0, x] (at least 1).
int bits_for(int64_t x) {
  int b = 1;
  while (b < 63 && (x >> b) != 0) b++;
  return b;
}


// Blocks of k_place_multi (grid-stride: the multi entries are few, their count is on the device).
constexpr unsigned kPlaceGrid = 256;

// Placement counters reset for a second pass over the multi entries.
__global__ void k_place_reset(int64_t* box, int64_t pool_top) {
  box[3] = 0;
  box[4] = pool_top;
  box[5] = 0;
}

MoveList move_list(Ctx& c) {
  DevGraph& g = c.g;
  return MoveList{g.moves.p, g.move_pos.p, g.move_src.p, g.move_k.p, c.d_box + 3};
}

// Applies lines [lo, hi): k_apply, then the new inserted-row entries placed as described above
// the placement kernels, with one read-back at the end. With `skip` (a device count of repeated
// pairs in the range, read back with the rest), nothing is applied when the count is not zero,
// and *skipped says so.
int apply_counts(Ctx& c, Channel& ch, int64_t lo, int64_t hi, int conditional, const int64_t* h, int skip_slot,
                 bool* skipped);

// defer: the counters are not read back here; they go to c.h_apply with the work queued next,
// and apply_settle finishes the apply (scout_commit's graph update).
int apply_range(Ctx& c, Channel& ch, int64_t lo, int64_t hi, int conditional, const int64_t* skip = nullptr,
                bool* skipped = nullptr, bool defer = false) {
  DevGraph& g = c.g;
  int64_t m = hi - lo;
  int64_t cap = 4 * m + 16;
  SCOUT_CHECK(g.nkey.reserve(cap));
  SCOUT_CHECK(g.ndst.reserve(cap));
  SCOUT_CHECK(g.nkey2.reserve(cap));
  SCOUT_CHECK(g.ndst2.reserve(cap));
  SCOUT_CHECK(g.mkey.reserve(cap));
  SCOUT_CHECK(g.mdst.reserve(cap));
  SCOUT_CHECK(g.moves.reserve(cap));
  SCOUT_CHECK(g.move_pos.reserve(cap));
  SCOUT_CHECK(g.move_src.reserve(cap));
  SCOUT_CHECK(g.move_k.reserve(cap));
  if (!g.ncnt.p || g.ncnt.cap < (int64_t)g.n + 1) {
    g.ncnt.release();
    SCOUT_CHECK(g.ncnt.reserve((int64_t)g.n + 1));
    SCOUT_TRY(hipMemsetAsync(g.ncnt.p, 0, (size_t)g.ncnt.cap * sizeof(int32_t), c.s));
  }
  // Room for the worst case before anything is placed: at most two new entries per line, and
  // every touched row moving to twice its new size (rows hold at most the pool's used part):
  // 2 * (A + new) + 4 per group <= 2 * A + 6 * new.
  int64_t bound = 2 * m;
  int64_t worst = 2 * g.A + 6 * bound + 1024;
  if (g.A + worst > pool_cap(g)) {
    if (g.A > 2 * g.live + (1 << 20)) SCOUT_CHECK(repack_inserted(c, 2 * g.live + 6 * bound));
    worst = 2 * g.A + 6 * bound + 1024;
    if (g.A + worst > pool_cap(g)) SCOUT_CHECK(repack_inserted(c, 2 * (g.A + worst)));
  }
  k_apply_begin<<<1, 64, 0, c.s>>>(c.d_box, g.A);
  SCOUT_CHECK(prof_begin(c, PROF_APPLY));
  const int64_t* pos = (ch.pos_gen == g.base_gen && ch.pos_lo <= lo && hi <= ch.pos_hi) ? ch.d_pos.p : nullptr;
  int ks = bits_for(std::max<int64_t>(m - 1, 0));  // key bits of a line offset
  if (m > 0)
    k_apply<<<grid_for(m, kBlock * kApplyLines), kBlock, 0, c.s>>>(g.view(), g.sorted, ch.d_uv.p, ch.d_ins.p, pos, lo,
                                                                   hi, conditional, ch.d_done.p, g.nkey.p, g.ndst.p,
                                                                   g.nkey.cap, ks, c.d_box, skip, g.ncnt.p);
  // algorithmic bytes per line: the pair and its label and done flag (10 B), both rows'
  // base and inserted offsets (32 B) and the entries flipped or appended (8 B)
  SCOUT_CHECK(prof_end(c, PROF_APPLY, 50 * m));
  MoveList mv = move_list(c);
  int64_t pcap = pool_cap(g);
  if (m > 0) {
    k_place_single<<<grid_for(bound), kBlock, 0, c.s>>>(g.ahead.p, g.acap.p, g.acol.p, g.aown.p, g.nkey.p,
                                                       g.ndst.p, c.d_box, cap, ks, g.ncnt.p, g.mkey.p, g.mdst.p,
                                                       c.d_box + kMultiSlot, c.d_box + 4, pcap, mv, c.d_box + 5);
    k_group_small<<<1, kSortThreads, 0, c.s>>>(g.mkey.p, g.mdst.p, c.d_box + kMultiSlot, cap, ks, g.nkey.p,
                                               g.nkey2.p, g.ndst2.p, c.d_box + kSortSlot);
    k_place_multi<<<kPlaceGrid, kBlock, 0, c.s>>>(g.ahead.p, g.acap.p, g.acol.p, g.aown.p, g.nkey2.p,
                                                  g.ndst2.p, c.d_box + kMultiSlot, cap, ks, g.ncnt.p, c.d_box + 4, pcap,
                                                  mv, c.d_box + 5, c.d_box + kSortSlot);
    k_move_rows<<<kListGrid, kBlock, 0, c.s>>>(g.ahead.p, g.acap.p, g.acol.p, g.aown.p, g.ndst.p,
                                               g.ndst2.p, mv);
  }
  if (defer) {
    // the counters travel to c.h_apply behind the placement; the gate tells the kernels queued
    // before the host has read them whether the apply is complete on the device
    SCOUT_TRY(hipMemcpyAsync(c.h_apply, c.d_box, (size_t)(kSortSlot + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                             c.s));
    if (!c.apply_ev) SCOUT_TRY(hipEventCreateWithFlags(&c.apply_ev, hipEventDisableTiming));
    SCOUT_TRY(hipEventRecord(c.apply_ev, c.s));
    k_apply_gate<<<1, 64, 0, c.s>>>(c.d_box, skip, c.d_box + kApplyGateSlot);
    c.apply_pend = ApplyPending{true, &ch, lo, hi, conditional, skip ? (int)(skip - c.d_box) : -1};
    return SCOUT_OK;
  }
  SCOUT_CHECK(box_read(c, kSortSlot + 1));
  return apply_counts(c, ch, lo, hi, conditional, c.h_box, skip ? (int)(skip - c.d_box) : -1, skipped);
}

// The host side of apply_range once its counters h (mailbox slots 0..kSortSlot) are on the
// host: the checks, the bookkeeping and, for more multi entries than one block groups, the
// library sort and placement.
int apply_counts(Ctx& c, Channel& ch, int64_t lo, int64_t hi, int conditional, const int64_t* h, int skip_slot,
                 bool* skipped) {
  (void)ch;
  (void)conditional;
  DevGraph& g = c.g;
  int64_t m = hi - lo;
  int64_t cap = 4 * m + 16;
  int ks = bits_for(std::max<int64_t>(m - 1, 0));
  MoveList mv = move_list(c);
  int64_t pcap = pool_cap(g);
  if (skip_slot >= 0 && h[skip_slot] != 0) {
    *skipped = true;
    return SCOUT_OK;
  }
  int64_t cnt_new = h[0];
  if (cnt_new > cap) {
    set_error("apply: inserted-entry staging overflow");
    return SCOUT_ERR_STATE;
  }
  if (h[5] != 0) {
    set_error("apply: inserted-row pool overflow");
    return SCOUT_ERR_STATE;
  }
  g.holes += h[1];
  g.A = h[4];
  if (h[kSortSlot] != 0) {
    // more multi entries than one block sorts: the library sorts them, then they are placed
    int64_t nm = h[kMultiSlot];
    SCOUT_CHECK(sort_pairs_u64(c, g.mkey.p, g.nkey2.p, g.mdst.p, g.ndst2.p, nm,
                               ks + bits_for(std::max<int64_t>(g.n - 1, 0))));
    k_place_reset<<<1, 1, 0, c.s>>>(c.d_box, g.A);
    k_place_multi<<<kPlaceGrid, kBlock, 0, c.s>>>(g.ahead.p, g.acap.p, g.acol.p, g.aown.p, g.nkey2.p,
                                                  g.ndst2.p, c.d_box + kMultiSlot, cap, ks, g.ncnt.p, c.d_box + 4, pcap,
                                                  mv, c.d_box + 5, nullptr);
    k_move_rows<<<kListGrid, kBlock, 0, c.s>>>(g.ahead.p, g.acap.p, g.acol.p, g.aown.p, g.ndst.p,
                                               g.ndst2.p, mv);
    SCOUT_CHECK(box_read(c, 6));
    if (c.h_box[5] != 0) {
      set_error("apply: inserted-row pool overflow");
      return SCOUT_ERR_STATE;
    }
    g.A = c.h_box[4];
  }
  g.live += cnt_new;
  if (g.holes > (g.live >> 1) + (1 << 20)) SCOUT_CHECK(repack_inserted(c, 2 * g.live + (1 << 20)));
  return SCOUT_OK;
}

// Empties size slots of the pair set (16 bytes per store) and zeroes its repeat counter.
__global__ void k_pair_set_clear(uint64_t* hkey, uint64_t size, int64_t* count) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *count = 0;
  ulonglong2* h2 = reinterpret_cast<ulonglong2*>(hkey);
  for (uint64_t k = i; 2 * k < size; k += (uint64_t)gridDim.x * blockDim.x) h2[k] = make_ulonglong2(kEmpty, kEmpty);
}

// Clears the pair set of k_dup_count and k_classify for m lines (a power of two of at least
// 4m slots, so it is at most a quarter full) and zeroes *count, in one launch. *size = its
// slot count.
int pair_set_clear(Ctx& c, int64_t m, uint64_t* size, int64_t* count) {
  uint64_t s = 2;
  while (s < (uint64_t)(4 * m)) s <<= 1;
  SCOUT_CHECK(c.g.hkey.reserve((int64_t)s));
  k_pair_set_clear<<<kListGrid, kBlock, 0, c.s>>>(c.g.hkey.p, s, count);
  *size = s;
  return SCOUT_OK;
}

}  // namespace

int graph_free(Ctx& c) {
  DevGraph& g = c.g;
  if (g.boff) (void)hipFree(g.boff);
  if (g.host_bcol) {
    if (g.hbcol) (void)hipHostFree(g.hbcol);
  } else if (g.bcol) {
    (void)hipFree(g.bcol);
  }
  g.boff = nullptr;
  g.bcol = nullptr;
  g.hbcol = nullptr;
  g.host_bcol = false;
  g.hoff.clear();
  g.hoff.shrink_to_fit();
  g.ahead.release();
  g.ahead2.release();
  g.acap.release();
  g.roff.release();
  g.rcap.release();
  g.acol.release();
  g.acol2.release();
  g.aown.release();
  g.aown2.release();
  g.moves.release();
  g.move_k.release();
  g.move_pos.release();
  g.move_src.release();
  g.nkey.release();
  g.nkey2.release();
  g.ndst.release();
  g.ndst2.release();
  g.hkey.release();
  g.ncnt.release();
  g.mkey.release();
  g.mdst.release();
  g.bsamp.release();
  g.twin.release();
  g.n = 0;
  g.E = g.A = g.live = g.holes = 0;
  g.base_gen++;
  g.shape_gen++;
  return SCOUT_OK;
}

int reset_inserted(Ctx& c) {
  DevGraph& g = c.g;
  size_t n1 = (size_t)g.n + 1;
  SCOUT_CHECK(g.ahead.reserve((int64_t)n1));
  SCOUT_CHECK(g.acap.reserve((int64_t)n1));
  k_reset_heads<<<grid_for((int64_t)n1), kBlock, 0, c.s>>>(g.ahead.p, g.acap.p, g.boff, g.n);
  SCOUT_CHECK(g.acol.reserve(1 << 16));
  SCOUT_CHECK(g.aown.reserve(g.acol.cap));
  g.A = g.live = g.holes = 0;
  return SCOUT_OK;
}

int check_rows(Ctx& c) {
  DevGraph& g = c.g;
  SCOUT_TRY(hipMemsetAsync(c.d_box, 0, 3 * sizeof(int64_t), c.s));
  DevBuf<int32_t> longs;
  SCOUT_CHECK(longs.reserve((int64_t)g.n + 1));
  if (g.n > 0) {
    k_check_rows<<<grid_for(g.n), kBlock, 0, c.s>>>(g.boff, g.bcol, g.n, c.d_box, longs.p, c.d_box + 2, longs.cap);
    k_check_long_rows<<<kListGrid, kBlock, 0, c.s>>>(g.boff, g.bcol, g.n, longs.p, c.d_box + 2, c.d_box);
  }
  SCOUT_TRY(hipStreamSynchronize(c.s));
  longs.release();
  SCOUT_CHECK(box_read(c, 2));
  if (c.h_box[1] != 0) {
    set_error("graph has adjacency entries outside [0, n)");
    return SCOUT_ERR_ARG;
  }
  g.sorted = c.h_box[0] == 0;
  return SCOUT_OK;
}

int graph_upload_egr(Ctx& c, int32_t n, int64_t E, const void* nindex, int wide, const int32_t* nlist,
                     bool host_resident) {
  if (n < 0 || E < 0) {
    set_error("graph_load: negative size");
    return SCOUT_ERR_ARG;
  }
  graph_free(c);
  DevGraph& g = c.g;
  g.n = n;
  g.E = E;
  g.base_gen++;
  g.shape_gen++;
  SCOUT_TRY(hipMalloc((void**)&g.boff, ((size_t)n + 1) * sizeof(int64_t)));
  if (host_resident) {
    // the base entries stay in pinned host memory, filled by OpenMP threads from the (mmap'd)
    // file; kernels read it over PCIe through its mapped device address
    size_t bytes = (size_t)std::max<int64_t>(E, 1) * sizeof(int32_t);
    SCOUT_TRY(hipHostMalloc((void**)&g.hbcol, bytes, hipHostMallocMapped | hipHostMallocNonCoherent));
    g.host_bcol = true;
    const char* src = reinterpret_cast<const char*>(nlist);
    char* dst = reinterpret_cast<char*>(g.hbcol);
    size_t total = (size_t)E * sizeof(int32_t), per = (size_t)64 << 20;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t k = 0; k < (int64_t)((total + per - 1) / per); k++) {
      size_t a = (size_t)k * per;
      memcpy(dst + a, src + a, std::min(per, total - a));
    }
    void* dp = nullptr;
    SCOUT_TRY(hipHostGetDevicePointer(&dp, g.hbcol, 0));
    g.bcol = static_cast<int32_t*>(dp);
    g.hoff.resize((size_t)n + 1);
    for (size_t v = 0; v <= (size_t)n; v++)
      g.hoff[v] = wide ? (int64_t)((const uint64_t*)nindex)[v] : (int64_t)((const int32_t*)nindex)[v];
  } else {
    SCOUT_TRY(hipMalloc((void**)&g.bcol, (size_t)std::max<int64_t>(E, 1) * sizeof(int32_t)));
  }
  if (wide) {
    SCOUT_CHECK(upload_chunked(c, g.boff, nindex, ((size_t)n + 1) * sizeof(int64_t)));
  } else {
    int32_t* tmp32 = nullptr;
    SCOUT_TRY(hipMalloc((void**)&tmp32, ((size_t)n + 1) * sizeof(int32_t)));
    int rc = upload_chunked(c, tmp32, nindex, ((size_t)n + 1) * sizeof(int32_t));
    if (rc == SCOUT_OK) {
      k_widen_offsets<<<grid_for((int64_t)n + 1), kBlock, 0, c.s>>>(tmp32, g.boff, (int64_t)n + 1);
      (void)hipStreamSynchronize(c.s);
    }
    (void)hipFree(tmp32);
    SCOUT_CHECK(rc);
  }
  if (!host_resident) SCOUT_CHECK(upload_chunked(c, g.bcol, nlist, (size_t)E * sizeof(int32_t)));
  int64_t first = 0, last = 0;
  SCOUT_TRY(hipMemcpy(&first, g.boff, sizeof(int64_t), hipMemcpyDeviceToHost));
  SCOUT_TRY(hipMemcpy(&last, g.boff + n, sizeof(int64_t), hipMemcpyDeviceToHost));
  if (first != 0 || last != E) {
    set_error("graph_load: nindex[0] must be 0 and nindex[n] the entry count");
    return SCOUT_ERR_ARG;
  }
  SCOUT_CHECK(reset_inserted(c));
  return check_rows(c);
}

int graph_upload(Ctx& c, int32_t n, int64_t E, const int64_t* off, const int32_t* col) {
  return graph_upload_egr(c, n, E, off, 1, col);
}

// Samples of the base entries for the searches of long sorted rows (taken again whenever the
// base arrays change; a view without them searches the rows directly).
int graph_samples(Ctx& c) {
  DevGraph& g = c.g;
  if (!g.sorted || g.E == 0 || g.bsamp_gen == g.base_gen) return SCOUT_OK;
  int64_t ns = (g.E + kSampleStride - 1) / kSampleStride;
  SCOUT_CHECK(g.bsamp.reserve(ns));
  k_base_samples<<<grid_for(ns), kBlock, 0, c.s>>>(g.bcol, g.E, g.bsamp.p);
  g.bsamp_gen = g.base_gen;
  return SCOUT_OK;
}

// Mirror offsets of the base entries (GView::twin): for the entry of row u holding v, the
// offset in row v of its entry for u (-1 if row v holds u other than exactly once). k_apply
// removes or restores both directions of a base pair; with the offset the second direction is
// one read instead of a search of row v. Built once per base layout (tombstone flips keep it
// valid) when its 4 bytes per entry take at most half the free device memory and no HBM budget
// is set (a budget bounds the graph and the solution state, which the index would crowd).
constexpr int kTwinSlot = 34;  // mailbox slot of the build's count of entries without a mirror

int graph_twins(Ctx& c) {
  DevGraph& g = c.g;
  if (!g.sorted || g.host_bcol || g.E == 0 || c.budget.hbm_bytes > 0 || g.twin_gen == g.shape_gen) return SCOUT_OK;
  if (g.twin.cap < g.E) {
    size_t free_b = 0, total_b = 0;
    SCOUT_TRY(hipMemGetInfo(&free_b, &total_b));
    if (4.0 * (double)g.E > 0.5 * (double)free_b) return SCOUT_OK;
    g.twin.release();
    SCOUT_CHECK(g.twin.reserve(g.E));
  }
  k_twin_rows<<<kListGrid, kBlock, 0, c.s>>>(g.boff, g.n, g.twin.p);
  SCOUT_TRY(hipMemsetAsync(c.d_box + kTwinSlot, 0, sizeof(int64_t), c.s));
  k_twin_search<<<4 * kListGrid, kBlock, 0, c.s>>>(g.view(), g.E, g.twin.p, c.d_box + kTwinSlot);
  SCOUT_TRY(hipMemcpyAsync(c.h_box + kTwinSlot, c.d_box + kTwinSlot, sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  g.twin_gen = g.shape_gen;
  g.twin_symmetric = c.h_box[kTwinSlot] == 0;
  return SCOUT_OK;
}

int graph_classify_lines(Ctx& c, Channel& ch, int64_t lo, int64_t hi, int rule) {
  if (!c.g.boff) {
    set_error("classify: no graph loaded");
    return SCOUT_ERR_STATE;
  }
  SCOUT_CHECK(graph_samples(c));
  SCOUT_CHECK(graph_twins(c));
  int64_t m = hi - lo;
  SCOUT_CHECK(ch.d_pos.reserve(2 * ch.m));
  if (ch.pos_gen != c.g.base_gen || ch.pos_hi != lo) ch.pos_lo = lo;  // else extend the valid range
  ch.pos_hi = hi;
  ch.pos_gen = c.g.base_gen;
  // A classify of every staged line also counts the repeated pairs for the apply that follows
  // (its skip test), into the channel's mailbox slot.
  bool count_repeats = lo == 0 && hi == ch.m && m > 0;
  uint64_t* hkey = nullptr;
  uint64_t size = 0;
  int64_t* repeated = c.d_box + kChannelRepeatSlot + (&ch - c.ch);
  if (count_repeats) {
    SCOUT_CHECK(pair_set_clear(c, m, &size, repeated));
    hkey = c.g.hkey.p;
  }
  SCOUT_CHECK(prof_begin(c, PROF_CLASSIFY));
  if (m > 0)
    k_classify<<<grid_for(m), kBlock, 0, c.s>>>(c.g.view(), c.g.sorted, ch.d_uv.p + 2 * lo, m, rule, ch.d_ins.p + lo,
                                                ch.d_pos.p + 2 * lo, hkey, size - 1, repeated,
                                                c.g.view().twin != nullptr && c.g.twin_symmetric);
  // per line: the pair (8 B), row u's base and inserted offsets (16 B), the located entry
  // (4 B), the label written (1 B) and, when counting repeats, the pair's set slot (8 B)
  SCOUT_CHECK(prof_end(c, PROF_CLASSIFY, (count_repeats ? 37 : 29) * m));
  ch.rep_hi = count_repeats ? hi : ch.rep_hi;
  return SCOUT_OK;
}

int graph_classify(Ctx& c, Channel& ch, int rule) { return graph_classify_lines(c, ch, 0, ch.m, rule); }

// Queues the count of the lines of [lo, hi) whose unordered pair an earlier line of the range
// holds into *out (device).
int queue_repeat_count(Ctx& c, Channel& ch, int64_t lo, int64_t hi, int64_t* out) {
  DevGraph& g = c.g;
  int64_t m = hi - lo;
  uint64_t size = 0;
  SCOUT_CHECK(pair_set_clear(c, m, &size, out));
  if (m > 0) k_dup_count<<<grid_for(m), kBlock, 0, c.s>>>(ch.d_uv.p + 2 * lo, m, g.hkey.p, size - 1, out);
  return SCOUT_OK;
}

// *repeated = the lines of [lo, hi) whose unordered pair an earlier line of the range holds.
int graph_repeated_pairs(Ctx& c, Channel& ch, int64_t lo, int64_t hi, int64_t* repeated) {
  *repeated = 0;
  if (hi <= lo) return SCOUT_OK;
  SCOUT_CHECK(queue_repeat_count(c, ch, lo, hi, c.d_box));
  SCOUT_CHECK(box_read(c, 1));
  *repeated = c.h_box[0];
  return SCOUT_OK;
}

// Applies lines [lo, hi) of the staged batch; pairs repeated inside the range are applied as
// runs of distinct pairs, in line order.
int graph_apply_lines(Ctx& c, Channel& ch, int64_t lo, int64_t hi, int conditional) {
  int64_t m = hi - lo;
  if (m <= 0) return SCOUT_OK;
  // The repeated-pair count goes to the device ahead of k_apply, which applies the range only
  // when no pair repeats; the count comes back with the new-entry count (one read-back). A
  // classify of every staged line has counted them already.
  int slot = kRepeatSlot;
  if (lo == 0 && hi == ch.m && ch.rep_hi == hi) {
    slot = kChannelRepeatSlot + (int)(&ch - c.ch);
  } else {
    SCOUT_CHECK(queue_repeat_count(c, ch, lo, hi, c.d_box + kRepeatSlot));
  }
  bool skipped = false;
  SCOUT_CHECK(apply_range(c, ch, lo, hi, conditional, c.d_box + slot, &skipped));
  ch.pairs_distinct = !skipped && lo == 0 && hi == ch.m;
  if (!skipped) return SCOUT_OK;
  return apply_runs(c, ch, lo, hi, conditional);
}

// Lines [lo, hi) that repeat a pair, applied as runs of distinct pairs in line order.
int apply_runs(Ctx& c, Channel& ch, int64_t lo, int64_t hi, int conditional) {
  int64_t m = hi - lo;
  // Repeated pairs: cut the range into runs of distinct pairs, applied in order.
  std::vector<int32_t> uv((size_t)(2 * m));
  SCOUT_TRY(hipMemcpy(uv.data(), ch.d_uv.p + 2 * lo, (size_t)(2 * m) * sizeof(int32_t), hipMemcpyDeviceToHost));
  std::unordered_set<uint64_t> seen;
  int64_t start = lo;
  for (int64_t i = 0; i < m; i++) {
    uint32_t a = (uint32_t)std::min(uv[2 * i], uv[2 * i + 1]), b = (uint32_t)std::max(uv[2 * i], uv[2 * i + 1]);
    uint64_t key = ((uint64_t)a << 32) | b;
    if (seen.count(key)) {
      SCOUT_CHECK(apply_range(c, ch, start, lo + i, conditional));
      seen.clear();
      start = lo + i;
    }
    seen.insert(key);
  }
  return apply_range(c, ch, start, hi, conditional);
}

int graph_apply(Ctx& c, Channel& ch, int conditional) {
  c.apply_seq++;
  return graph_apply_lines(c, ch, 0, ch.m, conditional);
}

int graph_apply_deferred(Ctx& c, Channel& ch) {
  c.apply_seq++;
  if (ch.m <= 0) return SCOUT_OK;
  int slot = kRepeatSlot;
  if (ch.rep_hi == ch.m) {
    slot = kChannelRepeatSlot + (int)(&ch - c.ch);
  } else {
    SCOUT_CHECK(queue_repeat_count(c, ch, 0, ch.m, c.d_box + kRepeatSlot));
  }
  ch.pairs_distinct = false;  // known once settled
  return apply_range(c, ch, 0, ch.m, 0, c.d_box + slot, nullptr, true);
}

int apply_settle(Ctx& c, bool* complete) {
  *complete = true;
  if (!c.apply_pend.on) return SCOUT_OK;
  ApplyPending a = c.apply_pend;
  c.apply_pend.on = false;
  Channel& ch = *a.ch;
  SCOUT_TRY(hipEventSynchronize(c.apply_ev));
  bool skipped = false;
  int64_t sorted_fallback = c.h_apply[kSortSlot];
  SCOUT_CHECK(apply_counts(c, ch, a.lo, a.hi, a.conditional, c.h_apply, a.skip_slot, &skipped));
  if (skipped) SCOUT_CHECK(apply_runs(c, ch, a.lo, a.hi, a.conditional));
  ch.pairs_distinct = !skipped && a.lo == 0 && a.hi == ch.m;
  *complete = !skipped && sorted_fallback == 0;
  // the host-side finish reused mailbox slots [4..7], which the work queued since expects
  // zeroed (the repair's seed counters)
  if (!*complete) k_apply_rezero<<<1, 64, 0, c.s>>>(c.d_box);
  return SCOUT_OK;
}

int graph_logical_degrees(Ctx& c, int64_t* d_deg) {
  if (c.g.n == 0) return SCOUT_OK;
  DevBuf<int32_t> longs;
  SCOUT_CHECK(longs.reserve((int64_t)c.g.n + 1));
  SCOUT_TRY(hipMemsetAsync(c.d_box + 40, 0, sizeof(int64_t), c.s));
  k_logical_degree<<<grid_for(c.g.n), kBlock, 0, c.s>>>(c.g.view(), d_deg, longs.p, c.d_box + 40, longs.cap);
  k_logical_degree_long<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), longs.p, c.d_box + 40, d_deg);
  SCOUT_TRY(hipStreamSynchronize(c.s));
  longs.release();
  return SCOUT_OK;
}

// Logical CSR into device buffers off2 (n+1) and col2; returns the entry count in *total.
int logical_csr_device(Ctx& c, DevBuf<int64_t>& off2, DevBuf<int32_t>& col2, int64_t* total) {
  DevGraph& g = c.g;
  int32_t n = g.n;
  DevBuf<int64_t> deg;
  SCOUT_CHECK(deg.reserve((int64_t)n + 1));
  SCOUT_TRY(hipMemsetAsync(deg.p + n, 0, sizeof(int64_t), c.s));
  SCOUT_CHECK(graph_logical_degrees(c, deg.p));
  SCOUT_CHECK(off2.reserve((int64_t)n + 1));
  SCOUT_CHECK(scan_exclusive_i64(c, deg.p, off2.p, (int64_t)n + 1));
  SCOUT_TRY(hipMemcpyAsync(total, off2.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  deg.release();
  SCOUT_CHECK(col2.reserve(*total + 1));
  if (n > 0) k_gather_live<<<kListGrid, kBlock, 0, c.s>>>(g.view(), off2.p, col2.p);
  return SCOUT_OK;
}

int graph_download(Ctx& c, int64_t* h_off, int32_t* h_col) {
  DevBuf<int64_t> off2;
  DevBuf<int32_t> col2;
  int64_t total = 0;
  int rc = logical_csr_device(c, off2, col2, &total);
  if (rc == SCOUT_OK) {
    if (hipMemcpyAsync(h_off, off2.p, ((size_t)c.g.n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, c.s) != hipSuccess)
      rc = SCOUT_ERR_HIP;
    if (rc == SCOUT_OK && h_col && total > 0 &&
        hipMemcpyAsync(h_col, col2.p, (size_t)total * sizeof(int32_t), hipMemcpyDeviceToHost, c.s) != hipSuccess)
      rc = SCOUT_ERR_HIP;
    if (hipStreamSynchronize(c.s) != hipSuccess) rc = SCOUT_ERR_HIP;
  }
  off2.release();
  col2.release();
  if (rc == SCOUT_ERR_HIP) set_error("graph_download: copy failed");
  return rc;
}

int graph_compact(Ctx& c) {
  DevGraph& g = c.g;
  if (g.host_bcol) {
    set_error("graph_compact: the base entries are host-resident; download the logical graph and write it on the "
              "host (scout_graph_download + the .egr writer)");
    return SCOUT_ERR_UNSUPPORTED;
  }
  DevBuf<int64_t> off2;
  DevBuf<int32_t> col2;
  int64_t total = 0;
  SCOUT_CHECK(logical_csr_device(c, off2, col2, &total));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  (void)hipFree(g.boff);
  (void)hipFree(g.bcol);
  g.boff = off2.p;
  g.bcol = col2.p;
  off2.p = nullptr;
  col2.p = nullptr;
  g.E = total;
  g.base_gen++;
  g.shape_gen++;
  SCOUT_CHECK(reset_inserted(c));
  return check_rows(c);
}

}  // namespace scout
