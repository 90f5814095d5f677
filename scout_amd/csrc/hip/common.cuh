// common.cuh: device helpers shared by the repair kernels (wave64 row walks, CAS-min,
// work-list appends).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.h"

namespace scout {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = kBlock / kWave;
// Rows longer than this are split into chunk tasks of this many entries, each walked by
// one wave, so a hub's row is spread over many CUs.
constexpr int64_t kChunk = 4096;
// Grid of the grid-stride work-list kernels: 8 blocks per CU on 256 CUs.
constexpr unsigned kListGrid = 2048;

struct Row {
  int64_t b0, b1, a0, a1;
};

__device__ __forceinline__ Row row_of(const GView& g, int32_t v) {
  if (g.touched) {
    uint32_t bit = 1u << (v & 31);
    if (!(g.touched[v >> 5] & bit)) atomicOr(&g.touched[v >> 5], bit);
  }
  AHead h = g.ahead[v];
  return Row{h.b0, h.b1(), h.off, h.off + h.len};
}
__device__ __forceinline__ int64_t row_len(const Row& r) { return (r.b1 - r.b0) + (r.a1 - r.a0); }
// k-th physical entry of the row (base segment first, then the inserted segment). Negative
// values are not neighbours (deleted base entry, tombstone, hole).
__device__ __forceinline__ int32_t row_at(const GView& g, const Row& r, int64_t k) {
  int64_t nb = r.b1 - r.b0;
  return k < nb ? g.bcol[r.b0 + k] : g.acol[r.a0 + (k - nb)];
}
__device__ __forceinline__ int32_t decode(int32_t x) { return x >= 0 ? x : ~x; }

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }
__device__ __forceinline__ int64_t global_wave() {
  return ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
}
__device__ __forceinline__ int64_t total_waves() { return ((int64_t)gridDim.x * blockDim.x) / kWave; }

// Adds a wave's count of a long-row code path to its device counter (one atomic per wave, none
// when the wave never took the path). cnt must be uniform across the wave.
__device__ __forceinline__ void count_path(int64_t cnt, int64_t* p) {
  if (lane_id() == 0 && cnt) atomicAdd((unsigned long long*)p, (unsigned long long)cnt);
}

// Lowers *p to cand if cand is smaller (-1 counts as infinity). True if this call lowered it.
__device__ __forceinline__ bool cas_min_level(int32_t* p, int32_t cand) {
  int32_t old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (old == -1 || old > cand) {
    int32_t prev = atomicCAS(p, old, cand);
    if (prev == old) return true;
    old = prev;
  }
  return false;
}

// Wave-aggregated append: every lane with `take` gets a slot in out[]; one atomic per wave.
__device__ __forceinline__ void wave_append(bool take, int32_t value, int32_t* out, int64_t* count,
                                            int64_t cap) {
  unsigned long long mask = __ballot(take);
  if (mask == 0) return;
  int leader = __ffsll((long long)mask) - 1;
  int64_t base = 0;
  if (lane_id() == leader) base = (int64_t)atomicAdd((unsigned long long*)count, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader);
  if (take) {
    int64_t rank = __popcll(mask & ((1ULL << lane_id()) - 1ULL));
    int64_t pos = base + rank;
    if (pos < cap) out[pos] = value;
  }
}

// Wave-wide append of up to K values per lane (value k kept if take[k]) with one atomic.
template <int K>
__device__ __forceinline__ void wave_append_k(const bool (&take)[K], const int32_t (&val)[K], int32_t* out,
                                              int64_t* count, int64_t cap) {
  int n = 0;
#pragma unroll
  for (int k = 0; k < K; k++) n += take[k];
  int x = n;  // inclusive scan across the wave
  for (int off = 1; off < kWave; off <<= 1) {
    int y = __shfl_up(x, off);
    if (lane_id() >= off) x += y;
  }
  int total = __shfl(x, kWave - 1);
  if (total == 0) return;
  int64_t base = 0;
  if (lane_id() == 0) base = (int64_t)atomicAdd((unsigned long long*)count, (unsigned long long)total);
  int64_t p = __shfl(base, 0) + (x - n);
#pragma unroll
  for (int k = 0; k < K; k++)
    if (take[k]) {
      if (p < cap) out[p] = val[k];
      p++;
    }
}

// Block-wide reductions and appends: one global atomic per block instead of one per thread or
// wave, so O(n) passes do not serialise on a single counter. Every thread of the block must
// call them (no early return before the call).
__device__ __forceinline__ int64_t wave_sum(int64_t v) {
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ void block_add(int64_t v, int64_t* out) {
  __shared__ int64_t part[kWavesPerBlock];
  v = wave_sum(v);
  int w = (int)(threadIdx.x / kWave);
  if (lane_id() == 0) part[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t s = 0;
    for (int i = 0; i < kWavesPerBlock; i++) s += part[i];
    if (s) atomicAdd((unsigned long long*)out, (unsigned long long)s);
  }
  __syncthreads();
}

// Block-wide append of up to K values per thread (value k kept if take[k]) with one atomic
// per block; pos[k] = its position in out, or -1 (not kept, or past cap).
template <int K, typename T>
__device__ __forceinline__ void block_append_k(const bool (&take)[K], const T (&val)[K], T* out, int64_t* count,
                                               int64_t cap, int64_t (&pos)[K]) {
  __shared__ int64_t wsum[kWavesPerBlock];
  __shared__ int64_t gbase;
  int64_t n = 0;
#pragma unroll
  for (int k = 0; k < K; k++) n += take[k];
  int64_t x = n;  // inclusive scan within the wave
  for (int off = 1; off < kWave; off <<= 1) {
    int64_t y = __shfl_up(x, off);
    if (lane_id() >= off) x += y;
  }
  int w = (int)(threadIdx.x / kWave);
  if (lane_id() == kWave - 1) wsum[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int i = 0; i < kWavesPerBlock; i++) {
      int64_t c = wsum[i];
      wsum[i] = t;
      t += c;
    }
    gbase = t ? (int64_t)atomicAdd((unsigned long long*)count, (unsigned long long)t) : 0;
  }
  __syncthreads();
  int64_t p = gbase + wsum[w] + x - n;
#pragma unroll
  for (int k = 0; k < K; k++) {
    pos[k] = -1;
    if (take[k]) {
      if (p < cap) {
        out[p] = val[k];
        pos[k] = p;
      }
      p++;
    }
  }
  __syncthreads();
}

// Block-wide reservation on N counters at once: each thread asks need[k] (< 2^31) slots of
// the k-th counter and gets its first slot in pos[k]. Lanes 0..N-1 of the first wave issue the
// N atomics together, so the block waits for one round trip instead of N in a row. Every
// thread of the block calls it. The counters come as separate arguments, picked per lane by a
// chain of selects: an array of them indexed by lane would live in memory.
template <int N, typename... Ctr>
__device__ __forceinline__ void block_reserve_n(const int32_t (&need)[N], int64_t (&pos)[N], Ctr... count) {
  static_assert(sizeof...(Ctr) == N, "one counter per need");
  __shared__ int64_t wsum[N * kWavesPerBlock];
  __shared__ int64_t gbase[N];
  int w = (int)(threadIdx.x / kWave);
  int32_t x[N];
#pragma unroll
  for (int k = 0; k < N; k++) x[k] = need[k];
  // inclusive scans within the wave, the N of them interleaved
  for (int off = 1; off < kWave; off <<= 1) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      int32_t y = __shfl_up(x[k], off);
      if (lane_id() >= off) x[k] += y;
    }
  }
  if (lane_id() == kWave - 1) {
#pragma unroll
    for (int k = 0; k < N; k++) wsum[k * kWavesPerBlock + w] = x[k];
  }
  __syncthreads();
  if (threadIdx.x < N) {
    int k = (int)threadIdx.x;
    int64_t t = 0;
    for (int i = 0; i < kWavesPerBlock; i++) {
      int64_t c = wsum[k * kWavesPerBlock + i];
      wsum[k * kWavesPerBlock + i] = t;
      t += c;
    }
    int64_t* ctr = nullptr;
    int j = 0;
    ((ctr = (k == j++) ? count : ctr), ...);
    gbase[k] = t ? (int64_t)atomicAdd((unsigned long long*)ctr, (unsigned long long)t) : 0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) pos[k] = gbase[k] + wsum[k * kWavesPerBlock + w] + x[k] - need[k];
  __syncthreads();
}

__device__ __forceinline__ void block_min_max(int64_t lo, int64_t hi, int64_t* out_min, int64_t* out_max) {
  __shared__ int64_t pmin[kWavesPerBlock], pmax[kWavesPerBlock];
  for (int off = kWave / 2; off > 0; off >>= 1) {
    lo = min(lo, (int64_t)__shfl_xor(lo, off));
    hi = max(hi, (int64_t)__shfl_xor(hi, off));
  }
  int w = (int)(threadIdx.x / kWave);
  if (lane_id() == 0) {
    pmin[w] = lo;
    pmax[w] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t a = pmin[0], b = pmax[0];
    for (int i = 1; i < kWavesPerBlock; i++) {
      a = min(a, pmin[i]);
      b = max(b, pmax[i]);
    }
    // only improvements go to the shared words (most blocks bring none once a few have)
    if (a < __hip_atomic_load(out_min, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMin((long long*)out_min, (long long)a);
    if (b > __hip_atomic_load(out_max, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax((long long*)out_max, (long long)b);
  }
  __syncthreads();
}

// Appends `value` for every thread with `take`, reserving the block's slots with one atomic.
// Returns the slot (or -1).
template <class T>
__device__ __forceinline__ int64_t block_append(bool take, T value, T* out, int64_t* count, int64_t cap) {
  __shared__ int64_t wbase[kWavesPerBlock];
  __shared__ int64_t gbase;
  unsigned long long mask = __ballot(take);
  int w = (int)(threadIdx.x / kWave);
  if (lane_id() == 0) wbase[w] = __popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t s = 0;
    for (int i = 0; i < kWavesPerBlock; i++) {
      int64_t c = wbase[i];
      wbase[i] = s;
      s += c;
    }
    gbase = s ? (int64_t)atomicAdd((unsigned long long*)count, (unsigned long long)s) : 0;
  }
  __syncthreads();
  int64_t pos = -1;
  if (take) {
    pos = gbase + wbase[w] + __popcll(mask & ((1ULL << lane_id()) - 1ULL));
    if (pos < cap) out[pos] = value;
    else pos = -1;
  }
  __syncthreads();
  return pos;
}

// Sparse appends over a long grid-stride loop: each wave stages its takes in an LDS segment
// of kSeg entries and global slots are reserved once per block at the end (or once per wave
// when its segment fills). A sparse list spread over many waves would otherwise cost one
// global atomic on the same counter per wave step, which serialises at one L2 channel.
// Every thread of the block must call stage_finish once, after the loop.
template <int kSeg>
struct WaveStage {
  int32_t* seg;  // this wave's segment in LDS
  int n;         // entries staged (uniform across the wave)
};

template <int kSeg>
__device__ __forceinline__ void stage_flush(WaveStage<kSeg>& ws, int32_t* out, int64_t* count, int64_t cap) {
  int64_t base = 0;
  if (lane_id() == 0) base = (int64_t)atomicAdd((unsigned long long*)count, (unsigned long long)ws.n);
  base = __shfl(base, 0);
  for (int k = lane_id(); k < ws.n; k += kWave)
    if (base + k < cap) out[base + k] = ws.seg[k];
  ws.n = 0;
}

template <int kSeg>
__device__ __forceinline__ void stage_push(WaveStage<kSeg>& ws, bool take, int32_t value, int32_t* out,
                                           int64_t* count, int64_t cap) {
  unsigned long long mask = __ballot(take);
  if (mask == 0) return;
  int cnt = __popcll(mask);
  if (ws.n + cnt > kSeg) {
    __threadfence_block();
    stage_flush(ws, out, count, cap);
  }
  if (take) ws.seg[ws.n + __popcll(mask & ((1ULL << lane_id()) - 1ULL))] = value;
  ws.n += cnt;
}

template <int kSeg>
__device__ __forceinline__ void stage_finish(WaveStage<kSeg>& ws, int32_t* out, int64_t* count, int64_t cap) {
  __shared__ int64_t wn[kWavesPerBlock];
  __shared__ int64_t gbase;
  __threadfence_block();
  int w = (int)(threadIdx.x / kWave);
  if (lane_id() == 0) wn[w] = ws.n;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int i = 0; i < kWavesPerBlock; i++) {
      int64_t c = wn[i];
      wn[i] = t;
      t += c;
    }
    gbase = t ? (int64_t)atomicAdd((unsigned long long*)count, (unsigned long long)t) : 0;
  }
  __syncthreads();
  int64_t base = gbase + wn[w];
  for (int k = lane_id(); k < ws.n; k += kWave)
    if (base + k < cap) out[base + k] = ws.seg[k];
  __syncthreads();
}

// Block-wide reservation: every thread asks for `need` slots of a shared counter; returns the
// start of this thread's slots. One global atomic per block.
__device__ __forceinline__ int64_t block_reserve(int64_t need, int64_t* top) {
  __shared__ int64_t wsum[kWavesPerBlock];
  __shared__ int64_t gbase;
  // inclusive scan within the wave
  int64_t x = need;
  for (int off = 1; off < kWave; off <<= 1) {
    int64_t y = __shfl_up(x, off);
    if (lane_id() >= off) x += y;
  }
  int w = (int)(threadIdx.x / kWave);
  if (lane_id() == kWave - 1) wsum[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int i = 0; i < kWavesPerBlock; i++) {
      int64_t c = wsum[i];
      wsum[i] = t;
      t += c;
    }
    gbase = t ? (int64_t)atomicAdd((unsigned long long*)top, (unsigned long long)t) : 0;
  }
  __syncthreads();
  int64_t pos = gbase + wsum[w] + x - need;
  __syncthreads();
  return pos;
}

// Per-thread append (when lanes of one wave are not converged on the same list).
__device__ __forceinline__ void thread_append(int32_t value, int32_t* out, int64_t* count, int64_t cap) {
  int64_t pos = (int64_t)atomicAdd((unsigned long long*)count, 1ULL);
  if (pos < cap) out[pos] = value;
}

// Wave walks the physical entries [lo, hi) of row r; f(x) for every live neighbour x.
template <class F>
__device__ __forceinline__ void wave_walk(const GView& g, const Row& r, int64_t lo, int64_t hi, F&& f) {
  for (int64_t k = lo + lane_id(); k < hi; k += kWave) {
    int32_t x = row_at(g, r, k);
    if (x >= 0) f(x);
  }
}

// wave_walk with kU entries per lane loaded before any is used (kU loads in flight per lane
// on long rows).
template <int kU, class F>
__device__ __forceinline__ void wave_walk_wide(const GView& g, const Row& r, int64_t lo, int64_t hi, F&& f) {
  for (int64_t base = lo; base < hi; base += (int64_t)kU * kWave) {
    int32_t xs[kU];
#pragma unroll
    for (int j = 0; j < kU; j++) {
      int64_t k = base + (int64_t)j * kWave + lane_id();
      xs[j] = k < hi ? row_at(g, r, k) : -1;
    }
#pragma unroll
    for (int j = 0; j < kU; j++)
      if (xs[j] >= 0) f(xs[j]);
  }
}

// Wave searches entries [lo, hi) of row r for a live neighbour satisfying pred. All lanes
// return the same answer: the index of the first such entry, or -1.
template <class P>
__device__ __forceinline__ int64_t wave_find_first(const GView& g, const Row& r, int64_t lo, int64_t hi, P&& pred) {
  for (int64_t base = lo; base < hi; base += kWave) {
    int64_t k = base + lane_id();
    bool hit = false;
    if (k < hi) {
      int32_t x = row_at(g, r, k);
      hit = x >= 0 && pred(x);
    }
    unsigned long long mask = __ballot(hit);
    if (mask) return base + (__ffsll((long long)mask) - 1);
  }
  return -1;
}

// Range [lo, hi) of the base row [b, e) whose entries decode to v (sorted rows: binary
// search). For unsorted rows returns the whole row and callers compare every entry.
__device__ __forceinline__ void base_range_in(const GView& g, bool sorted, int64_t b, int64_t e, int32_t v,
                                              int64_t& lo, int64_t& hi) {
  if (!sorted) {
    lo = b;
    hi = e;
    return;
  }
  int64_t l = b, h = e;
  while (l < h) {
    int64_t mid = l + ((h - l) >> 1);
    if (decode(g.bcol[mid]) < v) l = mid + 1;
    else h = mid;
  }
  lo = l;
  int64_t k = l;
  while (k < e && decode(g.bcol[k]) == v) k++;
  hi = k;
}

// base_range_in over base row u.
__device__ __forceinline__ void base_range(const GView& g, bool sorted, int32_t u, int32_t v, int64_t& lo,
                                           int64_t& hi) {
  base_range_in(g, sorted, g.boff[u], g.boff[u + 1], v, lo, hi);
}

// base_range_in through the entry samples when the row is long: the search first narrows to
// one sample stride, so a hub row costs a few cached sample reads and one stride of the row.
// (Kept apart from base_range_in, whose smaller code suits the kernels that inline it many
// times.)
__device__ __forceinline__ void base_range_sampled_in(const GView& g, bool sorted, int64_t b, int64_t e, int32_t v,
                                                      int64_t& lo, int64_t& hi) {
  if (!sorted || !g.bsamp || e - b <= 4 * kSampleStride) {
    base_range_in(g, sorted, b, e, v, lo, hi);
    return;
  }
  // first sample position p = kSampleStride * k in [b, e) whose entry is >= v: the lower
  // bound lies after the sample before it and at or before p (or after the last sample)
  int64_t k0 = (b + kSampleStride - 1) / kSampleStride, k1 = (e - 1) / kSampleStride;
  int64_t sl = k0, sh = k1 + 1;
  while (sl < sh) {
    int64_t mid = sl + ((sh - sl) >> 1);
    if (g.bsamp[mid] < v) sl = mid + 1;
    else sh = mid;
  }
  int64_t l = sl > k0 ? (sl - 1) * kSampleStride + 1 : b;
  int64_t h = sl <= k1 ? sl * kSampleStride : e;
  while (l < h) {
    int64_t mid = l + ((h - l) >> 1);
    if (decode(g.bcol[mid]) < v) l = mid + 1;
    else h = mid;
  }
  lo = l;
  int64_t k = l;
  while (k < e && decode(g.bcol[k]) == v) k++;
  hi = k;
}

// base_range_sampled_in over base row u.
__device__ __forceinline__ void base_range_sampled(const GView& g, bool sorted, int32_t u, int32_t v, int64_t& lo,
                                                   int64_t& hi) {
  base_range_sampled_in(g, sorted, g.boff[u], g.boff[u + 1], v, lo, hi);
}

struct PairState {
  bool base_live;     // base row u holds a live v
  bool base_tomb;     // base row u holds ~v
  bool base_any;      // base row u holds v or ~v
  bool added_live;    // inserted row u holds v
  bool added_tomb;    // inserted row u holds the tombstone ~v
  int64_t lo, hi;     // base entries of row u searched for v (base_range)
};

// pair_state over the base range [lo, hi) of row u already located; hu is u's header.
__device__ __forceinline__ PairState pair_state_range(const GView& g, const AHead& hu, int32_t v, int64_t lo,
                                                      int64_t hi) {
  PairState s{false, false, false, false, false, 0, 0};
  s.lo = lo;
  s.hi = hi;
  for (int64_t k = lo; k < hi; k++) {
    int32_t x = g.bcol[k];
    if (x == v) s.base_live = true;
    else if (x == ~v) s.base_tomb = true;
  }
  s.base_any = s.base_live || s.base_tomb;
  int64_t ab = hu.off, ae = ab + hu.len;
  for (int64_t k = ab; k < ae; k++) {
    int32_t x = g.acol[k];
    if (x == v) s.added_live = true;
    else if (x == ~v) s.added_tomb = true;
  }
  return s;
}

// pair_state over a base range [lo, hi) of row u already located.
__device__ __forceinline__ PairState pair_state_in(const GView& g, int32_t u, int32_t v, int64_t lo, int64_t hi) {
  return pair_state_range(g, g.ahead[u], v, lo, hi);
}

// pair_state with u's header hu already loaded.
__device__ __forceinline__ PairState pair_state_h(const GView& g, bool sorted, int32_t u, const AHead& hu,
                                                  int32_t v) {
  int64_t lo, hi;
  base_range_sampled_in(g, sorted, hu.b0, hu.b1(), v, lo, hi);
  return pair_state_range(g, hu, v, lo, hi);
}

__device__ __forceinline__ PairState pair_state(const GView& g, bool sorted, int32_t u, int32_t v) {
  return pair_state_h(g, sorted, u, g.ahead[u], v);
}

// The mirror offset of row u's one base entry for v (su.lo, when su found exactly one): the
// offset of v's entry for u in base row v, or -1.
__device__ __forceinline__ int32_t mirror_offset(const GView& g, const PairState& su) {
  return g.twin && su.hi - su.lo == 1 ? g.twin[su.lo] : -1;
}

// pair_state(g, sorted, u, v) when row u's one base entry for v is expected at offset t of the
// base row (a mirror offset): that entry is read instead of searching the row (the search runs
// when t is -1 or the entry holds another neighbour).
__device__ __forceinline__ PairState pair_state_twin(const GView& g, bool sorted, int32_t u, int32_t v, int32_t t) {
  AHead hu = g.ahead[u];
  if (t >= 0 && t < hu.blen && decode(g.bcol[hu.b0 + t]) == v)
    return pair_state_range(g, hu, v, hu.b0 + t, hu.b0 + t + 1);
  return pair_state_h(g, sorted, u, hu, v);
}

// HybridGraph::isAdjacent (core/FlashGraph.h:356-366): inserted copy -> true; tombstone ->
// false; else the base row.
__device__ __forceinline__ bool adjacent_now(const PairState& s) {
  if (s.added_live) return true;
  if (s.base_tomb || s.added_tomb) return false;
  return s.base_live;
}

// adjacent_now(pair_state(g, sorted, u, v)) computed by a whole wave, for long rows: lane 0
// searches a sorted base row (through the samples), an unsorted one and the inserted row are
// compared 64 entries at a time. Uniform across the wave.
__device__ __forceinline__ bool wave_adjacent(const GView& g, bool sorted, int32_t u, int32_t v) {
  AHead hu = g.ahead[u];
  int64_t b = hu.b0, e = hu.b1();
  unsigned long long live = 0, tomb = 0;
  if (sorted) {
    int flags = 0;
    if (lane_id() == 0) {
      int64_t lo, hi;
      base_range_sampled_in(g, true, b, e, v, lo, hi);
      for (int64_t k = lo; k < hi; k++) {
        int32_t x = g.bcol[k];
        flags |= (x == v ? 1 : 0) | (x == ~v ? 2 : 0);
      }
    }
    flags = __shfl(flags, 0);
    live = flags & 1;
    tomb = flags & 2;
  } else {
    for (int64_t base = b; base < e; base += kWave) {
      int64_t k = base + lane_id();
      int32_t x = k < e ? g.bcol[k] : HOLE;
      live |= __ballot(x == v);
      tomb |= __ballot(x == ~v);
    }
  }
  unsigned long long alive = 0, atomb = 0;
  int64_t ab = hu.off, ae = ab + hu.len;
  for (int64_t base = ab; base < ae; base += kWave) {
    int64_t k = base + lane_id();
    int32_t x = k < ae ? g.acol[k] : HOLE;
    alive |= __ballot(x == v);
    atomb |= __ballot(x == ~v);
  }
  if (alive) return true;
  if (tomb || atomb) return false;
  return live != 0;
}

}  // namespace scout
