// gen.hip: inputs built in HBM -- R-MAT graphs bit-identical to the host generator
// (csrc/host/gen.h), and greedy-ascending initial MIS / colouring (the closed forms of
// tools/compute_initial_mis.py:32-43 and tools/greedy_color.cpp:37-52) computed in rounds
// with vertex-id priority: a vertex decides once every lower-id neighbour has, which is
// exactly the ascending greedy order.
#include <algorithm>
#include <vector>

#include "../host/rng.h"
#include "common.cuh"

namespace scout {

int sort_keys_u64(Ctx& c, const uint64_t* kin, uint64_t* kout, int64_t count, int begin_bit, int end_bit);
int check_rows(Ctx& c);

namespace {

constexpr int kBins = 256;

// Pass 1: per-bin counts of directed keys (bin = top 8 bits of the source id).
__global__ void k_rmat_count(RmatParams p, uint64_t key, int64_t m, int shift, unsigned long long* bins) {
  __shared__ unsigned long long local[kBins];
  for (int i = threadIdx.x; i < kBins; i += blockDim.x) local[i] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t e = rmat_edge(p, key, (uint64_t)i);
    uint32_t u = (uint32_t)(e >> 32), v = (uint32_t)e;
    if (u == v) continue;
    atomicAdd(&local[u >> shift], 1ULL);
    atomicAdd(&local[v >> shift], 1ULL);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kBins; i += blockDim.x)
    if (local[i]) atomicAdd(&bins[i], local[i]);
}

// Pass 2: scatter keys into their bucket's region. Each block counts its keys per bucket in
// LDS and reserves every bucket's range with one global atomic, so the few bucket cursors are
// not contended by every thread.
__global__ void k_rmat_scatter(RmatParams p, uint64_t key, int64_t m, int shift, const int32_t* bin_bucket, int nb,
                               unsigned long long* cursor, uint64_t* out) {
  __shared__ unsigned int lcount[kBins];
  __shared__ unsigned long long lbase[kBins];
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < m; base += (int64_t)gridDim.x * blockDim.x) {
    for (int b = threadIdx.x; b < nb; b += blockDim.x) lcount[b] = 0;
    __syncthreads();
    int64_t i = base + threadIdx.x;
    bool valid = false;
    uint32_t u = 0, v = 0;
    int bu = 0, bv = 0;
    unsigned int ou = 0, ov = 0;
    if (i < m) {
      uint64_t e = rmat_edge(p, key, (uint64_t)i);
      u = (uint32_t)(e >> 32);
      v = (uint32_t)e;
      valid = u != v;
      if (valid) {
        bu = bin_bucket[u >> shift];
        bv = bin_bucket[v >> shift];
        ou = atomicAdd(&lcount[bu], 1u);
        ov = atomicAdd(&lcount[bv], 1u);
      }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
      lbase[b] = lcount[b] ? atomicAdd(&cursor[b], (unsigned long long)lcount[b]) : 0ULL;
    __syncthreads();
    if (valid) {
      out[lbase[bu] + ou] = ((uint64_t)u << 32) | v;
      out[lbase[bv] + ov] = ((uint64_t)v << 32) | u;
    }
    __syncthreads();
  }
}

// flag[i] = 1 if key i differs from key i-1 (first of a run).
__global__ void k_unique_degree(const uint64_t* keys, int64_t count, int64_t* deg) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) atomicAdd((unsigned long long*)&deg[k >> 32], 1ULL);
  }
}

__global__ void k_first_flags(const uint64_t* keys, int64_t count, int64_t* flags) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    flags[i] = (i == 0 || keys[i - 1] != keys[i]) ? 1 : 0;
}

// pos[i] (exclusive prefix of first-of-run flags) -> position of key i in the final nlist.
__global__ void k_place_by_prefix(const uint64_t* keys, int64_t count, const int64_t* pos, int32_t* col) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t k = keys[i];
    if (i > 0 && keys[i - 1] == k) continue;
    col[pos[i]] = (int32_t)(k & 0xffffffffu);
  }
}

// Greedy-ascending MIS and colouring in topological order of the id-ordered DAG (Kahn): a
// vertex becomes ready once all its lower-id neighbours are decided (a count of undecided
// lower neighbours reaches zero); ready vertices decide together, then release their higher-id
// neighbours. Rounds = the longest increasing path; work = O(entries) in total.

// cnt[v] = live entries w < v of v's row; v with none is ready at once.
__global__ void k_lower_count(GView g, int32_t* cnt, int32_t* ready, int64_t* ready_count, int64_t cap) {
  for (int64_t base = (int64_t)blockIdx.x * kWavesPerBlock; base < g.n; base += (int64_t)gridDim.x * kWavesPerBlock) {
    int64_t v = base + threadIdx.x / kWave;
    int64_t c = 0;
    if (v < g.n) {
      Row r = row_of(g, (int32_t)v);
      int64_t len = row_len(r);
      for (int64_t k = lane_id(); k < len; k += kWave) {
        int32_t w = row_at(g, r, k);
        c += (w >= 0 && w < v);
      }
    }
    c = wave_sum(c);
    bool take = v < g.n && c == 0 && lane_id() == 0;
    if (v < g.n && lane_id() == 0) cnt[v] = (int32_t)c;
    block_append(take, (int32_t)v, ready, ready_count, cap);
  }
}

// Wave per ready vertex: MIS member iff no lower-id neighbour is a member (state 1 member,
// 2 not); colouring: first fit over the lower-id neighbours' colours.
__global__ void k_greedy_decide(GView g, const int32_t* ready, const int64_t* ready_count, uint8_t* state,
                                int32_t* color) {
  int64_t nr = *ready_count;
  for (int64_t i = global_wave(); i < nr; i += total_waves()) {
    int32_t v = ready[i];
    Row r = row_of(g, v);
    int64_t len = row_len(r);
    if (state) {
      bool in_lower = false;
      for (int64_t k = lane_id(); k < len && !in_lower; k += kWave) {
        int32_t w = row_at(g, r, k);
        in_lower = w >= 0 && w < v && state[w] == 1;
        in_lower = __ballot(in_lower) != 0;
      }
      if (lane_id() == 0) state[v] = in_lower ? 2 : 1;
    } else {
      int32_t cbase = 0;
      while (true) {
        unsigned long long mask = 0;
        for (int64_t k = lane_id(); k < len; k += kWave) {
          int32_t w = row_at(g, r, k);
          if (w >= 0 && w < v) {
            int32_t cw = color[w] - cbase;
            if (cw >= 0 && cw < 64) mask |= 1ULL << cw;
          }
        }
        for (int off = kWave / 2; off > 0; off >>= 1) mask |= __shfl_xor(mask, off);
        if (~mask != 0ULL) {
          if (lane_id() == 0) color[v] = cbase + (__ffsll((long long)~mask) - 1);
          break;
        }
        cbase += 64;
      }
    }
  }
}

// Wave per decided vertex: every higher-id neighbour loses one undecided lower neighbour; the
// ones reaching zero are ready next round.
__global__ void k_greedy_release(GView g, const int32_t* ready, const int64_t* ready_count, int32_t* cnt,
                                 int32_t* next, int64_t* next_count, int64_t cap) {
  int64_t nr = *ready_count;
  for (int64_t i = global_wave(); i < nr; i += total_waves()) {
    int32_t v = ready[i];
    Row r = row_of(g, v);
    int64_t len = row_len(r);
    for (int64_t base = 0; base < len; base += kWave) {
      int64_t k = base + lane_id();
      bool take = false;
      int32_t w = -1;
      if (k < len) {
        w = row_at(g, r, k);
        take = w > v && atomicSub(&cnt[w], 1) == 1;
      }
      wave_append(take, w, next, next_count, cap);
    }
  }
}

__global__ void k_mis_from_state(const uint8_t* state, int32_t n, uint8_t* mis) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) mis[v] = state[v] == 1 ? 1 : 0;
}

__global__ void k_fill_i32(int32_t* p, int64_t count, int32_t value) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = value;
}

}  // namespace

int gen_rmat(Ctx& c, int scale, int ef, double a, double b, double cc, uint64_t seed) {
  if (scale < 1 || scale > 30 || ef < 1) {
    set_error("gen_rmat: scale must be in [1, 30] and edge_factor >= 1");
    return SCOUT_ERR_ARG;
  }
  RmatParams p = rmat_params(a, b, cc, scale);
  uint64_t key = stream_key(seed, RNG_RMAT);
  int64_t m = (int64_t)ef << scale;
  int32_t n = (int32_t)(1LL << scale);
  int shift = scale > 8 ? scale - 8 : 0;
  unsigned long long* bins = nullptr;
  SCOUT_TRY(hipMalloc((void**)&bins, 2 * kBins * sizeof(unsigned long long)));
  SCOUT_TRY(hipMemsetAsync(bins, 0, 2 * kBins * sizeof(unsigned long long), c.s));
  k_rmat_count<<<4096, kBlock, 0, c.s>>>(p, key, m, shift, bins);
  std::vector<unsigned long long> hb(kBins);
  SCOUT_TRY(hipMemcpyAsync(hb.data(), bins, kBins * sizeof(unsigned long long), hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  // Buckets of consecutive bins, each small enough for one radix sort call.
  const int64_t kMaxBucket = (int64_t)1 << 30;
  std::vector<int32_t> bin_bucket(kBins);
  std::vector<int64_t> bstart{0};
  int64_t acc = 0, total = 0;
  for (int i = 0; i < kBins; i++) {
    if (acc > 0 && acc + (int64_t)hb[(size_t)i] > kMaxBucket) {
      bstart.push_back(total);
      acc = 0;
    }
    bin_bucket[(size_t)i] = (int32_t)bstart.size() - 1;
    acc += (int64_t)hb[(size_t)i];
    total += (int64_t)hb[(size_t)i];
  }
  bstart.push_back(total);
  int nb = (int)bstart.size() - 1;
  int64_t maxb = 0;
  for (int i = 0; i < nb; i++) maxb = std::max(maxb, bstart[(size_t)i + 1] - bstart[(size_t)i]);
  uint64_t *keys = nullptr, *alt = nullptr;
  int32_t* d_bin_bucket = nullptr;
  SCOUT_TRY(hipMalloc((void**)&keys, (size_t)std::max<int64_t>(total, 1) * 8));
  SCOUT_TRY(hipMalloc((void**)&alt, (size_t)std::max<int64_t>(maxb, 1) * 8));
  SCOUT_TRY(hipMalloc((void**)&d_bin_bucket, kBins * sizeof(int32_t)));
  SCOUT_TRY(hipMemcpyAsync(d_bin_bucket, bin_bucket.data(), kBins * sizeof(int32_t), hipMemcpyHostToDevice, c.s));
  std::vector<unsigned long long> cur(bstart.begin(), bstart.end() - 1);
  unsigned long long* d_cur = bins + kBins;
  SCOUT_TRY(hipMemcpyAsync(d_cur, cur.data(), (size_t)nb * 8, hipMemcpyHostToDevice, c.s));
  k_rmat_scatter<<<4096, kBlock, 0, c.s>>>(p, key, m, shift, d_bin_bucket, nb, d_cur, keys);
  // Sort each bucket (source ids of different buckets never interleave).
  for (int i = 0; i < nb; i++) {
    int64_t lo = bstart[(size_t)i], cnt = bstart[(size_t)i + 1] - lo;
    if (cnt <= 0) continue;
    SCOUT_CHECK(sort_keys_u64(c, keys + lo, alt, cnt, 0, 32 + scale));
    SCOUT_TRY(hipMemcpyAsync(keys + lo, alt, (size_t)cnt * 8, hipMemcpyDeviceToDevice, c.s));
  }
  (void)hipFree(alt);
  // Unique keys -> degrees -> offsets -> columns.
  graph_free(c);
  DevGraph& g = c.g;
  g.n = n;
  int64_t* deg = nullptr;
  SCOUT_TRY(hipMalloc((void**)&deg, ((size_t)n + 1) * 8));
  SCOUT_TRY(hipMemsetAsync(deg, 0, ((size_t)n + 1) * 8, c.s));
  k_unique_degree<<<4096, kBlock, 0, c.s>>>(keys, total, deg);
  SCOUT_TRY(hipMalloc((void**)&g.boff, ((size_t)n + 1) * 8));
  SCOUT_CHECK(scan_exclusive_i64(c, deg, g.boff, (int64_t)n + 1));
  int64_t E = 0;
  SCOUT_TRY(hipMemcpyAsync(&E, g.boff + n, 8, hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  (void)hipFree(deg);
  SCOUT_TRY(hipMalloc((void**)&g.bcol, (size_t)std::max<int64_t>(E, 1) * 4));
  // Positions: exclusive prefix of first-of-run flags (reuses a 8-byte-per-key buffer).
  int64_t *flags = nullptr, *pos = nullptr;
  SCOUT_TRY(hipMalloc((void**)&flags, (size_t)std::max<int64_t>(total, 1) * 8));
  SCOUT_TRY(hipMalloc((void**)&pos, (size_t)std::max<int64_t>(total, 1) * 8));
  k_first_flags<<<4096, kBlock, 0, c.s>>>(keys, total, flags);
  SCOUT_CHECK(scan_exclusive_i64(c, flags, pos, total));
  k_place_by_prefix<<<4096, kBlock, 0, c.s>>>(keys, total, pos, g.bcol);
  SCOUT_TRY(hipStreamSynchronize(c.s));
  (void)hipFree(flags);
  (void)hipFree(pos);
  (void)hipFree(keys);
  (void)hipFree(bins);
  (void)hipFree(d_bin_bucket);
  g.E = E;
  g.base_gen++;
  g.shape_gen++;
  SCOUT_CHECK(reset_inserted(c));
  return check_rows(c);
}

namespace {

// Kahn rounds over the id-ordered DAG; state (MIS) or color (colouring) receives the result.
int greedy_rounds(Ctx& c, uint8_t* state, int32_t* color) {
  int32_t n = c.g.n;
  int64_t cap = (int64_t)n + 1;
  DevBuf<int32_t> cnt, la, lb;
  SCOUT_CHECK(cnt.reserve(cap));
  SCOUT_CHECK(la.reserve(cap));
  SCOUT_CHECK(lb.reserve(cap));
  int64_t* box = c.d_box;
  SCOUT_TRY(hipMemsetAsync(box, 0, 2 * sizeof(int64_t), c.s));
  k_lower_count<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), cnt.p, la.p, box, cap);
  int32_t* cur = la.p;
  int32_t* nxt = lb.p;
  int p = 0;
  while (true) {
    SCOUT_CHECK(box_read(c, 2));
    if (c.h_box[p] == 0) break;
    SCOUT_TRY(hipMemsetAsync(box + (p ^ 1), 0, sizeof(int64_t), c.s));
    k_greedy_decide<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), cur, box + p, state, color);
    k_greedy_release<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), cur, box + p, cnt.p, nxt, box + (p ^ 1), cap);
    std::swap(cur, nxt);
    p ^= 1;
  }
  SCOUT_TRY(hipStreamSynchronize(c.s));
  cnt.release();
  la.release();
  lb.release();
  return SCOUT_OK;
}

}  // namespace

int greedy_mis(Ctx& c) {
  int32_t n = c.g.n;
  SCOUT_CHECK(ensure_state(c, SCOUT_STATE_MIS));
  DevBuf<uint8_t> state;
  SCOUT_CHECK(state.reserve((int64_t)n + 1));
  SCOUT_TRY(hipMemsetAsync(state.p, 0, (size_t)n + 1, c.s));
  SCOUT_CHECK(greedy_rounds(c, state.p, nullptr));
  c.mis_version++;
  k_mis_from_state<<<grid_for(n), kBlock, 0, c.s>>>(state.p, n, c.mis);
  SCOUT_TRY(hipStreamSynchronize(c.s));
  state.release();
  return SCOUT_OK;
}

int greedy_color(Ctx& c) {
  int32_t n = c.g.n;
  SCOUT_CHECK(ensure_state(c, SCOUT_STATE_COLOR));
  k_fill_i32<<<grid_for(n), kBlock, 0, c.s>>>(c.color, n, -1);
  return greedy_rounds(c, nullptr, c.color);
}

}  // namespace scout

using namespace scout;

extern "C" {

int scout_gen_rmat(scout_ctx* x, int scale, int edge_factor, double a, double b, double c, uint64_t seed) {
  if (!x) return SCOUT_ERR_ARG;
  Ctx& ctx = x->c;
  SCOUT_TRY(hipSetDevice(ctx.dev));
  SCOUT_CHECK(gen_rmat(ctx, scale, edge_factor, a, b, c, seed));
  // fresh per-vertex scratch for the new vertex count
  int32_t n = ctx.g.n;
  void* ps[] = {ctx.level, ctx.parent, ctx.label, ctx.color, ctx.mis, ctx.stamp, ctx.aux, ctx.aux2,
                ctx.cc_owner, ctx.col_marker, ctx.mis_minw, ctx.cc_level, ctx.cc_parent};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  ctx.cc_level = ctx.cc_parent = ctx.parent = nullptr;
  ctx.level_seq = ctx.cc_level_seq = ctx.cc_components = ctx.label_seq = -1;
  ctx.level = ctx.label = ctx.color = nullptr;
  ctx.mis = nullptr;
  ctx.cc_owner = ctx.col_marker = ctx.mis_minw = nullptr;
  cc_release(ctx);
  mis_release(ctx);
  SCOUT_TRY(hipMalloc((void**)&ctx.stamp, ((size_t)n + 1) * 4));
  SCOUT_TRY(hipMalloc((void**)&ctx.aux, ((size_t)n + 1) * 4));
  SCOUT_TRY(hipMalloc((void**)&ctx.aux2, ((size_t)n + 1) * 4));
  SCOUT_TRY(hipMemsetAsync(ctx.stamp, 0, ((size_t)n + 1) * 4, ctx.s));
  SCOUT_TRY(hipMemsetAsync(ctx.aux, 0xff, ((size_t)n + 1) * 4, ctx.s));
  SCOUT_TRY(hipMemsetAsync(ctx.aux2, 0xff, ((size_t)n + 1) * 4, ctx.s));
  ctx.epoch = 0;
  ctx.hbm_used = ((int64_t)n + 1) * 28 + ctx.g.E * 4;
  SCOUT_TRY(hipStreamSynchronize(ctx.s));
  return SCOUT_OK;
}

int scout_mis_greedy(scout_ctx* x) {
  if (!x) return SCOUT_ERR_ARG;
  SCOUT_TRY(hipSetDevice(x->c.dev));
  if (!x->c.g.boff) {
    set_error("no graph loaded");
    return SCOUT_ERR_STATE;
  }
  return greedy_mis(x->c);
}

int scout_color_greedy(scout_ctx* x) {
  if (!x) return SCOUT_ERR_ARG;
  SCOUT_TRY(hipSetDevice(x->c.dev));
  if (!x->c.g.boff) {
    set_error("no graph loaded");
    return SCOUT_ERR_STATE;
  }
  return greedy_color(x->c);
}

}  // extern "C"
