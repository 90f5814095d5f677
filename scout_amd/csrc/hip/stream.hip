// stream.hip: synthetic update batches generated in HBM against the current graph and
// levels, and device-side snapshots for replaying one stream under several strategies.
//
// A batch holds `size` toggles: round(delete_frac * size) deletions of current edges, of which
// round(tree_frac * deletions) are BFS-tree edges (a reached vertex and a neighbour one level
// closer to the source) and the rest uniform over current adjacency entries, plus uniform
// insertions of current non-edges. As in the reference's generators
// (tools/generate_mixed_batches.py:94-141) every line is a canonical u < v pair and no pair is
// ever repeated within or across the batches of one context.
#include <algorithm>
#include <vector>

#include "../host/rng.h"
#include "common.cuh"

namespace scout {

namespace {

constexpr uint64_t kEmptyKey = ~0ULL;
// Longest probe sequence of a claim. The set is kept at most half full, where linear probing
// runs far shorter; the bound only caps the cost of a draw.
constexpr uint64_t kMaxProbe = 4096;

__device__ __forceinline__ uint64_t mix(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

// A draw proposes a canonical pair (smaller id first) in its own slot; kEmptyKey = no pair.
// Generation is deterministic: the draws of one attempt propose in parallel, a pair goes to
// the earliest draw that proposed it (and only if no earlier batch or attempt toggled it),
// and accepted pairs keep draw order.
__device__ __forceinline__ void propose(uint64_t* cand, int64_t i, uint32_t u, uint32_t v) {
  uint32_t a = u < v ? u : v, b = u < v ? v : u;
  cand[i] = ((uint64_t)a << 32) | b;
}

// Toggled-pair set: open addressing over canonical keys; owner[h] = the smallest draw tag
// that proposed the key (tags grow across attempts and batches).
__global__ void k_gen_claim(const uint64_t* cand, int64_t draws, uint64_t* set, uint64_t* owner, uint64_t mask,
                            uint64_t tag0, uint64_t* cslot) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= draws) return;
  uint64_t key = cand[i];
  if (key == kEmptyKey) return;
  uint64_t h = mix(key) & mask;
  for (uint64_t probe = 0; probe <= mask && probe < kMaxProbe; probe++) {
    uint64_t prev = atomicCAS((unsigned long long*)&set[h], (unsigned long long)kEmptyKey, (unsigned long long)key);
    if (prev == kEmptyKey || prev == key) {
      atomicMin((unsigned long long*)&owner[h], (unsigned long long)(tag0 + (uint64_t)i));
      cslot[i] = h;
      return;
    }
    h = (h + 1) & mask;
  }
  cslot[i] = kEmptyKey;  // no free slot within kMaxProbe: the draw is dropped
}

// Moves every pair of the old toggle set into a larger one. Their owner tag becomes 0, below
// every later draw's tag, so they are never accepted again.
__global__ void k_gen_rehash(const uint64_t* old, uint64_t old_slots, uint64_t* set, uint64_t* owner, uint64_t mask) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= old_slots) return;
  uint64_t key = old[i];
  if (key == kEmptyKey) return;
  uint64_t h = mix(key) & mask;
  for (uint64_t probe = 0; probe <= mask; probe++) {
    if (atomicCAS((unsigned long long*)&set[h], (unsigned long long)kEmptyKey, (unsigned long long)key) ==
        kEmptyKey) {
      owner[h] = 0;
      return;
    }
    h = (h + 1) & mask;
  }
}

__global__ void k_gen_accept(const uint64_t* cand, const uint64_t* cslot, const uint64_t* owner, uint64_t tag0,
                             int64_t draws, int32_t* acc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= draws) return;
  acc[i] = cand[i] != kEmptyKey && cslot[i] != kEmptyKey && owner[cslot[i]] == tag0 + (uint64_t)i;
}

// Accepted pairs in draw order, at most `room` of them; count = accepted pairs of the attempt.
__global__ void k_gen_emit(const uint64_t* cand, const int32_t* acc, const int64_t* pos, int64_t draws, int32_t* out,
                           int64_t room, int64_t* count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= draws) return;
  if (i == draws - 1) *count = pos[i] + acc[i];
  if (!acc[i] || pos[i] >= room) return;
  out[2 * pos[i]] = (int32_t)(cand[i] >> 32);
  out[2 * pos[i] + 1] = (int32_t)(cand[i] & 0xffffffffULL);
}

// The inserted entries in logical order (rows by id, entries in row order), independent of
// where rows sit in the pool: rowpre[u] = inserted-row slots of the rows before u
// (rowpre[n] = all of them), so entry j lies in the last row u with rowpre[u] <= j.
__global__ void k_rowpre_total(const AHead* ahead, int32_t n, int64_t* rowpre) {
  rowpre[n] = n > 0 ? rowpre[n - 1] + ahead[n - 1].len : 0;
}

// Inserted entry j of the logical order: its row and value; false if not live.
__device__ __forceinline__ bool pool_entry(const GView& g, const int64_t* rowpre, int64_t j, uint32_t* owner,
                                           int32_t* w) {
  int64_t lo = 0, hi = g.n;  // last u with rowpre[u] <= j
  while (lo < hi) {
    int64_t mid = (lo + hi + 1) >> 1;
    if (rowpre[mid] <= j) lo = mid;
    else hi = mid - 1;
  }
  *owner = (uint32_t)lo;
  *w = g.acol[g.ahead[lo].off + (j - rowpre[lo])];
  return *w >= 0;
}

// Tree deletions: a uniform vertex with level > 0 and its first neighbour one level lower,
// scanning its row cyclically from a random offset (thread per draw).
__global__ void k_gen_tree(GView g, const int32_t* level, uint64_t key, uint64_t base, int64_t draws, uint64_t* cand) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= draws) return;
  cand[i] = kEmptyKey;
  uint64_t x = draw(key, base + (uint64_t)i);
  uint32_t v = uniform_below(x, (uint32_t)g.n);
  int32_t lv = level[v];
  if (lv <= 0) return;
  Row r = row_of(g, (int32_t)v);
  int64_t len = row_len(r);
  if (len == 0) return;
  int64_t start = (int64_t)uniform_below(x << 32 | (x >> 32), (uint32_t)min<int64_t>(len, 0xffffffffLL));
  for (int64_t t = 0; t < len && t < 256; t++) {
    int64_t k = start + t;
    if (k >= len) k -= len;
    int32_t w = row_at(g, r, k);
    if (w >= 0 && level[w] == lv - 1) {
      propose(cand, i, v, (uint32_t)w);
      return;
    }
  }
}

// Uniform deletions: a uniform position over the base entries and the inserted entries (in
// logical pool order); the draw is kept if it holds a live entry (rejection keeps the choice
// uniform over the live adjacency entries).
__global__ void k_gen_del(GView g, int64_t E, int64_t V, const int64_t* rowpre, uint64_t key, uint64_t base,
                          int64_t draws, uint64_t* cand) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= draws) return;
  cand[i] = kEmptyKey;
  uint64_t x = draw(key, base + (uint64_t)i);
  uint64_t total = (uint64_t)(E + V);
  uint64_t pos = __umul64hi(x, total);
  int32_t w;
  uint32_t u;
  if (pos < (uint64_t)E) {
    int64_t p = (int64_t)pos;
    w = g.bcol[p];
    if (w < 0) return;
    int64_t l = 0, h = g.n;  // last vertex whose offset <= p
    while (l < h) {
      int64_t mid = (l + h + 1) >> 1;
      if (g.boff[mid] <= p) l = mid;
      else h = mid - 1;
    }
    u = (uint32_t)l;
  } else if (!pool_entry(g, rowpre, (int64_t)pos - E, &u, &w)) {
    return;
  }
  if (u == (uint32_t)w) return;
  propose(cand, i, u, (uint32_t)w);
}

// Uniform insertions of current non-edges.
__global__ void k_gen_ins(GView g, bool sorted, uint64_t key, uint64_t base, int64_t draws, uint64_t* cand) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= draws) return;
  cand[i] = kEmptyKey;
  uint64_t x = draw(key, base + (uint64_t)i);
  uint32_t u = uniform_below(x, (uint32_t)g.n), v = uniform_below(x << 32 | (x >> 32), (uint32_t)g.n);
  if (u == v) return;
  if (adjacent_now(pair_state(g, sorted, (int32_t)u, (int32_t)v))) return;
  propose(cand, i, u, v);
}

// A vertex drawn with probability proportional to its live degree + 1: a uniform position
// over the base entries, the inserted entries (logical pool order) and one slot per vertex;
// an entry position gives its row's vertex if the entry is live, else the draw is rejected.
__device__ __forceinline__ bool pick_degree_plus_one(const GView& g, int64_t E, int64_t V, const int64_t* rowpre,
                                                     uint64_t x, uint32_t* out) {
  uint64_t total = (uint64_t)(E + V + g.n);
  uint64_t pos = __umul64hi(x, total);
  if (pos < (uint64_t)E) {
    int64_t p = (int64_t)pos;
    if (g.bcol[p] < 0) return false;
    int64_t l = 0, h = g.n;  // last vertex whose offset <= p
    while (l < h) {
      int64_t mid = (l + h + 1) >> 1;
      if (g.boff[mid] <= p) l = mid;
      else h = mid - 1;
    }
    *out = (uint32_t)l;
    return true;
  }
  if (pos < (uint64_t)(E + V)) {
    int32_t w;
    return pool_entry(g, rowpre, (int64_t)pos - E, out, &w);
  }
  *out = (uint32_t)(pos - (uint64_t)(E + V));
  return true;
}

// Degree-biased deletions: u with probability proportional to degree + 1, then a uniform live
// entry of u's row (rejection over its base and inserted entries).
__global__ void k_gen_del_deg(GView g, int64_t E, int64_t V, const int64_t* rowpre, uint64_t key, uint64_t base,
                              int64_t draws, uint64_t* cand) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= draws) return;
  cand[i] = kEmptyKey;
  uint64_t x = draw(key, base + 2 * (uint64_t)i);
  uint32_t u;
  if (!pick_degree_plus_one(g, E, V, rowpre, x, &u)) return;
  Row r = row_of(g, (int32_t)u);
  int64_t len = row_len(r);
  if (len == 0 || len > 0xffffffffLL) return;
  int32_t w = row_at(g, r, (int64_t)uniform_below(draw(key, base + 2 * (uint64_t)i + 1), (uint32_t)len));
  if (w < 0 || (uint32_t)w == u) return;
  propose(cand, i, u, (uint32_t)w);
}

// Degree-biased insertions: both endpoints with probability proportional to degree + 1, kept
// if they are distinct and not adjacent.
__global__ void k_gen_ins_deg(GView g, bool sorted, int64_t E, int64_t V, const int64_t* rowpre, uint64_t key,
                              uint64_t base, int64_t draws, uint64_t* cand) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= draws) return;
  cand[i] = kEmptyKey;
  uint32_t u, v;
  if (!pick_degree_plus_one(g, E, V, rowpre, draw(key, base + 2 * (uint64_t)i), &u)) return;
  if (!pick_degree_plus_one(g, E, V, rowpre, draw(key, base + 2 * (uint64_t)i + 1), &v)) return;
  if (u == v) return;
  if (adjacent_now(pair_state(g, sorted, (int32_t)u, (int32_t)v))) return;
  propose(cand, i, u, v);
}

// Shuffle of a generated batch: key = (hash of the line index under the batch key) << 32 | index.
__global__ void k_shuffle_keys(uint64_t key, int64_t m, uint64_t* keys) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) keys[i] = (draw(key, (uint64_t)i) & 0xffffffff00000000ULL) | (uint64_t)i;
}

__global__ void k_gather_lines(const uint64_t* keys, int64_t m, const int32_t* in, int32_t* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int64_t k = (int64_t)(keys[i] & 0xffffffffULL);
  out[2 * i] = in[2 * k];
  out[2 * i + 1] = in[2 * k + 1];
}

}  // namespace

struct StreamState {
  int mode = SCOUT_GEN_UNIFORM;
  uint64_t batches = 0;  // batches generated so far
  DevBuf<uint64_t> set, owner;  // toggled pairs and the draw tag that owns each
  uint64_t mask = 0;
  int64_t used = 0;  // pairs in the set
  uint64_t draws_used[3] = {0, 0, 0};
  uint64_t next_tag = 0;           // draw tags, increasing over attempts and batches
  DevBuf<uint64_t> cand, cslot;    // per draw: proposed pair, its set slot
  DevBuf<int32_t> acc;             // per draw: accepted
  DevBuf<int64_t> pos;             // per draw: position among the accepted
  DevBuf<int64_t> rowpre;          // inserted-row slots before each row (n + 1)
  std::vector<DevBuf<int32_t>> slots;  // stored batches (pairs) for replay
  std::vector<int64_t> slot_m;
};

// Replaces the toggle set by one of at least `slots` slots holding the same pairs.
static int grow_set(Ctx& c, StreamState& s, uint64_t slots) {
  DevBuf<uint64_t> set2, own2;
  SCOUT_CHECK(set2.reserve((int64_t)slots));
  SCOUT_CHECK(own2.reserve((int64_t)slots));
  SCOUT_TRY(hipMemsetAsync(set2.p, 0xff, (size_t)set2.cap * 8, c.s));
  SCOUT_TRY(hipMemsetAsync(own2.p, 0xff, (size_t)own2.cap * 8, c.s));
  uint64_t mask = 1;
  while (mask * 2 <= (uint64_t)set2.cap) mask *= 2;
  mask -= 1;
  if (s.set.p) {
    k_gen_rehash<<<grid_for((int64_t)(s.mask + 1)), kBlock, 0, c.s>>>(s.set.p, s.mask + 1, set2.p, own2.p, mask);
    SCOUT_TRY(hipStreamSynchronize(c.s));
  }
  s.set.release();
  s.owner.release();
  s.set = set2;
  s.owner = own2;
  s.mask = mask;
  return SCOUT_OK;
}

int gen_batch(Ctx& c, Channel& ch, int64_t size, double delete_frac, double tree_frac, uint64_t seed,
              int64_t expected_total, int64_t* out_m) {
  if (!c.stream) c.stream = new StreamState();
  StreamState& s = *c.stream;
  // a batch fills at most about 3 * size slots (draws accepted past its room stay claimed);
  // the set stays at most half full
  uint64_t need = (uint64_t)(s.used + 4 * size);
  if (!s.set.p || 2 * need > s.mask + 1) {
    uint64_t want = std::max<uint64_t>((uint64_t)(4 * std::max<int64_t>(expected_total, size)), 4 * need);
    uint64_t slots = 1024;
    while (slots < want) slots <<= 1;
    SCOUT_CHECK(grow_set(c, s, slots));
  }
  int64_t ndel = (int64_t)llround(delete_frac * (double)size);
  int64_t ntree = (int64_t)llround(tree_frac * (double)ndel);
  int64_t nuni = ndel - ntree, nins = size - ndel;
  SCOUT_CHECK(ch.d_uv.reserve(2 * size + 2));
  SCOUT_CHECK(ch.d_ins.reserve(size + 1));
  SCOUT_CHECK(ch.d_done.reserve(size + 1));
  SCOUT_CHECK(channel_wait(c.s, ch));
  uint64_t keys[3] = {stream_key(seed, RNG_STREAM_TREE), stream_key(seed, RNG_STREAM_DEL), stream_key(seed, RNG_STREAM_INS)};
  int64_t want[3] = {ntree, nuni, nins};
  int64_t produced = 0;
  // the inserted entries in logical order, for the draws over adjacency entries
  int64_t V = 0;
  SCOUT_CHECK(s.rowpre.reserve((int64_t)c.g.n + 1));
  if (c.g.A > 0) {
    SCOUT_CHECK(scan_row_lens(c, c.g.ahead.p, s.rowpre.p, c.g.n));
    k_rowpre_total<<<1, 1, 0, c.s>>>(c.g.ahead.p, c.g.n, s.rowpre.p);
    SCOUT_TRY(hipMemcpyAsync(c.h_box, s.rowpre.p + c.g.n, sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
    SCOUT_TRY(hipStreamSynchronize(c.s));
    c.readbacks++;
    V = c.h_box[0];
  }
  for (int kind = 0; kind < 3; kind++) {
    if (want[kind] == 0) continue;
    if (kind == 0 && !c.level) {
      set_error("gen_batch: tree deletions need BFS levels");
      return SCOUT_ERR_STATE;
    }
    int32_t* out = ch.d_uv.p + 2 * produced;
    int64_t got = 0;
    for (int attempt = 0; attempt < 64 && got < want[kind]; attempt++) {
      int64_t draws = std::max<int64_t>(2 * (want[kind] - got), 1024);
      SCOUT_CHECK(s.cand.reserve(draws));
      SCOUT_CHECK(s.cslot.reserve(draws));
      SCOUT_CHECK(s.acc.reserve(draws));
      SCOUT_CHECK(s.pos.reserve(draws));
      unsigned grid = grid_for(draws);
      if (kind == 0)
        k_gen_tree<<<grid, kBlock, 0, c.s>>>(c.g.view(), c.level, keys[0], s.draws_used[0], draws, s.cand.p);
      else if (kind == 1 && s.mode == SCOUT_GEN_DEGREE)
        k_gen_del_deg<<<grid, kBlock, 0, c.s>>>(c.g.view(), c.g.E, V, s.rowpre.p, keys[1], s.draws_used[1], draws,
                                                s.cand.p);
      else if (kind == 1)
        k_gen_del<<<grid, kBlock, 0, c.s>>>(c.g.view(), c.g.E, V, s.rowpre.p, keys[1], s.draws_used[1], draws,
                                            s.cand.p);
      else if (s.mode == SCOUT_GEN_DEGREE)
        k_gen_ins_deg<<<grid, kBlock, 0, c.s>>>(c.g.view(), c.g.sorted, c.g.E, V, s.rowpre.p, keys[2],
                                                s.draws_used[2], draws, s.cand.p);
      else
        k_gen_ins<<<grid, kBlock, 0, c.s>>>(c.g.view(), c.g.sorted, keys[2], s.draws_used[2], draws, s.cand.p);
      // the degree-biased kernels use two draws per attempt
      s.draws_used[kind] += (uint64_t)draws * ((kind > 0 && s.mode == SCOUT_GEN_DEGREE) ? 2 : 1);
      uint64_t tag0 = s.next_tag;
      s.next_tag += (uint64_t)draws;
      k_gen_claim<<<grid, kBlock, 0, c.s>>>(s.cand.p, draws, s.set.p, s.owner.p, s.mask, tag0, s.cslot.p);
      k_gen_accept<<<grid, kBlock, 0, c.s>>>(s.cand.p, s.cslot.p, s.owner.p, tag0, draws, s.acc.p);
      SCOUT_CHECK(scan_exclusive_i32_to_i64(c, s.acc.p, s.pos.p, draws));
      k_gen_emit<<<grid, kBlock, 0, c.s>>>(s.cand.p, s.acc.p, s.pos.p, draws, out + 2 * got, want[kind] - got,
                                           c.d_box);
      SCOUT_CHECK(box_read(c, 1));
      s.used += c.h_box[0];
      got += std::min(c.h_box[0], want[kind] - got);
    }
    // A pair accepted past the batch's room stays claimed: such pairs are skipped by every
    // later batch, which keeps the no-repeat rule.
    produced += got;
  }
  if (s.mode == SCOUT_GEN_DEGREE && produced > 1) {
    // interleave deletions and insertions (deterministic shuffle)
    DevBuf<uint64_t> k1, k2;
    DevBuf<int32_t> tmp;
    SCOUT_CHECK(k1.reserve(produced));
    SCOUT_CHECK(k2.reserve(produced));
    SCOUT_CHECK(tmp.reserve(2 * produced));
    k_shuffle_keys<<<grid_for(produced), kBlock, 0, c.s>>>(stream_key(seed, RNG_STREAM_SIZE) + s.batches, produced,
                                                           k1.p);
    SCOUT_CHECK(sort_keys_u64(c, k1.p, k2.p, produced, 0, 64));
    SCOUT_TRY(hipMemcpyAsync(tmp.p, ch.d_uv.p, (size_t)(2 * produced) * 4, hipMemcpyDeviceToDevice, c.s));
    k_gather_lines<<<grid_for(produced), kBlock, 0, c.s>>>(k2.p, produced, tmp.p, ch.d_uv.p);
    SCOUT_TRY(hipStreamSynchronize(c.s));
    k1.release();
    k2.release();
    tmp.release();
  }
  s.batches++;
  SCOUT_TRY(hipEventRecord(ch.staged, c.s));
  channel_staged(ch, produced, true);
  if (out_m) *out_m = produced;
  return SCOUT_OK;
}

// Keeps a device copy of the channel's batch in slot k (batches stay resident in HBM).
int batch_store(Ctx& c, Channel& ch, int slot) {
  if (!c.stream) c.stream = new StreamState();
  StreamState& s = *c.stream;
  if (slot < 0) {
    set_error("batch_store: negative slot");
    return SCOUT_ERR_ARG;
  }
  if ((size_t)slot >= s.slots.size()) {
    s.slots.resize((size_t)slot + 1);
    s.slot_m.resize((size_t)slot + 1, 0);
  }
  SCOUT_CHECK(s.slots[(size_t)slot].reserve(2 * ch.m + 2));
  SCOUT_CHECK(channel_wait(c.s, ch));
  if (ch.m > 0)
    SCOUT_TRY(hipMemcpyAsync(s.slots[(size_t)slot].p, ch.d_uv.p, (size_t)(2 * ch.m) * 4, hipMemcpyDeviceToDevice, c.s));
  s.slot_m[(size_t)slot] = ch.m;
  return SCOUT_OK;
}

// Stages stored batch `slot` into the channel with a device-to-device copy on the copy
// stream (no PCIe transfer).
int batch_load(Ctx& c, Channel& ch, int slot) {
  if (!c.stream || slot < 0 || (size_t)slot >= c.stream->slots.size()) {
    set_error("batch_load: no such slot");
    return SCOUT_ERR_ARG;
  }
  int64_t m = c.stream->slot_m[(size_t)slot];
  SCOUT_CHECK(channel_wait(c.s, ch));  // an upload into the channel may still be in flight
  SCOUT_CHECK(ch.d_uv.reserve(2 * m + 2));
  SCOUT_CHECK(ch.d_ins.reserve(m + 1));
  SCOUT_CHECK(ch.d_done.reserve(m + 1));
  // a device-to-device copy on the compute stream: ordered after every earlier reader of the
  // channel's lines and before the work that follows, with no cross-stream wait
  if (m > 0)
    SCOUT_TRY(hipMemcpyAsync(ch.d_uv.p, c.stream->slots[(size_t)slot].p, (size_t)(2 * m) * 4, hipMemcpyDeviceToDevice,
                             c.s));
  SCOUT_TRY(hipEventRecord(ch.staged, c.s));
  channel_staged(ch, m, true);
  return SCOUT_OK;
}

// Grows a device buffer to `count` elements keeping its first `keep` elements (copied on the
// compute stream, which the caller has ordered after every user of the old buffer).
template <typename T>
int grow_keep(Ctx& c, DevBuf<T>& b, int64_t count, int64_t keep) {
  if (count <= b.cap) return SCOUT_OK;
  DevBuf<T> nb;
  SCOUT_CHECK(nb.reserve(count));
  if (keep > 0) SCOUT_TRY(hipMemcpyAsync(nb.p, b.p, (size_t)keep * sizeof(T), hipMemcpyDeviceToDevice, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  b.release();
  b = nb;
  return SCOUT_OK;
}

// Appends stored batch `slot` after the lines already staged in the channel (a serving step
// that takes several queued requests at once).
int batch_append(Ctx& c, Channel& ch, int slot) {
  if (!c.stream || slot < 0 || (size_t)slot >= c.stream->slots.size()) {
    set_error("batch_append: no such slot");
    return SCOUT_ERR_ARG;
  }
  int64_t add = c.stream->slot_m[(size_t)slot], old = ch.m;
  SCOUT_TRY(hipEventSynchronize(ch.staged));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  SCOUT_CHECK(grow_keep(c, ch.d_uv, 2 * (old + add) + 2, 2 * old));
  SCOUT_CHECK(grow_keep(c, ch.d_ins, old + add + 1, old));
  SCOUT_CHECK(grow_keep(c, ch.d_done, old + add + 1, 0));
  SCOUT_TRY(hipEventRecord(ch.staged, c.s));
  SCOUT_TRY(hipStreamWaitEvent(c.cs, ch.staged, 0));
  if (add > 0)
    SCOUT_TRY(hipMemcpyAsync(ch.d_uv.p + 2 * old, c.stream->slots[(size_t)slot].p, (size_t)(2 * add) * 4,
                             hipMemcpyDeviceToDevice, c.cs));
  SCOUT_TRY(hipEventRecord(ch.staged, c.cs));
  channel_staged(ch, old + add);
  return SCOUT_OK;
}

void stream_release(Ctx& c) {
  if (!c.stream) return;
  for (auto& b : c.stream->slots) b.release();
  c.stream->set.release();
  c.stream->owner.release();
  c.stream->cand.release();
  c.stream->cslot.release();
  c.stream->acc.release();
  c.stream->pos.release();
  c.stream->rowpre.release();
  delete c.stream;
  c.stream = nullptr;
}

// Device copies of the graph and every allocated solution, for replaying a stream.
struct Snapshot {
  DevBuf<int64_t> boff;
  DevBuf<AHead> ahead;
  DevBuf<int32_t> acap;
  DevBuf<int32_t> bcol, acol, aown;
  int64_t live = 0;
  DevBuf<int32_t> level, parent, label, color;
  DevBuf<uint8_t> mis;
  int64_t E = 0, A = 0, holes = 0;
  int64_t shape_gen = 0;
  bool sorted = false;
  bool has_level = false, has_label = false, has_color = false, has_mis = false;
};

int snapshot_save(Ctx& c) {
  if (c.g.host_bcol) {
    set_error("snapshot: not available for a host-resident graph");
    return SCOUT_ERR_UNSUPPORTED;
  }
  if (!c.snap) c.snap = new Snapshot();
  Snapshot& s = *c.snap;
  int64_t n = c.g.n;
  SCOUT_CHECK(s.boff.reserve(n + 1));
  SCOUT_CHECK(s.ahead.reserve(n + 1));
  SCOUT_CHECK(s.acap.reserve(n + 1));
  SCOUT_CHECK(s.bcol.reserve(c.g.E + 1));
  SCOUT_CHECK(s.acol.reserve(c.g.A + 1));
  SCOUT_CHECK(s.aown.reserve(c.g.A + 1));
  SCOUT_TRY(hipMemcpyAsync(s.boff.p, c.g.boff, (size_t)(n + 1) * 8, hipMemcpyDeviceToDevice, c.s));
  SCOUT_TRY(hipMemcpyAsync(s.ahead.p, c.g.ahead.p, (size_t)(n + 1) * sizeof(AHead), hipMemcpyDeviceToDevice, c.s));
  SCOUT_TRY(hipMemcpyAsync(s.acap.p, c.g.acap.p, (size_t)(n + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, c.s));
  SCOUT_TRY(hipMemcpyAsync(s.bcol.p, c.g.bcol, (size_t)c.g.E * 4, hipMemcpyDeviceToDevice, c.s));
  if (c.g.A > 0) {
    SCOUT_TRY(hipMemcpyAsync(s.acol.p, c.g.acol.p, (size_t)c.g.A * 4, hipMemcpyDeviceToDevice, c.s));
    SCOUT_TRY(hipMemcpyAsync(s.aown.p, c.g.aown.p, (size_t)c.g.A * 4, hipMemcpyDeviceToDevice, c.s));
  }
  s.E = c.g.E;
  s.shape_gen = c.g.shape_gen;
  s.A = c.g.A;
  s.live = c.g.live;
  s.holes = c.g.holes;
  s.sorted = c.g.sorted;
  auto save32 = [&](DevBuf<int32_t>& d, const int32_t* src, bool& has) -> int {
    has = src != nullptr;
    if (!src) return SCOUT_OK;
    SCOUT_CHECK(d.reserve(n));
    SCOUT_TRY(hipMemcpyAsync(d.p, src, (size_t)n * 4, hipMemcpyDeviceToDevice, c.s));
    return SCOUT_OK;
  };
  SCOUT_CHECK(save32(s.level, c.level, s.has_level));
  bool has_parent = false;
  SCOUT_CHECK(save32(s.parent, c.parent, has_parent));
  SCOUT_CHECK(save32(s.label, c.label, s.has_label));
  SCOUT_CHECK(save32(s.color, c.color, s.has_color));
  s.has_mis = c.mis != nullptr;
  if (c.mis) {
    SCOUT_CHECK(s.mis.reserve(n));
    SCOUT_TRY(hipMemcpyAsync(s.mis.p, c.mis, (size_t)n, hipMemcpyDeviceToDevice, c.s));
  }
  SCOUT_TRY(hipStreamSynchronize(c.s));
  return SCOUT_OK;
}

int snapshot_restore(Ctx& c) {
  if (!c.snap) {
    set_error("restore: no snapshot");
    return SCOUT_ERR_STATE;
  }
  Snapshot& s = *c.snap;
  int64_t n = c.g.n;
  if (c.g.host_bcol) {
    set_error("restore: not available for a host-resident graph");
    return SCOUT_ERR_UNSUPPORTED;
  }
  if (s.E != c.g.E) {
    // the base was compacted since the snapshot: restore its size
    (void)hipFree(c.g.bcol);
    SCOUT_TRY(hipMalloc((void**)&c.g.bcol, (size_t)std::max<int64_t>(s.E, 1) * 4));
  }
  SCOUT_TRY(hipMemcpyAsync(c.g.boff, s.boff.p, (size_t)(n + 1) * 8, hipMemcpyDeviceToDevice, c.s));
  SCOUT_TRY(hipMemcpyAsync(c.g.ahead.p, s.ahead.p, (size_t)(n + 1) * sizeof(AHead), hipMemcpyDeviceToDevice, c.s));
  SCOUT_TRY(hipMemcpyAsync(c.g.acap.p, s.acap.p, (size_t)(n + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, c.s));
  SCOUT_TRY(hipMemcpyAsync(c.g.bcol, s.bcol.p, (size_t)s.E * 4, hipMemcpyDeviceToDevice, c.s));

  SCOUT_CHECK(c.g.acol.reserve(s.A + 1));
  SCOUT_CHECK(c.g.aown.reserve(c.g.acol.cap));
  if (s.A > 0) {
    SCOUT_TRY(hipMemcpyAsync(c.g.acol.p, s.acol.p, (size_t)s.A * 4, hipMemcpyDeviceToDevice, c.s));
    SCOUT_TRY(hipMemcpyAsync(c.g.aown.p, s.aown.p, (size_t)s.A * 4, hipMemcpyDeviceToDevice, c.s));
  }
  c.g.E = s.E;
  c.g.base_gen++;
  c.g.shape_gen = s.shape_gen;
  c.g.A = s.A;
  c.g.live = s.live;
  c.g.holes = s.holes;
  c.g.sorted = s.sorted;
  if (s.has_level) {
    SCOUT_CHECK(ensure_state(c, SCOUT_STATE_LEVEL));
    SCOUT_TRY(hipMemcpyAsync(c.level, s.level.p, (size_t)n * 4, hipMemcpyDeviceToDevice, c.s));
    SCOUT_TRY(hipMemcpyAsync(c.parent, s.parent.p, (size_t)n * 4, hipMemcpyDeviceToDevice, c.s));
  }
  if (s.has_label) {
    SCOUT_CHECK(ensure_state(c, SCOUT_STATE_LABEL));
    SCOUT_TRY(hipMemcpyAsync(c.label, s.label.p, (size_t)n * 4, hipMemcpyDeviceToDevice, c.s));
  }
  if (s.has_color) {
    SCOUT_CHECK(ensure_state(c, SCOUT_STATE_COLOR));
    SCOUT_TRY(hipMemcpyAsync(c.color, s.color.p, (size_t)n * 4, hipMemcpyDeviceToDevice, c.s));
  }
  if (s.has_mis) {
    SCOUT_CHECK(ensure_state(c, SCOUT_STATE_MIS));
    SCOUT_TRY(hipMemcpyAsync(c.mis, s.mis.p, (size_t)n, hipMemcpyDeviceToDevice, c.s));
    c.mis_version++;
  }
  SCOUT_TRY(hipStreamSynchronize(c.s));
  return SCOUT_OK;
}

void snapshot_release(Ctx& c) {
  if (!c.snap) return;
  Snapshot& s = *c.snap;
  s.boff.release();
  s.ahead.release();
  s.acap.release();
  s.bcol.release();
  s.acol.release();
  s.aown.release();
  s.level.release();
  s.parent.release();
  s.label.release();
  s.color.release();
  s.mis.release();
  delete c.snap;
  c.snap = nullptr;
}

}  // namespace scout

using namespace scout;

extern "C" {

int scout_gen_batch(scout_ctx* x, int channel, int64_t size, double delete_frac, double tree_frac, uint64_t seed,
                    int64_t expected_total, int64_t* out_m) {
  if (!x || size < 0) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  Channel* ch = channel_of(c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  if (!c.g.boff) {
    set_error("gen_batch: no graph loaded");
    return SCOUT_ERR_STATE;
  }
  SCOUT_CHECK(gen_batch(c, *ch, size, delete_frac, tree_frac, seed, expected_total, out_m));
  return channel_used(c.s, *ch);
}

int scout_gen_mode(scout_ctx* x, int mode) {
  if (!x || (mode != SCOUT_GEN_UNIFORM && mode != SCOUT_GEN_DEGREE)) {
    set_error("gen_mode: unknown mode");
    return SCOUT_ERR_ARG;
  }
  Ctx& c = x->c;
  if (!c.stream) c.stream = new StreamState();
  c.stream->mode = mode;
  return SCOUT_OK;
}

int scout_batch_download(scout_ctx* x, int channel, int32_t* uv, int64_t cap_lines) {
  if (!x) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  Channel* ch = channel_of(c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  int64_t m = std::min(ch->m, cap_lines);
  SCOUT_CHECK(channel_wait(c.s, *ch));
  if (m > 0) SCOUT_TRY(hipMemcpyAsync(uv, ch->d_uv.p, (size_t)(2 * m) * 4, hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  return SCOUT_OK;
}

int scout_batch_store(scout_ctx* x, int channel, int slot) {
  if (!x) return SCOUT_ERR_ARG;
  SCOUT_TRY(hipSetDevice(x->c.dev));
  Channel* ch = channel_of(x->c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  SCOUT_CHECK(batch_store(x->c, *ch, slot));
  return channel_used(x->c.s, *ch);
}

int scout_batch_load(scout_ctx* x, int channel, int slot, int64_t* m) {
  if (!x) return SCOUT_ERR_ARG;
  SCOUT_TRY(hipSetDevice(x->c.dev));
  Channel* ch = channel_of(x->c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  SCOUT_CHECK(batch_load(x->c, *ch, slot));
  if (m) *m = ch->m;
  return SCOUT_OK;
}

int scout_batch_append(scout_ctx* x, int channel, int slot, int64_t* m) {
  if (!x) return SCOUT_ERR_ARG;
  SCOUT_TRY(hipSetDevice(x->c.dev));
  Channel* ch = channel_of(x->c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  SCOUT_CHECK(batch_append(x->c, *ch, slot));
  if (m) *m = ch->m;
  return SCOUT_OK;
}

int scout_snapshot(scout_ctx* x) {
  if (!x) return SCOUT_ERR_ARG;
  SCOUT_TRY(hipSetDevice(x->c.dev));
  return snapshot_save(x->c);
}

int scout_restore(scout_ctx* x) {
  if (!x) return SCOUT_ERR_ARG;
  SCOUT_TRY(hipSetDevice(x->c.dev));
  return snapshot_restore(x->c);
}

}  // extern "C"
