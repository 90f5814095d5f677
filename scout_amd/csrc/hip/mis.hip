// mis.hip: MIS maintenance with the reference's repair rules (LCL, radius 2), run as a
// sequential-equivalent parallel job.
//
// Rules (process_clusters_csr_kernel, src/mis/scout_mis.cu:290-362), per batch line in line
// order, on rows built as launchBatch builds them (:500-626):
//   row(x)   endpoint of the sub-batch     -> its 2-hop ball {x} + N(x) + N(N(x)),
//                                             in BFS discovery order
//            other vertex adjacent to one  -> N(x)
//            anything else                 -> no row (skipped)
//   insert (u, v), both members: r = max(u, v) leaves; then the first vertex c of row(r)
//            that is not a member, has a row and has no member in its row joins.
//   delete (u, v), exactly one member k: the other endpoint joins if its row holds no
//            member other than k.
// The reference runs these per cluster with racing threads; this file computes the result of
// running them one line after another in ascending line order.
//
// Sequential equivalence by speculative prefix commit: every pending line is evaluated
// against the committed state; minw[x] records the smallest pending line that would write x.
// A line whose reads include a vertex with minw < its own index may depend on an earlier
// line; the first such line k0 ends the round. Every line before k0 read nothing an earlier
// pending line writes, so its speculative outcome is its sequential outcome, and its writes
// commit. The first pending line never conflicts, so each round commits at least one line.
//
// The candidate order of row(r) is reproduced without building the ball: the sequence
// [r] + N(r) + N(y1) + N(y2) + ... (y_i in N(r) order) lists the ball's vertices in the same
// first-occurrence order, with repeats; a repeated candidate was already rejected at its first
// occurrence (nothing changes before the first acceptance), so the first accepted candidate
// is the same.
#include <algorithm>
#include <unordered_set>

#include "common.cuh"

namespace scout {

namespace {

constexpr int32_t kNoWriter = INT32_MAX;

struct MisView {
  GView g;
  const uint8_t* mis;
  const int32_t* tag;
  int32_t tag_ep;        // endpoint tag = 2*ep+1, hop-1 tag = 2*ep
  int32_t* minw;
  int32_t line;          // index of the line being evaluated
  bool check;            // pass B: report reads of vertices written by an earlier line
  int32_t dropped;       // vertex whose membership reads as 0 (the insertion's r), or -1
};

// Membership as seen by the evaluating line. In check mode a read of a vertex with an
// earlier pending writer sets *conflict.
__device__ __forceinline__ bool member(const MisView& s, int32_t x, bool* conflict) {
  if (s.check && s.minw[x] < s.line) *conflict = true;
  if (x == s.dropped) return false;
  return s.mis[x] != 0;
}

// Wave: is some vertex of row segment r[lo, hi) other than `skip` a member? (Any conflict in
// check mode also stops the scan.)
__device__ __forceinline__ bool wave_any_member(const MisView& s, const Row& r, int32_t skip, bool* conflict) {
  int64_t len = row_len(r);
  for (int64_t base = 0; base < len; base += kWave) {
    int64_t k = base + lane_id();
    bool hit = false, conf = false;
    if (k < len) {
      int32_t x = row_at(s.g, r, k);
      if (x >= 0 && x != skip) hit = member(s, x, &conf);
    }
    if (__ballot(conf)) {
      *conflict = true;
      return true;
    }
    if (__ballot(hit)) return true;
  }
  return false;
}

// Any member (other than skip) in the vertex's row: 2-hop ball for endpoints, N(x) for
// hop-1 vertices.
__device__ bool row_has_member(const MisView& s, int32_t x, bool two_hop, int32_t skip, bool* conflict) {
  if (x != skip) {
    bool conf = false;
    bool m = member(s, x, &conf);
    if (conf) {
      *conflict = true;
      return true;
    }
    if (m) return true;
  }
  Row r = row_of(s.g, x);
  if (wave_any_member(s, r, skip, conflict)) return true;
  if (!two_hop) return false;
  int64_t len = row_len(r);
  for (int64_t k = 0; k < len; k++) {
    int32_t y = row_at(s.g, r, k);
    if (y < 0) continue;
    if (wave_any_member(s, row_of(s.g, y), skip, conflict)) return true;
  }
  return false;
}

__device__ __forceinline__ int row_kind(const MisView& s, int32_t x) {
  int32_t t = s.tag[x];
  if (t == 2 * s.tag_ep + 1) return 2;
  if (t == 2 * s.tag_ep) return 1;
  return 0;
}

// Candidate c of the insertion rule: not a member, has a row, no member in its row. Returns
// 1 = accept, 0 = reject, -1 = conflict (check mode).
__device__ int try_candidate(const MisView& s, int32_t c) {
  bool conf = false;
  bool m = member(s, c, &conf);
  if (conf) return -1;
  if (m) return 0;
  int kind = row_kind(s, c);
  if (kind == 0) return 0;
  bool has = row_has_member(s, c, kind == 2, -1, &conf);
  if (conf) return -1;
  return has ? 0 : 1;
}

// Evaluates one line. Outputs the vertices it writes (w0 -> 0 or 1 by rule, w1 -> 1) or
// reports a conflict (check mode). Returns false on conflict.
__device__ bool eval_line(MisView s, int32_t u, int32_t v, bool ins, int32_t* w0, int32_t* w1) {
  *w0 = -1;
  *w1 = -1;
  bool conf = false;
  bool um = member(s, u, &conf);
  bool vm = member(s, v, &conf);
  if (conf) return false;
  if (ins) {
    if (!(um && vm)) return true;
    int32_t r = u > v ? u : v;
    *w0 = r;
    s.dropped = r;
    // candidate sequence [r] + N(r) + N(y) for y in N(r)
    int res = try_candidate(s, r);
    if (res < 0) return false;
    if (res > 0) {
      *w1 = r;
      return true;
    }
    Row rr = row_of(s.g, r);
    int64_t len = row_len(rr);
    for (int64_t k = 0; k < len; k++) {
      int32_t c = row_at(s.g, rr, k);
      if (c < 0) continue;
      res = try_candidate(s, c);
      if (res < 0) return false;
      if (res > 0) {
        *w1 = c;
        return true;
      }
    }
    for (int64_t k = 0; k < len; k++) {
      int32_t y = row_at(s.g, rr, k);
      if (y < 0) continue;
      Row ry = row_of(s.g, y);
      int64_t ly = row_len(ry);
      for (int64_t j = 0; j < ly; j++) {
        int32_t c = row_at(s.g, ry, j);
        if (c < 0) continue;
        res = try_candidate(s, c);
        if (res < 0) return false;
        if (res > 0) {
          *w1 = c;
          return true;
        }
      }
    }
    return true;
  }
  int32_t keep = -1, cand = -1;
  if (um && !vm) keep = u, cand = v;
  else if (!um && vm) keep = v, cand = u;
  if (cand < 0) return true;
  // the candidate is an endpoint: its row is its 2-hop ball
  bool has = row_has_member(s, cand, true, keep, &conf);
  if (conf) return false;
  if (!has) *w0 = cand;
  return true;
}

// Endpoint and hop-1 tags of a sub-batch.
__global__ void k_mis_tag_endpoints(const int32_t* uv, int64_t lo, int64_t hi, int32_t* tag, int32_t ep) {
  int64_t i = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  atomicMax(&tag[uv[2 * i]], 2 * ep + 1);
  atomicMax(&tag[uv[2 * i + 1]], 2 * ep + 1);
}

__global__ void k_mis_tag_hop1(GView g, const int32_t* uv, int64_t lo, int64_t hi, int32_t* tag, int32_t ep) {
  for (int64_t w = global_wave(); w < 2 * (hi - lo); w += total_waves()) {
    int32_t x = uv[2 * lo + w];
    Row r = row_of(g, x);
    wave_walk(g, r, 0, row_len(r), [&](int32_t y) { atomicMax(&tag[y], 2 * ep); });
  }
}

// Pass A: wave per pending line. Writes are recorded in w0/w1; minw[x] = smallest writer.
__global__ void k_mis_eval(GView g, const int32_t* uv, const uint8_t* ins, int64_t p, int64_t hi, const uint8_t* mis,
                           const int32_t* tag, int32_t ep, int32_t* minw, int32_t* w0s, int32_t* w1s) {
  for (int64_t k = p + global_wave(); k < hi; k += total_waves()) {
    MisView s{g, mis, tag, ep, minw, (int32_t)k, false, -1};
    int32_t w0, w1;
    eval_line(s, uv[2 * k], uv[2 * k + 1], ins[k] != 0, &w0, &w1);
    if (lane_id() == 0) {
      w0s[k] = w0;
      w1s[k] = w1;
      if (w0 >= 0) atomicMin(&minw[w0], (int32_t)k);
      if (w1 >= 0) atomicMin(&minw[w1], (int32_t)k);
    }
  }
}

// Pass B: the first pending line whose reads meet an earlier pending writer.
__global__ void k_mis_check(GView g, const int32_t* uv, const uint8_t* ins, int64_t p, int64_t hi, const uint8_t* mis,
                            const int32_t* tag, int32_t ep, int32_t* minw, int64_t* first_conflict) {
  for (int64_t k = p + 1 + global_wave(); k < hi; k += total_waves()) {
    if (k >= __hip_atomic_load(first_conflict, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) continue;
    MisView s{g, mis, tag, ep, minw, (int32_t)k, true, -1};
    int32_t w0, w1;
    bool ok = eval_line(s, uv[2 * k], uv[2 * k + 1], ins[k] != 0, &w0, &w1);
    if (!ok && lane_id() == 0) atomicMin((long long*)first_conflict, (long long)k);
  }
}

// Commit lines [p, k0); clear minw of every pending line's writes.
__global__ void k_mis_commit(const uint8_t* ins, int64_t p, int64_t k0, int64_t hi, const int32_t* w0s,
                             const int32_t* w1s, uint8_t* mis, int32_t* minw, int64_t* writes) {
  int64_t k = p + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= hi) return;
  int32_t w0 = w0s[k], w1 = w1s[k];
  if (k < k0) {
    if (w0 >= 0) mis[w0] = ins[k] ? 0 : 1;
    if (w1 >= 0) mis[w1] = 1;
    if (w0 >= 0) atomicAdd((unsigned long long*)writes, 1ULL);
  }
}

__global__ void k_mis_clear_minw(int64_t p, int64_t hi, const int32_t* w0s, const int32_t* w1s, int32_t* minw) {
  int64_t k = p + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= hi) return;
  if (w0s[k] >= 0) minw[w0s[k]] = kNoWriter;
  if (w1s[k] >= 0) minw[w1s[k]] = kNoWriter;
}

// Reference degree estimate per endpoint: base row length - |removed set| + |added list|
// (HybridGraph::degree, core/FlashGraph.h:267-274); removed entries are the deleted base
// entries plus the tombstones held in the inserted row.
__global__ void k_mis_degrees(GView g, const int32_t* uv, int64_t m, int64_t* deg) {
  for (int64_t w = global_wave(); w < 2 * m; w += total_waves()) {
    int32_t x = uv[w];
    Row row = row_of(g, x);
    int64_t b0 = row.b0, b1 = row.b1, a0 = row.a0, a1 = row.a1;
    int64_t removed = 0, added = 0;
    for (int64_t k = b0 + lane_id(); k < b1; k += kWave) removed += (g.bcol[k] < 0);
    for (int64_t k = a0 + lane_id(); k < a1; k += kWave) {
      int32_t y = g.acol[k];
      if (y >= 0) added++;
      else if (y != HOLE) removed++;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
      removed += __shfl_xor(removed, off);
      added += __shfl_xor(added, off);
    }
    if (lane_id() == 0) deg[w] = (b1 - b0) - removed + added;
  }
}

}  // namespace

struct MisState {
  DevBuf<int32_t> w0, w1;
  DevBuf<int64_t> deg;
  std::vector<int64_t> h_deg;
  int32_t tag_ep = 0;
};

void mis_release(Ctx& c) {
  if (!c.mis_state) return;
  c.mis_state->w0.release();
  c.mis_state->w1.release();
  c.mis_state->deg.release();
  delete c.mis_state;
  c.mis_state = nullptr;
}

int mis_repair(Ctx& c, Channel& ch, int64_t max_csr, scout_stats* st) {
  if (!c.mis_state) c.mis_state = new MisState();
  MisState& ms = *c.mis_state;
  int64_t m = ch.m;
  int32_t n = c.g.n;
  SCOUT_CHECK(ensure_state(c, SCOUT_STATE_MIS));
  SCOUT_CHECK(ensure_scratch(c, &c.mis_minw, kNoWriter));
  SCOUT_CHECK(ensure_scratch(c, &c.mis_tag, -1));
  if (m == 0) return SCOUT_OK;
  SCOUT_CHECK(ms.w0.reserve(m));
  SCOUT_CHECK(ms.w1.reserve(m));
  SCOUT_CHECK(ms.deg.reserve(2 * m));
  // Sub-batch cuts by the reference's estimate deg * 1000 against max_csr (:519-542).
  k_mis_degrees<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), ch.d_uv.p, m, ms.deg.p);
  ms.h_deg.resize((size_t)(2 * m));
  std::vector<int32_t> uv((size_t)(2 * m));
  std::vector<uint8_t> ins((size_t)m);
  SCOUT_TRY(hipMemcpyAsync(ms.h_deg.data(), ms.deg.p, (size_t)(2 * m) * 8, hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipMemcpyAsync(uv.data(), ch.d_uv.p, (size_t)(2 * m) * 4, hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipMemcpyAsync(ins.data(), ch.d_ins.p, (size_t)m, hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  std::vector<std::pair<int64_t, int64_t>> subs;
  {
    const int64_t MAX_BATCH = 1048576;
    int64_t idx = 0;
    std::unordered_set<int32_t> uniq;
    while (idx < m) {
      uniq.clear();
      int64_t cur = 0, se = 0, start = idx;
      while (idx < m && se < MAX_BATCH) {
        int32_t a = uv[(size_t)(2 * idx)], b = uv[(size_t)(2 * idx + 1)];
        int64_t est = 0;
        if (!uniq.count(a)) est += ms.h_deg[(size_t)(2 * idx)] * 1000;
        if (!uniq.count(b)) est += ms.h_deg[(size_t)(2 * idx + 1)] * 1000;
        if (se > 0 && cur + est >= max_csr) break;
        cur += est;
        uniq.insert(a);
        uniq.insert(b);
        se++;
        idx++;
      }
      subs.emplace_back(start, idx);
    }
  }
  for (int64_t k = 0; k < m; k++) (ins[(size_t)k] ? st->insertions : st->deletions)++;
  st->sub_batches += (int64_t)subs.size();
  int64_t rounds = 0;
  for (auto& sb : subs) {
    int64_t lo = sb.first, hi = sb.second;
    if (ms.tag_ep >= (1 << 29)) {
      SCOUT_TRY(hipMemsetAsync(c.mis_tag, 0xff, ((size_t)n + 1) * 4, c.s));
      ms.tag_ep = 0;
    }
    int32_t ep = ++ms.tag_ep;
    k_mis_tag_endpoints<<<grid_for(hi - lo), kBlock, 0, c.s>>>(ch.d_uv.p, lo, hi, c.mis_tag, ep);
    k_mis_tag_hop1<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), ch.d_uv.p, lo, hi, c.mis_tag, ep);
    int64_t p = lo;
    while (p < hi) {
      rounds++;
      k_mis_eval<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), ch.d_uv.p, ch.d_ins.p, p, hi, c.mis, c.mis_tag, ep,
                                                c.mis_minw, ms.w0.p, ms.w1.p);
      int64_t init[2] = {hi, 0};
      SCOUT_TRY(hipMemcpyAsync(c.d_box, init, sizeof(init), hipMemcpyHostToDevice, c.s));
      k_mis_check<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), ch.d_uv.p, ch.d_ins.p, p, hi, c.mis, c.mis_tag, ep,
                                                 c.mis_minw, c.d_box);
      SCOUT_CHECK(box_read(c, 1));
      int64_t k0 = c.h_box[0];
      k_mis_commit<<<grid_for(hi - p), kBlock, 0, c.s>>>(ch.d_ins.p, p, k0, hi, ms.w0.p, ms.w1.p, c.mis, c.mis_minw,
                                                         c.d_box + 1);
      k_mis_clear_minw<<<grid_for(hi - p), kBlock, 0, c.s>>>(p, hi, ms.w0.p, ms.w1.p, c.mis_minw);
      SCOUT_CHECK(box_read(c, 2));
      st->conflicts += c.h_box[1];
      p = k0;
    }
  }
  st->rounds += rounds;
  return SCOUT_OK;
}

}  // namespace scout

using namespace scout;

extern "C" int scout_repair_mis(scout_ctx* x, int channel, int64_t max_csr, scout_stats* st) {
  if (!x || !st) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  Channel* ch = channel_of(c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  if (!c.g.boff || !c.mis) {
    set_error("repair_mis: graph and MIS must be loaded first");
    return SCOUT_ERR_STATE;
  }
  if (max_csr <= 0) {
    set_error("repair_mis: max_csr must be positive");
    return SCOUT_ERR_ARG;
  }
  Span span;
  SCOUT_CHECK(span.begin(c.s));
  st->lines += ch->m;
  SCOUT_CHECK(mis_repair(c, *ch, max_csr, st));
  SCOUT_CHECK(span.end(c.s));
  st->repair_ms += span.ms();
  return SCOUT_OK;
}
