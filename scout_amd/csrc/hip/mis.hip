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
// Sequential equivalence by speculative prefix commit: the pending lines of a window are
// evaluated against the committed state; minw[x] records the smallest pending line that would
// write x. A line whose reads include a vertex with minw < its own index may depend on an
// earlier line; the first such line k0 ends the round. Every line before k0 read nothing an
// earlier pending line writes, so its speculative outcome is its sequential outcome, and its
// writes commit. The first pending line never conflicts, so each round commits at least one
// line. Lines past the window are not evaluated (they would be evaluated again after the
// conflict anyway): the window doubles while a round commits all of it and shrinks to twice
// the committed prefix after a conflict, so a round's work stays near twice its commits.
//
// Sub-batches: the reference cuts a batch into sub-batches and builds rows per sub-batch, so a
// vertex's row kind depends on the line's sub-batch. It is read per line: the endpoints of
// every sub-batch sit in one hash set of (vertex, sub-batch) keys, and a vertex has a row for
// a line iff it is an endpoint of the line's sub-batch or neighbours one. The graph does not
// change during the repair, so the lines of all sub-batches form one sequence, and one
// speculative window runs across sub-batch boundaries.
//
// Segments: the lines of an instance form one segment. Several segments (independent
// instances, which share no vertex) run their rounds in the same launches, each with its own
// window and first conflict. A single graph is the case of one instance.
//
// Counts instead of scans: mcount[x] = members among x's neighbours, kept exact across
// batches. "N(x) holds a member other than skip" is read from it in O(1) (less skip and the
// dropped vertex when they are member neighbours). Such a read depends on the neighbours'
// memberships. Pending lines mark, per vertex, the smallest line that adds a member
// neighbour (wn.x), the smallest that removes one (wn.y) and how many removals are pending
// (wncnt). Pass B takes "no member" as a conflict if an earlier line adds a neighbour, and
// "has a member" as a conflict only if an earlier line removes one and the pending removals
// could leave none. A 2-hop ball then costs one pass over N(x).
//
// Hop-2 candidates N(y), y in N(r), are searched by chunk tasks over many waves; the first
// decisive candidate in sequence order is the smallest (y position, entry position) key.
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
  bool sorted;
  const uint8_t* mis;
  const unsigned long long* eps;  // endpoint set: (vertex << 32 | sub-batch) keys, open addressing
  int64_t emask;                  // its capacity - 1 (a power of two)
  int32_t sb;                     // sub-batch of the line (its first line)
  const int32_t* mcount; // members among each vertex's neighbours (committed state)
  const int32_t* minw;   // smallest pending line writing the vertex
  const int2* wn;        // smallest pending line adding (x) / removing (y) a member neighbour
  const int32_t* wncnt;  // pending removals of member neighbours (any window line)
  int32_t line;          // index of the line being evaluated
  bool check;            // pass B: report reads of vertices written by an earlier line
  int32_t dropped;       // vertex whose membership reads as 0 (the insertion's r), or -1
  int64_t* scanned;      // row entries read by this wave (counted on every lane alike)
};

// Membership as seen by the evaluating line. In check mode a read of a vertex with an
// earlier pending writer sets *conflict.
__device__ __forceinline__ bool member(const MisView& s, int32_t x, bool* conflict) {
  if (s.check && s.minw[x] < s.line) *conflict = true;
  if (x == s.dropped) return false;
  return s.mis[x] != 0;
}

// Live entries equal to y in row x (base entries not marked deleted, inserted entries); the
// shorter of the two rows is searched (the adjacency is symmetric).
__device__ __forceinline__ int32_t live_count(const MisView& s, int32_t x, int32_t y) {
  int64_t lx = (s.g.boff[x + 1] - s.g.boff[x]) + s.g.ahead[x].len;
  int64_t ly = (s.g.boff[y + 1] - s.g.boff[y]) + s.g.ahead[y].len;
  if (ly < lx) {
    int32_t tmp = x;
    x = y;
    y = tmp;
  }
  int64_t lo, hi;
  base_range(s.g, s.sorted, x, y, lo, hi);
  int32_t c = 0;
  for (int64_t k = lo; k < hi; k++) c += s.g.bcol[k] == y;
  AHead hx = s.g.ahead[x];
  for (int64_t k = hx.off, e = k + hx.len; k < e; k++) c += s.g.acol[k] == y;
  return c;
}

// Does N(x) hold a member other than `skip`? The committed member-neighbour count of x,
// less the entries of skip when it is a member, less those of the dropped vertex (a member
// that reads as 0). In check mode, a conflict if pending earlier lines could change the
// answer (or write skip or the dropped vertex).
__device__ __forceinline__ bool nbr_has_member(const MisView& s, int32_t x, int32_t skip, bool* conflict) {
  int32_t c = s.mcount[x];
  int2 wn = s.check ? s.wn[x] : make_int2(kNoWriter, kNoWriter);
  // removals an earlier line may make (all pending removals: an upper bound)
  int32_t rem = (c > 0 && wn.y < s.line) ? s.wncnt[x] : 0;
  // a row holds a neighbour at most twice (an inserted copy next to a deleted base entry is
  // not live), so counts above 2 + rem stay positive without the adjacency tests
  if (skip >= 0) {
    if (s.check && s.minw[skip] < s.line) *conflict = true;
    if (c > 0 && c <= 2 + rem && s.mis[skip]) c -= live_count(s, x, skip);
  }
  if (s.dropped >= 0 && s.dropped != skip) {
    if (s.check && s.minw[s.dropped] < s.line) *conflict = true;
    if (c > 0 && c <= 2 + rem && s.mis[s.dropped]) c -= live_count(s, x, s.dropped);
  }
  bool has = c > 0;
  if (s.check && (has ? c <= rem : wn.x < s.line)) *conflict = true;
  return has;
}

// Any member (other than skip) in the vertex's row: 2-hop ball for endpoints, N(x) for
// hop-1 vertices. The hop-2 part takes y in N(x) order, 64 at a time, one lane each; the
// first y in order whose test conflicts or finds a member decides.
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
  {
    bool conf = false;
    bool has = nbr_has_member(s, x, skip, &conf);
    if (conf) {
      *conflict = true;
      return true;
    }
    if (has) return true;
  }
  if (!two_hop) return false;
  Row r = row_of(s.g, x);
  int64_t len = row_len(r);
  *s.scanned += len;
  // four chunks of 64 y's in flight per iteration, decided chunk by chunk in order
  constexpr int kU = 4;
  for (int64_t g0 = 0; g0 < len; g0 += kU * kWave) {
    bool conf[kU], has[kU];
#pragma unroll
    for (int j = 0; j < kU; j++) {
      int64_t k = g0 + j * kWave + lane_id();
      int32_t y = k < len ? row_at(s.g, r, k) : -1;
      conf[j] = false;
      has[j] = y >= 0 && nbr_has_member(s, y, skip, &conf[j]);
    }
#pragma unroll
    for (int j = 0; j < kU; j++) {
      unsigned long long cm = __ballot(conf[j]), hm = __ballot(has[j]);
      unsigned long long any = cm | hm;
      if (any) {
        int b = __ffsll((long long)any) - 1;
        if ((cm >> b) & 1ULL) *conflict = true;
        return true;
      }
    }
  }
  return false;
}

constexpr unsigned long long kNoKey = ~0ULL;

__device__ __forceinline__ unsigned long long ep_key(int32_t x, int32_t sb) {
  return ((unsigned long long)(uint32_t)x << 32) | (uint32_t)sb;
}

__device__ __forceinline__ int64_t ep_slot(unsigned long long key, int64_t mask) {
  unsigned long long h = key * 0x9E3779B97F4A7C15ULL;
  return (int64_t)((h ^ (h >> 31)) & (unsigned long long)mask);
}

// Is x an endpoint of the line's sub-batch? (The set is at most a quarter full, so every
// probe sequence meets an empty slot.)
__device__ __forceinline__ bool is_endpoint(const MisView& s, int32_t x) {
  unsigned long long key = ep_key(x, s.sb);
  for (int64_t h = ep_slot(key, s.emask);; h = (h + 1) & s.emask) {
    unsigned long long k = s.eps[h];
    if (k == key) return true;
    if (k == kNoKey) return false;
  }
}

// Does the non-endpoint c have a row for the line, i.e. does it neighbour an endpoint of the
// line's sub-batch? By the wave.
__device__ bool near_endpoint(const MisView& s, int32_t c) {
  Row r = row_of(s.g, c);
  int64_t len = row_len(r);
  *s.scanned += len;
  for (int64_t g0 = 0; g0 < len; g0 += kWave) {
    int64_t k = g0 + lane_id();
    int32_t z = k < len ? row_at(s.g, r, k) : -1;
    if (__ballot(z >= 0 && is_endpoint(s, z))) return true;
  }
  return false;
}

// Candidate c of the insertion rule (not a member, has a row, no member in its row), one
// lane: 1 = accept, 0 = reject, -1 = conflict (check mode), 3 = an endpoint (its row is its
// 2-hop ball, evaluated by the wave), 4 / 5 = accept / conflict if c has a row (the wave
// tests that with near_endpoint).
//
// Any candidate with a member neighbour (other than the dropped r) is rejected, whatever its
// kind (an endpoint's 2-hop ball contains N(c)), and so is a member; a reject needs only that
// its own read is settled (no earlier pending writer). The member-neighbour count is tested
// first: in a near-maximal set it rejects most candidates with one read.
//
// near_ep: c is known to neighbour an endpoint (c in N(r)). Otherwise whether c has a row
// costs a pass over N(c), so it is tested only when the outcome depends on it: a candidate
// without a row is rejected whatever it reads, so it neither accepts nor conflicts.
__device__ __forceinline__ int candidate_hop1(const MisView& s, int32_t c, bool near_ep) {
  bool conf_n = false, conf_m = false;
  bool has = nbr_has_member(s, c, -1, &conf_n);
  if (!conf_n && has) return 0;
  bool m = member(s, c, &conf_m);
  if (!conf_m && m) return 0;
  bool conf = conf_n || conf_m;
  if (is_endpoint(s, c)) return conf ? -1 : 3;
  if (near_ep) return conf ? -1 : 1;
  return conf ? 5 : 4;
}

// Outcome of a candidate code by the wave (codes 3 to 5 of candidate_hop1; 5 = a conflict
// if the candidate has a row, endpoint or not).
__device__ __forceinline__ int resolve_candidate(const MisView& s, int code, int32_t c);

// An endpoint candidate, by the wave.
__device__ int try_endpoint(const MisView& s, int32_t c) {
  bool conf = false;
  bool m = member(s, c, &conf);
  if (conf) return -1;
  if (m) return 0;
  bool has = row_has_member(s, c, true, -1, &conf);
  if (conf) return -1;
  return has ? 0 : 1;
}

__device__ __forceinline__ int resolve_candidate(const MisView& s, int code, int32_t c) {
  if (code == 3) return try_endpoint(s, c);
  if (code == 4) return near_endpoint(s, c) ? 1 : 0;
  if (code == 5) return (is_endpoint(s, c) || near_endpoint(s, c)) ? -1 : 0;
  return code;
}

// Candidates of row r in order, 256 at a time (one lane each for four chunks of 64, taken
// chunk by chunk in order; endpoints go through the wave when their turn comes). Returns 1
// with *accepted, 0 if none is accepted, -1 on conflict. near_ep: r is the row of an endpoint.
__device__ int first_candidate(const MisView& s, const Row& r, int32_t* accepted, bool near_ep) {
  constexpr int kU = 4;
  int64_t len = row_len(r);
  *s.scanned += len;
  for (int64_t g0 = 0; g0 < len; g0 += kU * kWave) {
    int32_t cs[kU];
    int res[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      int64_t k = g0 + u * kWave + lane_id();
      cs[u] = k < len ? row_at(s.g, r, k) : -1;
    }
#pragma unroll
    for (int u = 0; u < kU; u++) res[u] = cs[u] >= 0 ? candidate_hop1(s, cs[u], near_ep) : 0;
    for (int u = 0; u < kU; u++) {
      unsigned long long pending = __ballot(res[u] != 0);
      while (pending) {
        int j = __ffsll((long long)pending) - 1;
        pending &= pending - 1;
        int rj = __shfl(res[u], j);
        int32_t cj = __shfl(cs[u], j);
        rj = resolve_candidate(s, rj, cj);
        if (rj == -1) return -1;
        if (rj == 1) {
          *accepted = cj;
          return 1;
        }
      }
    }
  }
  return 0;
}

// Evaluates one line. Outputs the vertices it writes (w0 -> 0 or 1 by rule, w1 -> 1) or
// reports a conflict (check mode). Returns false on conflict.
//
// With defer_hop2, an insertion whose candidates [r] + N(r) are all rejected stops there with
// *hop2 = true (w0 = r known, w1 open): the hop-2 candidates N(y), y in N(r), are then
// searched by chunk tasks spread over many waves (k_mm_hop2_*).
__device__ bool eval_line(MisView s, int32_t u, int32_t v, bool ins, int32_t* w0, int32_t* w1, bool defer_hop2,
                          bool* hop2) {
  *w0 = -1;
  *w1 = -1;
  *hop2 = false;
  bool conf = false;
  bool um = member(s, u, &conf);
  bool vm = member(s, v, &conf);
  if (conf) return false;
  if (ins) {
    if (!(um && vm)) return true;
    int32_t r = u > v ? u : v;
    *w0 = r;
    s.dropped = r;
    // candidate sequence [r] + N(r) + N(y) for y in N(r); r is an endpoint
    int res = try_endpoint(s, r);
    if (res < 0) return false;
    if (res > 0) {
      *w1 = r;
      return true;
    }
    Row rr = row_of(s.g, r);
    int32_t acc = -1;
    res = first_candidate(s, rr, &acc, true);
    if (res < 0) return false;
    if (res > 0) {
      *w1 = acc;
      return true;
    }
    if (defer_hop2) {
      *hop2 = true;
      return true;
    }
    int64_t len = row_len(rr);
    for (int64_t k = 0; k < len; k++) {
      int32_t y = row_at(s.g, rr, k);
      if (y < 0) continue;
      res = first_candidate(s, row_of(s.g, y), &acc, false);
      if (res < 0) return false;
      if (res > 0) {
        *w1 = acc;
        return true;
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

// Distinct endpoints: the first slot to stamp x with this repair's generation appends x.
__global__ void k_mis_deg_claim(const int32_t* uv, int64_t m, int32_t* stamp, int32_t gen, int32_t* list,
                                int64_t* count) {
  int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= 2 * m) return;
  int32_t x = uv[w];
  if (atomicMax(&stamp[x], gen) != gen) list[atomicAdd((unsigned long long*)count, 1ULL)] = x;
}

// Reference degree estimate per distinct endpoint: base row length - |removed set| + |added
// list| (HybridGraph::degree, core/FlashGraph.h:267-274); removed entries are the deleted
// base entries plus the tombstones held in the inserted row. Block per vertex, so a hub's row
// is read by 256 threads.
__global__ void k_mis_degrees(GView g, const int32_t* list, const int64_t* count, int32_t* val) {
  __shared__ int32_t part[kBlock / kWave];
  int64_t nx = *count;
  for (int64_t i = blockIdx.x; i < nx; i += gridDim.x) {
    int32_t x = list[i];
    Row row = row_of(g, x);
    int32_t net = 0;  // added - removed
    for (int64_t k = row.b0 + threadIdx.x; k < row.b1; k += blockDim.x) net -= (g.bcol[k] < 0);
    for (int64_t k = row.a0 + threadIdx.x; k < row.a1; k += blockDim.x) {
      int32_t y = g.acol[k];
      net += y >= 0 ? 1 : (y != HOLE ? -1 : 0);
    }
    net = (int32_t)wave_sum(net);
    if (lane_id() == 0) part[threadIdx.x / kWave] = net;
    __syncthreads();
    if (threadIdx.x == 0) {
      int32_t t = 0;
      for (int j = 0; j < (int)(blockDim.x / kWave); j++) t += part[j];
      val[x] = (int32_t)(row.b1 - row.b0) + t;
    }
    __syncthreads();
  }
}

__global__ void k_mis_deg_gather(const int32_t* uv, int64_t m, const int32_t* val, int32_t* deg) {
  int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w < 2 * m) deg[w] = val[uv[w]];
}

// Endpoint set: key (x, sub-batch) for both endpoints of every line (linear probing; the
// capacity is at least four times the keys, so every insertion finds a slot).
__global__ void k_mis_ep_insert(const int32_t* uv, const int32_t* lsb, int64_t m, unsigned long long* eps,
                                int64_t mask) {
  for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < 2 * m; w += (int64_t)gridDim.x * blockDim.x) {
    unsigned long long key = ep_key(uv[w], lsb[w >> 1]);
    for (int64_t h = ep_slot(key, mask);; h = (h + 1) & mask) {
      unsigned long long prev = atomicCAS(&eps[h], kNoKey, key);
      if (prev == kNoKey || prev == key) break;
    }
  }
}

// --- round engine over segments ---------------------------------------------------------
// Segment table (int64, on the device): a header (slots below), then per segment: pending
// start p, end hi, first conflict k0, window size win, and woff (exclusive prefix of the
// current window sizes, S + 1).
enum : int {
  kHSegs = 0,      // segments S
  kHSlots = 1,     // window lines over all segments
  kHPending = 2,   // segments with lines left
  kHRounds = 3,    // non-empty rounds
  kHWrites = 4,    // committed lines that wrote
  kHEvals = 5,     // line evaluations of pass A
  kHEntries = 6,   // row entries read
  kHReqA = 7,      // hop-2 searches requested in pass A this round
  kHReqB = 8,      // ... and in pass B
  kHTasks = 9,     // hop-2 chunk tasks of the current pass
  kHHop2 = 10,     // hop-2 searches of pass A over all rounds
  kSegHdr = 12
};
constexpr int64_t kMinWindow = 64;
constexpr int64_t kMaxWindow = 1 << 20;

struct SegView {
  int64_t* t;
  int64_t stride;
  __device__ int64_t* p() const { return t + kSegHdr; }
  __device__ int64_t* hi() const { return t + kSegHdr + stride; }
  __device__ int64_t* k0() const { return t + kSegHdr + 2 * stride; }
  __device__ int64_t* win() const { return t + kSegHdr + 3 * stride; }
  __device__ int64_t* woff() const { return t + kSegHdr + 4 * stride; }
};

inline int64_t seg_table_size(int64_t stride) { return kSegHdr + 5 * stride + 1; }

// Window slot -> (segment, line): binary search over woff.
__device__ __forceinline__ int64_t slot_line(const SegView& sv, int64_t slot, int32_t* sg_out) {
  const int64_t* woff = sv.woff();
  int64_t lo = 0, hi = sv.t[kHSegs];  // find the last s with woff[s] <= slot
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) >> 1;
    if (woff[mid] <= slot) lo = mid;
    else hi = mid;
  }
  *sg_out = (int32_t)lo;
  return sv.p()[lo] + (slot - woff[lo]);
}

// Device arrays of one MIS repair.
struct MisDev {
  GView g;
  bool sorted;
  const int32_t* uv;
  const uint8_t* ins;
  uint8_t* mis;
  const int32_t* lsb;              // sub-batch of each line (its first line)
  const unsigned long long* eps;   // endpoint set of all sub-batches
  int64_t emask;
  int32_t* mcount;
  int32_t* minw;
  int2* wn;
  int32_t* wncnt;
  int32_t* w0s;
  int32_t* w1s;
  // hop-2 candidate searches of one pass: requests (line, segment), their best result, the
  // y-block prefix, overflow flags, and the chunk tasks
  int32_t* h2_line;
  int32_t* h2_seg;
  unsigned long long* h2_best;
  int64_t* h2_off;
  int32_t* h2_flag;
  uint64_t* h2_task;
  int32_t* h2_y;                // per task: its y
  int64_t* h2_rsb;              // per request: r << 32 | sub-batch of the line
  unsigned long long* h2_keyA;  // per line: the first accepted candidate pass A found at hop 2
  int64_t h2_rcap, h2_tcap;
};

__device__ __forceinline__ MisView view_of(const MisDev& d, int64_t k, bool check, int64_t* scanned) {
  return MisView{d.g,       d.sorted, d.mis,         d.eps,  d.emask, d.lsb[k], d.mcount,
                 d.minw,    d.wn,     d.wncnt,       (int32_t)k, check,  -1,       scanned};
}

constexpr int32_t kHop2Pending = -3;  // w1 of a line whose hop-2 search is still running
constexpr int64_t kHop2Requests = 65536;   // deferred hop-2 searches per pass (16-bit task field)
constexpr int64_t kHop2Tasks = 1 << 22;    // hop-2 chunk tasks per pass
constexpr int64_t kHop2Chunk = 512;  // row entries of N(y) per hop-2 task
constexpr unsigned long long kNoResult = ~0ULL;

// The insertion's r of line k.
__device__ __forceinline__ int32_t line_r(const MisDev& d, int64_t k) {
  int32_t u = d.uv[2 * k], v = d.uv[2 * k + 1];
  return u > v ? u : v;
}

// Marks the line's write w: minw[w] and, over w's row, the neighbour marks (a removal when
// w leaves the set).
__device__ __forceinline__ void mark_write(const MisDev& d, int32_t w, int64_t k, bool removal) {
  if (w < 0) return;
  if (lane_id() == 0) atomicMin(&d.minw[w], (int32_t)k);
  Row r = row_of(d.g, w);
  if (removal) {
    wave_walk_wide<8>(d.g, r, 0, row_len(r), [&](int32_t z) {
      atomicMin(&d.wn[z].y, (int32_t)k);
      atomicAdd(&d.wncnt[z], 1);
    });
  } else {
    wave_walk_wide<8>(d.g, r, 0, row_len(r), [&](int32_t z) { atomicMin(&d.wn[z].x, (int32_t)k); });
  }
}

// Records a deferred hop-2 search of line k (segment sg) in the pass's request list; false if
// the list is full (the caller then searches serially). best starts at `seed`: pass B needs
// only the candidates up to the one pass A accepted (the same reads in check mode end at a
// conflict or at that acceptance).
__device__ __forceinline__ bool hop2_request(const MisDev& d, int64_t* count, int64_t k, int32_t sg,
                                             unsigned long long seed) {
  int64_t idx = 0;
  if (lane_id() == 0) {
    idx = (int64_t)atomicAdd((unsigned long long*)count, 1ULL);
    if (idx < d.h2_rcap) {
      d.h2_line[idx] = (int32_t)k;
      d.h2_seg[idx] = sg;
      d.h2_rsb[idx] = ((int64_t)line_r(d, k) << 32) | (uint32_t)d.lsb[k];
      d.h2_best[idx] = seed;
      d.h2_flag[idx] = 0;
    }
  }
  idx = __shfl(idx, 0);
  return idx < d.h2_rcap;
}

// Pass A: wave per window line. Writes go to w0/w1 and are marked (minw[w] = smallest writer
// of w, and the neighbour marks of w's row). A deferred hop-2 search leaves w1 pending.
// Launched with kBlock threads: the bound lets the compiler give it more than the 128 VGPRs a
// 1024-thread block would allow (at 128 it spilled to scratch, 3x slower).
__global__ void __launch_bounds__(kBlock) k_mm_eval(MisDev d, SegView sv) {
  int64_t slots = sv.t[kHSlots];
  int64_t scanned = 0;
  for (int64_t t = global_wave(); t < slots; t += total_waves()) {
    int32_t sg;
    int64_t k = slot_line(sv, t, &sg);
    MisView s = view_of(d, k, false, &scanned);
    int32_t w0, w1;
    bool hop2 = false;
    eval_line(s, d.uv[2 * k], d.uv[2 * k + 1], d.ins[k] != 0, &w0, &w1, true, &hop2);
    if (hop2) {
      if (hop2_request(d, &sv.t[kHReqA], k, sg, kNoResult)) {
        w1 = kHop2Pending;
      } else {
        eval_line(s, d.uv[2 * k], d.uv[2 * k + 1], d.ins[k] != 0, &w0, &w1, false, &hop2);
        if (lane_id() == 0) d.h2_keyA[k] = kNoResult;
      }
    }
    if (lane_id() == 0) {
      d.w0s[k] = w0;
      d.w1s[k] = w1;
    }
    mark_write(d, w0, k, d.ins[k] != 0);
    mark_write(d, w1, k, false);
  }
  if (lane_id() == 0 && scanned) atomicAdd((unsigned long long*)&sv.t[kHEntries], (unsigned long long)scanned);
}

// Pass B: per segment, the first window line whose reads meet an earlier pending writer.
__global__ void k_mm_check(MisDev d, SegView sv) {
  int64_t slots = sv.t[kHSlots];
  int64_t* k0 = sv.k0();
  int64_t scanned = 0;
  for (int64_t t = global_wave(); t < slots; t += total_waves()) {
    int32_t sg;
    int64_t k = slot_line(sv, t, &sg);
    if (k == sv.p()[sg]) continue;  // the first pending line never conflicts
    if (k >= __hip_atomic_load(&k0[sg], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) continue;
    MisView s = view_of(d, k, true, &scanned);
    int32_t w0, w1;
    bool hop2 = false;
    bool ok = eval_line(s, d.uv[2 * k], d.uv[2 * k + 1], d.ins[k] != 0, &w0, &w1, true, &hop2);
    if (ok && hop2 && !hop2_request(d, &sv.t[kHReqB], k, sg, d.h2_keyA[k]))
      ok = eval_line(s, d.uv[2 * k], d.uv[2 * k + 1], d.ins[k] != 0, &w0, &w1, false, &hop2);
    if (!ok && lane_id() == 0) atomicMin((long long*)&k0[sg], (long long)k);
  }
  if (lane_id() == 0 && scanned) atomicAdd((unsigned long long*)&sv.t[kHEntries], (unsigned long long)scanned);
}

// Positions of N(r) that request i searches. Pass B skips a request whose line is already
// past its segment's first conflict, and stops at the y of the candidate pass A accepted.
__device__ __forceinline__ int64_t hop2_y_limit(const MisDev& d, const SegView& sv, int64_t i, int pass) {
  int64_t k = d.h2_line[i];
  int64_t len = row_len(row_of(d.g, line_r(d, k)));
  if (pass == 0) return len;
  if (k >= sv.k0()[d.h2_seg[i]]) return 0;
  unsigned long long seed = d.h2_best[i];
  return seed == kNoResult ? len : min(len, (int64_t)(seed >> 33) + 1);
}

// Hop-2 plan (one block): y-blocks of 256 entries of N(r) per request, prefix in h2_off; the
// task counter starts from 0.
__global__ void k_mm_hop2_plan(MisDev d, SegView sv, int pass) {
  __shared__ int64_t scan[kBlock];
  __shared__ int64_t carry;
  int64_t n = min(sv.t[pass == 0 ? kHReqA : kHReqB], d.h2_rcap);
  if (threadIdx.x == 0) {
    carry = 0;
    sv.t[kHTasks] = 0;
    if (pass == 0) sv.t[kHHop2] += n;
  }
  __syncthreads();
  for (int64_t base = 0; base < n; base += blockDim.x) {
    int64_t i = base + threadIdx.x;
    int64_t nb = 0;
    if (i < n) nb = (hop2_y_limit(d, sv, i, pass) + kBlock - 1) / kBlock;
    scan[threadIdx.x] = nb;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
      int64_t y = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 0;
      __syncthreads();
      scan[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < n) d.h2_off[i] = carry + scan[threadIdx.x] - nb;
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry += scan[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) d.h2_off[n] = carry;
}

// Hop-2 expand: block per y-block; thread per y in N(r): one task per kHop2Chunk entries of
// N(y) (none when y is a settled member). A request whose tasks do not fit is flagged and
// searched serially by the resume.
__global__ void k_mm_hop2_expand(MisDev d, SegView sv, int pass) {
  __shared__ int64_t scan[kBlock];
  __shared__ int64_t base_s;
  int64_t n = min(sv.t[pass == 0 ? kHReqA : kHReqB], d.h2_rcap);
  int64_t nblocks = n > 0 ? d.h2_off[n] : 0;
  for (int64_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
    int64_t lo = 0, hi = n;  // last request i with h2_off[i] <= b
    while (hi - lo > 1) {
      int64_t mid = (lo + hi) >> 1;
      if (d.h2_off[mid] <= b) lo = mid;
      else hi = mid;
    }
    int64_t i = lo;
    int32_t r = line_r(d, d.h2_line[i]);
    Row rr = row_of(d.g, r);
    int64_t yi = (b - d.h2_off[i]) * kBlock + threadIdx.x;
    int64_t nch = 0;
    int32_t y = -1;
    if (yi < hop2_y_limit(d, sv, i, pass)) {
      y = row_at(d.g, rr, yi);
      // a member y whose membership no earlier pending line changes gives every c in N(y) a
      // settled member neighbour: all of N(y) is rejected without reading it
      bool settled = y >= 0 && d.mis[y] && (pass == 0 || d.minw[y] >= d.h2_line[i]);
      if (y >= 0 && !settled) nch = (row_len(row_of(d.g, y)) + kHop2Chunk - 1) / kHop2Chunk;
    }
    scan[threadIdx.x] = nch;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
      int64_t y = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 0;
      __syncthreads();
      scan[threadIdx.x] += y;
      __syncthreads();
    }
    if (threadIdx.x == blockDim.x - 1) base_s = (int64_t)atomicAdd((unsigned long long*)&sv.t[kHTasks], (unsigned long long)scan[threadIdx.x]);
    __syncthreads();
    int64_t at = base_s + scan[threadIdx.x] - nch;
    if (at + nch > d.h2_tcap || nch > 0x1ffff) {
      if (nch > 0) d.h2_flag[i] = 1;
    } else {
      for (int64_t ch = 0; ch < nch; ch++) {
        d.h2_task[at + ch] = ((uint64_t)i << 48) | ((uint64_t)yi << 17) | (uint64_t)ch;
        d.h2_y[at + ch] = y;
      }
    }
    __syncthreads();
  }
}

__device__ __forceinline__ unsigned long long hop2_key(int64_t yi, int64_t pos, int code) {
  return ((unsigned long long)yi << 33) | ((unsigned long long)pos << 2) | (unsigned long long)code;
}

// Hop-2 scan: wave per task, the candidates of one chunk of N(y) in order, 256 at a time
// (four per lane, loaded together); the first decisive one (accept, or a conflict in pass B)
// lowers the request's best key (y index, position).
//
// A hop-2 candidate c with a member neighbour is rejected without the adjustments for r:
// if c is not in N(r), r is not among its member neighbours; if it is, c was a hop-1
// candidate, rejected there by the same reads (the search reached hop 2). The read is
// settled unless pending removals by earlier lines could leave c without member neighbours
// (then a conflict, if c has a row).
__global__ void k_mm_hop2_scan(MisDev d, SegView sv, int pass) {
  constexpr int kU = 4;
  int64_t ntask = min(sv.t[kHTasks], d.h2_tcap);
  bool check = pass == 1;
  int64_t scanned = 0;
  for (int64_t t = global_wave(); t < ntask; t += total_waves()) {
    uint64_t task = d.h2_task[t];
    int32_t y = d.h2_y[t];
    int64_t i = (int64_t)(task >> 48), yi = (int64_t)((task >> 17) & 0x7fffffffULL), ch = (int64_t)(task & 0x1ffffULL);
    if (d.h2_flag[i]) continue;
    int64_t c0 = ch * kHop2Chunk;
    unsigned long long* best = &d.h2_best[i];
    if (__shfl(__hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), 0) <= hop2_key(yi, c0, 0)) continue;
    int64_t k = d.h2_line[i];
    int64_t rsb = d.h2_rsb[i];
    Row ry = row_of(d.g, y);
    int64_t c1 = min(c0 + kHop2Chunk, row_len(ry));
    MisView s{d.g,     d.sorted, d.mis,      d.eps, d.emask, (int32_t)(uint32_t)rsb, d.mcount,
              d.minw,  d.wn,     d.wncnt,    (int32_t)k, check, (int32_t)(rsb >> 32),  &scanned};
    scanned += c1 - c0;
    bool done = false;
    for (int64_t g0 = c0; g0 < c1 && !done; g0 += kU * kWave) {
      if (__shfl(__hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), 0) <= hop2_key(yi, g0, 0)) break;
      int32_t cs[kU], mc[kU];
      int2 wn[kU];
      int res[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        int64_t q = g0 + u * kWave + lane_id();
        cs[u] = q < c1 ? row_at(d.g, ry, q) : -1;
      }
#pragma unroll
      for (int u = 0; u < kU; u++) {
        mc[u] = cs[u] >= 0 ? d.mcount[cs[u]] : 1;
        wn[u] = (check && cs[u] >= 0) ? d.wn[cs[u]] : make_int2(kNoWriter, kNoWriter);
      }
#pragma unroll
      for (int u = 0; u < kU; u++) {
        if (cs[u] < 0) res[u] = 0;
        else if (mc[u] > 0) res[u] = (wn[u].y < k && mc[u] <= d.wncnt[cs[u]]) ? 5 : 0;
        else res[u] = candidate_hop1(s, cs[u], false);
      }
      for (int u = 0; u < kU && !done; u++) {
        unsigned long long pending = __ballot(res[u] != 0);
        while (pending) {
          int j = __ffsll((long long)pending) - 1;
          pending &= pending - 1;
          int rj = __shfl(res[u], j);
          int32_t cj = __shfl(cs[u], j);
          rj = resolve_candidate(s, rj, cj);
          if (rj == -1 || rj == 1) {
            if (lane_id() == 0) atomicMin(best, hop2_key(yi, g0 + u * kWave + j, rj == 1 ? 1 : 0));
            done = true;
            break;
          }
        }
      }
    }
  }
  if (lane_id() == 0 && scanned) atomicAdd((unsigned long long*)&sv.t[kHEntries], (unsigned long long)scanned);
}

// Hop-2 resume: wave per request. Pass A: the accepted candidate becomes w1 and is marked.
// Pass B: a conflict before any acceptance ends the segment's prefix there. A flagged request
// is searched serially.
__global__ void k_mm_hop2_resume(MisDev d, SegView sv, int pass) {
  int64_t n = min(sv.t[pass == 0 ? kHReqA : kHReqB], d.h2_rcap);
  int64_t scanned = 0;
  for (int64_t i = global_wave(); i < n; i += total_waves()) {
    int64_t k = d.h2_line[i];
    int32_t sg = d.h2_seg[i];
    int32_t r = line_r(d, k);
    bool conflict = false;
    int32_t w1 = -1;
    if (d.h2_flag[i]) {
      MisView s = view_of(d, k, pass == 1, &scanned);
      int32_t w0;
      bool hop2 = false;
      conflict = !eval_line(s, d.uv[2 * k], d.uv[2 * k + 1], d.ins[k] != 0, &w0, &w1, false, &hop2);
    } else {
      unsigned long long key = d.h2_best[i];
      if (key != kNoResult) {
        if ((key & 3ULL) == 0) {
          conflict = true;
        } else {
          int64_t yi = (int64_t)(key >> 33), pos = (int64_t)((key >> 2) & 0x7fffffffULL);
          int32_t y = row_at(d.g, row_of(d.g, r), yi);
          w1 = row_at(d.g, row_of(d.g, y), pos);
        }
      }
    }
    if (pass == 0) {
      if (lane_id() == 0) {
        d.w1s[k] = w1;
        d.h2_keyA[k] = d.h2_flag[i] ? kNoResult : d.h2_best[i];
      }
      mark_write(d, w1, k, false);
    } else if (conflict && lane_id() == 0) {
      atomicMin((long long*)&sv.k0()[sg], (long long)k);
    }
  }
  if (lane_id() == 0 && scanned) atomicAdd((unsigned long long*)&sv.t[kHEntries], (unsigned long long)scanned);
}

// Wave per window line: commits each segment's prefix [p, k0) (memberships and the member
// counts of the written vertices' neighbours) and clears the pending marks of every window
// line's writes. Committed lines write distinct vertices (a second writer of a vertex reads
// it and conflicts), so the count updates do not overlap.
__global__ void k_mm_commit(MisDev d, SegView sv) {
  int64_t slots = sv.t[kHSlots];
  const int64_t* k0 = sv.k0();
  int64_t writes = 0;
  for (int64_t t = global_wave(); t < slots; t += total_waves()) {
    int32_t sg;
    int64_t k = slot_line(sv, t, &sg);
    int32_t w0 = d.w0s[k], w1 = d.w1s[k];
    bool commit = k < k0[sg];
    bool ins = d.ins[k] != 0;
    if (commit && lane_id() == 0) {
      if (w0 >= 0) d.mis[w0] = ins ? 0 : 1;
      if (w1 >= 0) d.mis[w1] = 1;
      writes += w0 >= 0;
    }
    if (lane_id() == 0) {
      if (w0 >= 0) d.minw[w0] = kNoWriter;
      if (w1 >= 0) d.minw[w1] = kNoWriter;
    }
    for (int j = 0; j < 2; j++) {
      int32_t w = j == 0 ? w0 : w1;
      if (w < 0) continue;
      int32_t delta = commit ? ((j == 0 && ins) ? -1 : 1) : 0;
      bool removal = j == 0 && ins;
      Row r = row_of(d.g, w);
      wave_walk_wide<8>(d.g, r, 0, row_len(r), [&](int32_t z) {
        d.wn[z] = make_int2(kNoWriter, kNoWriter);
        if (removal) d.wncnt[z] = 0;
        if (delta) atomicAdd(&d.mcount[z], delta);
      });
    }
  }
  if (lane_id() == 0 && writes) atomicAdd((unsigned long long*)&sv.t[kHWrites], (unsigned long long)writes);
}

// Member neighbours of every vertex (wave per vertex).
__global__ void k_mcount_init(GView g, const uint8_t* mis, int32_t* mcount) {
  for (int64_t v = global_wave(); v < g.n; v += total_waves()) {
    Row r = row_of(g, (int32_t)v);
    int64_t len = row_len(r);
    int64_t c = 0;
    for (int64_t k = lane_id(); k < len; k += kWave) {
      int32_t z = row_at(g, r, k);
      c += z >= 0 && mis[z];
    }
    c = wave_sum(c);
    if (lane_id() == 0) mcount[v] = (int32_t)c;
  }
}

// Member counts after one applied batch: every line that changed the graph adds (insertion)
// or removes (deletion) the other endpoint's membership.
__global__ void k_mcount_adjust(const int32_t* uv, const uint8_t* ins, const uint8_t* done, int64_t m,
                                const uint8_t* mis, int32_t* mcount) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !done[i]) return;
  int32_t u = uv[2 * i], v = uv[2 * i + 1];
  int32_t sgn = ins[i] ? 1 : -1;
  if (mis[v]) atomicAdd(&mcount[u], sgn);
  if (mis[u]) atomicAdd(&mcount[v], sgn);
}

// Block-wide exclusive scan of sizes[s] (s < S) into woff, in chunks of the block size;
// returns the total on every thread.
__device__ int64_t block_woff(int64_t S, int64_t* woff, SegView sv, int64_t* scan, int64_t* carry_s) {
  if (threadIdx.x == 0) *carry_s = 0;
  __syncthreads();
  for (int64_t base = 0; base < S; base += blockDim.x) {
    int64_t s = base + threadIdx.x;
    int64_t size = s < S ? sv.k0()[s] - sv.p()[s] : 0;
    scan[threadIdx.x] = size;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
      int64_t y = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 0;
      __syncthreads();
      scan[threadIdx.x] += y;
      __syncthreads();
    }
    if (s < S) woff[s] = *carry_s + scan[threadIdx.x] - size;
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) *carry_s += scan[threadIdx.x];
    __syncthreads();
  }
  int64_t total = *carry_s;
  __syncthreads();
  return total;
}

// One block: advances every segment (p = k0), resizes its window (k0 = window end), rebuilds
// woff and the slot count, and counts pending segments and non-empty rounds.
__global__ void k_mm_round_end(SegView sv) {
  __shared__ int64_t scan[kBlock];
  __shared__ int64_t carry_s, active_s, pending_s;
  int64_t* t = sv.t;
  int64_t S = t[kHSegs];
  int64_t *p = sv.p(), *hi = sv.hi(), *k0 = sv.k0(), *win = sv.win();
  if (threadIdx.x == 0) {
    active_s = pending_s = 0;
    t[kHReqA] = 0;
    t[kHReqB] = 0;
  }
  __syncthreads();
  for (int64_t s = threadIdx.x; s < S; s += blockDim.x) {
    int64_t ps = p[s], hs = hi[s], ks = k0[s], ws = win[s];
    if (ps < hs) {
      atomicAdd((unsigned long long*)&active_s, 1ULL);
      int64_t done = ks - ps;
      ws = (ks == ps + min(ws, hs - ps)) ? min(2 * ws, kMaxWindow) : max(kMinWindow, 2 * done);
    }
    ps = ks;
    if (ps < hs) atomicAdd((unsigned long long*)&pending_s, 1ULL);
    p[s] = ps;
    win[s] = ws;
    k0[s] = ps + min(ws, hs - ps);
  }
  __syncthreads();
  int64_t total = block_woff(S, sv.woff(), sv, scan, &carry_s);
  if (threadIdx.x == 0) {
    sv.woff()[S] = total;
    t[kHSlots] = total;
    t[kHPending] = pending_s;
    if (active_s) t[kHRounds] += 1;
    t[kHEvals] += total;
  }
}

}  // namespace

// Set of vertex ids for one sub-batch: open addressing with linear probing, at most a quarter
// full (it doubles as it fills), emptied in O(size). Small sub-batches stay in cache.
struct VertexSet {
  std::vector<int32_t> slot;  // -1 = empty
  std::vector<uint64_t> used; // occupied slots
  uint64_t mask = 0;
  VertexSet() : slot(1024, -1), mask(1023) {}
  uint64_t find(int32_t x) const {
    uint64_t h = ((uint64_t)(uint32_t)x * 0x9E3779B97F4A7C15ULL) >> 20;
    for (h &= mask;; h = (h + 1) & mask)
      if (slot[h] == -1 || slot[h] == x) return h;
  }
  bool has(int32_t x) const { return slot[find(x)] == x; }
  void insert(int32_t x) {
    uint64_t h = find(x);
    if (slot[h] == x) return;
    slot[h] = x;
    used.push_back(h);
    if (used.size() * 4 > slot.size()) grow();
  }
  void grow() {
    std::vector<int32_t> keys;
    keys.reserve(used.size());
    for (uint64_t h : used) keys.push_back(slot[h]);
    slot.assign(slot.size() * 2, -1);
    mask = slot.size() - 1;
    used.clear();
    for (int32_t x : keys) {
      uint64_t h = find(x);
      slot[h] = x;
      used.push_back(h);
    }
  }
  void clear() {
    for (uint64_t h : used) slot[h] = -1;
    used.clear();
  }
};

struct MisState {
  DevBuf<int32_t> w0, w1, uv2, lsb;
  DevBuf<uint8_t> ins2;
  DevBuf<int32_t> deg;
  DevBuf<int64_t> seg;
  DevBuf<int32_t> dstamp, dval;  // per-vertex degree stamps and values
  DevBuf<int32_t> dlist;          // distinct endpoints of the batch
  DevBuf<int64_t> dcount;         // their number
  int32_t dgen = 0;
  DevBuf<unsigned long long> eps;  // endpoint set
  int64_t ecap = 0;                // its slots (a power of two, at least 4 per key)
  PinBuf<int32_t> h_deg, h_uv, h_lsb;  // host copies for the sub-batch cuts
  PinBuf<uint8_t> h_ins;
  VertexSet uniq;
  unsigned grid = kListGrid;
  bool use_graph = true;
  int group_rounds = 16;
  int64_t window0 = 256;
  DevBuf<int32_t> mcount, wncnt;
  DevBuf<int2> wn;
  int64_t mc_version = -1, mc_apply = -1, mc_gen = -1;  // what mcount was computed for
  DevBuf<int32_t> h2_line, h2_seg, h2_flag;
  DevBuf<unsigned long long> h2_best;
  DevBuf<int64_t> h2_off;
  DevBuf<uint64_t> h2_task;
  DevBuf<int32_t> h2_y;
  DevBuf<int64_t> h2_rsb;
  DevBuf<unsigned long long> h2_keyA;
  int64_t h2_rcap = kHop2Requests, h2_tcap = kHop2Tasks;  // list capacities (fallback: serial)
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> key;
  // complete mode (scout_mis_mode): per-vertex completion state (0 none, 1 undecided
  // candidate, 2 joins, 3 stays out) and counts of undecided lower-id candidate neighbours
  bool complete = false;
  DevBuf<uint8_t> cst;
  DevBuf<int32_t> ccnt;
};

void mis_release(Ctx& c) {
  if (!c.mis_state) return;
  MisState& ms = *c.mis_state;
  ms.cst.release();
  ms.ccnt.release();
  ms.w0.release();
  ms.w1.release();
  ms.uv2.release();
  ms.lsb.release();
  ms.eps.release();
  ms.ins2.release();
  ms.deg.release();
  ms.h_deg.release();
  ms.h_uv.release();
  ms.h_lsb.release();
  ms.h_ins.release();
  ms.dstamp.release();
  ms.dval.release();
  ms.dlist.release();
  ms.dcount.release();
  ms.seg.release();
  ms.mcount.release();
  ms.wn.release();
  ms.wncnt.release();
  ms.h2_line.release();
  ms.h2_seg.release();
  ms.h2_flag.release();
  ms.h2_best.release();
  ms.h2_off.release();
  ms.h2_task.release();
  ms.h2_y.release();
  ms.h2_rsb.release();
  ms.h2_keyA.release();
  if (ms.exec) (void)hipGraphExecDestroy(ms.exec);
  delete c.mis_state;
  c.mis_state = nullptr;
}

namespace {

// Sub-batch cuts of lines [lo, hi) by the reference's estimate deg * 1000 per new endpoint
// against max_csr, at most 2^20 lines each (src/mis/scout_mis.cu:519-542). lsb[k] = the
// first line of k's sub-batch; returns the number of sub-batches.
int64_t mis_cuts(const int32_t* uv, const int32_t* deg, int64_t lo, int64_t hi, int64_t max_csr, VertexSet& uniq,
                 int32_t* lsb) {
  const int64_t MAX_BATCH = 1048576;
  int64_t idx = lo, subs = 0;
  while (idx < hi) {
    uniq.clear();
    int64_t cur = 0, se = 0, start = idx;
    while (idx < hi && se < MAX_BATCH) {
      int32_t a = uv[2 * idx], b = uv[2 * idx + 1];
      int64_t est = 0;
      if (!uniq.has(a)) est += (int64_t)deg[2 * idx] * 1000;
      if (!uniq.has(b)) est += (int64_t)deg[2 * idx + 1] * 1000;
      if (se > 0 && cur + est >= max_csr) break;
      cur += est;
      uniq.insert(a);
      uniq.insert(b);
      lsb[idx] = (int32_t)start;
      se++;
      idx++;
    }
    subs++;
  }
  return subs;
}

MisDev dev_of(Ctx& c, MisState& ms) {
  return MisDev{c.g.view(), c.g.sorted,   ms.uv2.p,     ms.ins2.p,    c.mis,        ms.lsb.p,
                ms.eps.p,   ms.ecap - 1,  ms.mcount.p, c.mis_minw,   ms.wn.p,      ms.wncnt.p,   ms.w0.p,
                ms.w1.p,    ms.h2_line.p, ms.h2_seg.p,  ms.h2_best.p, ms.h2_off.p,  ms.h2_flag.p,
                ms.h2_task.p, ms.h2_y.p,    ms.h2_rsb.p,  ms.h2_keyA.p, ms.h2_rcap,   ms.h2_tcap};
}

void mm_launch_round(Ctx& c, MisState& ms, SegView sv) {
  MisDev d = dev_of(c, ms);
  // pass A, its hop-2 searches, pass B, its hop-2 searches, commit, round end
  k_mm_eval<<<ms.grid, kBlock, 0, c.s>>>(d, sv);
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) k_mm_check<<<ms.grid, kBlock, 0, c.s>>>(d, sv);
    k_mm_hop2_plan<<<1, kBlock, 0, c.s>>>(d, sv, pass);
    k_mm_hop2_expand<<<ms.grid, kBlock, 0, c.s>>>(d, sv, pass);
    k_mm_hop2_scan<<<ms.grid, kBlock, 0, c.s>>>(d, sv, pass);
    k_mm_hop2_resume<<<ms.grid, kBlock, 0, c.s>>>(d, sv, pass);
  }
  k_mm_commit<<<ms.grid, kBlock, 0, c.s>>>(d, sv);
  k_mm_round_end<<<1, kBlock, 0, c.s>>>(sv);
}

void mm_launch_group(Ctx& c, MisState& ms, SegView sv) {
  for (int r = 0; r < ms.group_rounds; r++) mm_launch_round(c, ms, sv);
}

// Captures one group of rounds, unless the graph in hand holds the same buffers and settings.
int mm_graph(Ctx& c, MisState& ms, SegView sv) {
  GView g = c.g.view();
  std::vector<const void*> key = {ms.uv2.p, ms.ins2.p, sv.t, ms.w0.p, ms.w1.p, c.mis, ms.lsb.p, ms.eps.p,
                                  (const void*)(intptr_t)ms.ecap, c.mis_minw, ms.mcount.p, ms.wn.p, ms.wncnt.p,
                                  (const void*)(intptr_t)c.g.sorted, ms.h2_line.p, ms.h2_seg.p, ms.h2_flag.p,
                                  ms.h2_best.p, ms.h2_off.p, ms.h2_task.p, ms.h2_y.p, ms.h2_rsb.p, ms.h2_keyA.p,
                                  (const void*)(intptr_t)ms.h2_rcap,
                                  (const void*)(intptr_t)ms.h2_tcap, g.boff, g.bcol, g.ahead, g.acol, g.bsamp, g.touched,
                                  (const void*)(intptr_t)g.n,
                                  (const void*)(intptr_t)sv.stride, (const void*)(intptr_t)ms.grid,
                                  (const void*)(intptr_t)ms.group_rounds};
  if (ms.exec && key == ms.key) return SCOUT_OK;
  if (ms.exec) (void)hipGraphExecDestroy(ms.exec);
  ms.exec = nullptr;
  ms.key.clear();
  hipGraph_t graph = nullptr;
  SCOUT_TRY(hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal));
  mm_launch_group(c, ms, sv);
  hipError_t e = hipStreamEndCapture(c.s, &graph);
  if (e != hipSuccess) {
    set_error(std::string("mis: graph capture failed: ") + hipGetErrorString(e));
    return SCOUT_ERR_HIP;
  }
  e = hipGraphInstantiate(&ms.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    ms.exec = nullptr;
    set_error(std::string("mis: graph instantiation failed: ") + hipGetErrorString(e));
    return SCOUT_ERR_HIP;
  }
  ms.key = key;
  return SCOUT_OK;
}

// --- complete mode ----------------------------------------------------------------------
// After the rules, every vertex within two hops of a batch endpoint joins the set in ascending
// id unless it or a neighbour is a member (the completion the oracle restates as a sequential
// scan). Only undominated vertices of the region (no member neighbour: mcount == 0) can join,
// and the ascending scan over them is the greedy MIS of the subgraph they induce: it runs in
// Kahn rounds over the id-ordered DAG, a candidate deciding once every lower-id candidate
// neighbour has decided.

__global__ void k_zero_i64(int64_t* p, int count) {
  if ((int)threadIdx.x < count) p[threadIdx.x] = 0;
}

// The batch endpoints, listed once (stamp epoch).
__global__ void k_cm_endpoints(const int32_t* uv, int64_t m, int32_t* stamp, int32_t ep, int32_t* list, int64_t* count,
                               int64_t cap) {
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < 2 * m; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = base + threadIdx.x;
    bool take = false;
    int32_t x = 0;
    if (i < 2 * m) {
      x = uv[i];
      take = atomicExch(&stamp[x], ep) != ep;
    }
    block_append(take, x, list, count, cap);
  }
}

// Wave per listed vertex of list[lo, hi): its neighbours not listed before are appended.
__global__ void k_cm_expand(GView g, int32_t* list, int64_t lo, int64_t hi, int32_t* stamp, int32_t ep,
                            int64_t* count, int64_t cap) {
  for (int64_t i = lo + global_wave(); i < hi; i += total_waves()) {
    Row r = row_of(g, list[i]);
    int64_t len = row_len(r);
    for (int64_t base = 0; base < len; base += kWave) {
      int64_t k = base + lane_id();
      int32_t w = k < len ? row_at(g, r, k) : -1;
      bool take = w >= 0 && stamp[w] != ep && atomicExch(&stamp[w], ep) != ep;
      wave_append(take, w, list, count, cap);
    }
  }
}

// Region vertices that are neither members nor next to one become candidates.
__global__ void k_cm_candidates(const int32_t* list, int64_t nl, const uint8_t* mis, const int32_t* mcount,
                                uint8_t* cst, int32_t* cand, int64_t* count, int64_t cap) {
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nl; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = base + threadIdx.x;
    bool take = false;
    int32_t v = 0;
    if (i < nl) {
      v = list[i];
      take = !mis[v] && mcount[v] == 0;
      if (take) cst[v] = 1;
    }
    block_append(take, v, cand, count, cap);
  }
}

// Wave per candidate: its lower-id candidate neighbours; none makes it ready at once.
__global__ void k_cm_lower(GView g, const int32_t* cand, int64_t nc, const uint8_t* cst, int32_t* ccnt, int32_t* ready,
                           int64_t* ready_count, int64_t cap) {
  for (int64_t i = global_wave(); i < nc; i += total_waves()) {
    int32_t v = cand[i];
    Row r = row_of(g, v);
    int64_t len = row_len(r), cnt = 0;
    for (int64_t k = lane_id(); k < len; k += kWave) {
      int32_t w = row_at(g, r, k);
      cnt += w >= 0 && w < v && cst[w] == 1;
    }
    cnt = wave_sum(cnt);
    if (lane_id() == 0) ccnt[v] = (int32_t)cnt;
    wave_append(lane_id() == 0 && cnt == 0, v, ready, ready_count, cap);
  }
}

// Wave per ready candidate: it joins unless a lower-id candidate neighbour joined.
__global__ void k_cm_decide(GView g, const int32_t* ready, const int64_t* ready_count, uint8_t* cst) {
  int64_t nr = *ready_count;
  for (int64_t i = global_wave(); i < nr; i += total_waves()) {
    int32_t v = ready[i];
    Row r = row_of(g, v);
    int64_t len = row_len(r);
    bool blocked = false;
    for (int64_t base = 0; base < len && !blocked; base += kWave) {
      int64_t k = base + lane_id();
      int32_t w = k < len ? row_at(g, r, k) : -1;
      blocked = __ballot(w >= 0 && w < v && cst[w] == 2) != 0;
    }
    if (lane_id() == 0) cst[v] = blocked ? 3 : 2;
  }
}

// Wave per decided candidate: each higher-id undecided candidate neighbour has one lower
// neighbour fewer to wait for; those reaching zero are ready next round.
__global__ void k_cm_release(GView g, const int32_t* ready, const int64_t* ready_count, const uint8_t* cst,
                             int32_t* ccnt, int32_t* next, int64_t* next_count, int64_t cap) {
  int64_t nr = *ready_count;
  for (int64_t i = global_wave(); i < nr; i += total_waves()) {
    int32_t v = ready[i];
    Row r = row_of(g, v);
    int64_t len = row_len(r);
    for (int64_t base = 0; base < len; base += kWave) {
      int64_t k = base + lane_id();
      int32_t w = k < len ? row_at(g, r, k) : -1;
      bool take = w > v && cst[w] == 1 && atomicSub(&ccnt[w], 1) == 1;
      wave_append(take, w, next, next_count, cap);
    }
  }
}

// Wave per candidate: joiners become members and raise their neighbours' member counts; the
// completion state is cleared for the next batch. *added += joiners.
__global__ void k_cm_commit(GView g, const int32_t* cand, int64_t nc, uint8_t* cst, uint8_t* mis, int32_t* mcount,
                            int64_t* added) {
  int64_t joined = 0;
  for (int64_t i = global_wave(); i < nc; i += total_waves()) {
    int32_t v = cand[i];
    bool join = cst[v] == 2;
    if (join) {
      Row r = row_of(g, v);
      int64_t len = row_len(r);
      for (int64_t k = lane_id(); k < len; k += kWave) {
        int32_t w = row_at(g, r, k);
        if (w >= 0) atomicAdd(&mcount[w], 1);
      }
      joined++;
    }
    if (lane_id() == 0) {
      if (join) mis[v] = 1;
      cst[v] = 0;
    }
  }
  block_add(lane_id() == 0 ? joined : 0, added);
}

int mis_complete(Ctx& c, Channel& ch, MisState& ms, scout_stats* st) {
  int64_t m = ch.m;
  int32_t n = c.g.n;
  if (m == 0) return SCOUT_OK;
  int64_t cap = (int64_t)n + 1;
  SCOUT_CHECK(c.la.reserve(cap));
  SCOUT_CHECK(c.lb.reserve(cap));
  SCOUT_CHECK(c.lc.reserve(cap));
  SCOUT_CHECK(c.ld.reserve(cap));
  if (ms.cst.cap < cap) {
    ms.cst.release();
    SCOUT_CHECK(ms.cst.reserve(cap));
    SCOUT_TRY(hipMemsetAsync(ms.cst.p, 0, (size_t)ms.cst.cap, c.s));
  }
  SCOUT_CHECK(ms.ccnt.reserve(cap));
  int32_t ep = next_epoch(c);
  int64_t* box = c.d_box;
  // box: [0] region size, [1] candidates, [2]/[3] ready lists, [4] joiners
  SCOUT_TRY(hipMemsetAsync(box, 0, 5 * sizeof(int64_t), c.s));
  k_cm_endpoints<<<grid_for(2 * m), kBlock, 0, c.s>>>(ch.d_uv.p, m, c.stamp, ep, c.la.p, box, cap);
  int64_t lo = 0;
  for (int hop = 0; hop < 2; hop++) {
    SCOUT_CHECK(box_read(c, 1));
    int64_t hi = std::min(c.h_box[0], cap);
    k_cm_expand<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), c.la.p, lo, hi, c.stamp, ep, box, cap);
    lo = hi;
  }
  SCOUT_CHECK(box_read(c, 1));
  int64_t nr = std::min(c.h_box[0], cap);
  k_cm_candidates<<<grid_for(nr), kBlock, 0, c.s>>>(c.la.p, nr, c.mis, ms.mcount.p, ms.cst.p, c.lb.p, box + 1, cap);
  SCOUT_CHECK(box_read(c, 2));
  int64_t nc = c.h_box[1];
  int64_t rounds = 0;
  if (nc > 0) {
    k_cm_lower<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), c.lb.p, nc, ms.cst.p, ms.ccnt.p, c.lc.p, box + 2, cap);
    int32_t* cur = c.lc.p;
    int32_t* nxt = c.ld.p;
    int p = 0;
    while (true) {
      SCOUT_CHECK(box_read(c, 4));
      if (c.h_box[2 + p] == 0) break;
      rounds++;
      k_zero_i64<<<1, 64, 0, c.s>>>(box + 2 + (p ^ 1), 1);
      k_cm_decide<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), cur, box + 2 + p, ms.cst.p);
      k_cm_release<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), cur, box + 2 + p, ms.cst.p, ms.ccnt.p, nxt,
                                                   box + 2 + (p ^ 1), cap);
      std::swap(cur, nxt);
      p ^= 1;
    }
    k_cm_commit<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), c.lb.p, nc, ms.cst.p, c.mis, ms.mcount.p, box + 4);
    SCOUT_CHECK(box_read(c, 5));
    st->completed += c.h_box[4];
  }
  st->rounds += rounds;
  return SCOUT_OK;
}

}  // namespace

// Instances own disjoint vertex sets and their lines sit in the staged batch instance after
// instance: lines [off[i], off[i+1]) belong to instance i and form segment i. Every instance
// gets the sub-batch cuts it would get alone, so the result for each instance is the one its
// lines give when repaired alone.
int mis_repair_multi(Ctx& c, Channel& ch, int32_t ninst, const int64_t* off, int64_t max_csr, scout_stats* st) {
  if (!c.mis_state) c.mis_state = new MisState();
  MisState& ms = *c.mis_state;
  int64_t m = ch.m;
  int32_t n = c.g.n;
  SCOUT_CHECK(ensure_state(c, SCOUT_STATE_MIS));
  SCOUT_CHECK(ensure_scratch(c, &c.mis_minw, kNoWriter));
  if (m == 0) return SCOUT_OK;
  SCOUT_CHECK(ms.w0.reserve(m));
  SCOUT_CHECK(ms.w1.reserve(m));
  SCOUT_CHECK(ms.deg.reserve(2 * m));
  SCOUT_CHECK(ms.uv2.reserve(2 * m));
  SCOUT_CHECK(ms.ins2.reserve(m));
  SCOUT_CHECK(ms.lsb.reserve(m));
  SCOUT_CHECK(ms.h2_keyA.reserve(m));
  SCOUT_CHECK(ms.h2_line.reserve(ms.h2_rcap));
  SCOUT_CHECK(ms.h2_seg.reserve(ms.h2_rcap));
  SCOUT_CHECK(ms.h2_flag.reserve(ms.h2_rcap));
  SCOUT_CHECK(ms.h2_best.reserve(ms.h2_rcap));
  SCOUT_CHECK(ms.h2_off.reserve(ms.h2_rcap + 1));
  SCOUT_CHECK(ms.h2_task.reserve(ms.h2_tcap));
  SCOUT_CHECK(ms.h2_y.reserve(ms.h2_tcap));
  SCOUT_CHECK(ms.h2_rsb.reserve(ms.h2_rcap));
  // endpoint set: at least 4 slots per key; grows by doubling (its size is in the graph key)
  if (ms.ecap < 8 * m || !ms.eps.p) {
    if (ms.ecap < 1024) ms.ecap = 1024;
    while (ms.ecap < 8 * m) ms.ecap *= 2;
    ms.eps.release();
    SCOUT_CHECK(ms.eps.reserve(ms.ecap));
  }
  int64_t stride = std::max<int64_t>(ninst, 64);
  SCOUT_CHECK(ms.seg.reserve(seg_table_size(stride)));
  // the table stride is part of the captured graph: keep the largest seen
  stride = (ms.seg.cap - kSegHdr - 1) / 5;
  // member counts for the graph as applied: one batch on from the last repair, or recomputed
  bool fresh = ms.mcount.cap >= (int64_t)n + 1 && ms.mc_version == c.mis_version && ms.mc_gen == c.g.base_gen;
  SCOUT_CHECK(ms.mcount.reserve((int64_t)n + 1));
  if (ms.wn.cap < (int64_t)n + 1) {
    ms.wn.release();
    ms.wncnt.release();
    SCOUT_CHECK(ms.wn.reserve((int64_t)n + 1));
    SCOUT_CHECK(ms.wncnt.reserve((int64_t)n + 1));
    SCOUT_TRY(hipMemsetAsync(ms.wn.p, 0x7f, (size_t)ms.wn.cap * 8, c.s));  // 0x7f7f7f7f > any line
    SCOUT_TRY(hipMemsetAsync(ms.wncnt.p, 0, (size_t)ms.wncnt.cap * 4, c.s));
  }
  if (fresh && ms.mc_apply == c.apply_seq - 1) {
    k_mcount_adjust<<<grid_for(m), kBlock, 0, c.s>>>(ch.d_uv.p, ch.d_ins.p, ch.d_done.p, m, c.mis, ms.mcount.p);
  } else if (!(fresh && ms.mc_apply == c.apply_seq)) {
    k_mcount_init<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), c.mis, ms.mcount.p);
  }
  ms.mc_version = c.mis_version;
  ms.mc_gen = c.g.base_gen;
  ms.mc_apply = c.apply_seq;
  if (!ms.dstamp.p || ms.dstamp.cap < (int64_t)n + 1 || ms.dgen == INT32_MAX) {
    ms.dstamp.release();
    ms.dval.release();
    SCOUT_CHECK(ms.dstamp.reserve((int64_t)n + 1));
    SCOUT_CHECK(ms.dval.reserve((int64_t)n + 1));
    SCOUT_TRY(hipMemsetAsync(ms.dstamp.p, 0, (size_t)ms.dstamp.cap * 4, c.s));
    ms.dgen = 0;
  }
  ms.dgen++;
  SCOUT_CHECK(ms.dlist.reserve(2 * m));
  SCOUT_CHECK(ms.dcount.reserve(1));
  SCOUT_TRY(hipMemsetAsync(ms.dcount.p, 0, 8, c.s));
  k_mis_deg_claim<<<grid_for(2 * m), kBlock, 0, c.s>>>(ch.d_uv.p, m, ms.dstamp.p, ms.dgen, ms.dlist.p, ms.dcount.p);
  k_mis_degrees<<<4 * kListGrid, kBlock, 0, c.s>>>(c.g.view(), ms.dlist.p, ms.dcount.p, ms.dval.p);
  k_mis_deg_gather<<<grid_for(2 * m), kBlock, 0, c.s>>>(ch.d_uv.p, m, ms.dval.p, ms.deg.p);
  SCOUT_TRY(hipMemcpyAsync(ms.uv2.p, ch.d_uv.p, (size_t)(2 * m) * 4, hipMemcpyDeviceToDevice, c.s));
  SCOUT_TRY(hipMemcpyAsync(ms.ins2.p, ch.d_ins.p, (size_t)m, hipMemcpyDeviceToDevice, c.s));
  SCOUT_CHECK(ms.h_deg.reserve(2 * m));
  SCOUT_CHECK(ms.h_uv.reserve(2 * m));
  SCOUT_CHECK(ms.h_ins.reserve(m));
  SCOUT_CHECK(ms.h_lsb.reserve(m));
  SCOUT_TRY(hipMemcpyAsync(ms.h_deg.p, ms.deg.p, (size_t)(2 * m) * 4, hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipMemcpyAsync(ms.h_uv.p, ch.d_uv.p, (size_t)(2 * m) * 4, hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipMemcpyAsync(ms.h_ins.p, ch.d_ins.p, (size_t)m, hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  int64_t nins = 0;
  for (int64_t k = 0; k < m; k++) nins += ms.h_ins.p[k] != 0;
  st->insertions += nins;
  st->deletions += m - nins;
  for (int32_t i = 0; i < ninst; i++)
    st->sub_batches += mis_cuts(ms.h_uv.p, ms.h_deg.p, off[i], off[i + 1], max_csr, ms.uniq, ms.h_lsb.p);
  SCOUT_TRY(hipMemcpyAsync(ms.lsb.p, ms.h_lsb.p, (size_t)m * 4, hipMemcpyHostToDevice, c.s));
  SCOUT_TRY(hipMemsetAsync(ms.eps.p, 0xff, (size_t)ms.ecap * 8, c.s));
  k_mis_ep_insert<<<grid_for(2 * m), kBlock, 0, c.s>>>(ms.uv2.p, ms.lsb.p, m, ms.eps.p, ms.ecap - 1);
  SegView sv{ms.seg.p, stride};
  // segment table: one segment per instance with lines
  std::vector<int64_t> tab((size_t)seg_table_size(stride), 0);
  int64_t S = 0, slots0 = 0;
  for (int32_t i = 0; i < ninst; i++) {
    int64_t lo = off[i], hi = off[i + 1];
    if (lo == hi) continue;
    int64_t size = std::min(ms.window0, hi - lo);
    tab[(size_t)(kSegHdr + S)] = lo;
    tab[(size_t)(kSegHdr + stride + S)] = hi;
    tab[(size_t)(kSegHdr + 2 * stride + S)] = lo + size;
    tab[(size_t)(kSegHdr + 3 * stride + S)] = ms.window0;
    tab[(size_t)(kSegHdr + 4 * stride + S)] = slots0;
    slots0 += size;
    S++;
  }
  tab[(size_t)(kSegHdr + 4 * stride + S)] = slots0;
  tab[kHSegs] = S;
  tab[kHSlots] = slots0;
  tab[kHPending] = S;
  SCOUT_TRY(hipMemcpyAsync(ms.seg.p, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));  // the host tables are freed on return
  if (ms.use_graph) SCOUT_CHECK(mm_graph(c, ms, sv));
  SCOUT_CHECK(prof_begin(c, PROF_MIS_EVAL));
  while (true) {
    if (ms.use_graph) SCOUT_TRY(hipGraphLaunch(ms.exec, c.s));
    else mm_launch_group(c, ms, sv);
    SCOUT_TRY(hipMemcpyAsync(c.h_box, ms.seg.p, kSegHdr * sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
    SCOUT_TRY(hipStreamSynchronize(c.s));
    c.readbacks++;
    if (c.h_box[kHPending] == 0) break;
  }
  int64_t rounds = c.h_box[kHRounds], writes = c.h_box[kHWrites];
  int64_t evals = c.h_box[kHEvals] + slots0, entries = c.h_box[kHEntries];
  st->seeds += c.h_box[kHHop2];  // insertions whose candidate search reached the hop-2 rows
  // per line evaluation: the pair, label, row offsets of both endpoints, memberships and the
  // write slots (~48 B); per row entry read: the entry and the membership byte (5 B)
  SCOUT_CHECK(prof_end(c, PROF_MIS_EVAL, 48 * evals + 5 * entries));
  st->rounds += rounds;
  st->conflicts += writes;
  st->touched += evals;  // line evaluations of pass A over all rounds
  if (ms.complete) SCOUT_CHECK(mis_complete(c, ch, ms, st));
  return SCOUT_OK;
}

}  // namespace scout

using namespace scout;

namespace {

int mis_entry(scout_ctx* x, int channel, int32_t ninst, const int64_t* line_off, int64_t max_csr, scout_stats* st,
              const char* name) {
  if (!x || !st || !line_off || ninst < 1) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  Channel* ch = channel_of(c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  if (!c.g.boff || !c.mis) {
    set_error(std::string(name) + ": graph and MIS must be loaded first");
    return SCOUT_ERR_STATE;
  }
  if (max_csr <= 0) {
    set_error(std::string(name) + ": max_csr must be positive");
    return SCOUT_ERR_ARG;
  }
  if (ch->m >= INT32_MAX) {
    set_error(std::string(name) + ": at most 2^31 - 2 lines per batch");
    return SCOUT_ERR_ARG;
  }
  if (line_off[0] != 0 || line_off[ninst] != ch->m) {
    set_error(std::string(name) + ": line offsets must run from 0 to the staged line count");
    return SCOUT_ERR_ARG;
  }
  for (int32_t i = 0; i < ninst; i++)
    if (line_off[i + 1] < line_off[i]) {
      set_error(std::string(name) + ": line offsets must not decrease");
      return SCOUT_ERR_ARG;
    }
  TouchScope touch(c);
  Span span(c);
  SCOUT_CHECK(span.begin(c.s));
  st->lines += ch->m;
  int64_t rb0 = c.readbacks;
  SCOUT_CHECK(mis_repair_multi(c, *ch, ninst, line_off, max_csr, st));
  SCOUT_CHECK(span.end(c.s));
  st->repair_ms += span.ms();
  st->readbacks += c.readbacks - rb0;
  return channel_used(c.s, *ch);
}

}  // namespace

extern "C" int scout_repair_mis(scout_ctx* x, int channel, int64_t max_csr, scout_stats* st) {
  if (!x) return SCOUT_ERR_ARG;
  Channel* ch = channel_of(x->c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  int64_t off[2] = {0, ch->m};
  return mis_entry(x, channel, 1, off, max_csr, st, "repair_mis");
}

extern "C" int scout_repair_mis_multi(scout_ctx* x, int channel, int32_t ninst, const int64_t* line_off,
                                      int64_t max_csr, scout_stats* st) {
  return mis_entry(x, channel, ninst, line_off, max_csr, st, "repair_mis_multi");
}

extern "C" int scout_mis_mode(scout_ctx* x, int mode) {
  if (!x || (mode != SCOUT_MIS_RULES && mode != SCOUT_MIS_COMPLETE)) {
    set_error("mis_mode: unknown mode");
    return SCOUT_ERR_ARG;
  }
  Ctx& c = x->c;
  if (!c.mis_state) c.mis_state = new MisState();
  c.mis_state->complete = mode == SCOUT_MIS_COMPLETE;
  return SCOUT_OK;
}

extern "C" int scout_mis_config(scout_ctx* x, int grid, int group_rounds, int use_graph) {
  if (!x || grid < 1 || grid > 65535 || group_rounds < 1 || group_rounds > 64) {
    set_error("mis_config: need 1 <= grid <= 65535 and 1 <= group_rounds <= 64");
    return SCOUT_ERR_ARG;
  }
  Ctx& c = x->c;
  if (!c.mis_state) c.mis_state = new MisState();
  c.mis_state->grid = (unsigned)grid;
  c.mis_state->group_rounds = group_rounds;
  c.mis_state->use_graph = use_graph != 0;
  return SCOUT_OK;
}

extern "C" int scout_mis_get_config(scout_ctx* x, int* out) {
  if (!x || !out) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  if (!c.mis_state) c.mis_state = new MisState();
  out[0] = (int)c.mis_state->grid;
  out[1] = c.mis_state->group_rounds;
  out[2] = c.mis_state->use_graph ? 1 : 0;
  return SCOUT_OK;
}

// Times the staged, applied batch's MIS repair under each grid and round-group setting (the
// graph replay setting is kept) and keeps the fastest. Every trial starts from the same state:
// the membership, the member counts and what those counts were computed for are put back before
// it, and again at the end. Each trial's membership must equal the first one's.
extern "C" int scout_mis_autotune(scout_ctx* x, int channel, int32_t ninst, const int64_t* line_off, int64_t max_csr,
                                  double* best_ms) {
  if (!x) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  Channel* ch = channel_of(c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  if (!c.g.boff || !c.mis) {
    set_error("mis_autotune: graph, MIS and a staged batch are needed");
    return SCOUT_ERR_STATE;
  }
  if (!c.mis_state) c.mis_state = new MisState();
  MisState& ms = *c.mis_state;
  int32_t n = c.g.n;
  bool had_counts = ms.mcount.p && ms.mcount.cap >= (int64_t)n + 1;
  int64_t v0 = ms.mc_version, a0 = ms.mc_apply, g0 = ms.mc_gen;
  DevBuf<uint8_t> saved_mis;
  DevBuf<int32_t> saved_counts;
  SCOUT_CHECK(saved_mis.reserve((int64_t)n + 1));
  SCOUT_TRY(hipMemcpyAsync(saved_mis.p, c.mis, (size_t)n, hipMemcpyDeviceToDevice, c.s));
  if (had_counts) {
    SCOUT_CHECK(saved_counts.reserve((int64_t)n + 1));
    SCOUT_TRY(hipMemcpyAsync(saved_counts.p, ms.mcount.p, ((size_t)n + 1) * 4, hipMemcpyDeviceToDevice, c.s));
  }
  auto put_back = [&]() -> int {
    SCOUT_TRY(hipMemcpyAsync(c.mis, saved_mis.p, (size_t)n, hipMemcpyDeviceToDevice, c.s));
    if (had_counts)
      SCOUT_TRY(hipMemcpyAsync(ms.mcount.p, saved_counts.p, ((size_t)n + 1) * 4, hipMemcpyDeviceToDevice, c.s));
    ms.mc_version = v0;
    ms.mc_apply = a0;
    ms.mc_gen = g0;
    return SCOUT_OK;
  };
  const unsigned grids[] = {512, 1024, 2048, 4096};
  const int groups[] = {4, 8, 16, 32};
  unsigned keep_grid = ms.grid;
  int keep_group = ms.group_rounds;
  double best = 1e30;
  std::vector<uint8_t> first((size_t)n), got((size_t)n);
  bool have_first = false;
  int rc = SCOUT_OK;
  for (unsigned gr : grids) {
    for (int gp : groups) {
      ms.grid = gr;
      ms.group_rounds = gp;
      // the first run of a setting captures its graph; the second is timed
      for (int rep = 0; rep < 2 && rc == SCOUT_OK; rep++) {
        rc = put_back();
        scout_stats st{};
        if (rc == SCOUT_OK) rc = mis_entry(x, channel, ninst, line_off, max_csr, &st, "mis_autotune");
        if (rc != SCOUT_OK) break;
        if (rep == 1 && st.repair_ms < best) {
          best = st.repair_ms;
          keep_grid = gr;
          keep_group = gp;
        }
      }
      if (rc != SCOUT_OK) break;
      if (hipMemcpy(have_first ? got.data() : first.data(), c.mis, (size_t)n, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("mis_autotune: membership read-back failed");
        rc = SCOUT_ERR_HIP;
        break;
      }
      if (have_first && got != first) {
        set_error("mis_autotune: a launch setting changed the membership");
        rc = SCOUT_ERR_STATE;
        break;
      }
      have_first = true;
    }
    if (rc != SCOUT_OK) break;
  }
  ms.grid = keep_grid;
  ms.group_rounds = keep_group;
  int rc2 = put_back();
  if (hipStreamSynchronize(c.s) != hipSuccess && rc == SCOUT_OK) rc = SCOUT_ERR_HIP;
  saved_mis.release();
  saved_counts.release();
  if (rc != SCOUT_OK) return rc;
  if (rc2 != SCOUT_OK) return rc2;
  if (best_ms) *best_ms = best;
  return SCOUT_OK;
}

extern "C" int scout_mis_limits(scout_ctx* x, int64_t hop2_requests, int64_t hop2_tasks) {
  if (!x || hop2_requests < 1 || hop2_requests > kHop2Requests || hop2_tasks < 1 || hop2_tasks > (int64_t)1 << 30) {
    set_error("mis_limits: need 1 <= hop2_requests <= 65536 and 1 <= hop2_tasks <= 2^30");
    return SCOUT_ERR_ARG;
  }
  Ctx& c = x->c;
  if (!c.mis_state) c.mis_state = new MisState();
  c.mis_state->h2_rcap = hop2_requests;
  c.mis_state->h2_tcap = hop2_tasks;
  return SCOUT_OK;
}
