// cc.hip: connected-component labels (smallest vertex id of the component), exact after every
// batch.
//
// The reference keeps labels plus a union-find over labels: an insertion unites the two
// labels (lower wins), a deletion runs a host bidirectional BFS that expands the side with
// the smaller frontier until the sides meet (still connected), one side is exhausted (split:
// that side gets a fresh label) or the bound is passed (deferred to a recompute after the
// stream) -- src/cc/scout_cc.cu:300-441. One batch is one device job here:
//   1. insertions: lock-free union of the two labels, lower label wins, then every vertex
//      takes its label's root (labels stay component minima because minima win);
//   2. deletions: one search group per endpoint of every deleted pair, all expanding in
//      lock-step rounds over the updated graph. A group claims unclaimed vertices; touching
//      another group's vertex unites the two groups. Per component, every alive group
//      except the one with the largest frontier expands (the reference's "expand the
//      smaller side", generalised); a component is settled once at most one group in it
//      still has a frontier. A group whose frontier empties has explored a whole new
//      component. Every piece a component can split into contains an endpoint of a deleted
//      pair, so the pieces are exactly the exhausted groups plus one remainder;
//   3. relabel: exhausted groups take their smallest vertex; the remainder keeps the old
//      label unless its minimum moved into an exhausted group, in which case one pass over
//      the component finds the remainder's new minimum.
#include "common.cuh"

namespace scout {

namespace {

__device__ __forceinline__ int32_t uf_find(int32_t* p, int32_t x) {
  while (true) {
    int32_t q = __hip_atomic_load(&p[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (q == x) return x;
    int32_t r = __hip_atomic_load(&p[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (r != q) atomicCAS(&p[x], q, r);  // path halving
    x = q;
  }
}

// Unites the sets of a and b, the larger root hooked under the smaller. True if this call
// hooked (then *hooked = the root that stopped being a root).
__device__ __forceinline__ bool uf_union(int32_t* p, int32_t a, int32_t b, int32_t* hooked) {
  while (true) {
    a = uf_find(p, a);
    b = uf_find(p, b);
    if (a == b) return false;
    if (a < b) {
      int32_t t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(&p[a], a, b) == a) {
      *hooked = a;
      return true;
    }
  }
}

__global__ void k_iota(int32_t* p, int64_t count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = (int32_t)i;
}

// --- full recompute -------------------------------------------------------------------

// Wave per vertex v (rows up to kChunk; longer rows become chunk tasks): union(v, w) for
// every neighbour w > v.
__global__ void k_cc_hook(GView g, int32_t* parent, int64_t* tasks, int64_t* task_count, int64_t task_cap,
                          int64_t* scanned) {
  int64_t work = 0;
  for (int64_t v = global_wave(); v < g.n; v += total_waves()) {
    Row r = row_of(g, (int32_t)v);
    int64_t len = row_len(r);
    if (len > kChunk) {
      if (lane_id() == 0) {
        int64_t chunks = (len + kChunk - 1) / kChunk;
        int64_t pos = (int64_t)atomicAdd((unsigned long long*)task_count, (unsigned long long)chunks);
        for (int64_t k = 0; k < chunks && pos + k < task_cap; k++) tasks[pos + k] = (v << 32) | k;
      }
      continue;
    }
    work += len;
    wave_walk(g, r, 0, len, [&](int32_t w) {
      int32_t h;
      if (w > v) uf_union(parent, (int32_t)v, w, &h);
    });
  }
  if (lane_id() == 0 && work) atomicAdd((unsigned long long*)scanned, (unsigned long long)work);
}

__global__ void k_cc_hook_tasks(GView g, int32_t* parent, const int64_t* tasks, const int64_t* task_count,
                                int64_t task_cap, int64_t* scanned) {
  int64_t nt = min(*task_count, task_cap);
  int64_t work = 0;
  for (int64_t i = global_wave(); i < nt; i += total_waves()) {
    int64_t t = tasks[i];
    int32_t v = (int32_t)(t >> 32);
    int64_t chunk = t & 0xffffffffLL;
    Row r = row_of(g, v);
    int64_t lo = chunk * kChunk, hi = lo + kChunk, len = row_len(r);
    if (hi > len) hi = len;
    work += hi - lo;
    wave_walk(g, r, lo, hi, [&](int32_t w) {
      int32_t h;
      if (w > v) uf_union(parent, v, w, &h);
    });
  }
  if (lane_id() == 0 && work) atomicAdd((unsigned long long*)scanned, (unsigned long long)work);
}

__global__ void k_cc_flatten(int32_t* parent, int32_t n) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) parent[v] = uf_find(parent, (int32_t)v);
}

// --- batch repair ---------------------------------------------------------------------

// Insertion lines whose endpoints are adjacent after the batch unite their labels in the
// label forest lp. box[0] = hooks; hooked roots are listed for the reset.
__global__ void k_cc_insert(GView g, bool sorted, const int32_t* uv, const uint8_t* ins, int64_t m,
                            const int32_t* label, int32_t* lp, int32_t* hooked, int64_t cap, int64_t* box) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !ins[i]) return;
  atomicAdd((unsigned long long*)&box[1], 1ULL);
  int32_t u = uv[2 * i], v = uv[2 * i + 1];
  int32_t a = label[u], b = label[v];
  if (a == b) return;
  if (!adjacent_now(pair_state(g, sorted, u, v))) return;
  int32_t h;
  if (uf_union(lp, a, b, &h)) thread_append(h, hooked, &box[0], cap);
}

__global__ void k_cc_relabel_roots(int32_t* label, int32_t n, int32_t* lp) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int32_t l = label[v];
  if (lp[l] != l) label[v] = uf_find(lp, l);
}

__global__ void k_reset_identity(const int32_t* list, int64_t count, int32_t* p) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[list[i]] = list[i];
}

// Every deletion line whose endpoints still share a label and are no longer adjacent gives
// each endpoint a search group (id 2i, 2i+1); an endpoint already claimed joins the
// claimer's group. gsrc[gid] = the group's source vertex.
__global__ void k_cc_sources(GView g, bool sorted, const uint8_t* ins, const int32_t* uv, int64_t m,
                             const int32_t* label, int32_t* owner, int32_t* gp, int32_t* gsrc, int32_t* front,
                             int64_t* front_count, int32_t* claimed, int64_t* claimed_count, int64_t cap,
                             int64_t* box) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || ins[i]) return;
  atomicAdd((unsigned long long*)&box[2], 1ULL);
  int32_t e[2] = {uv[2 * i], uv[2 * i + 1]};
  if (label[e[0]] != label[e[1]]) return;
  if (adjacent_now(pair_state(g, sorted, e[0], e[1]))) return;
  atomicAdd((unsigned long long*)&box[3], 1ULL);
  for (int s = 0; s < 2; s++) {
    int32_t gid = (int32_t)(2 * i + s);
    int32_t prev = atomicCAS(&owner[e[s]], -1, gid);
    gsrc[gid] = e[s];
    if (prev == -1) {
      thread_append(e[s], front, front_count, cap);
      thread_append(e[s], claimed, claimed_count, cap);
    } else {
      int32_t h;
      uf_union(gp, gid, prev, &h);
    }
  }
}

// fcnt[root] += 1 for every frontier vertex.
__global__ void k_cc_count(const int32_t* front, const int64_t* nf, const int32_t* owner, int32_t* gp, int32_t* fcnt) {
  int64_t n = *nf;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&fcnt[uf_find(gp, owner[front[i]])], 1);
}

// Per component (indexed by label): number of alive groups and the group with the largest
// frontier (ties: smaller group id).
__global__ void k_cc_comp_stats(int64_t ngid, int32_t* gp, const int32_t* gsrc, const int32_t* fcnt,
                                const int32_t* label, int32_t* nalive, unsigned long long* best) {
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= ngid || gsrc[gid] < 0) return;
  if (gp[gid] != gid || fcnt[gid] == 0) return;
  int32_t comp = label[gsrc[gid]];
  atomicAdd(&nalive[comp], 1);
  unsigned long long key = ((unsigned long long)(uint32_t)fcnt[gid] << 32) | (unsigned long long)(0xffffffffu - (uint32_t)gid);
  atomicMax(&best[comp], key);
}

// Wave per frontier vertex. Settled components drop out; the largest group of an unsettled
// component waits (its vertices move to the next frontier); every other group expands.
__global__ void k_cc_expand(GView g, const int32_t* front, const int64_t* nf, int32_t* owner, int32_t* gp,
                            const int32_t* label, const int32_t* nalive, const unsigned long long* best,
                            int32_t* next, int64_t* next_count, int32_t* claimed, int64_t* claimed_count, int64_t cap,
                            int64_t* expanded, int64_t* scanned) {
  int64_t n = *nf;
  int64_t work = 0, exp_count = 0;
  for (int64_t i = global_wave(); i < n; i += total_waves()) {
    int32_t f = front[i];
    // one lane resolves the group root while other waves may be uniting groups
    int32_t r0 = 0;
    if (lane_id() == 0) r0 = uf_find(gp, owner[f]);
    int32_t r = __shfl(r0, 0);
    int32_t comp = label[f];
    if (nalive[comp] <= 1) continue;
    uint32_t lead = 0xffffffffu - (uint32_t)(best[comp] & 0xffffffffULL);
    if ((uint32_t)r == lead) {
      if (lane_id() == 0) thread_append(f, next, next_count, cap);
      continue;
    }
    exp_count++;
    Row row = row_of(g, f);
    int64_t len = row_len(row);
    work += len;
    for (int64_t base = 0; base < len; base += kWave) {
      int64_t k = base + lane_id();
      bool take = false;
      int32_t w = -1;
      if (k < len) {
        w = row_at(g, row, k);
        if (w >= 0) {
          int32_t o = __hip_atomic_load(&owner[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (o == -1) {
            o = atomicCAS(&owner[w], -1, r);
            take = (o == -1);
          }
          if (!take) {
            int32_t h;
            if (uf_find(gp, o) != uf_find(gp, r)) uf_union(gp, o, r, &h);
          }
        }
      }
      wave_append(take, w, next, next_count, cap);
      wave_append(take, w, claimed, claimed_count, cap);
    }
  }
  if (lane_id() == 0) {
    if (work) atomicAdd((unsigned long long*)scanned, (unsigned long long)work);
    if (exp_count) atomicAdd((unsigned long long*)expanded, (unsigned long long)exp_count);
  }
}

// Exhausted groups (root with no frontier) take their smallest claimed vertex as label.
__global__ void k_cc_piece_min(const int32_t* claimed, int64_t nc, const int32_t* owner, int32_t* gp,
                               const int32_t* fcnt, int32_t* gmin) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  int32_t v = claimed[i];
  int32_t r = uf_find(gp, owner[v]);
  if (fcnt[r] == 0) atomicMin(&gmin[r], v);
}

// Relabel exhausted pieces. A component whose old minimum fell into a piece needs a rescan
// for its remainder: rescan[old] = 1. box[4] counts pieces (one per exhausted root).
__global__ void k_cc_piece_relabel(const int32_t* claimed, int64_t nc, const int32_t* owner, int32_t* gp,
                                   const int32_t* fcnt, const int32_t* gmin, int32_t* label, int32_t* rescan) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  int32_t v = claimed[i];
  int32_t r = uf_find(gp, owner[v]);
  if (fcnt[r] != 0) return;
  int32_t old = label[v];
  if (v == old) rescan[old] = 1;
  label[v] = gmin[r];
}

__global__ void k_cc_count_pieces(int64_t ngid, const int32_t* gp, const int32_t* gsrc, const int32_t* fcnt,
                                  int64_t* box) {
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= ngid || gsrc[gid] < 0) return;
  if (gp[gid] == gid && fcnt[gid] == 0) atomicAdd((unsigned long long*)&box[4], 1ULL);
}

// Remainder minimum of components flagged for rescan: vertices still carrying the old label
// and not inside an exhausted piece. The piece holding the old minimum kept the old label
// value, so its vertices are excluded through their owner.
__global__ void k_cc_rescan_min(const int32_t* label, int32_t n, const int32_t* rescan, const int32_t* owner,
                                int32_t* gp, const int32_t* fcnt, int32_t* rmin) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int32_t l = label[v];
  if (!rescan[l]) return;
  int32_t o = owner[v];
  if (o >= 0 && fcnt[uf_find(gp, o)] == 0) return;
  atomicMin(&rmin[l], (int32_t)v);
}

__global__ void k_cc_rescan_apply(int32_t* label, int32_t n, const int32_t* rescan, const int32_t* owner, int32_t* gp,
                                  const int32_t* fcnt, const int32_t* rmin) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int32_t l = label[v];
  if (!rescan[l]) return;
  int32_t o = owner[v];
  if (o >= 0 && fcnt[uf_find(gp, o)] == 0) return;
  label[v] = rmin[l];
}

__global__ void k_cc_clear_comp(int64_t ngid, const int32_t* gsrc, const int32_t* label, int32_t* nalive,
                                unsigned long long* best) {
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= ngid || gsrc[gid] < 0) return;
  int32_t l = label[gsrc[gid]];
  nalive[l] = 0;
  best[l] = 0ULL;
}

__global__ void k_gather_labels(const int32_t* vs, int64_t nv, const int32_t* label, int32_t* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) out[i] = label[vs[i]];
}

__global__ void k_clear_owner(const int32_t* claimed, int64_t nc, int32_t* owner) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) owner[claimed[i]] = -1;
}

__global__ void k_clear_flags(const int32_t* vs, int64_t nv, const int32_t* labels, int32_t* rescan, int32_t* rmin) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  int32_t l = labels[i];
  rescan[l] = 0;
  rmin[l] = INT32_MAX;
}

struct CcScratch {
  DevBuf<int32_t> lp;         // label forest (identity between batches)
  DevBuf<int32_t> nalive;     // per label
  DevBuf<unsigned long long> best;  // per label
  DevBuf<int32_t> rescan;     // per label
  DevBuf<int32_t> rmin;       // per label
  DevBuf<int32_t> gp, gsrc, fcnt, gmin;  // per group id
  DevBuf<int32_t> hooked, front, next, claimed, claimed_label;
  DevBuf<int64_t> tasks;
  int32_t n = -1;
};

}  // namespace

struct CcState {
  CcScratch s;
};

static int cc_prepare(Ctx& c, CcScratch& s, int64_t m) {
  int32_t n = c.g.n;
  if (s.n != n) {
    s.n = n;
    SCOUT_CHECK(s.lp.reserve((int64_t)n + 1));
    k_iota<<<grid_for((int64_t)n + 1), kBlock, 0, c.s>>>(s.lp.p, (int64_t)n + 1);
    SCOUT_CHECK(s.nalive.reserve((int64_t)n + 1));
    SCOUT_TRY(hipMemsetAsync(s.nalive.p, 0, ((size_t)n + 1) * 4, c.s));
    SCOUT_CHECK(s.best.reserve((int64_t)n + 1));
    SCOUT_TRY(hipMemsetAsync(s.best.p, 0, ((size_t)n + 1) * 8, c.s));
    SCOUT_CHECK(s.rescan.reserve((int64_t)n + 1));
    SCOUT_TRY(hipMemsetAsync(s.rescan.p, 0, ((size_t)n + 1) * 4, c.s));
    SCOUT_CHECK(s.rmin.reserve((int64_t)n + 1));
    SCOUT_TRY(hipMemsetAsync(s.rmin.p, 0x7f, ((size_t)n + 1) * 4, c.s));
    SCOUT_CHECK(s.front.reserve((int64_t)n + 1));
    SCOUT_CHECK(s.next.reserve((int64_t)n + 1));
    SCOUT_CHECK(s.claimed.reserve((int64_t)n + 1));
    SCOUT_CHECK(s.claimed_label.reserve((int64_t)n + 1));
  }
  SCOUT_CHECK(s.tasks.reserve(2 * (c.g.E + c.g.A) / kChunk + 1024));
  int64_t ng = 2 * m + 2;
  SCOUT_CHECK(s.gp.reserve(ng));
  SCOUT_CHECK(s.gsrc.reserve(ng));
  SCOUT_CHECK(s.fcnt.reserve(ng));
  SCOUT_CHECK(s.gmin.reserve(ng));
  SCOUT_CHECK(s.hooked.reserve(m + 2));
  SCOUT_CHECK(ensure_state(c, SCOUT_STATE_LABEL));
  SCOUT_CHECK(ensure_scratch(c, &c.cc_owner, -1));
  return SCOUT_OK;
}

CcState* cc_state(Ctx& c);

int cc_full(Ctx& c, int64_t* scanned) {
  CcState* st = cc_state(c);
  if (!st) return SCOUT_ERR_HIP;
  SCOUT_CHECK(cc_prepare(c, st->s, 0));
  int32_t n = c.g.n;
  if (n == 0) return SCOUT_OK;
  k_iota<<<grid_for(n), kBlock, 0, c.s>>>(c.label, n);
  SCOUT_TRY(hipMemsetAsync(c.d_box, 0, 2 * sizeof(int64_t), c.s));
  k_cc_hook<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), c.label, st->s.tasks.p, c.d_box, st->s.tasks.cap, c.d_box + 1);
  k_cc_hook_tasks<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), c.label, st->s.tasks.p, c.d_box, st->s.tasks.cap,
                                                  c.d_box + 1);
  k_cc_flatten<<<grid_for(n), kBlock, 0, c.s>>>(c.label, n);
  SCOUT_CHECK(box_read(c, 2));
  if (c.h_box[0] > st->s.tasks.cap) {
    set_error("cc: chunk task overflow");
    return SCOUT_ERR_STATE;
  }
  if (scanned) *scanned += c.h_box[1];
  return SCOUT_OK;
}

int cc_local(Ctx& c, Channel& ch, scout_stats* st) {
  CcState* cs = cc_state(c);
  if (!cs) return SCOUT_ERR_HIP;
  CcScratch& s = cs->s;
  int64_t m = ch.m;
  SCOUT_CHECK(cc_prepare(c, s, m));
  int32_t n = c.g.n;
  int64_t cap = (int64_t)n + 1;
  int64_t* box = c.d_box;
  SCOUT_TRY(hipMemsetAsync(box, 0, 16 * sizeof(int64_t), c.s));
  // 1. insertions
  if (m > 0)
    k_cc_insert<<<grid_for(m), kBlock, 0, c.s>>>(c.g.view(), c.g.sorted, ch.d_uv.p, ch.d_ins.p, m, c.label, s.lp.p,
                                                 s.hooked.p, s.hooked.cap, box);
  SCOUT_CHECK(box_read(c, 2));
  int64_t nhook = c.h_box[0];
  st->insertions += c.h_box[1];
  if (nhook > 0) {
    k_cc_relabel_roots<<<grid_for(n), kBlock, 0, c.s>>>(c.label, n, s.lp.p);
    k_reset_identity<<<grid_for(nhook), kBlock, 0, c.s>>>(s.hooked.p, nhook, s.lp.p);
  }
  // 2. deletions
  int64_t ng = 2 * m;
  if (ng == 0) return SCOUT_OK;
  k_iota<<<grid_for(ng), kBlock, 0, c.s>>>(s.gp.p, ng);
  SCOUT_TRY(hipMemsetAsync(s.gsrc.p, 0xff, (size_t)ng * 4, c.s));
  SCOUT_TRY(hipMemsetAsync(s.gmin.p, 0x7f, (size_t)ng * 4, c.s));
  int64_t* bfront = box + 8;
  int64_t* bnext = box + 9;
  int64_t* bclaimed = box + 10;
  int64_t* bexp = box + 11;
  int64_t* bscan = box + 12;
  k_cc_sources<<<grid_for(m), kBlock, 0, c.s>>>(c.g.view(), c.g.sorted, ch.d_ins.p, ch.d_uv.p, m, c.label, c.cc_owner,
                                                s.gp.p, s.gsrc.p, s.front.p, bfront, s.claimed.p, bclaimed, cap, box);
  SCOUT_CHECK(box_read(c, 12));
  st->deletions += c.h_box[2];
  int64_t nsrc = c.h_box[8];
  int64_t rounds = 0;
  if (nsrc > 0) {
    int32_t* front = s.front.p;
    int32_t* nxt = s.next.p;
    int64_t scan_before = 0, claimed_before = c.h_box[10], front_before = nsrc;
    while (true) {
      SCOUT_TRY(hipMemsetAsync(s.fcnt.p, 0, (size_t)ng * 4, c.s));
      k_cc_count<<<kListGrid, kBlock, 0, c.s>>>(front, bfront, c.cc_owner, s.gp.p, s.fcnt.p);
      k_cc_comp_stats<<<grid_for(ng), kBlock, 0, c.s>>>(ng, s.gp.p, s.gsrc.p, s.fcnt.p, c.label, s.nalive.p, s.best.p);
      SCOUT_TRY(hipMemsetAsync(bnext, 0, sizeof(int64_t), c.s));
      SCOUT_TRY(hipMemsetAsync(bexp, 0, sizeof(int64_t), c.s));
      SCOUT_CHECK(prof_begin(c, PROF_CC_EXPAND));
      k_cc_expand<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), front, bfront, c.cc_owner, s.gp.p, c.label, s.nalive.p,
                                                  s.best.p, nxt, bnext, s.claimed.p, bclaimed, cap, bexp, bscan);
      SCOUT_CHECK(prof_end(c, PROF_CC_EXPAND, 0));
      rounds++;
      SCOUT_TRY(hipMemcpyAsync(c.h_box + 8, box + 8, 5 * sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
      SCOUT_TRY(hipStreamSynchronize(c.s));
      int64_t expanded = c.h_box[11];
      // algorithmic bytes: per frontier vertex id, owner, group root, label, component stats
      // and row offsets (64 B); per scanned entry the neighbour id and its owner (8 B); per
      // claim the owner write and the two list entries (12 B)
      prof_bytes(c, PROF_CC_EXPAND, 64 * front_before + 8 * (c.h_box[12] - scan_before) +
                                        12 * (c.h_box[10] - claimed_before));
      scan_before = c.h_box[12];
      claimed_before = c.h_box[10];
      front_before = c.h_box[9];
      // per-component stats are rebuilt each round; only the sources' components were set
      k_cc_clear_comp<<<grid_for(ng), kBlock, 0, c.s>>>(ng, s.gsrc.p, c.label, s.nalive.p, s.best.p);
      if (expanded == 0) break;
      std::swap(front, nxt);
      SCOUT_TRY(hipMemcpyAsync(bfront, bnext, sizeof(int64_t), hipMemcpyDeviceToDevice, c.s));
      if (c.h_box[10] > cap) {
        set_error("cc: claimed list overflow");
        return SCOUT_ERR_STATE;
      }
    }
    // fcnt holds the final frontier sizes: roots with none are the new pieces.
    int64_t nclaimed = c.h_box[10];
    st->touched += nclaimed;
    st->edges_scanned += c.h_box[12];
    k_cc_count_pieces<<<grid_for(ng), kBlock, 0, c.s>>>(ng, s.gp.p, s.gsrc.p, s.fcnt.p, box);
    k_cc_piece_min<<<grid_for(nclaimed), kBlock, 0, c.s>>>(s.claimed.p, nclaimed, c.cc_owner, s.gp.p, s.fcnt.p,
                                                           s.gmin.p);
    k_gather_labels<<<grid_for(nclaimed), kBlock, 0, c.s>>>(s.claimed.p, nclaimed, c.label, s.claimed_label.p);
    k_cc_piece_relabel<<<grid_for(nclaimed), kBlock, 0, c.s>>>(s.claimed.p, nclaimed, c.cc_owner, s.gp.p, s.fcnt.p,
                                                               s.gmin.p, c.label, s.rescan.p);
    k_cc_rescan_min<<<grid_for(n), kBlock, 0, c.s>>>(c.label, n, s.rescan.p, c.cc_owner, s.gp.p, s.fcnt.p, s.rmin.p);
    k_cc_rescan_apply<<<grid_for(n), kBlock, 0, c.s>>>(c.label, n, s.rescan.p, c.cc_owner, s.gp.p, s.fcnt.p, s.rmin.p);
    k_clear_flags<<<grid_for(nclaimed), kBlock, 0, c.s>>>(s.claimed.p, nclaimed, s.claimed_label.p, s.rescan.p,
                                                          s.rmin.p);
    k_clear_owner<<<grid_for(nclaimed), kBlock, 0, c.s>>>(s.claimed.p, nclaimed, c.cc_owner);
    SCOUT_CHECK(box_read(c, 5));
    st->splits += c.h_box[4];
  }
  st->rounds += rounds;
  st->seeds += nsrc;
  return SCOUT_OK;
}

CcState* cc_state(Ctx& c) {
  if (!c.cc) c.cc = new CcState();
  return c.cc;
}

void cc_release(Ctx& c) {
  if (!c.cc) return;
  CcScratch& s = c.cc->s;
  DevBuf<int32_t>* bs[] = {&s.lp, &s.nalive, &s.rescan, &s.rmin, &s.gp, &s.gsrc, &s.fcnt, &s.gmin,
                           &s.hooked, &s.front, &s.next, &s.claimed, &s.claimed_label};
  for (auto* b : bs) b->release();
  s.best.release();
  s.tasks.release();
  delete c.cc;
  c.cc = nullptr;
}

}  // namespace scout

using namespace scout;

extern "C" {

int scout_cc_full(scout_ctx* x) {
  if (!x) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  if (!c.g.boff) {
    set_error("no graph loaded");
    return SCOUT_ERR_STATE;
  }
  SCOUT_CHECK(ensure_state(c, SCOUT_STATE_LABEL));
  SCOUT_CHECK(cc_full(c, nullptr));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  return SCOUT_OK;
}

int scout_repair_cc(scout_ctx* x, int channel, int strategy, scout_stats* st) {
  if (!x || !st) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  Channel* ch = channel_of(c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  if (!c.g.boff || !c.label) {
    set_error("repair_cc: graph and labels must be loaded first");
    return SCOUT_ERR_STATE;
  }
  Span span;
  SCOUT_CHECK(span.begin(c.s));
  st->lines += ch->m;
  if (strategy == SCOUT_STRATEGY_RECOMPUTE) {
    st->fallback = 1;
    SCOUT_CHECK(cc_full(c, &st->edges_scanned));
  } else {
    SCOUT_CHECK(cc_local(c, *ch, st));
  }
  SCOUT_CHECK(span.end(c.s));
  st->repair_ms += span.ms();
  return SCOUT_OK;
}

}  // extern "C"
