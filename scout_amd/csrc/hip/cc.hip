// cc.hip: connected-component labels (smallest vertex id of the component), exact after every
// batch.
//
// The reference keeps labels plus a union-find over labels: an insertion unites the two
// labels (lower wins), a deletion runs a host bidirectional BFS that expands the side with
// the smaller frontier until the sides meet (still connected), one side is exhausted (split:
// that side gets a fresh label) or the bound is passed (deferred to a recompute after the
// stream) -- src/cc/scout_cc.cu:300-441. On the device one batch is settled in two parts:
//   1. vertex 0's component: the exact BFS levels from vertex 0 after the batch (the batch
//      BFS repair of bfs.hip: invalidated subtrees either re-attach or become unreached)
//      mark it; 0 is the smallest id, so every reached vertex has label 0. A deletion that
//      splits a piece off this component shows up as that piece becoming unreached;
//   2. every other component lies wholly in the unreached subgraph (a neighbour of an
//      unreached vertex is unreached). Only the components the batch can have changed are
//      relabelled: they are exactly the components of the unreached endpoints of its lines
//      (a piece cut off vertex 0's component or an edited component holds such an endpoint;
//      one without kept all its edges and vertices, so its label stands). A walk from those
//      endpoints lists them, and a lock-free union-find over their rows relabels them,
//      hooking larger roots under smaller so every root is its set's minimum. A second walk
//      from the reached endpoints with a non-zero label gives label 0 to the vertices that
//      joined vertex 0's component. The work follows the affected components, not n.
// When the context also maintains BFS from vertex 0 and this batch's BFS repair already ran,
// its levels are used directly; otherwise the CC state keeps its own level array.
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
                                int64_t task_cap, int64_t* scanned, int64_t* paths, const int64_t* gate) {
  if (gate && *gate != 0) return;
  int64_t nt = min(*task_count, task_cap);
  int64_t work = 0, ntask = 0;
  for (int64_t i = global_wave(); i < nt; i += total_waves()) {
    ntask++;
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
  count_path(ntask, &paths[PATH_CC_TASKS]);
  if (lane_id() == 0 && work) atomicAdd((unsigned long long*)scanned, (unsigned long long)work);
}

// k_cc_hook over the vertex range [v0, v1) only (one streamed chunk of the base entries).
__global__ void k_cc_hook_range(GView g, int32_t v0, int32_t v1, int32_t* parent, int64_t* tasks, int64_t* task_count,
                                int64_t task_cap, int64_t* scanned) {
  int64_t work = 0;
  for (int64_t v = v0 + global_wave(); v < v1; v += total_waves()) {
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

// Zeroes `count` counters with a kernel (no memset node, so a captured stream replays it).
__global__ void k_cc_zero_counters(int64_t* p, int count) {
  if ((int)threadIdx.x < count) p[threadIdx.x] = 0;
}

__global__ void k_cc_flatten(int32_t* parent, int32_t n) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) parent[v] = uf_find(parent, (int32_t)v);
}

// --- batch repair ---------------------------------------------------------------------

// Wave per listed long row v: union(v, w) for every neighbour w > v; rows longer than
// kChunk become chunk tasks.
__global__ void k_cc_unreached_hook(GView g, const int32_t* list, const int64_t* count, int32_t* par, int64_t* tasks,
                                    int64_t* task_count, int64_t task_cap, int64_t* scanned, const int64_t* gate) {
  if (gate && *gate != 0) return;
  int64_t nl = *count;
  int64_t work = 0;
  for (int64_t i = global_wave(); i < nl; i += total_waves()) {
    int32_t v = list[i];
    Row r = row_of(g, v);
    int64_t len = row_len(r);
    if (len > kChunk) {
      if (lane_id() == 0) {
        int64_t chunks = (len + kChunk - 1) / kChunk;
        int64_t pos = (int64_t)atomicAdd((unsigned long long*)task_count, (unsigned long long)chunks);
        for (int64_t k = 0; k < chunks && pos + k < task_cap; k++) tasks[pos + k] = ((int64_t)v << 32) | k;
      }
      continue;
    }
    work += len;
    wave_walk(g, r, 0, len, [&](int32_t w) {
      int32_t h;
      if (w > v) uf_union(par, v, w, &h);
    });
  }
  if (lane_id() == 0 && work) atomicAdd((unsigned long long*)scanned, (unsigned long long)work);
}

// --- walk mode -----------------------------------------------------------------------

// Batch repair by exploration from the batch's endpoints (work proportional to the affected
// components, not to n):
//  * reached side: an endpoint that is reached after the batch but still has a non-zero label
//    was in a component that the batch joined to vertex 0's; it and every reached vertex with a
//    non-zero label connected to it take label 0. Every newly reached vertex v is found: on a
//    path from v to vertex 0 the first vertex with an old label of 0 is entered over an inserted
//    edge, whose other endpoint is such a seed, and the path up to it has non-zero labels;
//  * unreached side: every component the batch can have changed among the unreached vertices
//    (a piece cut off vertex 0's component, a component edited inside, merged or split) holds
//    an unreached endpoint of a batch line -- a component without one kept every edge, so it
//    is an unchanged old component. The walk from the unreached endpoints lists exactly those
//    components (closed under adjacency); a union-find over their rows relabels them.
// The label write is the visited mark of the reached side; the unreached side uses stamps.

__global__ void k_cc_zero(int64_t* p, int count) {
  if ((int)threadIdx.x < count) p[threadIdx.x] = 0;
}

// The first round covers every seed: ranges [0, count) of both lists.
__global__ void k_cc_first_ranges(int64_t* box) {
  box[6] = 0;
  box[7] = box[0];
  box[9] = 0;
  box[10] = box[8];
}

// Seeds of both sides from the batch endpoints. roots_r / roots_u += old labels that were their
// own vertex (old component roots) among the seeds of each side.
__global__ void k_cc_seed(const int32_t* uv, int64_t m, const int32_t* L, int32_t* label, int32_t* stamp, int32_t ep,
                          int32_t* par, int32_t* rlist, int64_t* rcount, int32_t* ulist, int64_t* ucount, int64_t cap,
                          int64_t* roots_r, int64_t* roots_u) {
  int64_t rr = 0, ru = 0;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < 2 * m; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = base + threadIdx.x;
    bool tr = false, tu = false;
    int32_t x = 0;
    if (i < 2 * m) {
      x = uv[i];
      if (L[x] >= 0) {
        if (label[x] != 0) {
          int32_t old = atomicExch(&label[x], 0);
          tr = old != 0;
          rr += tr && old == x;
        }
      } else if (atomicExch(&stamp[x], ep) != ep) {
        tu = true;
        par[x] = x;
        ru += label[x] == x;
      }
    }
    block_append(tr, x, rlist, rcount, cap);
    block_append(tu, x, ulist, ucount, cap);
  }
  block_add(rr, roots_r);
  block_add(ru, roots_u);
}

// Rows up to this many entries are scanned by the thread that holds the vertex; longer rows by
// the whole wave.
constexpr int kShortRow = 16;
// Blocks of a walk launch: rounds hold tens of thousands of vertices, and every block ends on
// one counter.
constexpr unsigned kExploreGrid = 512;
// SCOUT_CC_AUTO walks while the batch's endpoints number at most n / kWalkRatio and scans
// beyond. Measured on R-MAT scale 26 (n = 67M, 70% deletions): the two cost the same at about
// 100K lines per batch (200K endpoints, n / 335); 1K lines walk in 0.10 ms against the scan's
// 0.16 ms, 1M lines take 0.45 ms walking and 0.31 ms scanning.
constexpr int64_t kWalkRatio = 256;

// Unreached side: w joins the list once per batch (the stamp is the visited mark), as its own
// union-find root. ru counts old component roots among the listed vertices.
__device__ __forceinline__ bool cc_take_unreached(int32_t w, const int32_t* label, int32_t* stamp, int32_t ep,
                                                  int32_t* par, int64_t& ru) {
  if (w < 0 || stamp[w] == ep || atomicExch(&stamp[w], ep) == ep) return false;
  par[w] = w;
  ru += label[w] == w;
  return true;
}

// Reached side: a reached w with a non-zero label takes label 0 and joins the list (the label
// write is the visited mark). rr counts old component roots among them.
__device__ __forceinline__ bool cc_take_reached(int32_t w, const int32_t* L, int32_t* label, int64_t& rr) {
  if (w < 0 || L[w] < 0 || label[w] == 0) return false;
  int32_t old = atomicExch(&label[w], 0);
  rr += old != 0 && old == w;
  return old != 0;
}

// One level of both walks in one launch. box[6..7] and box[9..10] hold the unreached and
// reached lists' last-round ranges; a thread per listed vertex scans its row when it has at
// most kShortRow entries, and the wave walks the longer rows together, one at a time. The
// last block to finish (box[12] counts finished blocks) publishes the next round's ranges, so
// a round is a single launch.
__global__ void k_cc_explore(GView g, const int32_t* L, int32_t* label, int32_t* stamp, int32_t ep, int32_t* par,
                             int32_t* ulist, int32_t* rlist, int64_t* box, int64_t cap, int64_t* paths) {
  const int64_t ulo = box[6], uhi = box[7], rlo = box[9], rhi = box[10];
  const int64_t nu = uhi - ulo, total = nu + (rhi - rlo);
  int64_t work = 0, ru = 0, rr = 0;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < total; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t t = base + threadIdx.x;
    bool have = t < total;
    bool unr = t < nu;
    Row r{0, 0, 0, 0};
    int64_t len = 0;
    if (have) {
      r = row_of(g, unr ? ulist[ulo + t] : rlist[rlo + (t - nu)]);
      len = row_len(r);
    }
    work += len;
    int steps = len <= kShortRow ? (int)len : 0;
    int smax = steps;
    for (int off = kWave / 2; off > 0; off >>= 1) smax = max(smax, __shfl_xor(smax, off));
    for (int k = 0; k < smax; k++) {
      int32_t w = k < steps ? row_at(g, r, k) : -1;
      bool tu = unr && cc_take_unreached(w, label, stamp, ep, par, ru);
      bool tr = !unr && cc_take_reached(w, L, label, rr);
      wave_append(tu, w, ulist, box + 0, cap);
      wave_append(tr, w, rlist, box + 8, cap);
    }
    unsigned long long longs = __ballot(len > kShortRow);
    count_path(__popcll(longs), paths + PATH_CC_WALK_ROWS);
    while (longs) {
      int j = __ffsll((long long)longs) - 1;
      longs &= longs - 1;
      Row lr{__shfl(r.b0, j), __shfl(r.b1, j), __shfl(r.a0, j), __shfl(r.a1, j)};
      bool lunr = __shfl((int)unr, j) != 0;
      int64_t llen = row_len(lr);
      for (int64_t b0 = 0; b0 < llen; b0 += kWave) {
        int64_t k = b0 + lane_id();
        int32_t w = k < llen ? row_at(g, lr, k) : -1;
        bool tu = lunr && cc_take_unreached(w, label, stamp, ep, par, ru);
        bool tr = !lunr && cc_take_reached(w, L, label, rr);
        wave_append(tu, w, ulist, box + 0, cap);
        wave_append(tr, w, rlist, box + 8, cap);
      }
    }
  }
  block_add(work, box + 2);
  block_add(ru, box + 4);
  block_add(rr, box + 11);
  __shared__ bool last;
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd((unsigned long long*)(box + 12), 1ULL) == (unsigned long long)gridDim.x - 1;
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    int64_t uc = __hip_atomic_load(box + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int64_t rc = __hip_atomic_load(box + 8, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    box[6] = uhi;
    box[7] = uc < cap ? uc : cap;
    box[9] = rhi;
    box[10] = rc < cap ? rc : cap;
    box[12] = 0;
  }
}

// Roots of the labels (vertices labelled with their own id): the component count.
__global__ void k_cc_count_roots(const int32_t* label, int32_t n, int64_t* roots) {
  int64_t cnt = 0;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t v = base + threadIdx.x;
    cnt += v < n && label[v] == v;
  }
  block_add(cnt, roots);
}

// --- scan mode -----------------------------------------------------------------------

// Marks (bitmap over label ids) the old label of every endpoint of a batch line that is
// unreached after the batch. Label 0 (vertex 0's old component) needs no mark: its unreached
// vertices are relabelled anyway. A bit already set is not written again, so the endpoints of
// one large component do not serialise on one word.
// A coarse bitmap (one bit per kCoarse labels, small enough to stay in L2) filters the
// lookups of the collect pass, so most unreached vertices never touch the fine bitmap.
constexpr int kCoarseShift = 4;

__global__ void k_cc_mark(const int32_t* uv, int64_t m, const int32_t* L, const int32_t* label, uint32_t* bits,
                          uint32_t* coarse, const int64_t* gate) {
  if (gate && *gate != 0) return;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * m) return;
  int32_t x = uv[i];
  if (L[x] >= 0) return;
  int32_t l = label[x];
  if (l == 0) return;
  uint32_t bit = 1u << (l & 31);
  if (!(bits[l >> 5] & bit)) atomicOr(&bits[l >> 5], bit);
  int32_t cl = l >> kCoarseShift;
  uint32_t cbit = 1u << (cl & 31);
  if (!(coarse[cl >> 5] & cbit)) atomicOr(&coarse[cl >> 5], cbit);
}

// Zeroes count words, and the first nbox counters of box (one launch for both).
__global__ void k_zero_u32(uint32_t* p, int64_t count, int64_t* box, int nbox, const int64_t* gate) {
  if (gate && *gate != 0) return;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
  if (i < nbox) box[i] = 0;
}

// Groups of four vertices each thread loads per iteration of the collect pass (scale 26:
// 0.148 ms per batch with 2; 4 took 0.167 ms, with fewer waves resident).
constexpr int kCollectGroups = 2;

// Grid-stride pass over all vertices: newly reached vertices take label 0; unreached vertices
// whose old component is marked, or that were in vertex 0's component (old label 0), are
// listed for relabelling with par[v] = v. kept counts the roots of untouched components.
// Four consecutive vertices per thread (one 16-byte load of levels and of labels).
__global__ void k_cc_collect(const int32_t* L, int32_t* label, const uint32_t* bits, const uint32_t* coarse, int32_t n,
                             int32_t* par, int32_t* list, int64_t* count, int64_t cap, int64_t* kept,
                             const int64_t* gate) {
  if (gate && *gate != 0) return;
  int64_t roots = 0;
  // listed vertices are sparse and spread over the whole range: staged per wave in LDS
  constexpr int kSeg = 512;
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  WaveStage<kSeg> ws{seg + (threadIdx.x / kWave) * kSeg, 0};
  int64_t n4 = ((int64_t)n + 3) / 4;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // kCollectGroups groups of four vertices per thread and iteration: every group's loads are
  // in flight before any is used
  constexpr int H = kCollectGroups;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n4; base += H * stride) {
    int32_t ls[H][4], lab[H][4];
#pragma unroll
    for (int h = 0; h < H; h++) {
      int64_t q = base + h * stride + threadIdx.x;
#pragma unroll
      for (int k = 0; k < 4; k++) ls[h][k] = -1, lab[h][k] = 0;
      if (q < n4) {
        if (4 * q + 4 <= n) {
          int4 a = reinterpret_cast<const int4*>(L)[q], b = reinterpret_cast<const int4*>(label)[q];
          ls[h][0] = a.x, ls[h][1] = a.y, ls[h][2] = a.z, ls[h][3] = a.w;
          lab[h][0] = b.x, lab[h][1] = b.y, lab[h][2] = b.z, lab[h][3] = b.w;
        } else {
          for (int k = 0; k < 4 && 4 * q + k < n; k++) ls[h][k] = L[4 * q + k], lab[h][k] = label[4 * q + k];
        }
      }
    }
    // the mark reads of an unreached vertex's old component, in two steps whose reads are all
    // in flight together: the coarse bitmap for every such vertex, then the fine bitmap where
    // the coarse bit is set (vertices past n read as reached)
    uint32_t cw[H][4], fw[H][4];
#pragma unroll
    for (int h = 0; h < H; h++)
#pragma unroll
      for (int k = 0; k < 4; k++) {
        int32_t l = lab[h][k];
        cw[h][k] = (ls[h][k] < 0 && l != 0) ? coarse[l >> (5 + kCoarseShift)] : 0u;
      }
#pragma unroll
    for (int h = 0; h < H; h++)
#pragma unroll
      for (int k = 0; k < 4; k++) {
        int32_t l = lab[h][k];
        fw[h][k] = ((cw[h][k] >> ((l >> kCoarseShift) & 31)) & 1u) ? bits[l >> 5] : 0u;
      }
#pragma unroll
    for (int h = 0; h < H; h++) {
      int64_t q = base + h * stride + threadIdx.x;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        int64_t v = 4 * q + k;
        int32_t l = lab[h][k];
        bool take = false;
        if (ls[h][k] >= 0) {
          if (l != 0) label[v] = 0;
        } else if (q < n4 && v < n) {
          if (l == 0 || ((fw[h][k] >> (l & 31)) & 1u)) {
            take = true;
            par[v] = (int32_t)v;
          } else {
            roots += l == v;
          }
        }
        stage_push(ws, take, (int32_t)v, list, count, cap);
      }
    }
  }
  stage_finish(ws, list, count, cap);
  block_add(roots, kept);
}

// Thread per listed vertex: union with neighbours w > v for rows up to 32 entries; longer
// rows are listed for k_cc_unreached_hook.
__global__ void k_cc_union_list(GView g, const int32_t* list, const int64_t* count, int32_t* par, int32_t* longs,
                                int64_t* nlong, int64_t cap, int64_t* scanned, const int64_t* gate) {
  if (gate && *gate != 0) return;
  int64_t nl = *count;
  int64_t work = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nl; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t v = list[i];
    Row r = row_of(g, v);
    int64_t len = row_len(r);
    if (len > 32) {
      thread_append(v, longs, nlong, cap);
      continue;
    }
    work += len;
    for (int64_t k = 0; k < len; k++) {
      int32_t w = row_at(g, r, k);
      int32_t h;
      if (w > v) uf_union(par, v, w, &h);
    }
  }
  block_add(work, scanned);
}

// Thread per listed vertex: its label is its set's root (the component minimum).
__global__ void k_cc_final_list(const int32_t* list, const int64_t* count, int32_t* par, int32_t* label,
                                int64_t* roots, const int64_t* gate) {
  if (gate && *gate != 0) return;
  int64_t nl = *count;
  int64_t cnt = 0;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nl; base += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = base + threadIdx.x;
    if (i < nl) {
      int32_t v = list[i];
      int32_t r = uf_find(par, v);
      label[v] = r;
      cnt += r == v;
    }
  }
  block_add(cnt, roots);
}

}  // namespace

int64_t cc_full_task_bound(const Ctx& c) { return 2 * (c.g.E + c.g.A) / kChunk + 1024; }

int cc_full_into(Ctx& c, int32_t* label, int64_t* tasks, int64_t task_cap, int64_t* counters) {
  int32_t n = c.g.n;
  k_iota<<<grid_for(n), kBlock, 0, c.s>>>(label, n);
  k_cc_zero_counters<<<1, 64, 0, c.s>>>(counters, 2);
  k_cc_hook<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), label, tasks, counters, task_cap, counters + 1);
  k_cc_hook_tasks<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), label, tasks, counters, task_cap, counters + 1,
                                                  c.d_paths, nullptr);
  k_cc_flatten<<<grid_for(n), kBlock, 0, c.s>>>(label, n);
  SCOUT_TRY(hipGetLastError());
  return SCOUT_OK;
}

// The full recompute of a storage-resident graph: the base entries are copied from pinned host
// memory in chunks of whole rows through two device buffers on the copy stream, the copy of
// chunk k+1 overlapping the hooking of chunk k on the compute stream (the chunked, two-stream
// upload of full_recompute_cc_gpu, src/cc/scout_cc.cu:157-215). The inserted rows are in HBM.
int cc_full_streamed(Ctx& c, int64_t* scanned) {
  DevGraph& g = c.g;
  int32_t n = g.n;
  const std::vector<int64_t>& off = g.hoff;
  // chunks of whole rows, at most kStreamEntries entries unless one row is longer
  constexpr int64_t kStreamEntries = int64_t(1) << 25;  // 128 MiB of entries per buffer
  std::vector<std::pair<int32_t, int32_t>> chunks;
  int64_t longest = 0;
  for (int32_t v0 = 0; v0 < n;) {
    int32_t v1 = v0 + 1;
    while (v1 < n && off[(size_t)v1 + 1] - off[(size_t)v0] <= kStreamEntries) v1++;
    chunks.emplace_back(v0, v1);
    longest = std::max(longest, off[(size_t)v1] - off[(size_t)v0]);
    v0 = v1;
  }
  int64_t cap_entries = std::max<int64_t>(longest, 1);
  int64_t task_cap = 2 * cap_entries / kChunk + 1024;
  DevBuf<int32_t> buf[2];
  DevBuf<int64_t> tasks;
  SCOUT_CHECK(buf[0].reserve(cap_entries));
  SCOUT_CHECK(buf[1].reserve(cap_entries));
  SCOUT_CHECK(tasks.reserve(task_cap));
  size_t nc = chunks.size();
  std::vector<hipEvent_t> ev(4 * nc + 4, nullptr);
  for (auto& e : ev) SCOUT_TRY(hipEventCreate(&e));
  // ev[4k] / ev[4k+1]: copy start / end of chunk k; ev[4k+2]: chunk k hooked
  k_iota<<<grid_for(n), kBlock, 0, c.s>>>(c.label, n);
  k_cc_zero_counters<<<1, 64, 0, c.s>>>(c.d_box, 2);
  SCOUT_TRY(hipEventRecord(ev[4 * nc], c.s));
  SCOUT_TRY(hipStreamWaitEvent(c.cs, ev[4 * nc], 0));  // the labels and counters are set first
  int64_t bytes = 0;
  for (size_t k = 0; k < nc; k++) {
    int32_t v0 = chunks[k].first, v1 = chunks[k].second;
    int b = (int)(k & 1);
    int64_t e0 = off[(size_t)v0], len = off[(size_t)v1] - e0;
    if (k >= 2) SCOUT_TRY(hipStreamWaitEvent(c.cs, ev[4 * (k - 2) + 2], 0));  // buffer b is free again
    SCOUT_TRY(hipEventRecord(ev[4 * k], c.cs));
    if (len > 0)
      SCOUT_TRY(hipMemcpyAsync(buf[b].p, g.hbcol + e0, (size_t)len * 4, hipMemcpyHostToDevice, c.cs));
    SCOUT_TRY(hipEventRecord(ev[4 * k + 1], c.cs));
    bytes += len * 4;
    SCOUT_TRY(hipStreamWaitEvent(c.s, ev[4 * k + 1], 0));
    // the chunk's rows read their base entries from the buffer: entry e of the graph sits at
    // buf[e - e0]
    GView gv = g.view();
    gv.bcol = reinterpret_cast<const int32_t*>(reinterpret_cast<uintptr_t>(buf[b].p) - (uintptr_t)e0 * 4);
    gv.bsamp = nullptr;
    k_cc_zero_counters<<<1, 64, 0, c.s>>>(c.d_box, 1);
    k_cc_hook_range<<<kListGrid, kBlock, 0, c.s>>>(gv, v0, v1, c.label, tasks.p, c.d_box, task_cap, c.d_box + 1);
    k_cc_hook_tasks<<<kListGrid, kBlock, 0, c.s>>>(gv, c.label, tasks.p, c.d_box, task_cap, c.d_box + 1, c.d_paths,
                                                    nullptr);
    SCOUT_TRY(hipEventRecord(ev[4 * k + 2], c.s));
  }
  k_cc_flatten<<<grid_for(n), kBlock, 0, c.s>>>(c.label, n);
  SCOUT_CHECK(box_read(c, 2));
  double ms = 0;
  for (size_t k = 0; k < nc; k++) {
    float t = 0.f;
    SCOUT_TRY(hipEventElapsedTime(&t, ev[4 * k], ev[4 * k + 1]));
    ms += t;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  c.fetch_ms += ms;
  c.fetch_bytes += bytes;
  if (scanned) *scanned += c.h_box[1];
  return SCOUT_OK;
}

int cc_full(Ctx& c, int64_t* scanned) {
  int32_t n = c.g.n;
  if (n == 0) return SCOUT_OK;
  if (c.g.host_bcol) {
    SCOUT_CHECK(cc_full_streamed(c, scanned));
    c.label_seq = c.apply_seq;
    c.cc_components = -1;
    return SCOUT_OK;
  }
  SCOUT_CHECK(c.wb.reserve(cc_full_task_bound(c)));
  SCOUT_CHECK(prof_begin(c, PROF_CC_HOOK));
  SCOUT_CHECK(cc_full_into(c, c.label, c.wb.p, c.wb.cap, c.d_box));
  SCOUT_CHECK(prof_end(c, PROF_CC_HOOK, 0));
  SCOUT_CHECK(box_read(c, 2));
  if (c.h_box[0] > c.wb.cap) {
    set_error("cc: chunk task overflow");
    return SCOUT_ERR_STATE;
  }
  // per entry: the neighbour id and the two roots' parents (12 B); per vertex its offsets,
  // parent and final label (40 B)
  prof_bytes(c, PROF_CC_HOOK, 12 * c.h_box[1] + 40 * (int64_t)n);
  if (scanned) *scanned += c.h_box[1];
  c.label_seq = c.apply_seq;
  c.cc_components = -1;
  return SCOUT_OK;
}

// Walk mode: lists the affected vertices by walking from the batch endpoints (work follows the
// affected components).
int cc_walk(Ctx& c, Channel& ch, const int32_t* L, scout_stats* st) {
  int32_t n = c.g.n;
  SCOUT_CHECK(ensure_scratch(c, &c.cc_owner, -1));
  int64_t cap = (int64_t)n + 1;
  SCOUT_CHECK(c.la.reserve(cap));
  SCOUT_CHECK(c.lb.reserve(cap));
  SCOUT_CHECK(c.lc.reserve(cap));
  SCOUT_CHECK(c.wb.reserve(2 * (c.g.E + c.g.A) / kChunk + 1024));
  int32_t* par = c.cc_owner;
  int64_t* box = c.d_box;
  // box: [0] unreached list, [1] chunk tasks, [2] scanned entries, [3] new roots among the
  // unreached list, [4] old roots among it, [5] long rows, [6..7] its round range,
  // [8] reached list, [9..10] its round range, [11] old roots among it, [12] finished blocks
  k_cc_zero<<<1, 64, 0, c.s>>>(box, 13);
  int32_t ep = next_epoch(c);
  SCOUT_CHECK(prof_begin(c, PROF_CC_EXPAND));
  if (ch.m > 0)
    k_cc_seed<<<grid_for(2 * ch.m, kBlock * 4), kBlock, 0, c.s>>>(ch.d_uv.p, ch.m, L, c.label, c.stamp, ep, par,
                                                                  c.lc.p, box + 8, c.la.p, box, cap, box + 11, box + 4);
  k_cc_first_ranges<<<1, 1, 0, c.s>>>(box);
  // both walks advance one level per launch, queued in groups between read-backs; the first
  // group is the previous batch's round count
  int group = std::max(2, std::min(c.hint_cc, 32));
  int64_t rounds = 0;
  while (true) {
    for (int k = 0; k < group; k++) {
      SCOUT_CHECK(prof_launch(c, PROF_K_CC_EXPLORE, k_cc_explore, dim3(kExploreGrid), dim3(kBlock), c.g.view(), L,
                              c.label, c.stamp, ep, par, c.la.p, c.lc.p, box, cap, c.d_paths));
    }
    rounds += group;
    SCOUT_CHECK(box_read(c, 13));
    if (c.h_box[6] == c.h_box[7] && c.h_box[9] == c.h_box[10]) break;
    group = std::min(2 * group, 16);
  }
  c.hint_cc = (int)rounds;
  // union-find over the unreached components' rows, then their roots as labels
  k_cc_union_list<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), c.la.p, box, par, c.lb.p, box + 5, cap, box + 2,
                                                  nullptr);
  k_cc_unreached_hook<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), c.lb.p, box + 5, par, c.wb.p, box + 1, c.wb.cap,
                                                      box + 2, nullptr);
  k_cc_hook_tasks<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), par, c.wb.p, box + 1, c.wb.cap, box + 2, c.d_paths,
                                                  nullptr);
  k_cc_final_list<<<kListGrid, kBlock, 0, c.s>>>(c.la.p, box, par, c.label, box + 3, nullptr);
  SCOUT_CHECK(prof_end(c, PROF_CC_EXPAND, 0));
  SCOUT_CHECK(box_read(c, 12));
  if (c.h_box[1] > c.wb.cap) {
    set_error("cc: chunk task overflow");
    return SCOUT_ERR_STATE;
  }
  int64_t nu = c.h_box[0], nr = c.h_box[8], scanned_e = c.h_box[2];
  // per listed vertex of either side its entry, row offsets and label or parent (32 B); per
  // scanned entry the neighbour and its level, label, stamp or parents (12 B); per batch line
  // endpoint its level, label and stamp (24 B per line)
  prof_bytes(c, PROF_CC_EXPAND, 32 * (nu + nr) + 12 * scanned_e + 24 * ch.m);
  // the walks alone: per listed vertex its entry and row offsets (32 B) and its take (stamp or
  // label, parent and list slot: 12 B); per row entry the neighbour and its stamp or level (8 B)
  prof_bytes(c, PROF_K_CC_EXPLORE, 44 * (nu + nr) + 8 * scanned_e);
  // components: old roots that joined vertex 0's or were relabelled go, new roots come
  if (c.cc_components >= 0) {
    int64_t comps = c.cc_components - c.h_box[11] - c.h_box[4] + c.h_box[3];
    if (comps > c.cc_components) st->splits += comps - c.cc_components;
    c.cc_components = comps;
  } else {
    k_cc_zero<<<1, 64, 0, c.s>>>(box, 1);
    k_cc_count_roots<<<kListGrid, kBlock, 0, c.s>>>(c.label, n, box);
    SCOUT_CHECK(box_read(c, 1));
    c.cc_components = c.h_box[0];
  }
  st->rounds += rounds;
  st->edges_scanned += scanned_e;
  st->touched += nu + nr;
  c.label_seq = c.apply_seq;
  return SCOUT_OK;
}

// Scan mode: one pass over every vertex lists the affected ones (cheaper than the walk's
// dependent rounds when the batch touches a large share of the graph).
// Queues the scan listing and the union-find (every kernel shut while *gate is non-zero, when
// a gate is given: the scan queued ahead of the BFS relax's last read-back).
int cc_scan_queue(Ctx& c, Channel& ch, const int32_t* L, const int64_t* gate) {
  int32_t n = c.g.n;
  SCOUT_CHECK(ensure_scratch(c, &c.cc_owner, -1));
  SCOUT_CHECK(c.la.reserve((int64_t)n + 1));
  SCOUT_CHECK(c.lb.reserve((int64_t)n + 1));
  SCOUT_CHECK(c.wb.reserve(2 * (c.g.E + c.g.A) / kChunk + 1024));
  int64_t words = ((int64_t)n + 31) / 32, cwords = (((int64_t)n >> kCoarseShift) + 32) / 32;
  SCOUT_CHECK(c.cc_bits.reserve(words + cwords + 1));
  uint32_t* bits = c.cc_bits.p;
  uint32_t* coarse = c.cc_bits.p + words;
  int32_t* par = c.cc_owner;
  int64_t* box = c.d_box;
  // box: [0] listed vertices, [1] chunk tasks, [2] scanned entries, [3] roots among the listed,
  // [4] roots of untouched components, [5] long rows (zeroed with the bitmaps)
  SCOUT_CHECK(prof_begin(c, PROF_CC_EXPAND));
  k_zero_u32<<<grid_for(std::max<int64_t>(words + cwords, 6)), kBlock, 0, c.s>>>(bits, words + cwords, box, 6, gate);
  if (ch.m > 0)
    k_cc_mark<<<grid_for(2 * ch.m), kBlock, 0, c.s>>>(ch.d_uv.p, ch.m, L, c.label, bits, coarse, gate);
  SCOUT_CHECK(prof_launch(c, PROF_K_CC_COLLECT, k_cc_collect, dim3(kListGrid), dim3(kBlock), L, c.label, bits, coarse,
                          n, par, c.la.p, box, (int64_t)n + 1, box + 4, gate));
  k_cc_union_list<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), c.la.p, box, par, c.lb.p, box + 5, (int64_t)n + 1,
                                                  box + 2, gate);
  k_cc_unreached_hook<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), c.lb.p, box + 5, par, c.wb.p, box + 1, c.wb.cap,
                                                      box + 2, gate);
  k_cc_hook_tasks<<<kListGrid, kBlock, 0, c.s>>>(c.g.view(), par, c.wb.p, box + 1, c.wb.cap, box + 2, c.d_paths,
                                                  gate);
  k_cc_final_list<<<kListGrid, kBlock, 0, c.s>>>(c.la.p, box, par, c.label, box + 3, gate);
  SCOUT_CHECK(prof_end(c, PROF_CC_EXPAND, 0));
  return SCOUT_OK;
}

// Reads back the queued scan's counters: overflow check, component count and statistics.
int cc_scan_finish(Ctx& c, Channel& ch, scout_stats* st) {
  int32_t n = c.g.n;
  SCOUT_CHECK(box_read(c, 6));
  if (c.h_box[1] > c.wb.cap) {
    set_error("cc: chunk task overflow");
    return SCOUT_ERR_STATE;
  }
  // per vertex its level and old label (8 B); per listed vertex its list entry, parent and
  // label (12 B); per scanned entry the neighbour and two parents (12 B); per batch line
  // endpoint its level, label and mark (12 B)
  prof_bytes(c, PROF_CC_EXPAND, 8 * (int64_t)n + 12 * c.h_box[0] + 12 * c.h_box[2] + 24 * ch.m);
  // k_cc_collect: every vertex's level and label (8 B), per listed vertex its parent and list
  // entry (8 B)
  prof_bytes(c, PROF_K_CC_COLLECT, 8 * (int64_t)n + 8 * c.h_box[0]);
  int64_t comps = 1 + c.h_box[4] + c.h_box[3];  // vertex 0's, untouched ones, relabelled ones
  if (c.cc_components >= 0 && comps > c.cc_components) st->splits += comps - c.cc_components;
  c.cc_components = comps;
  st->edges_scanned += c.h_box[2];
  st->touched += c.h_box[0];
  c.label_seq = c.apply_seq;
  return SCOUT_OK;
}

int cc_scan(Ctx& c, Channel& ch, const int32_t* L, scout_stats* st) {
  SCOUT_CHECK(cc_scan_queue(c, ch, L, nullptr));
  return cc_scan_finish(c, ch, st);
}

// The listing a local repair of this batch takes: the walk lists the endpoints' components;
// under the learned strategy the split head adds the vertices it expects the batch to cut off
// vertex 0's component (the walk must list them all, the scan's cost does not depend on them).
bool cc_walks(const Ctx& c, const Channel& ch, int strategy, bool levels_maintained) {
  int64_t est = 2 * ch.m;
  if (strategy == SCOUT_STRATEGY_LEARNED && c.pred_cut_seq == c.apply_seq && levels_maintained) est += c.pred_cut;
  return c.cc_mode == SCOUT_CC_WALK || (c.cc_mode == SCOUT_CC_AUTO && est * kWalkRatio <= (int64_t)c.g.n);
}

// True when repair_cc of this batch will scan with the maintained BFS levels from vertex 0
// (the local repair on exact labels): the scan may then be queued behind the BFS relax.
bool cc_scan_ahead_ok(const Ctx& c, const Channel& ch, int strategy) {
  return c.label && c.level && c.source == 0 && c.g.n > 0 && strategy != SCOUT_STRATEGY_RECOMPUTE &&
         c.label_seq == c.apply_seq - 1 && !cc_walks(c, ch, strategy, true);
}

// Exact labels after the batch: the BFS levels from vertex 0 decide vertex 0's component;
// the rest is a union-find over the unreached subgraph.
int cc_local(Ctx& c, Channel& ch, int strategy, scout_stats* st) {
  int32_t n = c.g.n;
  if (n == 0) return SCOUT_OK;
  if (c.cc_ahead_seq == c.apply_seq) {
    // the scan ran behind the BFS relax of this batch (scout_commit)
    c.cc_ahead_seq = -1;
    st->cc_mode = SCOUT_CC_SCAN;
    return cc_scan_finish(c, ch, st);
  }
  const int32_t* L = nullptr;
  if (c.level && c.source == 0 && c.level_seq == c.apply_seq) {
    L = c.level;  // this batch's BFS repair from vertex 0 already ran on this context
  } else {
    SCOUT_CHECK(ensure_scratch(c, &c.cc_level, -1));
    SCOUT_CHECK(ensure_scratch(c, &c.cc_parent, -1));
    scout_stats bst{};
    if (c.cc_level_seq >= 0 && c.cc_level_seq == c.apply_seq - 1) {
      SCOUT_CHECK(bfs_local_on(c, ch, c.cc_level, c.cc_parent, SCOUT_STRATEGY_LOCAL, &bst));
    } else {
      SCOUT_CHECK(bfs_full_on(c, c.cc_level, c.cc_parent, 0, &bst.rounds, &bst.edges_scanned));
    }
    c.cc_level_seq = c.apply_seq;
    st->rounds += bst.rounds;
    st->edges_scanned += bst.edges_scanned;
    st->seeds += bst.seeds;
    st->touched += bst.touched;
    st->insertions += bst.insertions;
    st->deletions += bst.deletions;
    st->fallback += bst.fallback;
    L = c.cc_level;
  }
  bool walk = cc_walks(c, ch, strategy, L == c.level);
  st->cc_mode = walk ? SCOUT_CC_WALK : SCOUT_CC_SCAN;
  return walk ? cc_walk(c, ch, L, st) : cc_scan(c, ch, L, st);
}

int repair_cc(Ctx& c, Channel& ch, int strategy, scout_stats* st) {
  if (!c.g.boff || !c.label) {
    set_error("repair_cc: graph and labels must be loaded first");
    return SCOUT_ERR_STATE;
  }
  TouchScope touch(c);
  Span span(c);
  SCOUT_CHECK(span.begin(c.s));
  st->lines += ch.m;
  int64_t rb0 = c.readbacks;
  // relabelling only the touched components needs labels exact for the graph before this batch
  bool stale = c.label_seq != c.apply_seq - 1;
  if (strategy == SCOUT_STRATEGY_RECOMPUTE || stale) {
    st->fallback = 1;
    SCOUT_CHECK(cc_full(c, &st->edges_scanned));
  } else {
    SCOUT_CHECK(cc_local(c, ch, strategy, st));
  }
  SCOUT_CHECK(span.end(c.s));
  st->repair_ms += span.ms();
  st->readbacks += c.readbacks - rb0;
  return SCOUT_OK;
}

void cc_release(Ctx& c) { (void)c; }

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

int scout_cc_mode(scout_ctx* x, int mode) {
  if (!x) return SCOUT_ERR_ARG;
  if (mode != SCOUT_CC_AUTO && mode != SCOUT_CC_WALK && mode != SCOUT_CC_SCAN) {
    set_error("cc mode must be SCOUT_CC_AUTO, SCOUT_CC_WALK or SCOUT_CC_SCAN");
    return SCOUT_ERR_ARG;
  }
  x->c.cc_mode = mode;
  return SCOUT_OK;
}

int scout_repair_cc(scout_ctx* x, int channel, int strategy, scout_stats* st) {
  if (!x || !st) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  Channel* ch = channel_of(c, channel);
  if (!ch) return SCOUT_ERR_ARG;
  SCOUT_CHECK(repair_cc(c, *ch, strategy, st));
  return channel_used(c.s, *ch);
}

}  // extern "C"
