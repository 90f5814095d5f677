// internal.h: shared declarations of libscout_hip (not part of the public C ABI).
//
// HBM layout of the dynamic graph (one copy per context):
//   boff[n+1] int64, bcol[E] int32   base CSR; a deleted base entry keeps its slot and is
//                                    stored as ~v (negative), so rows never move
//   ahead[n] {off int64, len int32, cap int32}, acol/aown[pool] int32
//                                    inserted edges of v in insertion order at
//                                    acol[off .. off + len) of ahead[v], inside a reserved
//                                    span of cap slots of a bump-allocated pool (the header is
//                                    one 16-byte load); aown
//                                    names the vertex a pool slot was written for. An entry is
//                                    v (live), ~v (a tombstone on a pair the base row does not
//                                    hold) or HOLE (a forgotten insertion). A row that
//                                    outgrows its span moves to fresh pool space of twice the
//                                    size, so a batch costs O(its edits), not O(n).
// The logical row of v is the live entries of its base row, then the live entries of its
// inserted row -- the order of core/FlashGraph.h:246-265.
#pragma once

#include <chrono>
#include <functional>
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../../include/scout_hip.h"

namespace scout {

constexpr int32_t HOLE = INT32_MIN;
constexpr int kBlock = 256;

#define SCOUT_TRY(expr)                                                              \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) {                                                          \
      ::scout::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));       \
      return SCOUT_ERR_HIP;                                                          \
    }                                                                                \
  } while (0)

#define SCOUT_CHECK(status)              \
  do {                                   \
    int _s = (status);                   \
    if (_s != SCOUT_OK) return _s;       \
  } while (0)

void set_error(const std::string& msg);

inline unsigned grid_for(int64_t items, int per_block = kBlock) {
  int64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 0x7fffffff) g = 0x7fffffff;
  return (unsigned)g;
}

// Device buffer with capacity growth (contents are not preserved on growth).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  int64_t cap = 0;
  int reserve(int64_t count) {
    if (count <= cap) return SCOUT_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    int64_t want = count + count / 4 + 1024;
    SCOUT_TRY(hipMalloc((void**)&p, (size_t)want * sizeof(T)));
    cap = want;
    return SCOUT_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Pinned host buffer with capacity growth.
template <typename T>
struct PinBuf {
  T* p = nullptr;
  int64_t cap = 0;
  int reserve(int64_t count) {
    if (count <= cap) return SCOUT_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    int64_t want = count + count / 4 + 1024;
    SCOUT_TRY(hipHostMalloc((void**)&p, (size_t)want * sizeof(T), hipHostMallocDefault));
    cap = want;
    return SCOUT_OK;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Header of a vertex's rows: its inserted row acol[off, off + len) and its base row
// bcol[b0, b0 + blen) (a copy of the base offsets). One 24-byte record, so a row lookup reads
// one cache line (two for the one record in eight that straddles a line boundary). The
// inserted row's reserved span capacity, read only when entries are placed, is kept apart
// (DevGraph::acap), so the headers take what the offsets and a 16-byte header took together.
struct alignas(8) AHead {
  int64_t off;
  int32_t len;
  int32_t blen;
  int64_t b0;
  __host__ __device__ int64_t b1() const { return b0 + blen; }
};

// Sets the inserted row of vertex u (start, length and span capacity), keeping its base row.
__host__ __device__ inline void set_inserted(AHead* ahead, int32_t* acap, int64_t u, int64_t off, int32_t len,
                                             int32_t cap) {
  ahead[u].off = off;
  ahead[u].len = len;
  acap[u] = cap;
}

// Read-only view of the logical graph for kernels.
struct GView {
  const int64_t* boff;
  const int32_t* bcol;
  const AHead* ahead;
  const int32_t* acol;
  int32_t n;
  const int32_t* bsamp;  // decoded bcol[kSampleStride * k], or nullptr (rows searched directly)
  uint32_t* touched;     // bitmap of vertices whose rows a repair read, or nullptr (not tracked)
  // per base entry (row u, neighbour v): the offset in row v of its one entry for u, or -1;
  // nullptr when not built (rows are searched instead)
  const int32_t* twin;
};

// Every kSampleStride-th base entry is sampled (decoded): a search in a long sorted row first
// narrows to one stride of the row through the samples, which stay cached for the hub rows
// many lines search.
constexpr int64_t kSampleStride = 64;

struct DevGraph {
  int32_t n = 0;
  int64_t E = 0;        // base entries
  int64_t A = 0;        // pool slots handed out (rows, their slack and moved-away spans)
  int64_t live = 0;     // entries inside rows (sum of the headers' len), holes included
  int64_t holes = 0;    // HOLE entries inside rows
  bool sorted = false;  // every base row ascending (by decoded id)
  int64_t base_gen = 0; // bumped whenever the base arrays are replaced or restored
  // bumped whenever the base rows' layout changes; a restore takes the snapshot's (its
  // tombstones differ, its layout is the snapshot's)
  int64_t shape_gen = 0;
  int64_t* boff = nullptr;
  int32_t* bcol = nullptr;      // device address of the base entries (HBM, or mapped host memory)
  // storage-resident mode: the base entries stay in pinned host memory that kernels read over
  // PCIe (hbcol is its host address, hoff a host copy of the offsets for streamed passes)
  bool host_bcol = false;
  int32_t* hbcol = nullptr;
  std::vector<int64_t> hoff;
  DevBuf<AHead> ahead, ahead2;
  DevBuf<int32_t> acap;  // per vertex: slots reserved for its inserted row
  DevBuf<int64_t> roff;   // re-pack scratch: new row starts
  DevBuf<int32_t> rcap;   // re-pack scratch: new row capacities
  DevBuf<int32_t> acol, acol2, aown, aown2;
  // rows of this batch's new entries that move to fresh pool space (row, span, entries)
  DevBuf<int32_t> moves, move_k;
  DevBuf<int64_t> move_pos, move_src;
  // new inserted-row entries of the current batch: key = src << 32 | line
  DevBuf<uint64_t> nkey, nkey2, mkey;
  DevBuf<int32_t> ndst, ndst2, mdst;
  // duplicate-pair detection of the current batch
  DevBuf<uint64_t> hkey;
  DevBuf<int32_t> ncnt;  // per vertex: new inserted-row entries of the batch being applied (0 between batches)
  // base entry samples and the base generation they were taken from
  DevBuf<int32_t> bsamp;
  int64_t bsamp_gen = -1;
  // mirror offsets of the base entries (GView::twin) and the layout they were built for
  DevBuf<int32_t> twin;
  int64_t twin_gen = -1;
  bool twin_symmetric = false;  // every base entry has exactly one mirror entry
  // set only while a repair runs with row tracking on (scout_touched_enable)
  uint32_t* touched = nullptr;
  GView view() const {
    return GView{boff, bcol, ahead.p, acol.p, n, (sorted && bsamp_gen == base_gen) ? bsamp.p : nullptr,
                 touched, (sorted && twin_gen == shape_gen) ? twin.p : nullptr};
  }
};

struct Channel;
// A graph update queued without its read-back (graph_apply_deferred).
struct ApplyPending {
  bool on = false;
  Channel* ch = nullptr;
  int64_t lo = 0, hi = 0;
  int conditional = 0;
  int skip_slot = -1;  // mailbox slot of the repeated-pair count, or -1
};
constexpr int kApplyGateSlot = 53;

// Per-batch edit staging: a channel holds one uploaded batch.
struct Channel {
  PinBuf<int32_t> h_uv;
  DevBuf<int32_t> d_uv;
  DevBuf<uint8_t> d_ins;    // 1 = insertion label
  DevBuf<uint8_t> d_done;   // 1 = the edit changed the graph
  // base range of row u searched by classify for line (u, v): lo, hi per line; valid for
  // lines [pos_lo, pos_hi) while the base generation is pos_gen
  DevBuf<int64_t> d_pos;
  int64_t pos_gen = -1, pos_lo = 0, pos_hi = 0;
  // the last apply covered every staged line and no pair repeated: an insertion line's pair
  // is then adjacent after the batch
  bool pairs_distinct = false;
  // a classify of lines [0, rep_hi) also counted the lines whose unordered pair an earlier line
  // holds, into the mailbox slot kChannelRepeatSlot + channel index (-1: no such count)
  int64_t rep_hi = -1;
  int64_t m = 0;
  hipEvent_t staged = nullptr;
  // the last staging ran on the compute stream, which orders the work after it by itself
  bool staged_on_compute = false;
  // recorded on the compute stream after the last work that reads this channel's lines: the
  // next upload into the channel waits for it (write after read)
  hipEvent_t consumed = nullptr;
  // the last upload: host copy into pinned staging (ms) and its copy-stream span (events)
  hipEvent_t h2d0 = nullptr, h2d1 = nullptr;
  double pack_ms = 0.0;
  int64_t h2d_bytes = 0;
};

// Marks the channel's lines as read by the work queued on the compute stream so far.
inline int channel_used(hipStream_t s, Channel& ch) {
  return hipEventRecord(ch.consumed, s) == hipSuccess ? SCOUT_OK : SCOUT_ERR_HIP;
}

// Orders stream s (the compute stream) after the channel's last staging.
inline int channel_wait(hipStream_t s, Channel& ch) {
  if (ch.staged_on_compute) return SCOUT_OK;
  return hipStreamWaitEvent(s, ch.staged, 0) == hipSuccess ? SCOUT_OK : SCOUT_ERR_HIP;
}

// Sets the staged line count; base ranges found by an earlier classify no longer apply.
inline void channel_staged(Channel& ch, int64_t m, bool on_compute = false) {
  ch.staged_on_compute = on_compute;
  ch.m = m;
  ch.pos_lo = ch.pos_hi = 0;
  ch.pairs_distinct = false;
  ch.rep_hi = -1;
}

struct MlpState;  // mlp.hip
struct CcState;   // cc.hip
struct MisState;  // mis.hip
struct ColorState;  // color.hip
struct StreamState;  // stream.hip
struct Snapshot;     // stream.hip
struct ProfState;    // prof in ctx.hip

struct Ctx {
  int dev = 0;
  hipStream_t s = nullptr;     // compute stream
  hipStream_t cs = nullptr;    // copy stream (fetch / staging)
  DevGraph g;
  Channel ch[2];
  // maintained solutions
  int32_t* level = nullptr;    // BFS levels, -1 unreached
  int32_t* parent = nullptr;   // BFS parent (a neighbour one level up), -1 for source/unreached
  int32_t* label = nullptr;    // CC min-id labels
  int32_t* color = nullptr;    // colours
  uint8_t* mis = nullptr;      // MIS membership
  int32_t source = 0;
  // n-sized scratch
  int32_t* stamp = nullptr;    // epoch stamps (dedup of frontiers)
  int32_t* aux = nullptr;      // per-problem scratch (owner / marker / minWriter)
  int32_t* aux2 = nullptr;
  int32_t epoch = 0;
  // problem-specific n-sized scratch (allocated on first use)
  int32_t* cc_owner = nullptr;   // search group of each claimed vertex (-1 = none)
  int32_t* col_marker = nullptr; // this round's proposal (-2 = none)
  int32_t* mis_minw = nullptr;   // smallest speculative writer line
  int32_t* cc_level = nullptr;   // BFS levels from vertex 0 kept for CC when BFS is not maintained
  int32_t* cc_parent = nullptr;  // parents for cc_level
  int64_t apply_seq = 0;         // batches applied to the graph
  int64_t bfs_bound = 0;          // scout_bfs_bound: invalidated vertices per deletion candidate past which a
                                 // batch's BFS repair is finished by a full recompute (0 = never)
  bool bfs_async = true;         // scout_bfs_mode: asynchronous invalidation rounds (else level-synchronous)
  bool relax_async = true;       // scout_bfs_mode: asynchronous relax rounds (else level-synchronous)
  double fetch_ms = 0;           // copy-stream time of streamed base-entry chunks (storage-resident mode)
  int64_t fetch_bytes = 0;       // bytes of those chunks
  int64_t level_seq = -1;        // apply_seq at which `level` was last made exact
  int64_t cc_level_seq = -1;     // same for cc_level
  int64_t cc_components = -1;    // component count after the last CC repair
  int64_t label_seq = -1;        // apply_seq at which `label` was last made exact
  int64_t mis_version = 0;       // bumped when `mis` is overwritten outside the MIS repair
  // small mailbox read back by the host after a round (kBoxSlots entries)
  int64_t* h_box = nullptr;    // pinned
  hipEvent_t box_ev = nullptr;  // marks the last box_read_async copy
  // a graph update whose counters have not been read yet (graph_apply_deferred); h_apply
  // (pinned) receives them, apply_ev marks the copy
  ApplyPending apply_pend;
  int64_t* h_apply = nullptr;
  hipEvent_t apply_ev = nullptr;
  // Set by scout_commit while its BFS repair runs: queues work that follows the BFS relax behind
  // each relax read-back, shut by the gate (zero once the relax has ended); *queued tells
  // whether it queued anything. cc_ahead_seq: the batch whose CC scan ran that way.
  std::function<int(const int64_t* gate, bool* queued)> after_relax;
  int64_t cc_ahead_seq = -1;
  int64_t* d_box = nullptr;
  // cumulative counts of the long-row code paths (kPathCount entries after the mailbox)
  int64_t* d_paths = nullptr;
  // generic work lists
  DevBuf<int32_t> la, lb, lc, ld, le, lf, lg, lh;
  DevBuf<int64_t> wa, wb, wc;
  DevBuf<int64_t> lvlb;        // BFS: per-level bucket offsets and cursors of candidates / seeds
  // rows read by repairs (SCOUT_DUMP_TOUCHED): one bit per vertex while tracking is on
  DevBuf<uint32_t> touched_bits;
  bool touched_on = false;
  DevBuf<uint8_t> tmp;         // library temp storage (scan/sort)
  std::vector<std::pair<hipEvent_t, hipEvent_t>> spans;  // timing
  MlpState* mlp = nullptr;
  CcState* cc = nullptr;
  MisState* mis_state = nullptr;
  ColorState* color_state = nullptr;
  StreamState* stream = nullptr;
  Snapshot* snap = nullptr;
  ProfState* prof = nullptr;
  scout_budget budget{};
  int64_t hbm_used = 0;
  int64_t pinned_used = 0;
  // last batch's kernel timings
  float last_repair_ms = 0.f;
  int64_t readbacks = 0;         // host read-backs of device counters so far
  double pred_over = 0;          // running mean of how far the predicted last level overshoots
  double pred_bias = 0;          // running mean of (invalidation rounds - predicted depth), asynchronous BFS
  double relax_bias = 0;         // running mean of (relax rounds - predicted depth), asynchronous BFS
  int64_t pred_cut = 0;          // split head: expected vertices the batch cuts off vertex 0's component
  int64_t pred_cut_seq = -1;     // the batch (apply_seq) pred_cut was predicted for
  int hint_inval = 0, hint_relax = 0;  // scout_bfs_schedule: first group sizes of the next repair
  int hint_cc = 0;                     // CC repair: walk rounds of the previous batch
  int cc_mode = SCOUT_CC_AUTO;         // CC repair: walk, scan or chosen by batch size
  DevBuf<uint32_t> cc_bits;            // CC scan mode: bitmaps of the labels the batch touched
};

constexpr int kBoxSlots = 64;
// Mailbox slots 44 and 45: each channel's repeated-pair count from its last full classify.
constexpr int kChannelRepeatSlot = 44;
// Code paths counted on the device (scout_path_counts): chunk tasks of split rows per kernel
// family, invalid rows whose children did not fit the LDS list, and vertices the asynchronous
// BFS invalidation made wait a round.
enum PathId { PATH_INVAL_TASKS = 0, PATH_INVAL_PUSH_TASKS, PATH_INVAL_KIDS_OVERFLOW, PATH_CC_TASKS,
              PATH_RELAX_TASKS, PATH_TENTATIVE_TASKS, PATH_CC_WALK_ROWS, PATH_INVAL_DEFERRED, kPathCount };

// prims.hip (library scans/sorts)
int scan_exclusive_i64(Ctx& c, const int64_t* in, int64_t* out, int64_t count);
int scan_exclusive_i32_to_i64(Ctx& c, const int32_t* in, int64_t* out, int64_t count);
// out[k] = the sum of the inserted-row lengths before row k (exclusive scan of ahead[].len).
int scan_row_lens(Ctx& c, const AHead* in, int64_t* out, int64_t count);
int sort_pairs_u64(Ctx& c, const uint64_t* kin, uint64_t* kout, const int32_t* vin, int32_t* vout, int64_t count,
                   int end_bit);
int sort_keys_u64(Ctx& c, const uint64_t* kin, uint64_t* kout, int64_t count, int begin_bit, int end_bit);
int reduce_max_i32(Ctx& c, const int32_t* in, int64_t count, int32_t* d_out);
int reduce_sum_i64(Ctx& c, const int64_t* in, int64_t count, int64_t* d_out);

// graph.hip
int graph_upload(Ctx& c, int32_t n, int64_t E, const int64_t* off, const int32_t* col);
int graph_upload_egr(Ctx& c, int32_t n, int64_t E, const void* nindex, int wide, const int32_t* nlist,
                     bool host_resident = false);
int graph_classify(Ctx& c, Channel& ch, int rule);
int graph_apply(Ctx& c, Channel& ch, int conditional);
// graph_apply (unconditional) without its read-back: the device work is queued, the counters
// follow to the host, and mailbox slot kApplyGateSlot becomes non-zero when the host must
// finish the apply (a repeated pair or more multi entries than one block groups). Kernels
// queued before apply_settle that read rows take that slot as their gate. apply_settle finishes
// the host side; *complete is false when the gated kernels did not run.
int graph_apply_deferred(Ctx& c, Channel& ch);
// Lines [lo, hi) that repeat a pair, applied as runs of distinct pairs in line order.
int apply_runs(Ctx& c, Channel& ch, int64_t lo, int64_t hi, int conditional);
int apply_settle(Ctx& c, bool* complete);
int graph_classify_lines(Ctx& c, Channel& ch, int64_t lo, int64_t hi, int rule);
int graph_apply_lines(Ctx& c, Channel& ch, int64_t lo, int64_t hi, int conditional);
int graph_repeated_pairs(Ctx& c, Channel& ch, int64_t lo, int64_t hi, int64_t* repeated);
int graph_logical_degrees(Ctx& c, int64_t* d_deg);
int graph_download(Ctx& c, int64_t* h_off, int32_t* h_col);
int graph_compact(Ctx& c);
int graph_free(Ctx& c);
// Empties the inserted rows (after a load, a generated graph or a compaction).
int reset_inserted(Ctx& c);

// per-problem modules
int bfs_full(Ctx& c, int32_t src, int64_t* rounds, int64_t* scanned);
int bfs_full_on(Ctx& c, int32_t* L, int32_t* P, int32_t src, int64_t* rounds, int64_t* scanned);
int bfs_local_on(Ctx& c, Channel& ch, int32_t* L, int32_t* P, int strategy, scout_stats* st);
// Parents for levels that came from outside (init file, upload); fails if L is not a BFS level
// array of the current graph.
int compute_parents(Ctx& c, const int32_t* L, int32_t* P);
int bfs_local(Ctx& c, Channel& ch, int strategy, scout_stats* st);
// scout_repair_bfs / scout_repair_cc on one channel: the local repair, or a full recompute when
// asked for or when the state is not exact for the graph before this batch; stats and timing
int repair_bfs(Ctx& c, Channel& ch, int strategy, scout_stats* st);
int repair_cc(Ctx& c, Channel& ch, int strategy, scout_stats* st);
int cc_full(Ctx& c, int64_t* scanned);
// Upper bound of the chunk tasks of a full CC recompute of the current graph.
int64_t cc_full_task_bound(const Ctx& c);
// Launches the full CC recompute into `label` on the compute stream with caller-provided
// chunk-task space and two counters ([0] tasks, [1] scanned entries); allocates nothing.
int cc_full_into(Ctx& c, int32_t* label, int64_t* tasks, int64_t task_cap, int64_t* counters);
// noalloc.hip: resolves the device address of the file-scope scratch array on device `dev`
// (once per device; called by scout_ctx_create with `dev` current).
int noalloc_init(int dev);
int cc_local(Ctx& c, Channel& ch, int strategy, scout_stats* st);
// The CC scan's kernels, shut while *gate is non-zero (gate may be null), and whether
// repair_cc of this batch will take that scan on the maintained levels from vertex 0.
int cc_scan_queue(Ctx& c, Channel& ch, const int32_t* L, const int64_t* gate);
bool cc_scan_ahead_ok(const Ctx& c, const Channel& ch, int strategy);
void cc_release(Ctx& c);
void mis_release(Ctx& c);
void stream_release(Ctx& c);
void snapshot_release(Ctx& c);

// mlp.hip: the learned scheduler (12 features -> 64 -> 64 -> 3 outputs, bf16 weights).
constexpr int kMlpIn = 12, kMlpHidden = 64, kMlpOut = 3;
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           kMlpHidden * kMlpOut + kMlpOut;
constexpr int64_t kMlpSample = 16384;  // candidates per batch logged for training
constexpr int64_t kMlpInfer = 512;     // candidates per batch the scheduler predicts for
bool mlp_logging(Ctx& c);
bool mlp_ready(Ctx& c);
// Features of the first min(ncand, kMlpSample) candidates (state before the repair); the
// counts are read on the device from the candidate pass's mailbox d_box ([0] candidates,
// [1] smallest level, [3] deletion lines).
int mlp_begin_batch(Ctx& c, const int32_t* cand, const int64_t* d_box, const int32_t* L, const int32_t* P,
                    int64_t sample);
// Root attribution arrays while logging (nullptr otherwise).
int32_t* mlp_root(Ctx& c);
int32_t* mlp_subtree(Ctx& c);
int32_t* mlp_maxlvl(Ctx& c);
// Predictions over the sample, aggregated on the device: d_out[0] = predicted cascade sizes
// summed over the sample, d_out[1] = the largest predicted last level.
int mlp_predict_async(Ctx& c, const int64_t* d_ncand, int64_t sample, int64_t* d_out);
// Appends the sample's features and outcomes (state after the repair) to the training store.
int mlp_end_batch(Ctx& c, const int32_t* cand, int64_t ns, const int32_t* L);
void mlp_release(Ctx& c);
void color_release(Ctx& c);

// Kernel timing for the roofline: events recorded on the compute stream around launches of a
// named kernel family while profiling is on. `bytes` is the algorithmic byte count the caller
// attributes to that launch.
int prof_begin(Ctx& c, int id);
int prof_end(Ctx& c, int id, int64_t bytes);
void prof_bytes(Ctx& c, int id, int64_t bytes);
// A span for one kernel launch: two events stamped by the dispatch itself (through
// hipExtLaunchKernelGGL), so the span is the kernel's execution and not the gap before it.
// False (no events) when profiling is off.
bool prof_kernel_events(Ctx& c, int id, hipEvent_t* start, hipEvent_t* stop);
// Launches kernel<<<grid, block, 0, c.s>>>(args...), timed as span `id` while profiling.
template <typename F, typename... Args>
int prof_launch(Ctx& c, int id, F kernel, dim3 grid, dim3 block, Args... args) {
  hipEvent_t a = nullptr, b = nullptr;
  if (prof_kernel_events(c, id, &a, &b))
    hipExtLaunchKernelGGL(kernel, grid, block, 0, c.s, a, b, 0, args...);
  else
    kernel<<<grid, block, 0, c.s>>>(args...);
  return SCOUT_OK;
}
void prof_release(Ctx& c);
enum ProfId { PROF_BFS_RELAX = 0, PROF_BFS_INVALIDATE, PROF_CC_EXPAND, PROF_APPLY, PROF_CLASSIFY,
              PROF_COLOR, PROF_MIS_EVAL, PROF_CC_HOOK, PROF_BFS_INVALIDATE_KERNEL, PROF_K_RELAX_SHORT,
              PROF_K_RELAX_ROWS, PROF_K_INVALIDATE_SHORT, PROF_K_INVAL_PUSH, PROF_K_CC_EXPLORE, PROF_K_CC_COLLECT, PROF_COUNT };

// helpers in ctx.hip
int box_read(Ctx& c, int count);  // copies d_box[0..count) into h_box, synchronises c.s
// box_read in two halves: the copy of d_box[first..first+count) is queued behind the work
// queued so far, and box_wait waits for that copy alone, so work queued after it runs on while
// the host reads the values
int box_read_async(Ctx& c, int first, int count);
int box_wait(Ctx& c);
int next_epoch(Ctx& c);
int ensure_state(Ctx& c, int kind);
// n-sized int32 scratch filled with `fill` on first allocation
int ensure_scratch(Ctx& c, int32_t** p, int32_t fill);
Channel* channel_of(Ctx& c, int ch);
int ensure_list(DevBuf<int32_t>& l, int64_t count);

// Row tracking for the duration of one repair entry point: the graph view carries the touched
// bitmap while tracking is on.
struct TouchScope {
  DevGraph& g;
  explicit TouchScope(Ctx& c) : g(c.g) { g.touched = c.touched_on ? c.touched_bits.p : nullptr; }
  ~TouchScope() { g.touched = nullptr; }
};

// Device time between two events recorded around a repair.
// Time of a repair call on the host clock, from its first launch to its end. Every repair ends
// with a read-back that waits for its device work (work queued behind that read-back for the
// next repair, scout_commit's CC scan, is not waited for); a caller timing other work
// synchronises before end(). The reference reports its stage times from host timers around
// each batch (src/bfs/scout_bfs.cu:578-589).
struct Span {
  Ctx& c;
  std::chrono::steady_clock::time_point t0;
  explicit Span(Ctx& ctx) : c(ctx) {}
  int begin(hipStream_t s);
  int end(hipStream_t s);
  double ms();
};

}  // namespace scout

// The opaque handle of the C ABI.
struct scout_ctx {
  scout::Ctx c;
};
