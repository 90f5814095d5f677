/* scout_hip.h -- C ABI of libscout_hip.so, the MI355X repair engine behind the scout_*
 * binaries.
 *
 * Plain pointers and sizes only. Every function returns SCOUT_OK (0) or an error code;
 * scout_last_error() describes the last failure of the calling thread. The library never
 * exits the process. Host pointers are caller-owned; device state and pinned staging are
 * owned by the context until scout_ctx_destroy.
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * reference repository). One host thread drives a context; calls on one context must not
 * overlap.
 */
#ifndef SCOUT_HIP_H
#define SCOUT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SCOUT_OK = 0,
  SCOUT_ERR_ARG = 1,         /* bad argument or size */
  SCOUT_ERR_HIP = 2,         /* a HIP runtime call failed */
  SCOUT_ERR_STATE = 3,       /* call out of order (e.g. repair before load) */
  SCOUT_ERR_BUDGET = 4,      /* the request exceeds the context's HBM or pinned budget */
  SCOUT_ERR_UNSUPPORTED = 5  /* input outside what the engine handles exactly */
};

/* Maintained solutions (scout_state_upload / scout_state_download). */
enum {
  SCOUT_STATE_MIS = 0,    /* uint8 per vertex, 1 = member */
  SCOUT_STATE_COLOR = 1,  /* int32 per vertex */
  SCOUT_STATE_LEVEL = 2,  /* int32 per vertex, -1 = unreached */
  SCOUT_STATE_LABEL = 3   /* int32 per vertex: smallest vertex id of its component */
};

/* Classification rules for scout_classify. */
enum {
  SCOUT_CLASSIFY_CURRENT = 0, /* insert iff absent from the current graph (coloring/bfs/cc) */
  SCOUT_CLASSIFY_BASE = 1     /* insert iff absent from the loaded base graph (mis: every
                                 batch is parsed against the initial graph) */
};

/* Repair strategies for scout_repair_bfs / scout_repair_cc. Every strategy returns the
 * same exact answer; they differ only in work. */
enum {
  SCOUT_STRATEGY_LOCAL = 0,     /* local repair of the affected region only */
  SCOUT_STRATEGY_RECOMPUTE = 1, /* recompute the whole solution on the device */
  SCOUT_STRATEGY_LEARNED = 2    /* let the trained scheduler choose per batch */
};

typedef struct scout_ctx scout_ctx;

typedef struct {
  int64_t hbm_bytes;    /* 0 = no limit beyond the device */
  int64_t pinned_bytes; /* 0 = no limit */
} scout_budget;

/* Per-batch statistics (fields a repair does not use stay 0). */
typedef struct {
  int64_t lines;        /* batch lines handed to the repair */
  int64_t insertions;   /* lines labelled insertion */
  int64_t deletions;    /* lines labelled deletion */
  int64_t noop;         /* lines that needed no repair work */
  int64_t seeds;        /* vertices the repair started from; mis: searches reaching hop-2 rows */
  int64_t touched;      /* vertices whose rows were scanned */
  int64_t rounds;       /* device rounds */
  int64_t fallback;     /* 1 if the batch was settled by a full recompute */
  int64_t splits;       /* CC: new components created by deletions */
  int64_t conflicts;    /* coloring: insertions with equal colours; mis: edits that wrote */
  int64_t sub_batches;  /* mis: sub-batches */
  int64_t edges_scanned;/* adjacency entries read by the repair kernels */
  double repair_ms;     /* time of the repair call on the host clock, until its device work is done */
  double apply_ms;      /* device time of classification + graph update */
  int64_t readbacks;    /* host read-backs of device counters (each waits for the stream) */
  int64_t inval_rounds; /* bfs: invalidation levels with work */
  int64_t relax_rounds; /* bfs: relax levels with work */
  int64_t completed;    /* mis complete mode: vertices the completion added */
  int64_t cc_mode;      /* cc: SCOUT_CC_WALK or SCOUT_CC_SCAN, how the local repair listed the
                           vertices to relabel (0 after a full recompute) */
} scout_stats;

const char* scout_last_error(void);

/* Device set-up, pinned and device pools. Replaces gpu_warmup_and_device_setup
 * (core/cli_setup.h:54-58) and the alloc_pair / alloc_and_upload pools
 * (core/gpu_common.h:43-72). budget may be NULL. */
int scout_ctx_create(int gpu_id, const scout_budget* budget, scout_ctx** out);
int scout_ctx_destroy(scout_ctx* ctx);
int scout_device_name(scout_ctx* ctx, char* buf, int len);
/* Bytes of device memory free / total (the reference's cudaMemGetInfo uses,
 * src/mis/scout_mis.cu:51-72,808-814). */
int scout_mem_info(scout_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);

/* Loads the base graph into HBM, streaming the (mmap'd) .egr arrays through pinned staging
 * on the copy stream. nindex is int32 (wide = 0) or uint64 (wide = 1), n+1 entries.
 * Replaces the per-round row fetch HybridGraph::neighbors_into (core/FlashGraph.h:376-385)
 * and the chunked nlist upload of full_recompute_cc_gpu (src/cc/scout_cc.cu:157-215).
 * Storage-resident mode: when the context's HBM budget cannot hold the base entries
 * (4 B each), they are copied by OpenMP threads into pinned host memory that the kernels read
 * over PCIe, and only the offsets, the inserted rows and the per-vertex state go to HBM; the
 * full CC recompute then streams the entries in chunks through two device buffers, the copy of
 * chunk k+1 overlapping the hooking of chunk k (src/cc/scout_cc.cu:86-215). Results are the
 * same in both modes. Snapshots and on-device compaction are not available in this mode.
 * Precondition: the CSR is symmetric (every entry u -> v has its v -> u), as the reference's
 * generated graphs are; the repairs read one endpoint's row per pair and the MIS member counts
 * assume both directions, so a one-direction graph gives answers that differ from the
 * reference's. */
int scout_graph_load(scout_ctx* ctx, int32_t n, int64_t entries, const void* nindex, int wide,
                     const int32_t* nlist);
/* out[0..7] = n, base entries, inserted-row entries, logical entries, rows sorted (0/1),
 * base entries host-resident (0/1), bytes streamed from host memory by full recomputes, and
 * the copy-stream time of those transfers in microseconds. */
enum { SCOUT_GRAPH_INFO_COUNT = 8 }; /* entries scout_graph_info writes */
int scout_graph_info(scout_ctx* ctx, int64_t* out);
/* The logical graph as a CSR in scratchpad-compaction row order (base rows minus deleted
 * entries, then inserted entries in insertion order): core/flash_compactor.h:23-95.
 * off has n+1 entries; col may be NULL to query off only. */
int scout_graph_download(scout_ctx* ctx, int64_t* off, int32_t* col);
/* Merges the inserted rows and deletions into a fresh base CSR in HBM (the scratchpad
 * compaction of core/flash_compactor.h:23-119, done on the device). */
int scout_graph_compact(scout_ctx* ctx);
/* Degree of every vertex of the logical graph into a host array of n int64. */
int scout_graph_degrees(scout_ctx* ctx, int64_t* deg);

/* Solution state. Replace setMISFromHost / getMISToHost (src/mis/scout_mis.cu:480-494) and
 * the colour / level uploads (src/coloring/scout_coloring.cu:313, src/bfs/scout_bfs.cu:562). */
int scout_state_upload(scout_ctx* ctx, int kind, const void* host, int64_t n);
int scout_state_download(scout_ctx* ctx, int kind, void* host, int64_t n);
int scout_set_source(scout_ctx* ctx, int32_t source);

/* Stages one batch of 0-indexed (u, v) pairs (2*m int32) into pinned memory and sends it to
 * the device on the copy stream; the compute stream waits for it before the next call that
 * uses the channel (channel 0 or 1). Replaces the pinned-pool H2D of launchBatch
 * (src/mis/scout_mis.cu:639-647) and the per-round H2D of gpu_relax_round
 * (src/bfs/scout_bfs.cu:223-229). */
int scout_stage_batch(scout_ctx* ctx, int channel, const int32_t* uv, int64_t m);
/* Same, from a device pointer already in HBM (no PCIe transfer). */
int scout_stage_batch_device(scout_ctx* ctx, int channel, const int32_t* d_uv, int64_t m);
/* The channel's last upload: host time copying the lines into pinned staging (the reference's
 * seed pack, src/mis/scout_mis.cu:614-637), device time of the host-to-device copy on the copy
 * stream (HIP events; the reference's h2d time, core/reporting.h:46-51) and bytes sent. Waits
 * for the copy. Any pointer may be NULL. scout_stage_batch on one channel may run on a second
 * host thread while the driving thread works on the other channel (a serving loop stages the
 * next step while the current one repairs); every other call on a context must not overlap. */
int scout_stage_info(scout_ctx* ctx, int channel, double* pack_ms, double* h2d_ms, int64_t* bytes);

/* Labels every staged line insertion (1) or deletion (0) against the rule's graph
 * (parse_batch_file, core/cli_setup.h:87-103). ins may be NULL. */
int scout_classify(scout_ctx* ctx, int channel, int rule, uint8_t* ins);
/* Applies the staged lines to the graph with the scratchpad rules of
 * DeltaOverlay::insertEdge/removeEdge (core/FlashGraph.h:227-310). conditional = 1 applies
 * an insertion only if absent and a deletion only if present at apply time (MIS,
 * src/mis/scout_mis.cu:743-749). Pairs repeated inside the batch are applied in line order. */
int scout_apply(scout_ctx* ctx, int channel, int conditional);

/* Batch repairs. Each runs after scout_apply on the same channel and leaves the solution
 * exact for the updated graph. */
/* BFS levels: replaces processBatch + gpu_relax_round + gpu_deletion_round
 * (src/bfs/scout_bfs.cu:177-272,325-515). */
int scout_repair_bfs(scout_ctx* ctx, int channel, int strategy, scout_stats* st);
/* One batch committed in one call: scout_classify (SCOUT_CLASSIFY_CURRENT, labels kept on the
 * device), scout_apply (unconditional), scout_repair_bfs and scout_repair_cc on the channel,
 * in that order, with the same results and statistics. Needs levels and labels. The
 * reference's BFS and CC binaries each run processBatch on their own solution
 * (src/bfs/scout_bfs.cu:177-272, src/cc/scout_cc.cu:300-441); this context keeps both
 * solutions, and one call lets the host go from each phase's last read-back straight to the
 * next phase's launches. */
int scout_commit(scout_ctx* ctx, int channel, int strategy, scout_stats* bfs_st, scout_stats* cc_st);
/* The cascade bound b of the reference (SCOUT_BFS_BOUND, src/bfs/scout_bfs.cu:475-488): a
 * batch whose invalidated vertices exceed b per deleted parent edge is finished by a full BFS
 * (st->fallback = 1) instead of the local tentative levels and relax. The answer is exact either
 * way, before the batch commits; 0 (the default) never falls back. */
int scout_bfs_bound(scout_ctx* ctx, int64_t bound);
/* How the BFS repair orders its rounds; levels are the same in both. SCOUT_BFS_ASYNC (the
 * default) decides the deletion candidates of every level in the first invalidation round and
 * relaxes every insertion seed in the first relax round: a vertex whose possible support is
 * still being decided waits one round, and a vertex lowered again is relaxed again, so the
 * rounds follow the longest cascade instead of the span of levels the batch touches.
 * SCOUT_BFS_LEVEL_SYNC runs one level per round, in increasing level order. No counterpart in the
 * reference, which repairs one edit at a time (src/bfs/scout_bfs.cu:436-515). */
enum { SCOUT_BFS_ASYNC = 0, SCOUT_BFS_LEVEL_SYNC = 1, SCOUT_BFS_ASYNC_INVALIDATION = 2 };
int scout_bfs_mode(scout_ctx* ctx, int mode);
/* Round groups of the next BFS repair only: its first invalidation group queues inval_rounds
 * levels and its first relax group relax_rounds levels before the first read-back of each
 * (at most 64; 0 keeps the strategy's own sizing). A measurement aid with no counterpart in the
 * reference: fed with the rounds a plain repair of the same batch reported (inval_rounds,
 * relax_rounds), it gives the schedule a perfect predictor would choose, the ceiling of what
 * any learned schedule can gain. Results do not depend on it. */
int scout_bfs_schedule(scout_ctx* ctx, int inval_rounds, int relax_rounds);
/* CC min-id labels: replaces processBatch + bidirectional_reachability_check
 * (src/cc/scout_cc.cu:300-441). */
int scout_repair_cc(scout_ctx* ctx, int channel, int strategy, scout_stats* st);
/* How the local CC repair finds the vertices a batch can relabel. SCOUT_CC_WALK walks from the
 * batch's endpoints over the affected components (work proportional to them, one device round
 * per level); SCOUT_CC_SCAN makes one pass over every vertex (fixed cost, no rounds);
 * SCOUT_CC_AUTO (the default) walks when the batch has at most n/256 endpoints and scans
 * otherwise; under SCOUT_STRATEGY_LEARNED the count adds the vertices the scheduler's split head
 * expects the batch to cut off vertex 0's component (when this batch's BFS repair ran learned
 * on the same context). Labels are the same in every mode. No counterpart in the reference, whose deletion
 * check is a bounded host BFS per line (src/cc/scout_cc.cu:380-441). */
enum { SCOUT_CC_AUTO = 0, SCOUT_CC_WALK = 1, SCOUT_CC_SCAN = 2 };
int scout_cc_mode(scout_ctx* ctx, int mode);
/* Colouring: replaces processBatch + gpu_color_round (src/coloring/scout_coloring.cu:172-277). */
int scout_repair_color(scout_ctx* ctx, int channel, scout_stats* st);
/* MIS: replaces launchBatch + syncAndCollect (src/mis/scout_mis.cu:500-720); max_csr is the
 * reference's MAX_CSR_ELEMENTS sub-batch budget (src/mis/scout_mis.cu:800-815). */
int scout_repair_mis(scout_ctx* ctx, int channel, int64_t max_csr, scout_stats* st);
/* Multi-instance MIS: no counterpart in the reference, which runs one graph per process. The
 * staged lines are grouped by instance (instances own disjoint vertex ranges of the loaded
 * graph): lines [line_off[i], line_off[i+1]) belong to instance i, line_off[0] = 0 and
 * line_off[ninst] = the staged line count. Each instance gets the sub-batch cuts and the result
 * scout_repair_mis gives it alone; all instances' rounds run in the same launches. */
int scout_repair_mis_multi(scout_ctx* ctx, int channel, int32_t ninst, const int64_t* line_off, int64_t max_csr,
                           scout_stats* st);
/* MIS mode. SCOUT_MIS_RULES (the default) applies the reference's per-line rules, which do not
 * restore maximality (a deletion's candidate is rejected by any member within two hops,
 * src/mis/scout_mis.cu:332-344; an insertion re-adds at most one vertex, :326).
 * SCOUT_MIS_COMPLETE then completes the set: every vertex within two hops of an endpoint of the
 * batch, in ascending id, joins unless it or a neighbour is a member, so a maximal set stays
 * maximal. No counterpart in the reference. */
enum { SCOUT_MIS_RULES = 0, SCOUT_MIS_COMPLETE = 1 };
int scout_mis_mode(scout_ctx* ctx, int mode);
/* Launch settings of the multi-instance MIS rounds: grid blocks, rounds per replay of the
 * captured hipGraph, graph on/off. No counterpart in the reference. */
int scout_mis_config(scout_ctx* ctx, int grid, int group_rounds, int use_graph);
/* out[0..2] = the MIS round settings in use: grid, rounds per replay, graph on/off. */
int scout_mis_get_config(scout_ctx* ctx, int* out);
/* Times the MIS repair of the staged, applied batch (lines grouped by instance as for
 * scout_repair_mis_multi) under grids {512, 1024, 2048, 4096} x round groups {4, 8, 16, 32}
 * and keeps the fastest. The membership and member counts are put back before each trial and
 * at the end, so the batch is then repaired as usual; a setting whose membership differs from
 * the first trial's is an error (SCOUT_ERR_STATE). best_ms (may be NULL) = the kept setting's
 * repair time. No counterpart in the reference. */
int scout_mis_autotune(scout_ctx* ctx, int channel, int32_t ninst, const int64_t* line_off, int64_t max_csr,
                       double* best_ms);
/* Capacities of the MIS hop-2 search lists per pass: deferred searches (at most 65536) and
 * chunk tasks. A search that does not fit runs serially in one wave; results do not depend on
 * them. No counterpart in the reference. */
int scout_mis_limits(scout_ctx* ctx, int64_t hop2_requests, int64_t hop2_tasks);
/* Colouring repair settings: rows of at most thread_max entries are handled per thread, rows
 * longer than block_min per workgroup, the rest per wavefront; grid = blocks of the round
 * kernels; use_graph = replay captured round pairs. The reference fixes one thread per vertex
 * (src/coloring/scout_coloring.cu:46-93); results do not depend on these settings. */
int scout_color_config(scout_ctx* ctx, int thread_max, int block_min, int grid, int use_graph);
/* out4 = {thread_max, block_min, grid, use_graph}. */
int scout_color_get_config(scout_ctx* ctx, int* out4);
/* Autotuner: times the colouring repair of the staged, already applied batch under each
 * candidate setting (colours restored between trials) and keeps the fastest. No counterpart
 * in the reference. */
int scout_color_autotune(scout_ctx* ctx, int channel, double* best_ms);

/* Full recomputes. BFS from `source` replaces gpu_full_bfs (src/bfs/scout_bfs.cu:274-323);
 * CC replaces full_recompute_cc_gpu (src/cc/scout_cc.cu:118-278). */
int scout_bfs_full(scout_ctx* ctx, int32_t source);
int scout_cc_full(scout_ctx* ctx);
/* scout_cc_full without any allocation: the min-id labels of the context's current graph go
 * to d_label, a caller-owned device array of n int32 on the context's GPU; the context's own
 * labels are left as they are. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc). Chunk tasks and counters live in a fixed 16 MiB device array per GPU (2^21
 * tasks: graphs of up to about 2^32 adjacency entries; larger ones return SCOUT_ERR_ARG).
 * Every call on one GPU shares that array, so calls that overlap on one GPU (two contexts,
 * two threads) race; calls from one thread in sequence are safe. */
int scout_cc_full_noalloc(scout_ctx* ctx, int32_t* d_label);
/* Greedy-ascending initial MIS and colouring (tools/compute_initial_mis.py:32-43,
 * tools/greedy_color.cpp:37-52), computed on the device. */
int scout_mis_greedy(scout_ctx* ctx);
int scout_color_greedy(scout_ctx* ctx);

/* Waits for all work of the context. */
int scout_sync(scout_ctx* ctx);

/* Synthetic inputs generated in HBM (no reference counterpart; the reference's generators
 * are Python tools). R-MAT with quadrant probabilities (a, b, c, 1-a-b-c), 2^scale
 * vertices, edge_factor * 2^scale generated edges, symmetrised, self-loops and duplicates
 * removed, rows sorted. Bit-identical to the host generator for the same seed. */
int scout_gen_rmat(scout_ctx* ctx, int scale, int edge_factor, double a, double b, double c, uint64_t seed);
/* Generates the next batch of a synthetic stream into a channel, against the current graph
 * and levels: round(delete_frac*size) deletions of current edges (tree_frac of them BFS-tree
 * edges, the rest uniform over adjacency entries) and uniform insertions of non-edges, as
 * canonical u < v pairs never repeated within or across the context's batches (the role of
 * tools/generate_mixed_batches.py). expected_total sizes the memory of toggled pairs. *m
 * receives the line count. */
int scout_gen_batch(scout_ctx* ctx, int channel, int64_t size, double delete_frac, double tree_frac, uint64_t seed,
                    int64_t expected_total, int64_t* m);
/* Endpoint distribution of the generated deletions and insertions (tree deletions are not
 * affected): SCOUT_GEN_UNIFORM = deletions uniform over adjacency entries, insertions uniform
 * pairs; SCOUT_GEN_DEGREE = every endpoint drawn with probability proportional to degree + 1
 * (a deletion takes a uniform live neighbour of its first endpoint), the convention of the
 * "degree" streams of the host generator. */
enum { SCOUT_GEN_UNIFORM = 0, SCOUT_GEN_DEGREE = 1 };
int scout_gen_mode(scout_ctx* ctx, int mode);
/* Copies a staged batch back to the host (cap_lines pairs at most). */
int scout_batch_download(scout_ctx* ctx, int channel, int32_t* uv, int64_t cap_lines);
/* Keeps a device copy of a channel's staged batch in slot k / stages slot k into a channel
 * with a device-to-device copy, so a replayed stream is resident in HBM. *m (may be NULL)
 * receives the loaded batch's line count, the size scout_classify's ins must have. */
int scout_batch_store(scout_ctx* ctx, int channel, int slot);
int scout_batch_load(scout_ctx* ctx, int channel, int slot, int64_t* m);
/* Appends stored batch `slot` after the lines already staged in the channel (a serving step
 * that takes several queued requests at once). *m (may be NULL) receives the total. No
 * counterpart in the reference, which processes one batch file at a time. */
int scout_batch_append(scout_ctx* ctx, int channel, int slot, int64_t* m);
/* Classifies and applies the staged lines as nseg consecutive segments [seg_off[k],
 * seg_off[k+1]) in order, each against the graph the earlier segments left: the same graph
 * and labels as staging the segments as separate batches (parse_batch_file +
 * processBatch's apply loop, core/cli_setup.h:87-103, per batch). The repairs that follow
 * treat all segments as one batch, which gives the same solution as repairing after each.
 * When no pair occurs twice among all the lines, the segments are classified and applied in
 * one pass, with the same result. ins (may be NULL) receives every line's label. */
int scout_apply_segments(scout_ctx* ctx, int channel, int32_t nseg, const int64_t* seg_off, int rule,
                         int conditional, uint8_t* ins);
/* Device copies of the graph and solutions, to replay one stream under several strategies. */
int scout_snapshot(scout_ctx* ctx);
int scout_restore(scout_ctx* ctx);

/* Kernel timing for roofline reports: while enabled, launches of the kernel families below
 * are bracketed by HIP events on the compute stream, and each family accumulates its
 * algorithmic bytes. scout_prof_read returns the summed device time, the launch count and
 * the bytes of one family. */
enum {
  SCOUT_PROF_BFS_RELAX = 0,
  SCOUT_PROF_BFS_INVALIDATE = 1,
  SCOUT_PROF_CC_EXPAND = 2,
  SCOUT_PROF_APPLY = 3,
  SCOUT_PROF_CLASSIFY = 4,
  SCOUT_PROF_COLOR = 5,
  SCOUT_PROF_MIS_EVAL = 6,
  SCOUT_PROF_CC_HOOK = 7,
  /* the k_invalidate launches alone (wave per medium-length candidate row) */
  SCOUT_PROF_BFS_INVALIDATE_KERNEL = 8,
  /* single kernels of the BFS and CC repairs, timed launch by launch */
  SCOUT_PROF_K_RELAX_SHORT = 9,
  SCOUT_PROF_K_RELAX_ROWS = 10,
  SCOUT_PROF_K_INVALIDATE_SHORT = 11,
  SCOUT_PROF_K_INVAL_PUSH = 12,
  SCOUT_PROF_K_CC_EXPLORE = 13, /* CC walk mode: the walks from the batch endpoints */
  SCOUT_PROF_K_CC_COLLECT = 14  /* CC scan mode: the pass over every vertex */
};
int scout_prof_enable(scout_ctx* ctx, int on);
int scout_prof_read(scout_ctx* ctx, int id, double* total_ms, int64_t* launches, int64_t* bytes);

/* Cumulative counts, since the context was created, of the long-row code paths the repairs
 * took (rows longer than one wave's share are split into chunk tasks walked by many waves).
 * No counterpart in the reference, whose kernels give every row to one thread; the counts let
 * tests show that a workload reached these paths. out[0..count), count <= SCOUT_PATH_COUNT. */
enum {
  SCOUT_PATH_INVAL_TASKS = 0,        /* BFS invalidation: chunk tasks searched for support */
  SCOUT_PATH_INVAL_PUSH_TASKS = 1,   /* BFS invalidation: chunk tasks of unsupported long rows */
  SCOUT_PATH_INVAL_KIDS_OVERFLOW = 2,/* BFS invalidation: invalid rows whose children were found by a
                                        pass over the row (too many for the LDS list, or a level-1
                                        row, which skips the support search) */
  SCOUT_PATH_CC_TASKS = 3,           /* CC union-find: chunk tasks of long rows */
  SCOUT_PATH_RELAX_TASKS = 4,        /* BFS relax: chunk tasks of long rows */
  SCOUT_PATH_TENTATIVE_TASKS = 5,    /* BFS tentative levels: chunk tasks of long rows */
  SCOUT_PATH_CC_WALK_ROWS = 6,       /* CC walk mode: rows walked by a whole wave */
  SCOUT_PATH_INVAL_DEFERRED = 7,     /* asynchronous BFS invalidation: vertices that waited a round
                                        for a pending neighbour one level up */
  SCOUT_PATH_COUNT = 8
};
int scout_path_counts(scout_ctx* ctx, int64_t* out, int count);

/* Rows read by the repairs, for the page-bound evaluation of SCOUT_DUMP_TOUCHED
 * (core/FlashGraph.h:318-329,376-385, where every vertex whose adjacency list is fetched is
 * recorded). While enabled, every repair marks each vertex whose row it reads (enabling clears
 * the set). scout_touched_download writes the marked ids in ascending order (at most cap) and
 * their count. */
int scout_touched_enable(scout_ctx* ctx, int on);
int scout_touched_download(scout_ctx* ctx, int32_t* ids, int64_t cap, int64_t* count);

/* Learned scheduler. No counterpart in the reference: it adds a small MLP (12 features, two
 * hidden layers of 64, 3 outputs; bf16 weights, fp32 accumulation) trained on the device on
 * samples logged by the plain BFS repair. Per deleted parent edge it predicts whether the
 * child ends unreached (a split of vertex 0's component), log1p of the cascade size and the
 * levels it spans. SCOUT_STRATEGY_LEARNED uses the predictions only to size work (rounds
 * queued before the first read-back, local repair versus full recompute); results equal the
 * plain strategy's bit for bit. Parameter layout: W1[12][64], b1[64], W2[64][64], b2[64],
 * W3[64][3], b3[3] (row-major, input index first). */
enum { SCOUT_MLP_FEATURES = 12, SCOUT_MLP_OUTPUTS = 3, SCOUT_MLP_PARAMS = 5187 };
/* He-uniform weights from the seed (counter-based, reproducible); not yet trained. */
int scout_mlp_init(scout_ctx* ctx, uint64_t seed);
/* Sets the fp32 master weights (SCOUT_MLP_PARAMS floats); the forward pass uses their
 * round-to-nearest-even bf16 values. Marks the network trained. */
int scout_mlp_set_params(scout_ctx* ctx, const float* params);
/* Master weights (rounded = 0) or the bf16 weights as floats (rounded = 1). */
int scout_mlp_get_params(scout_ctx* ctx, float* params, int rounded);
/* Outputs for n feature rows: out[n][3] = {split logit, log1p size, levels}. */
int scout_mlp_forward(scout_ctx* ctx, const float* feats, int64_t n, float* out);
/* Gradient of the mean loss over n <= 4096 samples (targets[n][3] = {split 0/1, log1p size,
 * levels}) with respect to the parameters, evaluated at the bf16 weights: BCE on the split
 * logit plus half squared errors on the two regressions. */
int scout_mlp_grad(scout_ctx* ctx, const float* feats, const float* targets, int64_t n, float* grad, float* loss);
/* While on, every local BFS repair of the maintained levels logs features and outcomes of up
 * to 16384 deletion candidates (at most 2^22 samples are kept). */
int scout_mlp_log(scout_ctx* ctx, int on);
/* Logged sample count, and up to cap rows of 15 floats (features, targets) if rows != NULL. */
int scout_mlp_samples(scout_ctx* ctx, float* rows, int64_t cap, int64_t* count);
/* Adam over the logged samples (batches of 4096, a seeded permutation); mean loss of the first
 * and the last epoch. Marks the network trained. */
int scout_mlp_train(scout_ctx* ctx, int epochs, float lr, uint64_t seed, float* first_loss, float* last_loss);

#ifdef __cplusplus
}
#endif
#endif
