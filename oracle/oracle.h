/* oracle.h -- CPU restatement of the SCOUT reference's hot path, used ONLY as the test oracle.
 *
 * This is test infrastructure. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it. The product path (libscout_hip.so and the scout_* binaries)
 * never links, loads or calls it.
 *
 * What each function restates (reference file:line) is written next to it below and in
 * oracle.cpp. Everything is integer arithmetic except the MLP. What pins each function, so
 * that a plausible mistake fails a check other than a retyped formula (tests/):
 *   - graph overlay, classify, apply, rows, degrees, logical CSR, compaction: equal to the
 *     reference's own HybridGraph and compact_flash_graph compiled from its headers
 *     (test_ref_core.py: test_scratchpad_restatement_equals_reference_hybrid_graph,
 *     test_compaction_is_byte_identical_to_reference,
 *     test_inconsistent_delta_fails_compaction_like_reference), plus
 *     test_scratchpad_row_order_and_toggles, test_scratchpad_toggles_are_symmetric_and_consistent;
 *   - orc_bfs, orc_cc: scipy shortest_path / connected_components after random updates
 *     (test_bfs_and_cc_match_scipy_after_random_updates); the sample fixtures
 *     (test_sample_fixtures_are_closed_forms);
 *   - orc_check_*: planted mistakes (test_validity_checks_catch_planted_mistakes);
 *   - orc_greedy_mis, orc_greedy_color: the sample fixtures and brute-force first fit
 *     (test_greedy_states_are_first_fit);
 *   - orc_color_batch: hand-worked rounds (test_coloring_rounds_hand_worked,
 *     test_coloring_path_insert) and properness on random streams;
 *   - orc_mis_batch: hand-worked rule cases (test_mis_deletion_*, test_mis_insertion_*), the
 *     multi-sub-batch cut case (test_mis_sub_batch_cut_decides_the_row_kinds) and
 *     independence on random streams. The line-order serialisation itself is "parity
 *     unpinned": the reference's own result is racy, so no reference output pins it;
 *   - orc_mis_max_csr: the formula with its clamps (test_mis_max_csr_formula);
 *   - orc_mis_complete: a hand case and maximality / region invariants on random streams
 *     (test_mis_complete_*);
 *   - orc_refcc_*: canonical labels equal the closed form every batch with no bound
 *     (test_per_edit_reference_cc_is_exact_every_batch);
 *   - orc_refbfs_*: its insertion relax equals the closed form on insertion-only streams
 *     (test_per_edit_reference_bfs_insertions_are_exact); its deletion worklist is "parity
 *     unpinned" beyond its code citation (the reference algorithm itself is not exact there);
 *   - orc_mlp_*, orc_bf16_round: numpy closed forms, central finite differences, a planted
 *     mistake and round-to-nearest-even cases (test_mlp_oracle_*, test_bf16_round_is_nearest_even).
 */
#ifndef SCOUT_ORACLE_H
#define SCOUT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_graph orc_graph;

/* Base graph + scratchpad overlay (core/FlashGraph.h DeltaOverlay/HybridGraph semantics). */
orc_graph* orc_graph_open(const char* egr_path);
orc_graph* orc_graph_from_csr(int32_t n, const int64_t* off, const int32_t* col);
void orc_graph_free(orc_graph* g);
int32_t orc_n(const orc_graph* g);
int64_t orc_base_edges(const orc_graph* g);

/* Classification "insert iff !isAdjacent(u,v)" of every pair (core/cli_setup.h:87-103). */
void orc_classify(orc_graph* g, const int32_t* uv, int64_t m, uint8_t* ins);
/* insertEdge/removeEdge per line, unconditionally (coloring/bfs/cc processBatch). */
void orc_apply(orc_graph* g, const int32_t* uv, const uint8_t* ins, int64_t m);
/* MIS's conditional apply (src/mis/scout_mis.cu:743-749). */
void orc_apply_conditional(orc_graph* g, const int32_t* uv, const uint8_t* ins, int64_t m);
int orc_is_adjacent(orc_graph* g, int32_t u, int32_t v);
int64_t orc_degree(orc_graph* g, int32_t v);
int64_t orc_row(orc_graph* g, int32_t v, int32_t* out, int64_t cap);
/* Logical CSR in compactor row order (core/flash_compactor.h:30-95). off has n+1 entries;
 * col may be NULL to query the size. Returns the entry count, or -1 on an inconsistent
 * delta (negative degree / degree mismatch, the compactor's error exits). */
int64_t orc_logical_csr(orc_graph* g, int64_t* off, int32_t* col);
/* Durable wide .egr of the logical graph (compact_flash_graph). 0 on success. */
int orc_compact_to(orc_graph* g, const char* path);

/* Closed forms. */
void orc_bfs(orc_graph* g, int32_t src, int32_t* level);
int32_t orc_cc(orc_graph* g, int32_t* label);

/* Per-batch repairs (graph already updated with the batch). */
int64_t orc_color_batch(orc_graph* g, const int32_t* uv, const uint8_t* ins, int64_t m, int32_t* color,
                        int64_t* stats /* [conflicts, rounds] */);
int64_t orc_mis_max_csr(double avg_degree, double free_bytes);
int orc_mis_batch(orc_graph* g, const int32_t* uv, const uint8_t* ins, int64_t m, int64_t max_csr, uint8_t* mis,
                  int64_t* stats /* [sub_batches, rows, writes] */);

/* Validity checks; each returns the number of violations. */
int64_t orc_check_bfs(orc_graph* g, int32_t src, const int32_t* level);
int64_t orc_check_cc(orc_graph* g, const int32_t* label);
int64_t orc_check_coloring(orc_graph* g, const int32_t* color);
int64_t orc_check_mis_independent(orc_graph* g, const uint8_t* mis);
int64_t orc_mis_maximality_violations(orc_graph* g, const uint8_t* mis);
/* MIS complete mode: greedy ascending completion over the batch's 2-hop region; returns the
 * vertices added. */
int64_t orc_mis_complete(orc_graph* g, const int32_t* uv, int64_t m, uint8_t* mis);

/* Initial states by greedy ascending id (tools/compute_initial_mis.py:32-43,
 * tools/greedy_color.cpp:37-52). */
void orc_greedy_mis(orc_graph* g, uint8_t* mis);
void orc_greedy_color(orc_graph* g, int32_t* color);

/* The reference's own per-edit algorithms restated on the CPU (CPU baseline):
 * BFS processBatch (src/bfs/scout_bfs.cu:436-515) and CC processBatch
 * (src/cc/scout_cc.cu:392-441). The states persist inside the handle. */
typedef struct orc_refbfs orc_refbfs;
orc_refbfs* orc_refbfs_create(orc_graph* g, const int32_t* level, int32_t source, int64_t bound, int defer);
void orc_refbfs_free(orc_refbfs* r);
void orc_refbfs_batch(orc_refbfs* r, const int32_t* uv, const uint8_t* ins, int64_t m, int64_t* stats /* 8 */);
void orc_refbfs_finish(orc_refbfs* r, int32_t* level);
/* OpenMP threads the per-edit baselines use for row fetches and full recomputes. */
int orc_threads(void);

typedef struct orc_refcc orc_refcc;
orc_refcc* orc_refcc_create(orc_graph* g, const int32_t* root, int64_t bound, int defer);
void orc_refcc_free(orc_refcc* r);
void orc_refcc_batch(orc_refcc* r, const int32_t* uv, const uint8_t* ins, int64_t m, int64_t* stats /* 8 */);
void orc_refcc_finish(orc_refcc* r, int32_t* label);

/* The learned scheduler's MLP (no reference counterpart): 12 -> 64 -> 64 -> 3, ReLU,
 * parameters W1[12][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3]. Restated in double
 * precision with the parameters taken as given (pass the bf16-rounded values the device
 * uses). Loss per sample: BCE(out0, t0) + 0.5 (out1 - t1)^2 + 0.5 (out2 - t2)^2; the mean over
 * n samples is returned and its gradient written to grad (5187 doubles) when grad != NULL. */
double orc_mlp_loss_grad(const float* params, const float* feats, const float* targets, int64_t n, double* grad);
void orc_mlp_forward(const float* params, const float* feats, int64_t n, double* out);
/* fp32 -> bf16 (round to nearest, ties to even) -> fp32. */
float orc_bf16_round(float x);

#ifdef __cplusplus
}
#endif
#endif
