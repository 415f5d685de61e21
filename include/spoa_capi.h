/*
 * spoa_capi.h -- C entry points over the C++ drop-in API (spoa::AlignmentEngine,
 * spoa::Graph, include/spoa/spoa.hpp), the batched and per-call window
 * drivers, and the seeded synthetic window generator.  For callers that
 * bind C (the Python package binds these with ctypes).
 *
 * Errors: functions returning int return a negative value (-1: the C++ call
 * threw, message in spoa_last_error(); -2: an output buffer is too small),
 * functions returning a handle return NULL.
 */
#ifndef SPOA_CAPI_H_
#define SPOA_CAPI_H_

#include <stdint.h>

#include "spoa_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* spoa_last_error(void);

/* AlignmentEngine::Create(type, m, n, g, e, q, c) (reference
 * include/spoa/alignment_engine.hpp:49-55); the engine runs on the GPU. */
void* spoa_engine_new(int type, int m, int n, int g, int e, int q, int c);
/* An engine that stands in for the reference's scalar engine
 * (SisdAlignmentEngine::Create, src/sisd_alignment_engine.cpp:16-27, what
 * AlignmentEngine::Create returns from a reference built without SIMD):
 * the same validation as spoa_engine_new, that engine's results and messages. */
void* spoa_sisd_engine_new(int type, int m, int n, int g, int e, int q, int c);
void spoa_engine_free(void* engine);
int spoa_engine_subtype(void* engine);
/* AlignmentEngine::Prealloc (alignment_engine.hpp:57-59) */
int spoa_engine_prealloc(void* engine, uint32_t max_len, uint32_t alphabet);
/* measurements of the engine's last Align */
int spoa_engine_stats(void* engine, spoa_hip_stats* out);
/* the C ABI handle behind an engine (tuning, memory budget) */
spoa_hip_engine* spoa_engine_device(void* engine);
/* AlignmentEngine::Align (alignment_engine.hpp:61-69): returns the number of
 * pairs written to pairs_out (2 int32 each); *score / *has_score as the
 * reference writes its score; *width_bits the score width used. */
int spoa_align(void* engine, void* graph, const char* seq, uint32_t len, int32_t* pairs_out, uint32_t cap,
               int32_t* score, int* has_score, int* width_bits);

/* spoa::Graph (reference include/spoa/graph.hpp:26-318) */
void* spoa_graph_new(void);
void spoa_graph_free(void* graph);
void spoa_graph_clear(void* graph);
uint32_t spoa_graph_num_nodes(void* graph);
/* Graph::AddAlignment with per-base weights (graph.hpp:154-157), or weight 1
 * when weights is NULL (graph.hpp:149-152) */
int spoa_graph_add_alignment(void* graph, const int32_t* pairs, uint32_t n_pairs, const char* seq, uint32_t len,
                             const uint32_t* weights, uint32_t n_weights);
/* Graph::GenerateConsensus (graph.hpp:167); returns the length */
int spoa_graph_consensus(void* graph, char* out, uint32_t cap);
/* Graph::GenerateMultipleSequenceAlignment (graph.hpp:164-165): rows back to back */
int spoa_graph_msa(void* graph, int include_consensus, char* out, uint64_t cap, uint32_t* rows, uint32_t* cols);
/* rank_to_node() ids (graph.hpp:110-112) */
int spoa_graph_ranks(void* graph, uint32_t* out, uint32_t cap);

/* Graph::Subgraph (reference include/spoa/graph.hpp:173-176): the part of
 * the graph node `end` reaches backwards down to node id `begin`, as a new
 * graph handle; sub_to_graph[k] (cap >= the graph's node count) gets the
 * graph node id of subgraph node k, -1 past the subgraph's nodes.  NULL on
 * error. */
void* spoa_graph_subgraph(void* graph, uint32_t begin, uint32_t end, int32_t* sub_to_graph, uint32_t cap);
/* Graph::UpdateAlignment (graph.hpp:178-180): node ids of n_pairs pairs
 * aligned against a subgraph mapped back to the graph, in place. */
int spoa_graph_update_alignment(void* graph, const int32_t* sub_to_graph, uint32_t n_map, int32_t* pairs,
                                uint32_t n_pairs);
/* Graph::PrintDot (graph.hpp:183): Graphviz text to the file at path. */
int spoa_graph_print_dot(void* graph, const char* path);
/* Graph::GenerateConsensus(summary, verbose) (graph.hpp:169-171): writes
 * the consensus (returns its length) and the coverage summary
 * (*summary_len values; -2 when a buffer is too small). */
int spoa_graph_consensus_summary(void* graph, int verbose, char* out, uint32_t cap, uint32_t* summary,
                                 uint32_t summary_cap, uint32_t* summary_len);
/* The command-line tool's GFA output (reference src/main.cpp:119-199)
 * after GenerateConsensus: headers holds one NUL-terminated name per
 * sequence back to back; reversed (n_reversed flags or NULL) marks reads
 * added reverse-complemented.  Returns the text length, or -2 - length
 * when cap is too small, -1 on error. */
int64_t spoa_graph_gfa(void* graph, const char* headers, const uint8_t* reversed, uint32_t n_reversed,
                       int include_consensus, char* out, uint64_t cap);

/* The flat partial-order graph of the batched form: the same operations
 * and results as spoa::Graph (AddAlignment, GenerateConsensus,
 * GenerateMultipleSequenceAlignment, rank_to_node). */
void* spoa_poa_graph_new(void);
void spoa_poa_graph_free(void* graph);
uint32_t spoa_poa_graph_num_nodes(void* graph);
int spoa_poa_graph_add_alignment(void* graph, const int32_t* pairs, uint32_t n_pairs, const char* seq, uint32_t len,
                                 const uint32_t* weights, uint32_t n_weights);
int spoa_poa_graph_consensus(void* graph, char* out, uint32_t cap);
int spoa_poa_graph_msa(void* graph, int include_consensus, char* out, uint64_t cap, uint32_t* rows, uint32_t* cols);
int spoa_poa_graph_ranks(void* graph, uint32_t* out, uint32_t cap);

/* ----------------------------------------------------------- windows -- */
enum { SPOA_BATCH_BATCHED = 0, SPOA_BATCH_PER_CALL = 1 };

typedef struct spoa_batch_opts {
  int32_t type, m, n, g, e, q, c;
  int32_t strand;  /* align both strands and keep the better (reference src/main.cpp:285-302) */
  int32_t msa;     /* also build every window's MSA */
  int32_t threads; /* host threads, 0 = all */
  int32_t device;
  int32_t overlap; /* batched form: the two halves' launches may run on the device at once */
  int32_t graph;   /* batched form: every pass launched as a captured HIP graph */
  int32_t resident; /* batched form: window graphs on the device (merges, consensus and MSA there) */
  int32_t sisd;     /* the reference's scalar engine's alignments (src/sisd_alignment_engine.cpp) */
} spoa_batch_opts;

typedef struct spoa_batch_stats {
  double wall_ms, align_ms, kernel_ms, h2d_ms, d2h_ms, graph_ms, finish_ms;
  double pack_ms, alloc_ms, wait_ms, unpack_ms;
  uint64_t cells, dp_bytes, align_calls;
  uint32_t launches, rounds;
  double kernel_ms16, kernel_ms32; /* kernel_ms of the int16 / int32 launches */
  uint64_t dp_bytes16, dp_bytes32; /* dp_bytes of the int16 / int32 jobs */
  uint32_t launches16, launches32; /* launches of int16 / int32 jobs */
  uint32_t graph_builds, graph_updates; /* graph mode: passes instantiated / updated in place and replayed */
} spoa_batch_stats;

/* Runs every window (reads [win_read_off[w], win_read_off[w+1]); read k is
 * seqs[read_off[k], read_off[k+1]), quals likewise or NULL) through
 * Align + AddAlignment in order, then GenerateConsensus (+ MSA).
 * mode SPOA_BATCH_BATCHED: one launch per round over all live windows;
 * SPOA_BATCH_PER_CALL: the drop-in Align, one read at a time. */
void* spoa_batch_run(int mode, const spoa_batch_opts* opts, uint32_t n_windows, const uint32_t* win_read_off,
                     const uint64_t* read_off, const char* seqs, const char* quals);
void spoa_batch_free(void* result);
/* A reusable batched-form context (engine, device buffers, host threads)
 * for repeated runs with one set of options; runs return results as
 * spoa_batch_run does. */
void* spoa_batch_context_new(const spoa_batch_opts* opts);
void spoa_batch_context_free(void* ctx);
spoa_hip_engine* spoa_batch_context_device(void* ctx);
/* the engine of one half of the windows (0 or 1; the batched form runs two) */
spoa_hip_engine* spoa_batch_context_half_device(void* ctx, int32_t half);
/* Launch configuration of the context's batched launches for rows of
 * reference width up to max_width at one score width (spoa_hip_set_config
 * on both engines); cols_per_lane 0 clears that width's overrides. */
int spoa_batch_context_set_config(void* ctx, int32_t width_bits, uint32_t max_width, int32_t cols_per_lane,
                                  int32_t waves);
/* The same with split (spoa_hip_set_config_split): rows that span several
 * tiles run each tile on a workgroup of its own. */
int spoa_batch_context_set_config_split(void* ctx, int32_t width_bits, uint32_t max_width, int32_t cols_per_lane,
                                        int32_t waves, int32_t split);
/* The autotuner (no counterpart in the reference): tries launch
 * configurations on the given windows (their first `rounds` rounds, or the
 * whole batch for a score width that those rounds do not reach), keeps the
 * fastest per score width and installs it on the context; cache_path (or
 * NULL) names a JSON cache.  Writes per width {width, cols_per_lane, waves,
 * flags} into out (4 ints each; flags bit 0: from the cache, bit 1: timed
 * on the whole batch) and the chosen configuration's device ms into
 * out_ms; returns the number of widths, -1 on error. */
int spoa_batch_context_autotune(void* ctx, uint32_t n_windows, const uint32_t* win_read_off, const uint64_t* read_off,
                                const char* seqs, uint32_t rounds, const char* cache_path, int32_t* out,
                                double* out_ms, uint32_t cap);
void* spoa_batch_context_run(void* ctx, uint32_t n_windows, const uint32_t* win_read_off, const uint64_t* read_off,
                             const char* seqs, const char* quals);
int spoa_batch_get_stats(void* result, spoa_batch_stats* out);
int spoa_batch_consensus(void* result, uint32_t window, char* out, uint32_t cap);
/* every window's consensus back to back; offs[n_windows + 1] gets the
 * offsets (filled even when cap is too small); returns the total length */
int64_t spoa_batch_consensus_all(void* result, char* out, uint64_t cap, uint64_t* offs);
int spoa_batch_msa(void* result, uint32_t window, char* out, uint64_t cap, uint32_t* rows, uint32_t* cols);
int spoa_batch_reversed(void* result, uint32_t window, uint8_t* out, uint32_t cap);

/* The reverse complement the strand-ambiguous mode aligns (reference
 * src/main.cpp:286, 296, biosoup Sequence::ReverseAndComplement, not
 * vendored by the reference): upper-case, then the Watson-Crick / IUPAC
 * complement (A-T, C-G, U-A, R-Y, K-M, B-V, D-H; S, W to themselves), every
 * other byte kept; written reversed into out[len].  Returns len. */
int spoa_reverse_complement(const char* seq, uint32_t len, char* out);

/* ----------------------------------------------------------- serving -- */
/* A serving loop on one GPU (continuous batching of consensus jobs, each a
 * window of reads with its own scoring parameters; DP memory admitted per
 * round).  No counterpart in the reference. */
typedef struct spoa_serve_opts {
  int32_t threads;          /* host threads, 0 = all */
  int32_t device;
  uint32_t max_active;      /* job slots */
  uint32_t max_param_sets;  /* scoring parameter sets served at once */
  double memory_fraction;   /* of the free device memory, for DP matrices */
} spoa_serve_opts;

typedef struct spoa_serve_stats {
  uint64_t submitted, completed, failed, rounds, launches, align_calls, deferred, cells, dp_bytes;
  double kernel_ms;
  uint32_t max_active_seen;
} spoa_serve_stats;

void* spoa_server_new(const spoa_serve_opts* opts);
/* finishes every submitted job, then stops */
void spoa_server_free(void* server);
/* One job: Create(type, m, n, g, e, q, c) parameters, reads
 * seqs[read_off[k], read_off[k+1]) (quals likewise or NULL), msa != 0 to
 * build the MSA too.  Returns the job id, 0 on error (spoa_last_error). */
uint64_t spoa_server_submit(void* server, int type, int m, int n, int g, int e, int q, int c, uint32_t n_reads,
                            const uint64_t* read_off, const char* seqs, const char* quals, int msa);
/* Waits for a job (timeout_s < 0: no limit); returns its result, NULL on
 * timeout.  Each result is returned once. */
void* spoa_server_wait(void* server, uint64_t id, double timeout_s);
int spoa_server_stats(void* server, spoa_serve_stats* out);
double spoa_server_now_ms(void* server);
/* result: status 0 ok / -1 failed; times in ms since the server started */
int spoa_serve_result_info(void* result, int* status, double* submit_ms, double* admit_ms, double* done_ms,
                           uint32_t* rounds);
/* the size-class lane the job ran in: 0..3 by first-read length (< 700,
 * < 2,000, < 5,000, longer) */
int spoa_serve_result_lane(void* result);
const char* spoa_serve_result_error(void* result);
int spoa_serve_result_consensus(void* result, char* out, uint32_t cap);
int spoa_serve_result_msa(void* result, char* out, uint64_t cap, uint32_t* rows, uint32_t* cols);
void spoa_serve_result_free(void* result);

/* ------------------------------------------------------------- synth -- */
typedef struct spoa_synth_spec {
  uint32_t template_min, template_max; /* template length range */
  uint32_t reads_min, reads_max;       /* reads per window */
  uint32_t read_len_min, read_len_max; /* 0, 0: every read covers the template */
  double err_min, err_max;             /* per-window error rate range */
  uint32_t revcomp_odd;                /* odd-numbered reads reverse-complemented */
  uint32_t log_uniform_len;            /* read length log-uniform in its range (the template's
                                          length when reads span the template) */
} spoa_synth_spec;

void* spoa_synth_new(const spoa_synth_spec* spec, uint64_t seed, uint32_t n_windows);
void spoa_synth_free(void* h);
void spoa_synth_sizes(void* h, uint32_t* n_windows, uint64_t* n_reads, uint64_t* seq_bytes, uint64_t* templ_bytes);
void spoa_synth_copy(void* h, uint32_t* win_read_off, uint64_t* read_off, char* seqs, uint64_t* templ_off,
                     char* templates);

#ifdef __cplusplus
}
#endif

#endif /* SPOA_CAPI_H_ */
