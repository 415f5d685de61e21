/*
 * spoa_hip.h -- the C ABI between the C++ host code (spoa::AlignmentEngine,
 * the batched window driver, the serving loop) and the HIP graph-to-sequence
 * alignment kernels for gfx950.
 *
 * Plain pointers and sizes only; every host buffer is owned by the caller.
 * Every entry point returns a status code (SPOA_HIP_OK on success); the
 * message of the last failure on the calling thread is returned by
 * spoa_hip_last_error().  Messages of the conditions the reference reports
 * are the reference's own strings.
 *
 * The graph is handed over as a rank-ordered CSR view of spoa::Graph:
 * rank_to_node() order, Node::inedges order of the predecessors,
 * Node::outedges emptiness and decoder(Node::code) (reference
 * include/spoa/graph.hpp:102-132).
 */
#ifndef SPOA_HIP_H_
#define SPOA_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum spoa_hip_status {
  SPOA_HIP_OK = 0,
  SPOA_HIP_INVALID_ARGUMENT = 1, /* bad parameter (the message says which) */
  SPOA_HIP_TOO_LARGE = 2,        /* "too large sequence!" */
  SPOA_HIP_OVERFLOW = 3,         /* "possible overflow!" */
  SPOA_HIP_NO_MEMORY = 4,        /* "insufficient memory!" */
  SPOA_HIP_DEVICE_ERROR = 5,     /* a HIP runtime call failed */
  SPOA_HIP_CAPACITY = 6,         /* an output buffer is too small */
  SPOA_HIP_INCONSISTENT = 7      /* the traceback found no consistent move */
};

/* Alignment type and gap subtype, numbered as spoa::AlignmentType
 * (reference include/spoa/alignment_engine.hpp:16-20) and
 * spoa::AlignmentSubtype (alignment_engine.hpp:22-26). */
enum spoa_hip_type { SPOA_HIP_SW = 0, SPOA_HIP_NW = 1, SPOA_HIP_OV = 2 };
enum spoa_hip_subtype { SPOA_HIP_LINEAR = 0, SPOA_HIP_AFFINE = 1, SPOA_HIP_CONVEX = 2 };

/* Scoring parameters after AlignmentEngine::Create's validation and
 * normalisation (reference src/alignment_engine.cpp:32-66): linear has
 * e == g, affine has q == g and c == e. */
typedef struct spoa_hip_params {
  int32_t type;
  int32_t subtype;
  int8_t m, n, g, e, q, c;
} spoa_hip_params;

/* Rank-ordered view of a POA graph with V nodes. */
typedef struct spoa_hip_graph {
  uint32_t num_nodes;        /* V */
  const uint8_t* row_char;   /* [V] decoder(code) of the node at each rank */
  const int32_t* node_id;    /* [V] node id at each rank */
  const uint32_t* pred_off;  /* [V + 1] CSR offsets into pred_rank */
  const uint32_t* pred_rank; /* predecessor ranks, in Node::inedges order */
  const uint8_t* is_sink;    /* [V] 1 if the node has no out-edges */
} spoa_hip_graph;

/* One alignment of a batch. */
typedef struct spoa_hip_job {
  spoa_hip_graph graph;
  const char* seq;
  uint32_t len;
  int32_t* pairs;      /* out: 2 * cap int32, (node id, column) pairs in path order */
  uint32_t cap;        /* capacity of pairs, in pairs (V + len suffices) */
  uint32_t n_pairs;    /* out */
  int32_t score;       /* out: valid when has_score != 0 */
  int32_t has_score;   /* out: 0 where the reference leaves *score untouched */
  int32_t width_bits;  /* out: 16 or 32 (the score width the range bound chose), 0 if no DP ran */
  int32_t status;      /* out: per-job status */
} spoa_hip_job;

typedef struct spoa_hip_engine spoa_hip_engine;

/* Replaces the device half of AlignmentEngine::Create (reference
 * include/spoa/alignment_engine.hpp:35-55, src/alignment_engine.cpp:32-73):
 * binds an engine with normalised parameters to a HIP device. */
int spoa_hip_engine_create(int device, const spoa_hip_params* params, spoa_hip_engine** out);

/* Replaces SisdAlignmentEngine::Create (reference
 * src/sisd_alignment_engine.hpp:26-34, src/sisd_alignment_engine.cpp:16-27),
 * the engine AlignmentEngine::Create returns from a reference built without
 * SIMD (src/alignment_engine.cpp:68-71, e.g. its default x86 CMake flags):
 * the same device path, with that engine's results and messages -- its
 * overflow bound on (n, V), int32 scores (width_bits 32), full-Gotoh convex
 * E / Q, its end cell (no padding columns, no 0 floor) and its traceback
 * rules (src/sisd_alignment_engine.cpp:255-926).  Every entry point taking
 * the engine follows them. */
int spoa_hip_engine_create_sisd(int device, const spoa_hip_params* params, spoa_hip_engine** out);

int spoa_hip_engine_destroy(spoa_hip_engine* engine);

/* Replaces AlignmentEngine::Prealloc (reference alignment_engine.hpp:57-59,
 * src/simd_alignment_engine_implementation.hpp:381-422): reserves device
 * memory for reads up to max_len over graphs of max_len * alphabet_size
 * nodes.  Errors: SPOA_HIP_TOO_LARGE, SPOA_HIP_NO_MEMORY. */
int spoa_hip_prealloc(spoa_hip_engine* engine, uint32_t max_len, uint32_t alphabet_size);

/* Replaces AlignmentEngine::Align (reference alignment_engine.hpp:66-69,
 * src/simd_alignment_engine_implementation.hpp:703-2131): aligns one read
 * to the graph on the device (DP + traceback) and returns the
 * (node id, column) pairs in path order; *score and *has_score as the
 * reference writes its score.  Errors: SPOA_HIP_TOO_LARGE,
 * SPOA_HIP_OVERFLOW, SPOA_HIP_NO_MEMORY, SPOA_HIP_CAPACITY. */
int spoa_hip_align(spoa_hip_engine* engine, const spoa_hip_graph* graph, const char* seq, uint32_t len,
                   int32_t* pairs, uint32_t cap, uint32_t* n_pairs, int32_t* score, int32_t* has_score);

/* No counterpart in the reference: the variant of spoa_hip_align for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  It calls no hipMalloc, hipFree or other
 * allocating routine and never synchronises: every device byte it needs
 * beyond its inputs and outputs comes from a fixed scratch array of 64 MiB
 * per GPU, held from the library's first use on that GPU until the process
 * ends, and shared by every call on that GPU -- calls that overlap on one
 * GPU race.  Inputs and outputs are device memory: the graph view's arrays,
 * seq[len], pairs (2 * cap int32, 8-byte aligned; (node id, column) pairs
 * in path order) and result[5] = {n_pairs, score, has_score, status,
 * width_bits}, written by kernels on `stream` (a hipStream_t; NULL: the
 * engine's stream).  Host-side errors as spoa_hip_align
 * (SPOA_HIP_TOO_LARGE, SPOA_HIP_OVERFLOW); an alignment whose matrices
 * and graph block need more than the scratch array fails with
 * SPOA_HIP_NO_MEMORY before any launch (or, when its predecessor lists do
 * not fit, with status SPOA_HIP_NO_MEMORY in result[3]); the per-alignment
 * status of spoa_hip_align goes to result[3]. */
int spoa_hip_align_noalloc(spoa_hip_engine* engine, const spoa_hip_graph* graph, const char* seq, uint32_t len,
                           int32_t* pairs, uint32_t cap, int32_t* result, void* stream);

/* No counterpart in the reference, which aligns one read per call: aligns
 * n_jobs independent (graph, read) pairs, one persistent launch per launch
 * configuration (and per memory-budget chunk), jobs queued by V x n.
 * Per-job results and status are written into the job records; the return
 * value is the first failure. */
int spoa_hip_align_batch(spoa_hip_engine* engine, spoa_hip_job* jobs, uint32_t n_jobs);

/* The batch form split in two, so the host can work while the device
 * aligns (no counterpart in the reference): submit packs and launches
 * (any memory-budget passes but the last run to completion), wait blocks
 * until the last pass is done and fills the job records; it returns what
 * spoa_hip_align_batch would have.  The jobs, their graphs and reads must
 * stay untouched until wait returns.  A further submit or align on the
 * engine finishes a pending batch first. */
int spoa_hip_align_batch_submit(spoa_hip_engine* engine, spoa_hip_job* jobs, uint32_t n_jobs);
int spoa_hip_align_batch_wait(spoa_hip_engine* engine);
/* No counterpart in the reference: 1 when a submitted batch has finished on
 * the device (or none is pending), so that wait will not block; 0 while it
 * runs; negative (a status) on a bad argument. */
int spoa_hip_align_batch_ready(spoa_hip_engine* engine);

/* Measurements of the last align / align_batch call: device time of the
 * alignment kernels from HIP events on the engine's stream, copies, launch
 * count, DP cells (sum of V * n) and algorithmic DP bytes (sum of
 * S * n * (W * V + Rd * R): every score matrix written once, the
 * predecessor rows read once per in-edge). */
typedef struct spoa_hip_stats {
  double kernel_ms;  /* sum of the alignment launches' device times (HIP events around each launch) */
  double h2d_ms;     /* device time of the input copy */
  double d2h_ms;     /* device time of the output copy */
  double host_ms;    /* host time packing the jobs */
  double alloc_ms;   /* host time growing device / pinned buffers */
  double wait_ms;    /* host time blocked on the stream */
  double unpack_ms;  /* host time unpacking the pairs */
  uint32_t launches;
  uint32_t jobs;
  uint64_t cells;
  uint64_t dp_bytes;
  double kernel_ms16;  /* kernel_ms of the int16-score launches */
  double kernel_ms32;  /* kernel_ms of the int32-score launches */
  uint64_t dp_bytes16, dp_bytes32;  /* dp_bytes of the int16 / int32 jobs */
  uint32_t launches16, launches32;  /* launches of int16 / int32 jobs */
  double merge_ms;   /* device-resident windows: device time of the graph merges */
  uint32_t merges;   /* device-resident windows: windows merged */
  /* graph mode: passes whose capture was instantiated as a new executable
   * graph, and passes that updated the previous pass's executable graph in
   * place (same shape) and replayed it */
  uint32_t graph_builds, graph_updates;
} spoa_hip_stats;
int spoa_hip_last_stats(spoa_hip_engine* engine, spoa_hip_stats* out);

/* Launch configuration (columns per lane, waves per workgroup) for a read
 * of length len at a score width; batch != 0 asks for the throughput
 * configuration of align_batch, 0 for the latency one of align. */
int spoa_hip_config(spoa_hip_engine* engine, uint32_t len, int32_t width_bits, int32_t batch,
                    int32_t* cols_per_lane, int32_t* waves);

/* Override the launch configuration for rows whose reference width
 * (ceil(len / L) * L) is at most max_width, at one score width and mode;
 * used by the autotuner.  cols_per_lane in {8, 16, 32} (convex int16 needs
 * at least 16; 32 with one wave only), waves in 1..8.  cols_per_lane == 0 clears all overrides of that width
 * and mode. */
int spoa_hip_set_config(spoa_hip_engine* engine, int32_t width_bits, int32_t batch, uint32_t max_width,
                        int32_t cols_per_lane, int32_t waves);
/* No counterpart in the reference.  As spoa_hip_config / spoa_hip_set_config,
 * with split: != 0 when a job whose row spans several tiles of waves x 64 x
 * cols_per_lane columns runs every tile on a workgroup of its own (each
 * tile hands every row's boundary -- its last H and the gap carry -- to the
 * next tile through device memory, the tiles of a job running side by
 * side), 0 when one workgroup runs a row's tiles in turn.  By default rows
 * wider than 8 waves are split into one-wave tiles.  spoa_hip_set_config
 * sets split = 0. */
int spoa_hip_config_split(spoa_hip_engine* engine, uint32_t len, int32_t width_bits, int32_t batch,
                          int32_t* cols_per_lane, int32_t* waves, int32_t* split);
int spoa_hip_set_config_split(spoa_hip_engine* engine, int32_t width_bits, int32_t batch, uint32_t max_width,
                              int32_t cols_per_lane, int32_t waves, int32_t split);
/* No counterpart in the reference.  The longest time one row tile of a
 * split job waits for another, in ticks of the device's 100 MHz clock
 * (default 10^9, 10 s); a job whose tiles wait longer fails with
 * SPOA_HIP_DEVICE_ERROR ("... timed out waiting for each other") and the
 * launch's other split jobs stop waiting.  For tests of that path. */
int spoa_hip_set_wait_limit(spoa_hip_engine* engine, uint32_t ticks);

/* No counterpart in the reference.  Makes engine launch its kernels on
 * owner's compute stream (owner NULL: back on its own), so the kernels of
 * several engines run one after another while their copies and host work
 * overlap; every launch then has the device to itself, which is what the
 * per-launch times in spoa_hip_stats measure.  Both engines must be on one
 * device, and owner must outlive engine. */
int spoa_hip_share_compute(spoa_hip_engine* engine, spoa_hip_engine* owner);
/* No counterpart in the reference.  The engine runs its copies and kernels
 * on owner's own stream: everything the engines sharing it do runs in order
 * on that one stream (one hardware queue), e.g. one size class of a serving
 * loop whose other classes run concurrently on theirs.  Both engines must
 * be on one device, and owner must outlive engine. */
int spoa_hip_share_stream(spoa_hip_engine* engine, spoa_hip_engine* owner);
/* No counterpart in the reference.  on != 0: every pass of align_batch is
 * captured into a HIP graph (copy in, counter reset, the launches, copy
 * out) and launched as one graph on the compute stream; the stats then
 * time the whole graph as one launch (copies included).  Default off. */
int spoa_hip_set_graph_mode(spoa_hip_engine* engine, int32_t on);
/* Device memory budget for DP matrices (bytes).  A batch whose matrices do
 * not fit is split into launches that do; a single job that does not fit
 * fails with SPOA_HIP_NO_MEMORY ("insufficient memory!"). */
int spoa_hip_set_memory_budget(spoa_hip_engine* engine, uint64_t bytes);
/* The current budget (by default 80% of the device memory free when the
 * engine was created). */
uint64_t spoa_hip_memory_budget(spoa_hip_engine* engine);
/* Device memory the engine holds for DP matrices (bytes); it stays within
 * the budget except while one pass needs more than the segments already
 * held (they are then released first). */
uint64_t spoa_hip_memory_in_use(spoa_hip_engine* engine);

/* No counterpart in the reference (measurement): the alignment kernels'
 * shader-clock cycles per row phase, summed over wave 0 of every job since
 * the engine was made or last reset -- query profile, vertical pass,
 * waiting for predecessor rows, local pass, scan, its barrier, fix-up and
 * stores, the wait before them, end cell and boundary hand-off, its
 * barrier, end-cell reduction, traceback, the counts of row tiles and jobs;
 * then the traceback's parts: tile staging (cycles, count), diagonal runs
 * (cycles, steps), general steps (cycles, count), left and up gap walks
 * (cycles) -- 22 values.  Only a library built with SPOA_ROW_PHASES
 * (make -C spoa_amd phases) keeps them; otherwise SPOA_HIP_INVALID_ARGUMENT.
 * Returns the number of values written. */
int spoa_hip_row_phases(spoa_hip_engine* engine, uint64_t* out, uint32_t n, int32_t reset);

/* DP-matrix bytes one job needs (the admission footprint used by the
 * serving loop), or 0 when the job runs no DP. */
uint64_t spoa_hip_job_footprint(spoa_hip_engine* engine, uint32_t num_nodes, uint32_t len, int32_t batch);

/* ------------------------------------------------------------------
 * Device-resident windows.  No counterpart in the reference, whose graphs
 * live on the host: every window's POA graph stays in device memory for the
 * run.  Round r aligns read r of every window that has one (the engine's
 * batched alignment kernels; both strands in strand mode, the better kept
 * as src/main.cpp:285-302 does), then merges each alignment into its
 * window's graph on the device -- Graph::AddAlignment, TopologicalSort and
 * the rank-ordered view of the next round (reference src/graph.cpp:76-303)
 * -- one wavefront a window.  The finish generates every window's consensus
 * and MSA on the device (src/graph.cpp:321-516).  Reads as
 * spoa_batch_run's: window w holds reads [win_read_off[w],
 * win_read_off[w + 1]), read k is seqs[read_off[k], read_off[k + 1]), quals
 * likewise (Phred+33 weights) or NULL.  Results equal the reference's
 * window by window.  Everything runs on the engine's stream. */
typedef struct spoa_hip_windows spoa_hip_windows;
int spoa_hip_windows_create(spoa_hip_engine* engine, uint32_t n_windows, const uint32_t* win_read_off,
                            const uint64_t* read_off, const char* seqs, const char* quals, int32_t strand,
                            spoa_hip_windows** out);
int spoa_hip_windows_destroy(spoa_hip_windows* windows);
/* the number of rounds: the most reads of one window */
uint32_t spoa_hip_windows_rounds(spoa_hip_windows* windows);
/* launches round r (rounds in order, each after the previous one's wait) */
int spoa_hip_windows_submit(spoa_hip_windows* windows, uint32_t round);
/* 1 when the submitted round has finished on the device (wait will not
 * block), 0 while it runs */
int spoa_hip_windows_ready(spoa_hip_windows* windows);
/* waits for the submitted round; a failed window's message names it */
int spoa_hip_windows_wait(spoa_hip_windows* windows);
/* device times and counts of the rounds so far */
int spoa_hip_windows_stats(spoa_hip_windows* windows, spoa_hip_stats* out);
/* consensus (and with msa != 0 the MSA) of every window, held by the
 * context until it is destroyed */
int spoa_hip_windows_finish(spoa_hip_windows* windows, int32_t msa);
int spoa_hip_windows_consensus(spoa_hip_windows* windows, uint32_t window, const char** out, uint32_t* len);
/* n_rows rows of width bytes back to back (sequences in insertion order) */
int spoa_hip_windows_msa(spoa_hip_windows* windows, uint32_t window, const char** rows, uint32_t* n_rows,
                         uint32_t* width);
/* strand mode: 1 per read where the reverse complement was added; returns
 * the read count (negative status on error) */
int spoa_hip_windows_reversed(spoa_hip_windows* windows, uint32_t window, uint8_t* out, uint32_t cap);
/* Graph::rank_to_node() ids of a window after the last finished round;
 * returns V (negative status on error) */
int spoa_hip_windows_ranks(spoa_hip_windows* windows, uint32_t window, uint32_t* out, uint32_t cap);
const char* spoa_hip_windows_last_error(spoa_hip_windows* windows);

const char* spoa_hip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SPOA_HIP_H_ */
