// Types shared by the alignment kernels (poa_kernels.hpp) and the host code
// that launches them: launch parameters, job records, the layout of a
// job's graph block and the kernel arguments.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spoa_hip_dev {

constexpr int kMaxWaves = 16;

// Scoring parameters and scan constants for one launch.
struct KParams {
  int32_t type;     // 0 SW, 1 NW, 2 OV
  int32_t m, mis, g, e, q, c;
  int32_t pad;      // padding-column score
  int32_t ninf;     // the score type's minus infinity (row 0 of F and O)
  int32_t decay;    // linear: K*g; affine: K*e (per-lane decay of the scan)
  int32_t mpow[8][4];  // convex: M^(2^t) for t = 0..6 (row-major 2x2), [7] unused
  int32_t mseg[4];     // convex: M of one L-column segment
  int32_t sisd;        // 1: the reference's scalar engine's end cell and traceback
};

// One (graph, read) job.  Offsets are in elements of the respective buffer.
struct DevJob {
  uint64_t mat_addr;   // device address of the job's DP matrices
  uint32_t graph_off;  // 32-bit words into the graph buffer
  uint32_t seq_off;    // bytes into the sequence buffer
  uint32_t V;          // graph nodes
  uint32_t n;          // read length
  uint32_t W;          // reference row width ceil(n / L) * L
  uint32_t stride;     // row stride of one matrix, a multiple of K
  uint32_t pair_off;   // pairs into the pair buffer
  uint32_t cap;        // pair capacity (V + n suffices)
  uint32_t nslots;     // distinct row bases with an LDS profile (0: formed per row)
  uint32_t slot_chars; // the base of each profile slot, a byte each
  // Split jobs: the number of row tiles, each aligned by its own workgroup
  // (0 or 1: one workgroup runs whole rows), and the job's hand-off words
  // in KArgs::xfer (64-bit words): per boundary b = 1 .. split - 1 and row
  // r, words 3 * ((b - 1) * (V + 1) + r) .. + 2 hold the H of tile b - 1's
  // last column and the carry (two words) entering tile b; after them two
  // words a tile with its end-cell candidate (score, row).
  uint32_t split;
  uint32_t xfer_off;
  uint32_t pad0, pad1;
};

// Graph block of a job (32-bit words from graph_off, 16-byte aligned):
//   rows[V] (uint4 per rank: decoded base | sink << 8 | profile slot << 9
//   (7: none) | predecessor count << 12 (1 for a source, whose predecessor
//   is row 0), first predecessor row, offset of the predecessor list in
//   preds, fcH of the row),
//   node_id[V], pred_off[V + 1] (relative to preds), fcH[V + 1],
//   fcF[V + 1], fcO[V + 1], preds[...] (predecessor rows, rank + 1, in
//   Node::inedges order).
// The graph block is never written while a kernel runs, so it is read
// through the constant address space: loads of one address for the whole
// wave become scalar loads, which are counted apart from the vector stores
// of the rows (waiting for a record or a predecessor index never waits for
// those stores).
template <typename T>
using Cst = const __attribute__((address_space(4))) T;
typedef uint32_t U32x4 __attribute__((ext_vector_type(4)));

struct GraphView {
  Cst<U32x4>* rows;
  Cst<int32_t>* node_id;
  Cst<uint32_t>* pred_off;
  Cst<int32_t>* fcH;
  Cst<int32_t>* fcF;
  Cst<int32_t>* fcO;
  Cst<uint32_t>* preds;
};

__device__ __forceinline__ GraphView MakeView(const uint32_t* base_g, uint32_t V) {
  Cst<uint32_t>* base = (Cst<uint32_t>*)base_g;
  GraphView g;
  g.rows = reinterpret_cast<Cst<U32x4>*>(base);
  g.node_id = reinterpret_cast<Cst<int32_t>*>(base + 4 * V);
  g.pred_off = base + 5 * V;
  g.fcH = reinterpret_cast<Cst<int32_t>*>(base + 6 * V + 1);
  g.fcF = g.fcH + (V + 1);
  g.fcO = g.fcF + (V + 1);
  g.preds = base + 9 * V + 4;
  return g;
}

__device__ __forceinline__ uint4 LoadRow(const GraphView& g, uint32_t r) {
  const U32x4 v = g.rows[r];
  return make_uint4(v.x, v.y, v.z, v.w);
}

struct KArgs {
  KParams prm;
  const DevJob* jobs;
  uint32_t n_jobs;
  uint32_t* queue;           // job counter, zero at launch
  const uint32_t* graphs;
  const uint8_t* seqs;
  int32_t* pairs;            // (node id, column) pairs in traceback order
  int32_t* results;          // per job: n_pairs, score, has_score, status
  uint32_t lds_bytes;        // dynamic LDS of the launch (at least the query profile's)
  uint32_t prio_jobs;        // the first prio_jobs jobs of the queue run at raised wave priority
  unsigned long long* phases;  // row-phase cycle counters (builds with SPOA_ROW_PHASES; else unused)
  // Queue items of a launch with split jobs: per item the job (index into
  // jobs) and the row tile it aligns, the tiles of a job consecutive and in
  // order (a tile only waits for items taken from the queue before it).
  // Null: item i is job i, whole rows.
  const uint2* items;
  uint32_t n_items;
  uint32_t epoch;       // tag of this launch's hand-off words (never 0)
  uint32_t wait_ticks;  // longest wait for another tile (100 MHz clock ticks)
  unsigned long long* xfer;  // hand-off words; word 0: the launch's abort word
};

enum DevStatus : int32_t { kDevOk = 0, kDevCapacity = 6, kDevInconsistent = 7, kDevTimeout = 8 };

// Hand-off words of split jobs start here (words 0-1: the abort word).
constexpr uint32_t kXferFirst = 2;

}  // namespace spoa_hip_dev
