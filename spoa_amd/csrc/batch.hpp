// Batched POA windows on one GPU: many independent windows (each its own
// graph and reads) advance one read per round; every round aligns the
// current read of every live window in one persistent-kernel launch, then
// merges the alignments into the graphs on a host thread pool.
//
// Also the plain per-call form (the drop-in AlignmentEngine::Align, one
// window and one read at a time) over the same inputs, so both forms can
// be compared on identical bytes.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "poa_graph.hpp"
#include "thread_pool.hpp"
#include "spoa/spoa.hpp"
#include "spoa_hip.h"

namespace spoa_amd {

// Flattened windows: window w holds reads [win_read_off[w], win_read_off[w+1]);
// read k is seqs[read_off[k], read_off[k+1]) (quals likewise, or null).
struct WindowSet {
  std::uint32_t n_windows = 0;
  const std::uint32_t* win_read_off = nullptr;
  const std::uint64_t* read_off = nullptr;
  const char* seqs = nullptr;
  const char* quals = nullptr;
};

struct BatchOptions {
  spoa::AlignmentType type = spoa::AlignmentType::kNW;
  std::int8_t m = 5, n = -4, g = -8, e = -6, q = -10, c = -4;
  bool strand = false;     // align both strands, keep the better (src/main.cpp:285-302)
  bool msa = true;         // also build the MSA of every window
  bool overlap = false;    // the two halves' launches may share the device (else they take turns)
  std::uint32_t max_rounds = 0;  // stop after this many rounds (0: all; the autotuner's short runs)
  bool graph = false;      // launch every pass as a captured HIP graph
  bool resident = false;   // window graphs on the device: merges, consensus and MSA there too
  bool sisd = false;       // the reference's scalar engine's alignments (SisdAlignmentEngine)
  unsigned threads = 0;    // host threads (0: all hardware threads)
  int device = 0;
};

struct WindowResult {
  std::string consensus;
  // the multiple sequence alignment as one block: msa_rows rows of msa_width
  // characters each, one after another (every read in order, then the
  // consensus row when asked for)
  std::string msa;
  std::uint32_t msa_rows = 0, msa_width = 0;
  std::vector<std::uint8_t> reversed;  // strand mode: read kept reverse-complemented
};

struct BatchStats {
  double wall_ms = 0;        // whole run
  double align_ms = 0;       // inside the alignment calls (pack + copies + kernels + unpack)
  double kernel_ms = 0;      // alignment kernels (HIP events)
  double kernel_ms16 = 0, kernel_ms32 = 0;  // of the int16 / int32 launches
  std::uint32_t graph_builds = 0, graph_updates = 0;  // graph mode: passes instantiated / updated in place
  std::uint64_t dp_bytes16 = 0, dp_bytes32 = 0;  // of the int16 / int32 jobs
  std::uint32_t launches16 = 0, launches32 = 0;
  double h2d_ms = 0, d2h_ms = 0;
  double graph_ms = 0;       // AddAlignment + topological sort
  double finish_ms = 0;      // consensus + MSA
  double pack_ms = 0, alloc_ms = 0, wait_ms = 0, unpack_ms = 0;  // inside the alignment calls
  std::uint64_t cells = 0;   // sum of V * n over the DPs
  std::uint64_t dp_bytes = 0;
  std::uint64_t align_calls = 0;
  std::uint32_t launches = 0;
  std::uint32_t rounds = 0;
};

// Long-lived engines and a host thread pool for repeated batched runs with
// one set of options (the engines keep their device buffers between runs).
// The windows are split in two halves with an engine each: while one half
// aligns on the device, the host merges the other half's alignments.
class BatchContext {
 public:
  explicit BatchContext(const BatchOptions& opt);
  const BatchOptions& options() const { return opt_; }
  spoa::AlignmentEngine* engine(int half = 0) { return engines_[half].get(); }
  BatchOptions& mutable_options() { return opt_; }
  ThreadPool& pool() { return pool_; }
  // The window graphs of a run: at least n, each cleared.  They keep their
  // storage from run to run (the context holds the largest batch's graph
  // memory until it is destroyed), so a steady stream of batches allocates
  // and frees no graph memory.
  std::vector<PoaGraph>& graphs(std::size_t n);

 private:
  BatchOptions opt_;
  std::unique_ptr<spoa::AlignmentEngine> engines_[2];
  ThreadPool pool_;
  std::vector<PoaGraph> graphs_;
};

// Batched form.  Throws std::invalid_argument with the engine's message on
// failure.  With options().resident the window graphs stay on the device
// (RunResident).
std::vector<WindowResult> RunBatched(const WindowSet& ws, BatchContext& ctx, BatchStats* stats);
std::vector<WindowResult> RunBatched(const WindowSet& ws, const BatchOptions& opt, BatchStats* stats);

// Batched form with device-resident window graphs (spoa_hip_windows_*): the
// windows split in two contiguous halves of about equal bases, an engine
// and a stream each, rounds alternating between them so that one half's
// merges run beside the other half's alignments.
std::vector<WindowResult> RunResident(const WindowSet& ws, BatchContext& ctx, BatchStats* stats);

// Autotuner of the batched launches: for each score width, tries launch
// configurations (columns per lane, waves per workgroup) on the windows
// (the first `rounds` rounds of the batched form, every candidate on the
// same bytes), keeps the one with the least device time of that width's
// launches and installs it on the context's engines.  Results are cached
// in cache_path (a JSON file; empty: no cache) under a key of the scoring
// subtype, alignment type, width, read-length and batch-size buckets, so
// a later run with the same shape installs them without measuring.
struct TuneResult {
  int width = 0;
  int k = 0, waves = 0;
  double kernel_ms = 0;      // of the chosen configuration (0: from the cache); for the
                             // int32 width, both widths' launches (they share the device)
  bool cached = false;
  bool whole_batch = false;  // the candidates ran the whole batch (no job of this width in the first rounds)
  std::vector<std::pair<std::pair<int, int>, double>> tried;  // ((k, waves), kernel ms)
};
std::vector<TuneResult> Autotune(const WindowSet& ws, BatchContext& ctx, std::uint32_t rounds,
                                 const std::string& cache_path);

// Plain per-call form: AlignmentEngine::Create(...)->Align for every read
// of every window, in order; on the calling thread, or with
// opt.threads > 1 on that many host threads with an engine each (windows
// dealt round-robin), as a multi-threaded caller of the reference runs.
// The stats' times are summed over the threads (wall_ms is the run's).
std::vector<WindowResult> RunPerCall(const WindowSet& ws, const BatchOptions& opt, BatchStats* stats);

std::string ReverseComplement(const char* s, std::size_t len);

}  // namespace spoa_amd
