// Serving loop: consensus jobs (a window of reads with its own scoring
// parameters) arrive at any time and are served by continuous batching on
// one GPU.
//
// Jobs are sorted into lanes by size class (the length of their first
// read).  A lane advances all its jobs by one read per round: one batched
// launch per scoring parameter set aligns the next read of every job of the
// lane, then the host merges the alignments into the jobs' graphs (thread
// pool), finishes completed jobs (consensus, MSA) and admits waiting jobs of
// its class into the free slots.  Each lane runs on its own stream (its own
// hardware queue) and a scheduler thread polls the lanes, so a lane of
// short jobs cycles at its own pace while a lane of long alignments is
// still on the device, and their launches share the GPU.  Device memory is admitted per round: a
// job's next alignment joins a round only while the DP-matrix footprints of
// the round's alignments fit the engine's budget (spoa_hip_job_footprint);
// otherwise it waits for the next round.
//
// Every job's result is the same as running its window alone: the graph
// of a job sees only its own reads, in order, and every alignment is
// bit-exact.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "spoa/spoa.hpp"

namespace spoa_amd {

struct ServeOptions {
  unsigned threads = 0;          // host threads of the merge pool (0: all)
  int device = 0;
  std::uint32_t max_active = 4096;  // job slots (both lanes together)
  std::uint32_t max_param_sets = 2; // distinct scoring parameter sets served at once
  double memory_fraction = 0.6;  // of the device memory free at start, for DP matrices
  // size classes: a job whose first read is shorter than lane_bounds[i]
  // (and not shorter than the previous bound) goes to lane i; longer ones
  // to the last lane
  std::vector<std::uint32_t> lane_bounds = {700, 2000, 5000};
};

struct ServeParams {
  spoa::AlignmentType type = spoa::AlignmentType::kNW;
  std::int8_t m = 5, n = -4, g = -8, e = -6, q = -10, c = -4;
  bool operator<(const ServeParams& o) const;
};

struct ServeJob {
  ServeParams params;
  std::vector<std::string> reads;
  std::vector<std::string> quals;  // empty, or one per read
  bool msa = true;
};

struct ServeResult {
  int status = 0;                  // 0 ok, -1 failed (error says why)
  std::string error;
  std::string consensus;
  std::vector<std::string> msa;
  double submit_ms = 0;            // times since the server started
  double admit_ms = 0;
  double done_ms = 0;
  std::uint32_t rounds = 0;        // rounds the job took part in
  std::uint32_t lane = 0;          // the size-class lane it ran in
};

struct ServeStats {
  std::uint64_t submitted = 0, completed = 0, failed = 0;
  std::uint64_t rounds = 0, launches = 0, align_calls = 0, deferred = 0;
  std::uint64_t cells = 0, dp_bytes = 0;
  double kernel_ms = 0;
  std::uint32_t max_active_seen = 0;
};

class Server {
 public:
  explicit Server(const ServeOptions& opt);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // Thread-safe.  Returns the job id; throws std::invalid_argument when the
  // parameters are invalid (AlignmentEngine::Create's checks) or the server
  // already serves max_param_sets other parameter sets.
  std::uint64_t Submit(ServeJob job);
  // Blocks until the job is done (timeout_s < 0: no limit); false on timeout.
  // The result is handed over once.
  bool Wait(std::uint64_t id, ServeResult* out, double timeout_s = -1);
  // Blocks until every submitted job is done.
  void Drain();
  ServeStats Stats();
  double NowMs() const;

 private:
  struct Active;
  struct Lane;
  struct Engines;

  void Loop();
  Engines* EnginesFor(const ServeParams& p);

  ServeOptions opt_;
  ThreadPool pool_;
  std::chrono::steady_clock::time_point t0_;
  std::mutex mu_;
  std::condition_variable cv_;       // scheduler wake-up
  std::condition_variable done_cv_;  // job completions
  std::vector<std::deque<std::pair<std::uint64_t, std::unique_ptr<ServeJob>>>> pending_;  // per lane
  std::uint64_t n_pending_ = 0;
  std::map<std::uint64_t, ServeResult> done_;
  std::map<std::uint64_t, double> submit_ms_;
  std::uint64_t next_id_ = 1;
  std::uint64_t in_flight_ = 0;      // submitted, not yet done
  bool stop_ = false;
  ServeStats stats_;

  // scheduler-thread state.  Every lane owns one stream (held by an engine
  // that never aligns, declared first so that it is destroyed last); the
  // engines of the lane run their copies and kernels on it in order.
  std::vector<std::unique_ptr<spoa::AlignmentEngine>> lane_owner_;
  std::map<ServeParams, std::unique_ptr<Engines>> engines_;
  std::uint64_t engine_budget_ = 0;
  std::vector<std::unique_ptr<Lane>> lanes_;
  std::thread thread_;

  std::size_t LaneOf(const ServeJob& job) const;
  bool Step(std::size_t h);  // one lane: collect, finish, admit, submit; false if nothing to do
};

}  // namespace spoa_amd
