// Serving loop (see serve.hpp).

#include "serve.hpp"

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <tuple>

#include "poa_graph.hpp"
#include "spoa_amd_host.hpp"

namespace spoa_amd {

namespace {

using Clock = std::chrono::steady_clock;

std::vector<std::uint32_t> QualWeights(const std::string& q) {
  std::vector<std::uint32_t> w(q.size());
  for (std::size_t i = 0; i < q.size(); ++i) w[i] = static_cast<std::uint32_t>(q[i] - 33);
  return w;
}

const char kNoMemory[] = "[spoa::SimdAlignmentEngine::Align] error: insufficient memory!";

}  // namespace

bool ServeParams::operator<(const ServeParams& o) const {
  return std::make_tuple(static_cast<int>(type), m, n, g, e, q, c) <
         std::make_tuple(static_cast<int>(o.type), o.m, o.n, o.g, o.e, o.q, o.c);
}

// One engine (device buffers) per lane for a parameter set, on its lane's
// stream.
struct Server::Engines {
  std::vector<std::unique_ptr<spoa::AlignmentEngine>> eng;
  std::vector<spoa_hip_engine*> dev;
};

// A job being served.
struct Server::Active {
  std::uint64_t id = 0;
  std::unique_ptr<ServeJob> job;
  Engines* engines = nullptr;
  PoaGraph graph;
  std::uint32_t next = 0;  // the read this round aligns
  std::uint32_t rounds = 0;
  double submit_ms = 0, admit_ms = 0;
  bool failed = false;
  std::string error;
  bool in_round = false;   // its next read takes part in the lane's current round
  int group = -1, slot = -1;
  std::vector<std::int32_t> pairs;
};

struct Server::Lane {
  struct Group {
    Engines* engines = nullptr;
    std::vector<spoa_hip_job> jobs;
  };
  std::vector<std::unique_ptr<Active>> active;
  std::vector<Group> groups;  // the round on the device (empty: none)
  bool in_round = false;      // a round was built (possibly with no DP at all)
};

Server::Server(const ServeOptions& opt) : opt_(opt), pool_(opt.threads), t0_(Clock::now()) {
  const std::size_t nl = opt_.lane_bounds.size() + 1;
  // the default budget of an engine is 80% of the free device memory
  {
    auto probe = spoa::AlignmentEngine::Create(spoa::AlignmentType::kNW, 1, -1, -1);
    auto* hip = dynamic_cast<spoa::HipAlignmentEngine*>(probe.get());
    if (!hip) throw std::invalid_argument("[spoa::hip] the serving loop needs the HIP engine");
    const double total = static_cast<double>(spoa_hip_memory_budget(hip->device())) / 0.8 * opt_.memory_fraction;
    const std::uint32_t sets = std::max<std::uint32_t>(1u, opt_.max_param_sets);
    engine_budget_ = static_cast<std::uint64_t>(total / (static_cast<double>(nl) * sets));
  }
  for (std::size_t h = 0; h < nl; ++h) {
    lanes_.emplace_back(new Lane());
    lane_owner_.push_back(spoa::AlignmentEngine::Create(spoa::AlignmentType::kNW, 1, -1, -1));
    spoa_hip_set_memory_budget(dynamic_cast<spoa::HipAlignmentEngine*>(lane_owner_.back().get())->device(), 0);
  }
  pending_.resize(nl);
  thread_ = std::thread([this] { Loop(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (thread_.joinable()) thread_.join();
}

double Server::NowMs() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0_).count(); }

std::size_t Server::LaneOf(const ServeJob& job) const {
  const std::size_t len = job.reads.empty() ? 0 : job.reads[0].size();
  std::size_t h = 0;
  while (h < opt_.lane_bounds.size() && len >= opt_.lane_bounds[h]) ++h;
  return h;
}

std::uint64_t Server::Submit(ServeJob job) {
  ServeParams& p = job.params;
  spoa::CheckCreateParams(p.type, &p.g, &p.e, &p.q, &p.c);
  if (!job.quals.empty() && job.quals.size() != job.reads.size()) {
    throw std::invalid_argument("[spoa::hip] one quality string per read");
  }
  const std::size_t h = LaneOf(job);
  std::lock_guard<std::mutex> lk(mu_);
  const std::uint64_t id = next_id_++;
  submit_ms_[id] = NowMs();
  pending_[h].emplace_back(id, std::unique_ptr<ServeJob>(new ServeJob(std::move(job))));
  ++n_pending_;
  ++in_flight_;
  ++stats_.submitted;
  cv_.notify_all();
  return id;
}

bool Server::Wait(std::uint64_t id, ServeResult* out, double timeout_s) {
  std::unique_lock<std::mutex> lk(mu_);
  auto ready = [&] { return done_.count(id) != 0; };
  if (timeout_s < 0) {
    done_cv_.wait(lk, ready);
  } else if (!done_cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
    return false;
  }
  auto it = done_.find(id);
  if (out) *out = std::move(it->second);
  done_.erase(it);
  return true;
}

void Server::Drain() {
  std::unique_lock<std::mutex> lk(mu_);
  done_cv_.wait(lk, [&] { return in_flight_ == 0; });
}

ServeStats Server::Stats() {
  std::lock_guard<std::mutex> lk(mu_);
  return stats_;
}

Server::Engines* Server::EnginesFor(const ServeParams& p) {
  auto it = engines_.find(p);
  if (it != engines_.end()) return it->second.get();
  if (engines_.size() >= std::max<std::uint32_t>(1u, opt_.max_param_sets)) {
    // retire a parameter set no active job uses
    for (auto jt = engines_.begin(); jt != engines_.end(); ++jt) {
      bool used = false;
      for (auto& lane : lanes_) {
        for (auto& a : lane->active) used = used || a->engines == jt->second.get();
      }
      if (!used) {
        engines_.erase(jt);
        break;
      }
    }
    if (engines_.size() >= std::max<std::uint32_t>(1u, opt_.max_param_sets)) return nullptr;
  }
  std::unique_ptr<Engines> en(new Engines());
  for (std::size_t h = 0; h < lanes_.size(); ++h) {
    en->eng.push_back(spoa::AlignmentEngine::Create(p.type, p.m, p.n, p.g, p.e, p.q, p.c));
    en->dev.push_back(dynamic_cast<spoa::HipAlignmentEngine*>(en->eng.back().get())->device());
    spoa_hip_set_memory_budget(en->dev.back(), engine_budget_);
    auto* owner = dynamic_cast<spoa::HipAlignmentEngine*>(lane_owner_[h].get())->device();
    if (spoa_hip_share_stream(en->dev.back(), owner) != SPOA_HIP_OK) throw std::invalid_argument(spoa_hip_last_error());
  }
  Engines* raw = en.get();
  engines_[p] = std::move(en);
  return raw;
}

bool Server::Step(std::size_t h) {
  Lane& lane = *lanes_[h];
  // the lane's round must have finished on the device
  for (auto& grp : lane.groups) {
    if (spoa_hip_align_batch_ready(grp.engines->dev[h]) == 0) return false;
  }
  bool any = lane.in_round;

  // 1. collect the round and merge it
  for (auto& grp : lane.groups) {
    (void)spoa_hip_align_batch_wait(grp.engines->dev[h]);  // per-job statuses below
    spoa_hip_stats hs{};
    spoa_hip_last_stats(grp.engines->dev[h], &hs);
    std::lock_guard<std::mutex> lk(mu_);
    stats_.kernel_ms += hs.kernel_ms;
    stats_.launches += hs.launches;
    stats_.cells += hs.cells;
    stats_.dp_bytes += hs.dp_bytes;
    stats_.align_calls += grp.jobs.size();
  }
  std::vector<Active*> merging;
  for (auto& a : lane.active) {
    if (a->in_round) merging.push_back(a.get());
  }
  pool_.For(merging.size(), [&](std::size_t x) {
    Active& a = *merging[x];
    a.in_round = false;
    const std::string& read = a.job->reads[a.next];
    const std::int32_t* pairs = nullptr;
    std::uint32_t np = 0;
    if (a.group >= 0) {
      const spoa_hip_job& j = lane.groups[a.group].jobs[a.slot];
      if (j.status != SPOA_HIP_OK) {
        a.failed = true;
        a.error = j.status == SPOA_HIP_NO_MEMORY ? std::string(kNoMemory)
                                                 : "[spoa::hip] alignment failed (status " + std::to_string(j.status) + ")";
        return;
      }
      pairs = j.pairs;
      np = j.n_pairs;
    }
    try {
      if (a.job->quals.empty()) {
        static thread_local std::vector<std::uint32_t> ones;
        if (ones.size() < read.size()) ones.assign(read.size(), 1u);
        a.graph.AddAlignment(pairs, np, read.data(), static_cast<std::uint32_t>(read.size()), ones.data());
      } else {
        const std::vector<std::uint32_t> w = QualWeights(a.job->quals[a.next]);
        a.graph.AddAlignment(pairs, np, read.data(), static_cast<std::uint32_t>(read.size()), w.data());
      }
      ++a.next;
      ++a.rounds;
    } catch (const std::exception& ex) {
      a.failed = true;
      a.error = ex.what();
    }
  });
  lane.groups.clear();
  lane.in_round = false;

  // 2. finish completed jobs
  std::vector<std::unique_ptr<Active>> finished;
  {
    std::vector<std::unique_ptr<Active>> keep;
    for (auto& a : lane.active) {
      if (a->failed || a->next >= a->job->reads.size()) {
        finished.push_back(std::move(a));
      } else {
        keep.push_back(std::move(a));
      }
    }
    lane.active.swap(keep);
  }
  if (!finished.empty()) {
    std::vector<ServeResult> res(finished.size());
    pool_.For(finished.size(), [&](std::size_t x) {
      Active& a = *finished[x];
      ServeResult& r = res[x];
      r.submit_ms = a.submit_ms;
      r.admit_ms = a.admit_ms;
      r.rounds = a.rounds;
      r.lane = static_cast<std::uint32_t>(h);
      if (a.failed) {
        r.status = -1;
        r.error = a.error;
        return;
      }
      r.consensus = a.graph.GenerateConsensus();
      if (a.job->msa) r.msa = a.graph.GenerateMultipleSequenceAlignment();
    });
    std::lock_guard<std::mutex> lk(mu_);
    const double now = NowMs();
    for (std::size_t x = 0; x < finished.size(); ++x) {
      res[x].done_ms = now;
      if (res[x].status != 0) {
        ++stats_.failed;
      } else {
        ++stats_.completed;
      }
      done_[finished[x]->id] = std::move(res[x]);
      --in_flight_;
    }
    done_cv_.notify_all();
  }

  // 3. admit waiting jobs of this lane's class into free slots
  {
    std::unique_lock<std::mutex> lk(mu_);
    std::size_t total = 0;
    for (auto& l : lanes_) total += l->active.size();
    auto& queue = pending_[h];
    while (!queue.empty() && total < opt_.max_active) {
      auto item = std::move(queue.front());
      queue.pop_front();
      --n_pending_;
      any = true;
      std::unique_ptr<Active> a(new Active());
      a->id = item.first;
      a->job = std::move(item.second);
      a->submit_ms = submit_ms_[a->id];
      submit_ms_.erase(a->id);
      a->admit_ms = NowMs();
      lk.unlock();
      try {
        a->engines = EnginesFor(a->job->params);
        if (!a->engines) {
          a->failed = true;
          a->error = "[spoa::hip] too many scoring parameter sets in flight";
        }
      } catch (const std::exception& ex) {
        a->failed = true;
        a->error = ex.what();
      }
      lk.lock();
      if (a->failed || a->job->reads.empty()) {
        ServeResult r;
        r.status = a->failed ? -1 : 0;
        r.error = a->error;
        r.submit_ms = a->submit_ms;
        r.admit_ms = a->admit_ms;
        r.done_ms = NowMs();
        r.lane = static_cast<std::uint32_t>(h);
        if (a->failed) {
          ++stats_.failed;
        } else {
          ++stats_.completed;
        }
        done_[a->id] = std::move(r);
        --in_flight_;
        done_cv_.notify_all();
        continue;
      }
      lane.active.push_back(std::move(a));
      ++total;
    }
    stats_.max_active_seen = std::max<std::uint32_t>(stats_.max_active_seen, static_cast<std::uint32_t>(total));
  }

  // 4. build and submit the lane's next round: oldest jobs first, while the
  // DP footprints of the round fit each engine's budget
  std::map<Engines*, std::uint64_t> used;
  std::map<Engines*, int> group_of;
  std::uint64_t deferred = 0;
  for (auto& ap : lane.active) {
    Active& a = *ap;
    const std::string& read = a.job->reads[a.next];
    a.group = -1;
    a.slot = -1;
    a.in_round = true;
    const std::uint32_t V = a.graph.num_nodes();
    if (V == 0 || read.empty()) continue;  // no DP: the read is merged as it is
    const std::uint64_t fp =
        spoa_hip_job_footprint(a.engines->dev[h], V, static_cast<std::uint32_t>(read.size()), 1);
    std::uint64_t& u = used[a.engines];
    if (u > 0 && u + fp > engine_budget_) {
      a.in_round = false;  // waits for memory
      ++deferred;
      continue;
    }
    u += fp;
    auto git = group_of.find(a.engines);
    if (git == group_of.end()) {
      git = group_of.emplace(a.engines, static_cast<int>(lane.groups.size())).first;
      lane.groups.emplace_back();
      lane.groups.back().engines = a.engines;
    }
    auto& grp = lane.groups[git->second];
    spoa_hip_job j{};
    j.graph = a.graph.View();
    j.seq = read.data();
    j.len = static_cast<std::uint32_t>(read.size());
    j.cap = V + j.len;
    a.pairs.resize(2ull * j.cap);
    j.pairs = a.pairs.data();
    a.group = git->second;
    a.slot = static_cast<int>(grp.jobs.size());
    grp.jobs.push_back(j);
  }
  for (auto& grp : lane.groups) {
    // per-job statuses carry any failure; a launched pass is always
    // collected before the jobs' buffers change
    (void)spoa_hip_align_batch_submit(grp.engines->dev[h], grp.jobs.data(),
                                      static_cast<std::uint32_t>(grp.jobs.size()));
  }
  lane.in_round = !lane.active.empty();
  {
    std::lock_guard<std::mutex> lk(mu_);
    stats_.deferred += deferred;
    if (lane.in_round) ++stats_.rounds;
  }
  return any;
}

void Server::Loop() {
  for (;;) {
    bool progressed = false, busy = false;
    for (std::size_t h = 0; h < lanes_.size(); ++h) {
      progressed = Step(h) || progressed;
      busy = busy || lanes_[h]->in_round;
    }
    std::unique_lock<std::mutex> lk(mu_);
    if (!busy && n_pending_ == 0) {
      if (stop_) return;
      cv_.wait(lk, [&] { return stop_ || n_pending_ > 0; });
      if (stop_ && n_pending_ == 0) return;
    } else if (!progressed) {
      // rounds on the device: poll again shortly (or on a new arrival)
      cv_.wait_for(lk, std::chrono::microseconds(200));
    }
  }
}

}  // namespace spoa_amd
