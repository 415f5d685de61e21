// Batched and per-call POA window drivers (see batch.hpp).

#include "batch.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <chrono>
#include <array>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "poa_graph.hpp"
#include "spoa_amd_host.hpp"

namespace spoa_amd {

namespace {

using Clock = std::chrono::steady_clock;

double Ms(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

// The strand-ambiguous caller's complement (src/main.cpp:285-302 uses
// biosoup::Sequence::ReverseAndComplement, a dependency the reference does
// not vendor): upper-case, then the Watson-Crick / IUPAC complement (A-T,
// C-G, U-A, R-Y, K-M, B-V, D-H; S and W map to themselves); every other
// character is kept as it is.  The oracle and the test harness follow the
// same rule.
char Complement(char b) {
  switch (b) {
    case 'A': case 'a': return 'T';
    case 'C': case 'c': return 'G';
    case 'G': case 'g': return 'C';
    case 'T': case 't': return 'A';
    case 'U': case 'u': return 'A';
    case 'R': case 'r': return 'Y';
    case 'Y': case 'y': return 'R';
    case 'K': case 'k': return 'M';
    case 'M': case 'm': return 'K';
    case 'B': case 'b': return 'V';
    case 'D': case 'd': return 'H';
    case 'H': case 'h': return 'D';
    case 'V': case 'v': return 'B';
    case 'S': case 's': return 'S';
    case 'W': case 'w': return 'W';
    default: return b;
  }
}

std::vector<std::uint32_t> Weights(const char* qual, std::size_t len) {
  std::vector<std::uint32_t> w(len);
  for (std::size_t i = 0; i < len; ++i) w[i] = static_cast<std::uint32_t>(qual[i] - 33);
  return w;
}

}  // namespace

std::string ReverseComplement(const char* s, std::size_t len) {
  std::string out(len, 'N');
  for (std::size_t i = 0; i < len; ++i) out[len - 1 - i] = Complement(s[i]);
  return out;
}

// --------------------------------------------------------------- drivers --

BatchContext::BatchContext(const BatchOptions& opt) : opt_(opt), pool_(opt.threads) {
  for (auto& e : engines_) e = spoa::CreateEngine(opt.type, opt.m, opt.n, opt.g, opt.e, opt.q, opt.c, opt.sisd);
  // the halves share the device memory: each engine gets half the budget;
  // their kernels take turns on one compute stream (engines_[0]'s, which is
  // destroyed last) while copies and host merging overlap.  Device-resident
  // runs use the first engine alone (RunResident), which keeps the budget.
  spoa_hip_engine* dev[2] = {nullptr, nullptr};
  for (int h = 0; h < 2; ++h) {
    auto* hip = dynamic_cast<spoa::HipAlignmentEngine*>(engines_[h].get());
    if (!hip) throw std::invalid_argument("[spoa::hip] the batched form needs the HIP engine");
    dev[h] = hip->device();
    if (!opt.resident || h == 1) spoa_hip_set_memory_budget(dev[h], spoa_hip_memory_budget(dev[h]) / 2);
  }
  if (!opt.overlap && spoa_hip_share_compute(dev[1], dev[0]) != SPOA_HIP_OK) {
    throw std::invalid_argument(spoa_hip_last_error());
  }
  if (opt.graph) {
    for (auto* d : dev) spoa_hip_set_graph_mode(d, 1);
  }
}

std::vector<PoaGraph>& BatchContext::graphs(std::size_t n) {
  if (graphs_.size() < n) graphs_.resize(n);
  pool_.For(n, [&](std::size_t w) { graphs_[w].Clear(); });
  return graphs_;
}

std::vector<WindowResult> RunBatched(const WindowSet& ws, const BatchOptions& opt, BatchStats* stats) {
  BatchContext ctx(opt);
  return RunBatched(ws, ctx, stats);
}

namespace {

// Grow-only int32 buffer that neither initialises nor copies its contents
// (a job's pair buffer grows with its graph every round; a vector would
// copy and zero-fill it each time).
struct PairBuf {
  std::unique_ptr<std::int32_t[]> p;
  std::size_t cap = 0;
  std::int32_t* Get(std::size_t n) {
    if (n > cap) {
      cap = n + n / 2;
      p.reset(new std::int32_t[cap]);
    }
    return p.get();
  }
};

// One half of the windows and its engine: the jobs of the current round
// and where their results go.
struct Half {
  spoa_hip_engine* dev = nullptr;
  std::vector<std::uint32_t> windows;  // window ids of this half
  std::vector<std::uint32_t> live;     // windows with a read in the current round
  std::vector<spoa_hip_job> jobs;
  std::vector<std::int32_t> job_of;    // per (live index, strand): job index or -1
  std::vector<PairBuf> pair_buf;
  std::uint32_t round = 0;
  bool pending = false;
};

}  // namespace

std::vector<WindowResult> RunResident(const WindowSet& ws, BatchContext& ctx, BatchStats* stats) {
  const auto t0 = Clock::now();
  BatchStats st;
  const BatchOptions& opt = ctx.options();
  const std::uint32_t W = ws.n_windows;
  std::vector<WindowResult> out(W);
  // All windows form one set on the first engine's stream: every round is
  // one alignment launch over all windows, then the merge kernel.  (Two
  // halves on two streams overlapped their rounds only some of the time --
  // 4,096 windows of 10-50 reads of 500 bp took 367 or 517 ms from run to
  // run, against 412 ms for one set every time -- and no host merge is left
  // to overlap.)  The first engine holds the context's
  // memory budget in this mode.
  const std::uint32_t split = W;
  struct Part {
    spoa_hip_engine* dev = nullptr;
    spoa_hip_windows* win = nullptr;
    std::uint32_t first = 0, count = 0, rounds = 0;
    bool pending = false;
  } part[2];
  auto fail = [&](spoa_hip_windows* x, int rc) {
    std::string msg = x ? spoa_hip_windows_last_error(x) : "";
    if (msg.empty()) msg = spoa_hip_last_error();
    for (auto& p : part) {
      if (p.win) spoa_hip_windows_destroy(p.win);
      p.win = nullptr;
    }
    throw std::invalid_argument(msg.empty() ? "[spoa::hip] device-resident windows failed (" + std::to_string(rc) + ")"
                                            : msg);
  };
  for (int h = 0; h < 2; ++h) {
    auto* hip = dynamic_cast<spoa::HipAlignmentEngine*>(ctx.engine(h));
    if (!hip) throw std::invalid_argument("[spoa::hip] the alignment engine is not the HIP engine");
    Part& p = part[h];
    p.dev = hip->device();
    p.first = h == 0 ? 0 : split;
    p.count = h == 0 ? split : W - split;
    if (p.count == 0) continue;
    const int rc = spoa_hip_windows_create(p.dev, p.count, ws.win_read_off + p.first, ws.read_off, ws.seqs, ws.quals,
                                           opt.strand ? 1 : 0, &p.win);
    if (rc != SPOA_HIP_OK) fail(nullptr, rc);
    p.rounds = spoa_hip_windows_rounds(p.win);
    if (opt.max_rounds > 0) p.rounds = std::min(p.rounds, opt.max_rounds);
  }
  const std::uint32_t rounds = std::max(part[0].rounds, part[1].rounds);
  for (std::uint32_t r = 0; r < rounds; ++r) {
    for (auto& p : part) {
      if (!p.win) continue;
      if (p.pending) {
        const auto ta = Clock::now();
        const int rc = spoa_hip_windows_wait(p.win);
        st.wait_ms += Ms(ta, Clock::now());
        p.pending = false;
        if (rc != SPOA_HIP_OK) fail(p.win, rc);
      }
      if (r < p.rounds) {
        const auto ta = Clock::now();
        const int rc = spoa_hip_windows_submit(p.win, r);
        st.align_ms += Ms(ta, Clock::now());
        if (rc != SPOA_HIP_OK) fail(p.win, rc);
        p.pending = true;
      }
    }
    ++st.rounds;
  }
  for (auto& p : part) {
    if (p.win && p.pending) {
      const auto ta = Clock::now();
      const int rc = spoa_hip_windows_wait(p.win);
      st.wait_ms += Ms(ta, Clock::now());
      if (rc != SPOA_HIP_OK) fail(p.win, rc);
    }
  }
  const auto tf = Clock::now();
  const bool finish = opt.max_rounds == 0;
  for (auto& p : part) {
    if (!p.win) continue;
    if (finish) {
      const int rc = spoa_hip_windows_finish(p.win, opt.msa ? 1 : 0);
      if (rc != SPOA_HIP_OK) fail(p.win, rc);
      // the windows' strings out of the finished set, on the host threads
      std::vector<char> bad(p.count, 0);
      ctx.pool().For(p.count, [&](std::size_t i) {
        WindowResult& o = out[p.first + i];
        const std::uint32_t wi = static_cast<std::uint32_t>(i);
        const char* c = nullptr;
        std::uint32_t len = 0;
        spoa_hip_windows_consensus(p.win, wi, &c, &len);
        o.consensus.assign(c, len);
        if (opt.msa) {
          const char* rows = nullptr;
          std::uint32_t nr = 0, width = 0;
          spoa_hip_windows_msa(p.win, wi, &rows, &nr, &width);
          o.msa.assign(rows, static_cast<std::size_t>(nr) * width);
          o.msa_rows = nr;
          o.msa_width = width;
        }
        if (opt.strand) {
          const std::uint32_t n = ws.win_read_off[p.first + i + 1] - ws.win_read_off[p.first + i];
          o.reversed.resize(n);
          if (spoa_hip_windows_reversed(p.win, wi, o.reversed.data(), n) < 0) bad[i] = 1;
        }
      });
      if (std::find(bad.begin(), bad.end(), 1) != bad.end()) fail(p.win, -1);
    }
    spoa_hip_stats hs{};
    spoa_hip_windows_stats(p.win, &hs);
    st.kernel_ms += hs.kernel_ms;
    st.kernel_ms16 += hs.kernel_ms16;
    st.kernel_ms32 += hs.kernel_ms32;
    st.graph_builds += hs.graph_builds;
    st.graph_updates += hs.graph_updates;
    st.dp_bytes16 += hs.dp_bytes16;
    st.dp_bytes32 += hs.dp_bytes32;
    st.launches16 += hs.launches16;
    st.launches32 += hs.launches32;
    st.launches += hs.launches;
    st.cells += hs.cells;
    st.dp_bytes += hs.dp_bytes;
    st.align_calls += hs.jobs;
    st.pack_ms += hs.host_ms;
    st.alloc_ms += hs.alloc_ms;
    st.graph_ms += hs.merge_ms;  // device time of the merges
    spoa_hip_windows_destroy(p.win);
    p.win = nullptr;
  }
  st.finish_ms = Ms(tf, Clock::now());
  st.wall_ms = Ms(t0, Clock::now());
  if (stats) *stats = st;
  return out;
}

std::vector<WindowResult> RunBatched(const WindowSet& ws, BatchContext& ctx, BatchStats* stats) {
  if (ctx.options().resident) return RunResident(ws, ctx, stats);
  const auto t0 = Clock::now();
  BatchStats st;
  const BatchOptions& opt = ctx.options();
  ThreadPool& pool = ctx.pool();
  const std::uint32_t W = ws.n_windows;
  std::vector<PoaGraph>& graphs = ctx.graphs(W);
  std::vector<WindowResult> out(W);
  std::vector<std::string> rc_seq(W);
  const int per_read = opt.strand ? 2 : 1;
  std::uint64_t max_len = 0;
  for (std::uint64_t k = 0; k + 1 < ws.win_read_off[W] + 1ull; ++k) {
    max_len = std::max<std::uint64_t>(max_len, ws.read_off[k + 1] - ws.read_off[k]);
  }
  const std::vector<std::uint32_t> ones(max_len, 1u);

  // halves balanced by total read bases, largest windows first
  Half half[2];
  for (int h = 0; h < 2; ++h) {
    auto* hip = dynamic_cast<spoa::HipAlignmentEngine*>(ctx.engine(h));
    if (!hip) throw std::invalid_argument("[spoa::hip] the alignment engine is not the HIP engine");
    half[h].dev = hip->device();
  }
  std::uint32_t max_reads = 0;
  {
    std::vector<std::pair<std::uint64_t, std::uint32_t>> size(W);
    for (std::uint32_t w = 0; w < W; ++w) {
      const std::uint32_t a = ws.win_read_off[w], b = ws.win_read_off[w + 1];
      max_reads = std::max(max_reads, b - a);
      size[w] = {ws.read_off[b] - ws.read_off[a], w};
    }
    std::sort(size.begin(), size.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
    std::uint64_t load[2] = {0, 0};
    for (const auto& sw : size) {
      const int h = load[0] <= load[1] ? 0 : 1;
      half[h].windows.push_back(sw.second);
      load[h] += sw.first;
    }
    for (auto& hf : half) std::sort(hf.windows.begin(), hf.windows.end());
  }

  auto read_of = [&](std::uint32_t w, std::uint32_t r, std::uint32_t* len) {
    const std::uint64_t k = ws.win_read_off[w] + r;
    *len = static_cast<std::uint32_t>(ws.read_off[k + 1] - ws.read_off[k]);
    return ws.seqs + ws.read_off[k];
  };

  // jobs of round r for half h (its graphs hold rounds < r)
  auto build = [&](Half& hf, std::uint32_t r) {
    hf.round = r;
    hf.live.clear();
    for (std::uint32_t w : hf.windows) {
      if (ws.win_read_off[w] + r < ws.win_read_off[w + 1]) hf.live.push_back(w);
    }
    if (opt.strand) {
      pool.For(hf.live.size(), [&](std::size_t x) {
        const std::uint32_t w = hf.live[x];
        std::uint32_t len;
        const char* s = read_of(w, r, &len);
        rc_seq[w] = ReverseComplement(s, len);
      });
    }
    hf.jobs.clear();
    hf.job_of.assign(hf.live.size() * per_read, -1);
    if (hf.pair_buf.size() < hf.live.size() * per_read) hf.pair_buf.resize(hf.live.size() * per_read);
    for (std::size_t x = 0; x < hf.live.size(); ++x) {
      const std::uint32_t w = hf.live[x];
      std::uint32_t len;
      const char* s = read_of(w, r, &len);
      if (graphs[w].num_nodes() == 0 || len == 0) continue;
      for (int d = 0; d < per_read; ++d) {
        spoa_hip_job j{};
        j.graph = graphs[w].View();
        j.seq = d == 0 ? s : rc_seq[w].data();
        j.len = len;
        const std::uint32_t cap = j.graph.num_nodes + len;
        j.pairs = hf.pair_buf[x * per_read + d].Get(2ull * cap);
        j.cap = cap;
        hf.job_of[x * per_read + d] = static_cast<std::int32_t>(hf.jobs.size());
        hf.jobs.push_back(j);
      }
    }
  };

  auto submit = [&](Half& hf) {
    hf.pending = false;
    if (hf.jobs.empty()) return;
    const auto ta = Clock::now();
    const int rc = spoa_hip_align_batch_submit(hf.dev, hf.jobs.data(), static_cast<std::uint32_t>(hf.jobs.size()));
    st.align_ms += Ms(ta, Clock::now());
    if (rc != SPOA_HIP_OK) throw std::invalid_argument(spoa_hip_last_error());
    hf.pending = true;
  };

  auto wait = [&](Half& hf) {
    if (!hf.pending) return;
    hf.pending = false;
    const auto ta = Clock::now();
    const int rc = spoa_hip_align_batch_wait(hf.dev);
    st.align_ms += Ms(ta, Clock::now());
    if (rc != SPOA_HIP_OK) throw std::invalid_argument(spoa_hip_last_error());
    spoa_hip_stats hs{};
    spoa_hip_last_stats(hf.dev, &hs);
    st.kernel_ms += hs.kernel_ms;
    st.kernel_ms16 += hs.kernel_ms16;
    st.kernel_ms32 += hs.kernel_ms32;
    st.graph_builds += hs.graph_builds;
    st.graph_updates += hs.graph_updates;
    st.dp_bytes16 += hs.dp_bytes16;
    st.dp_bytes32 += hs.dp_bytes32;
    st.launches16 += hs.launches16;
    st.launches32 += hs.launches32;
    st.h2d_ms += hs.h2d_ms;
    st.d2h_ms += hs.d2h_ms;
    st.pack_ms += hs.host_ms;
    st.alloc_ms += hs.alloc_ms;
    st.wait_ms += hs.wait_ms;
    st.unpack_ms += hs.unpack_ms;
    st.cells += hs.cells;
    st.dp_bytes += hs.dp_bytes;
    st.launches += hs.launches;
    st.align_calls += hf.jobs.size();
  };

  // merge the round's alignments into the half's graphs (src/main.cpp:275-314)
  auto merge = [&](Half& hf) {
    const auto tg = Clock::now();
    const std::uint32_t r = hf.round;
    std::mutex err_mu;
    std::string err;
    pool.For(hf.live.size(), [&](std::size_t x) {
      try {
        const std::uint32_t w = hf.live[x];
        std::uint32_t len;
        const char* s = read_of(w, r, &len);
        // the job's pairs (path order), or none where no DP ran
        auto alignment_of = [&](int d, std::int32_t* score, std::uint32_t* n) -> const std::int32_t* {
          *n = 0;
          const std::int32_t ji = hf.job_of[x * per_read + d];
          if (ji < 0) return nullptr;
          const spoa_hip_job& j = hf.jobs[ji];
          if (j.has_score) *score = j.score;
          *n = j.n_pairs;
          return j.pairs;
        };
        std::int32_t score = 0, score_rev = 0;
        std::uint32_t na = 0, nb = 0;
        const std::int32_t* a = alignment_of(0, &score, &na);
        bool reversed = false;
        if (opt.strand) {
          const std::int32_t* b = alignment_of(1, &score_rev, &nb);
          if (!(score >= score_rev)) {
            a = b;
            na = nb;
            reversed = true;
          }
          out[w].reversed.push_back(reversed ? 1 : 0);
        }
        const char* seq = reversed ? rc_seq[w].data() : s;
        if (ws.quals) {
          const char* q = ws.quals + (s - ws.seqs);
          std::vector<std::uint32_t> wt = Weights(q, len);
          if (reversed) std::reverse(wt.begin(), wt.end());
          graphs[w].AddAlignment(a, na, seq, len, wt.data());
        } else {
          graphs[w].AddAlignment(a, na, seq, len, ones.data());
        }
      } catch (const std::exception& ex) {
        std::lock_guard<std::mutex> lk(err_mu);
        if (err.empty()) err = ex.what();
      }
    });
    if (!err.empty()) throw std::invalid_argument(err);
    st.graph_ms += Ms(tg, Clock::now());
  };

  // Pipeline: the device aligns one half while the host merges the other.
  if (opt.max_rounds > 0) max_reads = std::min(max_reads, opt.max_rounds);
  for (std::uint32_t r = 0; r < max_reads; ++r) {
    for (int h = 0; h < 2; ++h) {
      Half& hf = half[h];
      if (r > 0) {
        wait(hf);
        merge(hf);
      }
      build(hf, r);
      submit(hf);
    }
    ++st.rounds;
  }
  for (int h = 0; h < 2; ++h) {
    if (max_reads > 0) {
      wait(half[h]);
      merge(half[h]);
    }
  }

  const auto tf = Clock::now();
  if (opt.max_rounds > 0) {  // a short run: no consensus
    st.wall_ms = Ms(t0, Clock::now());
    if (stats) *stats = st;
    return out;
  }
  pool.For(W, [&](std::size_t w) {
    out[w].consensus = graphs[w].GenerateConsensus();
    if (opt.msa) graphs[w].MultipleSequenceAlignmentBlock(false, &out[w].msa, &out[w].msa_rows, &out[w].msa_width);
  });
  st.finish_ms = Ms(tf, Clock::now());
  st.wall_ms = Ms(t0, Clock::now());
  if (stats) *stats = st;
  return out;
}

namespace {

// The windows w = first, first + step, ... through one engine on the
// calling thread, in order.
void PerCallWindows(const WindowSet& ws, const BatchOptions& opt, std::uint32_t first, std::uint32_t step,
                    std::vector<WindowResult>& out, BatchStats& st) {
  auto engine = spoa::CreateEngine(opt.type, opt.m, opt.n, opt.g, opt.e, opt.q, opt.c, opt.sisd);
  auto* hip = dynamic_cast<spoa::HipAlignmentEngine*>(engine.get());
  // the engines of the threads share the device's memory: each takes its
  // share of the budget an engine gets alone
  if (hip && step > 1) {
    spoa_hip_set_memory_budget(hip->device(), spoa_hip_memory_budget(hip->device()) / step);
  }
  for (std::uint32_t w = first; w < ws.n_windows; w += step) {
    spoa::Graph graph;
    for (std::uint64_t k = ws.win_read_off[w]; k < ws.win_read_off[w + 1]; ++k) {
      const char* s = ws.seqs + ws.read_off[k];
      const std::uint32_t len = static_cast<std::uint32_t>(ws.read_off[k + 1] - ws.read_off[k]);
      std::int32_t score = 0, score_rev = 0;
      const auto ta = Clock::now();
      spoa::Alignment a = engine->Align(s, len, graph, &score);
      auto account = [&] {
        if (!hip || graph.nodes().empty() || len == 0) return;
        spoa_hip_stats hs{};
        spoa_hip_last_stats(hip->device(), &hs);
        st.kernel_ms += hs.kernel_ms;
        st.h2d_ms += hs.h2d_ms;
        st.d2h_ms += hs.d2h_ms;
        st.cells += hs.cells;
        st.dp_bytes += hs.dp_bytes;
        st.launches += hs.launches;
        st.align_calls += 1;
      };
      account();
      bool reversed = false;
      std::string rc;
      if (opt.strand) {
        rc = ReverseComplement(s, len);
        spoa::Alignment b = engine->Align(rc.data(), len, graph, &score_rev);
        account();
        if (!(score >= score_rev)) {
          a.swap(b);
          reversed = true;
        }
        out[w].reversed.push_back(reversed ? 1 : 0);
      }
      st.align_ms += Ms(ta, Clock::now());
      const auto tg = Clock::now();
      const char* seq = reversed ? rc.data() : s;
      if (ws.quals) {
        std::vector<std::uint32_t> wt = Weights(ws.quals + ws.read_off[k], len);
        if (reversed) std::reverse(wt.begin(), wt.end());
        graph.AddAlignment(a, seq, len, wt);
      } else {
        graph.AddAlignment(a, seq, len);
      }
      st.graph_ms += Ms(tg, Clock::now());
    }
    const auto tf = Clock::now();
    out[w].consensus = graph.GenerateConsensus();
    if (opt.msa) {
      const std::vector<std::string> rows = graph.GenerateMultipleSequenceAlignment();
      WindowResult& o = out[w];
      o.msa_rows = static_cast<std::uint32_t>(rows.size());
      o.msa_width = rows.empty() ? 0u : static_cast<std::uint32_t>(rows[0].size());
      o.msa.clear();
      o.msa.reserve(static_cast<std::size_t>(o.msa_rows) * o.msa_width);
      for (const auto& row : rows) o.msa += row;
    }
    st.finish_ms += Ms(tf, Clock::now());
  }
}

}  // namespace

std::vector<WindowResult> RunPerCall(const WindowSet& ws, const BatchOptions& opt, BatchStats* stats) {
  const auto t0 = Clock::now();
  std::vector<WindowResult> out(ws.n_windows);
  const std::uint32_t nt = std::max<std::uint32_t>(1u, std::min<std::uint32_t>(opt.threads, ws.n_windows));
  std::vector<BatchStats> part(nt);
  if (nt == 1) {
    PerCallWindows(ws, opt, 0, 1, out, part[0]);
  } else {
    // one engine per host thread, the reference's threading model
    std::vector<std::thread> pool;
    std::vector<std::string> errors(nt);
    for (std::uint32_t t = 0; t < nt; ++t) {
      pool.emplace_back([&, t] {
        try {
          PerCallWindows(ws, opt, t, nt, out, part[t]);
        } catch (const std::exception& ex) {
          errors[t] = ex.what();
        }
      });
    }
    for (auto& th : pool) th.join();
    for (const auto& e : errors) {
      if (!e.empty()) throw std::invalid_argument(e);
    }
  }
  BatchStats st;
  for (const auto& p : part) {
    st.kernel_ms += p.kernel_ms;
    st.h2d_ms += p.h2d_ms;
    st.d2h_ms += p.d2h_ms;
    st.cells += p.cells;
    st.dp_bytes += p.dp_bytes;
    st.launches += p.launches;
    st.align_calls += p.align_calls;
    st.align_ms += p.align_ms;
    st.graph_ms += p.graph_ms;
    st.finish_ms += p.finish_ms;
  }
  st.wall_ms = Ms(t0, Clock::now());
  if (stats) *stats = st;
  return out;
}

// ------------------------------------------------------------- autotune --

namespace {

std::string TuneKey(const BatchOptions& o, int width, std::uint32_t max_len, std::uint32_t n_windows) {
  int subtype = 0;
  std::int8_t g = o.g, e = o.e, q = o.q, c = o.c;
  subtype = static_cast<int>(spoa::CheckCreateParams(o.type, &g, &e, &q, &c));
  int lb = 0, jb = 0;
  while ((1u << lb) < max_len) ++lb;
  while ((1u << jb) < n_windows) ++jb;
  return "sub" + std::to_string(subtype) + "_type" + std::to_string(static_cast<int>(o.type)) + "_w" +
         std::to_string(width) + "_len" + std::to_string(lb) + "_win" + std::to_string(jb);
}

// The cache: one line per entry, `"key": [k, waves, kernel_ms]`, inside braces.
std::map<std::string, std::array<double, 3>> LoadTuneCache(const std::string& path) {
  std::map<std::string, std::array<double, 3>> out;
  if (path.empty()) return out;
  std::FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return out;
  char line[512];
  while (std::fgets(line, sizeof(line), f)) {
    char key[256];
    double k = 0, w = 0, ms = 0;
    if (std::sscanf(line, " \"%255[^\"]\": [%lf, %lf, %lf]", key, &k, &w, &ms) == 4) out[key] = {k, w, ms};
  }
  std::fclose(f);
  return out;
}

void SaveTuneCache(const std::string& path, const std::map<std::string, std::array<double, 3>>& c) {
  if (path.empty()) return;
  const std::string tmp = path + ".tmp";
  std::FILE* f = std::fopen(tmp.c_str(), "w");
  if (!f) return;
  std::fprintf(f, "{\n");
  std::size_t i = 0;
  for (const auto& kv : c) {
    std::fprintf(f, "  \"%s\": [%d, %d, %.3f]%s\n", kv.first.c_str(), static_cast<int>(kv.second[0]),
                 static_cast<int>(kv.second[1]), kv.second[2], ++i < c.size() ? "," : "");
  }
  std::fprintf(f, "}\n");
  std::fclose(f);
  std::rename(tmp.c_str(), path.c_str());
}

}  // namespace

std::vector<TuneResult> Autotune(const WindowSet& ws, BatchContext& ctx, std::uint32_t rounds,
                                 const std::string& cache_path) {
  BatchOptions& opt = ctx.mutable_options();
  spoa_hip_engine* dev[2];
  for (int h = 0; h < 2; ++h) {
    auto* hip = dynamic_cast<spoa::HipAlignmentEngine*>(ctx.engine(h));
    if (!hip) throw std::invalid_argument("[spoa::hip] the autotuner needs the HIP engine");
    dev[h] = hip->device();
  }
  std::int8_t g = opt.g, e = opt.e, q = opt.q, c = opt.c;
  const bool convex = spoa::CheckCreateParams(opt.type, &g, &e, &q, &c) == spoa::AlignmentSubtype::kConvex;
  std::uint32_t max_len = 0;
  for (std::uint64_t k = 0; k < ws.win_read_off[ws.n_windows]; ++k) {
    max_len = std::max<std::uint32_t>(max_len, static_cast<std::uint32_t>(ws.read_off[k + 1] - ws.read_off[k]));
  }
  auto cache = LoadTuneCache(cache_path);
  auto set = [&](int width, int k, int waves) {
    for (int h = 0; h < 2; ++h) {
      if (spoa_hip_set_config(dev[h], width, 1, 0xffffffffu, k, waves) != SPOA_HIP_OK) {
        throw std::invalid_argument(spoa_hip_last_error());
      }
    }
  };
  const std::uint32_t saved_rounds = opt.max_rounds;
  std::vector<TuneResult> results;
  try {
    for (int width : {16, 32}) {
      TuneResult tr;
      tr.width = width;
      const std::string key = TuneKey(opt, width, max_len, ws.n_windows);
      auto it = cache.find(key);
      if (it != cache.end()) {
        tr.k = static_cast<int>(it->second[0]);
        tr.waves = static_cast<int>(it->second[1]);
        tr.cached = true;
        set(width, tr.k, tr.waves);
        results.push_back(tr);
        continue;
      }
      // candidates: lanes of 8 / 16 / 32 columns (convex int16 at least 16,
      // int32 at most 16), as many waves as a row needs and twice that
      const int L = width == 16 ? 16 : 8;
      const std::uint32_t W = (max_len + L - 1) / L * L;
      std::vector<std::pair<int, int>> cands;
      for (int k : {8, 16, 32}) {
        if (width == 16 && convex && k < 16) continue;
        if (width == 32 && k > 16) continue;
        const int need = static_cast<int>(std::min<std::uint32_t>(8u, (W + 64u * k - 1) / (64u * k)));
        for (int wv : {need, 2 * need}) {
          if (wv < 1 || wv > 8 || (k == 32 && wv != 1)) continue;
          if (std::find(cands.begin(), cands.end(), std::make_pair(k, wv)) == cands.end()) cands.push_back({k, wv});
        }
      }
      // The first rounds decide; where they hold no job of this width (the
      // int32-width jobs of growing windows appear only once their graphs
      // pass the int16 bound, in the late rounds), the candidates run the
      // whole batch instead.
      double best = -1;
      for (int pass = 0; pass < 2 && best < 0; ++pass) {
        opt.max_rounds = pass == 0 ? std::max<std::uint32_t>(2u, rounds) : 0u;
        if (pass == 1) tr.whole_batch = true;
        for (const auto& cd : cands) {
          set(width, cd.first, cd.second);
          BatchStats st;
          (void)RunBatched(ws, ctx, &st);
          // An int32-width launch runs beside the int16 launches and takes
          // part of the device from them, so its shapes are judged by both
          // widths' launch time, which is what a step pays.
          const double own = width == 16 ? st.kernel_ms16 : st.kernel_ms32;
          const double ms = width == 16 ? own : st.kernel_ms16 + st.kernel_ms32;
          tr.tried.push_back({cd, ms});
          if (own <= 0) continue;  // no job of this width
          if (best < 0 || ms < best) {
            best = ms;
            tr.k = cd.first;
            tr.waves = cd.second;
            tr.kernel_ms = ms;
          }
        }
      }
      if (best < 0) {
        set(width, 0, 0);  // no job of this width: keep the defaults
      } else {
        set(width, tr.k, tr.waves);
        cache[key] = {static_cast<double>(tr.k), static_cast<double>(tr.waves), tr.kernel_ms};
      }
      results.push_back(tr);
    }
  } catch (...) {
    opt.max_rounds = saved_rounds;
    throw;
  }
  opt.max_rounds = saved_rounds;
  SaveTuneCache(cache_path, cache);
  return results;
}

}  // namespace spoa_amd
