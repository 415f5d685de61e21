// CPU check of spoa_amd/csrc/window_graph.hpp (the device-resident window
// graph: AddAlignment, TopologicalSort, the alignment kernels' graph block,
// consensus and MSA), run with a one-lane team against the product's flat
// host graph (PoaGraph) and the oracle, after every read of seeded random
// windows in all nine scoring modes, with and without Phred weights.  Every
// other read the topological sort also runs on a 4-lane team of threads in
// lockstep, so that the lane-parallel steps of its walk (list chunks,
// ballots, prefix counts, lane minima) are checked against the one-lane
// result.
//
// Built and run by tests/test_window_graph.py:
//   g++ -std=c++17 -O2 -pthread -I include -I spoa_amd/csrc window_graph_test.cpp
//       spoa_amd/csrc/poa_graph.cpp oracle/spoa_oracle.cpp

#define __host__
#define __device__
#include "window_graph.hpp"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "poa_graph.hpp"

extern "C" {
void* oracle_engine_new(int type, int m, int n, int g, int e, int q, int c);
void oracle_engine_free(void* p);
void* oracle_graph_new();
void oracle_graph_free(void* p);
int oracle_align(void* eng, void* graph, const char* seq, std::uint32_t len, std::int32_t* pairs_out,
                 std::uint32_t cap, std::int32_t* score, int* has_score, int* width_bits);
int oracle_graph_add_alignment(void* graph, const std::int32_t* pairs, std::uint32_t n_pairs, const char* seq,
                               std::uint32_t len, const std::uint32_t* weights, std::uint32_t n_weights);
int oracle_graph_consensus(void* graph, char* out, std::uint32_t cap);
std::uint32_t oracle_graph_num_nodes(void* p);
}

using namespace spoa_amd;
using namespace spoa_amd::wg;

struct SerialTeam {
  std::uint32_t lane = 0;
  static constexpr std::uint32_t size = 1;
  void sync() const {}
  std::uint32_t scan(std::uint32_t v, std::uint32_t* total) const {
    *total = v;
    return 0;
  }
  bool any(bool b) const { return b; }
  std::uint32_t uniform(std::uint32_t v) const { return v; }
  std::uint64_t ballot(bool b) const { return b ? 1u : 0u; }
  std::uint32_t below(std::uint64_t) const { return 0; }
  std::uint32_t readlane(std::uint32_t v, std::uint32_t) const { return v; }
  std::uint32_t min16(std::uint32_t v) const { return v; }
  std::uint32_t min_all(std::uint32_t v) const { return v; }
  std::uint32_t max_all(std::uint32_t v) const { return v; }
  std::uint32_t sum_all(std::uint32_t v) const { return v; }
};

// A team of L lanes as L threads in lockstep: every team operation is a
// barrier-separated exchange, so the lanes run the same steps between
// barriers as a wave's lanes run the same instructions.  Between two
// barriers every lane writes the same values or distinct locations, which
// is what the wave code relies on too.
struct LaneShared {
  explicit LaneShared(std::uint32_t n) : L(n) {}
  std::uint32_t L;
  std::atomic<std::uint32_t> arrived{0}, gen{0};
  std::uint64_t buf[64] = {};
  void barrier() {
    const std::uint32_t g = gen.load(std::memory_order_acquire);
    if (arrived.fetch_add(1, std::memory_order_acq_rel) + 1 == L) {
      arrived.store(0, std::memory_order_relaxed);
      gen.fetch_add(1, std::memory_order_acq_rel);
    } else {
      while (gen.load(std::memory_order_acquire) == g) std::this_thread::yield();
    }
  }
};

template <std::uint32_t L>
struct LaneTeam {
  std::uint32_t lane;
  LaneShared* sh;
  static constexpr std::uint32_t size = L;
  void sync() const { sh->barrier(); }
  // every lane's value, then a reduction of them (the same on every lane)
  template <class F>
  std::uint64_t all(std::uint64_t v, F f) const {
    sh->buf[lane] = v;
    sh->barrier();
    const std::uint64_t x = f(sh->buf);
    sh->barrier();
    return x;
  }
  std::uint32_t scan(std::uint32_t v, std::uint32_t* total) const {
    const std::uint64_t packed = all(v, [&](const std::uint64_t* b) {
      std::uint64_t ex = 0, tot = 0;
      for (std::uint32_t i = 0; i < L; ++i) {
        if (i < lane) ex += b[i];
        tot += b[i];
      }
      return ex | tot << 32;
    });
    *total = static_cast<std::uint32_t>(packed >> 32);
    return static_cast<std::uint32_t>(packed);
  }
  bool any(bool b) const {
    return all(b, [](const std::uint64_t* x) {
             std::uint64_t a = 0;
             for (std::uint32_t i = 0; i < L; ++i) a |= x[i];
             return a;
           }) != 0;
  }
  std::uint32_t uniform(std::uint32_t v) const {
    return static_cast<std::uint32_t>(all(v, [](const std::uint64_t* x) { return x[0]; }));
  }
  std::uint64_t ballot(bool b) const {
    return all(b, [](const std::uint64_t* x) {
      std::uint64_t m = 0;
      for (std::uint32_t i = 0; i < L; ++i) m |= (x[i] ? 1ull : 0ull) << i;
      return m;
    });
  }
  std::uint32_t below(std::uint64_t m) const {
    return static_cast<std::uint32_t>(__builtin_popcountll(m & ((1ull << lane) - 1)));
  }
  std::uint32_t readlane(std::uint32_t v, std::uint32_t l) const {
    return static_cast<std::uint32_t>(all(v, [&](const std::uint64_t* x) { return x[l]; }));
  }
  std::uint32_t reduce(std::uint32_t v, std::uint32_t n, int op) const {
    return static_cast<std::uint32_t>(all(v, [&](const std::uint64_t* x) {
      std::uint64_t a = x[0];
      for (std::uint32_t i = 1; i < n; ++i) a = op == 0 ? std::min(a, x[i]) : op == 1 ? std::max(a, x[i]) : a + x[i];
      return a;
    }));
  }
  std::uint32_t min16(std::uint32_t v) const { return reduce(v, L < 16 ? L : 16, 0); }
  std::uint32_t min_all(std::uint32_t v) const { return reduce(v, L, 0); }
  std::uint32_t max_all(std::uint32_t v) const { return reduce(v, L, 1); }
  std::uint32_t sum_all(std::uint32_t v) const { return reduce(v, L, 2); }
};

// The topological sort run by an L-lane team of threads.
template <std::uint32_t L>
static bool SortWithLanes(const WinView& gv, const SortSpace<std::uint32_t>& sp) {
  LaneShared sh(L);
  std::vector<std::thread> th;
  std::vector<int> ok(L, 0);
  for (std::uint32_t l = 0; l < L; ++l) {
    th.emplace_back([&, l] { ok[l] = wg::TopologicalSort(LaneTeam<L>{l, &sh}, gv, sp) ? 1 : 0; });
  }
  for (auto& x : th) x.join();
  for (std::uint32_t l = 0; l < L; ++l) {
    if (!ok[l]) return false;
  }
  return true;
}

static int g_fail = 0;
#define CHECK(c, ...)                          \
  do {                                         \
    if (!(c)) {                                \
      std::fprintf(stderr, "FAIL %s: ", #c);   \
      std::fprintf(stderr, __VA_ARGS__);       \
      std::fprintf(stderr, "\n");              \
      ++g_fail;                                \
      return;                                  \
    }                                          \
  } while (0)

static std::string Mutate(std::mt19937_64& rng, const std::string& t, double p, const char* alpha) {
  std::uniform_real_distribution<double> u(0, 1);
  const int na = static_cast<int>(std::strlen(alpha));
  std::string out;
  for (char b : t) {
    const double x = u(rng);
    if (x < p / 3) {
      out.push_back(alpha[rng() % na]);
    } else if (x < 2 * p / 3) {
      out.push_back(b);
      out.push_back(alpha[rng() % na]);
    } else if (x < p) {
      continue;
    } else {
      out.push_back(b);
    }
  }
  return out.empty() ? t : out;
}

// first columns by the reference's recurrence over the rank view (simd...:568-694)
static void FirstColumnsRef(int type, int sub, int g, int e, int q, int c, const spoa_hip_graph& gr,
                            std::vector<std::int32_t>& fcH, std::vector<std::int32_t>& fcF,
                            std::vector<std::int32_t>& fcO) {
  const std::uint32_t V = gr.num_nodes;
  fcH.assign(V + 1, 0);
  fcF.assign(V + 1, 0);
  fcO.assign(V + 1, 0);
  if (type != 1) return;
  const std::int32_t kNinf = -2147483647 - 1 + 1024;
  for (std::uint32_t i = 1; i <= V; ++i) {
    const std::uint32_t b = gr.pred_off[i - 1], en = gr.pred_off[i];
    std::int32_t pf = b == en ? g - e : kNinf, po = b == en ? q - c : kNinf, ph = b == en ? 0 : kNinf;
    for (std::uint32_t k = b; k < en; ++k) {
      const std::uint32_t row = gr.pred_rank[k] + 1;
      pf = std::max(pf, fcF[row]);
      po = std::max(po, fcO[row]);
      ph = std::max(ph, fcH[row]);
    }
    fcF[i] = pf + e;
    fcO[i] = po + c;
    fcH[i] = sub == 2 ? std::max(fcF[i], fcO[i]) : (sub == 1 ? fcF[i] : ph + g);
  }
}

static void RunWindow(std::uint64_t seed, int mode, bool quality, bool small_stack) {
  static const int kParams[9][7] = {
      {0, 5, -4, -8, -8, -8, -8}, {0, 5, -4, -8, -6, -8, -6}, {0, 5, -4, -8, -6, -10, -2},
      {1, 5, -4, -8, -8, -8, -8}, {1, 5, -4, -8, -6, -8, -6}, {1, 5, -4, -8, -6, -10, -2},
      {2, 5, -4, -8, -8, -8, -8}, {2, 5, -4, -8, -6, -8, -6}, {2, 5, -4, -8, -6, -10, -2}};
  const int* P = kParams[mode];
  // normalised gaps (src/alignment_engine.cpp:57-66)
  int g = P[3], e = P[4], q = P[5], c = P[6];
  int sub = g >= e ? 0 : ((g <= q || e >= c) ? 1 : 2);
  if (sub == 0) e = g, q = g, c = g;
  if (sub == 1) q = g, c = e;
  std::mt19937_64 rng(seed);
  const char* alpha = (seed % 3 == 0) ? "ACGTN" : "ACGT";
  std::string templ;
  const int tl = 40 + static_cast<int>(rng() % 160);
  for (int i = 0; i < tl; ++i) templ.push_back("ACGT"[rng() % 4]);
  const int nreads = 3 + static_cast<int>(rng() % 12);
  std::vector<std::string> reads;
  std::uint64_t total = 0;
  for (int k = 0; k < nreads; ++k) {
    std::string r = Mutate(rng, templ, 0.15, alpha);
    if (seed % 5 == 1 && k % 3 == 2) {
      for (auto& ch : r) ch = static_cast<char>(std::tolower(ch));
    }
    if (k == 1 && seed % 7 == 3) r = r.substr(r.size() / 3);  // an unaligned prefix / suffix
    reads.push_back(r);
    total += r.size();
  }
  WinCaps caps{static_cast<std::uint32_t>(total), static_cast<std::uint32_t>(total),
               static_cast<std::uint32_t>(total), static_cast<std::uint32_t>(nreads)};
  const WinLayout lay = MakeLayout(caps);
  std::vector<std::uint64_t> region(lay.total / 8 + 1);
  WinView gv = MakeWinView(reinterpret_cast<std::uint8_t*>(region.data()), caps);
  SerialTeam t;
  wg::InitWindow(t, gv);
  std::vector<std::uint32_t> block(BlockWords(caps.V, caps.E) + 16);
  std::vector<std::uint32_t> lds;  // a small "LDS" sort space
  void* oe = oracle_engine_new(P[0], P[1], P[2], P[3], P[4], P[5], P[6]);
  void* og = oracle_graph_new();
  PoaGraph pg;
  for (int k = 0; k < nreads; ++k) {
    const std::string& r = reads[k];
    std::vector<std::int32_t> pairs(2 * (oracle_graph_num_nodes(og) + r.size() + 8));
    std::int32_t score = 0;
    int has = 0, width = 0;
    const int np = oracle_align(oe, og, r.data(), static_cast<std::uint32_t>(r.size()), pairs.data(),
                                static_cast<std::uint32_t>(pairs.size() / 2), &score, &has, &width);
    CHECK(np >= 0, "oracle align");
    std::vector<std::int32_t> stored(2 * np);
    for (int i = 0; i < np; ++i) {
      stored[2 * (np - 1 - i)] = pairs[2 * i];
      stored[2 * (np - 1 - i) + 1] = pairs[2 * i + 1];
    }
    std::string qual;
    std::vector<std::uint32_t> w(r.size(), 1);
    if (quality) {
      for (std::size_t i = 0; i < r.size(); ++i) {
        qual.push_back(static_cast<char>(33 + rng() % 40));
        w[i] = static_cast<std::uint32_t>(qual[i] - 33);
      }
    }
    wg::AlignIn in{stored.data(), static_cast<std::uint32_t>(np), reinterpret_cast<const std::uint8_t*>(r.data()),
                   static_cast<std::uint32_t>(r.size()),
                   quality ? reinterpret_cast<const std::uint8_t*>(qual.data()) : nullptr, false};
    const std::int32_t st = wg::AddAlignment(t, gv, in);
    CHECK(st == wg::kWinOk, "AddAlignment status %d (seed %llu read %d)", st, (unsigned long long)seed, k);
    const std::uint32_t V = gv.hdr->V, E = gv.hdr->E;
    std::uint32_t A = 0;
    for (std::uint32_t v = 0; v < V; ++v) A += gv.al_cnt[v];
    const bool nw = P[0] == 1;
    bool sorted = false;
    if (small_stack && V < 65535 && E < 65535) {
      // 16-bit space with a tiny stack first: overflow must be reported
      const std::uint32_t K = 3;
      lds.assign(SortSpaceBytes(V, E, A, K, nw, 2) / 4 + 2, 0);
      auto sp = wg::CarveSortSpace<std::uint16_t>(lds.data(), V, E, A, K, nw);
      sorted = wg::TopologicalSort(t, gv, sp);
      if (sorted) {
        wg::BuildBlock(t, gv, sp, block.data(), P[0], sub, g, e, q, c);
      }
    }
    if (!sorted) {
      auto sp = wg::CarveSortSpace<std::uint32_t>(gv.scratch, V, E, A, V + E + A + 1, nw);
      CHECK(wg::TopologicalSort(t, gv, sp), "sort overflow with a full stack");
      wg::BuildBlock(t, gv, sp, block.data(), P[0], sub, g, e, q, c);
    }
    pg.AddAlignment(pairs.data(), static_cast<std::uint32_t>(np), r.data(), static_cast<std::uint32_t>(r.size()),
                    w.data());
    oracle_graph_add_alignment(og, pairs.data(), static_cast<std::uint32_t>(np), r.data(),
                               static_cast<std::uint32_t>(r.size()), w.data(), static_cast<std::uint32_t>(w.size()));
    CHECK(V == pg.num_nodes(), "V %u vs %u", V, pg.num_nodes());
    const auto& ranks = pg.rank_to_node();
    for (std::uint32_t i = 0; i < V; ++i) {
      CHECK(gv.rank_to_node[i] == ranks[i], "rank %u: %u vs %u (seed %llu read %d)", i, gv.rank_to_node[i], ranks[i],
            (unsigned long long)seed, k);
    }
    // the graph block against the host view (and the reference first columns)
    const spoa_hip_graph hv = pg.View();
    std::vector<std::int32_t> fcH, fcF, fcO;
    FirstColumnsRef(P[0], sub, g, e, q, c, hv, fcH, fcF, fcO);
    const std::uint32_t* rows = block.data();
    const std::int32_t* node_id = reinterpret_cast<const std::int32_t*>(block.data() + 4ull * V);
    const std::uint32_t* poff = block.data() + 5ull * V;
    const std::int32_t* bH = reinterpret_cast<const std::int32_t*>(block.data() + 6ull * V + 1);
    const std::int32_t* bF = bH + (V + 1);
    const std::int32_t* bO = bF + (V + 1);
    const std::uint32_t* preds = block.data() + 9ull * V + 4;
    for (std::uint32_t i = 0; i < V; ++i) {
      const std::uint32_t pb = hv.pred_off[i] - hv.pred_off[0], pe = hv.pred_off[i + 1] - hv.pred_off[0];
      CHECK(node_id[i] == hv.node_id[i], "node id at rank %u", i);
      CHECK(poff[i] == pb, "pred offset at rank %u", i);
      CHECK((rows[4 * i] & 0xffu) == hv.row_char[i], "row char at rank %u", i);
      CHECK(((rows[4 * i] >> 8) & 1u) == hv.is_sink[i], "sink at rank %u", i);
      CHECK((rows[4 * i] >> 12) == std::max<std::uint32_t>(1, pe - pb), "npred at rank %u", i);
      CHECK(rows[4 * i + 1] == (pe > pb ? hv.pred_rank[hv.pred_off[i]] + 1 : 0u), "first pred at rank %u", i);
      for (std::uint32_t x = pb; x < pe; ++x) CHECK(preds[x] == hv.pred_rank[hv.pred_off[0] + x] + 1, "pred");
    }
    for (std::uint32_t i = 0; i <= V; ++i) {
      CHECK(bH[i] == fcH[i] && bF[i] == fcF[i] && bO[i] == fcO[i], "first columns at row %u (mode %d)", i, mode);
      if (i > 0) CHECK(static_cast<std::int32_t>(rows[4 * (i - 1) + 3]) == fcH[i], "row fcH %u", i);
    }
    CHECK(E == static_cast<std::uint32_t>(hv.pred_off[V] - hv.pred_off[0]), "E");
    if (k % 2 == 0) {
      // the sort again by a 4-lane team (a node's lists then span several
      // lane chunks): the same ranks, and the same graph block (its first
      // columns follow the sort's source distances)
      const std::vector<std::uint32_t> block1(block.begin(), block.begin() + BlockWords(V, E));
      std::vector<std::uint32_t> space(SortSpaceBytes(V, E, A, V + E + A + 1, nw, 4) / 4 + 2, 0);
      auto sp4 = wg::CarveSortSpace<std::uint32_t>(space.data(), V, E, A, V + E + A + 1, nw);
      CHECK(SortWithLanes<4>(gv, sp4), "4-lane sort overflow");
      for (std::uint32_t i = 0; i < V; ++i) {
        CHECK(gv.rank_to_node[i] == ranks[i], "4-lane rank %u: %u vs %u (seed %llu read %d)", i, gv.rank_to_node[i],
              ranks[i], (unsigned long long)seed, k);
        CHECK(gv.rank_of[ranks[i]] == i, "4-lane rank_of (seed %llu read %d)", (unsigned long long)seed, k);
      }
      wg::BuildBlock(t, gv, sp4, block.data(), P[0], sub, g, e, q, c);
      CHECK(std::equal(block1.begin(), block1.end(), block.begin()), "4-lane graph block (seed %llu read %d)",
            (unsigned long long)seed, k);
    }
  }
  // consensus and MSA
  std::vector<std::uint32_t> cons(gv.hdr->V + 1), col(gv.hdr->V + 1);
  const std::uint32_t n = wg::Consensus(t, gv, gv.scratch, cons.data());
  std::string c1;
  for (std::uint32_t i = 0; i < n; ++i) c1.push_back(static_cast<char>(gv.hdr->decoder[gv.code[cons[i]]]));
  const std::string c2 = pg.GenerateConsensus();
  char buf[1 << 14];
  const int nc = oracle_graph_consensus(og, buf, sizeof(buf));
  CHECK(c1 == c2 && c1 == std::string(buf, nc), "consensus (seed %llu mode %d)", (unsigned long long)seed, mode);
  const std::uint32_t width = wg::MsaColumns(t, gv, col.data());
  std::vector<std::uint8_t> rows((gv.hdr->nseq + 1) * static_cast<std::size_t>(width));
  wg::WriteMsa(t, gv, col.data(), width, cons.data(), n, rows.data());
  const auto msa = pg.GenerateMultipleSequenceAlignment(true);
  CHECK(msa.size() == gv.hdr->nseq + 1, "msa rows");
  for (std::size_t s = 0; s < msa.size(); ++s) {
    CHECK(msa[s] == std::string(reinterpret_cast<char*>(rows.data()) + s * width, width), "msa row %zu", s);
  }
  oracle_graph_free(og);
  oracle_engine_free(oe);
}

int main() {
  int windows = 0;
  for (std::uint64_t seed = 1; seed <= 60; ++seed) {
    for (int mode = 0; mode < 9; ++mode) {
      RunWindow(seed * 131 + mode, mode, seed % 2 == 0, seed % 4 == 1);
      ++windows;
      if (g_fail) {
        std::printf("failed after %d windows\n", windows);
        return 1;
      }
    }
  }
  std::printf("ok %d windows\n", windows);
  return 0;
}
