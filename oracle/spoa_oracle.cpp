// TEST INFRASTRUCTURE ONLY -- the oracle.
//
// A scalar CPU restatement of the reference spoa v4.0.8 partial-order
// alignment path: AlignmentEngine::Create / Align (SIMD engine semantics,
// AVX2 lane counts) and Graph::AddAlignment / TopologicalSort /
// GenerateConsensus / GenerateMultipleSequenceAlignment.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product library never links or calls it.
//
// Every function cites the reference file:line it restates.  The DP is
// restated on logical (V+1) x W matrices, W = ceil(n / L) * L, where L is
// the AVX2 lane count of the chosen score width (16 for int16, 8 for
// int32).  L matters for the stored E/Q contents of convex mode only: the
// SIMD engine opens E/Q gaps from the pre-gap H inside a vector and from
// the final H at vector boundaries
// (src/simd_alignment_engine_implementation.hpp:1733-1769).
//
// SisdAlign restates the reference's scalar engine
// (src/sisd_alignment_engine.cpp), selected per engine
// (oracle_sisd_engine_new).
//
// Pinning: every function here is checked against the compiled reference
// (oracle/_ref, built from /root/reference by oracle/Makefile) on seeded
// random inputs and against the reference's 18 golden consensus strings
// (test/spoa_test.cpp:154-440); see tests/test_oracle.py.

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

enum Type { kSW = 0, kNW = 1, kOV = 2 };
enum Subtype { kLinear = 0, kAffine = 1, kConvex = 2 };

thread_local std::string g_error;

// ---------------------------------------------------------------- graph --
// include/spoa/graph.hpp:26-318, src/graph.cpp:13-660 (the parts on the
// path).  Nodes and edges are kept as indices instead of pointers.
struct Graph {
  std::vector<std::uint32_t> code;                 // Node::code
  std::vector<std::vector<std::uint32_t>> in, out; // edge indices
  std::vector<std::vector<std::uint32_t>> aligned; // node ids
  std::vector<std::uint32_t> tail, head;
  std::vector<std::int64_t> weight;
  std::vector<std::vector<std::uint32_t>> labels;
  std::vector<std::int64_t> seq_begin;             // sequences_ (-1: null)
  std::int32_t coder[256];
  std::int32_t decoder[256];
  std::uint32_t num_codes = 0;
  std::vector<std::uint32_t> rank_to_node;
  std::vector<std::uint32_t> consensus;

  Graph() { Clear(); }

  // src/graph.cpp:649-658
  void Clear() {
    for (int i = 0; i < 256; ++i) coder[i] = decoder[i] = -1;
    num_codes = 0;
    code.clear(); in.clear(); out.clear(); aligned.clear();
    tail.clear(); head.clear(); weight.clear(); labels.clear();
    seq_begin.clear(); rank_to_node.clear(); consensus.clear();
  }

  std::uint32_t V() const { return static_cast<std::uint32_t>(code.size()); }

  // src/graph.cpp:76-79
  std::uint32_t AddNode(std::uint32_t c) {
    code.push_back(c);
    in.emplace_back(); out.emplace_back(); aligned.emplace_back();
    return V() - 1;
  }

  // src/graph.cpp:81-91: an existing tail->head edge gains the label.
  void AddEdge(std::uint32_t t, std::uint32_t h, std::uint32_t w) {
    for (auto e : out[t]) {
      if (head[e] == h) {
        labels[e].push_back(static_cast<std::uint32_t>(seq_begin.size()));
        weight[e] += w;
        return;
      }
    }
    std::uint32_t e = static_cast<std::uint32_t>(tail.size());
    tail.push_back(t); head.push_back(h); weight.push_back(w);
    labels.push_back({static_cast<std::uint32_t>(seq_begin.size())});
    out[t].push_back(e);
    in[h].push_back(e);
  }

  // src/graph.cpp:93-110
  std::int64_t AddSequence(const char* s, const std::vector<std::uint32_t>& w,
                           std::uint32_t b, std::uint32_t en) {
    if (b == en) return -1;
    std::int64_t prev = -1;
    for (std::uint32_t i = b; i < en; ++i) {
      std::uint32_t curr = AddNode(coder[static_cast<std::uint8_t>(s[i])]);
      if (prev >= 0) AddEdge(static_cast<std::uint32_t>(prev), curr, w[i - 1] + w[i]);
      prev = curr;
    }
    return static_cast<std::int64_t>(V()) - (en - b);
  }

  // src/graph.cpp:155-247
  void AddAlignment(const std::int32_t* pairs, std::uint32_t n_pairs,
                    const char* s, std::uint32_t len,
                    const std::vector<std::uint32_t>& w) {
    if (len == 0) return;
    if (len != w.size()) {
      throw std::invalid_argument(
          "[spoa::Graph::AddAlignment] error: "
          "sequence and weights are of unequal size!");
    }
    for (std::uint32_t i = 0; i < len; ++i) {
      std::uint8_t ch = static_cast<std::uint8_t>(s[i]);
      if (coder[ch] == -1) {
        coder[ch] = num_codes;
        decoder[num_codes++] = s[i];
      }
    }
    if (n_pairs == 0) {
      seq_begin.push_back(AddSequence(s, w, 0, len));
      TopologicalSort();
      return;
    }
    std::vector<std::uint32_t> valid;
    for (std::uint32_t k = 0; k < n_pairs; ++k) {
      std::int32_t sj = pairs[2 * k + 1];
      if (sj != -1) {
        if (sj < 0 || sj >= static_cast<std::int32_t>(len)) {
          throw std::invalid_argument(
              "[spoa::Graph::AddAlignment] error: invalid alignment");
        }
        valid.push_back(sj);
      }
    }
    if (valid.empty()) {
      throw std::invalid_argument(
          "[spoa::Graph::AddAlignment] error: missing sequence in alignment");
    }
    std::int64_t begin = AddSequence(s, w, 0, valid.front());
    std::int64_t prev = begin >= 0 ? static_cast<std::int64_t>(V()) - 1 : -1;
    std::int64_t last = AddSequence(s, w, valid.back() + 1, len);
    for (std::uint32_t k = 0; k < n_pairs; ++k) {
      std::int32_t ni = pairs[2 * k], sj = pairs[2 * k + 1];
      if (sj == -1) continue;
      std::uint32_t c = coder[static_cast<std::uint8_t>(s[sj])];
      std::int64_t curr = -1;
      if (ni == -1) {
        curr = AddNode(c);
      } else {
        std::uint32_t jt = static_cast<std::uint32_t>(ni);
        if (code[jt] == c) {
          curr = jt;
        } else {
          for (auto kt : aligned[jt]) {
            if (code[kt] == c) { curr = kt; break; }
          }
          if (curr < 0) {
            std::uint32_t nn = AddNode(c);
            for (auto kt : aligned[jt]) {
              aligned[kt].push_back(nn);
              aligned[nn].push_back(kt);
            }
            aligned[jt].push_back(nn);
            aligned[nn].push_back(jt);
            curr = nn;
          }
        }
      }
      if (begin < 0) begin = curr;
      if (prev >= 0) {
        AddEdge(static_cast<std::uint32_t>(prev), static_cast<std::uint32_t>(curr),
                w[sj - 1] + w[sj]);
      }
      prev = curr;
    }
    if (last >= 0) {
      AddEdge(static_cast<std::uint32_t>(prev), static_cast<std::uint32_t>(last),
              w[valid.back()] + w[valid.back() + 1]);
    }
    seq_begin.push_back(begin);
    TopologicalSort();
  }

  // src/graph.cpp:249-303: iterative DFS; an aligned group is emitted as a
  // block right after its first finished member.
  void TopologicalSort() {
    rank_to_node.clear();
    std::vector<std::uint8_t> marks(V(), 0);
    std::vector<bool> ignored(V(), false);
    std::vector<std::uint32_t> stack;
    for (std::uint32_t id = 0; id < V(); ++id) {
      if (marks[id] != 0) continue;
      stack.push_back(id);
      while (!stack.empty()) {
        std::uint32_t curr = stack.back();
        bool is_valid = true;
        if (marks[curr] != 2) {
          for (auto e : in[curr]) {
            if (marks[tail[e]] != 2) { stack.push_back(tail[e]); is_valid = false; }
          }
          if (!ignored[curr]) {
            for (auto a : aligned[curr]) {
              if (marks[a] != 2) { stack.push_back(a); ignored[a] = true; is_valid = false; }
            }
          }
          if (is_valid) {
            marks[curr] = 2;
            if (!ignored[curr]) {
              rank_to_node.push_back(curr);
              for (auto a : aligned[curr]) rank_to_node.push_back(a);
            }
          } else {
            marks[curr] = 1;
          }
        }
        if (is_valid) stack.pop_back();
      }
    }
  }

  // src/graph.cpp:478-516
  std::int64_t BranchCompletion(std::uint32_t rank, std::vector<std::int64_t>& sc,
                                std::vector<std::int64_t>& pr) {
    std::uint32_t start = rank_to_node[rank];
    for (auto e : out[start]) {
      for (auto e2 : in[head[e]]) {
        if (tail[e2] != start) sc[tail[e2]] = -1;
      }
    }
    std::int64_t mx = -1;
    for (std::uint32_t i = rank + 1; i < rank_to_node.size(); ++i) {
      std::uint32_t it = rank_to_node[i];
      sc[it] = -1;
      pr[it] = -1;
      for (auto e : in[it]) {
        if (sc[tail[e]] == -1) continue;
        if (sc[it] < weight[e] ||
            (sc[it] == weight[e] && sc[pr[it]] <= sc[tail[e]])) {
          sc[it] = weight[e];
          pr[it] = tail[e];
        }
      }
      if (pr[it] >= 0) sc[it] += sc[pr[it]];
      if (mx < 0 || sc[mx] < sc[it]) mx = it;
    }
    return mx;
  }

  // src/graph.cpp:433-476
  void TraverseHeaviestBundle() {
    if (rank_to_node.empty()) return;
    std::vector<std::int64_t> pr(V(), -1), sc(V(), -1);
    std::int64_t mx = -1;
    for (auto it : rank_to_node) {
      for (auto e : in[it]) {
        if (sc[it] < weight[e] ||
            (sc[it] == weight[e] && sc[pr[it]] <= sc[tail[e]])) {
          sc[it] = weight[e];
          pr[it] = tail[e];
        }
      }
      if (pr[it] >= 0) sc[it] += sc[pr[it]];
      if (mx < 0 || sc[mx] < sc[it]) mx = it;
    }
    if (!out[mx].empty()) {
      std::vector<std::uint32_t> rank_of(V(), 0);
      for (std::uint32_t i = 0; i < rank_to_node.size(); ++i) rank_of[rank_to_node[i]] = i;
      while (!out[mx].empty()) mx = BranchCompletion(rank_of[mx], sc, pr);
    }
    consensus.clear();
    while (pr[mx] >= 0) {
      consensus.push_back(static_cast<std::uint32_t>(mx));
      mx = pr[mx];
    }
    consensus.push_back(static_cast<std::uint32_t>(mx));
    std::reverse(consensus.begin(), consensus.end());
  }

  // src/graph.cpp:368-375
  std::string GenerateConsensus() {
    TraverseHeaviestBundle();
    std::string s;
    for (auto id : consensus) s += static_cast<char>(decoder[code[id]]);
    return s;
  }

  // src/graph.cpp:22-30
  std::int64_t Successor(std::uint32_t node, std::uint32_t label) const {
    for (auto e : out[node]) {
      for (auto l : labels[e]) if (l == label) return head[e];
    }
    return -1;
  }

  // src/graph.cpp:321-366
  std::vector<std::string> GenerateMsa(bool include_consensus) {
    std::vector<std::uint32_t> col(V(), 0);
    std::uint32_t j = 0;
    for (std::uint32_t i = 0; i < rank_to_node.size(); ++i, ++j) {
      std::uint32_t it = rank_to_node[i];
      col[it] = j;
      for (auto a : aligned[it]) { col[a] = j; ++i; }
    }
    std::vector<std::string> rows;
    for (std::uint32_t s = 0; s < seq_begin.size(); ++s) {
      std::string row(j, '-');
      std::int64_t it = seq_begin[s];
      while (it >= 0) {
        row[col[it]] = static_cast<char>(decoder[code[it]]);
        it = Successor(static_cast<std::uint32_t>(it), s);
      }
      rows.push_back(row);
    }
    if (include_consensus) {
      TraverseHeaviestBundle();
      std::string row(j, '-');
      for (auto id : consensus) row[col[id]] = static_cast<char>(decoder[code[id]]);
      rows.push_back(row);
    }
    return rows;
  }
};

// --------------------------------------------------------------- engine --
struct Engine {
  int type, subtype;
  std::int8_t m, n, g, e, q, c;
  bool sisd = false;  // the scalar engine's semantics (SisdAlign)

  // src/alignment_engine.cpp:101-110
  std::int64_t WorstCase(std::int64_t i, std::int64_t j) const {
    auto gap = [&](std::int64_t len) -> std::int64_t {
      return len == 0 ? 0 : std::min<std::int64_t>(g + (len - 1) * e, q + (len - 1) * c);
    };
    return std::min<std::int64_t>(-1 * (m * std::min(i, j) + gap(std::llabs(i - j))),
                                  gap(i) + gap(j));
  }
};

// src/alignment_engine.cpp:32-73 (validation, subtype choice, normalisation)
Engine* CreateEngine(int type, std::int8_t m, std::int8_t n, std::int8_t g,
                     std::int8_t e, std::int8_t q, std::int8_t c) {
  if (type != kSW && type != kNW && type != kOV) {
    throw std::invalid_argument(
        "[spoa::AlignmentEngine::Create] error: invalid alignment type!");
  }
  if (g > 0 || q > 0) {
    throw std::invalid_argument(
        "[spoa::AlignmentEngine::Create] error: "
        "gap opening penalty must be non-positive!");
  }
  if (e > 0 || c > 0) {
    throw std::invalid_argument(
        "[spoa::AlignmentEngine::Create] error: "
        "gap extension penalty must be non-positive!");
  }
  int sub = g >= e ? kLinear : ((g <= q || e >= c) ? kAffine : kConvex);
  if (sub == kLinear) {
    e = g;
  } else if (sub == kAffine) {
    q = g;
    c = e;
  }
  return new Engine{type, sub, m, n, g, e, q, c};
}

struct AlignResult {
  std::vector<std::int32_t> pairs;  // flattened (node id, column) pairs
  bool has_score = false;
  std::int32_t score = 0;
  int width_bits = 0;               // 16 or 32; 0 when no DP ran
};

// src/simd_alignment_engine_implementation.hpp:703-780 (Align), 518-701
// (Initialize), 782-2131 (Linear / Affine / Convex DP and traceback).
AlignResult Align(const Engine& en, const Graph& gr, const char* seq,
                  std::uint32_t len) {
  AlignResult res;
  if (len > static_cast<std::uint32_t>(std::numeric_limits<std::int32_t>::max())) {
    throw std::invalid_argument(
        "[spoa::SimdAlignmentEngine::Align] error: too large sequence!");
  }
  const std::uint32_t V = gr.V();
  if (V == 0 || len == 0) return res;

  // :719-774 score width from the worst-case bound
  std::int64_t worst = en.WorstCase(static_cast<std::int64_t>(len) + 8, V);
  if (worst < std::numeric_limits<std::int32_t>::min() + 1024) {
    throw std::invalid_argument(
        "[spoa::SimdAlignmentEngine::Align] error: possible overflow!");
  }
  const bool w16 = worst >= std::numeric_limits<std::int16_t>::min() + 1024;
  res.width_bits = w16 ? 16 : 32;
  const std::int64_t L = w16 ? 16 : 8;
  const std::int32_t NINF = w16 ? std::numeric_limits<std::int16_t>::min() + 1024
                                : std::numeric_limits<std::int32_t>::min() + 1024;
  const std::int64_t W = (len + L - 1) / L * L;
  const std::int64_t R = V + 1;
  const int T = en.type, S = en.subtype;
  const std::int32_t m = en.m, n = en.n, g = en.g, e = en.e, q = en.q, c = en.c;

  // rank views: row i = rank + 1; row 0 is the virtual source row
  std::vector<std::uint32_t> rank_of(V, 0);
  for (std::uint32_t r = 0; r < V; ++r) rank_of[gr.rank_to_node[r]] = r;
  std::vector<std::vector<std::uint32_t>> preds(R);  // inedges order
  std::vector<std::vector<std::uint32_t>> real_preds(R);
  std::vector<char> row_char(R, 0);
  std::vector<bool> sink(R, false);
  for (std::uint32_t i = 1; i < R; ++i) {
    std::uint32_t id = gr.rank_to_node[i - 1];
    for (auto ed : gr.in[id]) real_preds[i].push_back(rank_of[gr.tail[ed]] + 1);
    preds[i] = real_preds[i];
    if (preds[i].empty()) preds[i].push_back(0);
    row_char[i] = static_cast<char>(gr.decoder[gr.code[id]]);
    sink[i] = gr.out[id].empty();
  }

  // :522-542 query profile (padding lanes get the padding penalty)
  const std::int32_t pad = -1 * std::max(std::max(std::abs(m), std::abs(n)),
                                         std::max(std::abs(g), std::abs(q)));
  auto prof = [&](std::uint32_t i, std::int64_t col) -> std::int32_t {
    return col < static_cast<std::int64_t>(len) ? (row_char[i] == seq[col] ? m : n) : pad;
  };

  std::vector<std::int32_t> H(R * W), F, E, O, Q;
  if (S != kLinear) { F.assign(R * W, 0); E.assign(R * W, 0); }
  if (S == kConvex) { O.assign(R * W, 0); Q.assign(R * W, 0); }
  std::vector<std::int32_t> fcH(R, 0), fcF(R, 0), fcO(R, 0);

  // :556-605 secondary first columns, row 0 of E/F/O/Q
  if (S == kConvex) {
    for (std::int64_t j = 0; j < W; ++j) { O[j] = NINF; Q[j] = q + static_cast<std::int32_t>(j) * c; }
    fcO[0] = 0;
    for (std::uint32_t i = 1; i < R; ++i) {
      std::int32_t pen = real_preds[i].empty() ? q - c : NINF;
      for (auto p : real_preds[i]) pen = std::max(pen, fcO[p]);
      fcO[i] = pen + c;
    }
  }
  if (S != kLinear) {
    for (std::int64_t j = 0; j < W; ++j) { F[j] = NINF; E[j] = g + static_cast<std::int32_t>(j) * e; }
    fcF[0] = 0;
    for (std::uint32_t i = 1; i < R; ++i) {
      std::int32_t pen = real_preds[i].empty() ? g - e : NINF;
      for (auto p : real_preds[i]) pen = std::max(pen, fcF[p]);
      fcF[i] = pen + e;
    }
  }
  // :607-694 primary matrix row 0 and first column
  if (T == kSW) {
    for (std::int64_t j = 0; j < W; ++j) H[j] = 0;
    std::fill(fcH.begin(), fcH.end(), 0);
  } else {
    for (std::int64_t j = 0; j < W; ++j) {
      if (S == kConvex) H[j] = std::max(E[j], Q[j]);
      else if (S == kAffine) H[j] = E[j];
      else H[j] = g + static_cast<std::int32_t>(j) * g;
    }
    if (T == kNW) {
      if (S == kConvex) {
        for (std::uint32_t i = 0; i < R; ++i) fcH[i] = std::max(fcF[i], fcO[i]);
      } else if (S == kAffine) {
        fcH = fcF;
      } else {
        fcH[0] = 0;
        for (std::uint32_t i = 1; i < R; ++i) {
          std::int32_t pen = real_preds[i].empty() ? 0 : NINF;
          for (auto p : real_preds[i]) pen = std::max(pen, fcH[p]);
          fcH[i] = pen + g;
        }
      }
    } else {
      std::fill(fcH.begin(), fcH.end(), 0);
    }
  }

  auto hp = [&](std::uint32_t p, std::int64_t col) -> std::int32_t {
    return col < 0 ? fcH[p] : H[p * W + col];
  };

  // rows in rank order (:824-953, :1179-1322, :1638-1824)
  std::int32_t max_score = T == kSW ? 0 : NINF;
  std::int64_t max_i = -1, max_j = -1;
  std::vector<std::int32_t> H1(W), Hfin(W);
  for (std::uint32_t i = 1; i < R; ++i) {
    std::int32_t* Hr = &H[i * W];
    for (std::int64_t j = 0; j < W; ++j) {
      std::int32_t d = std::numeric_limits<std::int32_t>::min();
      std::int32_t f = d, o = d, up = d;
      for (auto p : preds[i]) {
        d = std::max(d, hp(p, j - 1) + prof(i, j));
        if (S == kLinear) {
          up = std::max(up, H[p * W + j] + g);
        } else {
          f = std::max(f, std::max(H[p * W + j] + g, F[p * W + j] + e));
          if (S == kConvex) o = std::max(o, std::max(H[p * W + j] + q, O[p * W + j] + c));
        }
      }
      if (S == kLinear) {
        H1[j] = std::max(d, up);
      } else {
        F[i * W + j] = f;
        H1[j] = std::max(d, f);
        if (S == kConvex) { O[i * W + j] = o; H1[j] = std::max(H1[j], o); }
      }
    }
    // horizontal pass; carries use the value before the local clamp
    for (std::int64_t j = 0; j < W; ++j) {
      std::int32_t h;
      if (S == kLinear) {
        std::int32_t left = j == 0 ? fcH[i] : Hfin[j - 1];
        h = std::max(H1[j], left + g);
      } else {
        std::int32_t ev, qv = 0;
        if (j == 0) {
          ev = fcH[i] + g;
          qv = fcH[i] + q;
        } else {
          bool seg_start = (j % L) == 0;
          std::int32_t open_from = seg_start ? Hfin[j - 1] : H1[j - 1];
          ev = std::max(open_from + g, E[i * W + j - 1] + e);
          if (S == kConvex) qv = std::max(open_from + q, Q[i * W + j - 1] + c);
        }
        E[i * W + j] = ev;
        h = std::max(H1[j], ev);
        if (S == kConvex) { Q[i * W + j] = qv; h = std::max(h, qv); }
      }
      Hfin[j] = h;
      Hr[j] = T == kSW ? std::max(h, 0) : h;
    }
    // end-cell candidates (:252-288, :903-930)
    if (T == kSW || (T == kOV && sink[i])) {
      std::int32_t row_max = 0;
      for (std::int64_t j = 0; j < W; ++j) row_max = std::max(row_max, Hr[j]);
      if (max_score < row_max) { max_score = row_max; max_i = i; }
    } else if (T == kNW && sink[i]) {
      std::int32_t v = Hr[len - 1];
      if (max_score < v) { max_score = v; max_i = i; }
    }
  }

  if (max_i == -1 && max_j == -1) return res;
  res.has_score = true;
  res.score = max_score;
  if (T == kNW) {
    max_j = len - 1;
  } else {
    max_j = -1;
    for (std::int64_t j = 0; j < W; ++j) {
      if (H[max_i * W + j] == max_score) { max_j = j; break; }
    }
  }

  // traceback (:955-1124 linear, :1324-1574 affine, :1826-2121 convex)
  auto node_id = [&](std::int64_t row) { return static_cast<std::int32_t>(gr.rank_to_node[row - 1]); };
  std::vector<std::int32_t>& al = res.pairs;
  auto emit = [&](std::int32_t a, std::int32_t b) { al.push_back(a); al.push_back(b); };
  std::int64_t i = max_i, j = max_j, prev_i = 0, prev_j = 0;
  while (true) {
    if (j == -1 || i == 0) break;
    const std::int32_t h = H[i * W + j];
    if (T == kSW && h == 0) break;
    bool found = false, ext_left = false, ext_up = false;
    for (auto p : preds[i]) {
      if (h == hp(p, j - 1) + prof(i, j)) { prev_i = p; prev_j = j - 1; found = true; break; }
    }
    if (!found) {
      for (auto p : preds[i]) {
        const std::int32_t hu = H[p * W + j];
        bool hit;
        if (S == kLinear) {
          hit = h == hu + g;
        } else if (S == kAffine) {
          hit = (ext_up = h == F[p * W + j] + e) || h == hu + g;
        } else {
          hit = (ext_up = h == F[p * W + j] + e) || h == hu + g ||
                (ext_up = h == O[p * W + j] + c) || h == hu + q;
        }
        if (hit) { prev_i = p; prev_j = j; found = true; break; }
      }
    }
    if (!found) {
      const std::int32_t hl = j > 0 ? H[i * W + j - 1] : fcH[i];
      bool hit;
      if (S == kLinear) {
        hit = h == hl + g;
      } else {
        const std::int32_t el = j > 0 ? E[i * W + j - 1] : fcH[i];
        if (S == kAffine) {
          hit = (ext_left = h == el + e) || h == hl + g;
        } else {
          const std::int32_t ql = j > 0 ? Q[i * W + j - 1] : fcH[i];
          hit = (ext_left = h == el + e) || h == hl + g ||
                (ext_left = h == ql + c) || h == hl + q;
        }
      }
      if (hit) { prev_i = i; prev_j = j - 1; found = true; }
    }
    emit(i == prev_i ? -1 : node_id(i), j == prev_j ? -1 : static_cast<std::int32_t>(j));
    i = prev_i;
    j = prev_j;
    if (ext_left) {
      // :1484-1510 / :1988-2037; the convex loop stops when either the E
      // or the Q chain breaks
      while (true) {
        emit(-1, static_cast<std::int32_t>(j));
        --j;
        if (j == -1) break;
        if (E[i * W + j] + e != E[i * W + j + 1]) break;
        if (S == kConvex && Q[i * W + j] + c != Q[i * W + j + 1]) break;
      }
    } else if (ext_up) {
      // :1511-1547 affine / :2038-2093 convex; sources end at row 0
      while (true) {
        bool stop;
        prev_i = 0;
        const std::int32_t fv = F[i * W + j];
        if (S == kAffine) {
          stop = false;
          for (auto p : real_preds[i]) {
            if ((stop = fv == H[p * W + j] + g) || fv == F[p * W + j] + e) { prev_i = p; break; }
          }
        } else {
          const std::int32_t ov = O[i * W + j];
          stop = true;
          for (auto p : real_preds[i]) {
            if (fv == F[p * W + j] + e || ov == O[p * W + j] + c) { prev_i = p; stop = false; break; }
          }
          if (stop) {
            for (auto p : real_preds[i]) {
              if (fv == H[p * W + j] + g || ov == H[p * W + j] + q) { prev_i = p; break; }
            }
          }
        }
        emit(node_id(i), -1);
        i = prev_i;
        if (stop || i == 0) break;
      }
    }
  }

  // NW tail: finish along row 0 / column 0 (:1102-1124, :1552-1574, :2098-2121)
  if (T == kNW) {
    while (i == 0 && j != -1) { emit(-1, static_cast<std::int32_t>(j)); --j; }
    while (i != 0 && j == -1) {
      emit(node_id(i), -1);
      if (real_preds[i].empty()) { i = 0; continue; }
      std::int64_t next = i;
      for (auto p : real_preds[i]) {
        bool hit;
        if (S == kLinear) hit = fcH[i] == fcH[p] + g;
        else if (S == kAffine) hit = fcH[i] == fcH[p] + e;
        else hit = fcF[i] == fcF[p] + e || fcO[i] == fcO[p] + c;
        if (hit) { next = p; break; }
      }
      if (next == i) break;  // no consistent predecessor (cannot happen on valid input)
      i = next;
    }
  }

  // reverse pair order
  std::vector<std::int32_t> rev(al.size());
  const std::size_t np = al.size() / 2;
  for (std::size_t k = 0; k < np; ++k) {
    rev[2 * k] = al[2 * (np - 1 - k)];
    rev[2 * k + 1] = al[2 * (np - 1 - k) + 1];
  }
  al.swap(rev);
  return res;
}

// ---------------------------------------------------------- SISD engine --
// The reference's scalar engine, SisdAlignmentEngine
// (src/sisd_alignment_engine.cpp), which AlignmentEngine::Create returns
// when the library is built without SIMD (src/alignment_engine.cpp:68-71).
// Its DP is on (V+1) x (n+1) int32 matrices: column 0 is the first column,
// row 0 the virtual source row.  Against the SIMD engine it differs in
// three places: convex E/Q open from the final H of every column (full
// Gotoh, :751-768), the end cell (no 0 floor, no padding columns,
// :296-311 and :354-364), and the traceback (the convex left walk goes on
// while either gap chain holds, :880-887; the NW walk runs through row 0
// and column 0 in the main loop, not a separate tail).
AlignResult SisdAlign(const Engine& en, const Graph& gr, const char* seq, std::uint32_t len) {
  AlignResult res;
  // :258-267
  if (len > static_cast<std::uint32_t>(std::numeric_limits<std::int32_t>::max())) {
    throw std::invalid_argument("[spoa::SisdAlignmentEngine::Align] error: too large sequence!");
  }
  const std::uint32_t V = gr.V();
  if (V == 0 || len == 0) return res;
  // :269-272: the bound on (n, V), not (n + 8, V); always int32
  if (en.WorstCase(len, V) < std::numeric_limits<std::int32_t>::min() + 1024) {
    throw std::invalid_argument("[spoa::SisdAlignmentEngine::Align] error: possible overflow!");
  }
  res.width_bits = 32;
  const std::int32_t NINF = std::numeric_limits<std::int32_t>::min() + 1024;  // :13-14
  const std::int64_t W = static_cast<std::int64_t>(len) + 1, R = V + 1;
  const int T = en.type, S = en.subtype;
  const std::int32_t m = en.m, n = en.n, g = en.g, e = en.e, q = en.q, c = en.c;

  std::vector<std::uint32_t> rank_of(V, 0);
  for (std::uint32_t r = 0; r < V; ++r) rank_of[gr.rank_to_node[r]] = r;
  // predecessor rows in in-edge order; a source's only predecessor is row 0
  std::vector<std::vector<std::int64_t>> in_rows(R), preds(R);
  std::vector<bool> sink(R, false);
  std::vector<std::int32_t> code(R, 0);
  for (std::int64_t i = 1; i < R; ++i) {
    const std::uint32_t id = gr.rank_to_node[i - 1];
    for (auto ed : gr.in[id]) in_rows[i].push_back(rank_of[gr.tail[ed]] + 1);
    preds[i] = in_rows[i];
    if (preds[i].empty()) preds[i].push_back(0);
    code[i] = static_cast<std::int32_t>(gr.code[id]);
    sink[i] = gr.out[id].empty();
  }
  auto node_id = [&](std::int64_t row) { return static_cast<std::int32_t>(gr.rank_to_node[row - 1]); };
  // :123-130 profile: column 0 scores 0, column j the match of seq[j - 1]
  std::vector<std::int32_t> profile(static_cast<std::size_t>(gr.num_codes) * W, 0);
  for (std::uint32_t k = 0; k < gr.num_codes; ++k) {
    const char b = static_cast<char>(gr.decoder[k]);
    for (std::int64_t j = 1; j < W; ++j) profile[k * W + j] = b == seq[j - 1] ? m : n;
  }
  auto prof = [&](std::int64_t i, std::int64_t j) { return profile[code[i] * W + j]; };

  std::vector<std::int32_t> H(R * W, 0), F, E, O, Q;
  if (S != kLinear) { F.assign(R * W, 0); E.assign(R * W, 0); }
  if (S == kConvex) { O.assign(R * W, 0); Q.assign(R * W, 0); }
  auto at = [&](std::vector<std::int32_t>& M, std::int64_t i, std::int64_t j) -> std::int32_t& { return M[i * W + j]; };

  // :137-182 secondary matrices: row 0 and column 0
  if (S == kConvex) {
    O[0] = 0;
    Q[0] = 0;
    for (std::int64_t j = 1; j < W; ++j) { O[j] = NINF; Q[j] = q + static_cast<std::int32_t>(j - 1) * c; }
    for (std::int64_t i = 1; i < R; ++i) {
      std::int32_t pen = in_rows[i].empty() ? q - c : NINF;
      for (auto p : in_rows[i]) pen = std::max(pen, at(O, p, 0));
      at(O, i, 0) = pen + c;
      at(Q, i, 0) = NINF;
    }
  }
  if (S != kLinear) {
    F[0] = 0;
    E[0] = 0;
    for (std::int64_t j = 1; j < W; ++j) { F[j] = NINF; E[j] = g + static_cast<std::int32_t>(j - 1) * e; }
    for (std::int64_t i = 1; i < R; ++i) {
      std::int32_t pen = in_rows[i].empty() ? g - e : NINF;
      for (auto p : in_rows[i]) pen = std::max(pen, at(F, p, 0));
      at(F, i, 0) = pen + e;
      at(E, i, 0) = NINF;
    }
  }
  H[0] = 0;
  // :185-252 primary matrix: row 0 and column 0 per type
  if (T == kSW) {
    // all zero already
  } else {
    for (std::int64_t j = 1; j < W; ++j) {
      if (S == kConvex) H[j] = std::max(Q[j], E[j]);
      else if (S == kAffine) H[j] = E[j];
      else H[j] = static_cast<std::int32_t>(j) * g;
    }
    for (std::int64_t i = 1; i < R; ++i) {
      if (T == kOV) {
        at(H, i, 0) = 0;
      } else if (S == kConvex) {
        at(H, i, 0) = std::max(at(O, i, 0), at(F, i, 0));
      } else if (S == kAffine) {
        at(H, i, 0) = at(F, i, 0);
      } else {
        std::int32_t pen = in_rows[i].empty() ? 0 : NINF;
        for (auto p : in_rows[i]) pen = std::max(pen, at(H, p, 0));
        at(H, i, 0) = pen + g;
      }
    }
  }

  // DP rows in rank order (:321-352 linear, :487-542 affine, :703-772 convex)
  std::int32_t max_score = T == kSW ? 0 : NINF;
  std::int64_t max_i = 0, max_j = 0;
  for (std::int64_t i = 1; i < R; ++i) {
    for (std::size_t k = 0; k < preds[i].size(); ++k) {
      const std::int64_t p = preds[i][k];
      for (std::int64_t j = 1; j < W; ++j) {
        const std::int32_t dg = at(H, p, j - 1) + prof(i, j);
        if (S == kLinear) {
          const std::int32_t v = std::max(dg, at(H, p, j) + g);
          at(H, i, j) = k == 0 ? v : std::max(at(H, i, j), v);
        } else {
          const std::int32_t fv = std::max(at(H, p, j) + g, at(F, p, j) + e);
          at(F, i, j) = k == 0 ? fv : std::max(at(F, i, j), fv);
          if (S == kConvex) {
            const std::int32_t ov = std::max(at(H, p, j) + q, at(O, p, j) + c);
            at(O, i, j) = k == 0 ? ov : std::max(at(O, i, j), ov);
          }
          at(H, i, j) = k == 0 ? dg : std::max(at(H, i, j), dg);
        }
      }
    }
    for (std::int64_t j = 1; j < W; ++j) {
      // the horizontal gap opens from the final (clamped) H of column j - 1
      std::int32_t h;
      if (S == kLinear) {
        h = std::max(at(H, i, j - 1) + g, at(H, i, j));
      } else {
        at(E, i, j) = std::max(at(H, i, j - 1) + g, at(E, i, j - 1) + e);
        if (S == kAffine) {
          h = std::max(at(H, i, j), std::max(at(F, i, j), at(E, i, j)));
        } else {
          at(Q, i, j) = std::max(at(H, i, j - 1) + q, at(Q, i, j - 1) + c);
          h = std::max(at(H, i, j), std::max(std::max(at(F, i, j), at(E, i, j)), std::max(at(O, i, j), at(Q, i, j))));
        }
      }
      if (T == kSW) h = std::max(h, 0);
      at(H, i, j) = h;
      // first strictly better cell in row-major order
      const bool cand = T == kSW || (sink[i] && (T == kOV || j == W - 1));
      if (cand && max_score < h) {
        max_score = h;
        max_i = i;
        max_j = j;
      }
    }
  }
  if (max_i == 0 && max_j == 0) return res;  // :354-356: nothing, score untouched
  res.has_score = true;
  res.score = max_score;

  // traceback (:361-437 linear, :547-676 affine, :777-923 convex)
  std::vector<std::int32_t>& al = res.pairs;
  std::int64_t i = max_i, j = max_j, prev_i = 0, prev_j = 0;
  auto go_on = [&]() {
    if (T == kSW) return at(H, i, j) != 0;
    if (T == kNW) return !(i == 0 && j == 0);
    return !(i == 0 || j == 0);
  };
  // guards a walk that finds no move (stale prev_*, as the reference keeps
  // them) from looping forever on an inconsistent matrix
  std::size_t steps = 0;
  const std::size_t max_steps = static_cast<std::size_t>(R + W) * 4;
  while (go_on()) {
    if (++steps > max_steps) throw std::runtime_error("oracle: inconsistent SISD traceback");
    const std::int32_t h = at(H, i, j);
    bool found = false, ext_left = false, ext_up = false;
    if (i != 0 && j != 0) {
      for (auto p : preds[i]) {
        if (h == at(H, p, j - 1) + prof(i, j)) {
          prev_i = p;
          prev_j = j - 1;
          found = true;
          break;
        }
      }
    }
    if (!found && i != 0) {
      for (auto p : preds[i]) {
        bool hit;
        if (S == kLinear) {
          hit = h == at(H, p, j) + g;
        } else if (S == kAffine) {
          hit = (ext_up = h == at(F, p, j) + e) || h == at(H, p, j) + g;
        } else {
          hit = (ext_up = ext_up || h == at(F, p, j) + e) || h == at(H, p, j) + g ||
                (ext_up = ext_up || h == at(O, p, j) + c) || h == at(H, p, j) + q;
        }
        if (hit) {
          prev_i = p;
          prev_j = j;
          found = true;
          break;
        }
      }
    }
    if (!found && (S == kLinear || j != 0)) {
      // linear tests the left cell without the column guard (:416); on a
      // consistent matrix the up move always wins in column 0
      const std::int32_t hl = H[i * W + j - 1];
      bool hit;
      if (S == kLinear) {
        hit = h == hl + g;
      } else if (S == kAffine) {
        hit = (ext_left = h == at(E, i, j - 1) + e) || h == hl + g;
      } else {
        hit = (ext_left = h == at(E, i, j - 1) + e) || h == hl + g ||
              (ext_left = ext_left || h == at(Q, i, j - 1) + c) || h == hl + q;
      }
      if (hit) {
        prev_i = i;
        prev_j = j - 1;
        found = true;
      }
    }
    al.push_back(i == prev_i ? -1 : node_id(i));
    al.push_back(j == prev_j ? -1 : static_cast<std::int32_t>(j - 1));
    i = prev_i;
    j = prev_j;
    if (ext_left) {
      while (true) {
        al.push_back(-1);
        al.push_back(static_cast<std::int32_t>(j - 1));
        --j;
        const bool e_breaks = at(E, i, j) + e != at(E, i, j + 1);
        const bool q_breaks = S != kConvex || at(Q, i, j) + c != at(Q, i, j + 1);
        if (e_breaks && q_breaks) break;
      }
    } else if (ext_up) {
      while (true) {
        bool stop;
        prev_i = 0;
        if (S == kAffine) {
          stop = false;
          for (auto p : in_rows[i]) {
            if ((stop = at(F, i, j) == at(H, p, j) + g) || at(F, i, j) == at(F, p, j) + e) {
              prev_i = p;
              break;
            }
          }
        } else {
          stop = true;
          for (auto p : in_rows[i]) {
            if (at(F, i, j) == at(F, p, j) + e || at(O, i, j) == at(O, p, j) + c) {
              prev_i = p;
              stop = false;
              break;
            }
          }
          if (stop) {
            for (auto p : in_rows[i]) {
              if (at(F, i, j) == at(H, p, j) + g || at(O, i, j) == at(H, p, j) + q) {
                prev_i = p;
                break;
              }
            }
          }
        }
        al.push_back(node_id(i));
        al.push_back(-1);
        i = prev_i;
        if (stop || i == 0) break;
      }
    }
  }
  // reverse pair order
  std::vector<std::int32_t> rev(al.size());
  const std::size_t np = al.size() / 2;
  for (std::size_t k = 0; k < np; ++k) {
    rev[2 * k] = al[2 * (np - 1 - k)];
    rev[2 * k + 1] = al[2 * (np - 1 - k) + 1];
  }
  al.swap(rev);
  return res;
}

// The strand-ambiguous caller's reverse complement (src/main.cpp:286, 296
// call biosoup 0.10.0 Sequence::ReverseAndComplement, which the reference
// does not vendor).  One rule shared by the product, this oracle, the
// reference shim and the Python harness: upper-case, then the Watson-Crick
// / IUPAC complement (A-T, C-G, U-A, R-Y, K-M, B-V, D-H; S and W map to
// themselves); every other character is kept as it is.  Parity with
// biosoup is unpinned.
char Complement(char b) {
  switch (b) {
    case 'A': case 'a': return 'T';
    case 'C': case 'c': return 'G';
    case 'G': case 'g': return 'C';
    case 'T': case 't': case 'U': case 'u': return 'A';
    case 'R': case 'r': return 'Y';
    case 'Y': case 'y': return 'R';
    case 'K': case 'k': return 'M';
    case 'M': case 'm': return 'K';
    case 'B': case 'b': return 'V';
    case 'V': case 'v': return 'B';
    case 'D': case 'd': return 'H';
    case 'H': case 'h': return 'D';
    case 'S': case 's': return 'S';
    case 'W': case 'w': return 'W';
    default: return b;
  }
}

std::string ReverseComplement(const char* s, std::size_t len) {
  std::string out(len, 'N');
  for (std::size_t i = 0; i < len; ++i) out[len - 1 - i] = Complement(s[i]);
  return out;
}

}  // namespace oracle

// ------------------------------------------------------------- C ABI ----
using oracle::Engine;
using oracle::Graph;

extern "C" {

const char* oracle_last_error() { return oracle::g_error.c_str(); }

void* oracle_engine_new(int type, int m, int n, int g, int e, int q, int c) {
  try {
    return oracle::CreateEngine(type, m, n, g, e, q, c);
  } catch (std::exception& ex) {
    oracle::g_error = ex.what();
    return nullptr;
  }
}

// The scalar engine: the same validation and normalisation as Create, then
// SisdAlignmentEngine semantics for every Align (src/alignment_engine.cpp:68-71).
void* oracle_sisd_engine_new(int type, int m, int n, int g, int e, int q, int c) {
  try {
    Engine* en = oracle::CreateEngine(type, m, n, g, e, q, c);
    en->sisd = true;
    return en;
  } catch (std::exception& ex) {
    oracle::g_error = ex.what();
    return nullptr;
  }
}

void oracle_engine_free(void* p) { delete static_cast<Engine*>(p); }

int oracle_engine_subtype(void* p) { return static_cast<Engine*>(p)->subtype; }

// src/simd_alignment_engine_implementation.hpp:381-422 (Prealloc): a length
// beyond int32 is rejected; a bound past int32 returns without sizing
// anything; otherwise the reference sizes its buffers for that width (the
// oracle holds none).  Returns the width it would size for (0, 16 or 32),
// -1 on error.
int oracle_engine_prealloc(void* p, std::uint32_t max_len, std::uint32_t alphabet) {
  if (max_len > static_cast<std::uint32_t>(std::numeric_limits<std::int32_t>::max())) {
    oracle::g_error = "[spoa::SimdAlignmentEngine::Prealloc] error: too large sequence!";
    return -1;
  }
  const std::int64_t w = static_cast<Engine*>(p)->WorstCase(static_cast<std::int64_t>(max_len) + 8,
                                                            static_cast<std::int64_t>(max_len) * (alphabet & 0xffu));
  if (w < std::numeric_limits<std::int32_t>::min() + 1024) return 0;
  return w < std::numeric_limits<std::int16_t>::min() + 1024 ? 32 : 16;
}

void* oracle_graph_new() { return new Graph(); }
void oracle_graph_free(void* p) { delete static_cast<Graph*>(p); }
void oracle_graph_clear(void* p) { static_cast<Graph*>(p)->Clear(); }
std::uint32_t oracle_graph_num_nodes(void* p) { return static_cast<Graph*>(p)->V(); }

// Align; pairs_out holds cap pairs (2*cap int32).  Returns the number of
// pairs, or -1 on error (message in oracle_last_error), or -2 if cap is too
// small.  *score is written only when the reference writes it.
int oracle_align(void* eng, void* graph, const char* seq, std::uint32_t len,
                 std::int32_t* pairs_out, std::uint32_t cap, std::int32_t* score,
                 int* has_score, int* width_bits) {
  try {
    const Engine& en = *static_cast<Engine*>(eng);
    auto r = en.sisd ? oracle::SisdAlign(en, *static_cast<Graph*>(graph), seq, len)
                     : oracle::Align(en, *static_cast<Graph*>(graph), seq, len);
    std::uint32_t np = static_cast<std::uint32_t>(r.pairs.size() / 2);
    if (np > cap) return -2;
    if (np > 0) std::memcpy(pairs_out, r.pairs.data(), r.pairs.size() * sizeof(std::int32_t));
    if (has_score) *has_score = r.has_score ? 1 : 0;
    if (r.has_score && score) *score = r.score;
    if (width_bits) *width_bits = r.width_bits;
    return static_cast<int>(np);
  } catch (std::exception& ex) {
    oracle::g_error = ex.what();
    return -1;
  }
}

// weights: len entries, or null for weight 1 everywhere
int oracle_graph_add_alignment(void* graph, const std::int32_t* pairs, std::uint32_t n_pairs,
                               const char* seq, std::uint32_t len,
                               const std::uint32_t* weights, std::uint32_t n_weights) {
  try {
    std::vector<std::uint32_t> w;
    if (weights) w.assign(weights, weights + n_weights);
    else w.assign(len, 1);
    static_cast<Graph*>(graph)->AddAlignment(pairs, n_pairs, seq, len, w);
    return 0;
  } catch (std::exception& ex) {
    oracle::g_error = ex.what();
    return -1;
  }
}

// Returns the consensus length; writes up to cap bytes.
int oracle_graph_consensus(void* graph, char* out, std::uint32_t cap) {
  std::string s = static_cast<Graph*>(graph)->GenerateConsensus();
  if (s.size() > cap) return -2;
  std::memcpy(out, s.data(), s.size());
  return static_cast<int>(s.size());
}

// MSA rows are written back to back (rows * cols bytes).
int oracle_graph_msa(void* graph, int include_consensus, char* out, std::uint64_t cap,
                     std::uint32_t* rows, std::uint32_t* cols) {
  auto m = static_cast<Graph*>(graph)->GenerateMsa(include_consensus != 0);
  *rows = static_cast<std::uint32_t>(m.size());
  *cols = m.empty() ? 0 : static_cast<std::uint32_t>(m[0].size());
  std::uint64_t need = static_cast<std::uint64_t>(*rows) * *cols;
  if (need > cap) return -2;
  for (std::size_t r = 0; r < m.size(); ++r) std::memcpy(out + r * (*cols), m[r].data(), *cols);
  return 0;
}

// rank_to_node ids (V entries)
int oracle_graph_ranks(void* graph, std::uint32_t* out, std::uint32_t cap) {
  auto& g = *static_cast<Graph*>(graph);
  if (g.rank_to_node.size() > cap) return -2;
  std::copy(g.rank_to_node.begin(), g.rank_to_node.end(), out);
  return static_cast<int>(g.rank_to_node.size());
}

}  // extern "C"

// One whole window in the oracle, for the CPU baseline: Align +
// AddAlignment for every read in order (with the strand-ambiguous rule of
// src/main.cpp:285-302 when strand != 0), then the consensus.  Reads are
// seqs[read_off[k], read_off[k+1]).  Returns the consensus length, -1 on
// error, -2 if cap is too small.
extern "C" int oracle_reverse_complement(const char* seq, std::uint32_t len, char* out) {
  const std::string rc = oracle::ReverseComplement(seq, len);
  if (len) std::memcpy(out, rc.data(), len);
  return static_cast<int>(len);
}

extern "C" int oracle_run_window(int type, int m, int n, int g, int e, int q, int c, std::uint32_t n_reads,
                                 const std::uint64_t* read_off, const char* seqs, int strand, char* out,
                                 std::uint32_t cap) {
  try {
    std::unique_ptr<Engine> en(oracle::CreateEngine(type, m, n, g, e, q, c));
    Graph gr;
    std::vector<std::int32_t> rc_pairs;
    for (std::uint32_t k = 0; k < n_reads; ++k) {
      const char* s = seqs + read_off[k];
      const std::uint32_t len = static_cast<std::uint32_t>(read_off[k + 1] - read_off[k]);
      auto r = oracle::Align(*en, gr, s, len);
      const char* use = s;
      std::string rev;
      if (strand) {
        rev = oracle::ReverseComplement(s, len);
        auto rr = oracle::Align(*en, gr, rev.data(), len);
        const std::int32_t sf = r.has_score ? r.score : 0;
        const std::int32_t sr = rr.has_score ? rr.score : 0;
        if (!(sf >= sr)) {
          r = rr;
          use = rev.data();
        }
      }
      std::vector<std::uint32_t> w(len, 1);
      gr.AddAlignment(r.pairs.data(), static_cast<std::uint32_t>(r.pairs.size() / 2), use, len, w);
    }
    std::string cons = gr.GenerateConsensus();
    volatile std::size_t rows = gr.GenerateMsa(false).size();  // a window ends with its MSA too
    (void)rows;
    if (cons.size() > cap) return -2;
    std::memcpy(out, cons.data(), cons.size());
    return static_cast<int>(cons.size());
  } catch (std::exception& ex) {
    oracle::g_error = ex.what();
    return -1;
  }
}

// Graph structure for brute-force checks: decoded base of every node id and
// the edges (tail, head) in creation order.  Returns the edge count, -2 if
// cap (edges) is too small.
extern "C" int oracle_graph_structure(void* graph, char* node_chars, std::uint32_t* tails, std::uint32_t* heads,
                                      std::uint32_t cap) {
  auto& g = *static_cast<Graph*>(graph);
  for (std::uint32_t v = 0; v < g.V(); ++v) node_chars[v] = static_cast<char>(g.decoder[g.code[v]]);
  if (g.tail.size() > cap) return -2;
  for (std::size_t k = 0; k < g.tail.size(); ++k) {
    tails[k] = g.tail[k];
    heads[k] = g.head[k];
  }
  return static_cast<int>(g.tail.size());
}
