// TEST INFRASTRUCTURE ONLY.
//
// A C shim over the reference spoa library (built from the sources under
// /root/reference by oracle/Makefile into oracle/_ref/).  It exposes the
// same entry points as the oracle (ref_* instead of oracle_*), so tests can
// drive the reference, the oracle and the product through one harness.
// Bound reference interfaces: AlignmentEngine::Create / Align
// (include/spoa/alignment_engine.hpp:35-69), Graph::AddAlignment /
// GenerateConsensus / GenerateMultipleSequenceAlignment / rank_to_node
// (include/spoa/graph.hpp:106-171).

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "spoa/spoa.hpp"
#include "sisd_alignment_engine.hpp"  // src/: the scalar engine, reached as the reference library does

namespace {
thread_local std::string g_error;

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

}  // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

void* ref_engine_new(int type, int m, int n, int g, int e, int q, int c) {
  try {
    return spoa::AlignmentEngine::Create(static_cast<spoa::AlignmentType>(type), m, n, g, e, q, c)
        .release();
  } catch (std::exception& ex) {
    g_error = ex.what();
    return nullptr;
  }
}

// The reference's scalar engine (src/sisd_alignment_engine.cpp:16-27), the
// one AlignmentEngine::Create returns when the library is built without
// SIMD (src/alignment_engine.cpp:68-71, e.g. the default x86 CMake flags).
// Create itself validates the parameters (same messages); the subtype
// choice and normalisation below are those of src/alignment_engine.cpp:57-66.
void* ref_sisd_engine_new(int type, int m, int n, int g, int e, int q, int c) {
  try {
    spoa::AlignmentEngine::Create(static_cast<spoa::AlignmentType>(type), m, n, g, e, q, c);
    auto sub = g >= e ? spoa::AlignmentSubtype::kLinear
                      : ((g <= q || e >= c) ? spoa::AlignmentSubtype::kAffine : spoa::AlignmentSubtype::kConvex);
    if (sub == spoa::AlignmentSubtype::kLinear) {
      e = g;
    } else if (sub == spoa::AlignmentSubtype::kAffine) {
      q = g;
      c = e;
    }
    return spoa::SisdAlignmentEngine::Create(static_cast<spoa::AlignmentType>(type), sub, m, n, g, e, q, c)
        .release();
  } catch (std::exception& ex) {
    g_error = ex.what();
    return nullptr;
  }
}

void ref_engine_free(void* p) { delete static_cast<spoa::AlignmentEngine*>(p); }

int ref_engine_prealloc(void* p, std::uint32_t max_len, std::uint32_t alphabet) {
  try {
    static_cast<spoa::AlignmentEngine*>(p)->Prealloc(max_len, static_cast<std::uint8_t>(alphabet));
    return 0;
  } catch (std::exception& ex) {
    g_error = ex.what();
    return -1;
  }
}

void* ref_graph_new() { return new spoa::Graph(); }
void ref_graph_free(void* p) { delete static_cast<spoa::Graph*>(p); }
void ref_graph_clear(void* p) { static_cast<spoa::Graph*>(p)->Clear(); }
std::uint32_t ref_graph_num_nodes(void* p) {
  return static_cast<std::uint32_t>(static_cast<spoa::Graph*>(p)->nodes().size());
}

int ref_align(void* eng, void* graph, const char* seq, std::uint32_t len,
              std::int32_t* pairs_out, std::uint32_t cap, std::int32_t* score,
              int* has_score, int* width_bits) {
  try {
    const std::int32_t kUnset = INT32_MIN;
    std::int32_t s = kUnset;
    auto a = static_cast<spoa::AlignmentEngine*>(eng)->Align(
        seq, len, *static_cast<spoa::Graph*>(graph), &s);
    if (a.size() > cap) return -2;
    for (std::size_t k = 0; k < a.size(); ++k) {
      pairs_out[2 * k] = a[k].first;
      pairs_out[2 * k + 1] = a[k].second;
    }
    if (has_score) *has_score = s != kUnset ? 1 : 0;
    if (s != kUnset && score) *score = s;
    if (width_bits) *width_bits = 0;  // not observable through the public API
    return static_cast<int>(a.size());
  } catch (std::exception& ex) {
    g_error = ex.what();
    return -1;
  }
}

int ref_graph_add_alignment(void* graph, const std::int32_t* pairs, std::uint32_t n_pairs,
                            const char* seq, std::uint32_t len,
                            const std::uint32_t* weights, std::uint32_t n_weights) {
  try {
    spoa::Alignment a(n_pairs);
    for (std::uint32_t k = 0; k < n_pairs; ++k) a[k] = {pairs[2 * k], pairs[2 * k + 1]};
    auto& g = *static_cast<spoa::Graph*>(graph);
    if (weights) {
      std::vector<std::uint32_t> w(weights, weights + n_weights);
      g.AddAlignment(a, seq, len, w);
    } else {
      g.AddAlignment(a, seq, len);
    }
    return 0;
  } catch (std::exception& ex) {
    g_error = ex.what();
    return -1;
  }
}

int ref_graph_consensus(void* graph, char* out, std::uint32_t cap) {
  std::string s = static_cast<spoa::Graph*>(graph)->GenerateConsensus();
  if (s.size() > cap) return -2;
  std::memcpy(out, s.data(), s.size());
  return static_cast<int>(s.size());
}

int ref_graph_msa(void* graph, int include_consensus, char* out, std::uint64_t cap,
                  std::uint32_t* rows, std::uint32_t* cols) {
  auto m = static_cast<spoa::Graph*>(graph)->GenerateMultipleSequenceAlignment(include_consensus != 0);
  *rows = static_cast<std::uint32_t>(m.size());
  *cols = m.empty() ? 0 : static_cast<std::uint32_t>(m[0].size());
  std::uint64_t need = static_cast<std::uint64_t>(*rows) * *cols;
  if (need > cap) return -2;
  for (std::size_t r = 0; r < m.size(); ++r) std::memcpy(out + r * (*cols), m[r].data(), *cols);
  return 0;
}

int ref_graph_ranks(void* graph, std::uint32_t* out, std::uint32_t cap) {
  const auto& r = static_cast<spoa::Graph*>(graph)->rank_to_node();
  if (r.size() > cap) return -2;
  for (std::size_t k = 0; k < r.size(); ++k) out[k] = r[k]->id;
  return static_cast<int>(r.size());
}

int ref_reverse_complement(const char* seq, std::uint32_t len, char* out) {
  const std::string rc = ReverseComplement(seq, len);
  if (len) std::memcpy(out, rc.data(), len);
  return static_cast<int>(len);
}

// Graph::Subgraph through the reference (include/spoa/graph.hpp:173-176).
void* ref_graph_subgraph(void* graph, std::uint32_t begin, std::uint32_t end, std::int32_t* sub_to_graph,
                         std::uint32_t cap) {
  try {
    const auto& g = *static_cast<spoa::Graph*>(graph);
    if (cap < g.nodes().size()) return nullptr;
    std::vector<const spoa::Graph::Node*> map;
    auto* sub = new spoa::Graph(g.Subgraph(begin, end, &map));
    for (std::size_t k = 0; k < map.size(); ++k) sub_to_graph[k] = map[k] ? static_cast<std::int32_t>(map[k]->id) : -1;
    return sub;
  } catch (std::exception& ex) {
    g_error = ex.what();
    return nullptr;
  }
}

// Graph::UpdateAlignment through the reference (graph.hpp:178-180).
int ref_graph_update_alignment(void* graph, const std::int32_t* sub_to_graph, std::uint32_t n_map,
                               std::int32_t* pairs, std::uint32_t n_pairs) {
  const auto& g = *static_cast<spoa::Graph*>(graph);
  std::vector<const spoa::Graph::Node*> map(n_map, nullptr);
  for (std::uint32_t k = 0; k < n_map; ++k) {
    if (sub_to_graph[k] >= 0) map[k] = g.nodes()[sub_to_graph[k]].get();
  }
  spoa::Alignment a(n_pairs);
  for (std::uint32_t k = 0; k < n_pairs; ++k) a[k] = {pairs[2 * k], pairs[2 * k + 1]};
  g.UpdateAlignment(map, &a);
  for (std::uint32_t k = 0; k < n_pairs; ++k) pairs[2 * k] = a[k].first;
  return 0;
}

int ref_graph_print_dot(void* graph, const char* path) {
  static_cast<spoa::Graph*>(graph)->PrintDot(path);
  return 0;
}

int ref_graph_consensus_summary(void* graph, int verbose, char* out, std::uint32_t cap, std::uint32_t* summary,
                                std::uint32_t summary_cap, std::uint32_t* summary_len) {
  std::vector<std::uint32_t> sum;
  const std::string s = static_cast<spoa::Graph*>(graph)->GenerateConsensus(&sum, verbose != 0);
  *summary_len = static_cast<std::uint32_t>(sum.size());
  if (s.size() > cap || sum.size() > summary_cap) return -2;
  std::memcpy(out, s.data(), s.size());
  if (!sum.empty()) std::memcpy(summary, sum.data(), sum.size() * sizeof(std::uint32_t));
  return static_cast<int>(s.size());
}

// The GFA of the reference's command-line tool, restated (its printer lives
// in src/main.cpp:119-199, which needs the un-vendored parsers to build):
// segments, links with weights, one path per sequence, the consensus path.
std::int64_t ref_graph_gfa(void* graph, const char* headers, const std::uint8_t* reversed, std::uint32_t n_reversed,
                           int include_consensus, char* out, std::uint64_t cap) {
  auto& g = *static_cast<spoa::Graph*>(graph);
  g.GenerateConsensus();
  std::vector<std::string> names;
  const char* h = headers;
  for (std::size_t k = 0; k < g.sequences().size(); ++k) {
    names.emplace_back(h);
    h += names.back().size() + 1;
  }
  std::vector<bool> cons(g.nodes().size(), false);
  for (const auto* nd : g.consensus()) cons[nd->id] = true;
  std::string t = "H\tVN:Z:1.0\n";
  for (const auto& nd : g.nodes()) {
    t += "S\t" + std::to_string(nd->id + 1) + "\t" + static_cast<char>(g.decoder(nd->code));
    if (cons[nd->id]) t += "\tic:Z:true";
    t += "\n";
    for (const auto* e : nd->outedges) {
      t += "L\t" + std::to_string(nd->id + 1) + "\t+\t" + std::to_string(e->head->id + 1) + "\t+\tOM\tew:f:" +
           std::to_string(e->weight);
      if (cons[nd->id] && cons[e->head->id]) t += "\tic:Z:true";
      t += "\n";
    }
  }
  for (std::uint32_t i = 0; i < g.sequences().size(); ++i) {
    t += "P\t" + names[i] + "\t";
    std::vector<std::uint32_t> path;
    for (auto* nd = g.sequences()[i]; nd; nd = nd->Successor(i)) path.push_back(nd->id + 1);
    const bool rev = reversed && i < n_reversed && reversed[i];
    if (rev) std::reverse(path.begin(), path.end());
    for (std::size_t j = 0; j < path.size(); ++j) {
      if (j) t += ",";
      t += std::to_string(path[j]) + (rev ? "-" : "+");
    }
    t += "\t*\n";
  }
  if (include_consensus) {
    t += "P\tConsensus\t";
    for (std::size_t j = 0; j < g.consensus().size(); ++j) {
      if (j) t += ",";
      t += std::to_string(g.consensus()[j]->id + 1) + "+";
    }
    t += "\t*\n";
  }
  if (t.size() > cap) return -2 - static_cast<std::int64_t>(t.size());
  std::memcpy(out, t.data(), t.size());
  return static_cast<std::int64_t>(t.size());
}

// One window through the reference: every read aligned and added in order
// (strand mode: the read or its reverse complement, whichever aligns with
// the higher score, forward on ties, as src/main.cpp:285-302 does), then
// the consensus and the MSA.  Writes the consensus; returns its length,
// -2 if cap is too small, -1 on error.  The CPU-baseline timing entry.
int ref_run_window(int type, int m, int n, int g, int e, int q, int c, std::uint32_t n_reads,
                   const std::uint64_t* read_off, const char* seqs, int strand, char* out, std::uint32_t cap) {
  try {
    auto eng = spoa::AlignmentEngine::Create(static_cast<spoa::AlignmentType>(type), m, n, g, e, q, c);
    spoa::Graph gr;
    for (std::uint32_t k = 0; k < n_reads; ++k) {
      const char* s = seqs + read_off[k];
      const std::uint32_t len = static_cast<std::uint32_t>(read_off[k + 1] - read_off[k]);
      std::int32_t sf = 0;
      auto a = eng->Align(s, len, gr, &sf);
      if (!strand) {
        gr.AddAlignment(a, s, len);
        continue;
      }
      const std::string rev = ReverseComplement(s, len);
      std::int32_t sr = 0;
      auto ar = eng->Align(rev.data(), len, gr, &sr);
      if (sf >= sr) {
        gr.AddAlignment(a, s, len);
      } else {
        gr.AddAlignment(ar, rev.data(), len);
      }
    }
    std::string cons = gr.GenerateConsensus();
    volatile std::size_t rows = gr.GenerateMultipleSequenceAlignment(false).size();
    (void)rows;
    if (cons.size() > cap) return -2;
    std::memcpy(out, cons.data(), cons.size());
    return static_cast<int>(cons.size());
  } catch (std::exception& ex) {
    g_error = ex.what();
    return -1;
  }
}

}  // extern "C"
