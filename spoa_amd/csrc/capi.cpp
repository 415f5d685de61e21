// C entry points over the C++ drop-in (spoa::AlignmentEngine, spoa::Graph),
// the batched / per-call window drivers and the synthetic window
// generator, for callers that bind C (the Python package, ctypes).
// Declared in include/spoa_capi.h.

#include "spoa_capi.h"

#include <cstring>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "batch.hpp"
#include "io.hpp"
#include "serve.hpp"
#include "poa_graph.hpp"
#include "spoa/spoa.hpp"
#include "spoa_amd_host.hpp"
#include "synth.hpp"

namespace {

thread_local std::string g_error;

int Error(const std::exception& ex) {
  g_error = ex.what();
  return -1;
}

spoa::AlignmentEngine* Eng(void* p) { return static_cast<spoa::AlignmentEngine*>(p); }
spoa::Graph* Gr(void* p) { return static_cast<spoa::Graph*>(p); }

struct BatchHandle {
  std::vector<spoa_amd::WindowResult> results;
  spoa_amd::BatchStats stats;
};

struct SynthHandle {
  std::vector<std::uint32_t> win_read_off;
  std::vector<std::uint64_t> read_off;
  std::string seqs;
  std::string templates;
  std::vector<std::uint64_t> templ_off;
};

spoa_amd::BatchOptions Options(const spoa_batch_opts* o) {
  spoa_amd::BatchOptions b;
  b.type = static_cast<spoa::AlignmentType>(o->type);
  b.m = static_cast<std::int8_t>(o->m);
  b.n = static_cast<std::int8_t>(o->n);
  b.g = static_cast<std::int8_t>(o->g);
  b.e = static_cast<std::int8_t>(o->e);
  b.q = static_cast<std::int8_t>(o->q);
  b.c = static_cast<std::int8_t>(o->c);
  b.strand = o->strand != 0;
  b.msa = o->msa != 0;
  b.threads = o->threads > 0 ? static_cast<unsigned>(o->threads) : 0u;
  b.device = o->device;
  b.overlap = o->overlap != 0;
  b.graph = o->graph != 0;
  b.resident = o->resident != 0;
  b.sisd = o->sisd != 0;
  return b;
}

}  // namespace

extern "C" {

const char* spoa_last_error(void) { return g_error.c_str(); }

// ------------------------------------------------------------ engine --

void* spoa_engine_new(int type, int m, int n, int g, int e, int q, int c) {
  try {
    return spoa::AlignmentEngine::Create(static_cast<spoa::AlignmentType>(type), static_cast<std::int8_t>(m),
                                         static_cast<std::int8_t>(n), static_cast<std::int8_t>(g),
                                         static_cast<std::int8_t>(e), static_cast<std::int8_t>(q),
                                         static_cast<std::int8_t>(c))
        .release();
  } catch (const std::exception& ex) {
    Error(ex);
    return nullptr;
  }
}

void* spoa_sisd_engine_new(int type, int m, int n, int g, int e, int q, int c) {
  try {
    return spoa::CreateEngine(static_cast<spoa::AlignmentType>(type), static_cast<std::int8_t>(m),
                              static_cast<std::int8_t>(n), static_cast<std::int8_t>(g), static_cast<std::int8_t>(e),
                              static_cast<std::int8_t>(q), static_cast<std::int8_t>(c), true)
        .release();
  } catch (const std::exception& ex) {
    Error(ex);
    return nullptr;
  }
}

void spoa_engine_free(void* p) { delete Eng(p); }

int spoa_engine_subtype(void* p) { return static_cast<int>(Eng(p)->subtype()); }

int spoa_engine_prealloc(void* p, uint32_t max_len, uint32_t alphabet) {
  try {
    Eng(p)->Prealloc(max_len, static_cast<std::uint8_t>(alphabet));
    return 0;
  } catch (const std::exception& ex) {
    return Error(ex);
  }
}

int spoa_engine_stats(void* p, spoa_hip_stats* out) {
  auto* hip = dynamic_cast<spoa::HipAlignmentEngine*>(Eng(p));
  if (!hip || !out) return -1;
  return spoa_hip_last_stats(hip->device(), out) == SPOA_HIP_OK ? 0 : -1;
}

spoa_hip_engine* spoa_engine_device(void* p) {
  auto* hip = dynamic_cast<spoa::HipAlignmentEngine*>(Eng(p));
  return hip ? hip->device() : nullptr;
}

int spoa_align(void* eng, void* graph, const char* seq, uint32_t len, int32_t* pairs_out, uint32_t cap,
               int32_t* score, int* has_score, int* width_bits) {
  try {
    const std::int32_t kUnset = INT32_MIN;
    std::int32_t s = kUnset;
    auto a = Eng(eng)->Align(seq, len, *Gr(graph), &s);
    if (a.size() > cap) return -2;
    for (std::size_t k = 0; k < a.size(); ++k) {
      pairs_out[2 * k] = a[k].first;
      pairs_out[2 * k + 1] = a[k].second;
    }
    if (has_score) *has_score = s != kUnset ? 1 : 0;
    if (s != kUnset && score) *score = s;
    if (width_bits) {
      auto* hip = dynamic_cast<spoa::HipAlignmentEngine*>(Eng(eng));
      *width_bits = hip ? hip->last_width() : 0;
    }
    return static_cast<int>(a.size());
  } catch (const std::exception& ex) {
    return Error(ex);
  }
}

// ------------------------------------------------------------- graph --

void* spoa_graph_new(void) { return new spoa::Graph(); }
void spoa_graph_free(void* p) { delete Gr(p); }
void spoa_graph_clear(void* p) { Gr(p)->Clear(); }
uint32_t spoa_graph_num_nodes(void* p) { return static_cast<uint32_t>(Gr(p)->nodes().size()); }

int spoa_graph_add_alignment(void* graph, const int32_t* pairs, uint32_t n_pairs, const char* seq, uint32_t len,
                             const uint32_t* weights, uint32_t n_weights) {
  try {
    spoa::Alignment a(n_pairs);
    for (uint32_t k = 0; k < n_pairs; ++k) a[k] = {pairs[2 * k], pairs[2 * k + 1]};
    if (weights) {
      std::vector<std::uint32_t> w(weights, weights + n_weights);
      Gr(graph)->AddAlignment(a, seq, len, w);
    } else {
      Gr(graph)->AddAlignment(a, seq, len);
    }
    return 0;
  } catch (const std::exception& ex) {
    return Error(ex);
  }
}

int spoa_graph_consensus(void* graph, char* out, uint32_t cap) {
  const std::string s = Gr(graph)->GenerateConsensus();
  if (s.size() > cap) return -2;
  std::memcpy(out, s.data(), s.size());
  return static_cast<int>(s.size());
}

int spoa_graph_msa(void* graph, int include_consensus, char* out, uint64_t cap, uint32_t* rows, uint32_t* cols) {
  auto m = Gr(graph)->GenerateMultipleSequenceAlignment(include_consensus != 0);
  *rows = static_cast<uint32_t>(m.size());
  *cols = m.empty() ? 0 : static_cast<uint32_t>(m[0].size());
  const uint64_t need = static_cast<uint64_t>(*rows) * *cols;
  if (need > cap) return -2;
  for (std::size_t r = 0; r < m.size(); ++r) std::memcpy(out + r * (*cols), m[r].data(), *cols);
  return 0;
}

int spoa_graph_ranks(void* graph, uint32_t* out, uint32_t cap) {
  const auto& r = Gr(graph)->rank_to_node();
  if (r.size() > cap) return -2;
  for (std::size_t k = 0; k < r.size(); ++k) out[k] = r[k]->id;
  return static_cast<int>(r.size());
}

void* spoa_graph_subgraph(void* graph, uint32_t begin, uint32_t end, int32_t* sub_to_graph, uint32_t cap) {
  try {
    const spoa::Graph& g = *Gr(graph);
    if (cap < g.nodes().size()) throw std::invalid_argument("[spoa::Graph::Subgraph] error: map buffer too small");
    std::vector<const spoa::Graph::Node*> map;
    auto* sub = new spoa::Graph(g.Subgraph(begin, end, &map));
    for (std::size_t k = 0; k < map.size(); ++k) sub_to_graph[k] = map[k] ? static_cast<int32_t>(map[k]->id) : -1;
    return sub;
  } catch (const std::exception& ex) {
    Error(ex);
    return nullptr;
  }
}

int spoa_graph_update_alignment(void* graph, const int32_t* sub_to_graph, uint32_t n_map, int32_t* pairs,
                                uint32_t n_pairs) {
  try {
    const spoa::Graph& g = *Gr(graph);
    std::vector<const spoa::Graph::Node*> map(n_map, nullptr);
    for (uint32_t k = 0; k < n_map; ++k) {
      if (sub_to_graph[k] >= 0) {
        if (static_cast<std::size_t>(sub_to_graph[k]) >= g.nodes().size()) {
          throw std::invalid_argument("[spoa::Graph::UpdateAlignment] error: node out of range");
        }
        map[k] = g.nodes()[sub_to_graph[k]].get();
      }
    }
    spoa::Alignment a(n_pairs);
    for (uint32_t k = 0; k < n_pairs; ++k) {
      a[k] = {pairs[2 * k], pairs[2 * k + 1]};
      if (a[k].first >= static_cast<int32_t>(n_map) || (a[k].first >= 0 && !map[a[k].first])) {
        throw std::invalid_argument("[spoa::Graph::UpdateAlignment] error: node not in the subgraph");
      }
    }
    g.UpdateAlignment(map, &a);
    for (uint32_t k = 0; k < n_pairs; ++k) pairs[2 * k] = a[k].first;
    return 0;
  } catch (const std::exception& ex) {
    return Error(ex);
  }
}

int spoa_graph_print_dot(void* graph, const char* path) {
  try {
    Gr(graph)->PrintDot(path ? path : "");
    return 0;
  } catch (const std::exception& ex) {
    return Error(ex);
  }
}

int spoa_graph_consensus_summary(void* graph, int verbose, char* out, uint32_t cap, uint32_t* summary,
                                 uint32_t summary_cap, uint32_t* summary_len) {
  try {
    std::vector<std::uint32_t> sum;
    const std::string s = Gr(graph)->GenerateConsensus(&sum, verbose != 0);
    *summary_len = static_cast<uint32_t>(sum.size());
    if (s.size() > cap || sum.size() > summary_cap) return -2;
    std::memcpy(out, s.data(), s.size());
    std::memcpy(summary, sum.data(), sum.size() * sizeof(uint32_t));
    return static_cast<int>(s.size());
  } catch (const std::exception& ex) {
    return Error(ex);
  }
}

int64_t spoa_graph_gfa(void* graph, const char* headers, const uint8_t* reversed, uint32_t n_reversed,
                       int include_consensus, char* out, uint64_t cap) {
  try {
    spoa::Graph& g = *Gr(graph);
    std::vector<std::string> names;
    const char* h = headers;
    for (std::size_t k = 0; k < g.sequences().size(); ++k) {
      names.emplace_back(h);
      h += names.back().size() + 1;
    }
    std::vector<bool> rev(reversed ? n_reversed : 0);
    for (std::size_t k = 0; k < rev.size(); ++k) rev[k] = reversed[k] != 0;
    g.GenerateConsensus();
    std::ostringstream os;
    if (!spoa_amd::WriteGfa(os, g, names, rev, include_consensus != 0)) {
      throw std::invalid_argument("[spoa::PrintGfa] error: missing header(s)");
    }
    const std::string text = os.str();
    if (text.size() > cap) return -2 - static_cast<int64_t>(text.size());
    std::memcpy(out, text.data(), text.size());
    return static_cast<int64_t>(text.size());
  } catch (const std::exception& ex) {
    return Error(ex);
  }
}

// ----------------------------------------------------- flat POA graph --

void* spoa_poa_graph_new(void) { return new spoa_amd::PoaGraph(); }
void spoa_poa_graph_free(void* g) { delete static_cast<spoa_amd::PoaGraph*>(g); }
uint32_t spoa_poa_graph_num_nodes(void* g) { return static_cast<spoa_amd::PoaGraph*>(g)->num_nodes(); }

int spoa_poa_graph_add_alignment(void* g, const int32_t* pairs, uint32_t n_pairs, const char* seq, uint32_t len,
                                 const uint32_t* weights, uint32_t n_weights) {
  try {
    if (weights && n_weights != len) {
      throw std::invalid_argument(
          "[spoa::Graph::AddAlignment] error: "
          "sequence and weights are of unequal size!");
    }
    std::vector<uint32_t> ones;
    if (!weights) {
      ones.assign(len, 1u);
      weights = ones.data();
    }
    static_cast<spoa_amd::PoaGraph*>(g)->AddAlignment(pairs, n_pairs, seq, len, weights);
    return 0;
  } catch (const std::exception& ex) {
    return Error(ex);
  }
}

int spoa_poa_graph_consensus(void* g, char* out, uint32_t cap) {
  const std::string s = static_cast<spoa_amd::PoaGraph*>(g)->GenerateConsensus();
  if (s.size() > cap) return -2;
  std::memcpy(out, s.data(), s.size());
  return static_cast<int>(s.size());
}

int spoa_poa_graph_msa(void* g, int include_consensus, char* out, uint64_t cap, uint32_t* rows, uint32_t* cols) {
  auto m = static_cast<spoa_amd::PoaGraph*>(g)->GenerateMultipleSequenceAlignment(include_consensus != 0);
  *rows = static_cast<uint32_t>(m.size());
  *cols = m.empty() ? 0 : static_cast<uint32_t>(m[0].size());
  const uint64_t need = static_cast<uint64_t>(*rows) * *cols;
  if (need > cap) return -2;
  for (std::size_t r = 0; r < m.size(); ++r) std::memcpy(out + r * (*cols), m[r].data(), *cols);
  return 0;
}

int spoa_poa_graph_ranks(void* g, uint32_t* out, uint32_t cap) {
  const auto& r = static_cast<spoa_amd::PoaGraph*>(g)->rank_to_node();
  if (r.size() > cap) return -2;
  std::memcpy(out, r.data(), r.size() * sizeof(uint32_t));
  return static_cast<int>(r.size());
}

// ------------------------------------------------------------- batch --

void* spoa_batch_run(int mode, const spoa_batch_opts* opts, uint32_t n_windows, const uint32_t* win_read_off,
                     const uint64_t* read_off, const char* seqs, const char* quals) {
  try {
    spoa_amd::WindowSet ws;
    ws.n_windows = n_windows;
    ws.win_read_off = win_read_off;
    ws.read_off = read_off;
    ws.seqs = seqs;
    ws.quals = quals;
    auto h = std::make_unique<BatchHandle>();
    const spoa_amd::BatchOptions o = Options(opts);
    if (mode == SPOA_BATCH_PER_CALL) {
      h->results = spoa_amd::RunPerCall(ws, o, &h->stats);
    } else {
      h->results = spoa_amd::RunBatched(ws, o, &h->stats);
    }
    return h.release();
  } catch (const std::exception& ex) {
    Error(ex);
    return nullptr;
  }
}

void spoa_batch_free(void* h) { delete static_cast<BatchHandle*>(h); }

void* spoa_batch_context_new(const spoa_batch_opts* opts) {
  try {
    return new spoa_amd::BatchContext(Options(opts));
  } catch (const std::exception& ex) {
    Error(ex);
    return nullptr;
  }
}

void spoa_batch_context_free(void* ctx) { delete static_cast<spoa_amd::BatchContext*>(ctx); }

spoa_hip_engine* spoa_batch_context_device(void* ctx) {
  auto* hip = dynamic_cast<spoa::HipAlignmentEngine*>(static_cast<spoa_amd::BatchContext*>(ctx)->engine());
  return hip ? hip->device() : nullptr;
}

spoa_hip_engine* spoa_batch_context_half_device(void* ctx, int32_t half) {
  if (half < 0 || half > 1) return nullptr;
  auto* hip = dynamic_cast<spoa::HipAlignmentEngine*>(static_cast<spoa_amd::BatchContext*>(ctx)->engine(half));
  return hip ? hip->device() : nullptr;
}

int spoa_batch_context_set_config_split(void* ctx, int32_t width_bits, uint32_t max_width, int32_t cols_per_lane,
                                        int32_t waves, int32_t split) {
  auto* c = static_cast<spoa_amd::BatchContext*>(ctx);
  for (int h = 0; h < 2; ++h) {
    auto* hip = dynamic_cast<spoa::HipAlignmentEngine*>(c->engine(h));
    if (!hip) return -1;
    if (spoa_hip_set_config_split(hip->device(), width_bits, 1, max_width, cols_per_lane, waves, split) !=
        SPOA_HIP_OK) {
      g_error = spoa_hip_last_error();
      return -1;
    }
  }
  return 0;
}

int spoa_batch_context_set_config(void* ctx, int32_t width_bits, uint32_t max_width, int32_t cols_per_lane,
                                  int32_t waves) {
  return spoa_batch_context_set_config_split(ctx, width_bits, max_width, cols_per_lane, waves, 0);
}

int spoa_batch_context_autotune(void* ctx, uint32_t n_windows, const uint32_t* win_read_off, const uint64_t* read_off,
                                const char* seqs, uint32_t rounds, const char* cache_path, int32_t* out,
                                double* out_ms, uint32_t cap) {
  try {
    spoa_amd::WindowSet ws;
    ws.n_windows = n_windows;
    ws.win_read_off = win_read_off;
    ws.read_off = read_off;
    ws.seqs = seqs;
    auto res = spoa_amd::Autotune(ws, *static_cast<spoa_amd::BatchContext*>(ctx), rounds,
                                  cache_path ? std::string(cache_path) : std::string());
    uint32_t n = 0;
    for (const auto& r : res) {
      if (n >= cap) break;
      out[4 * n + 0] = r.width;
      out[4 * n + 1] = r.k;
      out[4 * n + 2] = r.waves;
      out[4 * n + 3] = (r.cached ? 1 : 0) | (r.whole_batch ? 2 : 0);
      out_ms[n] = r.kernel_ms;
      ++n;
    }
    return static_cast<int>(n);
  } catch (const std::exception& ex) {
    Error(ex);
    return -1;
  }
}

void* spoa_batch_context_run(void* ctx, uint32_t n_windows, const uint32_t* win_read_off, const uint64_t* read_off,
                             const char* seqs, const char* quals) {
  try {
    spoa_amd::WindowSet ws;
    ws.n_windows = n_windows;
    ws.win_read_off = win_read_off;
    ws.read_off = read_off;
    ws.seqs = seqs;
    ws.quals = quals;
    auto h = std::make_unique<BatchHandle>();
    h->results = spoa_amd::RunBatched(ws, *static_cast<spoa_amd::BatchContext*>(ctx), &h->stats);
    return h.release();
  } catch (const std::exception& ex) {
    Error(ex);
    return nullptr;
  }
}

int spoa_batch_get_stats(void* h, spoa_batch_stats* out) {
  const auto& s = static_cast<BatchHandle*>(h)->stats;
  out->wall_ms = s.wall_ms;
  out->align_ms = s.align_ms;
  out->kernel_ms = s.kernel_ms;
  out->h2d_ms = s.h2d_ms;
  out->d2h_ms = s.d2h_ms;
  out->graph_ms = s.graph_ms;
  out->finish_ms = s.finish_ms;
  out->pack_ms = s.pack_ms;
  out->alloc_ms = s.alloc_ms;
  out->wait_ms = s.wait_ms;
  out->unpack_ms = s.unpack_ms;
  out->cells = s.cells;
  out->dp_bytes = s.dp_bytes;
  out->align_calls = s.align_calls;
  out->launches = s.launches;
  out->rounds = s.rounds;
  out->kernel_ms16 = s.kernel_ms16;
  out->graph_builds = s.graph_builds;
  out->graph_updates = s.graph_updates;
  out->kernel_ms32 = s.kernel_ms32;
  out->dp_bytes16 = s.dp_bytes16;
  out->dp_bytes32 = s.dp_bytes32;
  out->launches16 = s.launches16;
  out->launches32 = s.launches32;
  return 0;
}

int spoa_batch_consensus(void* h, uint32_t w, char* out, uint32_t cap) {
  const auto& r = static_cast<BatchHandle*>(h)->results;
  if (w >= r.size()) return -1;
  const std::string& s = r[w].consensus;
  if (s.size() > cap) return -2;
  std::memcpy(out, s.data(), s.size());
  return static_cast<int>(s.size());
}

int64_t spoa_batch_consensus_all(void* h, char* out, uint64_t cap, uint64_t* offs) {
  const auto& r = static_cast<BatchHandle*>(h)->results;
  uint64_t total = 0;
  for (std::size_t w = 0; w < r.size(); ++w) {
    offs[w] = total;
    total += r[w].consensus.size();
  }
  offs[r.size()] = total;
  if (total > cap) return -2;
  for (std::size_t w = 0; w < r.size(); ++w) {
    std::memcpy(out + offs[w], r[w].consensus.data(), r[w].consensus.size());
  }
  return static_cast<int64_t>(total);
}

int spoa_batch_msa(void* h, uint32_t w, char* out, uint64_t cap, uint32_t* rows, uint32_t* cols) {
  const auto& r = static_cast<BatchHandle*>(h)->results;
  if (w >= r.size()) return -1;
  const auto& o = r[w];
  *rows = o.msa_rows;
  *cols = o.msa_width;
  if (o.msa.size() > cap) return -2;
  if (!o.msa.empty()) std::memcpy(out, o.msa.data(), o.msa.size());
  return 0;
}

int spoa_batch_reversed(void* h, uint32_t w, uint8_t* out, uint32_t cap) {
  const auto& r = static_cast<BatchHandle*>(h)->results;
  if (w >= r.size()) return -1;
  const auto& v = r[w].reversed;
  if (v.size() > cap) return -2;
  std::memcpy(out, v.data(), v.size());
  return static_cast<int>(v.size());
}

int spoa_reverse_complement(const char* seq, uint32_t len, char* out) {
  if ((!seq || !out) && len) return -1;
  const std::string rc = spoa_amd::ReverseComplement(seq, len);
  if (len) std::memcpy(out, rc.data(), len);
  return static_cast<int>(len);
}

// ----------------------------------------------------------- serving --

void* spoa_server_new(const spoa_serve_opts* o) {
  try {
    spoa_amd::ServeOptions so;
    if (o) {
      so.threads = o->threads > 0 ? static_cast<unsigned>(o->threads) : 0u;
      so.device = o->device;
      if (o->max_active) so.max_active = o->max_active;
      if (o->max_param_sets) so.max_param_sets = o->max_param_sets;
      if (o->memory_fraction > 0) so.memory_fraction = o->memory_fraction;
    }
    return new spoa_amd::Server(so);
  } catch (const std::exception& ex) {
    Error(ex);
    return nullptr;
  }
}

void spoa_server_free(void* srv) { delete static_cast<spoa_amd::Server*>(srv); }

uint64_t spoa_server_submit(void* srv, int type, int m, int n, int g, int e, int q, int c, uint32_t n_reads,
                            const uint64_t* read_off, const char* seqs, const char* quals, int msa) {
  try {
    spoa_amd::ServeJob job;
    job.params.type = static_cast<spoa::AlignmentType>(type);
    job.params.m = static_cast<std::int8_t>(m);
    job.params.n = static_cast<std::int8_t>(n);
    job.params.g = static_cast<std::int8_t>(g);
    job.params.e = static_cast<std::int8_t>(e);
    job.params.q = static_cast<std::int8_t>(q);
    job.params.c = static_cast<std::int8_t>(c);
    job.msa = msa != 0;
    job.reads.reserve(n_reads);
    for (uint32_t k = 0; k < n_reads; ++k) {
      job.reads.emplace_back(seqs + read_off[k], seqs + read_off[k + 1]);
      if (quals) job.quals.emplace_back(quals + read_off[k], quals + read_off[k + 1]);
    }
    return static_cast<spoa_amd::Server*>(srv)->Submit(std::move(job));
  } catch (const std::exception& ex) {
    Error(ex);
    return 0;
  }
}

void* spoa_server_wait(void* srv, uint64_t id, double timeout_s) {
  std::unique_ptr<spoa_amd::ServeResult> r(new spoa_amd::ServeResult());
  if (!static_cast<spoa_amd::Server*>(srv)->Wait(id, r.get(), timeout_s)) return nullptr;
  return r.release();
}

int spoa_server_stats(void* srv, spoa_serve_stats* out) {
  const spoa_amd::ServeStats s = static_cast<spoa_amd::Server*>(srv)->Stats();
  out->submitted = s.submitted;
  out->completed = s.completed;
  out->failed = s.failed;
  out->rounds = s.rounds;
  out->launches = s.launches;
  out->align_calls = s.align_calls;
  out->deferred = s.deferred;
  out->cells = s.cells;
  out->dp_bytes = s.dp_bytes;
  out->kernel_ms = s.kernel_ms;
  out->max_active_seen = s.max_active_seen;
  return 0;
}

double spoa_server_now_ms(void* srv) { return static_cast<spoa_amd::Server*>(srv)->NowMs(); }

int spoa_serve_result_info(void* res, int* status, double* submit_ms, double* admit_ms, double* done_ms,
                           uint32_t* rounds) {
  const auto& r = *static_cast<spoa_amd::ServeResult*>(res);
  if (status) *status = r.status;
  if (submit_ms) *submit_ms = r.submit_ms;
  if (admit_ms) *admit_ms = r.admit_ms;
  if (done_ms) *done_ms = r.done_ms;
  if (rounds) *rounds = r.rounds;
  return 0;
}

int spoa_serve_result_lane(void* res) { return static_cast<int>(static_cast<spoa_amd::ServeResult*>(res)->lane); }

const char* spoa_serve_result_error(void* res) { return static_cast<spoa_amd::ServeResult*>(res)->error.c_str(); }

int spoa_serve_result_consensus(void* res, char* out, uint32_t cap) {
  const std::string& s = static_cast<spoa_amd::ServeResult*>(res)->consensus;
  if (s.size() > cap) return -2;
  std::memcpy(out, s.data(), s.size());
  return static_cast<int>(s.size());
}

int spoa_serve_result_msa(void* res, char* out, uint64_t cap, uint32_t* rows, uint32_t* cols) {
  const auto& m = static_cast<spoa_amd::ServeResult*>(res)->msa;
  *rows = static_cast<uint32_t>(m.size());
  *cols = m.empty() ? 0 : static_cast<uint32_t>(m[0].size());
  const uint64_t need = static_cast<uint64_t>(*rows) * *cols;
  if (need > cap) return -2;
  for (std::size_t k = 0; k < m.size(); ++k) std::memcpy(out + k * (*cols), m[k].data(), *cols);
  return 0;
}

void spoa_serve_result_free(void* res) { delete static_cast<spoa_amd::ServeResult*>(res); }

// ------------------------------------------------------------- synth --

void* spoa_synth_new(const spoa_synth_spec* spec, uint64_t seed, uint32_t n_windows) {
  try {
    spoa_synth::WindowSpec sp;
    sp.template_min = spec->template_min;
    sp.template_max = spec->template_max;
    sp.reads_min = spec->reads_min;
    sp.reads_max = spec->reads_max;
    sp.read_len_min = spec->read_len_min;
    sp.read_len_max = spec->read_len_max;
    sp.err_min = spec->err_min;
    sp.err_max = spec->err_max;
    sp.revcomp_odd = spec->revcomp_odd;
    sp.log_uniform_len = spec->log_uniform_len;
    auto h = std::make_unique<SynthHandle>();
    h->win_read_off.push_back(0);
    h->read_off.push_back(0);
    h->templ_off.push_back(0);
    for (uint32_t w = 0; w < n_windows; ++w) {
      const spoa_synth::Window win = spoa_synth::MakeWindow(spoa_synth::WindowSeed(seed, w), sp);
      for (const auto& r : win.reads) {
        h->seqs += r;
        h->read_off.push_back(h->seqs.size());
      }
      h->win_read_off.push_back(static_cast<uint32_t>(h->read_off.size() - 1));
      h->templates += win.templ;
      h->templ_off.push_back(h->templates.size());
    }
    return h.release();
  } catch (const std::exception& ex) {
    Error(ex);
    return nullptr;
  }
}

void spoa_synth_free(void* h) { delete static_cast<SynthHandle*>(h); }

void spoa_synth_sizes(void* h, uint32_t* n_windows, uint64_t* n_reads, uint64_t* seq_bytes, uint64_t* templ_bytes) {
  const auto* s = static_cast<SynthHandle*>(h);
  *n_windows = static_cast<uint32_t>(s->win_read_off.size() - 1);
  *n_reads = s->read_off.size() - 1;
  *seq_bytes = s->seqs.size();
  *templ_bytes = s->templates.size();
}

void spoa_synth_copy(void* h, uint32_t* win_read_off, uint64_t* read_off, char* seqs, uint64_t* templ_off,
                     char* templates) {
  const auto* s = static_cast<SynthHandle*>(h);
  std::memcpy(win_read_off, s->win_read_off.data(), s->win_read_off.size() * sizeof(uint32_t));
  std::memcpy(read_off, s->read_off.data(), s->read_off.size() * sizeof(uint64_t));
  std::memcpy(seqs, s->seqs.data(), s->seqs.size());
  if (templ_off) std::memcpy(templ_off, s->templ_off.data(), s->templ_off.size() * sizeof(uint64_t));
  if (templates) std::memcpy(templates, s->templates.data(), s->templates.size());
}

}  // extern "C"
