// spoa::AlignmentEngine for the MI355X path.  Create() follows the
// reference factory (src/alignment_engine.cpp:15-73: validation, subtype
// choice, normalisation) and returns a HipAlignmentEngine, the analogue of
// the reference's CreateSimdAlignmentEngine hook (:68); the DP and the
// traceback run on the GPU through include/spoa_hip.h.  There is no CPU
// engine behind it: without a usable HIP device Create() throws.

#include "spoa/alignment_engine.hpp"

#include <algorithm>
#include <cstdlib>
#include <limits>
#include <stdexcept>
#include <string>

#include "spoa/graph.hpp"
#include "spoa_amd_host.hpp"
#include "spoa_hip.h"

namespace spoa {

namespace {

[[noreturn]] void Rethrow(int status) {
  throw std::invalid_argument(spoa_hip_last_error());
  (void)status;
}

}  // namespace

std::unique_ptr<AlignmentEngine> AlignmentEngine::Create(AlignmentType type, std::int8_t m, std::int8_t n,
                                                         std::int8_t g) {
  return Create(type, m, n, g, g);
}

std::unique_ptr<AlignmentEngine> AlignmentEngine::Create(AlignmentType type, std::int8_t m, std::int8_t n,
                                                         std::int8_t g, std::int8_t e) {
  return Create(type, m, n, g, e, g, e);
}

std::unique_ptr<AlignmentEngine> AlignmentEngine::Create(AlignmentType type, std::int8_t m, std::int8_t n,
                                                         std::int8_t g, std::int8_t e, std::int8_t q,
                                                         std::int8_t c) {
  return CreateEngine(type, m, n, g, e, q, c, DefaultSisd());
}

std::unique_ptr<AlignmentEngine> CreateEngine(AlignmentType type, std::int8_t m, std::int8_t n, std::int8_t g,
                                              std::int8_t e, std::int8_t q, std::int8_t c, bool sisd) {
  const AlignmentSubtype subtype = CheckCreateParams(type, &g, &e, &q, &c);
  return std::unique_ptr<AlignmentEngine>(new HipAlignmentEngine(type, subtype, m, n, g, e, q, c, sisd));
}

bool DefaultSisd() {
  const char* env = std::getenv("SPOA_HIP_ENGINE");
  return env && std::string(env) == "sisd";
}

AlignmentSubtype CheckCreateParams(AlignmentType type, std::int8_t* g_, std::int8_t* e_, std::int8_t* q_,
                                   std::int8_t* c_) {
  std::int8_t& g = *g_;
  std::int8_t& e = *e_;
  std::int8_t& q = *q_;
  std::int8_t& c = *c_;
  if (type != AlignmentType::kSW && type != AlignmentType::kNW && type != AlignmentType::kOV) {
    throw std::invalid_argument("[spoa::AlignmentEngine::Create] error: invalid alignment type!");
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
  // :57-66 -- linear when g >= e; affine when one of the pieces dominates
  AlignmentSubtype subtype = g >= e ? AlignmentSubtype::kLinear
                                    : ((g <= q || e >= c) ? AlignmentSubtype::kAffine : AlignmentSubtype::kConvex);
  if (subtype == AlignmentSubtype::kLinear) {
    e = g;
  } else if (subtype == AlignmentSubtype::kAffine) {
    q = g;
    c = e;
  }
  return subtype;
}

AlignmentEngine::AlignmentEngine(AlignmentType type, AlignmentSubtype subtype, std::int8_t m, std::int8_t n,
                                 std::int8_t g, std::int8_t e, std::int8_t q, std::int8_t c)
    : type_(type), subtype_(subtype), m_(m), n_(n), g_(g), e_(e), q_(q), c_(c) {}

Alignment AlignmentEngine::Align(const std::string& sequence, const Graph& graph, std::int32_t* score) {
  return Align(sequence.c_str(), static_cast<std::uint32_t>(sequence.size()), graph, score);
}

// src/alignment_engine.cpp:101-110
std::int64_t AlignmentEngine::WorstCaseAlignmentScore(std::int64_t i, std::int64_t j) const {
  auto gap = [&](std::int64_t len) -> std::int64_t {
    return len == 0 ? 0 : std::min<std::int64_t>(g_ + (len - 1) * e_, q_ + (len - 1) * c_);
  };
  return std::min<std::int64_t>(-1 * (m_ * std::min(i, j) + gap(std::llabs(i - j))), gap(i) + gap(j));
}

// ------------------------------------------------------------------------

int DefaultDevice() {
  const char* env = std::getenv("SPOA_HIP_DEVICE");
  return env ? std::atoi(env) : 0;
}

HipAlignmentEngine::HipAlignmentEngine(AlignmentType type, AlignmentSubtype subtype, std::int8_t m, std::int8_t n,
                                       std::int8_t g, std::int8_t e, std::int8_t q, std::int8_t c, bool sisd)
    : AlignmentEngine(type, subtype, m, n, g, e, q, c), sisd_(sisd) {
  spoa_hip_params p{};
  p.type = static_cast<std::int32_t>(type);
  p.subtype = static_cast<std::int32_t>(subtype);
  p.m = m; p.n = n; p.g = g; p.e = e; p.q = q; p.c = c;
  const int rc = sisd ? spoa_hip_engine_create_sisd(DefaultDevice(), &p, &device_)
                      : spoa_hip_engine_create(DefaultDevice(), &p, &device_);
  if (rc != SPOA_HIP_OK) Rethrow(rc);
}

HipAlignmentEngine::~HipAlignmentEngine() { spoa_hip_engine_destroy(device_); }

// :381-422
void HipAlignmentEngine::Prealloc(std::uint32_t max_sequence_len, std::uint8_t alphabet_size) {
  const int rc = spoa_hip_prealloc(device_, max_sequence_len, alphabet_size);
  if (rc != SPOA_HIP_OK) Rethrow(rc);
}

spoa_hip_graph HipAlignmentEngine::View(const Graph& graph) {
  const auto& v = graph.rank_view();
  spoa_hip_graph gv{};
  gv.num_nodes = static_cast<std::uint32_t>(v.node_id.size());
  gv.row_char = v.row_char.data();
  gv.node_id = v.node_id.data();
  gv.pred_off = v.pred_off.data();
  gv.pred_rank = v.pred_rank.data();
  gv.is_sink = v.is_sink.data();
  return gv;
}

// :703-780 (the scalar engine: sisd...:255-291)
Alignment HipAlignmentEngine::Align(const char* sequence, std::uint32_t sequence_len, const Graph& graph,
                                    std::int32_t* score) {
  const std::string who = sisd_ ? "[spoa::SisdAlignmentEngine::Align]" : "[spoa::SimdAlignmentEngine::Align]";
  if (sequence_len > static_cast<std::uint32_t>(std::numeric_limits<std::int32_t>::max())) {
    throw std::invalid_argument(who + " error: too large sequence!");
  }
  last_width_ = 0;
  if (graph.nodes().empty() || sequence_len == 0) return Alignment();
  {
    // the scalar engine bounds (n, V) and always scores in int32
    const std::int64_t w = WorstCaseAlignmentScore(static_cast<std::int64_t>(sequence_len) + (sisd_ ? 0 : 8),
                                                   static_cast<std::int64_t>(graph.nodes().size()));
    last_width_ = w < std::numeric_limits<std::int32_t>::min() + 1024
                      ? 0
                      : (sisd_ || w < std::numeric_limits<std::int16_t>::min() + 1024 ? 32 : 16);
    // :723-725, before anything is sized for the call
    if (last_width_ == 0) throw std::invalid_argument(who + " error: possible overflow!");
  }
  const spoa_hip_graph gv = View(graph);
  const std::size_t cap = static_cast<std::size_t>(gv.num_nodes) + sequence_len;
  pairs_.resize(2 * cap);
  std::uint32_t np = 0;
  std::int32_t s = 0, has = 0;
  const int rc = spoa_hip_align(device_, &gv, sequence, sequence_len, pairs_.data(), static_cast<std::uint32_t>(cap),
                                &np, &s, &has);
  if (rc != SPOA_HIP_OK) Rethrow(rc);
  if (has && score) *score = s;
  Alignment out(np);
  for (std::uint32_t k = 0; k < np; ++k) out[k] = {pairs_[2 * k], pairs_[2 * k + 1]};
  return out;
}

}  // namespace spoa
