// Internal host-side declarations of the MI355X alignment path.
#pragma once

#include <cstdint>
#include <vector>

#include "spoa/alignment_engine.hpp"
#include "spoa/graph.hpp"
#include "spoa_hip.h"

namespace spoa {

// The engine AlignmentEngine::Create returns: the reference's
// SimdAlignmentEngine slot (src/alignment_engine.cpp:68-72) filled by the
// HIP kernels behind include/spoa_hip.h.
//
// With sisd it stands in for the reference's scalar engine instead
// (SisdAlignmentEngine, src/sisd_alignment_engine.cpp, the engine of a
// reference built without SIMD): its results and messages.
class HipAlignmentEngine : public AlignmentEngine {
 public:
  HipAlignmentEngine(AlignmentType type, AlignmentSubtype subtype, std::int8_t m, std::int8_t n, std::int8_t g,
                     std::int8_t e, std::int8_t q, std::int8_t c, bool sisd = false);
  ~HipAlignmentEngine() override;

  void Prealloc(std::uint32_t max_sequence_len, std::uint8_t alphabet_size) override;

  using AlignmentEngine::Align;
  Alignment Align(const char* sequence, std::uint32_t sequence_len, const Graph& graph,
                  std::int32_t* score = nullptr) override;

  spoa_hip_engine* device() const { return device_; }
  bool sisd() const { return sisd_; }

  // score width (16 or 32) the range bound chose for the last Align, 0 if
  // it ran no DP
  int last_width() const { return last_width_; }

  // the rank-ordered CSR view of a graph, as the C ABI takes it
  static spoa_hip_graph View(const Graph& graph);

 private:
  spoa_hip_engine* device_ = nullptr;
  std::vector<std::int32_t> pairs_;
  int last_width_ = 0;
  bool sisd_ = false;
};

// AlignmentEngine::Create with the engine chosen explicitly: sisd selects
// the reference's scalar engine's results.  Create itself passes
// DefaultSisd().
std::unique_ptr<AlignmentEngine> CreateEngine(AlignmentType type, std::int8_t m, std::int8_t n, std::int8_t g,
                                              std::int8_t e, std::int8_t q, std::int8_t c, bool sisd);

// The engine AlignmentEngine::Create stands in for: the reference picks it
// when it is built (its SIMD engine when built with SSE4.1 / AVX2, else the
// scalar one); here SPOA_HIP_ENGINE=sisd selects the scalar engine's
// results at run time, anything else (or nothing) the SIMD engine's.
bool DefaultSisd();

// AlignmentEngine::Create's checks and normalisation
// (src/alignment_engine.cpp:32-66): throws std::invalid_argument with the
// reference's messages, normalises g/e/q/c in place, returns the subtype.
AlignmentSubtype CheckCreateParams(AlignmentType type, std::int8_t* g, std::int8_t* e, std::int8_t* q,
                                   std::int8_t* c);

// Device ordinal for engines created through AlignmentEngine::Create
// (SPOA_HIP_DEVICE, default 0).
int DefaultDevice();

}  // namespace spoa
