// spoa::AlignmentEngine -- drop-in for the reference interface
// (include/spoa/alignment_engine.hpp:14-96 of spoa v4.0.8): same enums,
// same factory overloads, same virtual Prealloc / Align, same error
// messages.  Create() returns an engine whose DP and traceback run on a
// gfx950 GPU through the C ABI of spoa_hip.h; there is no CPU fallback --
// Create() throws when no HIP device can be opened.
#ifndef SPOA_AMD_ALIGNMENT_ENGINE_HPP_
#define SPOA_AMD_ALIGNMENT_ENGINE_HPP_

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace spoa {

enum class AlignmentType {
  kSW,  // local (Smith-Waterman)
  kNW,  // global (Needleman-Wunsch)
  kOV   // semi-global (overlap)
};

enum class AlignmentSubtype {
  kLinear,  // g * i
  kAffine,  // g + (i - 1) * e
  kConvex   // min(g + (i - 1) * e, q + (i - 1) * c)
};

class Graph;

// (node id or -1, sequence position or -1) pairs in path order
using Alignment = std::vector<std::pair<std::int32_t, std::int32_t>>;

class AlignmentEngine {
 public:
  virtual ~AlignmentEngine() = default;

  static std::unique_ptr<AlignmentEngine> Create(
      AlignmentType type, std::int8_t m, std::int8_t n, std::int8_t g);

  static std::unique_ptr<AlignmentEngine> Create(
      AlignmentType type, std::int8_t m, std::int8_t n, std::int8_t g,
      std::int8_t e);

  static std::unique_ptr<AlignmentEngine> Create(
      AlignmentType type, std::int8_t m, std::int8_t n, std::int8_t g,
      std::int8_t e, std::int8_t q, std::int8_t c);

  virtual void Prealloc(std::uint32_t max_sequence_len,
                        std::uint8_t alphabet_size) = 0;

  Alignment Align(const std::string& sequence, const Graph& graph,
                  std::int32_t* score = nullptr);

  virtual Alignment Align(const char* sequence, std::uint32_t sequence_len,
                          const Graph& graph,
                          std::int32_t* score = nullptr) = 0;

  AlignmentType type() const { return type_; }
  AlignmentSubtype subtype() const { return subtype_; }

 protected:
  AlignmentEngine(AlignmentType type, AlignmentSubtype subtype, std::int8_t m,
                  std::int8_t n, std::int8_t g, std::int8_t e, std::int8_t q,
                  std::int8_t c);

  std::int64_t WorstCaseAlignmentScore(std::int64_t sequence_len,
                                       std::int64_t graph_len) const;

  AlignmentType type_;
  AlignmentSubtype subtype_;
  std::int8_t m_;
  std::int8_t n_;
  std::int8_t g_;
  std::int8_t e_;
  std::int8_t q_;
  std::int8_t c_;
};

}  // namespace spoa

#endif  // SPOA_AMD_ALIGNMENT_ENGINE_HPP_
