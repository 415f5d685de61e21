// Splitting a batch of alignment jobs into passes (host only, no HIP).
//
// A pass is one staging block copied to the device and one persistent
// launch per launch group.  Its jobs' DP matrices must fit the engine's
// memory budget together, and the job records address the pass's graph
// blocks, reads and pair slots with 32-bit offsets, so the running totals
// of those must stay within 32 bits as well.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>

namespace spoa_amd {

struct PassItem {
  std::uint64_t mat_bytes;    // DP matrices
  std::uint64_t graph_words;  // 32-bit words of the graph block
  std::uint64_t seq_bytes;    // bytes of the padded read
  std::uint64_t pairs;        // pair slots (V + n)
};

// True when one job alone fits a pass (the budget and the 32-bit offsets).
inline bool PassItemFits(const PassItem& it, std::uint64_t budget) {
  constexpr std::uint64_t kMax32 = std::numeric_limits<std::uint32_t>::max();
  return it.mat_bytes <= budget && it.graph_words <= kMax32 && it.seq_bytes <= kMax32 && it.pairs <= kMax32;
}

// The end of the pass that starts at lo over items[0, n): every job from lo
// on that still fits beside the ones before it.  Every item must fit alone
// (PassItemFits); the pass then holds at least one job.
inline std::size_t PassEnd(const PassItem* items, std::size_t lo, std::size_t n, std::uint64_t budget) {
  constexpr std::uint64_t kMax32 = std::numeric_limits<std::uint32_t>::max();
  std::uint64_t bytes = 0, words = 0, seq = 0, pairs = 0;
  std::size_t hi = lo;
  while (hi < n) {
    const PassItem& it = items[hi];
    if (hi > lo && (bytes + it.mat_bytes > budget || words + it.graph_words > kMax32 ||
                    seq + it.seq_bytes > kMax32 || pairs + it.pairs > kMax32)) {
      break;
    }
    bytes += it.mat_bytes;
    words += it.graph_words;
    seq += it.seq_bytes;
    pairs += it.pairs;
    ++hi;
  }
  return hi;
}

}  // namespace spoa_amd
