// Kernel instances for engines that reproduce the reference's scalar engine
// (src/sisd_alignment_engine.cpp).  Its linear and affine rows are the SIMD
// engine's, so they run the same kernels; its convex E / Q open from the
// final H of every column (:751-768), which is the SIMD engine's segment
// rule with one-column segments: the convex kernel with L = 1.  Kept apart
// from poa_hip.hip so the two build in parallel.

#include "pick_kernel.hpp"

namespace spoa_hip_host {

KernelFn PickConvexFullGotoh(int bits, int k, int waves) {
  return bits == 16 ? PickK<2, int16_t, 1>(k, waves) : PickK<2, int32_t, 1>(k, waves);
}

}  // namespace spoa_hip_host
