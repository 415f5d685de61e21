// One family of alignment-kernel instances (KernelFamily in
// pick_kernel.hpp), chosen by SPOA_KERNEL_FAMILY when the unit is compiled:
// the Makefile builds this file once per family, so the families' many
// instances compile in parallel.
//
// Linear and affine rows depend on the lane count only through the row
// width each job carries, so each of their widths has one family, with and
// without the scalar engine's behaviour.  Convex rows reproduce the SIMD
// engine's 16- or 8-lane segments (never the scalar engine), or the scalar
// engine's one-column segments (src/sisd_alignment_engine.cpp:751-768:
// E / Q open from the final H of every column).

#include "pick_kernel.hpp"

#ifndef SPOA_KERNEL_FAMILY
#error "compile with -DSPOA_KERNEL_FAMILY=<family index>"
#endif

namespace spoa_hip_host {

namespace {

// a template, so that only the compiled family's branch is instantiated
template <int F>
KernelFn FamilyInstances(int bits, int lanes, int k, int waves, bool split, bool sisd) {
  (void)bits;
  (void)lanes;
  if constexpr (F == kFamLinear16) {
    return sisd ? PickSplit<0, int16_t, 16, true>(k, waves, split) : PickSplit<0, int16_t, 16, false>(k, waves, split);
  } else if constexpr (F == kFamLinear32) {
    return sisd ? PickSplit<0, int32_t, 8, true>(k, waves, split) : PickSplit<0, int32_t, 8, false>(k, waves, split);
  } else if constexpr (F == kFamAffine16) {
    return sisd ? PickSplit<1, int16_t, 16, true>(k, waves, split) : PickSplit<1, int16_t, 16, false>(k, waves, split);
  } else if constexpr (F == kFamAffine32) {
    return sisd ? PickSplit<1, int32_t, 8, true>(k, waves, split) : PickSplit<1, int32_t, 8, false>(k, waves, split);
  } else if constexpr (F == kFamConvex16) {
    if (sisd) return nullptr;
    return lanes == 8 ? PickSplit<2, int16_t, 8, false>(k, waves, split)
                      : PickSplit<2, int16_t, 16, false>(k, waves, split);
  } else if constexpr (F == kFamConvex32) {
    if (sisd) return nullptr;
    return PickSplit<2, int32_t, 8, false>(k, waves, split);
  } else {
    static_assert(F == kFamConvexGotoh, "unknown kernel family");
    if (!sisd) return nullptr;
    return bits == 16 ? PickSplit<2, int16_t, 1, true>(k, waves, split)
                      : PickSplit<2, int32_t, 1, true>(k, waves, split);
  }
}

}  // namespace

template <>
KernelFn PickFamily<SPOA_KERNEL_FAMILY>(int bits, int lanes, int k, int waves, bool split, bool sisd) {
  return FamilyInstances<SPOA_KERNEL_FAMILY>(bits, lanes, k, waves, split, sisd);
}

}  // namespace spoa_hip_host
