// The alignment kernel instance of a launch configuration: subtype, stored
// score type, columns per lane (K), workgroup size and the reference lane
// count (L) whose row width and convex gap segments the kernel reproduces.
#pragma once

#include "engine.hpp"
#include "poa_kernels.hpp"

namespace spoa_hip_host {

template <int SUB, typename ST, int K, int L>
KernelFn PickMaxT(int waves) {
  if (waves == 1) return spoa_hip_dev::PoaAlignKernel<SUB, ST, K, 64, L>;
  if constexpr (K == 32) {
    return nullptr;  // 32 columns a lane only in single-wave workgroups
  } else {
    return spoa_hip_dev::PoaAlignKernel<SUB, ST, K, 64 * kMaxGroupWaves, L>;
  }
}

template <int SUB, typename ST, int L>
KernelFn PickK(int k, int waves) {
  if (k == 32) return PickMaxT<SUB, ST, 32, L>(waves);
  if (k == 16) return PickMaxT<SUB, ST, 16, L>(waves);
  if constexpr (SUB == 2 && L == 16) {
    return nullptr;  // convex lanes hold whole 16-column segments
  } else {
    return PickMaxT<SUB, ST, 8, L>(waves);
  }
}

// Convex kernels whose E / Q open from the final H of every column (the
// scalar engine's full Gotoh rows, one-column segments); built in their
// own translation unit (poa_sisd.hip).
KernelFn PickConvexFullGotoh(int bits, int k, int waves);

}  // namespace spoa_hip_host
