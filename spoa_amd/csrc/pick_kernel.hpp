// The alignment kernel instance of a launch configuration: subtype, stored
// score type, columns per lane (K), workgroup size, the reference lane count
// (L) whose row width and convex gap segments the kernel reproduces, and
// whether the launch holds split jobs or reproduces the scalar engine.
//
// The instances of one (subtype, stored width, lane count) family are built
// in a translation unit of their own (poa_family.hip compiled once per
// family), so the families compile in parallel; each unit defines its
// family's PickFamily specialisation.
#pragma once

#include "engine.hpp"
#include "poa_kernels.hpp"

namespace spoa_hip_host {

template <int SUB, typename ST, int K, int L, bool SPLIT, bool SISD>
KernelFn PickMaxT(int waves) {
  if (waves == 1) return spoa_hip_dev::PoaAlignKernel<SUB, ST, K, 64, L, SPLIT, SISD>;
  if constexpr (K == 32) {
    return nullptr;  // 32 columns a lane only in single-wave workgroups
  } else {
    return spoa_hip_dev::PoaAlignKernel<SUB, ST, K, 64 * kMaxGroupWaves, L, SPLIT, SISD>;
  }
}

template <int SUB, typename ST, int L, bool SPLIT, bool SISD>
KernelFn PickK(int k, int waves) {
  if (k == 32) return PickMaxT<SUB, ST, 32, L, SPLIT, SISD>(waves);
  if (k == 16) return PickMaxT<SUB, ST, 16, L, SPLIT, SISD>(waves);
  if constexpr (SUB == 2 && L == 16) {
    return nullptr;  // convex lanes hold whole 16-column segments
  } else {
    return PickMaxT<SUB, ST, 8, L, SPLIT, SISD>(waves);
  }
}

template <int SUB, typename ST, int L, bool SISD>
KernelFn PickSplit(int k, int waves, bool split) {
  return split ? PickK<SUB, ST, L, true, SISD>(k, waves) : PickK<SUB, ST, L, false, SISD>(k, waves);
}

// One family of instances: FAM names (subtype, stored width, lane count) as
// below; nullptr for a configuration the family does not build.
enum KernelFamily {
  kFamLinear16 = 0,   // linear, int16 rows (16 lanes)
  kFamLinear32,       // linear, int32 rows (8 lanes)
  kFamAffine16,       // affine, int16 rows
  kFamAffine32,       // affine, int32 rows
  kFamConvex16,       // convex, int16 rows, 16- and 8-lane segments
  kFamConvex32,       // convex, int32 rows, 8-lane segments
  kFamConvexGotoh,    // convex, one-column segments (the scalar engine), both widths
};

template <int FAM>
KernelFn PickFamily(int bits, int lanes, int k, int waves, bool split, bool sisd);

template <>
KernelFn PickFamily<kFamLinear16>(int, int, int, int, bool, bool);
template <>
KernelFn PickFamily<kFamLinear32>(int, int, int, int, bool, bool);
template <>
KernelFn PickFamily<kFamAffine16>(int, int, int, int, bool, bool);
template <>
KernelFn PickFamily<kFamAffine32>(int, int, int, int, bool, bool);
template <>
KernelFn PickFamily<kFamConvex16>(int, int, int, int, bool, bool);
template <>
KernelFn PickFamily<kFamConvex32>(int, int, int, int, bool, bool);
template <>
KernelFn PickFamily<kFamConvexGotoh>(int, int, int, int, bool, bool);

}  // namespace spoa_hip_host
