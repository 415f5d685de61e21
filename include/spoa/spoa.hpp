// Umbrella header, as the reference's include/spoa/spoa.hpp:6-7.
#ifndef SPOA_AMD_SPOA_HPP_
#define SPOA_AMD_SPOA_HPP_

#include "spoa/alignment_engine.hpp"
#include "spoa/graph.hpp"

#endif  // SPOA_AMD_SPOA_HPP_
