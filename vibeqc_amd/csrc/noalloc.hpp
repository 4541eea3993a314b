// Device workspace of the allocation-free J/K entry point.
#pragma once

#include "common.hpp"

namespace vqc {

// Size of the fixed per-device workspace (each device holds its own copy
// until the process ends).
constexpr std::size_t kNoAllocScratchBytes = std::size_t{128} << 20;

// Device address of `device`'s copy of the workspace (resolved once per device
// by hipGetSymbolAddress with that device current).
void* noalloc_scratch(int device);

}  // namespace vqc
