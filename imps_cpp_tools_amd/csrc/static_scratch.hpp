// Fixed per-device scratch of the allocation-free H_AC entry point (static_scratch.hip).
#pragma once

#include <cstddef>

namespace imps {

constexpr size_t kStaticScratchBytes = size_t(64) << 20;  // 64 MiB per GPU, held until the process ends

// Device address of the scratch array on `device` (hipGetSymbolAddress on first use per device).
void* static_scratch_address(int device);

}  // namespace imps
