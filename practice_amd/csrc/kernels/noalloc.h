// Static device scratch of the allocation-free execute (noalloc.hip).
#pragma once

#include <cstddef>

// bytes of the file-scope device array (per GPU)
#define CQ_NOALLOC_SCRATCH_BYTES (32u << 20)

namespace db {
// Device address of this GPU's copy of the scratch array, looked up once per
// device with that device current (hipGetSymbolAddress); throws on failure.
void* noAllocScratch(int device);
inline constexpr size_t noAllocScratchBytes() { return CQ_NOALLOC_SCRATCH_BYTES; }
}  // namespace db
