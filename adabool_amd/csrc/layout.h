// layout.h -- constants of the resident index layout shared by host and device code.
#pragma once

#include <cstdint>

namespace ab {
// Records per key-stream block of the value-tiled layout; a wave64 decodes a
// block with 16 records per lane (record s*64 + lane in slot s).
constexpr uint32_t kBlockRecs = 1024;
constexpr uint32_t kSlots = kBlockRecs / 64;
// seg8 row layout: 8 u32 per key boundary.
//   [0..5] per-type record prefix (key_segments_), [6] all-type record prefix,
//   [7] value (nnz) prefix.
constexpr int kSegStride = 8;
}  // namespace ab
