// Device-wide primitives shared by the tuple and path stages.
#pragma once

#include "common.hpp"

namespace palz {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;
static_assert(kScanItems % 4 == 0, "the apply pass moves full tiles as uint4s");

// Exclusive scan of n u32 values (reduce-then-scan, three launches).  `n_dev`
// (optional) overrides n with a device-side count, in which case n is the
// capacity.  `total` (optional, device) receives the sum.  `block_sums` needs
// div_up(n, kScanTile) + 1 entries.  In-place (in == out) is allowed.
void launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, const uint32_t* n_dev,
                           uint32_t* total, uint32_t* block_sums, hipStream_t stream);

inline uint32_t scan_scratch_entries(uint32_t n) { return div_up(n, kScanTile) + 2u; }

}  // namespace palz
