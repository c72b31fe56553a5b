// Look-back words of the onesweep digit passes: a tag, a flag and a count in
// one 64-bit word, so that one atomic load sees all three.  Host and device
// code (the CPU tests compile this header with g++).
//
//   bits 63..32  tag = sort epoch * 8 + pass (the table is never cleared
//                between sorts: a word from an earlier sort or pass carries
//                another tag and reads as unpublished)
//   bits 31..30  flag (1 = the tile's own count, 2 = inclusive prefix)
//   bits 29..0   count (a sort holds fewer than 2^30 keys)
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define PALZ_LB_HD __host__ __device__ inline
#else
#define PALZ_LB_HD inline
#endif

namespace palz {

constexpr uint32_t kLbCountBits = 30;
constexpr uint32_t kLbCountMask = (1u << kLbCountBits) - 1u;
constexpr uint32_t kLbPassesPerSort = 8;
constexpr uint32_t kLbFlagAggregate = 1u;
constexpr uint32_t kLbFlagInclusive = 2u;
// Sorts a table serves between two clears: the tags epoch * 8 + pass of
// epochs 1 .. 2^28 stay below 2^31 + 8, short of the 32-bit wrap, so no tag
// repeats between clears.
constexpr uint64_t kLbSortsPerClear = 1ull << 28;

PALZ_LB_HD uint32_t lb_tag(uint32_t epoch, uint32_t pass) { return epoch * kLbPassesPerSort + pass; }
PALZ_LB_HD uint64_t lb_pack(uint32_t tag, uint32_t flag, uint32_t count) {
    return ((uint64_t)tag << 32) | ((uint64_t)(flag & 3u) << kLbCountBits) | (uint64_t)(count & kLbCountMask);
}
PALZ_LB_HD bool lb_current(uint64_t w, uint32_t tag) { return (uint32_t)(w >> 32) == tag; }
PALZ_LB_HD uint32_t lb_flag(uint64_t w) { return (uint32_t)(w >> kLbCountBits) & 3u; }
PALZ_LB_HD uint32_t lb_count(uint64_t w) { return (uint32_t)w & kLbCountMask; }

}  // namespace palz
