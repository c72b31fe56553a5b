// Onesweep radix sort for gfx950 (wave64): internal interface shared by the
// sort entry points and the LBVH builder, which fuses the digit count into its
// Morton kernel.
#pragma once

#include "common.hpp"

namespace palz {

constexpr int kSortThreads = 256;                  // 4 waves of 64
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 16;                     // keys per lane
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096 keys per tile
constexpr int kSortMaxPasses = 8;                  // 64-bit words, 8-bit digits
constexpr uint32_t kSortMaxN = (1u << 30) - 1;     // look-back counts are 30-bit (lookback.hpp)

struct SortScratch {
    uint32_t* hist = nullptr;      // [kSortMaxPasses][256] digit histograms
    uint32_t* tickets = nullptr;   // [kSortMaxPasses] dynamic tile tickets
    uint32_t* epoch = nullptr;     // [1] bumped once per sort
    uint64_t* lookback = nullptr;  // [tiles * 256] (epoch, flag, count) words
    uint32_t lookback_tiles = 0;
};

// Clear a look-back table: every word zero and the epoch 0 (lookback.hpp).
void launch_sort_clear(const SortScratch& s, hipStream_t stream);

// Zero the histograms and tile tickets, bump the epoch.  If `aabb` is not null
// it is reset to the empty box (the LBVH build uses the same launch).
void launch_sort_prepare(const SortScratch& s, palz_aabb* aabb, hipStream_t stream);

// Count the digits of every pass in one read of the words.
// `n_dev` (optional): device-side count; n is then the capacity.
void launch_sort_count(const SortScratch& s, const uint64_t* words, uint32_t n, int start_bit,
                       int end_bit, hipStream_t stream, const uint32_t* n_dev = nullptr);

// Run the digit passes after the histograms are complete.  Returns the number
// of passes; when it is odd the sorted words are in `tmp`.
// `pass_events`, if not null, receives one event record after each pass.
int launch_sort_passes(const SortScratch& s, uint64_t* words, uint64_t* tmp, uint32_t* values,
                       uint32_t* tmp_values, uint32_t n, int start_bit, int end_bit,
                       hipStream_t stream, hipEvent_t* pass_events = nullptr,
                       const uint32_t* n_dev = nullptr);

}  // namespace palz

struct palz_sorter {
    palz::SortScratch scratch;
    uint64_t uses = 0;        // reserve() calls since the look-back table was last cleared
    int reserve(uint32_t n);  // grows the look-back table; allocates, so not graph-safe
    ~palz_sorter();
};
