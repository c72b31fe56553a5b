// Onesweep LSD radix sort for gfx950 (wave64).
//
// Semantics follow tinyhiponesweep::OnesweepSort::sort (tinyhiponesweep.h:
// 592-668): a stable LSD sort with 8-bit digits over [startBit, endBit) of
// 64-bit words or 32-bit keys, optional u32 payload.  The design is
// MI355X-native:
//   * one digit count pass reads the words once for every digit position
//     (LDS histograms, one global atomic per non-empty bin and block);
//   * each digit pass is ONE kernel: a 4096-key tile per 256-lane block, ranked
//     by a wave64 multisplit (8 64-bit ballots per 64 keys, leader-lane LDS
//     counters), a decoupled look-back across tiles in a table of 64-bit
//     (epoch, flag, count) words, LDS staging and a digit-run coalesced scatter;
//   * tiles are handed out by a ticket counter, so a block only ever waits on
//     tiles owned by blocks that started before it;
//   * look-back words carry an epoch that a device-side counter bumps at the
//     start of every sort, so the table is never cleared between sorts and the
//     sort can be captured in a hipGraph and replayed (the word layout is in
//     lookback.hpp; a sorter clears its table once per 2^27 uses, long before
//     a tag could repeat).
#include "sort.hpp"

#include <algorithm>

#include "lookback.hpp"

namespace palz {

namespace {

constexpr int kLookbackWindow = 8;  // predecessor tiles read per look-back step

template <class KeyT>
__device__ __forceinline__ uint32_t digit_of(KeyT w, int shift, uint32_t mask) {
    return (uint32_t)(w >> shift) & mask;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* s_tot) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_tot[wave] = x;
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < wave; ++w) add += s_tot[w];
    __syncthreads();
    return x - v + add;
}

__global__ void __launch_bounds__(256) k_sort_prepare(uint32_t* hist, uint32_t* tickets, uint32_t* epoch,
                                                      palz_aabb* aabb) {
    for (int i = threadIdx.x; i < kSortMaxPasses * 256; i += blockDim.x) hist[i] = 0u;
    if (threadIdx.x < kSortMaxPasses) tickets[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
        epoch[0] = epoch[0] + 1u;
        if (aabb) {
            aabb->lower = {kFltMax, kFltMax, kFltMax};
            aabb->upper = {-kFltMax, -kFltMax, -kFltMax};
        }
    }
}

// A fresh look-back table: every word zero (tag 0, which no sort uses) and
// the epoch back at 0.
__global__ void __launch_bounds__(256) k_sort_clear(uint64_t* __restrict__ lookback, size_t n_words,
                                                    uint32_t* __restrict__ epoch) {
    for (size_t i = blockIdx.x * 256u + threadIdx.x; i < n_words; i += (size_t)gridDim.x * 256u) lookback[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) epoch[0] = 0u;
}

template <class KeyT>
__global__ void __launch_bounds__(256) k_sort_count(const KeyT* __restrict__ words, uint32_t n,
                                                    int start_bit, int end_bit, int npass,
                                                    uint32_t* __restrict__ hist,
                                                    const uint32_t* __restrict__ n_dev) {
    __shared__ uint32_t s_h[kSortMaxPasses * 256];
    if (n_dev) n = min(n, *n_dev);
    for (int i = threadIdx.x; i < npass * 256; i += 256) s_h[i] = 0u;
    __syncthreads();
    const uint32_t stride = gridDim.x * 256u;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += stride) {
        const KeyT w = words[i];
        for (int p = 0; p < npass; ++p) {
            const int shift = start_bit + 8 * p;
            const int bits = min(8, end_bit - shift);
            atomicAdd(&s_h[p * 256 + digit_of(w, shift, (1u << bits) - 1u)], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < npass * 256; i += 256) {
        const uint32_t c = s_h[i];
        if (c) atomicAdd(&hist[i], c);
    }
}

template <class KeyT, bool kValues>
__global__ void __launch_bounds__(kSortThreads)
    k_sort_pass(const KeyT* __restrict__ src, KeyT* __restrict__ dst,
                const uint32_t* __restrict__ vsrc, uint32_t* __restrict__ vdst, uint32_t n,
                int shift, uint32_t mask, int pass, const uint32_t* __restrict__ hist,
                uint64_t* __restrict__ lookback, uint32_t* __restrict__ ticket,
                const uint32_t* __restrict__ epoch, const uint32_t* __restrict__ n_dev) {
    __shared__ uint32_t s_wave_hist[kSortWaves][256];
    __shared__ uint32_t s_count[256];
    __shared__ int32_t s_local_start[256];
    __shared__ int32_t s_scatter_base[256];
    __shared__ uint32_t s_tot[kSortWaves];
    __shared__ uint32_t s_tile;
    __shared__ KeyT s_keys[kSortTile];
    __shared__ uint32_t s_vals[kValues ? kSortTile : 1];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;

    if (tid == 0) s_tile = atomicAdd(ticket, 1u);
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) s_wave_hist[w][tid] = 0u;
    s_count[tid] = 0u;
    __syncthreads();

    const uint32_t tile = s_tile;
    if (n_dev) {
        n = min(n, *n_dev);  // blocks past the device-side count have nothing to do
        if (tile * (uint32_t)kSortTile >= n) return;
    }
    const uint32_t ep = lb_tag(epoch[0], (uint32_t)pass);
    const uint32_t tile_base = tile * (uint32_t)kSortTile;
    const uint32_t wave_base = tile_base + (uint32_t)wave * (64u * kSortItems);

    KeyT key[kSortItems];
    uint32_t val[kValues ? kSortItems : 1];
    uint32_t rank[kSortItems];

#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t idx = wave_base + (uint32_t)j * 64u + (uint32_t)lane;
        key[j] = idx < n ? src[idx] : (KeyT)0;
        if constexpr (kValues) val[j] = idx < n ? vsrc[idx] : 0u;
    }

    // The tile's digit counts first, so that they are published (and later
    // tiles can look back past this one) while the keys are being ranked.
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t idx = wave_base + (uint32_t)j * 64u + (uint32_t)lane;
        if (idx < n) atomicAdd(&s_count[digit_of(key[j], shift, mask)], 1u);
    }
    __syncthreads();
    const uint32_t d = (uint32_t)tid;  // one digit per thread from here
    const uint32_t count = s_count[d];
    uint64_t* my = lookback + (size_t)tile * 256u + d;
    __hip_atomic_store(my, lb_pack(ep, tile == 0 ? kLbFlagInclusive : kLbFlagAggregate, count), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);

    // Wave64 multisplit: rank each key among equal digits, in key order.
    const uint64_t lt_mask = __lanemask_lt();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t idx = wave_base + (uint32_t)j * 64u + (uint32_t)lane;
        const bool valid = idx < n;
        const uint32_t dj = digit_of(key[j], shift, mask);
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (dj >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            m &= bit ? bal : ~bal;
        }
        const int leader = valid ? (__ffsll((unsigned long long)m) - 1) : lane;
        uint32_t old = 0u;
        if (valid && lane == leader) {
            old = s_wave_hist[wave][dj];
            s_wave_hist[wave][dj] = old + (uint32_t)__popcll(m);
        }
        old = (uint32_t)__shfl((int)old, leader, 64);
        rank[j] = old + (uint32_t)__popcll(m & lt_mask);
    }
    __syncthreads();

    // Per digit: exclusive prefix over waves.
    {
        uint32_t acc = 0u;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            const uint32_t c = s_wave_hist[w][d];
            s_wave_hist[w][d] = acc;
            acc += c;
        }
    }

    // Decoupled look-back over previous tiles for this digit.
    uint32_t excl = 0u;
    if (tile != 0) {
        // Walk back kLookbackWindow tiles per step: their entries are loaded
        // together, then taken in order up to the nearest inclusive one (an
        // unpublished entry ends the step; tiles before the first count as an
        // inclusive zero).
        int p = (int)tile - 1;
        for (;;) {
            uint64_t e[kLookbackWindow];
#pragma unroll
            for (int w = 0; w < kLookbackWindow; ++w)
                e[w] = p - w >= 0 ? __hip_atomic_load(lookback + (size_t)(p - w) * 256u + d, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : lb_pack(ep, kLbFlagInclusive, 0u);
            int used = 0;
            bool done = false, stalled = false;
#pragma unroll
            for (int w = 0; w < kLookbackWindow; ++w) {
                if (done || stalled) continue;
                if (!lb_current(e[w], ep)) {
                    stalled = true;
                    continue;
                }
                excl += lb_count(e[w]);
                ++used;
                done = lb_flag(e[w]) == kLbFlagInclusive;
            }
            if (done) break;
            p -= used;
            if (stalled) __builtin_amdgcn_s_sleep(1);
        }
        __hip_atomic_store(my, lb_pack(ep, kLbFlagInclusive, excl + count), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }

    const uint32_t global_start = block_excl_scan256(hist[d], s_tot);
    const uint32_t local_start = block_excl_scan256(count, s_tot);
    s_local_start[d] = (int32_t)local_start;
    s_scatter_base[d] = (int32_t)(global_start + excl) - (int32_t)local_start;
    __syncthreads();

    // Stage the tile in LDS in digit order.
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t idx = wave_base + (uint32_t)j * 64u + (uint32_t)lane;
        if (idx < n) {
            const uint32_t dj = digit_of(key[j], shift, mask);
            const uint32_t pos = (uint32_t)s_local_start[dj] + s_wave_hist[wave][dj] + rank[j];
            s_keys[pos] = key[j];
            if constexpr (kValues) s_vals[pos] = val[j];
        }
    }
    __syncthreads();

    // Digit runs are contiguous in LDS and in the output: coalesced scatter.
    const uint32_t tile_n = min((uint32_t)kSortTile, n - tile_base);
    for (uint32_t i = (uint32_t)tid; i < tile_n; i += kSortThreads) {
        const KeyT k = s_keys[i];
        const uint32_t di = digit_of(k, shift, mask);
        const uint32_t g = (uint32_t)(s_scatter_base[di] + (int32_t)i);
        dst[g] = k;
        if constexpr (kValues) vdst[g] = s_vals[i];
    }
}

template <class KeyT>
void sort_count(const SortScratch& s, const KeyT* words, uint32_t n, int start_bit, int end_bit,
                hipStream_t stream, const uint32_t* n_dev) {
    const int npass = (int)div_up((uint32_t)(end_bit - start_bit), 8u);
    const uint32_t blocks = std::max(1u, std::min(div_up(n, 256u * 16u), 2048u));
    k_sort_count<KeyT><<<blocks, 256, 0, stream>>>(words, n, start_bit, end_bit, npass, s.hist, n_dev);
}

template <class KeyT>
int sort_passes(const SortScratch& s, KeyT* words, KeyT* tmp, uint32_t* values, uint32_t* tmp_values, uint32_t n,
                int start_bit, int end_bit, hipStream_t stream, hipEvent_t* pass_events, const uint32_t* n_dev) {
    const int npass = (int)div_up((uint32_t)(end_bit - start_bit), 8u);
    const uint32_t tiles = div_up(n, (uint32_t)kSortTile);
    KeyT* src = words;
    KeyT* dst = tmp;
    uint32_t* vsrc = values;
    uint32_t* vdst = tmp_values;
    for (int p = 0; p < npass; ++p) {
        const int shift = start_bit + 8 * p;
        const int bits = std::min(8, end_bit - shift);
        const uint32_t mask = (1u << bits) - 1u;
        if (values) {
            k_sort_pass<KeyT, true><<<tiles, kSortThreads, 0, stream>>>(
                src, dst, vsrc, vdst, n, shift, mask, p, s.hist + p * 256, s.lookback, s.tickets + p, s.epoch, n_dev);
        } else {
            k_sort_pass<KeyT, false><<<tiles, kSortThreads, 0, stream>>>(
                src, dst, nullptr, nullptr, n, shift, mask, p, s.hist + p * 256, s.lookback, s.tickets + p, s.epoch,
                n_dev);
        }
        if (pass_events) (void)hipEventRecord(pass_events[p], stream);
        std::swap(src, dst);
        std::swap(vsrc, vdst);
    }
    return npass;
}

}  // namespace

void launch_sort_prepare(const SortScratch& s, palz_aabb* aabb, hipStream_t stream) {
    k_sort_prepare<<<1, 256, 0, stream>>>(s.hist, s.tickets, s.epoch, aabb);
}

void launch_sort_clear(const SortScratch& s, hipStream_t stream) {
    const size_t words = (size_t)s.lookback_tiles * 256u;
    const uint32_t blocks = (uint32_t)std::max<size_t>(1, std::min<size_t>((words + 255u) / 256u, 4096));
    k_sort_clear<<<blocks, 256, 0, stream>>>(s.lookback, words, s.epoch);
}

void launch_sort_count(const SortScratch& s, const uint64_t* words, uint32_t n, int start_bit,
                       int end_bit, hipStream_t stream, const uint32_t* n_dev) {
    sort_count<uint64_t>(s, words, n, start_bit, end_bit, stream, n_dev);
}

int launch_sort_passes(const SortScratch& s, uint64_t* words, uint64_t* tmp, uint32_t* values,
                       uint32_t* tmp_values, uint32_t n, int start_bit, int end_bit,
                       hipStream_t stream, hipEvent_t* pass_events, const uint32_t* n_dev) {
    return sort_passes<uint64_t>(s, words, tmp, values, tmp_values, n, start_bit, end_bit, stream, pass_events,
                                 n_dev);
}

// A whole sort on the stream: prepare, count, passes, and the copy back when
// the number of passes is odd.  Sizes and bits are checked by the callers.
template <class KeyT>
int sort_all(const SortScratch& s, KeyT* words, KeyT* tmp, uint32_t* values, uint32_t* tmp_values, uint32_t n,
             int start_bit, int end_bit, hipStream_t st) {
    launch_sort_prepare(s, nullptr, st);
    sort_count<KeyT>(s, words, n, start_bit, end_bit, st, nullptr);
    const int npass = sort_passes<KeyT>(s, words, tmp, values, tmp_values, n, start_bit, end_bit, st, nullptr,
                                        nullptr);
    PALZ_LAUNCH_CHECK();
    if (npass & 1) {
        PALZ_HIP(hipMemcpyAsync(words, tmp, sizeof(KeyT) * n, hipMemcpyDeviceToDevice, st));
        if (values) PALZ_HIP(hipMemcpyAsync(values, tmp_values, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    }
    return PALZ_OK;
}

}  // namespace palz

int palz_sorter::reserve(uint32_t n) {
    using namespace palz;
    if (!scratch.hist) {
        PALZ_HIP(hipMalloc(&scratch.hist, sizeof(uint32_t) * kSortMaxPasses * 256));
        PALZ_HIP(hipMalloc(&scratch.tickets, sizeof(uint32_t) * kSortMaxPasses));
        PALZ_HIP(hipMalloc(&scratch.epoch, sizeof(uint32_t)));
        PALZ_HIP(hipMemset(scratch.epoch, 0, sizeof(uint32_t)));
    }
    // Each use runs at most two sorts; the table is cleared long before the
    // tags epoch * 8 + pass could come round again.
    if (++uses >= kLbSortsPerClear / 2 && scratch.lookback) {
        PALZ_HIP(hipDeviceSynchronize());
        launch_sort_clear(scratch, nullptr);
        PALZ_HIP(hipDeviceSynchronize());
        uses = 0;
    }
    const uint32_t tiles = std::max(1u, div_up(n, (uint32_t)kSortTile));
    if (tiles > scratch.lookback_tiles) {
        // Grow geometrically; fresh words are zero (epoch 0 never matches).
        const uint32_t want = std::max(tiles, scratch.lookback_tiles * 2u);
        uint64_t* table = nullptr;
        PALZ_HIP(hipMalloc(&table, sizeof(uint64_t) * 256u * want));
        PALZ_HIP(hipMemset(table, 0, sizeof(uint64_t) * 256u * want));
        PALZ_HIP(hipDeviceSynchronize());
        if (scratch.lookback) PALZ_HIP(hipFree(scratch.lookback));
        scratch.lookback = table;
        scratch.lookback_tiles = want;
    }
    return PALZ_OK;
}

palz_sorter::~palz_sorter() {
    if (scratch.hist) (void)hipFree(scratch.hist);
    if (scratch.tickets) (void)hipFree(scratch.tickets);
    if (scratch.epoch) (void)hipFree(scratch.epoch);
    if (scratch.lookback) (void)hipFree(scratch.lookback);
}

extern "C" {

int palz_sorter_create(palz_sorter** out) {
    if (!out) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sorter_create: out is NULL");
    palz_sorter* s = new palz_sorter();
    const int rc = s->reserve(1u << 20);
    if (rc != PALZ_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return PALZ_OK;
}

int palz_sorter_destroy(palz_sorter* s) {
    delete s;
    return PALZ_OK;
}

int palz_sorter_reserve(palz_sorter* s, uint32_t n) {
    if (!s) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sorter_reserve: sorter is NULL");
    if (n > palz::kSortMaxN) return palz::fail(PALZ_ERR_CAPACITY, "palz_sorter_reserve: n >= 2^30");
    return s->reserve(n);
}

int palz_sort_u64(palz_sorter* s, uint64_t* words, uint64_t* tmp, uint32_t* values,
                  uint32_t* tmp_values, uint32_t n, int start_bit, int end_bit, void* stream) {
    using namespace palz;
    if (!s) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sort_u64: sorter is NULL");
    if (start_bit < 0 || end_bit > 64 || start_bit >= end_bit)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sort_u64: need 0 <= start_bit < end_bit <= 64");
    if (n > kSortMaxN) return fail(PALZ_ERR_CAPACITY, "palz_sort_u64: n must be < 2^30");
    if (n == 0) return PALZ_OK;
    if (!words || !tmp) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sort_u64: words/tmp is NULL");
    if ((values == nullptr) != (tmp_values == nullptr))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sort_u64: values and tmp_values go together");
    const int rc = s->reserve(n);
    if (rc != PALZ_OK) return rc;
    return sort_all<uint64_t>(s->scratch, words, tmp, values, tmp_values, n, start_bit, end_bit,
                              (hipStream_t)stream);
}

int palz_sort_u32(palz_sorter* s, uint32_t* keys, uint32_t* tmp, uint32_t* values, uint32_t* tmp_values,
                  uint32_t n, int start_bit, int end_bit, void* stream) {
    using namespace palz;
    if (!s) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sort_u32: sorter is NULL");
    if (start_bit < 0 || end_bit > 32 || start_bit >= end_bit)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sort_u32: need 0 <= start_bit < end_bit <= 32");
    if (n > kSortMaxN) return fail(PALZ_ERR_CAPACITY, "palz_sort_u32: n must be < 2^30");
    if (n == 0) return PALZ_OK;
    if (!keys || !tmp) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sort_u32: keys/tmp is NULL");
    if ((values == nullptr) != (tmp_values == nullptr))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sort_u32: values and tmp_values go together");
    const int rc = s->reserve(n);
    if (rc != PALZ_OK) return rc;
    return sort_all<uint32_t>(s->scratch, keys, tmp, values, tmp_values, n, start_bit, end_bit,
                              (hipStream_t)stream);
}

}  // extern "C"
