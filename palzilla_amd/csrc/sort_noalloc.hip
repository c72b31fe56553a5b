// palz_sort_u64_noalloc: the onesweep sort of palz_sort_u64 with its scratch
// (look-back table, digit histograms, tile tickets, epoch) in one file-scope
// device array instead of a sorter's allocations, for callers that cannot
// allocate: code under stream capture, tight serving loops, harnesses that
// forbid hipMalloc.  The kernels are palz_sort_u64's (sort.hip).
//
// Each device holds its own copy of the array until the process ends, and
// every call on a device shares it: calls that overlap on one device race.
// The histograms and tickets are cleared by the sort's first kernel on every
// call; the look-back table is epoch-tagged (lookback.hpp) and cleared only on
// the first call on a device and then once per 2^27 calls.
#include <mutex>

#include "lookback.hpp"
#include "sort.hpp"

namespace palz {

namespace {

constexpr uint32_t kNoAllocMaxN = 1u << 24;                        // keys per call
constexpr uint32_t kNoAllocTiles = kNoAllocMaxN / (uint32_t)kSortTile;  // 4096 look-back tiles
// Layout in 64-bit words, every part starting on a 256-byte boundary.
constexpr size_t kLookbackWords = (size_t)kNoAllocTiles * 256u;        // 8 MiB
constexpr size_t kHistWords = (size_t)kSortMaxPasses * 256u / 2u;      // u32 [8][256]
constexpr size_t kTicketWords = 32;                                    // u32 [8], padded
constexpr size_t kEpochWords = 32;                                     // u32 [1], padded
constexpr size_t kScratchWords = kLookbackWords + kHistWords + kTicketWords + kEpochWords;

static __device__ __align__(256) uint64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
struct DeviceScratch {
    uint64_t* ptr = nullptr;  // the device address of `scratch` on that device
    uint64_t uses = 0;        // calls since its look-back table was last cleared
};
DeviceScratch g_devices[kMaxDevices];
std::mutex g_mutex;

SortScratch view(uint64_t* base) {
    SortScratch s;
    s.lookback = base;
    s.hist = reinterpret_cast<uint32_t*>(base + kLookbackWords);
    s.tickets = reinterpret_cast<uint32_t*>(base + kLookbackWords + kHistWords);
    s.epoch = reinterpret_cast<uint32_t*>(base + kLookbackWords + kHistWords + kTicketWords);
    s.lookback_tiles = kNoAllocTiles;
    return s;
}

// The current device's scratch.  The first call on a device looks up the
// array's device address and clears its look-back table and epoch on `st`
// (then waits for the stream, so that the set-up is done before any other
// stream's call); later calls clear it again once per 2^27 calls.
int device_scratch(hipStream_t st, SortScratch* out) {
    int dev = 0;
    PALZ_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sort_u64_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceScratch& d = g_devices[dev];
    bool clear = false;
    if (!d.ptr) {
        void* p = nullptr;
        PALZ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        d.ptr = static_cast<uint64_t*>(p);
        clear = true;
    } else if (++d.uses >= kLbSortsPerClear / 2) {
        clear = true;
    }
    *out = view(d.ptr);
    if (clear) {
        launch_sort_clear(*out, st);
        PALZ_LAUNCH_CHECK();
        PALZ_HIP(hipStreamSynchronize(st));
        d.uses = 0;
    }
    return PALZ_OK;
}

}  // namespace

}  // namespace palz

extern "C" int palz_sort_u64_noalloc(uint64_t* words, uint64_t* tmp, uint32_t* values, uint32_t* tmp_values,
                                     uint32_t n, int start_bit, int end_bit, void* stream) {
    using namespace palz;
    if (start_bit < 0 || end_bit > 64 || start_bit >= end_bit)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sort_u64_noalloc: need 0 <= start_bit < end_bit <= 64");
    if (n > kNoAllocMaxN) return fail(PALZ_ERR_CAPACITY, "palz_sort_u64_noalloc: n must be <= 2^24");
    if (n == 0) return PALZ_OK;
    if (!words || !tmp) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sort_u64_noalloc: words/tmp is NULL");
    if ((values == nullptr) != (tmp_values == nullptr))
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_sort_u64_noalloc: values and tmp_values go together");
    hipStream_t st = (hipStream_t)stream;
    SortScratch s;
    const int rc = device_scratch(st, &s);
    if (rc != PALZ_OK) return rc;
    launch_sort_prepare(s, nullptr, st);
    launch_sort_count(s, words, n, start_bit, end_bit, st);
    const int npass = launch_sort_passes(s, words, tmp, values, tmp_values, n, start_bit, end_bit, st);
    PALZ_LAUNCH_CHECK();
    if (npass & 1) {
        PALZ_HIP(hipMemcpyAsync(words, tmp, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, st));
        if (values) PALZ_HIP(hipMemcpyAsync(values, tmp_values, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    }
    return PALZ_OK;
}
