// Node-level path cuts: the types shared by the host entry (pathcuts.hip) and
// the kernels, which are compiled once per chain length K in
// pathcuts_k1.hip .. pathcuts_k3.hip (the interval tests inline into every
// kernel; one translation unit per K lets the build compile them in parallel).
#pragma once

#include "interval.hpp"
#include "prims.hpp"

namespace palz {
namespace cuts {

constexpr int kCutsThreads = 128;
constexpr uint32_t kCutsStack = 192;                      // stack entries per thread
constexpr uint32_t kCutsMaxThreads = 2048u * kCutsThreads;  // resident walkers
constexpr uint32_t kCutsExpandTarget = 1u << 25;  // listed tuples before the walks (2^20: 1.13 s, 2^25: 0.36 s)
constexpr uint32_t kCutsListMax = 1u << 26;       // largest expansion level (16 B per entry)
constexpr uint32_t kCutsListCap = 1u << 24;       // receiver points per call
constexpr int kCutsMaxLevels = 64;
constexpr uint32_t kCutsOnePassCap = 1u << 26;   // records: larger capacities take the two-pass walk

struct CutsArgs {
    const palz_triangle* tris;
    const palz_triangle_attrib* attribs;
    const palz_internal_node* nodes;
    const palz_node_index* root;
    const palz_normal_bound* nb;
    const palz_first_diffuse* points;
    uint32_t n_points;
    uint32_t event;
    uint32_t scene;
    uint32_t n_threads;
    float eta_lo, eta_hi;
    float lx, ly, lz;
    uint32_t* stack;
    uint32_t* counts;  // counting pass: tuples per point; writing pass: their offsets
    palz_path* out;
    uint32_t capacity;
    uint32_t* status;
    unsigned long long* total;  // 64-bit sum of the counts; all ones once it does not fit 32 bits
    uint64_t* words;            // one-pass mode: (tuple << 32 | record position) per record
};

// A pending tuple of the traversal: receiver point and one node per slot.
struct alignas(16) CutEntry {
    uint32_t q;
    uint32_t n[3];
};

// Walk modes: kCount (per-tuple counts and their 64-bit sum), kWrite (after
// the scan: records at the tuples' offsets), kEmit (one pass: records in
// arrival order with their (tuple, position) words).
enum WalkMode { kCount = 0, kWrite = 1, kEmit = 2 };

// Launchers, instantiated for K = 1, 2, 3 in pathcuts_k<K>*.hip.  launch_walk
// expects *queue == 0.
template <int K>
void launch_expand(const CutsArgs& a, const CutEntry* in, uint32_t n_in, uint32_t* counts, CutEntry* out,
                   uint32_t* expandable, hipStream_t st, bool write);
template <int K, int kMode>
void launch_walk(const CutsArgs& a, const CutEntry* entries, uint32_t n_entries, uint32_t* queue, hipStream_t st);

}  // namespace cuts
}  // namespace palz
