// Node-level path cuts: the host entry palz_path_cuts_enumerate and its small
// kernels (seeding, counters, the gather of the one-pass walk); the traversal
// kernels are in pathcuts_kernels.hpp, compiled per chain length.
#include <algorithm>

#include "pathcuts.hpp"

namespace palz {

namespace {

using namespace cuts;

// receiver points without a diffuse hit are skipped by the kernels; the seed
// lists every point with its slots at the root
__global__ void __launch_bounds__(256) k_seed(CutEntry* __restrict__ out, uint32_t n_points,
                                              const palz_node_index* __restrict__ root) {
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q >= n_points) return;
    const palz_node_index r = *root;
    out[q] = {q, {r, r, r}};
}

__global__ void k_clear_total(unsigned long long* total) { *total = 0ull; }

// a total past 32 bits: report it and skip the writing pass
__global__ void k_check_total(unsigned long long* total, uint32_t* count, uint32_t* status) {
    if (*total > 0xFFFFFFFFull) {
        *count = 0xFFFFFFFFu;
        *total = ~0ull;
        if (status) atomicMax(status, (uint32_t)PALZ_ERR_CAPACITY);
    }
}

__global__ void k_write_zero(uint32_t* p) { *p = 0u; }

// records in sorted-word order: out[i] = scratch[position of word i]
__global__ void __launch_bounds__(256) k_gather_records(const uint64_t* __restrict__ words,
                                                        const palz_path* __restrict__ scratch, uint32_t n,
                                                        palz_path* __restrict__ out) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        out[i] = scratch[(uint32_t)words[i]];
}

__global__ void k_write_count(uint32_t* count, const unsigned long long* total) { *count = (uint32_t)*total; }

}  // namespace

}  // namespace palz

struct palz_path_cuts {
    uint32_t cap_list[2] = {0, 0};
    palz::cuts::CutEntry* list[2] = {nullptr, nullptr};  // pending tuples, ping-pong
    uint32_t cap_counts = 0;
    uint32_t* counts = nullptr;
    uint32_t* scan_sums = nullptr;
    size_t cap_stack = 0;  // words
    uint32_t* stack = nullptr;
    unsigned long long* total = nullptr;
    uint32_t* scalars = nullptr;  // [0] expandable flag, [1] level total, [2] walk queue
    // one-pass mode: unsorted records, their words and the sort's tmp
    uint32_t cap_emit = 0;
    palz_path* emit_records = nullptr;
    uint64_t* emit_words = nullptr;
    uint64_t* emit_tmp = nullptr;
    palz_sorter* sorter = nullptr;

    int ensure_emit(uint32_t n) {
        if (!sorter && palz_sorter_create(&sorter) != PALZ_OK) return PALZ_ERR_OUT_OF_MEMORY;
        if (n <= cap_emit) return PALZ_OK;
        for (void* p : {(void*)emit_records, (void*)emit_words, (void*)emit_tmp})
            if (p) (void)hipFree(p);
        emit_records = nullptr;
        emit_words = emit_tmp = nullptr;
        cap_emit = 0;
        PALZ_HIP(hipMalloc(&emit_records, sizeof(palz_path) * n));
        PALZ_HIP(hipMalloc(&emit_words, sizeof(uint64_t) * n));
        PALZ_HIP(hipMalloc(&emit_tmp, sizeof(uint64_t) * n));
        cap_emit = n;
        return PALZ_OK;
    }

    // Buffers grow to the next power of two of what a level needs (contents
    // are not kept: a list is grown only before it is written).
    static uint32_t grown(uint32_t n) {
        uint32_t c = 1u << 16;
        while (c < n) c <<= 1;
        return c;
    }
    int ensure_list(int i, uint32_t n_entries) {
        if (n_entries > cap_list[i]) {
            if (list[i]) (void)hipFree(list[i]);
            list[i] = nullptr;
            cap_list[i] = 0;
            const uint32_t c = grown(n_entries);
            PALZ_HIP(hipMalloc(&list[i], sizeof(palz::cuts::CutEntry) * c));
            cap_list[i] = c;
        }
        return PALZ_OK;
    }
    int ensure_counts(uint32_t n_entries) {
        if (n_entries > cap_counts) {
            for (void* p : {(void*)counts, (void*)scan_sums})
                if (p) (void)hipFree(p);
            counts = scan_sums = nullptr;
            cap_counts = 0;
            const uint32_t c = grown(n_entries);
            PALZ_HIP(hipMalloc(&counts, sizeof(uint32_t) * c));
            PALZ_HIP(hipMalloc(&scan_sums, sizeof(uint32_t) * palz::scan_scratch_entries(c)));
            cap_counts = c;
        }
        if (!total) PALZ_HIP(hipMalloc(&total, sizeof(unsigned long long)));
        if (!scalars) PALZ_HIP(hipMalloc(&scalars, sizeof(uint32_t) * 4));
        return PALZ_OK;
    }
    int ensure_stack(size_t stack_words) {
        if (stack_words > cap_stack) {
            if (stack) (void)hipFree(stack);
            stack = nullptr;
            cap_stack = 0;
            PALZ_HIP(hipMalloc(&stack, sizeof(uint32_t) * stack_words));
            cap_stack = stack_words;
        }
        return PALZ_OK;
    }
    ~palz_path_cuts() {
        for (void* p : {(void*)list[0], (void*)list[1], (void*)counts, (void*)scan_sums, (void*)stack, (void*)total,
                        (void*)scalars, (void*)emit_records, (void*)emit_words, (void*)emit_tmp})
            if (p) (void)hipFree(p);
        if (sorter) palz_sorter_destroy(sorter);
    }
};

extern "C" {

int palz_path_cuts_create(palz_path_cuts** out) {
    if (!out) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_path_cuts_create: NULL");
    *out = new palz_path_cuts();
    return PALZ_OK;
}

int palz_path_cuts_destroy(palz_path_cuts* pc) {
    delete pc;
    return PALZ_OK;
}

int palz_path_cuts_enumerate(palz_path_cuts* pc, const palz_scene_desc* scene, const palz_normal_bound* d_normal_bounds,
                             uint32_t event, float eta_lo, float eta_hi, uint32_t scene_index, palz_path* d_paths,
                             uint32_t capacity, uint32_t* d_count, uint32_t* d_status, void* stream) {
    using namespace palz;
    using namespace palz::cuts;
    if (!pc || !scene || !d_count) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_path_cuts_enumerate: NULL argument");
    const int K = ev_size(event);
    if (K < 1 || K > 3) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_path_cuts_enumerate: chain length must be 1..3");
    if (capacity && !d_paths) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_path_cuts_enumerate: NULL d_paths");
    if (!(eta_lo <= eta_hi)) return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_path_cuts_enumerate: eta_lo > eta_hi");
    hipStream_t st = (hipStream_t)stream;
    if (scene->n_points == 0) {
        k_write_zero<<<1, 1, 0, st>>>(d_count);
        PALZ_LAUNCH_CHECK();
        return PALZ_OK;
    }
    if (!scene->tris || !scene->attribs || !scene->nodes || !scene->root || !scene->points || !d_normal_bounds)
        return fail(PALZ_ERR_INVALID_ARGUMENT, "palz_path_cuts_enumerate: incomplete scene");
    if (scene->n_points > kCutsListCap)
        return fail(PALZ_ERR_CAPACITY, "palz_path_cuts_enumerate: more than 2^24 receiver points");
    // the seed list, and the counts with the call's scalars and total (whose
    // addresses go into the kernel arguments below)
    int rc = pc->ensure_list(0, scene->n_points);
    if (rc == PALZ_OK) rc = pc->ensure_counts(scene->n_points);
    if (rc != PALZ_OK) return rc;
    CutsArgs a;
    a.tris = scene->tris;
    a.attribs = scene->attribs;
    a.nodes = scene->nodes;
    a.root = scene->root;
    a.nb = d_normal_bounds;
    a.points = scene->points;
    a.n_points = scene->n_points;
    a.event = event;
    a.scene = scene_index;
    a.eta_lo = eta_lo;
    a.eta_hi = eta_hi;
    a.lx = scene->p_light.x;
    a.ly = scene->p_light.y;
    a.lz = scene->p_light.z;
    a.stack = nullptr;  // set once the walkers are counted
    a.counts = nullptr;  // set per level, as the counts buffer grows
    a.out = d_paths;
    a.capacity = capacity;
    a.status = d_status;
    a.total = pc->total;
    a.words = nullptr;

    // Breadth-first expansion until the list holds enough tuples that the
    // walks below are short: one host read of the level's total per level.
    // Every lane of an expansion tests one tuple's children, so the lanes stay
    // busy; the walks' lanes diverge (each walks a subtree of its own).
    int cur = 0;
    uint32_t n = scene->n_points;
    k_seed<<<div_up(n, 256u), 256, 0, st>>>(pc->list[0], n, scene->root);
    PALZ_LAUNCH_CHECK();
    for (int level = 0; level < kCutsMaxLevels && n > 0 && n < kCutsExpandTarget; ++level) {
        if ((rc = pc->ensure_counts(n)) != PALZ_OK) return rc;
        a.counts = pc->counts;
        k_write_zero<<<1, 1, 0, st>>>(pc->scalars);
        if (K == 1) launch_expand<1>(a, pc->list[cur], n, pc->counts, nullptr, pc->scalars, st, false);
        if (K == 2) launch_expand<2>(a, pc->list[cur], n, pc->counts, nullptr, pc->scalars, st, false);
        if (K == 3) launch_expand<3>(a, pc->list[cur], n, pc->counts, nullptr, pc->scalars, st, false);
        PALZ_LAUNCH_CHECK();
        launch_exclusive_scan(pc->counts, pc->counts, n, nullptr, pc->scalars + 1, pc->scan_sums, st);
        PALZ_LAUNCH_CHECK();
        uint32_t h[2];
        PALZ_HIP(hipMemcpyAsync(h, pc->scalars, sizeof(h), hipMemcpyDeviceToHost, st));
        PALZ_HIP(hipStreamSynchronize(st));
        if (!h[0] || h[1] > kCutsListMax) break;  // nothing left to cut, or the list would not fit
        if ((rc = pc->ensure_list(cur ^ 1, h[1])) != PALZ_OK) return rc;
        if (K == 1) launch_expand<1>(a, pc->list[cur], n, pc->counts, pc->list[cur ^ 1], pc->scalars, st, true);
        if (K == 2) launch_expand<2>(a, pc->list[cur], n, pc->counts, pc->list[cur ^ 1], pc->scalars, st, true);
        if (K == 3) launch_expand<3>(a, pc->list[cur], n, pc->counts, pc->list[cur ^ 1], pc->scalars, st, true);
        PALZ_LAUNCH_CHECK();
        cur ^= 1;
        n = h[1];
    }
    if (n == 0) {
        k_write_zero<<<1, 1, 0, st>>>(d_count);
        PALZ_LAUNCH_CHECK();
        return PALZ_OK;
    }

    // walk every listed tuple's subtree: count, scan, write
    if ((rc = pc->ensure_counts(n)) != PALZ_OK) return rc;
    a.counts = pc->counts;
    a.n_threads = std::min(div_up(n, (uint32_t)kCutsThreads) * kCutsThreads, kCutsMaxThreads);
    rc = pc->ensure_stack((size_t)kCutsStack * K * a.n_threads);
    if (rc != PALZ_OK) return rc;
    a.stack = pc->stack;
    auto walk = [&](int mode) {
        k_write_zero<<<1, 1, 0, st>>>(pc->scalars + 2);  // the walkers' queue
        uint32_t* q = pc->scalars + 2;
        const CutEntry* l = pc->list[cur];
        if (K == 1 && mode == kCount) launch_walk<1, kCount>(a, l, n, q, st);
        if (K == 1 && mode == kWrite) launch_walk<1, kWrite>(a, l, n, q, st);
        if (K == 1 && mode == kEmit) launch_walk<1, kEmit>(a, l, n, q, st);
        if (K == 2 && mode == kCount) launch_walk<2, kCount>(a, l, n, q, st);
        if (K == 2 && mode == kWrite) launch_walk<2, kWrite>(a, l, n, q, st);
        if (K == 2 && mode == kEmit) launch_walk<2, kEmit>(a, l, n, q, st);
        if (K == 3 && mode == kCount) launch_walk<3, kCount>(a, l, n, q, st);
        if (K == 3 && mode == kWrite) launch_walk<3, kWrite>(a, l, n, q, st);
        if (K == 3 && mode == kEmit) launch_walk<3, kEmit>(a, l, n, q, st);
    };
    if (capacity && capacity <= kCutsOnePassCap) {
        // one walk into scratch, a stable sort of its words by listed tuple and a
        // gather; a total past `capacity` falls back to the two-pass walk below
        rc = pc->ensure_emit(capacity);
        if (rc != PALZ_OK) return rc;
        a.out = pc->emit_records;
        a.words = pc->emit_words;
        k_clear_total<<<1, 1, 0, st>>>(pc->total);
        walk(kEmit);
        PALZ_LAUNCH_CHECK();
        unsigned long long total = 0;
        PALZ_HIP(hipMemcpyAsync(&total, pc->total, sizeof(total), hipMemcpyDeviceToHost, st));
        PALZ_HIP(hipStreamSynchronize(st));
        a.out = d_paths;
        a.words = nullptr;
        if (total <= capacity) {
            const uint32_t m = (uint32_t)total;
            int bits = 1;
            while (bits < 32 && (n - 1u) >> bits) ++bits;
            if (m > 1) {
                rc = palz_sort_u64(pc->sorter, pc->emit_words, pc->emit_tmp, nullptr, nullptr, m, 32, 32 + bits, st);
                if (rc != PALZ_OK) return rc;
            }
            if (m) {
                k_gather_records<<<std::min(div_up(m, 256u), 65536u), 256, 0, st>>>(pc->emit_words, pc->emit_records,
                                                                                 m, d_paths);
                PALZ_LAUNCH_CHECK();
            }
            k_write_count<<<1, 1, 0, st>>>(d_count, pc->total);
            PALZ_LAUNCH_CHECK();
            return PALZ_OK;
        }
    }
    k_clear_total<<<1, 1, 0, st>>>(pc->total);
    walk(kCount);
    PALZ_LAUNCH_CHECK();
    launch_exclusive_scan(pc->counts, pc->counts, n, nullptr, d_count, pc->scan_sums, st);
    k_check_total<<<1, 1, 0, st>>>(pc->total, d_count, d_status);
    PALZ_LAUNCH_CHECK();
    if (capacity) {
        walk(kWrite);
        PALZ_LAUNCH_CHECK();
    }
    return PALZ_OK;
}

}  // extern "C"
