// Reduce-then-scan exclusive prefix sum over u32 (256 lanes x 16 items per
// tile; the per-tile sums are scanned by one block).
#include "prims.hpp"

namespace palz {

namespace {

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total_out.
__device__ __forceinline__ uint32_t block_excl(uint32_t v, uint32_t* s_w, uint32_t* total_out) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    uint32_t add = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
        const uint32_t s = s_w[w];
        if (w < wave) add += s;
        tot += s;
    }
    __syncthreads();
    *total_out = tot;
    return inc - v + add;
}

__device__ __forceinline__ uint32_t count_of(uint32_t n, const uint32_t* n_dev) { return n_dev ? min(n, *n_dev) : n; }

__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n,
                                                              const uint32_t* __restrict__ n_dev,
                                                              uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_w[kScanThreads / 64];
    const uint32_t m = count_of(n, n_dev);
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint32_t i = base + j * kScanThreads + threadIdx.x;
        if (i < m) acc += in[i];
    }
    uint32_t tot;
    (void)block_excl(acc, s_w, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_scan_sums(uint32_t* __restrict__ sums, uint32_t nblocks,
                                                    uint32_t* __restrict__ total) {
    __shared__ uint32_t s_w[16];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nblocks; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nblocks ? sums[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl(v, s_w, &tot);
        if (i < nblocks) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const uint32_t* in, uint32_t* out, uint32_t n,
                                                             const uint32_t* __restrict__ n_dev,
                                                             const uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_w[kScanThreads / 64];
    const uint32_t m = count_of(n, n_dev);
    const uint32_t base = blockIdx.x * kScanTile;
    if (base >= m) return;
    // each thread owns kScanItems consecutive items; a whole tile of 16-byte
    // aligned buffers moves as uint4s (a wave's loads then cover whole lines
    // instead of one word per line and instruction)
    uint32_t v[kScanItems];
    const uint32_t t0 = base + threadIdx.x * kScanItems;
    const bool vec = base + kScanTile <= m && ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0u;
    uint32_t acc = 0;
    if (vec) {
        const uint4* in4 = reinterpret_cast<const uint4*>(in + t0);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            const uint4 x = in4[q];
            v[4 * q + 0] = x.x;
            v[4 * q + 1] = x.y;
            v[4 * q + 2] = x.z;
            v[4 * q + 3] = x.w;
        }
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) acc += v[j];
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const uint32_t i = t0 + j;
            v[j] = i < m ? in[i] : 0u;
            acc += v[j];
        }
    }
    uint32_t tot;
    uint32_t run = block_excl(acc, s_w, &tot) + sums[blockIdx.x];
    if (vec) {
        uint4* out4 = reinterpret_cast<uint4*>(out + t0);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            uint4 y;
            y.x = run;
            run += v[4 * q + 0];
            y.y = run;
            run += v[4 * q + 1];
            y.z = run;
            run += v[4 * q + 2];
            y.w = run;
            run += v[4 * q + 3];
            out4[q] = y;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const uint32_t i = t0 + j;
            if (i < m) out[i] = run;
            run += v[j];
        }
    }
}

}  // namespace

void launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, const uint32_t* n_dev,
                           uint32_t* total, uint32_t* block_sums, hipStream_t stream) {
    const uint32_t nb = div_up(n, kScanTile);
    if (nb == 0) {
        // still define the total
        k_scan_sums<<<1, 1024, 0, stream>>>(block_sums, 0, total);
        return;
    }
    k_scan_reduce<<<nb, kScanThreads, 0, stream>>>(in, n, n_dev, block_sums);
    k_scan_sums<<<1, 1024, 0, stream>>>(block_sums, nb, total);
    k_scan_apply<<<nb, kScanThreads, 0, stream>>>(in, out, n, n_dev, block_sums);
}

}  // namespace palz
