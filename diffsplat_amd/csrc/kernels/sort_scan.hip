// Device-wide exclusive scan and one stable LSD radix pass, wave64.
//
// The radix pass replaces cub::DeviceRadixSort::SortPairs as called at
// src/layers/tile_rasterize_layer.cu:667-683 (stable, ascending). Each pass
// is reduce-then-scan: a per-chunk digit histogram, an exclusive scan of the
// digit-major histogram matrix, then a scatter in which each chunk ranks its
// keys stably with wave64 ballot matching, reorders them in LDS by digit and
// writes runs of equal digits contiguously.
#include "sort_scan.hpp"

namespace ds {
namespace {

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t *__restrict__ in,
                                                              const uint32_t *__restrict__ gidx,
                                                              ScanLen L, uint32_t *part) {
    __shared__ uint32_t s_tmp[kScanThreads / kWave];
    const uint32_t len = scan_len(L);
    const uint32_t nblk = (len + kScanChunk - 1) / kScanChunk;
    for (uint32_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const uint32_t base = b * kScanChunk;
        uint32_t sum = 0;
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const uint32_t e = base + j * kScanThreads + threadIdx.x;
            if (e < len) sum += gidx ? in[gidx[e]] : in[e];
        }
        uint32_t tot;
        block_excl_scan<kScanThreads>(sum, s_tmp, &tot);
        if (threadIdx.x == 0) part[b] = tot;
    }
}

__global__ __launch_bounds__(kSpineThreads) void k_scan_spine(uint32_t *part, ScanLen L,
                                                              uint32_t *total_out,
                                                              uint32_t *clamped_out,
                                                              uint32_t cap) {
    __shared__ uint32_t s_tmp[kSpineThreads / kWave];
    const uint32_t len = scan_len(L);
    const uint32_t nblk = (len + kScanChunk - 1) / kScanChunk;
    uint32_t v[kSpineItems];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kSpineItems; ++j) {
        const uint32_t e = threadIdx.x * kSpineItems + j;
        v[j] = (e < nblk) ? part[e] : 0u;
        sum += v[j];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan<kSpineThreads>(sum, s_tmp, &tot);
#pragma unroll
    for (int j = 0; j < kSpineItems; ++j) {
        const uint32_t e = threadIdx.x * kSpineItems + j;
        if (e < nblk) part[e] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) {
        if (total_out) *total_out = tot;
        if (clamped_out) *clamped_out = tot < cap ? tot : cap;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(const uint32_t *in,
                                                            const uint32_t *__restrict__ gidx,
                                                            ScanLen L,
                                                            const uint32_t *__restrict__ part,
                                                            uint32_t *out) {
    __shared__ uint32_t s_tmp[kScanThreads / kWave];
    const uint32_t len = scan_len(L);
    const uint32_t nblk = (len + kScanChunk - 1) / kScanChunk;
    for (uint32_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const uint32_t base = b * kScanChunk + threadIdx.x * kScanItems;
        uint32_t v[kScanItems];
        uint32_t sum = 0;
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const uint32_t e = base + j;
            v[j] = (e < len) ? (gidx ? in[gidx[e]] : in[e]) : 0u;
            sum += v[j];
        }
        uint32_t run = block_excl_scan<kScanThreads>(sum, s_tmp, nullptr) + part[b];
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const uint32_t e = base + j;
            if (e < len) out[e] = run;
            run += v[j];
        }
    }
}

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K k, int shift) {
    return (uint32_t)(k >> shift) & (kRadix - 1);
}

template <typename K>
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const K *__restrict__ kin,
                                                             const uint32_t *n_dev, int shift,
                                                             uint32_t *hist) {
    __shared__ uint32_t s_h[kRadix];
    const uint32_t n = *n_dev;
    const uint32_t nch = (n + kSortChunk - 1) / kSortChunk;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        s_h[threadIdx.x] = 0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
            const uint32_t e = c * kSortChunk + j * kSortThreads + threadIdx.x;
            if (e < n) atomicAdd(&s_h[digit_of(kin[e], shift)], 1u);
        }
        __syncthreads();
        hist[threadIdx.x * nch + c] = s_h[threadIdx.x];
        __syncthreads();
    }
}

template <typename K>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const K *__restrict__ kin, const uint32_t *__restrict__ vin, K *__restrict__ kout,
    uint32_t *__restrict__ vout, const uint32_t *n_dev, int shift,
    const uint32_t *__restrict__ hist) {
    constexpr int NW = kSortThreads / kWave;
    __shared__ K s_k[kSortChunk];
    __shared__ uint32_t s_v[kSortChunk];
    __shared__ uint32_t s_w[NW][kRadix];
    __shared__ uint32_t s_run[kRadix];
    __shared__ uint32_t s_start[kRadix];
    __shared__ uint32_t s_gb[kRadix];
    __shared__ uint32_t s_tmp[NW];

    const uint32_t n = *n_dev;
    const uint32_t nch = (n + kSortChunk - 1) / kSortChunk;
    const int tid = threadIdx.x;
    const int w = tid / kWave;
    const uint64_t lt = lanemask_lt();

    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const uint32_t base = c * kSortChunk;
        const uint32_t cnt = min((uint32_t)kSortChunk, n - base);
        K key[kSortItems];
        uint32_t val[kSortItems];
        uint32_t rank[kSortItems];
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
            const uint32_t e = j * kSortThreads + tid;
            if (e < cnt) {
                key[j] = kin[base + e];
                val[j] = vin[base + e];
            } else {
                key[j] = 0;
                val[j] = 0;
            }
        }
        s_run[tid] = 0;
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
#pragma unroll
            for (int q = 0; q < NW; ++q) s_w[q][tid] = 0;
            __syncthreads();
            const bool valid = (uint32_t)(j * kSortThreads + tid) < cnt;
            const uint32_t d = digit_of(key[j], shift);
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < kRadixBits; ++bit) {
                const bool on = (d >> bit) & 1u;
                const uint64_t b = __ballot(on);
                peers &= on ? b : ~b;
            }
            const uint32_t r = (uint32_t)__popcll(peers & lt);
            if (valid && r == 0) s_w[w][d] = (uint32_t)__popcll(peers);
            __syncthreads();
            {
                uint32_t run = s_run[tid];
#pragma unroll
                for (int q = 0; q < NW; ++q) {
                    const uint32_t cq = s_w[q][tid];
                    s_w[q][tid] = run;
                    run += cq;
                }
                s_run[tid] = run;
            }
            __syncthreads();
            rank[j] = valid ? s_w[w][d] + r : 0u;
            __syncthreads();
        }
        const uint32_t start = block_excl_scan<kSortThreads>(s_run[tid], s_tmp, nullptr);
        s_start[tid] = start;
        s_gb[tid] = hist[tid * nch + c];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
            if ((uint32_t)(j * kSortThreads + tid) < cnt) {
                const uint32_t p = s_start[digit_of(key[j], shift)] + rank[j];
                s_k[p] = key[j];
                s_v[p] = val[j];
            }
        }
        __syncthreads();
        for (uint32_t p = tid; p < cnt; p += kSortThreads) {
            const K k = s_k[p];
            const uint32_t d = digit_of(k, shift);
            const uint32_t pos = s_gb[d] + (p - s_start[d]);
            kout[pos] = k;
            vout[pos] = s_v[p];
        }
        __syncthreads();
    }
}

int clamp_grid(uint64_t blocks, int cap) {
    if (blocks < 1) return 1;
    return blocks > (uint64_t)cap ? cap : (int)blocks;
}

}  // namespace

ds_status scan_exclusive_u32(const uint32_t *in, const uint32_t *gather_idx, uint32_t *out,
                             ScanLen len, uint32_t max_len, uint32_t *part, uint32_t *total_out,
                             uint32_t *clamped_out, uint32_t clamp_cap, hipStream_t s) {
    const uint64_t max_blocks = ((uint64_t)max_len + kScanChunk - 1) / kScanChunk;
    if (max_blocks > (uint64_t)kSpineMax)
        return set_error(DS_ERR_CAPACITY, "scan longer than " +
                                              std::to_string((uint64_t)kSpineMax * kScanChunk));
    const int grid = clamp_grid(max_blocks, 1024);
    if (max_blocks > 0) {
        hipLaunchKernelGGL(k_scan_reduce, dim3(grid), dim3(kScanThreads), 0, s, in, gather_idx,
                           len, part);
        DS_CHECK_LAUNCH("scan_reduce");
    }
    hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(kSpineThreads), 0, s, part, len, total_out,
                       clamped_out, clamp_cap);
    DS_CHECK_LAUNCH("scan_spine");
    if (max_blocks > 0) {
        hipLaunchKernelGGL(k_scan_down, dim3(grid), dim3(kScanThreads), 0, s, in, gather_idx,
                           len, (const uint32_t *)part, out);
        DS_CHECK_LAUNCH("scan_down");
    }
    return DS_OK;
}

template <typename K>
ds_status radix_pass(const K *kin, const uint32_t *vin, K *kout, uint32_t *vout,
                     const uint32_t *n_dev, uint32_t max_n, int shift, uint32_t *hist,
                     uint32_t *part, hipStream_t s) {
    const uint64_t max_chunks = ((uint64_t)max_n + kSortChunk - 1) / kSortChunk;
    if (max_chunks == 0) return DS_OK;
    const int grid = clamp_grid(max_chunks, 2048);
    hipLaunchKernelGGL(k_radix_hist<K>, dim3(grid), dim3(kSortThreads), 0, s, kin, n_dev, shift,
                       hist);
    DS_CHECK_LAUNCH("radix_hist");
    DS_TRY(scan_exclusive_u32(hist, nullptr, hist, ScanLen{n_dev, (uint32_t)kSortChunk,
                                                           (uint32_t)kRadix},
                              (uint32_t)(max_chunks * kRadix), part, nullptr, nullptr, 0, s));
    hipLaunchKernelGGL(k_radix_scatter<K>, dim3(grid), dim3(kSortThreads), 0, s, kin, vin, kout,
                       vout, n_dev, shift, (const uint32_t *)hist);
    DS_CHECK_LAUNCH("radix_scatter");
    return DS_OK;
}

template ds_status radix_pass<uint32_t>(const uint32_t *, const uint32_t *, uint32_t *,
                                        uint32_t *, const uint32_t *, uint32_t, int, uint32_t *,
                                        uint32_t *, hipStream_t);
template ds_status radix_pass<uint64_t>(const uint64_t *, const uint32_t *, uint64_t *,
                                        uint32_t *, const uint32_t *, uint32_t, int, uint32_t *,
                                        uint32_t *, hipStream_t);

}  // namespace ds
