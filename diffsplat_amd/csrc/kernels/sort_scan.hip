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
                                                            uint32_t *out, ScanHeads H) {
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
            mark_heads(H, e, run, v[j]);
            run += v[j];
        }
    }
}

template <typename K, int RB = kRadixBits>
__device__ __forceinline__ uint32_t digit_of(K k, int shift) {
    return (uint32_t)(k >> shift) & ((1u << RB) - 1u);
}

template <typename K, int RB>
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const K *__restrict__ kin,
                                                             const uint32_t *n_dev, int shift,
                                                             uint32_t *hist) {
    constexpr int R = 1 << RB;
    __shared__ uint32_t s_h[R];
    const uint32_t n = *n_dev;
    const uint32_t nch = (n + kSortChunk - 1) / kSortChunk;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        for (int d = threadIdx.x; d < R; d += kSortThreads) s_h[d] = 0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
            const uint32_t e = c * kSortChunk + j * kSortThreads + threadIdx.x;
            if (e < n) atomicAdd(&s_h[digit_of<K, RB>(kin[e], shift)], 1u);
        }
        __syncthreads();
        for (int d = threadIdx.x; d < R; d += kSortThreads) hist[d * nch + c] = s_h[d];
        __syncthreads();
    }
}

// LDS of one workgroup ranking and scattering a chunk of kSortChunk keys by
// RB-bit digits (2^RB digits: one per thread at 8 bits, eight at 11).
template <typename K, int RB = kRadixBits>
struct ChunkLds {
    static constexpr int NW = kSortThreads / kWave;
    static constexpr int R = 1 << RB;
    K k[kSortChunk];
    uint32_t v[kSortChunk];
    uint32_t w[NW][R];  // per-wave digit counts, then per-wave digit offsets
    uint32_t run[R];    // digit counts of the chunk
    uint32_t start[R];  // exclusive prefix of run over digits
    uint32_t gb[R];     // global position of the chunk's first key of each digit
    uint32_t tmp[NW];
    uint32_t ticket;
};

// Loads the cnt keys and values of the chunk starting at base: wave w owns the
// contiguous keys [w * kSortChunk / NW, (w + 1) * kSortChunk / NW) of the
// chunk, item j of a lane at w * kSortChunk / NW + 64 j + lane (zero past cnt).
template <typename K>
__device__ __forceinline__ void chunk_load(const K *__restrict__ kin,
                                           const uint32_t *__restrict__ vin, uint32_t base,
                                           uint32_t cnt, K (&key)[kSortItems],
                                           uint32_t (&val)[kSortItems]) {
    constexpr int kWaveKeys = kSortChunk / (kSortThreads / kWave);
    const int w = threadIdx.x / kWave, lane = __lane_id();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t e = (uint32_t)(w * kWaveKeys + j * kWave + lane);
        if (e < cnt) {
            key[j] = kin[base + e];
            val[j] = vin[base + e];
        } else {
            key[j] = 0;
            val[j] = 0;
        }
    }
}

// Ranks the loaded keys of a chunk (chunk_load) stably within their digit:
// rank[j] is the number of earlier keys of the chunk with the same digit.
// Leaves the digit counts in L.run and their exclusive prefix in L.start.
// Each wave walks its keys 64 at a time in order, matching the RB digit bits
// across the wave with ballots and keeping a running count per digit in its
// own LDS row (lanes of one digit share one first lane, so the row's words
// written in one step are distinct); the waves' counts are then scanned per
// digit. Four barriers per chunk.
template <typename K, int RB = kRadixBits>
__device__ __forceinline__ void chunk_rank_loaded(uint32_t cnt, int shift,
                                                  const K (&key)[kSortItems],
                                                  uint32_t (&rank)[kSortItems],
                                                  ChunkLds<K, RB> &L) {
    constexpr int NW = ChunkLds<K, RB>::NW;
    constexpr int R = ChunkLds<K, RB>::R;
    constexpr int DPT = R / kSortThreads;  // digits per thread
    constexpr int kWaveKeys = kSortChunk / NW;
    const int tid = threadIdx.x;
    const int w = tid / kWave, lane = __lane_id();
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int i = 0; i < DPT; ++i)
#pragma unroll
        for (int q = 0; q < NW; ++q) L.w[q][i * kSortThreads + tid] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const bool valid = (uint32_t)(w * kWaveKeys + j * kWave + lane) < cnt;
        const uint32_t d = digit_of<K, RB>(key[j], shift);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < RB; ++bit) {
            const bool on = (d >> bit) & 1u;
            const uint64_t b = __ballot(on);
            peers &= on ? b : ~b;
        }
        const uint32_t r = (uint32_t)__popcll(peers & lt);
        uint32_t before = 0;
        if (valid) before = L.w[w][d];
        if (valid && r == 0) L.w[w][d] = before + (uint32_t)__popcll(peers);
        rank[j] = before + r;  // among this wave's keys of digit d
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < DPT; ++i) {  // per digit: the waves' first ranks, the count
        const int d = i * kSortThreads + tid;
        uint32_t run = 0;
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            const uint32_t cq = L.w[q][d];
            L.w[q][d] = run;
            run += cq;
        }
        L.run[d] = run;
    }
    // the prefixes of every digit are written before any thread reads them
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) rank[j] += L.w[w][digit_of<K, RB>(key[j], shift)];
    if constexpr (DPT == 1) {
        L.start[tid] = block_excl_scan<kSortThreads>(L.run[tid], L.tmp, nullptr);
    } else {  // thread t scans its DPT consecutive digits
        uint32_t x[DPT], sum = 0;
#pragma unroll
        for (int i = 0; i < DPT; ++i) {
            x[i] = L.run[tid * DPT + i];
            sum += x[i];
        }
        uint32_t run = block_excl_scan<kSortThreads>(sum, L.tmp, nullptr);
#pragma unroll
        for (int i = 0; i < DPT; ++i) {
            L.start[tid * DPT + i] = run;
            run += x[i];
        }
    }
}

// chunk_load then chunk_rank_loaded.
template <typename K, int RB = kRadixBits>
__device__ __forceinline__ void chunk_rank(const K *__restrict__ kin,
                                           const uint32_t *__restrict__ vin, uint32_t base,
                                           uint32_t cnt, int shift, K (&key)[kSortItems],
                                           uint32_t (&val)[kSortItems],
                                           uint32_t (&rank)[kSortItems], ChunkLds<K, RB> &L) {
    chunk_load(kin, vin, base, cnt, key, val);
    chunk_rank_loaded<K, RB>(cnt, shift, key, rank, L);
}

// Reorders the ranked chunk by digit in LDS, then writes each run of equal
// digits contiguously at L.gb[digit] (which the caller has filled).
template <typename K, int RB = kRadixBits>
__device__ __forceinline__ void chunk_scatter(const K (&key)[kSortItems],
                                              const uint32_t (&val)[kSortItems],
                                              const uint32_t (&rank)[kSortItems], uint32_t cnt,
                                              int shift, ChunkLds<K, RB> &L, K *__restrict__ kout,
                                              uint32_t *__restrict__ vout) {
    const int tid = threadIdx.x;
    constexpr int kWaveKeys = kSortChunk / ChunkLds<K, RB>::NW;
    const int w = tid / kWave, lane = __lane_id();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        if ((uint32_t)(w * kWaveKeys + j * kWave + lane) < cnt) {
            const uint32_t p = L.start[digit_of<K, RB>(key[j], shift)] + rank[j];
            L.k[p] = key[j];
            L.v[p] = val[j];
        }
    }
    __syncthreads();
    for (uint32_t p = tid; p < cnt; p += kSortThreads) {
        const K k = L.k[p];
        const uint32_t d = digit_of<K, RB>(k, shift);
        const uint32_t pos = L.gb[d] + (p - L.start[d]);
        kout[pos] = k;
        vout[pos] = L.v[p];
    }
    __syncthreads();
}

template <typename K, int RB>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const K *__restrict__ kin, const uint32_t *__restrict__ vin, K *__restrict__ kout,
    uint32_t *__restrict__ vout, const uint32_t *n_dev, int shift,
    const uint32_t *__restrict__ hist) {
    constexpr int R = 1 << RB;
    __shared__ ChunkLds<K, RB> L;
    const uint32_t n = *n_dev;
    const uint32_t nch = (n + kSortChunk - 1) / kSortChunk;
    const int tid = threadIdx.x;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const uint32_t base = c * kSortChunk;
        const uint32_t cnt = min((uint32_t)kSortChunk, n - base);
        K key[kSortItems];
        uint32_t val[kSortItems], rank[kSortItems];
        chunk_rank<K, RB>(kin, vin, base, cnt, shift, key, val, rank, L);
        for (int d = tid; d < R; d += kSortThreads) L.gb[d] = hist[d * nch + c];
        __syncthreads();
        chunk_scatter<K, RB>(key, val, rank, cnt, shift, L, kout, vout);
    }
}

// ---------------------------------------------------------------- look-back

enum : uint32_t { kLbAggregate = 1u, kLbPrefix = 2u };
// Polls of one status word before giving up (each poll sleeps ~64 cycles);
// a running predecessor publishes within microseconds.
constexpr uint32_t kLbSpinLimit = 1u << 20;

__device__ __forceinline__ void lb_publish(uint64_t *word, uint32_t epoch, uint32_t state,
                                           uint32_t value) {
    const uint64_t w = ((uint64_t)epoch << 32) | ((uint64_t)state << 30) | value;
    __hip_atomic_store(word, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of the values of the chunks before chunk c in one digit column
// (words at status[p * stride] for p < c), walking back until a chunk
// with its inclusive prefix published. Every chunk before c has published
// at least its aggregate, or is about to (see LookbackSlot).
// The sum saturates at kLookbackMaxCount. The words of kLbWindow
// predecessors are loaded at once (one memory round trip per window rather
// than per chunk) and consumed nearest first.
constexpr int kLbWindow = 16;

__device__ __forceinline__ uint32_t lb_exclusive(const uint64_t *status, size_t stride,
                                                 uint32_t c, uint32_t epoch, uint32_t *err) {
    uint64_t excl = 0;
    uint32_t spins = 0;
    int64_t p = (int64_t)c - 1;  // nearest chunk not yet added
    while (p >= 0) {
        uint64_t w[kLbWindow];
#pragma unroll
        for (int j = 0; j < kLbWindow; ++j)
            w[j] = (p - j >= 0) ? __hip_atomic_load(status + (size_t)(p - j) * stride,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                : 0ull;
        int used = 0;  // words consumed from the window
        bool done = false, stalled = false;
#pragma unroll
        for (int j = 0; j < kLbWindow; ++j) {
            if (done || stalled || p - j < 0) continue;
            const uint32_t state = (uint32_t)(w[j] >> 30) & 3u;
            if ((uint32_t)(w[j] >> 32) != epoch || state == 0u) {
                stalled = true;  // not published yet: reload from here
                continue;
            }
            excl += (uint32_t)w[j] & kLookbackMaxCount;
            ++used;
            done = (state == kLbPrefix);
        }
        if (done) break;
        p -= used;
        if (used > 0) {
            spins = 0;
        } else if (++spins > kLbSpinLimit) {
            atomicOr(err, 1u);
            break;
        } else {
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return (uint32_t)min(excl, (uint64_t)kLookbackMaxCount);
}

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
    return (uint32_t)min((uint64_t)a + b, (uint64_t)kLookbackMaxCount);
}

// LDS of the pair-offset side output: the chunk's weights in digit order,
// then their exclusive scan (entry kSortChunk: the chunk's total), and per
// digit the output offset of the chunk's first key of that digit.
struct OffsetLds {
    uint32_t w[kSortChunk];
    uint32_t s[kSortChunk + 1];
    uint32_t wb[kRadix];
    unsigned long long tmp[kSortThreads / kWave];
};
template <bool ON>
struct OffsetLdsIf {
    OffsetLds v;
};
template <>
struct OffsetLdsIf<false> {};

// Exclusive scan of one 64-bit value per thread over a 256-thread block.
__device__ __forceinline__ unsigned long long block_excl_scan64(unsigned long long v,
                                                                unsigned long long *s_tmp,
                                                                unsigned long long *total) {
    constexpr int NW = kSortThreads / kWave;
    const int w = threadIdx.x / kWave, lane = __lane_id();
    unsigned long long inc = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const unsigned long long y = __shfl_up(inc, d, kWave);
        if (lane >= d) inc += y;
    }
    if (lane == kWave - 1) s_tmp[w] = inc;
    __syncthreads();
    unsigned long long base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        base += (k < w) ? s_tmp[k] : 0ull;
        tot += s_tmp[k];
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

// One stable radix pass in one kernel (see LookbackSlot). With OFF it also
// writes the exclusive scan of weight[value] in the output order (see
// PairOffsets): a key's offset is the weight of all keys of smaller digits
// (from whist) plus that of the keys of its digit in earlier chunks (a
// second look-back over per-digit weight sums) plus that of the keys of its
// digit before it in its chunk (a scan of the chunk's weights in digit
// order, in LDS).
template <bool OFF>
__global__ __launch_bounds__(kSortThreads) void k_onesweep(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, const uint32_t *n_dev, int shift,
    const uint32_t *__restrict__ ghist, int hist_reps, uint32_t hist_rep_stride,
    LookbackSlot lb, PairOffsets po) {
    __shared__ ChunkLds<uint32_t> L;
    __shared__ OffsetLdsIf<OFF> O;
    const uint32_t n = *n_dev;
    const uint32_t nch = (n + kSortChunk - 1) / kSortChunk;
    const uint32_t epoch = *lb.epoch;
    const int tid = threadIdx.x;
    // The digit's count over all keys; its loads and the first ticket share
    // one memory wait.
    uint32_t count = 0;
    for (int r = 0; r < hist_reps; ++r) count += ghist[(size_t)r * hist_rep_stride + tid];
    unsigned long long wdig = 0;  // weight of all keys of digit tid
    if constexpr (OFF)
        for (int r = 0; r < po.whist_reps; ++r) wdig += po.whist[(size_t)r * kRadix + tid];
    // The first ticket is usually blockIdx.x: that chunk's keys load while the
    // ticket is in flight (and load again if the ticket differs).
    const uint32_t guess = blockIdx.x;
    uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
    if (guess < nch)
        chunk_load(kin, vin, guess * kSortChunk, min((uint32_t)kSortChunk, n - guess * kSortChunk),
                   key, val);
    if (tid == 0) L.ticket = atomicAdd(lb.ticket, 1u);
    __syncthreads();
    // first position of each digit in the output
    const uint32_t dbase = block_excl_scan<kSortThreads>(count, L.tmp, nullptr);
    unsigned long long wbase = 0;  // weight of all keys of smaller digits
    if constexpr (OFF) {
        unsigned long long wtot;
        wbase = block_excl_scan64(wdig, O.v.tmp, &wtot);
        if (blockIdx.x == 0 && tid == 0) {
            const uint32_t total = (uint32_t)min(wtot, (unsigned long long)kLookbackMaxCount);
            if (po.total_out) *po.total_out = total;
            if (po.clamped_out) *po.clamped_out = min(total, po.cap);
        }
    }
    for (bool first = true;; first = false) {
        if (!first) {
            // a grid of at least one workgroup per chunk: the first tickets
            // covered every chunk
            if (gridDim.x >= nch) break;
            if (tid == 0) L.ticket = atomicAdd(lb.ticket, 1u);
            __syncthreads();
        }
        const uint32_t c = L.ticket;
        if (c >= nch) break;  // uniform
        const uint32_t base = c * kSortChunk;
        const uint32_t cnt = min((uint32_t)kSortChunk, n - base);
        if (!first || c != guess) chunk_load(kin, vin, base, cnt, key, val);
        chunk_rank_loaded(cnt, shift, key, rank, L);
        uint64_t *col = lb.status + (size_t)c * kRadix + tid;
        const uint32_t mine = L.run[tid];
        uint32_t excl = 0;
        if constexpr (!OFF) {
            if (c == 0) {
                lb_publish(col, epoch, kLbPrefix, mine);
            } else {
                lb_publish(col, epoch, kLbAggregate, mine);
                excl = lb_exclusive(lb.status + tid, kRadix, c, epoch, lb.err);
                lb_publish(col, epoch, kLbPrefix, excl + mine);
            }
            L.gb[tid] = dbase + excl;
            __syncthreads();
            chunk_scatter(key, val, rank, cnt, shift, L, kout, vout);
        } else {
            OffsetLds &Q = O.v;
            if (c == 0)
                lb_publish(col, epoch, kLbPrefix, mine);
            else
                lb_publish(col, epoch, kLbAggregate, mine);
            // the chunk's keys, values and weights in digit order
            constexpr int kWaveKeys = kSortChunk / ChunkLds<uint32_t>::NW;
            const int w = tid / kWave, lane = __lane_id();
            uint32_t wt[kSortItems];
#pragma unroll
            for (int j = 0; j < kSortItems; ++j) {
                const bool valid = (uint32_t)(w * kWaveKeys + j * kWave + lane) < cnt;
                wt[j] = valid ? po.weight[val[j]] : 0u;
            }
            __syncthreads();  // every digit's L.start is written
#pragma unroll
            for (int j = 0; j < kSortItems; ++j) {
                if ((uint32_t)(w * kWaveKeys + j * kWave + lane) < cnt) {
                    const uint32_t p = L.start[digit_of(key[j], shift)] + rank[j];
                    L.k[p] = key[j];
                    L.v[p] = val[j];
                    Q.w[p] = wt[j];
                }
            }
            __syncthreads();
            {  // exclusive scan of the weights over the digit order
                uint32_t x[kSortItems], sum = 0;
#pragma unroll
                for (int j = 0; j < kSortItems; ++j) {
                    const uint32_t p = (uint32_t)(tid * kSortItems + j);
                    x[j] = p < cnt ? Q.w[p] : 0u;
                    sum += x[j];
                }
                uint32_t tot;
                uint32_t run = block_excl_scan<kSortThreads>(sum, L.tmp, &tot);
#pragma unroll
                for (int j = 0; j < kSortItems; ++j) {
                    Q.s[tid * kSortItems + j] = run;
                    run += x[j];
                }
                if (tid == 0) Q.s[kSortChunk] = tot;
            }
            __syncthreads();
            // this chunk's weight of digit tid: its keys occupy [start, start + run)
            const uint32_t d0 = L.start[tid];
            const uint32_t wmine = Q.s[min(d0 + mine, cnt)] - Q.s[min(d0, cnt)];
            uint64_t *wcol = po.status + (size_t)c * kRadix + tid;
            uint32_t wexcl = 0;
            if (c == 0) {
                lb_publish(wcol, epoch, kLbPrefix, min(wmine, kLookbackMaxCount));
            } else {
                lb_publish(wcol, epoch, kLbAggregate, min(wmine, kLookbackMaxCount));
                excl = lb_exclusive(lb.status + tid, kRadix, c, epoch, lb.err);
                lb_publish(col, epoch, kLbPrefix, excl + mine);
                wexcl = lb_exclusive(po.status + tid, kRadix, c, epoch, lb.err);
                lb_publish(wcol, epoch, kLbPrefix, sat_add(wexcl, wmine));
            }
            L.gb[tid] = dbase + excl;
            Q.wb[tid] = (uint32_t)min(wbase + wexcl, (unsigned long long)kLookbackMaxCount);
            __syncthreads();
            for (uint32_t p = tid; p < cnt; p += kSortThreads) {
                const uint32_t k = L.k[p];
                const uint32_t d = digit_of(k, shift);
                const uint32_t pos = L.gb[d] + (p - L.start[d]);
                kout[pos] = k;
                vout[pos] = L.v[p];
                const uint32_t off = sat_add(Q.wb[d], Q.s[p] - Q.s[L.start[d]]);
                po.out[pos] = off;
                mark_heads(po.heads, pos, off, Q.s[p + 1] - Q.s[p]);
            }
            __syncthreads();
        }
    }
}

int clamp_grid(uint64_t blocks, int cap) {
    if (blocks < 1) return 1;
    return blocks > (uint64_t)cap ? cap : (int)blocks;
}

}  // namespace

ds_status scan_exclusive_u32(const uint32_t *in, const uint32_t *gather_idx, uint32_t *out,
                             ScanLen len, uint32_t max_len, uint32_t *part, uint32_t *total_out,
                             uint32_t *clamped_out, uint32_t clamp_cap, hipStream_t s,
                             ScanHeads heads) {
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
                           len, (const uint32_t *)part, out, heads);
        DS_CHECK_LAUNCH("scan_down");
    }
    return DS_OK;
}

template <typename K, int RB>
static ds_status radix_pass_rb(const K *kin, const uint32_t *vin, K *kout, uint32_t *vout,
                               const uint32_t *n_dev, uint32_t max_n, int shift, uint32_t *hist,
                               uint32_t *part, hipStream_t s) {
    constexpr int R = 1 << RB;
    const uint64_t max_chunks = ((uint64_t)max_n + kSortChunk - 1) / kSortChunk;
    if (max_chunks == 0) return DS_OK;
    if (max_chunks * R > (uint64_t)kSpineMax * kScanChunk)
        return set_error(DS_ERR_CAPACITY, "radix pass: too many digit counts to scan");
    const int grid = clamp_grid(max_chunks, 2048);
    hipLaunchKernelGGL((k_radix_hist<K, RB>), dim3(grid), dim3(kSortThreads), 0, s, kin, n_dev,
                       shift, hist);
    DS_CHECK_LAUNCH("radix_hist");
    DS_TRY(scan_exclusive_u32(hist, nullptr, hist, ScanLen{n_dev, (uint32_t)kSortChunk,
                                                           (uint32_t)R},
                              (uint32_t)(max_chunks * R), part, nullptr, nullptr, 0, s));
    hipLaunchKernelGGL((k_radix_scatter<K, RB>), dim3(grid), dim3(kSortThreads), 0, s, kin, vin,
                       kout, vout, n_dev, shift, (const uint32_t *)hist);
    DS_CHECK_LAUNCH("radix_scatter");
    return DS_OK;
}

template <typename K>
ds_status radix_pass(const K *kin, const uint32_t *vin, K *kout, uint32_t *vout,
                     const uint32_t *n_dev, uint32_t max_n, int shift, uint32_t *hist,
                     uint32_t *part, hipStream_t s, int bits) {
    if (bits == kWideRadixBits)
        return radix_pass_rb<K, kWideRadixBits>(kin, vin, kout, vout, n_dev, max_n, shift, hist,
                                                part, s);
    if (bits != kRadixBits) return set_error(DS_ERR_INVALID_ARG, "radix digits of 8 or 11 bits");
    return radix_pass_rb<K, kRadixBits>(kin, vin, kout, vout, n_dev, max_n, shift, hist, part,
                                        s);
}

template ds_status radix_pass<uint32_t>(const uint32_t *, const uint32_t *, uint32_t *,
                                        uint32_t *, const uint32_t *, uint32_t, int, uint32_t *,
                                        uint32_t *, hipStream_t, int);
template ds_status radix_pass<uint64_t>(const uint64_t *, const uint32_t *, uint64_t *,
                                        uint32_t *, const uint32_t *, uint32_t, int, uint32_t *,
                                        uint32_t *, hipStream_t, int);

ds_status sort_pairs_u64_scratch(const uint64_t *keys_in, const uint32_t *vals_in,
                                 uint64_t *keys_out, uint32_t *vals_out, uint32_t n,
                                 int key_bits, const SortPairsScratch &scr, hipStream_t s) {
    if (n == 0) return DS_OK;
    DS_TRY(fill_u32(scr.count, n, 1, s));
    const int passes = (key_bits + kRadixBits - 1) / kRadixBits;
    // ping-pong so the last pass lands in the output buffers
    const uint64_t *ksrc = keys_in;
    const uint32_t *vsrc = vals_in;
    for (int p = 0; p < passes; ++p) {
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint64_t *kd = to_out ? keys_out : scr.keys;
        uint32_t *vd = to_out ? vals_out : scr.vals;
        DS_TRY(radix_pass<uint64_t>(ksrc, vsrc, kd, vd, scr.count, n, kRadixBits * p, scr.hist,
                                    scr.part, s));
        ksrc = kd;
        vsrc = vd;
    }
    return DS_OK;
}

ds_status onesweep_pass(const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                        uint32_t *vout, const uint32_t *n_dev, uint32_t max_n, int shift,
                        const uint32_t *ghist, int hist_reps, uint32_t hist_rep_stride,
                        LookbackSlot lb, hipStream_t s, PairOffsets po) {
    if (max_n > kLookbackMaxCount)
        return set_error(DS_ERR_CAPACITY, "radix pass longer than 2^30 - 1 keys");
    const uint64_t max_chunks = ((uint64_t)max_n + kSortChunk - 1) / kSortChunk;
    if (max_chunks == 0) return DS_OK;
    const dim3 grid(clamp_grid(max_chunks, 2048));
    if (po.weight)
        hipLaunchKernelGGL(k_onesweep<true>, grid, dim3(kSortThreads), 0, s, kin, vin, kout, vout,
                           n_dev, shift, ghist, hist_reps, hist_rep_stride, lb, po);
    else
        hipLaunchKernelGGL(k_onesweep<false>, grid, dim3(kSortThreads), 0, s, kin, vin, kout,
                           vout, n_dev, shift, ghist, hist_reps, hist_rep_stride, lb, po);
    DS_CHECK_LAUNCH("onesweep");
    return DS_OK;
}

}  // namespace ds
