// Device-wide exclusive scan and stable LSD radix sort passes (wave64,
// reduce-then-scan).  Element counts are read from device memory so that a
// pass can sit in a captured graph without a host round trip; grids are
// grid-stride over chunks.
#pragma once

#include "../ds_device.hpp"
#include "../ds_internal.hpp"

namespace ds {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanChunk = kScanThreads * kScanItems;   // 2048
constexpr int kSpineThreads = 1024;
constexpr int kSpineItems = 8;
constexpr int kSpineMax = kSpineThreads * kSpineItems;  // 8192 chunks -> 16M elements

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortChunk = kSortThreads * kSortItems;   // 2048
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;                  // 256 == kSortThreads
// The wide digit of radix_pass: three passes cover 32-bit keys instead of four.
constexpr int kWideRadixBits = 11;

// Length of a scan: ceil(n / div) * mul with n read from device memory.
struct ScanLen {
    const uint32_t *n_dev;
    uint32_t div;
    uint32_t mul;
};

__device__ __forceinline__ uint32_t scan_len(const ScanLen &L) {
    const uint32_t n = *L.n_dev;
    return ((n + L.div - 1) / L.div) * L.mul;
}

// Inclusive wave64 scan.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const int l = __lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t y = __shfl_up(x, d, kWave);
        if (l >= d) x += y;
    }
    return x;
}

// Exclusive scan of one value per thread across a block of NT threads;
// returns the exclusive prefix, *total receives the block sum.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_tmp,
                                                    uint32_t *total) {
    constexpr int NW = NT / kWave;
    const int w = threadIdx.x / kWave;
    const uint32_t inc = wave_incl_scan(v);
    if (__lane_id() == kWave - 1) s_tmp[w] = inc;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const uint32_t s = s_tmp[k];
        wbase += (k < w) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    if (total) *total = tot;
    return wbase + inc - v;
}

// Optional scan side output: for every multiple b * 2^shift (b < count) of
// the output positions that falls in an element's range [out[e], out[e] +
// in[e]), first[b] = e. With first == nullptr nothing is written.
struct ScanHeads {
    uint32_t *first;
    uint32_t shift;
    uint32_t count;
};

__device__ __forceinline__ void mark_heads(const ScanHeads &H, uint32_t e, uint32_t start,
                                           uint32_t len) {
    if (!H.first || len == 0) return;
    const uint64_t end = (uint64_t)start + len;
    for (uint64_t b = ((uint64_t)start + (1u << H.shift) - 1) >> H.shift;
         b < H.count && (b << H.shift) < end; ++b)
        H.first[b] = e;
}

ds_status scan_exclusive_u32(const uint32_t *in, const uint32_t *gather_idx, uint32_t *out,
                             ScanLen len, uint32_t max_len, uint32_t *part, uint32_t *total_out,
                             uint32_t *clamped_out, uint32_t clamp_cap, hipStream_t s,
                             ScanHeads heads = ScanHeads{nullptr, 0, 0});

// One stable LSD pass: sort (key, val) by bits [shift, shift + bits) of the
// key, bits 8 or 11 (kWideRadixBits). n_dev: element count; hist: scratch of
// 2^bits * ceil(max_n / kSortChunk) words (sort_hist_words); part: scan
// scratch of kSpineMax words. At 11 bits the digit counts (2048 per chunk of
// 2048 keys) must fit one scan: max_n up to 2^24.
template <typename K>
ds_status radix_pass(const K *kin, const uint32_t *vin, K *kout, uint32_t *vout,
                     const uint32_t *n_dev, uint32_t max_n, int shift, uint32_t *hist,
                     uint32_t *part, hipStream_t s, int bits = kRadixBits);

inline uint32_t sort_hist_words(uint64_t max_n, int bits = kRadixBits) {
    return (uint32_t)((1u << bits) * ((max_n + kSortChunk - 1) / kSortChunk));
}

// Scratch of a stand-alone (u64 key, u32 value) sort of up to n pairs: key
// and value ping-pong buffers of n each, sort_hist_words(n) histogram words,
// kSpineMax scan words and one count word. Nothing in it needs clearing: the
// count word is set and every other word is written before it is read.
struct SortPairsScratch {
    uint64_t *keys;
    uint32_t *vals;
    uint32_t *hist;
    uint32_t *part;
    uint32_t *count;
};

// Stable LSD sort of n pairs on the low key_bits bits, 8 bits per pass,
// ping-ponging through the scratch so the last pass lands in keys_out /
// vals_out. Asynchronous on s; allocates nothing.
ds_status sort_pairs_u64_scratch(const uint64_t *keys_in, const uint32_t *vals_in,
                                 uint64_t *keys_out, uint32_t *vals_out, uint32_t n,
                                 int key_bits, const SortPairsScratch &scr, hipStream_t s);

// ------------------------------------------------------------------
// One-kernel passes with decoupled look-back.
//
// A radix pass is one kernel: each workgroup draws the next chunk
// index from a ticket counter, publishes its chunk's per-digit counts, and
// looks back over earlier chunks for its exclusive prefix. A workgroup only
// waits on chunks with lower tickets, whose workgroups are running and
// publish their counts before they wait themselves, so the chain always
// advances. Status words are {epoch, state | value}: a word written in an
// earlier forward carries an older epoch and reads as "not ready", so the
// status memory is never cleared (it is zero at allocation; epochs start at
// 1). The spin is bounded: if it ever gives up it sets *err and the result
// is wrong but the kernel ends.
struct LookbackSlot {
    uint64_t *status;       // kRadix words per chunk
    uint32_t *ticket;       // zero before the pass
    const uint32_t *epoch;  // current forward's epoch, never 0
    uint32_t *err;
};

// Values in status words are below 2^30.
constexpr uint32_t kLookbackMaxCount = (1u << 30) - 1;

inline size_t lookback_words(uint64_t max_n, int digits) {
    return (size_t)digits * (size_t)((max_n + kSortChunk - 1) / kSortChunk);
}

// Optional side output of the last radix pass: the exclusive scan of a
// per-value weight (weight[v] for the value v of a key) in the pass's output
// order, so that no separate scan pass over the sorted values is needed.
// whist holds, per digit of the pass, the sum of the weights of all keys
// with that digit (reps partial copies of kRadix words, copy r at whist +
// r * kRadix). The scan, total and heads follow scan_chained: prefixes and the
// total saturate at kLookbackMaxCount, clamped = min(total, cap), and
// heads marks the owner of every multiple of 2^heads.shift.
struct PairOffsets {
    const uint32_t *weight;  // nullptr: no side output
    const unsigned long long *whist;
    int whist_reps;
    uint32_t *out;
    uint64_t *status;  // look-back words of the weight sums, kRadix per chunk
    uint32_t *total_out, *clamped_out;
    uint32_t cap;
    ScanHeads heads;
};

// Stable pass on bits [shift, shift + 8) of 32-bit keys. The pass's 256-bin
// digit histogram over all n keys is the sum of hist_reps partial copies,
// copy r at ghist + r * hist_rep_stride.
ds_status onesweep_pass(const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                        uint32_t *vout, const uint32_t *n_dev, uint32_t max_n, int shift,
                        const uint32_t *ghist, int hist_reps, uint32_t hist_rep_stride,
                        LookbackSlot lb, hipStream_t s,
                        PairOffsets po = PairOffsets{nullptr, nullptr, 0, nullptr, nullptr,
                                                     nullptr, nullptr, 0,
                                                     ScanHeads{nullptr, 0, 0}});

}  // namespace ds
