// The stand-alone pair sort without any allocation, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops).
//
// Its scratch is one file-scope device array of fixed size. Each GPU holds
// its own copy of the array (at its own address) for the life of the
// process. The host never passes the array's name to a kernel: on the host
// that name is a stand-in, not a device address. The first call on a device
// looks up the device address with hipGetSymbolAddress, with that device
// current, and keeps it per device; every kernel argument and fill uses that
// address.
//
// Calls on one GPU share that GPU's array: calls that overlap in time on one
// GPU (on different streams, or from different host threads) race and give
// wrong results. Nothing in the array needs clearing between calls: the count
// word is set by a fill kernel at the start of every call and every other
// word is written before it is read.
#include <mutex>

#include "../ds_device.hpp"
#include "../ds_internal.hpp"
#include "sort_scan.hpp"

namespace ds {
namespace {

// Largest pair count of one call.
constexpr uint64_t kNoAllocMaxPairs = uint64_t(1) << 20;

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kOffKeys = 0;
constexpr size_t kOffVals = kOffKeys + align256(sizeof(uint64_t) * kNoAllocMaxPairs);
constexpr size_t kOffHist = kOffVals + align256(sizeof(uint32_t) * kNoAllocMaxPairs);
constexpr size_t kHistWords = kRadix * ((kNoAllocMaxPairs + kSortChunk - 1) / kSortChunk);
constexpr size_t kOffPart = kOffHist + align256(sizeof(uint32_t) * kHistWords);
constexpr size_t kOffCount = kOffPart + align256(sizeof(uint32_t) * kSpineMax);
constexpr size_t kScratchBytes = kOffCount + 256;  // about 12.6 MiB

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// Device address of the current device's copy of the scratch array.
ds_status scratch_of_current_device(unsigned char **out) {
    int dev = 0;
    DS_TRY_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(DS_ERR_INVALID_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return set_error(DS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                             hipGetErrorString(e));
        g_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_ptr[dev];
    return DS_OK;
}

}  // namespace
}  // namespace ds

extern "C" ds_status ds_sort_pairs_u64_noalloc(const uint64_t *keys_in, const uint32_t *vals_in,
                                               uint64_t *keys_out, uint32_t *vals_out, int64_t n,
                                               int key_bits, ds_stream_t s) {
    using namespace ds;
    clear_error();
    DS_REQUIRE(n >= 0, "n out of range");
    if ((uint64_t)n > kNoAllocMaxPairs)
        return set_error(DS_ERR_CAPACITY,
                         "ds_sort_pairs_u64_noalloc sorts at most 1048576 pairs per call; "
                         "use ds_sort_pairs_u64 for more");
    DS_REQUIRE(key_bits >= 1 && key_bits <= 64, "key_bits out of range");
    if (n == 0) return DS_OK;
    DS_REQUIRE(keys_in && vals_in && keys_out && vals_out, "NULL argument");
    unsigned char *base = nullptr;
    DS_TRY(scratch_of_current_device(&base));
    const SortPairsScratch scr{reinterpret_cast<uint64_t *>(base + kOffKeys),
                               reinterpret_cast<uint32_t *>(base + kOffVals),
                               reinterpret_cast<uint32_t *>(base + kOffHist),
                               reinterpret_cast<uint32_t *>(base + kOffPart),
                               reinterpret_cast<uint32_t *>(base + kOffCount)};
    return sort_pairs_u64_scratch(keys_in, vals_in, keys_out, vals_out, (uint32_t)n, key_bits,
                                  scr, as_stream(s));
}
