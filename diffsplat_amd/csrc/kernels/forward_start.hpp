// Device side of ds::ForwardStart: the work of a rasterizer forward's first
// kernel, done by the threads of whichever kernel runs it (thread index i
// of `stride` threads).
#pragma once
#include "../ds_internal.hpp"

namespace ds {

__device__ __forceinline__ void forward_start(const ForwardStart &f, int64_t i,
                                              int64_t stride) {
    if (i == 0) {
        if (f.tick) *f.tick += 1;
        if (f.epoch) {
            const uint32_t e = *f.epoch + 1u;
            *f.epoch = e ? e : 1u;
        }
    }
#pragma unroll
    for (int r = 0; r < kStartRanges; ++r)
        for (int64_t t = i; t < (int64_t)f.nzero[r]; t += stride) {
            uint32_t *w = f.zero[r] + t;
            *w = (w == f.set_word) ? f.set_value : 0u;
        }
    for (int64_t t = i; t < f.ngrad; t += stride) {
#pragma unroll
        for (int g = 0; g < 10; ++g)
            if (f.grad[g]) f.grad[g][t] = 0.f;
    }
}

}  // namespace ds
