// JPEG decoding for target images (host code; see jpeg.cpp).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace dsh {

struct JpegImage {
    int w = 0, h = 0;
    std::vector<uint8_t> rgb;  // w * h * 3, row-major
};

// Decodes a baseline, extended-sequential or progressive Huffman JPEG with 1
// or 3 components of 8-bit samples. Returns false with a message otherwise.
bool decode_jpeg(const uint8_t *data, size_t size, JpegImage &out, std::string &err);

}  // namespace dsh
