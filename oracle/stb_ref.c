/* Builds the reference's own image path (stb_image decoding and
 * stb_image_resize2's linear-space resize, as src/io/image_io.cpp:26 and :44
 * call them) from the headers where they lie in the reference tree, into
 * oracle/_ref/libstbref.so: test infrastructure only, used to check the
 * engine's JPEG decoder and resampler against the reference's. */
#define STB_IMAGE_IMPLEMENTATION
#include "stb_image.h"
#define STB_IMAGE_RESIZE_IMPLEMENTATION
#include "stb_image_resize2.h"

/* stbi_load(path, &w, &h, &channels, 3); NULL on failure. */
unsigned char *ref_load_rgb(const char *path, int *w, int *h) {
    int channels = 0;
    return stbi_load(path, w, h, &channels, 3);
}

void ref_free(unsigned char *p) { stbi_image_free(p); }

/* stbir_resize_uint8_linear(src, sw, sh, 0, dst, dw, dh, 0, STBIR_RGB). */
int ref_resize_rgb(const unsigned char *src, int sw, int sh, unsigned char *dst, int dw, int dh) {
    return stbir_resize_uint8_linear(src, sw, sh, 0, dst, dw, dh, 0, STBIR_RGB) != NULL;
}
