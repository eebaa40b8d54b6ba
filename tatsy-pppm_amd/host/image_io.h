// image_io.h -- image file containers of libpppm_hip.so (image_io.cc), shared
// by the host writer pppm_image_write and the device encoder behind
// pppm_image_save.  Pixels arrive encoded: rgb8 [h][w][3] and rgbe [h][w][4]
// top row first, f32 [h][w][3] top row first.  Each returns a PPPM_* status.
#pragma once

#include <stdint.h>

int image_write_png8(const char* path, const unsigned char* rgb8, int32_t w, int32_t h);
int image_write_bmp8(const char* path, const unsigned char* rgb8, int32_t w, int32_t h);
int image_write_rgbe(const char* path, const unsigned char* rgbe, int32_t w, int32_t h);
int image_write_pfm(const char* path, const float* rgb, int32_t w, int32_t h);
