// Film writers for the image formats the reference's Image::save_to_file
// accepts (texture.cu:219-230): .exr (HALF channels), .hdr (Radiance RGBE) and
// the 8-bit formats .png, .bmp, .jpg, plus .pfm and .ppm.  `px(row, col)`
// gives the RGBA floats of the pixel at `row` from the top, `col` from the
// left (the caller applies the reference's x/y flip).
#pragma once

#include <cstdint>
#include <functional>
#include <ostream>

namespace srt {

using PixelFn = std::function<const float*(int row_from_top, int col)>;

// 8-bit channel value as the reference's Image::f_to_u (texture.cu:11-21):
// clamp to [0, 1], times 255, truncated
uint8_t to_ldr(float x);
// IEEE binary16 of x, round to nearest even (OpenEXR HALF)
uint16_t to_half(float x);

void write_exr(std::ostream& f, int W, int H, const PixelFn& px);
void write_hdr(std::ostream& f, int W, int H, const PixelFn& px);
void write_png(std::ostream& f, int W, int H, const PixelFn& px);
void write_bmp(std::ostream& f, int W, int H, const PixelFn& px);
void write_jpg(std::ostream& f, int W, int H, const PixelFn& px);

}  // namespace srt
