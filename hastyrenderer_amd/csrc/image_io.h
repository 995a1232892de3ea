// Image files for scene textures and the environment map, read into float
// pixels with row 0 at the top (the order OpenImageIO hands the reference,
// core/src/Image.cpp:74-138), and the renderer's output files.  Pure C++17,
// no third-party dependency:
//   PFM  "PF" (RGB) / "Pf" (grey) 32-bit floats
//   PPM / PGM  binary "P6" / "P5" and ASCII "P3" / "P2", 8 or 16 bits per
//        sample, scaled by 1 / maxval as OIIO converts integer data to float
//   PNG  grey, grey + alpha, RGB, RGBA and palette images of 8 or 16 bits
//        (1, 2, 4 bits for grey and palette), non-interlaced, decoded with
//        this file's own inflate (RFC 1950 / 1951) and PNG row filters;
//        16-bit samples / 65535, 8-bit / 255
//   EXR  single-part scanline files with NONE, RLE, ZIPS or ZIP compression
//        and HALF, FLOAT or UINT channels; R, G, B, A are returned in that
//        order (or the one channel of a single-channel file).  Written
//        uncompressed as RGB 32-bit float scanlines
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace hasty {

struct FloatImage {
  uint32_t width = 0, height = 0, channels = 0;  // channels 1 to 4
  std::vector<float> px;                          // row-major, row 0 = top, interleaved channels
};

// Reads any of the formats above (chosen by the file's signature); throws
// std::runtime_error naming the file on a malformed or unsupported file.
FloatImage readImage(const std::string& path);

// Inflates a zlib stream (RFC 1950 wrapper, RFC 1951 deflate blocks); throws
// once the output grows past maxOut bytes (by at most one match).
std::vector<uint8_t> zlibInflate(const uint8_t* data, size_t size, size_t maxOut = SIZE_MAX);

// OpenEXR scanline file, no compression, three FLOAT channels B, G, R
// (alphabetical channel order, as the format requires), row 0 = top.
void writeEXR(const std::string& path, uint32_t width, uint32_t height, const float* rgbTopRowFirst);

}  // namespace hasty
