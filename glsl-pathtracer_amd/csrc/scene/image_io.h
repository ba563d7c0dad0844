// Image decoding for scene files: PNG (textures) and Radiance HDR
// (environment maps), producing what the reference's texture and env-map
// loaders hand to the scene (src/core/Texture.cpp:40-51 asks stb_image for
// RGBA8; src/core/EnvironmentMap.cpp:63-73 for RGB float): rows top-down, no
// flip, 16-bit samples reduced to their high byte, RGBE mantissa m and
// exponent e decoded as m * 2^(e - 136).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace glslpt {

// 8/16-bit, non-interlaced PNG of any colour type (grey, RGB, palette,
// grey+alpha, RGBA; tRNS honoured) -> RGBA8.
bool LoadPNG(const std::string& path, int& width, int& height, std::vector<uint8_t>& rgba, std::string& err);

// Radiance RGBE (.hdr, "-Y H +X W" orientation, flat or new-style RLE
// scanlines) -> RGB float.
bool LoadHDR(const std::string& path, int& width, int& height, std::vector<float>& rgb, std::string& err);

}  // namespace glslpt
