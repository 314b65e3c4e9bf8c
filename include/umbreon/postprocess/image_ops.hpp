// Image helpers of the public API (source-compatible with libumbreon's
// postprocess/image_ops.hpp). These operate on host buffers a caller already
// holds; inside render() the same resolve runs on the GPU.
#pragma once

#include <cstdint>
#include <vector>

#include "../render/render_types.hpp"

namespace umbreon {

// ss x ss box average (non-finite samples contribute 0).
std::vector<float> boxDownsample(const std::vector<float>& src, int w, int h, int channels, int ss);
// RGB ^ g in place (g == 1 is a no-op; non-positive values become 0).
void applyAssumedGamma(FrameResult& frame, float g);
// Linear RGBA -> 8-bit sRGB (channels 3 or 4; alpha stored linear).
std::vector<std::uint8_t> srgbEncode8(const FrameResult& frame, int channels);
float srgbEncodeF(float v);
float srgbDecodeF(float s);

}  // namespace umbreon
