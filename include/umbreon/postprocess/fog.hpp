// OpenGL linear fog over view-z (libumbreon postprocess/fog.hpp). render()
// applies it on the GPU, fused into the pixel resolve; these host versions
// serve callers that fog a frame themselves.
#pragma once

#include "../scene.hpp"

namespace umbreon {

// f = clamp((end - z) / (end - start), 0, 1); 1 when fog is disabled.
float fogFactor(const Fog& fog, float z);
// Mix RGB toward fog.color by (1 - f) where viewZ > 0 (transparent background:
// scale alpha by f instead).
void applyFog(const Fog& fog, int width, int height, int channels, float* color, const float* viewZ,
              bool transparentBackground);

}  // namespace umbreon
