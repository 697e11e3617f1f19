// The material image textures as the kernels read them: every texture with
// its mip chain, RGBA32F, packed into one texel array.
//
// The chain follows CudaMipMap (noor/cuda/texture.cuh:226-400): level l is
// max(1, w >> l) by max(1, h >> l) texels, levels = bit length of max(w, h),
// and each texel of level l + 1 is the mean of four bilinear samples of level
// l taken at the normalized coordinates ((x + i) / W, (y + j) / H), i, j in
// {0, 1}, with (W, H) the size of level l + 1 and wrap addressing
// (noor/cuda/texture.cuh:204-221).  Sizes must be powers of two: the
// reference resamples other sizes to the nearest power of two first
// (noor/cuda/texture.cuh:141-200), which is not reproduced.
#ifndef NOOR_TEXTURE_STORE_H
#define NOOR_TEXTURE_STORE_H

#include <stdint.h>

#include <vector>

namespace noor {

struct TexLevels {
    uint32_t offset, w, h, levels;  // texel offset of level 0; the levels follow it
};

struct TextureStore {
    std::vector<float> texels;  // RGBA
    std::vector<TexLevels> desc;
};

// Bit length of max(w, h) (CudaMipMap::getMipMapLevels).
uint32_t mip_levels(uint32_t w, uint32_t h);

// Appends a w x h RGBA texture and its mip chain; returns its index.
uint32_t store_add_texture(TextureStore& ts, const float* rgba, uint32_t w, uint32_t h);

// Bilinear sample of one RGBA level with wrap addressing at normalized
// (s, q): texel centres at half-integers, weights in binary32.
void level_bilinear(const float* lvl, uint32_t w, uint32_t h, float s, float q, float out[4]);

}  // namespace noor

#endif
