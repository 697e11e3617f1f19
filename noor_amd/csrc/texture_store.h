// The material image textures as the kernels read them: every texture with
// its mip chain, RGBA32F, packed into one texel array.
//
// The chain follows CudaMipMap (noor/cuda/texture.cuh:226-400): level l is
// max(1, w >> l) by max(1, h >> l) texels, levels = bit length of max(w, h),
// and each texel of level l + 1 is the mean of four bilinear samples of level
// l taken at the normalized coordinates ((x + i) / W, (y + j) / H), i, j in
// {0, 1}, with (W, H) the size of level l + 1 and wrap addressing
// (noor/cuda/texture.cuh:204-221).  Other sizes are first resampled to the
// nearest power of two per side with a Lanczos filter (radius 2, tau 2,
// clamp addressing; noor/cuda/texture.cuh:141-200, noor/cuda/image.cuh:28-125):
// with `resize_sum` as the reference does it, the sum of a horizontal pass
// along the source row and a vertical pass along the source column under
// each output texel (integer centres); otherwise composed separably, first
// along x then along y, with the centres (i + 1/2) * old / new.
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

// The power of two nearest n, the smaller one on a tie (NOOR::nearestPow2,
// noor/cuda/utils.cuh:129-135).
uint32_t nearest_pow2(uint32_t n);

// Lanczos resampling of a w x h RGBA image to W x H (see above).
std::vector<float> lanczos_resize(const float* rgba, uint32_t w, uint32_t h, uint32_t W, uint32_t H,
                                  bool resize_sum);

// Appends a w x h RGBA texture (resampled to powers of two first when
// needed) and its mip chain; returns its index.
uint32_t store_add_texture(TextureStore& ts, const float* rgba, uint32_t w, uint32_t h, bool resize_sum);

// Bilinear sample of one RGBA level with wrap addressing at normalized
// (s, q): texel centres at half-integers, weights in binary32.
void level_bilinear(const float* lvl, uint32_t w, uint32_t h, float s, float q, float out[4]);

}  // namespace noor

#endif
