// Mip chains of the material image textures (noor/cuda/texture.cuh:204-400).
#include "texture_store.h"

#include <cmath>

namespace noor {

uint32_t mip_levels(uint32_t w, uint32_t h) {
    uint32_t sz = w > h ? w : h, levels = 0;
    while (sz) {
        sz /= 2;
        ++levels;
    }
    return levels;
}

static int wrap(int x, int n) {
    const int r = x % n;
    return r < 0 ? r + n : r;
}

// Same operations, in the same order, as the kernels' tex_level().
void level_bilinear(const float* lvl, uint32_t w, uint32_t h, float s, float q, float out[4]) {
    if (!(std::fabs(s) < 1e30f)) s = 0.0f;
    if (!(std::fabs(q) < 1e30f)) q = 0.0f;
    s = s - std::floor(s);
    q = q - std::floor(q);
    const float x = s * (float)w - 0.5f, y = q * (float)h - 0.5f;
    const float fx = std::floor(x), fy = std::floor(y);
    const float ax = x - fx, ay = y - fy;
    const int x0 = wrap((int)fx, (int)w), x1 = wrap((int)fx + 1, (int)w);
    const int y0 = wrap((int)fy, (int)h), y1 = wrap((int)fy + 1, (int)h);
    const float* t00 = lvl + 4 * ((size_t)y0 * w + x0);
    const float* t10 = lvl + 4 * ((size_t)y0 * w + x1);
    const float* t01 = lvl + 4 * ((size_t)y1 * w + x0);
    const float* t11 = lvl + 4 * ((size_t)y1 * w + x1);
    const float bx = 1.0f - ax, by = 1.0f - ay;
    for (int c = 0; c < 4; ++c)
        out[c] = by * (bx * t00[c] + ax * t10[c]) + ay * (bx * t01[c] + ax * t11[c]);
}

uint32_t store_add_texture(TextureStore& ts, const float* rgba, uint32_t w, uint32_t h) {
    TexLevels d;
    d.offset = (uint32_t)(ts.texels.size() / 4);
    d.w = w;
    d.h = h;
    d.levels = mip_levels(w, h);
    ts.texels.insert(ts.texels.end(), rgba, rgba + 4 * (size_t)w * h);
    size_t src = 4 * (size_t)d.offset;
    uint32_t sw = w, sh = h;
    for (uint32_t l = 1; l < d.levels; ++l) {
        const uint32_t W = sw > 1 ? sw / 2 : 1, H = sh > 1 ? sh / 2 : 1;
        const size_t dst = ts.texels.size();
        ts.texels.resize(dst + 4 * (size_t)W * H);
        const float px = (float)(1.0 / (double)W), py = (float)(1.0 / (double)H);
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) {
                float a[4], b[4], c[4], e[4];
                const float* lvl = ts.texels.data() + src;
                level_bilinear(lvl, sw, sh, (float)(x + 0) * px, (float)(y + 0) * py, a);
                level_bilinear(lvl, sw, sh, (float)(x + 1) * px, (float)(y + 0) * py, b);
                level_bilinear(lvl, sw, sh, (float)(x + 1) * px, (float)(y + 1) * py, c);
                level_bilinear(lvl, sw, sh, (float)(x + 0) * px, (float)(y + 1) * py, e);
                float* o = ts.texels.data() + dst + 4 * ((size_t)y * W + x);
                for (int k = 0; k < 4; ++k) o[k] = (((a[k] + b[k]) + c[k]) + e[k]) / 4.0f;
            }
        src = dst;
        sw = W;
        sh = H;
    }
    ts.desc.push_back(d);
    return (uint32_t)(ts.desc.size() - 1);
}

}  // namespace noor
