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

uint32_t nearest_pow2(uint32_t n) {
    uint32_t next = 1;
    while (next < n) next <<= 1;
    uint32_t prev = next > n ? next >> 1 : next;
    if (prev < 1) prev = 1;
    return (next - n < n - prev) ? next : prev;
}

// Lanczos windowed sinc with tau 2 (noor/cuda/image.cuh:41-50), evaluated in
// binary64 and rounded.
static float lanczos(float x) {
    double a = std::fabs((double)x);
    if (a < 1e-5) return 1.0f;
    if (a > 1.0) return 0.0f;
    a *= 3.14159265358979323846;
    const double s = std::sin(a / 2.0) / (a / 2.0);
    const double l = std::sin(a) / a;
    return (float)(s * l);
}

// The four taps of one output texel: first source texel and normalised
// weights (ResampleWeight, noor/cuda/image.cuh:52-98).
static void taps(float center, int& first, float wt[4]) {
    first = (int)std::floor(center - 2.0f + 0.5f);
    float sum = 0.0f;
    for (int j = 0; j < 4; ++j) {
        const float pos = (float)(first + j) + 0.5f;
        wt[j] = lanczos((pos - center) / 2.0f);
        sum += wt[j];
    }
    for (int j = 0; j < 4; ++j) wt[j] /= sum;
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

std::vector<float> lanczos_resize(const float* src, uint32_t w, uint32_t h, uint32_t W, uint32_t H,
                                  bool resize_sum) {
    std::vector<float> out(4 * (size_t)W * H, 0.0f);
    const float xs = (float)w / (float)W, ys = (float)h / (float)H;
    if (resize_sum) {
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) {
                const int cx = (int)(((float)x + 0.5f) * xs), cy = (int)(((float)y + 0.5f) * ys);
                int fx, fy;
                float wx[4], wy[4];
                taps((float)cx, fx, wx);
                taps((float)cy, fy, wy);
                float* o = out.data() + 4 * ((size_t)y * W + x);
                for (int j = 0; j < 4; ++j) {
                    const float* t = src + 4 * ((size_t)cy * w + clampi(fx + j, 0, (int)w - 1));
                    for (int c = 0; c < 4; ++c) o[c] += wx[j] * t[c];
                }
                for (int j = 0; j < 4; ++j) {
                    const float* t = src + 4 * ((size_t)clampi(fy + j, 0, (int)h - 1) * w + cx);
                    for (int c = 0; c < 4; ++c) o[c] += wy[j] * t[c];
                }
            }
        return out;
    }
    std::vector<float> tmp(4 * (size_t)W * h, 0.0f);
    for (uint32_t x = 0; x < W; ++x) {
        int fx;
        float wx[4];
        taps(((float)x + 0.5f) * xs, fx, wx);
        for (uint32_t r = 0; r < h; ++r) {
            float* o = tmp.data() + 4 * ((size_t)r * W + x);
            for (int j = 0; j < 4; ++j) {
                const float* t = src + 4 * ((size_t)r * w + clampi(fx + j, 0, (int)w - 1));
                for (int c = 0; c < 4; ++c) o[c] += wx[j] * t[c];
            }
        }
    }
    for (uint32_t y = 0; y < H; ++y) {
        int fy;
        float wy[4];
        taps(((float)y + 0.5f) * ys, fy, wy);
        for (uint32_t x = 0; x < W; ++x) {
            float* o = out.data() + 4 * ((size_t)y * W + x);
            for (int j = 0; j < 4; ++j) {
                const float* t = tmp.data() + 4 * ((size_t)clampi(fy + j, 0, (int)h - 1) * W + x);
                for (int c = 0; c < 4; ++c) o[c] += wy[j] * t[c];
            }
        }
    }
    return out;
}

uint32_t store_add_texture(TextureStore& ts, const float* rgba, uint32_t w, uint32_t h, bool resize_sum) {
    const uint32_t W = nearest_pow2(w), H = nearest_pow2(h);
    std::vector<float> resized;
    if (W != w || H != h) {
        resized = lanczos_resize(rgba, w, h, W, H, resize_sum);
        rgba = resized.data();
        w = W;
        h = H;
    }
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
