/*
 * oracle_scene.c -- scene preparation for the CPU oracle (TEST INFRASTRUCTURE).
 *
 * Restates, on the CPU, the load-time work the reference does on the device:
 * material texture resolution (noor/cuda/materialimp.cuh:134-214), the
 * physical-sky bake (noor/cuda/skydome.cuh:27-37), the environment 2D
 * distribution build (noor/cuda/distribution2d.cuh:63-228), plus the RNG
 * primitives (noor/cuda/rng.cuh:37-76) and the Hosek-Wilkie evaluation
 * (noor/cuda/hosek.cuh:43-104).  All arithmetic is float, in the reference's
 * operation order; transcendentals come from omath.h.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "oracle.h"
#include "omath.h"

#define O_PI 3.141592654f
#define O_2PI (2.0f * O_PI)
#define O_PI_OVER_2 (O_PI / 2.0f)

static float clampf_(float v, float a, float b) { return fmaxf(a, fminf(v, b)); }
static int clampi_(int v, int a, int b) { return v < a ? a : (v > b ? b : v); }

/* ---- RNG --------------------------------------------------------------- */

/* TEA with `rounds` cycles (noor/cuda/rng.cuh:48-61; Zafar, Olano & Curtis,
 * "GPU Random Numbers via the Tiny Encryption Algorithm", HPG 2010). */
uint32_t oracle_tea(uint32_t v0, uint32_t v1, const uint32_t key[4], int rounds, uint32_t* v1_out) {
    uint32_t s0 = 0;
    for (int n = 0; n < rounds; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + key[0]) ^ (v1 + s0) ^ ((v1 >> 5) + key[1]);
        v1 += ((v0 << 4) + key[2]) ^ (v0 + s0) ^ ((v0 >> 5) + key[3]);
    }
    if (v1_out) *v1_out = v1;
    return v0;
}

uint32_t oracle_tea16(uint32_t v0, uint32_t v1) {
    static const uint32_t key[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};
    return oracle_tea(v0, v1, key, 16, NULL);
}

/* PCG32 (O'Neill, pcg-random.org, pcg32_srandom_r / pcg32_random_r). */
uint32_t oracle_pcg32_next(uint64_t* state, uint64_t inc) {
    uint64_t old = *state;
    *state = old * 6364136223846793005ULL + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}

void oracle_pcg32_seed(uint64_t* state, uint64_t* inc, uint64_t initstate, uint64_t initseq) {
    *state = 0u;
    *inc = (initseq << 1u) | 1u;
    oracle_pcg32_next(state, *inc);
    *state += initstate;
    oracle_pcg32_next(state, *inc);
}

/* ---- textures ------------------------------------------------------------ */

/* tex2D with normalized coordinates, linear filtering and clamp addressing
 * (noor/cuda/texture.cuh:107-113, descriptor at 265-278), restated with FP32
 * weights: x = u*w - 0.5, lerp in x then in y. */
float oracle_tex_bilinear(const float* rgba, int w, int h, float u, float v, int c) {
    if (u != u) u = 0.0f;
    if (v != v) v = 0.0f;
    float x = u * (float)w - 0.5f;
    float y = v * (float)h - 0.5f;
    x = clampf_(x, -2.0f, (float)w + 1.0f);
    y = clampf_(y, -2.0f, (float)h + 1.0f);
    float fx0 = floorf(x), fy0 = floorf(y);
    float ax = x - fx0, ay = y - fy0;
    int ix = (int)fx0, iy = (int)fy0;
    int i0 = clampi_(ix, 0, w - 1), i1 = clampi_(ix + 1, 0, w - 1);
    int j0 = clampi_(iy, 0, h - 1), j1 = clampi_(iy + 1, 0, h - 1);
    float t00 = rgba[((size_t)j0 * w + i0) * 4 + c];
    float t10 = rgba[((size_t)j0 * w + i1) * 4 + c];
    float t01 = rgba[((size_t)j1 * w + i0) * 4 + c];
    float t11 = rgba[((size_t)j1 * w + i1) * 4 + c];
    float bx = 1.0f - ax, by = 1.0f - ay;
    float r0 = bx * t00 + ax * t10;
    float r1 = bx * t01 + ax * t11;
    return by * r0 + ay * r1;
}

/* ---- Hosek-Wilkie (noor/cuda/hosek.cuh:55-101) ----------------------------- */

float oracle_hosek_radiance(const noor_hosek_state* st, float theta, float gamma, int channel) {
    const float* c = st->configs[channel];
    const float cos_gamma = om_cosf(gamma);
    const float cos_gamma_2 = cos_gamma * cos_gamma;
    const float cos_theta = om_cosf(theta);
    const float expM = om_expf(c[4] * gamma);
    const float rayM = cos_gamma_2;
    const float mieM = (1.0f + cos_gamma_2) /
        om_powf((1.0f + c[8] * c[8] - 2.0f * c[8] * cos_gamma), 1.5f);
    const float zenith = sqrtf(cos_theta);
    const float r = (1.0f + c[0] * om_expf(c[1] / (cos_theta + 0.01f))) *
        (c[2] + c[3] * expM + c[5] * rayM + c[6] * mieM + c[7] * zenith);
    return r * st->radiances[channel];
}

static float inverse_gamma(float value) {
    /* noor/cuda/utils.cuh:214-217 */
    if (value <= 0.04045f) return value * 1.f / 12.92f;
    return om_powf((value + 0.055f) * 1.f / 1.055f, 2.4f);
}

void oracle_hosek_query(const noor_hosek* hk, const float v[3], float rgb[3]) {
    if (v[1] < 0.f) { rgb[0] = rgb[1] = rgb[2] = 0.0f; return; }
    const float scale = .025f;
    const float theta = clampf_(om_acosf(clampf_(v[1], -1.f, 1.f)), 0.f, O_PI_OVER_2);
    const float d = v[0] * hk->solar_dir.x + v[1] * hk->solar_dir.y + v[2] * hk->solar_dir.z;
    const float gamma = om_acosf(clampf_(d, -1.f, 1.f));
    float c[3];
    c[0] = oracle_hosek_radiance(&hk->state_r, theta, gamma, 0);
    c[1] = oracle_hosek_radiance(&hk->state_g, theta, gamma, 1);
    c[2] = oracle_hosek_radiance(&hk->state_b, theta, gamma, 2);
    if (gamma < hk->solar_radius) {
        rgb[0] = hk->solar_scale * c[0];
        rgb[1] = hk->solar_scale * c[1];
        rgb[2] = hk->solar_scale * c[2];
        return;
    }
    rgb[0] = inverse_gamma(c[0] * scale);
    rgb[1] = inverse_gamma(c[1] * scale);
    rgb[2] = inverse_gamma(c[2] * scale);
}

/* ---- 2D distribution (noor/cuda/distribution2d.cuh) ----------------------- */

/* nearestPow2 (noor/cuda/utils.cuh:106-135), integer form */
static int nearest_pow2(int n) {
    int next = 1;
    while (next < n) next <<= 1;
    int prev = next;
    if (prev > n) prev >>= 1;
    if (prev < 1) prev = 1;
    return (abs(n - next) < abs(n - prev)) ? next : prev;
}

/* prescan: Blelloch work-efficient exclusive scan of func/n, in the tree
 * order of the reference kernel (noor/cuda/distribution2d.cuh:85-141). */
static void prescan_row(float* func, float* cdf, int n, float* temp) {
    for (int i = 0; i < n; ++i) temp[i] = func[i] / (float)n;
    int offset = 1;
    for (int d = n >> 1; d > 0; d >>= 1) {
        for (int tid = 0; tid < d; ++tid) {
            int ai = offset * (2 * tid + 1) - 1;
            int bi = offset * (2 * tid + 2) - 1;
            temp[bi] += temp[ai];
        }
        offset *= 2;
    }
    temp[n - 1] = 0.0f;
    for (int d = 1; d < n; d *= 2) {
        offset >>= 1;
        for (int tid = 0; tid < d; ++tid) {
            int ai = offset * (2 * tid + 1) - 1;
            int bi = offset * (2 * tid + 2) - 1;
            float t = temp[ai];
            temp[ai] = temp[bi];
            temp[bi] += t;
        }
    }
    for (int i = 0; i < n; ++i) cdf[i] = temp[i];
    const float funcInt = cdf[n - 1] + func[n - 1] / (float)n;
    cdf[n] = funcInt;
    func[n] = funcInt;
}

/* process_cdf (noor/cuda/distribution2d.cuh:143-163) for one row */
static void process_row(float* func, float* cdf, int width, int height, int row) {
    const float funcInt = cdf[row * (width + 1) + width];
    for (int col = 0; col < width; ++col) {
        if (funcInt == 0)
            cdf[row * (width + 1) + col + 1] = (float)(col + 1) / (float)width;
        else
            cdf[row * (width + 1) + col + 1] /= funcInt;
    }
    if (height > 1) func[height * (width + 1) + row] = funcInt;
}

int oracle_dist_build(const float* rgba, int tw, int th, orc_dist2d* d) {
    int W = nearest_pow2(tw), H = nearest_pow2(th);
    W = clampi_(W, 32, 1024);
    H = clampi_(H, 32, 1024);
    size_t n = (size_t)H * (W + 2) + 1;
    d->w = W;
    d->h = H;
    d->func = (float*)calloc(n, sizeof(float));
    d->cdf = (float*)calloc(n, sizeof(float));
    float* temp = (float*)malloc(sizeof(float) * (size_t)(W > H ? W : H));
    if (!d->func || !d->cdf || !temp) { free(temp); return -1; }
    /* update_func: 3x3 box of bilinear luminance (distribution2d.cuh:63-83) */
    for (int row = 0; row < H; ++row) {
        for (int col = 0; col < W; ++col) {
            float func = 0.0f;
            const float nbox = 9.0f;
            for (int y = -1; y <= 1; ++y) {
                for (int x = -1; x <= 1; ++x) {
                    float u = (float)(x + col) / (float)W;
                    float v = (float)(y + row) / (float)H;
                    float r = oracle_tex_bilinear(rgba, tw, th, u, v, 0);
                    float g = oracle_tex_bilinear(rgba, tw, th, u, v, 1);
                    float b = oracle_tex_bilinear(rgba, tw, th, u, v, 2);
                    func += r * 0.2126f + g * 0.7152f + b * 0.0722f;
                }
            }
            d->func[row * (W + 1) + col] = func / nbox;
        }
    }
    for (int row = 0; row < H; ++row)
        prescan_row(&d->func[row * (W + 1)], &d->cdf[row * (W + 1)], W, temp);
    for (int row = 0; row < H; ++row)
        process_row(d->func, d->cdf, W, H, row);
    float* mf = &d->func[H * (W + 1)];
    float* mc = &d->cdf[H * (W + 1)];
    prescan_row(mf, mc, H, temp);
    process_row(mf, mc, H, 1, 0);
    free(temp);
    return 0;
}

void oracle_dist_free(orc_dist2d* d) {
    free(d->func);
    free(d->cdf);
    d->func = d->cdf = NULL;
}

const float* oracle_scene_dist_func(const orc_scene* s, int* w, int* h) {
    if (w) *w = s->dist.w;
    if (h) *h = s->dist.h;
    return s->dist.func;
}
const float* oracle_scene_dist_cdf(const orc_scene* s) { return s->dist.cdf; }

/* The physical sky's baked texture (sin(theta) * sky, RGBA), or NULL when the
   scene's sky is not PHYSICAL. */
const float* oracle_scene_sky_bake(const orc_scene* s, int* w, int* h) {
    if (w) *w = (int)s->sky_w;
    if (h) *h = (int)s->sky_h;
    return s->sky_bake;
}

/* ---- scene ----------------------------------------------------------------- */

static void tex_rgb(const noor_payload_view* p, uint32_t idx, float out[3]) {
    out[0] = out[1] = out[2] = 0.0f;
    if (idx >= p->num_textures) return;
    const noor_texture_view* t = &p->textures[idx];
    if (!t->rgba || t->width == 0 || t->height == 0) return;
    /* constant textures are 1x1; larger ones are read at their first texel */
    out[0] = t->rgba[0];
    out[1] = t->rgba[1];
    out[2] = t->rgba[2];
}

/* ---- material texture store (noor/cuda/texture.cuh:204-400) ------------------ */

static uint32_t mip_levels(uint32_t w, uint32_t h) {
    uint32_t sz = w > h ? w : h, levels = 0;
    while (sz) { sz /= 2; ++levels; }   /* CudaMipMap::getMipMapLevels */
    return levels;
}

static int wrapi(int x, int n) {
    const int r = x % n;
    return r < 0 ? r + n : r;
}

/* tex2D, linear filtering, wrap addressing, normalized (s, q) on one level */
static void level_bilinear(const float* lvl, uint32_t w, uint32_t h, float s, float q, float out[4]) {
    if (!(fabsf(s) < 1e30f)) s = 0.0f;
    if (!(fabsf(q) < 1e30f)) q = 0.0f;
    s = s - floorf(s);
    q = q - floorf(q);
    const float x = s * (float)w - 0.5f, y = q * (float)h - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    const int x0 = wrapi((int)fx, (int)w), x1 = wrapi((int)fx + 1, (int)w);
    const int y0 = wrapi((int)fy, (int)h), y1 = wrapi((int)fy + 1, (int)h);
    const float* t00 = lvl + 4 * ((size_t)y0 * w + x0);
    const float* t10 = lvl + 4 * ((size_t)y0 * w + x1);
    const float* t01 = lvl + 4 * ((size_t)y1 * w + x0);
    const float* t11 = lvl + 4 * ((size_t)y1 * w + x1);
    const float bx = 1.0f - ax, by = 1.0f - ay;
    for (int c = 0; c < 4; ++c)
        out[c] = by * (bx * t00[c] + ax * t10[c]) + ay * (bx * t01[c] + ax * t11[c]);
}

/* The power of two nearest n, the smaller on a tie (NOOR::nearestPow2,
 * noor/cuda/utils.cuh:129-135). */
static uint32_t nearest_pow2u(uint32_t n) {
    uint32_t next = 1;
    while (next < n) next <<= 1;
    uint32_t prev = next > n ? next >> 1 : next;
    if (prev < 1) prev = 1;
    return (next - n < n - prev) ? next : prev;
}

/* Lanczos windowed sinc, tau 2 (noor/cuda/image.cuh:41-50), in binary64, rounded */
static float lanczos2(float x) {
    double a = fabs((double)x);
    if (a < 1e-5) return 1.0f;
    if (a > 1.0) return 0.0f;
    a *= 3.14159265358979323846;
    const double s = sin(a / 2.0) / (a / 2.0);
    const double l = sin(a) / a;
    return (float)(s * l);
}

/* ResampleWeight (noor/cuda/image.cuh:52-98): first tap, normalised weights */
static void resample_taps(float center, int* first, float wt[4]) {
    *first = (int)floorf(center - 2.0f + 0.5f);
    float sum = 0.0f;
    for (int j = 0; j < 4; ++j) {
        const float pos = (float)(*first + j) + 0.5f;
        wt[j] = lanczos2((pos - center) / 2.0f);
        sum += wt[j];
    }
    for (int j = 0; j < 4; ++j) wt[j] /= sum;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* resize_kernel (noor/cuda/image.cuh:61-123), clamp addressing.  `sum`: as
 * the reference, the horizontal taps along the source row and the vertical
 * taps along the source column under each output texel (integer centres) are
 * added; otherwise the two passes are composed, x first, with centres
 * (i + 1/2) * old / new. */
static float* lanczos_resize(const float* src, uint32_t w, uint32_t h, uint32_t W, uint32_t H, int sum) {
    float* out = (float*)calloc(4 * (size_t)W * H, sizeof(float));
    const float xs = (float)w / (float)W, ys = (float)h / (float)H;
    if (sum) {
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) {
                const int cx = (int)(((float)x + 0.5f) * xs), cy = (int)(((float)y + 0.5f) * ys);
                int fx, fy;
                float wx[4], wy[4];
                resample_taps((float)cx, &fx, wx);
                resample_taps((float)cy, &fy, wy);
                float* o = out + 4 * ((size_t)y * W + x);
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
    float* tmp = (float*)calloc(4 * (size_t)W * h, sizeof(float));
    for (uint32_t x = 0; x < W; ++x) {
        int fx;
        float wx[4];
        resample_taps(((float)x + 0.5f) * xs, &fx, wx);
        for (uint32_t r = 0; r < h; ++r) {
            float* o = tmp + 4 * ((size_t)r * W + x);
            for (int j = 0; j < 4; ++j) {
                const float* t = src + 4 * ((size_t)r * w + clampi(fx + j, 0, (int)w - 1));
                for (int c = 0; c < 4; ++c) o[c] += wx[j] * t[c];
            }
        }
    }
    for (uint32_t y = 0; y < H; ++y) {
        int fy;
        float wy[4];
        resample_taps(((float)y + 0.5f) * ys, &fy, wy);
        for (uint32_t x = 0; x < W; ++x) {
            float* o = out + 4 * ((size_t)y * W + x);
            for (int j = 0; j < 4; ++j) {
                const float* t = tmp + 4 * ((size_t)clampi(fy + j, 0, (int)h - 1) * W + x);
                for (int c = 0; c < 4; ++c) o[c] += wy[j] * t[c];
            }
        }
    }
    free(tmp);
    return out;
}

/* Appends a texture and its mip chain: each texel of the next level is the
 * mean of the four samples of this level at ((x + i) / W, (y + j) / H),
 * i, j in {0, 1}, (W, H) the next level's size (the mipmap kernel,
 * noor/cuda/texture.cuh:204-221, driven by generateMipMaps, 322-388). */
static uint32_t store_add(orc_scene* s, const noor_texture_view* t0, size_t* cap) {
    /* sides resampled to the nearest powers of two first (CudaTexture::load,
     * noor/cuda/texture.cuh:141-165) */
    noor_texture_view tv = *t0;
    const noor_texture_view* t = &tv;
    float* resized = NULL;
    const uint32_t W = nearest_pow2u(t0->width), H = nearest_pow2u(t0->height);
    if (W != t0->width || H != t0->height) {
        resized = lanczos_resize(t0->rgba, t0->width, t0->height, W, H,
                                 (s->opt.compat & NOOR_COMPAT_RESIZE_SUM) != 0);
        tv.width = W;
        tv.height = H;
        tv.rgba = resized;
    }
    const uint32_t levels = mip_levels(t->width, t->height);
    size_t need = 0;
    uint32_t w = t->width, h = t->height;
    for (uint32_t l = 0; l < levels; ++l) {
        need += (size_t)w * h;
        w = w > 1 ? w / 2 : 1;
        h = h > 1 ? h / 2 : 1;
    }
    orc_texdesc d = {0, t->width, t->height, levels};
    size_t used = 0;
    for (uint32_t i = 0; i < s->num_tex; ++i) {
        size_t n = 0;
        uint32_t a = s->tex_desc[i].w, b = s->tex_desc[i].h;
        for (uint32_t l = 0; l < s->tex_desc[i].levels; ++l) {
            n += (size_t)a * b;
            a = a > 1 ? a / 2 : 1;
            b = b > 1 ? b / 2 : 1;
        }
        used += n;
    }
    if (used + need > *cap) {
        *cap = 2 * (used + need);
        s->tex_texels = (float*)realloc(s->tex_texels, sizeof(float) * 4 * *cap);
    }
    s->tex_desc = (orc_texdesc*)realloc(s->tex_desc, sizeof(orc_texdesc) * (s->num_tex + 1));
    d.offset = (uint32_t)used;
    float* base = s->tex_texels + 4 * used;
    memcpy(base, t->rgba, sizeof(float) * 4 * (size_t)t->width * t->height);
    float* src = base;
    uint32_t sw = t->width, sh = t->height;
    for (uint32_t l = 1; l < levels; ++l) {
        const uint32_t W = sw > 1 ? sw / 2 : 1, H = sh > 1 ? sh / 2 : 1;
        float* dst = src + 4 * (size_t)sw * sh;
        const float px = (float)(1.0 / (double)W), py = (float)(1.0 / (double)H);
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) {
                float a[4], b[4], c[4], e[4];
                level_bilinear(src, sw, sh, (float)(x + 0) * px, (float)(y + 0) * py, a);
                level_bilinear(src, sw, sh, (float)(x + 1) * px, (float)(y + 0) * py, b);
                level_bilinear(src, sw, sh, (float)(x + 1) * px, (float)(y + 1) * py, c);
                level_bilinear(src, sw, sh, (float)(x + 0) * px, (float)(y + 1) * py, e);
                float* o = dst + 4 * ((size_t)y * W + x);
                for (int k = 0; k < 4; ++k) o[k] = (((a[k] + b[k]) + c[k]) + e[k]) / 4.0f;
            }
        src = dst;
        sw = W;
        sh = H;
    }
    s->tex_desc[s->num_tex] = d;
    free(resized);
    return s->num_tex++;
}

int oracle_texture_desc(const orc_scene* s, uint32_t i, orc_texdesc* out) {
    if (i >= s->num_tex) return -1;
    *out = s->tex_desc[i];
    return 0;
}
const float* oracle_texture_texels(const orc_scene* s) { return s->tex_texels; }

static void resolve_material(orc_scene* s, const noor_payload_view* p, const noor_material* m, orc_material* o,
                             int64_t* store_idx, size_t* cap) {
    float tmp[3];
    o->type = m->type;
    const float one[2] = {1.0f, 1.0f};
    const uint32_t tex[ORC_TX_COUNT] = {m->diffuse_tex_idx, m->specular_tex_idx, m->transmission_tex_idx,
                                        m->emittance_tex_idx, m->ior_tex_idx, m->k_tex_idx,
                                        m->roughness_tex_idx, m->bump_tex_idx, m->transparency_tex_idx};
    const float* uvs[ORC_TX_COUNT] = {&m->diffuse_uvscale.x, &m->specular_uvscale.x, one,
                                      &m->emittance_uvscale.x, one, one, &m->roughness_uvscale.x,
                                      &m->bump_uvscale.x, &m->transparency_uvscale.x};
    for (int f = 0; f < ORC_TX_COUNT; ++f) {
        if (f == ORC_TX_BUMP && !(m->type & NOOR_MAT_BUMP)) continue;
        if (f == ORC_TX_ALPHA && !(m->type & NOOR_MAT_ALPHA)) continue;
        if (tex[f] >= p->num_textures) continue;
        const noor_texture_view* t = &p->textures[tex[f]];
        if (!t->rgba || t->width == 0 || t->height == 0) continue;
        if (f < ORC_TX_BUMP && t->width * t->height == 1) continue;
        if (store_idx[tex[f]] < 0) store_idx[tex[f]] = store_add(s, t, cap);
        o->tex_mask |= 1u << f;
        o->tex[f] = (uint32_t)store_idx[tex[f]];
        o->uvs[f][0] = uvs[f][0];
        o->uvs[f][1] = uvs[f][1];
    }
    o->bumpfactor = m->bumpfactor;
    o->emitter_scale = m->emitter_scale;
    if (o->tex_mask) s->textured = 1;
    tex_rgb(p, m->diffuse_tex_idx, o->diffuse);
    tex_rgb(p, m->specular_tex_idx, o->specular);
    tex_rgb(p, m->transmission_tex_idx, o->transmission);
    tex_rgb(p, m->emittance_tex_idx, tmp);
    for (int c = 0; c < 3; ++c) o->emittance[c] = m->emitter_scale * tmp[c];
    tex_rgb(p, m->ior_tex_idx, o->eta);
    tex_rgb(p, m->k_tex_idx, o->k);
    tex_rgb(p, m->roughness_tex_idx, tmp);
    o->rough[0] = tmp[0];
    o->rough[1] = tmp[1];
    o->coat_ior = m->coat_ior;
    o->coat_weight = m->coat_weight;
    o->coat_thickness = m->coat_thickness;
    o->coat_sigma = m->coat_sigma;
}

static int bake_sky(orc_scene* s) {
    const int W = 2048, H = 2048;  /* skydome_res (noor/cuda/path_tracer.cuh:77) */
    s->sky_w = W;
    s->sky_h = H;
    s->sky_bake = (float*)malloc(sizeof(float) * 4 * (size_t)W * H);
    if (!s->sky_bake) return -1;
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            const float u0 = (float)x / (float)W, u1 = (float)y / (float)H;
            const float phi = u0 * O_2PI;
            const float theta = u1 * O_PI;
            const float cosT = om_cosf(theta);
            const float sinT = om_sinf(theta);
            float d[3] = {sinT * om_cosf(phi), cosT, sinT * om_sinf(phi)};
            float l2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            float inv = 1.0f / sqrtf(l2);
            d[0] *= inv; d[1] *= inv; d[2] *= inv;
            float rgb[3];
            oracle_hosek_query(&s->hosek, d, rgb);
            float* o = &s->sky_bake[((size_t)y * W + x) * 4];
            o[0] = sinT * rgb[0];
            o[1] = sinT * rgb[1];
            o[2] = sinT * rgb[2];
            o[3] = sinT * 1.0f;
        }
    }
    return 0;
}

static int build_env(orc_scene* s) {
    oracle_dist_free(&s->dist);
    if (s->spec.skydome_type == NOOR_SKY_PHYSICAL) {
        if (!s->sky_bake && bake_sky(s) != 0) return -1;
        return oracle_dist_build(s->sky_bake, (int)s->sky_w, (int)s->sky_h, &s->dist);
    }
    if (!s->env) return 0;
    return oracle_dist_build(s->env, (int)s->env_w, (int)s->env_h, &s->dist);
}

int oracle_scene_create(const noor_payload_view* p, const noor_hosek* hosek,
                        const noor_camera* cam, const noor_spec* spec,
                        const noor_options* opt, orc_scene** out) {
    orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
    if (!s) return -1;
    s->p = *p;
    if (hosek) s->hosek = *hosek;
    s->cam = *cam;
    s->spec = *spec;
    s->opt = *opt;
    s->mats = (orc_material*)calloc(p->num_materials ? p->num_materials : 1, sizeof(orc_material));
    {
        int64_t* store_idx = (int64_t*)malloc(sizeof(int64_t) * (p->num_textures ? p->num_textures : 1));
        for (uint32_t i = 0; i < p->num_textures; ++i) store_idx[i] = -1;
        size_t cap = 0;
        for (uint32_t i = 0; i < p->num_materials; ++i)
            resolve_material(s, p, &p->materials[i], &s->mats[i], store_idx, &cap);
        free(store_idx);
    }
    if (p->num_textures > 0) {
        /* the environment is a CudaMipMap too (noor/cuda/skydome.cuh:40-52):
         * sides resampled to the nearest power of two */
        const noor_texture_view* t = &p->textures[p->num_textures - 1];
        const uint32_t W = nearest_pow2u(t->width), H = nearest_pow2u(t->height);
        s->env_w = W;
        s->env_h = H;
        s->env = t->rgba;
        if (t->rgba && (W != t->width || H != t->height)) {
            s->env_resized = lanczos_resize(t->rgba, t->width, t->height, W, H,
                                            (opt->compat & NOOR_COMPAT_RESIZE_SUM) != 0);
            s->env = s->env_resized;
        }
    }
    s->num_lights = (int)(p->num_area_lights + p->num_point_lights + p->num_spot_lights +
                          p->num_distant_lights + p->num_infinite_lights);
    s->sky_enabled = (spec->lighting_type & NOOR_LIGHT_SKY_MASK) != 0;
    s->stack_size = 256;
    if (build_env(s) != 0) { oracle_scene_destroy(s); return -1; }
    *out = s;
    return 0;
}

void oracle_scene_destroy(orc_scene* s) {
    if (!s) return;
    oracle_dist_free(&s->dist);
    free(s->sky_bake);
    free(s->mats);
    free(s->tex_texels);
    free(s->tex_desc);
    free(s->env_resized);
    free(s);
}

void oracle_scene_set_camera(orc_scene* s, const noor_camera* cam) { s->cam = *cam; }
void oracle_scene_set_spec(orc_scene* s, const noor_spec* spec) {
    s->spec = *spec;
    s->sky_enabled = (spec->lighting_type & NOOR_LIGHT_SKY_MASK) != 0;
}
void oracle_scene_set_options(orc_scene* s, const noor_options* opt) { s->opt = *opt; }

int oracle_scene_set_env(orc_scene* s, const float* rgba) {
    s->env = rgba;
    return build_env(s);
}

void oracle_scene_set_material_params(orc_scene* s, int mat, const float* diffuse, const float* rough) {
    if (mat < 0 || (uint32_t)mat >= s->p.num_materials) return;
    if (diffuse) memcpy(s->mats[mat].diffuse, diffuse, 3 * sizeof(float));
    if (rough) memcpy(s->mats[mat].rough, rough, 2 * sizeof(float));
}
