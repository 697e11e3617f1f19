/* FP32 instantiation of the oracle integrator (TEST INFRASTRUCTURE). */
#include "oracle.h"
#include "omath.h"
#define REAL float
#define REAL_IS_FLOAT 1
#define SFX(name) name##_f32
#define R_SIN om_sinf
#define R_COS om_cosf
#define R_ACOS om_acosf
#define R_ATAN2 om_atan2f
#define R_EXP om_expf
#define R_LOG om_logf
#define R_POW om_powf
#define R_SQRT sqrtf
#define R_FLOOR floorf
#define R_FABS fabsf
#define R_FMAX fmaxf
#define R_FMIN fminf
#include "oracle_core.inc"

/* ---- unit-test hooks into the FP32 integrator -------------------------------- */

/* n floats of the per-pixel stream (gid, frame) */
void oracle_rng_sequence(int kind, uint32_t gid, uint32_t frame, int n, float* out) {
    rng_t r;
    rng_init(&r, kind, gid, frame);
    for (int i = 0; i < n; ++i) out[i] = rng_next(&r);
}

static orc_material mat_from(const float* p) {
    /* p: diffuse[3] specular[3] eta[3] k[3] rough[2] coat_ior transmission[3]
     *    coat_weight coat_thickness coat_sigma */
    orc_material m;
    memset(&m, 0, sizeof(m));
    memcpy(m.diffuse, p, 12);
    memcpy(m.specular, p + 3, 12);
    memcpy(m.eta, p + 6, 12);
    memcpy(m.k, p + 9, 12);
    m.rough[0] = p[12];
    m.rough[1] = p[13];
    m.coat_ior = p[14];
    memcpy(m.transmission, p + 15, 12);
    m.coat_weight = p[18];
    m.coat_thickness = p[19];
    m.coat_sigma = p[20];
    return m;
}

/* BSDF of material type `mtype` in a local frame (n = +z, gn = +z):
 * f(wo, wi) and pdf(wo, wi) over all non-specular lobes. */
void oracle_bsdf_eval(uint32_t mtype, const float* params, const float* wo, const float* wi, float* f, float* pdf) {
    orc_material m = mat_from(params);
    m.type = mtype;
    bsdf_t b;
    make_bsdf(mtype, &b);
    isect_t I;
    memset(&I, 0, sizeof(I));
    I.sn = mk3(0, 0, 1); I.gn = mk3(0, 0, 1); I.ss = mk3(1, 0, 0); I.ts = mk3(0, 1, 0);
    const v3 o = mk3(wo[0], wo[1], wo[2]), i = mk3(wi[0], wi[1], wi[2]);
    const v3 v = bsdf_eval(&b, &I, &m, o, i, B_ALL & ~B_SPECULAR);
    f[0] = v.x; f[1] = v.y; f[2] = v.z;
    *pdf = bsdf_pdf(&b, &I, &m, o, i, B_ALL & ~B_SPECULAR);
}

/* One BSDF sample with the stream (gid, frame); returns the sampled flags. */
int oracle_bsdf_sample(uint32_t mtype, const float* params, const float* wo, uint32_t gid, uint32_t frame,
                       float* wi, float* f, float* pdf) {
    orc_material m = mat_from(params);
    m.type = mtype;
    bsdf_t b;
    make_bsdf(mtype, &b);
    isect_t I;
    memset(&I, 0, sizeof(I));
    I.sn = mk3(0, 0, 1); I.gn = mk3(0, 0, 1); I.ss = mk3(1, 0, 0); I.ts = mk3(0, 1, 0);
    tctx c;
    memset(&c, 0, sizeof(c));
    rng_init(&c.rng, NOOR_RNG_PCG32, gid, frame);
    v3 w = mk3(0, 0, 0), fs;
    real p = 0;
    int sampled = 0;
    const v3 v = bsdf_sample(&b, &I, &m, &m, mk3(wo[0], wo[1], wo[2]), &w, &p, B_ALL, &sampled, &fs, &c);
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    f[0] = v.x; f[1] = v.y; f[2] = v.z;
    *pdf = p;
    return sampled;
}

float oracle_fresnel_dielectric(float cosI, float etaI, float etaT) { return fr_dielectric(cosI, etaI, etaT); }
void oracle_fresnel_conductor(float cosI, const float* eta, const float* k, float* out) {
    const v3 r = fr_conductor(cosI, mk3(1, 1, 1), ld3(eta), ld3(k));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
float oracle_ggx_D(float ax, float ay, const float* wh) {
    ggx_t g = {ax, ay};
    return ggx_D(g, mk3(wh[0], wh[1], wh[2]));
}

/* Environment importance sample with (u1, u2): direction and solid-angle pdf,
 * plus the pdf the light's pdf_Li reports for that direction. */
void oracle_env_sample(orc_scene* s, float u1, float u2, float* dir, float* pdf, float* pdf_li) {
    real map_pdf = 1;
    v2 u = {u1, u2};
    u = d_sample2d(s, u, &map_pdf);
    const real phi = u.x * K_2PI, theta = u.y * K_PI;
    const real cosT = R_COS(theta), sinT = R_SIN(theta);
    const v3 d = nrm3(mk3(sinT * R_COS(phi), cosT, sinT * R_SIN(phi)));
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    *pdf = (sinT == 0) ? 0.0f : map_pdf / ((real)2 * K_PI * K_PI * sinT);
    isect_t I;
    memset(&I, 0, sizeof(I));
    *pdf_li = light_pdf_Li(s, &I, d, s->num_lights - 1);
}

/* One sample of light idx seen from point p with normal n (PCG32 stream gid,
 * frame): the unscaled Li, the direction and distance to the sampled point,
 * the sample's pdf and pdf_Li of the same direction. */
void oracle_light_sample(orc_scene* s, int idx, const float* p, const float* n, uint32_t gid, uint32_t frame,
                         float* Li, float* wi, float* dist, float* pdf, float* pdf_li) {
    isect_t I;
    memset(&I, 0, sizeof(I));
    I.p = mk3(p[0], p[1], p[2]);
    I.gn = I.sn = mk3(n[0], n[1], n[2]);
    tctx c;
    memset(&c, 0, sizeof(c));
    rng_init(&c.rng, NOOR_RNG_PCG32, gid, frame);
    vis_t v;
    memset(&v, 0, sizeof(v));
    real q = 0;
    const v3 L = light_sample_one(s, s, &I, idx, &v, &q, &c);
    Li[0] = L.x; Li[1] = L.y; Li[2] = L.z;
    wi[0] = v.wi.x; wi[1] = v.wi.y; wi[2] = v.wi.z;
    *dist = v.dist;
    *pdf = q;
    *pdf_li = (q > 0) ? light_pdf_Li(s, &I, v.wi, idx) : 0.0f;
}

/* exported portable math for the accuracy tests */
float oracle_sinf(float x) { return om_sinf(x); }
float oracle_cosf(float x) { return om_cosf(x); }
float oracle_atan2f(float y, float x) { return om_atan2f(y, x); }
float oracle_acosf(float x) { return om_acosf(x); }
float oracle_expf(float x) { return om_expf(x); }
float oracle_logf(float x) { return om_logf(x); }
float oracle_powf(float x, float y) { return om_powf(x, y); }
