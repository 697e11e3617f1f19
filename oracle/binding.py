"""ctypes binding of oracle/liboracle.so (TEST INFRASTRUCTURE).

Used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the reference results of parity checks or as the CPU baseline.  Never
imported by noor_amd/.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from noor_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


class Counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("extension_rays", "shadow_rays", "node_visits", "tri_tests",
                                          "tri_second_fetch", "instance_entries", "shd_node_visits",
                                          "shd_tri_tests", "shd_tri_second_fetch", "shd_instance_entries")]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = C.CDLL(LIB)
    P = C.POINTER
    vp = C.c_void_p
    L.oracle_scene_create.argtypes = [P(abi.PayloadView), P(abi.Hosek), P(abi.Camera), P(abi.Spec), P(abi.Options),
                                      P(vp)]
    L.oracle_scene_destroy.argtypes = [vp]
    L.oracle_scene_set_camera.argtypes = [vp, P(abi.Camera)]
    L.oracle_scene_set_spec.argtypes = [vp, P(abi.Spec)]
    L.oracle_scene_set_options.argtypes = [vp, P(abi.Options)]
    L.oracle_scene_set_env.argtypes = [vp, P(C.c_float)]
    L.oracle_scene_set_material_params.argtypes = [vp, C.c_int, P(C.c_float), P(C.c_float)]
    L.oracle_scene_dist_func.restype = P(C.c_float)
    L.oracle_scene_dist_func.argtypes = [vp, P(C.c_int), P(C.c_int)]
    L.oracle_scene_dist_cdf.restype = P(C.c_float)
    L.oracle_scene_dist_cdf.argtypes = [vp]
    L.oracle_scene_sky_bake.restype = P(C.c_float)
    L.oracle_scene_sky_bake.argtypes = [vp, P(C.c_int), P(C.c_int)]
    for sfx, real in (("f32", C.c_float), ("f64", C.c_double)):
        f = getattr(L, f"oracle_render_{sfx}")
        f.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P(real),
                      P(Counters)]
        f = getattr(L, f"oracle_render_sig_{sfx}")
        f.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P(real),
                      P(Counters), P(C.c_uint64)]
        f = getattr(L, f"oracle_trace_{sfx}")
        f.argtypes = [vp, P(abi.Ray), C.c_uint32, C.c_int, P(abi.Hit), P(Counters)]
    L.oracle_trace_brute_f32.argtypes = [vp, P(abi.Ray), C.c_uint32, P(abi.Hit)]
    L.oracle_camera_ray_f32.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, P(abi.Ray)]
    L.oracle_render_detached_f64.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_int, P(C.c_double)]
    L.oracle_texture_desc.argtypes = [vp, C.c_uint32, P(C.c_uint32 * 4)]
    L.oracle_texture_texels.restype = P(C.c_float)
    L.oracle_texture_texels.argtypes = [vp]
    L.oracle_tex_lookup_f32.argtypes = [vp, C.c_uint32, C.c_int, C.c_float, C.c_float, P(C.c_float), P(C.c_float)]
    L.oracle_camera_ray_diff_f32.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_float)]
    L.oracle_hit_diff_f32.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_float)]
    L.oracle_tea16.restype = C.c_uint32
    L.oracle_tea16.argtypes = [C.c_uint32, C.c_uint32]
    L.oracle_tea.restype = C.c_uint32
    L.oracle_tea.argtypes = [C.c_uint32, C.c_uint32, P(C.c_uint32), C.c_int, P(C.c_uint32)]
    L.oracle_pcg32_seed.argtypes = [P(C.c_uint64), P(C.c_uint64), C.c_uint64, C.c_uint64]
    L.oracle_pcg32_next.restype = C.c_uint32
    L.oracle_pcg32_next.argtypes = [P(C.c_uint64), C.c_uint64]
    L.oracle_hosek_radiance.restype = C.c_float
    L.oracle_hosek_radiance.argtypes = [P(abi.HosekState), C.c_float, C.c_float, C.c_int]
    L.oracle_hosek_query.argtypes = [P(abi.Hosek), P(C.c_float), P(C.c_float)]
    L.oracle_tex_bilinear.restype = C.c_float
    L.oracle_tex_bilinear.argtypes = [P(C.c_float), C.c_int, C.c_int, C.c_float, C.c_float, C.c_int]
    for name in ("oracle_sinf", "oracle_cosf", "oracle_acosf", "oracle_expf", "oracle_logf"):
        f = getattr(L, name)
        f.restype = C.c_float
        f.argtypes = [C.c_float]
    for name in ("oracle_atan2f", "oracle_powf"):
        f = getattr(L, name)
        f.restype = C.c_float
        f.argtypes = [C.c_float, C.c_float]
    _lib = L
    return L


class OracleScene:
    """The oracle's view of a workload (payload + camera + spec + options + sky)."""

    def __init__(self, payload, camera, spec, options, hosek=None):
        self.payload = payload  # keep alive
        self._h = C.c_void_p()
        hk = hosek if hosek is not None else abi.Hosek()
        self._hosek = hk
        r = lib().oracle_scene_create(C.byref(payload.view), C.byref(hk), C.byref(camera), C.byref(spec),
                                      C.byref(options), C.byref(self._h))
        if r != 0:
            raise RuntimeError("oracle_scene_create failed")
        self.camera = camera

    @classmethod
    def from_workload(cls, wl):
        return cls(wl.payload, wl.camera, wl.spec, wl.options, wl.hosek)

    def set_camera(self, cam):
        self.camera = cam
        lib().oracle_scene_set_camera(self._h, C.byref(cam))

    def set_spec(self, spec):
        lib().oracle_scene_set_spec(self._h, C.byref(spec))

    def set_options(self, opt):
        lib().oracle_scene_set_options(self._h, C.byref(opt))

    def set_env(self, rgba):
        self._env = np.ascontiguousarray(rgba, np.float32)
        if lib().oracle_scene_set_env(self._h, self._env.ctypes.data_as(C.POINTER(C.c_float))) != 0:
            raise RuntimeError("oracle_scene_set_env failed")

    def set_material(self, m, diffuse=None, rough=None):
        d = None if diffuse is None else (C.c_float * 3)(*map(float, diffuse))
        r = None if rough is None else (C.c_float * 2)(*map(float, rough))
        lib().oracle_scene_set_material_params(self._h, m, d, r)

    def render(self, frame0=1, spp=1, threads=0, rect=None, fp64=False, film=None, signature=None):
        """Film (H, W, 4) after spp frames starting at frame0, plus counters.
        A film passed in is continued (its running mean is updated in place).
        `signature`: a uint64 (H, W) array that receives each pixel's path
        signature (the hash of the discrete decisions of its last sample)."""
        if threads <= 0:
            threads = os.cpu_count() or 1
        x0, y0, x1, y1 = rect if rect else (0, 0, self.camera.w, self.camera.h)
        w, h = x1 - x0, y1 - y0
        if film is None:
            film = np.zeros((h, w, 4), np.float64 if fp64 else np.float32)
        assert film.shape == (h, w, 4) and film.flags.c_contiguous
        cnt = Counters()
        if fp64:
            f = lib().oracle_render_f64
            ptr = film.ctypes.data_as(C.POINTER(C.c_double))
        else:
            f = lib().oracle_render_f32
            ptr = film.ctypes.data_as(C.POINTER(C.c_float))
        if signature is not None:
            assert signature.shape == (h, w) and signature.dtype == np.uint64 and signature.flags.c_contiguous
            f = lib().oracle_render_sig_f64 if fp64 else lib().oracle_render_sig_f32
            r = f(self._h, frame0, spp, threads, x0, y0, x1, y1, ptr, C.byref(cnt),
                  signature.ctypes.data_as(C.POINTER(C.c_uint64)))
        else:
            r = f(self._h, frame0, spp, threads, x0, y0, x1, y1, ptr, C.byref(cnt))
        if r != 0:
            raise RuntimeError(f"oracle render failed ({r})")
        return film, cnt

    def trace(self, rays, any_hit=False, fp64=False, brute=False):
        n = len(rays)
        hits = (abi.Hit * n)()
        if brute:
            lib().oracle_trace_brute_f32(self._h, rays, n, hits)
            return hits, None
        cnt = Counters()
        f = lib().oracle_trace_f64 if fp64 else lib().oracle_trace_f32
        if f(self._h, rays, n, 1 if any_hit else 0, hits, C.byref(cnt)) != 0:
            raise RuntimeError("oracle trace failed")
        return hits, cnt

    def camera_ray(self, x, y, frame):
        r = abi.Ray()
        lib().oracle_camera_ray_f32(self._h, x, y, frame, C.byref(r))
        return r

    def camera_ray_diff(self, x, y, frame):
        """(o, d, ox, oy, dx, dy) of the camera ray with its differentials."""
        out = np.zeros(18, np.float32)
        lib().oracle_camera_ray_diff_f32(self._h, x, y, frame, _fp(out))
        return out.reshape(6, 3)

    def hit_diff(self, x, y, frame):
        """First hit of the camera ray with its differentials as a dict, or None."""
        out = np.zeros(18, np.float32)
        if not lib().oracle_hit_diff_f32(self._h, x, y, frame, _fp(out)):
            return None
        return dict(uv=out[0:2], g=out[2:6], p=out[6:9], dpdx=out[9:12], dpdy=out[12:15], gn=out[15:18])

    def texture_levels(self, i):
        """The mip chain of store texture i: a list of (h, w, 4) arrays."""
        d = (C.c_uint32 * 4)()
        if lib().oracle_texture_desc(self._h, i, C.byref(d)) != 0:
            raise IndexError(i)
        off, w, h, levels = list(d)
        n, ww, hh = 0, w, h
        for _ in range(levels):
            n += ww * hh
            ww, hh = max(1, ww // 2), max(1, hh // 2)
        flat = np.ctypeslib.as_array(lib().oracle_texture_texels(self._h), shape=(4 * (off + n),))
        out, o, ww, hh = [], 4 * off, w, h
        for _ in range(levels):
            out.append(flat[o:o + 4 * ww * hh].reshape(hh, ww, 4).copy())
            o += 4 * ww * hh
            ww, hh = max(1, ww // 2), max(1, hh // 2)
        return out

    def tex_lookup(self, i, kind, u, v, g=(0.0, 0.0, 0.0, 0.0)):
        """kind 0: tex2D (level 0), 1: tex2DGrad with g = (dudx, dvdx, dudy, dvdy), 2: point."""
        gg = np.asarray(g, np.float32)
        out = np.zeros(4, np.float32)
        lib().oracle_tex_lookup_f32(self._h, i, kind, float(u), float(v), _fp(gg), _fp(out))
        return out

    def dist(self):
        w = C.c_int()
        h = C.c_int()
        fp = lib().oracle_scene_dist_func(self._h, C.byref(w), C.byref(h))
        n = h.value * (w.value + 2) + 1
        func = np.ctypeslib.as_array(fp, shape=(n,)).copy()
        cdf = np.ctypeslib.as_array(lib().oracle_scene_dist_cdf(self._h), shape=(n,)).copy()
        return w.value, h.value, func, cdf

    def sky_bake(self):
        """The baked physical sky (H, W, 4): sin(theta) * sky at texel corners."""
        w = C.c_int()
        h = C.c_int()
        p = lib().oracle_scene_sky_bake(self._h, C.byref(w), C.byref(h))
        if not p:
            return None
        return np.ctypeslib.as_array(p, shape=(h.value, w.value, 4)).copy()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().oracle_scene_destroy(h)
            except Exception:
                pass
            self._h = None


def render_detached_f64(scene_eval, scene_frozen, frame0=1, spp=1, threads=0):
    """FP64 render sampling with `scene_frozen`, evaluating with `scene_eval`."""
    if threads <= 0:
        threads = os.cpu_count() or 1
    cam = scene_eval.camera
    film = np.zeros((cam.h, cam.w, 4), np.float64)
    r = lib().oracle_render_detached_f64(scene_eval._h, scene_frozen._h, frame0, spp, threads,
                                         film.ctypes.data_as(C.POINTER(C.c_double)))
    if r != 0:
        raise RuntimeError("detached render failed")
    return film


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def rng_sequence(kind, gid, frame, n):
    out = np.zeros(n, np.float32)
    L = lib()
    L.oracle_rng_sequence.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_float)]
    L.oracle_rng_sequence(kind, gid, frame, n, _fp(out))
    return out


def material_params(diffuse=(0.5, 0.5, 0.5), specular=(1, 1, 1), eta=(1, 1, 1), k=(0, 0, 0), rough=(0.3, 0.3),
                    coat_ior=1.5, transmission=(1, 1, 1), coat_weight=0.5, coat_thickness=0.0, coat_sigma=0.0):
    return np.array([*diffuse, *specular, *eta, *k, *rough, coat_ior, *transmission, coat_weight, coat_thickness,
                     coat_sigma], np.float32)


def bsdf_eval(mtype, params, wo, wi):
    L = lib()
    L.oracle_bsdf_eval.argtypes = [C.c_uint32] + [C.POINTER(C.c_float)] * 5
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    wo = np.ascontiguousarray(wo, np.float32)
    wi = np.ascontiguousarray(wi, np.float32)
    L.oracle_bsdf_eval(mtype, _fp(params), _fp(wo), _fp(wi), _fp(f), _fp(pdf))
    return f, float(pdf[0])


def bsdf_sample(mtype, params, wo, gid, frame=1):
    L = lib()
    L.oracle_bsdf_sample.restype = C.c_int
    L.oracle_bsdf_sample.argtypes = [C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32,
                                     C.c_uint32] + [C.POINTER(C.c_float)] * 3
    wi = np.zeros(3, np.float32)
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    wo = np.ascontiguousarray(wo, np.float32)
    flags = L.oracle_bsdf_sample(mtype, _fp(params), _fp(wo), gid, frame, _fp(wi), _fp(f), _fp(pdf))
    return wi, f, float(pdf[0]), flags


def fresnel_dielectric(cos_i, eta_i, eta_t):
    L = lib()
    L.oracle_fresnel_dielectric.restype = C.c_float
    L.oracle_fresnel_dielectric.argtypes = [C.c_float] * 3
    return L.oracle_fresnel_dielectric(cos_i, eta_i, eta_t)


def fresnel_conductor(cos_i, eta, k):
    L = lib()
    L.oracle_fresnel_conductor.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                           C.POINTER(C.c_float)]
    out = np.zeros(3, np.float32)
    e = np.asarray(eta, np.float32)
    kk = np.asarray(k, np.float32)
    L.oracle_fresnel_conductor(cos_i, _fp(e), _fp(kk), _fp(out))
    return out


def ggx_D(ax, ay, wh):
    L = lib()
    L.oracle_ggx_D.restype = C.c_float
    L.oracle_ggx_D.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float)]
    w = np.ascontiguousarray(wh, np.float32)
    return L.oracle_ggx_D(ax, ay, _fp(w))


def env_sample(scene, u1, u2):
    L = lib()
    L.oracle_env_sample.argtypes = [C.c_void_p, C.c_float, C.c_float] + [C.POINTER(C.c_float)] * 3
    d = np.zeros(3, np.float32)
    p = np.zeros(1, np.float32)
    q = np.zeros(1, np.float32)
    L.oracle_env_sample(scene._h, u1, u2, _fp(d), _fp(p), _fp(q))
    return d, float(p[0]), float(q[0])


def light_sample(scene, idx, p, n, gid, frame=1):
    """One sample of light `idx` (area, point, spot, distant, infinite in that
    order) seen from point p with normal n: (Li without the light-count
    factor, unit direction, distance, pdf, pdf_Li of that direction)."""
    L = lib()
    f = C.POINTER(C.c_float)
    L.oracle_light_sample.argtypes = [C.c_void_p, C.c_int, f, f, C.c_uint32, C.c_uint32, f, f, f, f, f]
    pp = np.ascontiguousarray(p, np.float32)
    nn = np.ascontiguousarray(n, np.float32)
    Li = np.zeros(3, np.float32)
    wi = np.zeros(3, np.float32)
    d = np.zeros(1, np.float32)
    q = np.zeros(1, np.float32)
    ql = np.zeros(1, np.float32)
    L.oracle_light_sample(scene._h, idx, _fp(pp), _fp(nn), gid, frame, _fp(Li), _fp(wi), _fp(d), _fp(q), _fp(ql))
    return Li, wi, float(d[0]), float(q[0]), float(ql[0])
