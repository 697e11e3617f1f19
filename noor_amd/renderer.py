"""Python handle on a render context of the C ABI (include/noor_hip.h).

Mirrors the reference's render calls: Scene::initCudaContext -> noor_create,
Scene::path_tracer -> noor_render, Camera::updateCudaCamera ->
noor_update_camera, and so on (noor/source/scene.cpp:124-160).  There is no
CPU fallback: without the built library or a GPU the constructor raises.
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import check, lib


class Renderer:
    def __init__(self, payload, camera, spec, options=None, hosek=None):
        self.payload = payload
        self._h = C.c_void_p()
        self.camera = camera
        opt = options if options is not None else abi.Options(abi.RNG_PCG32, abi.TRI_WATERTIGHT, abi.COMPAT_ALL, 0)
        self.options = opt
        hk = C.byref(hosek) if hosek is not None else None
        st = lib().noor_create(C.byref(payload.view), hk, C.byref(camera), C.byref(spec), C.byref(opt),
                               C.byref(self._h))
        if st != 0:
            raise abi.NoorError(st, "noor_create")
        self.frame = 1

    @classmethod
    def from_workload(cls, wl):
        return cls(wl.payload, wl.camera, wl.spec, wl.options, wl.hosek)

    def _check(self, st):
        check(st, self._h)

    def render(self, spp=1, frame=None):
        f = C.c_uint32(self.frame if frame is None else frame)
        self._check(lib().noor_render(self._h, C.byref(f), spp))
        self.frame = f.value

    def render_wavefront(self, spp=1, frame=None):
        f = C.c_uint32(self.frame if frame is None else frame)
        self._check(lib().noor_render_wavefront(self._h, C.byref(f), spp))
        self.frame = f.value

    def reset(self):
        self.frame = 1

    def film(self):
        out = np.empty((self.camera.h, self.camera.w, 4), np.float32)
        self._check(lib().noor_read_film(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), out.size))
        return out

    def film_device_ptr(self):
        p = C.c_void_p()
        self._check(lib().noor_film_device_ptr(self._h, C.byref(p)))
        return p.value

    def update_camera(self, cam):
        self._check(lib().noor_update_camera(self._h, C.byref(cam)))
        self.camera = cam
        self.frame = 1

    def update_spec(self, spec):
        self._check(lib().noor_update_spec(self._h, C.byref(spec)))
        self.frame = 1

    def update_hosek(self, hk):
        self._check(lib().noor_update_hosek(self._h, C.byref(hk)))
        self.frame = 1

    def update_options(self, opt):
        self._check(lib().noor_update_options(self._h, C.byref(opt)))
        self.options = opt

    def lookat(self):
        v = abi.float4()
        self._check(lib().noor_get_lookat(self._h, C.byref(v)))
        return np.array([v.x, v.y, v.z, v.w], np.float32)

    def stats(self):
        s = abi.Stats()
        self._check(lib().noor_get_stats(self._h, C.byref(s)))
        return {n: getattr(s, n) for n, _ in abi.Stats._fields_}

    def reset_stats(self):
        self._check(lib().noor_reset_stats(self._h))

    def synchronize(self):
        self._check(lib().noor_synchronize(self._h))

    def set_timing(self, enable=True):
        self._check(lib().noor_set_timing(self._h, 1 if enable else 0))

    def kernel_times(self):
        """{kernel name: (launches, total ms)} measured with HIP events on the context's stream."""
        n = C.c_uint32()
        self._check(lib().noor_get_kernel_times(self._h, None, 0, C.byref(n)))
        arr = (abi.KernelTime * max(1, n.value))()
        self._check(lib().noor_get_kernel_times(self._h, arr, n.value, C.byref(n)))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(n.value)}

    def reset_kernel_times(self):
        self._check(lib().noor_reset_kernel_times(self._h))

    def set_wavefront_params(self, queue_capacity=0, block=64):
        self._check(lib().noor_set_wavefront_params(self._h, queue_capacity, block))

    def set_stack_lds(self, entries=0):
        self._check(lib().noor_set_stack_lds(self._h, entries))

    @staticmethod
    def _views(views):
        """views: list of (camera, frame, spp, seed) tuples or abi.View."""
        arr = (abi.View * len(views))()
        for i, v in enumerate(views):
            if isinstance(v, abi.View):
                arr[i] = v
            else:
                cam, frame, spp, seed = v
                arr[i] = abi.View(cam, frame, spp, seed, 0)
        return arr

    def render_batch(self, views):
        """Renders several views in one wavefront pass (noor_render_batch)."""
        arr = self._views(views)
        self._batch = [(arr[i].camera.w, arr[i].camera.h) for i in range(len(arr))]
        self._check(lib().noor_render_batch(self._h, arr, len(arr)))

    def batch_film(self, i):
        w, h = self._batch[i]
        out = np.empty((h, w, 4), np.float32)
        self._check(lib().noor_read_batch_film(self._h, i, out.ctypes.data_as(C.POINTER(C.c_float)), out.size))
        return out

    def autotune(self, views):
        """Times the batch for each wavefront setting and keeps the fastest;
        returns [(queue_capacity, block, stack_lds, ms)]."""
        arr = self._views(views)
        self._batch = [(arr[i].camera.w, arr[i].camera.h) for i in range(len(arr))]
        out = (abi.TuneEntry * 64)()
        n = C.c_uint32()
        self._check(lib().noor_autotune(self._h, arr, len(arr), out, 64, C.byref(n)))
        return [(out[i].queue_capacity, out[i].block, out[i].stack_lds, out[i].ms) for i in range(n.value)]

    def trace(self, rays, any_hit=False):
        n = len(rays)
        hits = (abi.Hit * n)()
        self._check(lib().noor_trace(self._h, rays, n, 1 if any_hit else 0, hits))
        return hits

    def close(self):
        if self._h:
            lib().noor_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hits_to_numpy(hits):
    a = np.ctypeslib.as_array(C.cast(hits, C.POINTER(C.c_uint32)), shape=(len(hits), 8)).copy()
    return a


def rays_from_numpy(o, d, tmax=None):
    n = len(o)
    rays = (abi.Ray * n)()
    arr = np.ctypeslib.as_array(C.cast(rays, C.POINTER(C.c_float)), shape=(n, 8))
    arr[:, 0:3] = o
    arr[:, 3] = np.inf if tmax is None else tmax
    arr[:, 4:7] = d
    arr[:, 7] = 0
    return rays
