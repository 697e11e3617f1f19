"""Loader for the in-tree native library noor_amd/libnoor_hip.so.

The library holds the HIP kernels (gfx950), the C ABI of include/noor_hip.h
and the host scene builder of include/noor_scene.h.  There is no fallback: if
the library is missing, every entry point raises.
"""
import ctypes as C
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnoor_hip.so")
_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    # torch bundles its own HIP runtime under the same soname as the system's:
    # loading it first makes both use one runtime in this process
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    sigs = {
        "noor_status_string": (C.c_char_p, [C.c_int]),
        "noor_abi_version": (C.c_int, []),
        "noor_device_count": (C.c_int, []),
        "noor_create": (C.c_int, [P(abi.PayloadView), P(abi.Hosek), P(abi.Camera), P(abi.Spec),
                                  P(abi.Options), P(C.c_void_p)]),
        "noor_destroy": (C.c_int, [C.c_void_p]),
        "noor_update_camera": (C.c_int, [C.c_void_p, P(abi.Camera)]),
        "noor_update_hosek": (C.c_int, [C.c_void_p, P(abi.Hosek)]),
        "noor_update_spec": (C.c_int, [C.c_void_p, P(abi.Spec)]),
        "noor_update_options": (C.c_int, [C.c_void_p, P(abi.Options)]),
        "noor_last_error": (C.c_char_p, [C.c_void_p]),
        "noor_render": (C.c_int, [C.c_void_p, P(C.c_uint32), C.c_uint32]),
        "noor_render_wavefront": (C.c_int, [C.c_void_p, P(C.c_uint32), C.c_uint32]),
        "noor_debug_skydome": (C.c_int, [C.c_void_p, P(C.c_uint32)]),
        "noor_get_lookat": (C.c_int, [C.c_void_p, P(abi.float4)]),
        "noor_read_film": (C.c_int, [C.c_void_p, P(C.c_float), C.c_size_t]),
        "noor_film_device_ptr": (C.c_int, [C.c_void_p, P(C.c_void_p)]),
        "noor_get_stats": (C.c_int, [C.c_void_p, P(abi.Stats)]),
        "noor_reset_stats": (C.c_int, [C.c_void_p]),
        "noor_synchronize": (C.c_int, [C.c_void_p]),
        "noor_trace": (C.c_int, [C.c_void_p, P(abi.Ray), C.c_uint32, C.c_int, P(abi.Hit)]),
        "noor_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
        "noor_get_kernel_times": (C.c_int, [C.c_void_p, P(abi.KernelTime), C.c_uint32, P(C.c_uint32)]),
        "noor_reset_kernel_times": (C.c_int, [C.c_void_p]),
        "noor_set_wavefront_params": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int32]),
        "noor_set_stack_lds": (C.c_int, [C.c_void_p, C.c_uint32]),
        "noor_render_batch": (C.c_int, [C.c_void_p, P(abi.View), C.c_uint32]),
        "noor_read_batch_film": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_float), C.c_size_t]),
        "noor_batch_film_device_ptr": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_void_p)]),
        "noor_autotune": (C.c_int, [C.c_void_p, P(abi.View), C.c_uint32, P(abi.TuneEntry), C.c_uint32,
                                    P(C.c_uint32)]),
        "noor_set_params_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                             C.c_uint32]),
        "noor_render_sample": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
        "noor_adjoint": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                   C.c_void_p]),
        "noor_trace_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]),
        "noor_wide_bvh_check": (C.c_int, [P(abi.PayloadView), P(C.c_uint32), P(C.c_uint32)]),
        "noor_abi_struct_size": (C.c_size_t, [C.c_char_p]),
        "noor_scene_build": (C.c_int, [P(abi.SceneDesc), P(abi.BvhParams), P(C.c_void_p)]),
        "noor_scene_build_device": (C.c_int, [P(abi.SceneDesc), P(abi.BvhParams), C.c_int, P(C.c_void_p)]),
        "noor_payload_view_get": (C.c_int, [C.c_void_p, P(abi.PayloadView)]),
        "noor_payload_stats": (C.c_int, [C.c_void_p, P(abi.BuildStats)]),
        "noor_payload_free": (None, [C.c_void_p]),
        "noor_camera_look_at": (C.c_int, [P(C.c_float), P(C.c_float), P(C.c_float), C.c_float, C.c_float,
                                          C.c_float, C.c_uint32, C.c_uint32, C.c_int32, P(abi.Camera)]),
        "noor_spec_fill": (C.c_int, [C.c_void_p, C.c_uint16, C.c_uint16, C.c_int32, C.c_uint8, P(abi.Spec)]),
        "noor_hosek_cook_state": (C.c_int, [C.c_float, C.c_float, C.c_float, P(abi.HosekState)]),
        "noor_hosek_cook": (C.c_int, [C.c_float, P(C.c_float), C.c_float, C.c_float, C.c_float, C.c_float,
                                      P(abi.Hosek)]),
        "noor_obj_load": (C.c_int, [C.c_char_p, C.c_int, C.c_float, P(C.c_void_p)]),
        "noor_obj_view_get": (C.c_int, [C.c_void_p, P(abi.ObjView)]),
        "noor_obj_free": (None, [C.c_void_p]),
        "noor_loader_last_error": (C.c_char_p, []),
        "noor_image_load": (C.c_int, [C.c_char_p, P(C.c_uint32), P(C.c_uint32), P(P(C.c_float))]),
        "noor_image_free": (None, [P(C.c_float)]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status, ctx=None):
    if status != 0:
        msg = ""
        if ctx:
            m = lib().noor_last_error(ctx)
            msg = m.decode() if m else ""
        raise abi.NoorError(status, msg)


def has_symbol(name):
    try:
        getattr(lib(), name)
        return True
    except AttributeError:
        return False
