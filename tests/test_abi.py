"""The C ABI library without a GPU: it loads, exports every function that
include/*.h declares, its struct layouts agree with the ctypes mirror, and
the argument/payload checks of noor_create reject bad input before any
device work (on a host without a GPU a valid payload ends in NO_DEVICE)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from noor_amd import abi, scenes
from noor_amd._lib import LIB_PATH, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "noor_camera": abi.Camera, "noor_hosek": abi.Hosek, "noor_hosek_state": abi.HosekState,
    "noor_spec": abi.Spec, "noor_material": abi.Material, "noor_shape": abi.Shape,
    "noor_area_light": abi.AreaLight, "noor_point_light": abi.PointLight, "noor_spot_light": abi.SpotLight,
    "noor_distant_light": abi.DistantLight, "noor_texture_view": abi.TextureView,
    "noor_payload_view": abi.PayloadView, "noor_options": abi.Options, "noor_stats": abi.Stats,
    "noor_ray": abi.Ray, "noor_hit": abi.Hit, "noor_transform": abi.Transform, "noor_kernel_time": abi.KernelTime,
    "noor_mesh_desc": abi.MeshDesc, "noor_instance_desc": abi.InstanceDesc, "noor_bvh_params": abi.BvhParams,
    "noor_scene_desc": abi.SceneDesc, "noor_build_stats": abi.BuildStats,
    "noor_view": abi.View, "noor_tune_entry": abi.TuneEntry,
    "noor_obj_material": abi.ObjMaterial, "noor_obj_view": abi.ObjView,
}


def header_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(noor_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_built_in_tree():
    assert os.path.exists(LIB_PATH)
    assert os.path.dirname(LIB_PATH) == os.path.join(ROOT, "noor_amd")


def test_every_declared_function_is_exported():
    names = header_functions()
    assert len(names) >= 30
    L = C.CDLL(LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize("name", sorted(STRUCTS))
def test_struct_layout_matches_ctypes(name):
    assert lib().noor_abi_struct_size(name.encode()) == C.sizeof(STRUCTS[name])


def test_status_strings_and_version():
    L = lib()
    assert L.noor_abi_version() >= 1
    seen = set()
    for s in range(9):
        txt = L.noor_status_string(s)
        assert txt and txt not in seen
        seen.add(txt)
    assert L.noor_abi_struct_size(b"no_such_struct") == 0


def _create(view, wl):
    out = C.c_void_p()
    st = lib().noor_create(C.byref(view), C.byref(wl.hosek) if wl.hosek is not None else None,
                           C.byref(wl.camera), C.byref(wl.spec), C.byref(wl.options), C.byref(out))
    return st, out


@pytest.fixture(scope="module")
def small():
    return scenes.cornell(subdiv=2, width=16, height=16)


def _copy_view(wl):
    return abi.PayloadView.from_buffer_copy(wl.payload.view)


def test_create_rejects_null_arguments(small):
    out = C.c_void_p()
    assert lib().noor_create(None, None, C.byref(small.camera), C.byref(small.spec), None, C.byref(out)) == \
        abi.ERR_INVALID_ARG
    v = _copy_view(small)
    assert lib().noor_create(C.byref(v), None, None, C.byref(small.spec), None, C.byref(out)) == abi.ERR_INVALID_ARG
    assert lib().noor_create(C.byref(v), None, C.byref(small.camera), C.byref(small.spec), None, None) == \
        abi.ERR_INVALID_ARG


def test_create_rejects_bad_payloads(small):
    arr = small.payload.arrays()
    keep = []

    def ptr_u4(a):
        a = np.ascontiguousarray(a, np.uint32)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(abi.uint4))

    # vertex index out of range
    v = _copy_view(small)
    attr = arr["attr"].copy()
    attr[3, 1] = v.num_vertices
    v.attribute_indices = ptr_u4(attr)
    assert _create(v, small)[0] == abi.ERR_INVALID_PAYLOAD
    # material index out of range
    v = _copy_view(small)
    attr = arr["attr"].copy()
    attr[0, 3] = v.num_materials
    v.attribute_indices = ptr_u4(attr)
    assert _create(v, small)[0] == abi.ERR_INVALID_PAYLOAD
    # root out of range
    v = _copy_view(small)
    v.bvh_root_node = v.num_bvh_nodes
    assert _create(v, small)[0] == abi.ERR_INVALID_PAYLOAD
    # inner node whose right child points past the array
    v = _copy_view(small)
    a = arr["bvh_a"].copy()
    b = arr["bvh_b"]
    inner = [i for i in range(len(a)) if (int(b[i, 3]) & 0xF8000000) == 0][0]
    a[inner, 3] = v.num_bvh_nodes + 5
    v.bvh_min_right_start = ptr_u4(a)
    assert _create(v, small)[0] == abi.ERR_INVALID_PAYLOAD
    # triangle leaf overrunning the triangle array
    v = _copy_view(small)
    a = arr["bvh_a"].copy()
    leaf = [i for i in range(len(a)) if int(b[i, 3]) & 0x80000000][0]
    a[leaf, 3] = v.num_triangles
    v.bvh_min_right_start = ptr_u4(a)
    assert _create(v, small)[0] == abi.ERR_INVALID_PAYLOAD
    # non-finite vertex
    v = _copy_view(small)
    vert = arr["vertices"].copy()
    vert[2, 1] = np.nan
    keep.append(vert)
    v.vertices = vert.ctypes.data_as(C.POINTER(abi.float4))
    assert _create(v, small)[0] == abi.ERR_INVALID_PAYLOAD
    # missing arrays
    v = _copy_view(small)
    v.num_materials = 0
    assert _create(v, small)[0] == abi.ERR_INVALID_ARG


def test_valid_payload_without_device(small):
    if lib().noor_device_count() > 0:
        pytest.skip("a HIP device is present")
    st, out = _create(_copy_view(small), small)
    assert st == abi.ERR_NO_DEVICE and not out.value


def test_context_calls_reject_null_context():
    L = lib()
    f = (C.c_uint32 * 1)(1)
    assert L.noor_render(None, f, 1) == abi.ERR_INVALID_ARG
    assert L.noor_render_wavefront(None, f, 1) == abi.ERR_INVALID_ARG
    assert L.noor_read_film(None, None, 0) == abi.ERR_INVALID_ARG
    assert L.noor_trace(None, None, 0, 0, None) == abi.ERR_INVALID_ARG
    assert L.noor_synchronize(None) == abi.ERR_INVALID_ARG
    assert L.noor_destroy(None) == abi.ERR_INVALID_ARG
    assert L.noor_render_batch(None, None, 0) == abi.ERR_INVALID_ARG
    assert L.noor_read_batch_film(None, 0, None, 0) == abi.ERR_INVALID_ARG
    assert L.noor_autotune(None, None, 0, None, 0, None) == abi.ERR_INVALID_ARG
    assert L.noor_set_stack_lds(None, 0) == abi.ERR_INVALID_ARG
    assert L.noor_set_params_device(None, None, None, 0, None, 0, 0) == abi.ERR_INVALID_ARG
    assert L.noor_render_sample(None, 1, None) == abi.ERR_INVALID_ARG
    assert L.noor_adjoint(None, 1, None, None, None, None, 0, None) == abi.ERR_INVALID_ARG


@pytest.mark.parametrize("name", ["cornell", "procedural", "terrain"])
def test_wide_bvh_covers_every_triangle_once(name):
    """The four-wide BVH of the traversal kernels, built on the host from the
    payload's binary BVH: exact triangle coverage per BLAS, nested boxes,
    leaf boxes around their triangles, and each 8-bit quantised box (as the
    GPU decodes it) containing the exact box of its child."""
    kw = dict(width=16, height=16)
    if name == "terrain":
        kw.update(grid=200, n_spheres=50, env_w=64, env_h=32)
    if name == "procedural":
        kw.update(target_tris=80000, n_meshes=40)
    wl = getattr(scenes, name)(**kw)
    n, need = C.c_uint32(), C.c_uint32()
    assert lib().noor_wide_bvh_check(C.byref(wl.payload.view), C.byref(n), C.byref(need)) == abi.OK
    binary = wl.payload.view.num_bvh_nodes
    assert 0 < n.value < binary  # four-wide: far fewer nodes than the binary tree
    assert 1 <= need.value <= 3 * 64


def test_wide_bvh_check_rejects_corrupt_payload(small):
    arr = small.payload.arrays()
    v = _copy_view(small)
    a = arr["bvh_a"].copy()
    b = arr["bvh_b"]
    inner = [i for i in range(len(a)) if (int(b[i, 3]) & 0xF8000000) == 0][0]
    a[inner, 3] = v.num_bvh_nodes + 5
    keep = np.ascontiguousarray(a, np.uint32)
    v.bvh_min_right_start = keep.ctypes.data_as(C.POINTER(abi.uint4))
    assert lib().noor_wide_bvh_check(C.byref(v), None, None) == abi.ERR_INVALID_PAYLOAD


def test_pod_offsets_match_reference_layouts(tmp_path):
    """tests/abi_offsets.c asserts every field offset of the camera, spec,
    material and Hosek PODs against the reference's CUDA declarations; a
    planted offset (the spec's sky type packed after the shorts, as a 2-byte
    aligned enum would sit) must fail the same check."""
    import subprocess
    inc = os.path.join(ROOT, "include")
    r = subprocess.run(["gcc", "-fsyntax-only", "-std=c11", "-I", inc, os.path.join(ROOT, "tests", "abi_offsets.c")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    bad = tmp_path / "bad.c"
    bad.write_text('#include <stddef.h>\n#include "noor_hip.h"\n'
                   '_Static_assert(offsetof(noor_spec, skydome_type) == 54, "planted");\n')
    r = subprocess.run(["gcc", "-fsyntax-only", "-std=c11", "-I", inc, str(bad)], capture_output=True, text=True)
    assert r.returncode != 0 and "planted" in r.stderr

