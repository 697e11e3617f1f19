"""Payload producers for files (include/noor_loaders.h): Wavefront OBJ + MTL
models into a SceneBuilder and float images as arrays.

Mirrors Model::load with ObjLoader (noor/source/model.cpp:132-330,
noor/source/objLoader.cpp:28-147, noor/header/objLoader.h:90-169): one mesh
with the file's materials (constants or image textures, BUMP with a bump map,
ALPHA with a transparency map), quad area lights for groups named "light",
coat parameters at the AssetLoader defaults (noor/header/assetLoader.h:143-148).
Parsing is the library's (no CPU fallback: the native library must be built).
"""
import ctypes as C
import os

import numpy as np

from . import abi
from ._lib import lib
from .scene import Mesh, SceneBuilder, material


def _err(status, what):
    msg = lib().noor_loader_last_error()
    raise abi.NoorError(status, f"{what}: {msg.decode() if msg else ''}")


def load_image(path):
    """(H, W, 4) float32, row 0 first: Radiance .hdr, PFM, binary PGM / PPM,
    PNG, TGA, OpenEXR, baseline JPEG (include/noor_loaders.h)."""
    w, h = C.c_uint32(), C.c_uint32()
    p = C.POINTER(C.c_float)()
    st = lib().noor_image_load(os.fsencode(path), C.byref(w), C.byref(h), C.byref(p))
    if st != 0:
        _err(st, "noor_image_load")
    try:
        return np.ctypeslib.as_array(p, shape=(h.value, w.value, 4)).copy()
    finally:
        lib().noor_image_free(p)


class ObjModel:
    """The arrays of a loaded OBJ (copies) and its materials and lights."""

    def __init__(self, path, unitize=True, crease_deg=30.0):
        h = C.c_void_p()
        st = lib().noor_obj_load(os.fsencode(path), 1 if unitize else 0, float(crease_deg), C.byref(h))
        if st != 0:
            _err(st, "noor_obj_load")
        try:
            v = abi.ObjView()
            st = lib().noor_obj_view_get(h, C.byref(v))
            if st != 0:
                _err(st, "noor_obj_view_get")
            nv, nt = v.num_vertices, v.num_triangles
            arr = np.ctypeslib.as_array
            self.positions = arr(v.positions, shape=(nv, 3)).copy()
            self.normals = arr(v.normals, shape=(nv, 3)).copy()
            self.uvs = arr(v.uvs, shape=(nv, 2)).copy()
            self.indices = arr(v.indices, shape=(nt, 3)).copy()
            self.material = arr(v.material, shape=(nt,)).copy()
            self.materials = []
            for i in range(v.num_materials):
                m = v.materials[i]
                self.materials.append(dict(
                    name=m.name.decode(), type=m.type, diffuse=tuple(m.diffuse), specular=tuple(m.specular),
                    transmission=tuple(m.transmission), emittance=tuple(m.emittance), ior=m.ior,
                    roughness=m.roughness, transparency=m.transparency, bumpfactor=m.bumpfactor,
                    uvscale=tuple(m.uvscale), map_kd=m.map_kd.decode(), map_ks=m.map_ks.decode(),
                    map_ke=m.map_ke.decode(), map_d=m.map_d.decode(), map_bump=m.map_bump.decode()))
            self.lights = arr(v.lights, shape=(v.num_lights, 15)).copy() if v.num_lights else np.zeros((0, 15))
            self.bbox_min = np.array(v.bbox_min)
            self.bbox_max = np.array(v.bbox_max)
            self.unitize_scale = v.unitize_scale
        finally:
            lib().noor_obj_free(h)
        self.path = path


def obj_scene(path, sb=None, unitize=True, crease_deg=30.0, texture_dir=None):
    """Adds the model in `path` to a SceneBuilder (new if None) as
    Model::loadMaterial / loadLights do; image maps are read from
    `texture_dir` (default: the OBJ's directory).  Returns (sb, ObjModel)."""
    m = ObjModel(path, unitize, crease_deg)
    sb = sb if sb is not None else SceneBuilder()
    tdir = texture_dir if texture_dir is not None else os.path.dirname(os.path.abspath(path))
    images = {}

    def img(name):
        if name not in images:
            images[name] = load_image(os.path.join(tdir, name))
        return images[name]

    ids = []
    for d in m.materials:
        kw = dict(diffuse=img(d["map_kd"]) if d["map_kd"] else d["diffuse"],
                  specular=img(d["map_ks"]) if d["map_ks"] else d["specular"],
                  transmission=d["transmission"],
                  emittance=img(d["map_ke"]) if d["map_ke"] else d["emittance"],
                  eta=(d["ior"],) * 3, k=(1.0, 1.0, 1.0), roughness=(d["roughness"], d["roughness"]),
                  coat_ior=1.5, coat_weight=1.0, coat_thickness=1.0, coat_sigma=0.0, emitter_scale=1.0,
                  uvscale={"diffuse": d["uvscale"]})
        if d["map_d"]:
            kw["alpha"] = img(d["map_d"])
        if d["map_bump"]:
            kw["bump"] = img(d["map_bump"])
            kw["bumpfactor"] = d["bumpfactor"]
        ids.append(sb.add_material(material(d["type"], **kw)))
    mat = np.asarray(ids, np.uint32)[m.material]
    sb.add_mesh(Mesh(m.positions, m.indices.astype(np.uint32), normals=m.normals, uvs=m.uvs, material=mat))
    for L in m.lights:
        sb.add_quad_light(L[0:3], L[3:6], L[6:9], L[12:15])
    return sb, m
