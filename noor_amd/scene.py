"""Host-side scene API: meshes, materials, lights, environment, camera, spec.

Mirrors the reference's host layer (Model / Scene / Camera / HosekSky of
noor/header/*.h, noor/source/*.cpp) on top of the native builder
(include/noor_scene.h).  Materials are stored as the reference stores them:
every parameter is a texture index, constants are 1x1 textures
(noor/source/model.cpp:184-330); the last texture is the environment.
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import check, lib


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Mesh:
    """One source mesh in object space (triangles counter-clockwise)."""

    def __init__(self, positions, indices, normals=None, uvs=None, material=0):
        self.positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        self.normals = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        self.uvs = None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2)
        if np.isscalar(material):
            self.material_all = int(material)
            self.material = None
        else:
            self.material = np.ascontiguousarray(material, dtype=np.uint32).reshape(-1)
            self.material_all = 0

    @property
    def num_triangles(self):
        return int(self.indices.shape[0])


class Payload:
    """A built scene (BVH + triangle records + tables) owned by the native side."""

    def __init__(self, handle, keep):
        self._h = handle
        self._keep = keep
        self.view = abi.PayloadView()
        check(lib().noor_payload_view_get(self._h, C.byref(self.view)))
        self.stats = abi.BuildStats()
        check(lib().noor_payload_stats(self._h, C.byref(self.stats)))

    @property
    def handle(self):
        return self._h

    def arrays(self):
        """numpy views of the payload arrays (no copy)."""
        v = self.view
        nt, nn = v.num_triangles, v.num_bvh_nodes

        def u4(p, n):
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), shape=(n, 4))

        def f4(p, n):
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(n, 4))

        return {
            "attr": u4(v.attribute_indices, nt),
            "wald0": u4(v.wald_ax_nu_nv_nd, nt),
            "wald1": u4(v.wald_bnu_bnv_au_av, nt),
            "wald2": u4(v.wald_cnu_cnv_deg_mat, nt),
            "prim_ids": np.ctypeslib.as_array(v.prim_ids, shape=(nt,)),
            "vertices": f4(v.vertices, v.num_vertices),
            "bvh_a": u4(v.bvh_min_right_start, nn),
            "bvh_b": u4(v.bvh_max_axis_count, nn),
        }

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().noor_payload_free(h)
            except Exception:
                pass
            self._h = None


def material(mtype, diffuse=(0.5, 0.5, 0.5), specular=(1.0, 1.0, 1.0), transmission=(1.0, 1.0, 1.0),
             emittance=(0.0, 0.0, 0.0), eta=(1.0, 1.0, 1.0), k=(0.0, 0.0, 0.0), roughness=(0.5, 0.5),
             coat_ior=1.5, emitter_scale=1.0, coat_weight=0.5, coat_thickness=0.0, coat_sigma=0.0,
             bump=None, bumpfactor=1.0, alpha=None, uvscale=None):
    """Material description; textures are created by SceneBuilder.add_material.

    Every colour (and `roughness`) is a constant or an image: an (H, W), (H, W, 3)
    or (H, W, 4) array with power-of-two sides, row 0 at v = 0.  `bump` is a
    displacement image (its first channel, scaled by `bumpfactor`) and sets the
    BUMP type bit; `alpha` is a transparency image (first channel; texels
    <= 1e-4 cut the surface away) and sets the ALPHA bit
    (noor/source/model.cpp:286-311).  `uvscale` maps "diffuse", "specular",
    "emittance", "roughness", "bump" or "transparency" to the (su, sv) scale
    of that texture's coordinates; transmission, eta and k always read at
    scale 1 (noor/cuda/materialimp.cuh:164-186)."""
    t = int(mtype)
    if bump is not None:
        t |= abi.MAT_BUMP
    if alpha is not None:
        t |= abi.MAT_ALPHA
    return dict(type=t, diffuse=diffuse, specular=specular, transmission=transmission,
                emittance=emittance, eta=eta, k=k, roughness=roughness, coat_ior=coat_ior,
                emitter_scale=emitter_scale, coat_weight=coat_weight, coat_thickness=coat_thickness,
                coat_sigma=coat_sigma, bump=bump, bumpfactor=bumpfactor, alpha=alpha, uvscale=uvscale or {})


def _is_image(v):
    return isinstance(v, np.ndarray) and v.ndim >= 2 and v.shape[0] * v.shape[1] > 1


def image_rgba(img):
    """(H, W), (H, W, 1..4) -> contiguous (H, W, 4) float32; grey images fill
    rgb, missing alpha is 1."""
    a = np.asarray(img, np.float32)
    if a.ndim == 2:
        a = a[:, :, None]
    h, w, c = a.shape
    out = np.ones((h, w, 4), np.float32)
    if c == 1:
        out[:, :, :3] = a
    else:
        out[:, :, :min(c, 4)] = a[:, :, :4]
    return np.ascontiguousarray(out)


class SceneBuilder:
    def __init__(self):
        self.meshes = []
        self.instances = []
        self.materials = []
        self.textures = [np.full((1, 1, 4), 0.5, np.float32)]  # texture 0: default grey
        self.area_lights = []
        self.point_lights = []
        self.spot_lights = []
        self.distant_lights = []
        self.env = None

    def add_mesh(self, mesh):
        self.meshes.append(mesh)
        return len(self.meshes) - 1

    def add_instance(self, mesh_index, object_to_world=None):
        M = np.eye(4, dtype=np.float64)[:3] if object_to_world is None else np.asarray(object_to_world, np.float64)[:3, :4]
        self.instances.append((mesh_index, M))
        return len(self.instances) - 1

    def _const_tex(self, rgb, alpha=1.0):
        t = np.zeros((1, 1, 4), np.float32)
        t[0, 0, :3] = np.asarray(rgb, np.float32)[:3]
        t[0, 0, 3] = alpha
        self.textures.append(t)
        return len(self.textures) - 1

    def add_texture(self, img):
        """Adds an image texture (see material()); returns its index."""
        self.textures.append(image_rgba(img))
        return len(self.textures) - 1

    def add_material(self, m):
        mt = abi.Material()
        for f in ("diffuse_uvscale", "specular_uvscale", "transparency_uvscale", "emittance_uvscale",
                  "bump_uvscale", "roughness_uvscale", "metalness_uvscale"):
            getattr(mt, f).x = 1.0
            getattr(mt, f).y = 1.0
        for name, (sx, sy) in m.get("uvscale", {}).items():
            f = getattr(mt, name + "_uvscale")
            f.x, f.y = float(sx), float(sy)
        tex = lambda v: self.add_texture(v) if _is_image(v) else self._const_tex(v)
        mt.transmission_tex_idx = tex(m["transmission"])
        mt.ior_tex_idx = tex(m["eta"])
        mt.k_tex_idx = tex(m["k"])
        mt.diffuse_tex_idx = tex(m["diffuse"])
        r = m["roughness"]
        if _is_image(r):
            ra = np.asarray(r, np.float32)
            if ra.ndim == 2:
                ra = np.stack([ra, ra], -1)
            mt.roughness_tex_idx = self.add_texture(np.concatenate([ra[:, :, :2], np.ones_like(ra[:, :, :1])], -1))
        else:
            mt.roughness_tex_idx = self._const_tex((r[0], r[1], 1.0))
        mt.metalness_tex_idx = self._const_tex((1.0, 1.0, 1.0))
        mt.specular_tex_idx = tex(m["specular"])
        mt.emittance_tex_idx = tex(m["emittance"])
        alpha = m.get("alpha")
        mt.transparency_tex_idx = self.add_texture(alpha) if alpha is not None else self._const_tex((1.0, 1.0, 1.0))
        bump = m.get("bump")
        mt.bump_tex_idx = self.add_texture(bump) if bump is not None else 0
        mt.bumpfactor = float(m.get("bumpfactor", 1.0))
        mt.emitter_scale = m["emitter_scale"]
        mt.coat_roughness = 0.001
        mt.coat_ior = m["coat_ior"]
        mt.coat_weight = m.get("coat_weight", 0.5)
        mt.coat_thickness = m.get("coat_thickness", 0.0)
        mt.coat_sigma = m.get("coat_sigma", 0.0)
        mt.type = m["type"]
        self.materials.append(mt)
        return len(self.materials) - 1

    def add_quad_light(self, corner, u, v, ke, two_sided=False):
        """Area quad light (CudaShape QUAD, noor/cuda/shape.cuh:44-64)."""
        corner, u, v = (np.asarray(a, np.float64) for a in (corner, u, v))
        n = np.cross(u, v)
        n = n / np.linalg.norm(n)
        L = abi.AreaLight()
        L.ke = abi.float3(*map(float, ke))
        L.type = abi.LIGHT_AREA
        s = L.shape
        s.center = abi.float3(*map(float, corner))
        s.u = abi.float3(*map(float, u))
        s.v = abi.float3(*map(float, v))
        s.n = abi.float3(*map(float, n))
        s.radius = 0.5 * min(np.linalg.norm(u), np.linalg.norm(v))
        s.radius2 = s.radius * s.radius
        s.area = float(np.float32(np.linalg.norm(np.cross(u, v))))
        s.type = abi.SHAPE_QUAD
        L.two_sided = 1 if two_sided else 0
        L.power = abi.float3(0, 0, 0)
        self.area_lights.append(L)
        return len(self.area_lights) - 1

    def _area(self, ke, shape_type, center, radius, normal=(0.0, 0.0, 1.0), area=0.0, two_sided=False):
        L = abi.AreaLight()
        L.ke = abi.float3(*map(float, ke))
        L.type = abi.LIGHT_AREA
        s = L.shape
        s.center = abi.float3(*map(float, center))
        s.u = abi.float3(0.0, 0.0, 0.0)
        s.v = abi.float3(0.0, 0.0, 0.0)
        s.n = abi.float3(*map(float, normal))
        s.radius = float(radius)
        s.radius2 = float(np.float32(radius) * np.float32(radius))
        s.area = float(np.float32(area))
        s.type = shape_type
        L.two_sided = 1 if two_sided else 0
        L.power = abi.float3(0, 0, 0)
        self.area_lights.append(L)
        return len(self.area_lights) - 1

    def add_sphere_light(self, center, radius, ke):
        """Spherical area light, emitting on its whole surface (CudaShape SPHERE,
        noor/cuda/shape.cuh; sampled by cone, noor/cuda/geometry.cuh:48-85)."""
        return self._area(ke, abi.SHAPE_SPHERE, center, radius, area=4.0 * np.pi * radius * radius)

    def add_disk_light(self, center, normal, radius, ke, two_sided=False):
        """Disk area light facing `normal` (CudaShape DISK, noor/cuda/geometry.cuh:86-103)."""
        n = np.asarray(normal, np.float64)
        n = n / np.linalg.norm(n)
        return self._area(ke, abi.SHAPE_DISK, center, radius, normal=n, area=np.pi * radius * radius,
                          two_sided=two_sided)

    def add_point_light(self, position, ke):
        """Point light of intensity ke (CudaPointLight, noor/cuda/lightimp.cuh:190-205)."""
        L = abi.PointLight()
        L.ke = abi.float3(*map(float, ke))
        L.power = abi.float3(0, 0, 0)
        L.type = abi.LIGHT_POINT
        L.position = abi.float3(*map(float, position))
        self.point_lights.append(L)
        return len(self.point_lights) - 1

    def add_spot_light(self, position, direction, cos_inner, cos_outer, ke):
        """Spot light (CudaSpotLight, noor/cuda/light.cuh:185-226): full intensity
        where dot(direction, w) >= cos_inner for w from the lit point towards the
        light, none below cos_outer, a quartic falloff between."""
        d = np.asarray(direction, np.float64)
        d = d / np.linalg.norm(d)
        L = abi.SpotLight()
        L.ke = abi.float3(*map(float, ke))
        L.power = abi.float3(0, 0, 0)
        L.type = abi.LIGHT_SPOT
        L.position = abi.float3(*map(float, position))
        L.direction = abi.float3(*map(float, d))
        L.cosinner = float(cos_inner)
        L.cosouter = float(cos_outer)
        self.spot_lights.append(L)
        return len(self.spot_lights) - 1

    def add_distant_light(self, direction, ke):
        """Directional light arriving from `direction` (towards the light), which
        the reference normalises in its constructor (noor/cuda/light.cuh:228-240)."""
        d = np.asarray(direction, np.float64)
        d = d / np.linalg.norm(d)
        L = abi.DistantLight()
        L.ke = abi.float3(*map(float, ke))
        L.power = abi.float3(0, 0, 0)
        L.type = abi.LIGHT_DISTANT
        L.direction = abi.float3(*map(float, d))
        self.distant_lights.append(L)
        return len(self.distant_lights) - 1

    def set_env(self, rgba):
        """Environment texture (RGBA32F, H x W x 4), stored as the last texture."""
        self.env = np.ascontiguousarray(rgba, np.float32)

    def build(self, num_bins=16, max_height=32, min_leaf=2, max_leaf=2, Ci=1.0, Ct=0.125, threads=0, device=None):
        """The payload (noor_scene_build); with `device` (a HIP device index) the
        trees of large meshes are built on that GPU (noor_scene_build_device)."""
        keep = []
        mds = (abi.MeshDesc * max(1, len(self.meshes)))()
        for i, m in enumerate(self.meshes):
            d = mds[i]
            d.num_vertices = m.positions.shape[0]
            d.positions = _fptr(m.positions)
            d.normals = _fptr(m.normals) if m.normals is not None else None
            d.uvs = _fptr(m.uvs) if m.uvs is not None else None
            d.num_triangles = m.num_triangles
            d.indices = _uptr(m.indices)
            d.material = _uptr(m.material) if m.material is not None else None
            d.material_all = m.material_all
            keep.append(m)
        insts = (abi.InstanceDesc * max(1, len(self.instances)))()
        for i, (mi, M) in enumerate(self.instances):
            insts[i].mesh = mi
            for k, val in enumerate(M.reshape(-1)):
                insts[i].object_to_world[k] = float(val)
        texs = list(self.textures)
        env = self.env if self.env is not None else np.full((1, 1, 4), 0.25, np.float32)
        texs.append(env)
        tvs = (abi.TextureView * len(texs))()
        for i, t in enumerate(texs):
            t = np.ascontiguousarray(t, np.float32)
            keep.append(t)
            tvs[i].height, tvs[i].width = t.shape[0], t.shape[1]
            tvs[i].rgba = _fptr(t)
        mats = (abi.Material * max(1, len(self.materials)))(*self.materials)
        lights = (abi.AreaLight * max(1, len(self.area_lights)))(*self.area_lights)
        d = abi.SceneDesc()
        d.num_meshes = len(self.meshes)
        d.meshes = mds
        d.num_instances = len(self.instances)
        d.instances = insts
        d.num_materials = len(self.materials)
        d.materials = mats
        d.num_textures = len(texs)
        d.textures = tvs
        d.num_area_lights = len(self.area_lights)
        d.area_lights = lights
        for name, T in (("point", abi.PointLight), ("spot", abi.SpotLight), ("distant", abi.DistantLight)):
            ls = getattr(self, name + "_lights")
            arr = (T * max(1, len(ls)))(*ls)
            keep.append(arr)
            setattr(d, f"num_{name}_lights", len(ls))
            setattr(d, f"{name}_lights", arr)
        bp = abi.BvhParams(num_bins, max_height, min_leaf, max_leaf, Ci, Ct, threads)
        h = C.c_void_p()
        if device is None:
            check(lib().noor_scene_build(C.byref(d), C.byref(bp), C.byref(h)))
        else:
            check(lib().noor_scene_build_device(C.byref(d), C.byref(bp), int(device), C.byref(h)))
        return Payload(h, keep)


def make_camera(eye, look_at, up=(0.0, 1.0, 0.0), fov=np.pi / 4, w=512, h=512, lens_radius=0.0,
                ortho_zoom=1.0, camera_type=abi.CAMERA_PERSP):
    cam = abi.Camera()
    e = (C.c_float * 3)(*map(float, eye))
    la = (C.c_float * 3)(*map(float, look_at))
    u = (C.c_float * 3)(*map(float, up))
    check(lib().noor_camera_look_at(e, la, u, float(fov), float(ortho_zoom), float(lens_radius), w, h,
                                    camera_type, C.byref(cam)))
    return cam


def make_spec(payload, bounces=5, rr=3, skydome_type=abi.SKY_HDR, mis=False):
    s = abi.Spec()
    check(lib().noor_spec_fill(payload.handle, bounces, rr, skydome_type, 1 if mis else 0, C.byref(s)))
    return s


def make_options(rng=abi.RNG_PCG32, tri_test=abi.TRI_WATERTIGHT, compat=abi.COMPAT_ALL, stats=0):
    return abi.Options(rng, tri_test, compat, stats)


def cook_hosek(turbidity=3.5, ground_albedo=(1.0, 1.0, 1.0), sun_lat=np.pi / 4, sun_lon=np.pi / 4, scale=250.0,
               solar_radius=2.0):
    """The sky HosekSky::update cooks (noor/header/hosek.h:61-83), with the
    reference's defaults: turbidity 3.5, white ground, sun at 45 degrees."""
    hk = abi.Hosek()
    alb = (C.c_float * 3)(*map(float, ground_albedo))
    check(lib().noor_hosek_cook(float(turbidity), alb, float(sun_lat), float(sun_lon), float(scale),
                                float(solar_radius), C.byref(hk)))
    return hk


def cook_hosek_state(turbidity, albedo, elevation):
    """One channel-set state of arhosek_rgb_skymodelstate_alloc_update."""
    st = abi.HosekState()
    check(lib().noor_hosek_cook_state(float(turbidity), float(albedo), float(elevation), C.byref(st)))
    return st


def make_hosek(configs, radiances, solar_radius, turbidity=3.0, albedo=0.3, elevation=np.pi / 6,
               sun_lat=np.pi / 3, sun_lon=np.pi / 4, solar_scale=250.0, solar_radius_factor=2.0):
    """CudaHosekSky from a cooked Hosek-Wilkie state (HosekSky::update,
    noor/header/hosek.h:61-83).  `configs` is 3 x 9, `radiances` 3; the same
    state is used for the three channel states, as the reference does when the
    ground albedo is grey."""
    hk = abi.Hosek()
    for st in (hk.state_r, hk.state_g, hk.state_b):
        for c in range(3):
            for j in range(9):
                st.configs[c][j] = float(configs[c][j])
            st.radiances[c] = float(radiances[c])
        st.turbidity = turbidity
        st.solar_radius = solar_radius
        st.albedo = albedo
        st.elevation = elevation
    # sphericalDirection(theta = sun_lat, phi = sun_lon), y up (noor/cuda/utils.cuh:437-446)
    d = np.array([np.sin(sun_lat) * np.cos(sun_lon), np.cos(sun_lat), np.sin(sun_lat) * np.sin(sun_lon)])
    d = d / np.linalg.norm(d)
    hk.solar_dir = abi.float3(*map(float, d))
    hk.ground_albedo = abi.float3(albedo, albedo, albedo)
    hk.solar_scale = solar_scale
    hk.solar_radius = float(np.float32(solar_radius_factor) * np.float32(solar_radius))
    return hk
