"""Oracle pins for the image-texture pipeline (noor/cuda/texture.cuh:30-423,
noor/cuda/materialimp.cuh:120-186, noor/cuda/intersection.cuh:115-440,
noor/cuda/cameraimp.cuh:175-254, noor/cuda/accumulate.cuh:47-99,
noor/cuda/triangle.cuh:124-135).

- mip chains against an independent box-filter statement of the reference's
  mipmap kernel (four bilinear samples at texel corners = a 4x4 box with a
  half-texel offset, wrap addressing), which fails for a dropped offset, a
  clamp instead of a wrap or a wrong level size;
- lookups against closed forms: bilinear at level 0, trilinear at integer
  and half-integer lod, the last level for a huge footprint, point sampling;
- camera differentials against the pixel geometry (the auxiliary rays pass
  through the neighbouring pixel), for pinhole, thin-lens, orthographic and
  environment cameras;
- hit differentials against the analytic intersection of the auxiliary rays
  with a plane of affine uv (a transposed 2x2 system or a wrong projection
  axis fails);
- alpha: an all-zero transparency map removes a surface, camera and shadow
  rays alike (the film equals the film without it, bit for bit), an all-one
  map leaves it opaque; a constant bump map changes the frame only by
  rounding.

The reference samples through CUDA texture units (filter weights in
hardware fixed point, 64x anisotropic footprints); their exact output is
unspecified, so the pins are the closed forms above, not reference images
(parity with the reference's hardware filtering is unpinned).
"""
import numpy as np
import pytest

from noor_amd import abi
from noor_amd.scene import Mesh, SceneBuilder, make_camera, make_options, make_spec, material
from noor_amd.scenes import Workload, quad_uv, textured
from oracle.binding import OracleScene


def _plane_scene(mat, cam=None, extra=(), res=(32, 24), mis=False, env=None, lights=None, spec_bounces=3):
    sb = SceneBuilder()
    m = sb.add_material(mat)
    pos, tri, uv = quad_uv((-2.0, 0.0, -3.0), (4.0, 0.0, 0.0), (0.0, 0.0, 4.0))  # y = 0, normal +y
    sb.add_mesh(Mesh(pos, tri, uvs=uv, material=m))
    for (p, t, u), mm in extra:
        sb.add_mesh(Mesh(p, t, uvs=u, material=sb.add_material(mm)))
    if lights:
        lights(sb)
    if env is not None:
        sb.set_env(env)
    payload = sb.build()
    if cam is None:
        cam = make_camera((0.3, 1.2, 1.5), (0.0, 0.0, -1.0), (0, 1, 0), np.radians(50.0), res[0], res[1])
    spec = make_spec(payload, bounces=spec_bounces, rr=10, skydome_type=abi.SKY_HDR, mis=mis)
    return Workload("plane", payload, cam, spec, make_options(compat=abi.COMPAT_ALL), None, 1, info={})


def _box_mip(src):
    """Next level of the reference's mipmap kernel, restated as a separable
    filter in float64: along an axis of size n > 1 the texel x of the next
    level averages source texels 2x-1 .. 2x+2 (mod n) with weight 1/4; an axis
    of size 1 stays."""
    def axis(a, ax):
        n = a.shape[ax]
        if n == 1:
            return a
        m = max(1, n // 2)
        out = np.zeros_like(np.take(a, np.arange(m), axis=ax))
        for k in (-1, 0, 1, 2):
            out += 0.25 * np.take(a, (2 * np.arange(m) + k) % n, axis=ax)
        return out
    return axis(axis(src.astype(np.float64), 0), 1)


@pytest.mark.parametrize("h,w", [(8, 16), (32, 4), (1, 8), (16, 16)])
def test_mip_chain_is_the_reference_box_filter(h, w):
    rng = np.random.default_rng(h * 100 + w)
    img = rng.random((h, w, 4)).astype(np.float32)
    o = OracleScene.from_workload(_plane_scene(material(abi.MAT_DIFFUSE, diffuse=img)))
    levels = o.texture_levels(0)
    assert len(levels) == int(max(h, w)).bit_length()
    np.testing.assert_array_equal(levels[0], img)
    ref = img.astype(np.float64)
    for lv in levels[1:]:
        ref = _box_mip(ref)
        assert lv.shape == ref.shape
        np.testing.assert_allclose(lv, ref, rtol=1e-5, atol=1e-6)
    assert levels[-1].shape == (1, 1, 4)


def _bilinear(img, u, v):
    """tex2D with linear filtering and wrap addressing, float64."""
    h, w, _ = img.shape
    u -= np.floor(u)
    v -= np.floor(v)
    x, y = u * w - 0.5, v * h - 0.5
    x0, y0 = int(np.floor(x)), int(np.floor(y))
    ax, ay = x - x0, y - y0
    t = lambda i, j: img[j % h, i % w].astype(np.float64)
    return ((1 - ay) * ((1 - ax) * t(x0, y0) + ax * t(x0 + 1, y0))
            + ay * ((1 - ax) * t(x0, y0 + 1) + ax * t(x0 + 1, y0 + 1)))


def test_texture_lookups_against_closed_forms():
    rng = np.random.default_rng(5)
    img = rng.random((16, 32, 4)).astype(np.float32)
    o = OracleScene.from_workload(_plane_scene(material(abi.MAT_DIFFUSE, diffuse=img)))
    levels = o.texture_levels(0)
    for _ in range(50):
        u, v = (float(np.float32(a)) for a in rng.uniform(-3, 3, 2))  # the lookup takes binary32 uv
        b0 = _bilinear(img, u, v)
        np.testing.assert_allclose(o.tex_lookup(0, 0, u, v), b0, atol=4e-6)
        # zero footprint: tex2DGrad reads level 0
        np.testing.assert_allclose(o.tex_lookup(0, 1, u, v), b0, atol=4e-6)
        # footprint of 4 texels along x (width 32): lod 2 exactly
        np.testing.assert_allclose(o.tex_lookup(0, 1, u, v, (4.0 / 32, 0, 0, 1.0 / 16)),
                                   _bilinear(levels[2], u, v), atol=4e-6)
        # lod 2.5: halfway between levels 2 and 3 (footprint 4 sqrt 2 texels, along y here)
        f = 4.0 * np.sqrt(2.0) / 16
        np.testing.assert_allclose(o.tex_lookup(0, 1, u, v, (0, 0, 0, f)),
                                   0.5 * _bilinear(levels[2], u, v) + 0.5 * _bilinear(levels[3], u, v), atol=1e-5)
        # a footprint beyond the texture: the last level, the mean of the image's box chain
        np.testing.assert_allclose(o.tex_lookup(0, 1, u, v, (1e3, 0, 0, 0)), levels[-1][0, 0], atol=1e-6)
        # point sampling: the texel under (u, v)
        x = int(np.floor((u - np.floor(u)) * 32)) % 32
        y = int(np.floor((v - np.floor(v)) * 16)) % 16
        np.testing.assert_array_equal(o.tex_lookup(0, 2, u, v), img[y, x])


def _cam_frame(cam):
    R = np.array([[cam.camera_to_world.row[i].x, cam.camera_to_world.row[i].y, cam.camera_to_world.row[i].z]
                  for i in range(3)], np.float64)
    t = np.array([cam.camera_to_world.t.x, cam.camera_to_world.t.y, cam.camera_to_world.t.z])
    dxc = np.array([cam.dx_camera.x, cam.dx_camera.y, cam.dx_camera.z])
    dyc = np.array([cam.dy_camera.x, cam.dy_camera.y, cam.dy_camera.z])
    return R, t, dxc, dyc


@pytest.mark.parametrize("lens", [0.0, 0.05])
def test_perspective_camera_differentials(lens):
    """The auxiliary directions point through the raster point one pixel to
    the right / below: in camera space, scaled to the image plane z of the
    main ray, they differ from it by dx_camera / dy_camera (pinhole); with a
    lens they leave the same lens point and meet their own focus points."""
    wl = _plane_scene(material(abi.MAT_DIFFUSE))
    cam = make_camera((0.3, 1.2, 1.5), (0.0, 0.0, -1.0), (0, 1, 0), np.radians(50.0), 32, 24, lens_radius=lens)
    wl.camera = cam
    o = OracleScene.from_workload(wl)
    R, t, dxc, dyc = _cam_frame(cam)
    Rinv = np.linalg.inv(R)
    for (x, y, fr) in ((3, 4, 1), (20, 11, 2), (31, 23, 3)):
        ro, rd, ox, oy, dx, dy = o.camera_ray_diff(x, y, fr).astype(np.float64)
        oc, dc = Rinv @ (ro - t), Rinv @ rd
        dxcam, dycam = Rinv @ dx, Rinv @ dy
        if lens == 0.0:
            np.testing.assert_allclose(ox, ro, atol=1e-6)
            np.testing.assert_allclose(oy, ro, atol=1e-6)
            # scaled to the image plane, the auxiliary directions are one pixel step away
            zp = _image_plane_z(cam)
            np.testing.assert_allclose(dxcam * (zp / dxcam[2]) - dc * (zp / dc[2]), dxc, rtol=1e-4, atol=1e-6)
            np.testing.assert_allclose(dycam * (zp / dycam[2]) - dc * (zp / dc[2]), dyc, rtol=1e-4, atol=1e-6)
        else:
            np.testing.assert_allclose(ox, ro, atol=1e-6)
            np.testing.assert_allclose(oy, ro, atol=1e-6)
            # the main ray and its x neighbour meet the focal plane one focal-plane pixel apart
            f = cam.focal_length
            pf = oc + dc * (-f / dc[2])
            pfx = oc + dxcam * (-f / dxcam[2])
            pfy = oc + dycam * (-f / dycam[2])
            zp = _image_plane_z(cam)
            np.testing.assert_allclose((pfx - pf)[:2], (dxc * (-f / zp))[:2], rtol=2e-3, atol=1e-6)
            np.testing.assert_allclose((pfy - pf)[:2], (dyc * (-f / zp))[:2], rtol=2e-3, atol=1e-6)


def _image_plane_z(cam):
    r = cam.raster_to_camera
    return float(r.t.z)  # raster (x, y, 0) -> camera z of the image plane


def test_ortho_and_env_camera_differentials():
    wl = _plane_scene(material(abi.MAT_DIFFUSE))
    cam = make_camera((0.3, 1.2, 1.5), (0.0, 0.0, -1.0), (0, 1, 0), np.radians(50.0), 32, 24,
                      camera_type=abi.CAMERA_ORTHO, ortho_zoom=2.0)
    wl.camera = cam
    o = OracleScene.from_workload(wl)
    R, t, dxc, dyc = _cam_frame(cam)
    ro, rd, ox, oy, dx, dy = o.camera_ray_diff(7, 9, 1).astype(np.float64)
    np.testing.assert_allclose(ox - ro, R @ dxc, atol=1e-6)
    np.testing.assert_allclose(oy - ro, R @ dyc, atol=1e-6)
    np.testing.assert_allclose(dx, rd, atol=1e-6)
    np.testing.assert_allclose(dy, rd, atol=1e-6)
    cam = make_camera((0.3, 1.2, 1.5), (0.0, 0.0, -1.0), (0, 1, 0), np.radians(50.0), 64, 32,
                      camera_type=abi.CAMERA_ENV)
    wl.camera = cam
    o = OracleScene.from_workload(wl)
    ro, rd, ox, oy, dx, dy = o.camera_ray_diff(10, 20, 1).astype(np.float64)
    # in camera space (this camera's frame is not orthonormal) one pixel is
    # 2 pi / 64 in azimuth and pi / 32 in polar angle
    R, t, dxc, dyc = _cam_frame(cam)
    Rinv = np.linalg.inv(R)
    unit = lambda a: a / np.linalg.norm(a)
    dc, dxcam, dycam = unit(Rinv @ rd), unit(Rinv @ dx), unit(Rinv @ dy)
    np.testing.assert_allclose(ox, ro, atol=1e-6)
    cos_x = np.dot(dc, dxcam)
    cos_y = np.dot(dc, dycam)
    th = np.arccos(np.clip(dc[1], -1, 1))
    np.testing.assert_allclose(np.arccos(np.clip(cos_y, -1, 1)), np.pi / 32, rtol=1e-3)
    np.testing.assert_allclose(1 - cos_x, (1 - np.cos(2 * np.pi / 64)) * np.sin(th) ** 2, rtol=2e-2)


def test_hit_uv_gradients_on_a_plane():
    """The floor y = 0 carries u = (x + 2) / 4, v = 1 - (z + 3) / 4 (the
    getter flips v).  The uv change per pixel is where the auxiliary rays
    meet the plane, mapped to uv."""
    wl = _plane_scene(material(abi.MAT_DIFFUSE, diffuse=np.zeros((4, 4, 3), np.float32) + 0.5))
    o = OracleScene.from_workload(wl)
    uvw = lambda p: np.array([(p[0] + 2.0) / 4.0, 1.0 - (p[2] + 3.0) / 4.0])
    n = 0
    for (x, y) in ((5, 20), (16, 18), (28, 22), (10, 15), (5, 2), (16, 4), (28, 6), (12, 9)):
        hd = o.hit_diff(x, y, 1)
        if hd is None:
            continue
        n += 1
        ro, rd, ox, oy, dx, dy = o.camera_ray_diff(x, y, 1).astype(np.float64)
        tx = -ox[1] / dx[1]
        ty = -oy[1] / dy[1]
        px, py = ox + tx * dx, oy + ty * dy
        p = hd["p"].astype(np.float64)
        np.testing.assert_allclose(uvw(p), hd["uv"], atol=1e-5)
        np.testing.assert_allclose(hd["dpdx"], px - p, atol=1e-5)
        np.testing.assert_allclose(hd["dpdy"], py - p, atol=1e-5)
        dudx, dvdx, dudy, dvdy = hd["g"]
        np.testing.assert_allclose([dudx, dvdx], uvw(px) - uvw(p), rtol=1e-3, atol=1e-6)
        np.testing.assert_allclose([dudy, dvdy], uvw(py) - uvw(p), rtol=1e-3, atol=1e-6)
        assert abs(dudx) > 1e-4 and abs(dvdy) > 1e-4
    assert n >= 4


def _panel_scene(alpha, mis=False):
    """A lit floor and, above it, a panel that (with alpha) may be cut away."""
    panel = (quad_uv((-0.8, 0.4, -1.6), (1.6, 0.0, 0.0), (0.0, 0.0, 1.2)), None)
    def lights(sb):
        sb.add_quad_light((-0.3, 1.5, -1.3), (0.6, 0, 0), (0, 0, 0.6), (8.0, 8.0, 8.0))
    extra = []
    if alpha is not False:
        pm = material(abi.MAT_DIFFUSE, diffuse=(0.7, 0.2, 0.2)) if alpha is None else \
            material(abi.MAT_DIFFUSE, diffuse=(0.7, 0.2, 0.2), alpha=alpha)
        extra = [(panel[0], pm)]
    cam = make_camera((0.0, 2.5, 1.0), (0.0, 0.0, -1.0), (0, 1, 0), np.radians(60.0), 24, 20)
    return _plane_scene(material(abi.MAT_DIFFUSE, diffuse=(0.6, 0.6, 0.6)), cam=cam, extra=extra, mis=mis,
                        lights=lights)


@pytest.mark.parametrize("mis", [False, True])
def test_alpha_maps_cut_and_keep_surfaces(mis):
    none, _ = OracleScene.from_workload(_panel_scene(False, mis)).render(1, 8)
    zero, _ = OracleScene.from_workload(_panel_scene(np.zeros((4, 4), np.float32), mis)).render(1, 8)
    opaque, _ = OracleScene.from_workload(_panel_scene(None, mis)).render(1, 8)
    one, _ = OracleScene.from_workload(_panel_scene(np.ones((4, 4), np.float32), mis)).render(1, 8)
    np.testing.assert_array_equal(zero, none)
    np.testing.assert_array_equal(one, opaque)
    assert np.abs(opaque - none).max() > 0.05  # the panel is in view and casts a shadow


def test_constant_bump_changes_the_frame_only_by_rounding():
    flat = np.full((4, 4), 0.3, np.float32)
    env = np.ones((8, 16, 4), np.float32)
    a, _ = OracleScene.from_workload(_plane_scene(material(abi.MAT_GLOSSY, diffuse=(0.5, 0.5, 0.5),
                                                           roughness=(0.3, 0.3)), env=env)).render(1, 4, fp64=True)
    b, _ = OracleScene.from_workload(_plane_scene(material(abi.MAT_GLOSSY, diffuse=(0.5, 0.5, 0.5),
                                                           roughness=(0.3, 0.3), bump=flat, bumpfactor=3.0),
                                                  env=env)).render(1, 4, fp64=True)
    np.testing.assert_allclose(b, a, rtol=1e-9, atol=1e-9)
    ramp = np.tile(np.linspace(0, 1, 16, dtype=np.float32), (16, 1))
    c, _ = OracleScene.from_workload(_plane_scene(material(abi.MAT_GLOSSY, diffuse=(0.5, 0.5, 0.5),
                                                           roughness=(0.3, 0.3), bump=ramp, bumpfactor=0.5),
                                                  env=env)).render(1, 4, fp64=True)
    assert np.abs(c - a).max() > 1e-3


def test_textured_workload_renders_and_textures_matter():
    wl = textured(width=48, height=40)
    f, _ = OracleScene.from_workload(wl).render(1, 2)
    assert np.isfinite(f).all() and f[..., :3].mean() > 0.05


def _lanczos_taps(center):
    """ResampleWeight in float64 (noor/cuda/image.cuh:41-98): first tap and the
    four normalised Lanczos (radius 2, tau 2) weights."""
    first = int(np.floor(center - 2.0 + 0.5))
    def L(x):
        x = abs(x)
        if x < 1e-5:
            return 1.0
        if x > 1:
            return 0.0
        x *= np.pi
        return (np.sin(x / 2) / (x / 2)) * (np.sin(x) / x)
    w = np.array([L((first + j + 0.5 - center) / 2.0) for j in range(4)])
    return first, w / w.sum()


@pytest.mark.parametrize("summed", [True, False])
def test_lanczos_resize_of_non_power_of_two_textures(summed):
    """Sides go to the nearest power of two (ties to the smaller): 12 -> 8,
    40 -> 32, 48 -> 32, 100 -> 128.  With the reference's summed passes
    (compat RESIZE_SUM) a constant image doubles and, at integer centres c,
    f(x) = x^2 along x gives (c - 1/2)^2 + 1/4 + 4 w0 + c^2 with w0 the outer
    tap weight (about 0.1058); composed, a constant stays and a texel equals
    the separable float64 filter."""
    h, w = 12, 100
    compat = abi.COMPAT_ALL if summed else 0
    const = np.full((h, w, 4), 0.3, np.float32)
    wl = _plane_scene(material(abi.MAT_DIFFUSE, diffuse=const))
    wl.options.compat = compat
    lv = OracleScene.from_workload(wl).texture_levels(0)
    assert lv[0].shape == (8, 128, 4)
    np.testing.assert_allclose(lv[0], 0.6 if summed else 0.3, rtol=2e-6)
    x = np.arange(w, dtype=np.float64)
    quad = np.broadcast_to((x * x)[None, :, None], (h, w, 4)).astype(np.float32)
    wl = _plane_scene(material(abi.MAT_DIFFUSE, diffuse=quad))
    wl.options.compat = compat
    got = OracleScene.from_workload(wl).texture_levels(0)[0].astype(np.float64)
    W, H = 128, 8
    _, wt = _lanczos_taps(10.0)
    w0 = wt[0]
    assert 0.10 < w0 < 0.11
    for X in (5, 40, 77, 120):
        for Y in (0, 3, 7):
            if summed:
                c = int((X + 0.5) * (w / W))
                if c - 2 < 0 or c + 1 > w - 1:
                    continue
                expect = (c - 0.5) ** 2 + 0.25 + 4 * w0 + c * c
            else:
                first, wx = _lanczos_taps((X + 0.5) * (w / W))
                xs = np.clip(first + np.arange(4), 0, w - 1).astype(np.float64)
                expect = float(np.dot(wx, xs * xs))  # constant along y: the vertical pass keeps it
            np.testing.assert_allclose(got[Y, X, 0], expect, rtol=2e-5)


@pytest.mark.parametrize("summed", [True, False])
def test_environment_is_resampled_like_textures(summed):
    """A 40 x 96 environment becomes 32 x 64 (nearest powers of two, ties to
    the smaller); a constant sky seen directly doubles under the summed passes."""
    env = np.full((40, 96, 4), 0.5, np.float32)
    cam = make_camera((0.0, 1.0, 0.0), (0.0, 5.0, 0.0), (0, 0, -1), np.radians(40.0), 8, 6)
    wl = _plane_scene(material(abi.MAT_DIFFUSE), cam=cam, env=env)
    wl.options.compat = abi.COMPAT_ALL if summed else 0
    wl.spec.lighting_type |= abi.LIGHT_SKY_MASK
    o = OracleScene.from_workload(wl)
    w, h, func, cdf = o.dist()
    assert (w, h) == (64, 32)
    f, _ = o.render(1, 1)
    np.testing.assert_allclose(f[..., :3], 1.0 if summed else 0.5, rtol=2e-6)


def test_payload_texture_validation():
    """Material textures are at most 32768 texels a side; the check runs
    before any device call."""
    from noor_amd.renderer import Renderer
    wl = _plane_scene(material(abi.MAT_DIFFUSE, diffuse=np.zeros((2, 32769, 3), np.float32)))
    with pytest.raises(abi.NoorError) as e:
        Renderer.from_workload(wl)
    assert e.value.status == abi.ERR_UNSUPPORTED
