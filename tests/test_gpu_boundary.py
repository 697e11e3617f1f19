"""GPU: the drop-in boundary beyond noor_render.

- noor_update_camera / noor_update_spec / noor_update_hosek followed by a
  render from frame 1 (the reference resets the frame on every dirty update,
  noor/source/scene.cpp:124-144; a new sky is re-cooked, re-baked and its
  distribution rebuilt, noor/cuda/skydome.cuh:58-71) against the oracle;
- noor_get_lookat (the autofocus point, noor/cuda/path_tracer.cu:40-42,97-99)
  against the oracle's first hit of the centre pixel's last sample;
- noor_debug_skydome (noor/cuda/path_tracer.cu:102-130) against the bilinear
  lookup of a float64 statement of the baked sky;
- the wavefront renderer with zero bounces and a pool smaller than the frame.
"""
import ctypes as C

import numpy as np
import pytest

from noor_amd import abi, scenes
from noor_amd._lib import lib
from noor_amd.renderer import Renderer
from noor_amd.scene import cook_hosek, make_camera, make_spec
from oracle.binding import OracleScene
from tests.helpers import mean_rel_err
from tests.test_oracle_sky import sky64

pytestmark = pytest.mark.gpu

TOL_IMAGE = 1e-4


@pytest.fixture(scope="module")
def sky_wl():
    """A small open scene of the 1M-triangle kind (spheres, tori, heightfield
    patches, Lambert / copper / glass) under the Hosek sky."""
    return scenes.procedural(width=64, height=48, spp=2, n_meshes=24, target_tris=60000)


def _vs_oracle(r, wl, hosek, cam, spec, spp):
    o = OracleScene(wl.payload, cam, spec, wl.options, hosek)
    c, _ = o.render(1, spp)
    g = r.film()
    return mean_rel_err(g, c), g, c


def test_updates_then_render_match_oracle(sky_wl):
    wl = sky_wl
    r = Renderer.from_workload(wl)
    r.render(spp=2)
    assert r.frame == 3
    err, g, _ = _vs_oracle(r, wl, wl.hosek, wl.camera, wl.spec, 2)
    assert err <= TOL_IMAGE
    # a new camera: the caller restarts at frame 1, which overwrites the film
    cam = make_camera((30.0, 60.0, 120.0), (50.0, 30.0, 50.0), w=64, h=48, fov=np.radians(50.0))
    r.update_camera(cam)
    r.render(spp=2)
    err, g2, c2 = _vs_oracle(r, wl, wl.hosek, cam, wl.spec, 2)
    print(f"after update_camera: {err:.3e}")
    assert err <= TOL_IMAGE and not np.array_equal(g2, g)
    # a new spec (fewer bounces, MIS on)
    spec = make_spec(wl.payload, bounces=3, rr=2, skydome_type=abi.SKY_PHYSICAL, mis=True)
    r.update_spec(spec)
    r.render(spp=2)
    err, g3, _ = _vs_oracle(r, wl, wl.hosek, cam, spec, 2)
    print(f"after update_spec: {err:.3e}")
    assert err <= TOL_IMAGE and not np.array_equal(g3, g2)
    # a new sky: turbidity 6, coloured ground, sun lower in another direction;
    # the bake and its importance-sampling distribution follow it
    hk = cook_hosek(6.0, (0.2, 0.4, 0.6), np.float32(np.radians(70.0)), np.float32(np.radians(200.0)))
    r.update_hosek(hk)
    r.render(spp=2)
    err, g4, c4 = _vs_oracle(r, wl, hk, cam, spec, 2)
    print(f"after update_hosek: {err:.3e}")
    assert err <= TOL_IMAGE and not np.array_equal(g4, g3)
    # a fresh context on the final state renders the same film
    f = Renderer(wl.payload, cam, spec, wl.options, hk)
    f.render(spp=2)
    np.testing.assert_array_equal(f.film(), g4)
    # continuing the running mean (frames 3-4) matches the oracle's frames 1-4
    r.render(spp=2)
    o = OracleScene(wl.payload, cam, spec, wl.options, hk)
    c, _ = o.render(1, 4)
    assert mean_rel_err(r.film(), c) <= TOL_IMAGE


def _lookat_expect(o, cam, frame):
    cx, cy = cam.center % cam.w, cam.center // cam.w
    h = o.hit_diff(cx, cy, frame)
    if h is None:
        return np.zeros(4, np.float32)
    return np.array([h["p"][0], h["p"][1], h["p"][2], 1.0], np.float32)


def test_lookat_matches_oracle_first_hit(sky_wl):
    wl = sky_wl
    # a camera 12 units in front of the third instance: its centre pixel hits it
    v = wl.payload.view
    pos = np.array([v.object_to_world_row[k][2].w for k in range(3)], np.float64)
    cam = make_camera(tuple(pos + np.array([0.0, 3.0, 12.0])), tuple(pos), w=64, h=48)
    r = Renderer(wl.payload, cam, wl.spec, wl.options, wl.hosek)
    o = OracleScene(wl.payload, cam, wl.spec, wl.options, wl.hosek)
    r.render(spp=3)
    got = r.lookat()
    want = _lookat_expect(o, cam, 3)
    assert want[3] == 1.0
    np.testing.assert_array_equal(got, want)
    # a camera whose centre pixel sees the sky: no hit, (0, 0, 0, 0)
    cam = make_camera((50.0, 200.0, 50.0), (50.0, 400.0, 51.0), w=64, h=48)
    r.update_camera(cam)
    o.set_camera(cam)
    r.render(spp=1)
    np.testing.assert_array_equal(r.lookat(), np.zeros(4, np.float32))
    np.testing.assert_array_equal(_lookat_expect(o, cam, 1), np.zeros(4, np.float32))


def _bilinear64(tex64, w, h, u, v):
    """CUDA-style clamp-addressed bilinear lookup at normalised (u, v), float64."""
    x = u * tex64.shape[1] - 0.5
    y = v * tex64.shape[0] - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    ax, ay = x - x0, y - y0
    W, H = tex64.shape[1], tex64.shape[0]
    i0 = np.clip(x0.astype(int), 0, W - 1)
    i1 = np.clip(x0.astype(int) + 1, 0, W - 1)
    j0 = np.clip(y0.astype(int), 0, H - 1)
    j1 = np.clip(y0.astype(int) + 1, 0, H - 1)
    t = lambda j, i: tex64[j, i]
    return ((1 - ay)[..., None] * ((1 - ax)[..., None] * t(j0, i0) + ax[..., None] * t(j0, i1)) +
            ay[..., None] * ((1 - ax)[..., None] * t(j1, i0) + ax[..., None] * t(j1, i1)))


def test_debug_skydome_physical_sky(sky_wl):
    """A 512 x 256 view of the 2048^2 bake: pixel (x, y) reads between texel
    columns 4x-1, 4x and rows 8y-1, 8y... of sin(theta) sky(direction)."""
    wl = sky_wl
    w, h = 512, 256
    cam = make_camera((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), w=w, h=h)
    r = Renderer(wl.payload, cam, wl.spec, wl.options, wl.hosek)
    frame = C.c_uint32(1)
    assert lib().noor_debug_skydome(r._h, C.byref(frame)) == 0 and frame.value == 2
    got = r.film().astype(np.float64)
    # the texels the lookups touch, in float64
    Wt = Ht = 2048
    xs = np.arange(Wt)
    ys = np.unique(np.clip(np.concatenate([8 * np.arange(h) - 1, 8 * np.arange(h)]), 0, Ht - 1))
    X, Y = np.meshgrid(xs, ys)
    th, ph = np.pi * Y / Ht, 2 * np.pi * X / Wt
    d = np.stack([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)], -1).reshape(-1, 3)
    tex = np.zeros((Ht, Wt, 4))
    tex[ys] = np.concatenate([np.sin(th)[..., None] * sky64(wl.hosek, d).reshape(len(ys), Wt, 3),
                              np.sin(th)[..., None]], -1)
    # binary32 rounds the horizon row below the horizon (fl(pi) > pi)
    tex[Ht // 2:, :, :3] = 0.0
    U, V = np.meshgrid(np.arange(w) / w, np.arange(h) / h)
    want = _bilinear64(tex, Wt, Ht, U, V)
    sun = np.array([wl.hosek.solar_dir.x, wl.hosek.solar_dir.y, wl.hosek.solar_dir.z])
    dpix = np.stack([np.sin(np.pi * V) * np.cos(2 * np.pi * U), np.cos(np.pi * V),
                     np.sin(np.pi * V) * np.sin(2 * np.pi * U)], -1)
    far = np.arccos(np.clip(dpix @ sun, -1, 1)) > 0.05
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-6)
    print(f"debug skydome: max rel err {rel[far].max():.3e}")
    assert rel[far].max() < 5e-5
    assert np.all(got[h // 2 + 1:, :, :3] == 0.0)
    # MIS on: the importance-sample dots are red, every other pixel keeps the sky
    spec = make_spec(wl.payload, bounces=2, rr=5, skydome_type=abi.SKY_PHYSICAL, mis=True)
    r.update_spec(spec)
    frame = C.c_uint32(1)
    assert lib().noor_debug_skydome(r._h, C.byref(frame)) == 0
    g2 = r.film().astype(np.float64)
    red = np.all(g2 == [1.0, 0.0, 0.0, 1.0], axis=-1)
    assert 0 < red.sum() < red.size
    np.testing.assert_array_equal(g2[~red], got[~red])
    # the dots follow the sky's luminance: nearly all above the horizon
    assert red[: h // 2 + 1].sum() >= 0.99 * red.sum()


def test_debug_skydome_hdr_environment():
    wl = scenes.lights(width=48, height=40)
    cam = make_camera((0.0, 0.0, 0.95), (0.0, -0.1, -1.0), w=96, h=80)
    r = Renderer(wl.payload, cam, wl.spec, wl.options, wl.hosek)
    frame = C.c_uint32(1)
    assert lib().noor_debug_skydome(r._h, C.byref(frame)) == 0
    env = scenes.env_map(128, 64, 6 + 3, lobes=2, sigma_deg=8.0, peak=20.0).astype(np.float64)
    # the lookup's texel coordinates are binary32 (u = x / w, u W - 1/2, their
    # fractions): near u W = 100 they carry ~4e-6 of rounding, which the
    # texel contrast of the log-normal map turns into ~1e-5 of relative
    # difference; with the coordinates rounded as binary32 the blend agrees to
    # float precision
    f32 = np.float32
    X, Y = np.meshgrid(np.arange(96, dtype=np.float32), np.arange(80, dtype=np.float32))
    U, V = X / f32(96), Y / f32(80)
    xs, ys = U * f32(128) - f32(0.5), V * f32(64) - f32(0.5)
    want = _bilinear64(env, 128, 64, (xs.astype(np.float64) + 0.5) / 128, (ys.astype(np.float64) + 0.5) / 64)
    np.testing.assert_allclose(r.film(), want, rtol=2e-6, atol=1e-7)
    # and against the exact coordinates, the rounding bound above
    want64 = _bilinear64(env, 128, 64, np.arange(96)[None, :] / 96 + 0 * Y, np.arange(80)[:, None] / 80 + 0 * X)
    np.testing.assert_allclose(r.film(), want64, rtol=5e-5)


def test_wavefront_zero_bounces_small_pool():
    """bounces = 0: every sample is the (clamped) zero radiance; a pool of 256
    slots for 64 x 64 pixels must still drain (slots are reused)."""
    wl = scenes.cornell(subdiv=3, width=64, height=64)
    wl.spec.bounces = 0
    a = Renderer.from_workload(wl)
    a.render(spp=3)
    b = Renderer.from_workload(wl)
    b.set_wavefront_params(queue_capacity=256, block=64)
    b.render_wavefront(spp=3)
    np.testing.assert_array_equal(b.film(), a.film())
    f = a.film()
    assert np.all(f[..., :3] == 0.0) and np.all(f[..., 3] == 1.0)
