"""GPU parity: the HIP kernels against the CPU oracle on identical inputs.

Bit-exact for hit queries (t, barycentrics, primitive id, instance); images
with identical per-pixel RNG streams within a mean relative error of 1e-4
(the tolerance the specification states for floating point).
"""
import ctypes as C

import numpy as np
import pytest

from noor_amd import abi, scenes
from noor_amd.renderer import Renderer, hits_to_numpy, rays_from_numpy
from oracle.binding import OracleScene
from tests.helpers import mean_rel_err, random_rays, vertex_rays

pytestmark = pytest.mark.gpu

TOL_IMAGE = 1e-4


@pytest.fixture(scope="module")
def cornell_small():
    wl = scenes.cornell(subdiv=5, width=96, height=80)
    return wl


def _both(wl):
    return Renderer.from_workload(wl), OracleScene.from_workload(wl)


@pytest.mark.parametrize("tri", [abi.TRI_WATERTIGHT, abi.TRI_WALD])
def test_trace_closest_bit_exact(cornell_small, tri):
    wl = cornell_small
    wl.options.tri_test = tri
    r, o = _both(wl)
    st = wl.payload.stats
    c = (np.array(st.scene_min) + np.array(st.scene_max)) / 2
    rays = random_rays(20000, c, 2.5 * st.scene_radius, seed=11)
    g = hits_to_numpy(r.trace(rays))
    h, _ = o.trace(rays)
    c_ = hits_to_numpy(h)
    assert (g[:, 6] > 0).sum() > 1000
    np.testing.assert_array_equal(g, c_)


def test_trace_any_bit_exact(cornell_small):
    wl = cornell_small
    r, o = _both(wl)
    st = wl.payload.stats
    c = (np.array(st.scene_min) + np.array(st.scene_max)) / 2
    rays = random_rays(20000, c, 2.5 * st.scene_radius, seed=12, tmax=np.float32(st.scene_radius))
    g = hits_to_numpy(r.trace(rays, any_hit=True))
    h, _ = o.trace(rays, any_hit=True)
    np.testing.assert_array_equal(g[:, 6], hits_to_numpy(h)[:, 6])


@pytest.mark.parametrize("any_hit", [False, True])
def test_trace_instanced_bit_exact(any_hit):
    """Two-level scene (24 meshes under a TLAS): the GPU's four-wide BVH
    returns the hits of the oracle's binary traversal of the payload BVH."""
    wl = scenes.procedural(width=16, height=16, n_meshes=24, target_tris=60000)
    r, o = _both(wl)
    # rays aimed near the instances' world positions (the scene is sparse)
    v = wl.payload.view
    T = np.stack([np.ctypeslib.as_array(C.cast(v.object_to_world_row[k], C.POINTER(C.c_float)),
                                        shape=(v.num_instances, 4))[:, 3] for k in range(3)], axis=1)
    rng = np.random.default_rng(21)
    tgt = T[rng.integers(0, v.num_instances, 30000)] + rng.normal(size=(30000, 3)) * 0.7
    org = tgt + rng.normal(size=(30000, 3)) * 20
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = rays_from_numpy(org.astype(np.float32), d.astype(np.float32))
    g = hits_to_numpy(r.trace(rays, any_hit=any_hit))
    h = hits_to_numpy(o.trace(rays, any_hit=any_hit)[0])
    if any_hit:
        np.testing.assert_array_equal(g[:, 6], h[:, 6])
    else:
        assert (g[:, 6] > 0).sum() > 10000
        np.testing.assert_array_equal(g, h)


def test_trace_vertex_targeted_bit_exact(cornell_small):
    wl = cornell_small
    r, o = _both(wl)
    rays = vertex_rays(wl.payload.arrays()["vertices"], 20000, seed=13)
    g = hits_to_numpy(r.trace(rays))
    h, _ = o.trace(rays)
    np.testing.assert_array_equal(g, hits_to_numpy(h))


@pytest.mark.parametrize("rng_kind", [abi.RNG_PCG32, abi.RNG_TEA_LCG])
def test_megakernel_image_matches_oracle(cornell_small, rng_kind):
    wl = cornell_small
    wl.options.rng = rng_kind
    wl.options.tri_test = abi.TRI_WATERTIGHT
    r, o = _both(wl)
    r.render(spp=4)
    g = r.film()
    c, _ = o.render(1, 4)
    err = mean_rel_err(g, c)
    print(f"mean rel err {err:.3e}, identical pixels {np.mean(np.all(g == c, axis=-1)):.4f}")
    assert err <= TOL_IMAGE


@pytest.mark.parametrize("graph", [True, False])
def test_wavefront_matches_megakernel(cornell_small, graph, monkeypatch):
    wl = cornell_small
    wl.options.rng = abi.RNG_PCG32
    wl.options.tri_test = abi.TRI_WATERTIGHT
    if not graph:
        monkeypatch.setenv("NOOR_NO_GRAPH", "1")
    a = Renderer.from_workload(wl)
    a.render(spp=3)
    b = Renderer.from_workload(wl)
    b.render_wavefront(spp=3)
    fa, fb = a.film(), b.film()
    err = mean_rel_err(fb, fa)
    print(f"wavefront vs megakernel mean rel err {err:.3e}, identical {np.mean(np.all(fa == fb, axis=-1)):.4f}")
    assert err <= TOL_IMAGE
    sa, sb = a.stats(), b.stats()
    assert sa["extension_rays"] == sb["extension_rays"]
    assert sa["shadow_rays"] == sb["shadow_rays"]


@pytest.mark.parametrize("capacity", [0, 1024])
def test_wavefront_counters_equal_megakernel(cornell_small, capacity):
    """Both renderers trace the same rays through the same traversal code, so
    with counting on every counter agrees: rays, hits, node visits, triangle
    tests and instance entries, the rays the wavefront's drain kernel traced
    (the last paths of each group) included."""
    wl = cornell_small
    wl.options.stats = 1
    bounces = wl.spec.bounces
    wl.spec.bounces = 8  # paths outlive a replay of iterations, so groups end in the drain
    try:
        a = Renderer.from_workload(wl)
        a.render(spp=3)
        b = Renderer.from_workload(wl)
        b.set_wavefront_params(queue_capacity=capacity)
        b.render_wavefront(spp=3)
    finally:
        wl.options.stats = 0
        wl.spec.bounces = bounces
    np.testing.assert_array_equal(b.film(), a.film())
    sa, sb = a.stats(), b.stats()
    print({k: (sa[k], sb[k]) for k in sa})
    for k in ("extension_rays", "shadow_rays", "extension_hits", "node_visits", "tri_tests", "instance_entries",
              "shadow_node_visits", "shadow_tri_tests", "shadow_instance_entries"):
        assert sa[k] == sb[k], k
    assert sa["drain_extension_rays"] == 0 and sa["drain_shadow_rays"] == 0  # the megakernel has no drain
    assert 0 < sb["drain_extension_rays"] < sb["extension_rays"]
    assert sb["drain_shadow_rays"] <= sb["shadow_rays"]
    assert 0 < sb["extension_hits"] < sb["extension_rays"]


@pytest.mark.parametrize("capacity", [0, 1024])
def test_wavefront_pixel_major_items_odd_sizes(capacity):
    """With one view the items of a group run pixel by pixel (a pixel's
    samples are consecutive items), and the finish kernel reads them back in
    blocks of 256 pixels and 8 layers: a frame of 37 x 29 pixels (no multiple
    of the block) at 11 spp (no multiple of 8), continuing a 2-spp render,
    equals the megakernel, with the pool holding every item or far fewer."""
    wl = scenes.cornell(subdiv=3, width=37, height=29)
    a = Renderer.from_workload(wl)
    a.render(spp=2, frame=1)
    a.render(spp=11)
    b = Renderer.from_workload(wl)
    b.set_wavefront_params(queue_capacity=capacity)
    b.render_wavefront(spp=2, frame=1)
    b.render_wavefront(spp=11)
    assert a.frame == b.frame == 14
    np.testing.assert_array_equal(b.film(), a.film())


def test_wavefront_small_queue_chunks(cornell_small):
    """Queue capacity below the pixel count: the frame is rendered in chunks."""
    wl = cornell_small
    a = Renderer.from_workload(wl)
    a.render(spp=2)
    b = Renderer.from_workload(wl)
    b.set_wavefront_params(queue_capacity=1024, block=128)
    b.render_wavefront(spp=2)
    assert mean_rel_err(b.film(), a.film()) <= TOL_IMAGE


@pytest.mark.parametrize("name", ["procedural", "terrain"])
def test_large_scene_megakernel_and_wavefront_vs_oracle(name):
    """The 1M-triangle Hosek-sky scene and the 5M-triangle env-lit terrain at
    reduced resolution: megakernel and wavefront against the oracle."""
    fn = getattr(scenes, name)
    wl = fn(width=160, height=90, spp=2)
    r = Renderer.from_workload(wl)
    r.render(spp=2)
    g = r.film()
    w = Renderer.from_workload(wl)
    w.render_wavefront(spp=2)
    c, _ = OracleScene.from_workload(wl).render(1, 2)
    e1, e2 = mean_rel_err(g, c), mean_rel_err(w.film(), c)
    print(f"{name}: megakernel {e1:.3e} wavefront {e2:.3e}")
    assert e1 <= TOL_IMAGE and e2 <= TOL_IMAGE


@pytest.mark.parametrize("mis", [False, True])
def test_secondary_materials_vs_oracle(mis):
    """Shadow catcher, rough glass, substrate and clear coat (the rest of the
    reference's material factory) under an HDR environment: megakernel
    against the oracle and the wavefront renderer against the megakernel
    (with MIS a path's two shadow rays are one task of the shadow kernel)."""
    wl = scenes.showcase(mis=mis)
    r = Renderer.from_workload(wl)
    r.render(spp=4)
    g = r.film()
    c, _ = OracleScene.from_workload(wl).render(1, 4)
    err = mean_rel_err(g, c)
    print(f"mis={mis}: megakernel vs oracle {err:.3e}, identical {np.mean(np.all(g == c, axis=-1)):.4f}")
    assert err <= TOL_IMAGE
    w = Renderer.from_workload(wl)
    w.render_wavefront(spp=4)
    np.testing.assert_array_equal(w.film(), g)


@pytest.mark.parametrize("mis,sky,compat", [(False, True, abi.COMPAT_ALL), (True, True, abi.COMPAT_ALL),
                                             (False, False, abi.COMPAT_ALL), (True, True, 0)])
def test_all_light_types_vs_oracle(mis, sky, compat):
    """Sphere and disk area lights, point, spot and distant lights (and the
    environment when `sky`), with the reference's quirks and with all of them
    corrected: megakernel against the oracle and the wavefront renderer
    against the megakernel."""
    wl = scenes.lights(mis=mis, sky=sky, compat=compat)
    r = Renderer.from_workload(wl)
    r.render(spp=4)
    g = r.film()
    c, _ = OracleScene.from_workload(wl).render(1, 4)
    err = mean_rel_err(g, c)
    print(f"mis={mis} sky={sky} compat={compat}: megakernel vs oracle {err:.3e}, identical {np.mean(np.all(g == c, axis=-1)):.4f}")
    assert err <= TOL_IMAGE
    assert float(g[..., :3].mean()) > 0.0
    w = Renderer.from_workload(wl)
    w.render_wavefront(spp=4)
    np.testing.assert_array_equal(w.film(), g)
