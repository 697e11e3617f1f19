"""GPU: multi-view batches and the wavefront autotuner.

A batch renders several cameras of one scene in one wavefront pass; each
view's film must be identical to rendering that camera alone with the
megakernel (same frames, same RNG streams)."""
import numpy as np
import pytest

from noor_amd import abi, scenes
from noor_amd.renderer import Renderer
from noor_amd.scene import make_camera

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wl():
    return scenes.cornell(subdiv=4, width=64, height=48)


def _cams():
    return [
        make_camera((0.0, 1.0, 3.4), (0.0, 1.0, 0.0), w=64, h=48),
        make_camera((0.8, 1.2, 3.0), (0.0, 0.8, 0.0), w=40, h=40),
        make_camera((-0.6, 1.5, 2.8), (0.1, 0.9, 0.0), w=33, h=17),
    ]


def test_batch_views_equal_single_megakernel_renders(wl):
    cams = _cams()
    spps = [3, 1, 2]
    r = Renderer.from_workload(wl)
    r.render_batch([(c, 1, s, 0) for c, s in zip(cams, spps)])
    films = [r.batch_film(i) for i in range(len(cams))]
    for i, (c, s) in enumerate(zip(cams, spps)):
        m = Renderer(wl.payload, c, wl.spec, wl.options, wl.hosek)
        m.render(spp=s, frame=1)
        ref = m.film()
        same = np.mean(np.all(films[i] == ref, axis=-1))
        print(f"view {i}: identical pixels {same:.4f}")
        np.testing.assert_array_equal(films[i], ref)


def test_batch_continues_running_mean(wl):
    cams = _cams()[:2]
    r = Renderer.from_workload(wl)
    r.render_batch([(c, 1, 2, 0) for c in cams])
    r.render_batch([(c, 3, 2, 0) for c in cams])
    for i, c in enumerate(cams):
        m = Renderer(wl.payload, c, wl.spec, wl.options, wl.hosek)
        m.render(spp=4, frame=1)
        np.testing.assert_array_equal(r.batch_film(i), m.film())


def test_batch_with_small_pool_and_lds_split(wl):
    """A pool far smaller than the batch, 128-lane blocks and a 4-entry LDS
    stack (the rest of the stack in device memory) change nothing."""
    cams = _cams()
    r = Renderer.from_workload(wl)
    r.set_wavefront_params(queue_capacity=4096, block=128)
    r.set_stack_lds(4)
    r.render_batch([(c, 1, 2, 0) for c in cams])
    for i, c in enumerate(cams):
        m = Renderer(wl.payload, c, wl.spec, wl.options, wl.hosek)
        m.render(spp=2, frame=1)
        np.testing.assert_array_equal(r.batch_film(i), m.film())


@pytest.mark.parametrize("mis", [False, True])
def test_textured_batch_with_mis(mis):
    """Batches of the textured scene (per-view camera differentials, bump,
    alpha, textures), with and without MIS."""
    tw = scenes.textured(mis=mis)
    cams = [make_camera((0.0, 0.0, 0.95), (0.0, -0.15, -1.0), fov=np.radians(60.0), w=48, h=40),
            make_camera((0.3, 0.2, 0.8), (-0.2, -0.3, -1.0), fov=np.radians(45.0), w=32, h=24, lens_radius=0.03),
            make_camera((-0.2, 0.1, 0.9), (0.1, -0.2, -1.0), w=20, h=30, camera_type=abi.CAMERA_ORTHO)]
    r = Renderer.from_workload(tw)
    r.render_batch([(c, 1, 2, 0) for c in cams])
    for i, c in enumerate(cams):
        m = Renderer(tw.payload, c, tw.spec, tw.options, tw.hosek)
        m.render(spp=2, frame=1)
        np.testing.assert_array_equal(r.batch_film(i), m.film())


@pytest.mark.parametrize("mis", [False, True])
def test_drain_of_textured_paths_with_mis(mis):
    """The last paths of each group finished by the drain kernel (eight
    bounces and a pool smaller than the frame, so paths are in flight when
    the last items start), in the textured instantiation with and without
    MIS: identical to the megakernel, and the drain traced some rays."""
    tw = scenes.textured(mis=mis, bounces=8)
    tw.options.stats = 1
    r = Renderer.from_workload(tw)
    r.set_wavefront_params(queue_capacity=2048)
    r.render_wavefront(spp=3, frame=1)
    st = r.stats()
    m = Renderer.from_workload(tw)
    m.render(spp=3, frame=1)
    np.testing.assert_array_equal(r.film(), m.film())
    sm = m.stats()
    print({k: (st[k], sm[k]) for k in ("extension_rays", "shadow_rays", "drain_extension_rays", "drain_shadow_rays")})
    assert st["extension_rays"] == sm["extension_rays"] and st["shadow_rays"] == sm["shadow_rays"]
    assert st["drain_extension_rays"] > 0


def test_autotune_picks_a_measured_setting(wl):
    cams = _cams()
    r = Renderer.from_workload(wl)
    res = r.autotune([(c, 1, 1, 0) for c in cams])
    # pools from 2^20 up to the first that holds every item (these few views:
    # 2^20 only), three block sizes, the LDS stack splits the BVH allows
    assert len(res) >= 9
    assert all(ms > 0 for *_, ms in res)
    blocks = {b for _, b, _, _ in res}
    assert blocks == {64, 128, 256}
    assert {q for q, *_ in res} == {1 << 20}
    # the tuned context still renders the batch identically
    r.render_batch([(c, 1, 1, 0) for c in cams])
    m = Renderer(wl.payload, cams[0], wl.spec, wl.options, wl.hosek)
    m.render(spp=1, frame=1)
    np.testing.assert_array_equal(r.batch_film(0), m.film())


def test_batch_rejects_bad_views(wl):
    r = Renderer.from_workload(wl)
    c = _cams()[0]
    with pytest.raises(abi.NoorError):
        r.render_batch([(c, 0, 1, 0)])  # frame must be >= 1
    with pytest.raises(abi.NoorError):
        r.render_batch([(c, 1, 0, 0)])  # spp must be >= 1
    with pytest.raises(abi.NoorError):
        r.set_wavefront_params(queue_capacity=(1 << 28) + 1)  # above the hit record's queue positions
    with pytest.raises(abi.NoorError):
        r.set_wavefront_params(queue_capacity=32)  # below one wave
