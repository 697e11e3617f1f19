"""GPU: the image-texture pipeline (mip-mapped lookups at the ray-differential
footprint, bump maps, alpha-cut surfaces, differentials through glossy
bounces) against the oracle, and the wavefront renderer against the
megakernel.  Images within the parity tolerance of the other image tests
(mean relative error 1e-4); the wavefront film equal to the megakernel's bit
for bit; hit queries through alpha-cut surfaces bit-exact."""
import numpy as np
import pytest

from noor_amd import abi, scenes
from noor_amd.renderer import Renderer, hits_to_numpy, rays_from_numpy
from oracle.binding import OracleScene
from tests.helpers import mean_rel_err

pytestmark = pytest.mark.gpu

TOL_IMAGE = 1e-4


@pytest.mark.parametrize("lens,mis,tri", [(0.0, False, abi.TRI_WATERTIGHT), (0.04, False, abi.TRI_WATERTIGHT),
                                          (0.0, True, abi.TRI_WATERTIGHT), (0.0, False, abi.TRI_WALD)])
def test_textured_megakernel_and_wavefront_vs_oracle(lens, mis, tri):
    wl = scenes.textured(lens_radius=lens, mis=mis)
    wl.options.tri_test = tri
    r = Renderer.from_workload(wl)
    r.render(spp=4)
    g = r.film()
    c, _ = OracleScene.from_workload(wl).render(1, 4)
    err = mean_rel_err(g, c)
    print(f"lens={lens} mis={mis} tri={tri}: megakernel vs oracle {err:.3e}, "
          f"identical {np.mean(np.all(g == c, axis=-1)):.4f}", flush=True)
    assert err <= TOL_IMAGE
    assert float(g[..., :3].mean()) > 0.05
    w = Renderer.from_workload(wl)
    w.render_wavefront(spp=4)
    np.testing.assert_array_equal(w.film(), g)


@pytest.mark.parametrize("compat", [abi.COMPAT_ALL, 0])
def test_resized_textures_vs_oracle(compat):
    """48x48 and 24x24 images and a 96x40 environment, resampled to 32x32,
    16x16 and 64x32 at creation (the reference's summed Lanczos passes under
    compat, composed passes without): the device's copies and the oracle's
    agree, so the films do."""
    wl = scenes.textured(tex=48, compat=compat, env_size=(96, 40))
    r = Renderer.from_workload(wl)
    r.render(spp=2)
    g = r.film()
    c, _ = OracleScene.from_workload(wl).render(1, 2)
    err = mean_rel_err(g, c)
    print(f"compat={compat}: megakernel vs oracle {err:.3e}, identical {np.mean(np.all(g == c, axis=-1)):.4f}",
          flush=True)
    assert err <= TOL_IMAGE


@pytest.mark.parametrize("camera", [abi.CAMERA_ORTHO, abi.CAMERA_ENV])
def test_textured_other_cameras_vs_oracle(camera):
    """Orthographic (auxiliary origins one pixel apart) and environment
    (auxiliary directions one pixel apart in angle) camera differentials."""
    wl = scenes.textured(camera_type=camera)
    r = Renderer.from_workload(wl)
    r.render(spp=2)
    g = r.film()
    c, _ = OracleScene.from_workload(wl).render(1, 2)
    err = mean_rel_err(g, c)
    print(f"camera={camera}: megakernel vs oracle {err:.3e}, identical {np.mean(np.all(g == c, axis=-1)):.4f}",
          flush=True)
    assert err <= TOL_IMAGE and float(g[..., :3].mean()) > 0.0
    w = Renderer.from_workload(wl)
    w.render_wavefront(spp=2)
    np.testing.assert_array_equal(w.film(), g)


def test_alpha_cut_hit_queries_bit_exact():
    """Closest and any hit through the alpha-cut panel of the textured scene."""
    wl = scenes.textured()
    r, o = Renderer.from_workload(wl), OracleScene.from_workload(wl)
    rng = np.random.default_rng(3)
    # rays from the camera side through the panel (x in [-0.55, -0.05], y in [-0.45, 0.15], z = -0.6)
    n = 20000
    tgt = np.stack([rng.uniform(-0.6, 0.0, n), rng.uniform(-0.5, 0.2, n), np.full(n, -0.6)], -1)
    org = np.stack([rng.uniform(-0.3, 0.3, n), rng.uniform(-0.2, 0.2, n), np.full(n, 0.8)], -1)
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rr = rays_from_numpy(org, d)
    g = hits_to_numpy(r.trace(rr))
    h, _ = o.trace(rr)
    np.testing.assert_array_equal(g, hits_to_numpy(h))
    rr = rays_from_numpy(org, d, np.float32(1.45))  # just past the panel
    ga = hits_to_numpy(r.trace(rr, any_hit=True))
    ha, _ = o.trace(rr, any_hit=True)
    np.testing.assert_array_equal(ga[:, 6], hits_to_numpy(ha)[:, 6])
    frac = float(np.mean(ga[:, 6] > 0))
    print(f"any-hit fraction through the panel {frac:.3f}", flush=True)
    assert 0.05 < frac < 0.95  # some rays pass the holes, some hit the panel


def test_adjoint_rejects_textured_scenes():
    """The adjoint has no texture derivatives: a textured scene is refused
    before any buffer is read."""
    import ctypes as C
    from noor_amd._lib import lib
    wl = scenes.textured()
    wl.spec.mis = 0
    wl.options.compat = 0
    r = Renderer.from_workload(wl)
    dummy = np.zeros(4, np.float32)
    p = dummy.ctypes.data_as(C.c_void_p)
    st = lib().noor_adjoint(r._h, 1, p, p, None, None, 0, None)
    assert st == abi.ERR_UNSUPPORTED
