"""GPU: the benchmark configurations at their stated sizes against the oracle.

- C0: the Cornell box with the 6x subdivided displaced icosphere (81,944
  triangles), 512 x 512, 16 spp, 5 bounces, RR from bounce 3: megakernel
  against the FP32 oracle (mean relative error <= 1e-4), wavefront against the
  megakernel;
- C0 against the FP64 oracle at 256 spp on the same RNG streams, with the
  RMSE stated (a 96 x 96 crop through the sphere and the light);
- C3: 48 views of the 1M-triangle Hosek-sky scene at uneven sizes
  ({256^2, 512^2, 1280 x 720, 1920 x 1080}) rendered in one wavefront batch
  after autotuning, at one sample per view: every view equals the megakernel's
  render of that camera, and two full views (one of them 1920 x 1080) match
  the oracle.
"""
import numpy as np
import pytest

from noor_amd import scenes
from noor_amd.renderer import Renderer
from noor_amd.scene import make_camera
from oracle.binding import OracleScene
from tests.helpers import mean_rel_err

pytestmark = pytest.mark.gpu

TOL_IMAGE = 1e-4


@pytest.fixture(scope="module")
def c0():
    wl = scenes.cornell()  # subdivision 6, 512 x 512, 16 spp, 5 bounces, rr 3
    assert wl.payload.view.num_triangles == 81944
    assert (wl.camera.w, wl.camera.h, wl.spp, wl.spec.bounces, wl.spec.rr) == (512, 512, 16, 5, 3)
    return wl


def test_c0_full_size_megakernel_and_wavefront(c0):
    wl = c0
    r = Renderer.from_workload(wl)
    r.render(spp=wl.spp)
    g = r.film()
    c, cnt = OracleScene.from_workload(wl).render(1, wl.spp)
    err = mean_rel_err(g, c)
    same = float(np.mean(np.all(g == c, axis=-1)))
    print(f"C0 megakernel vs oracle: mean rel err {err:.3e}, identical pixels {same:.4f}, "
          f"{cnt.extension_rays + cnt.shadow_rays} rays")
    assert err <= TOL_IMAGE
    w = Renderer.from_workload(wl)
    w.render_wavefront(spp=wl.spp)
    np.testing.assert_array_equal(w.film(), g)


def test_c0_fp64_oracle_rmse(c0):
    """FP32 GPU against the FP64 oracle on identical RNG streams at 256 spp.
    Paths agree until a binary32 and a binary64 decision part ways (a
    Russian-roulette or lobe threshold, a grazing hit), so the error is the
    precision of the estimator, not its variance."""
    wl = c0
    spp = 256
    r = Renderer.from_workload(wl)
    r.render(spp=spp)
    g = r.film()[..., :3].astype(np.float64)
    rect = (208, 288, 304, 384)  # x0, y0, x1, y1
    f64, _ = OracleScene.from_workload(wl).render(1, spp, rect=rect, fp64=True)
    x0, y0, x1, y1 = rect
    crop = g[y0:y1, x0:x1]
    ref = f64[..., :3]
    rmse = float(np.sqrt(np.mean((crop - ref) ** 2)))
    rel_rmse = rmse / float(np.sqrt(np.mean(ref ** 2)))
    mre = mean_rel_err(crop, ref)
    print(f"C0 {spp} spp, crop {rect}: RMSE vs FP64 oracle {rmse:.3e} (relative {rel_rmse:.3e}), "
          f"mean rel err {mre:.3e}, mean radiance {ref.mean():.4f}")
    assert rel_rmse < 1e-3
    assert mre < 1e-4


@pytest.fixture(scope="module")
def c3():
    wl = scenes.procedural(width=256, height=256, spp=1)
    views = [(cam, 1, 1, seed) for cam, _, _, seed in scenes.batch_views(48)]
    return wl, views


def _scaled(rgb, frame):
    """A sample c seen through a fresh film at `frame`: 0 + (1/frame) (c - 0),
    in binary32 as the film update computes it."""
    inv = np.float32(1.0) / np.float32(frame)
    return (inv * rgb.astype(np.float32)).astype(np.float32)


def test_c3_batch_of_48_views(c3):
    """Each view of the batch draws its samples from RNG streams frame0 + s +
    seed and folds them into its film with weight 1/(frame0 + s).  The
    megakernel and the oracle render stream 1 + seed at frame 1 + seed into a
    zeroed film, which holds the sample scaled by 1/(1 + seed): the batch film
    scaled the same way must equal it."""
    wl, views = c3
    sizes = {(c.w, c.h) for c, *_ in views}
    assert sizes == {(256, 256), (512, 512), (1280, 720), (1920, 1080)}
    assert len({seed for *_, seed in views}) == 48
    r = Renderer.from_workload(wl)
    tune = r.autotune(views)
    best = min(tune, key=lambda x: x[3])
    print(f"C3 autotune: {len(tune)} settings, best capacity {best[0]} block {best[1]} lds {best[2]}")
    r.render_batch(views)
    m = Renderer.from_workload(wl)
    tiny = make_camera((0.0, 0.0, 1.0), (0.0, 0.0, 0.0), w=8, h=8)
    for i, (cam, frame, spp, seed) in enumerate(views):
        fb = r.batch_film(i)
        assert np.all(fb[..., 3] == 1.0)
        m.update_camera(tiny)  # a size change reallocates (and zeroes) the film
        m.update_camera(cam)
        m.render(spp=spp, frame=frame + seed)
        ref = m.film()
        np.testing.assert_array_equal(_scaled(fb[..., :3], frame + seed), ref[..., :3], err_msg=f"view {i}")
    # two full views against the oracle, a 1920 x 1080 one among them
    big = next(i for i, v in enumerate(views) if (v[0].w, v[0].h) == (1920, 1080))
    small = next(i for i, v in enumerate(views) if (v[0].w, v[0].h) == (512, 512))
    for i in (big, small):
        cam, frame, spp, seed = views[i]
        o = OracleScene(wl.payload, cam, wl.spec, wl.options, wl.hosek)
        c, _ = o.render(frame + seed, 1)
        g = _scaled(r.batch_film(i)[..., :3], frame + seed)
        same = float(np.mean(np.all(g == c[..., :3], axis=-1)))
        # relative error at the radiance scale (both undone by the same factor)
        k = float(frame + seed)
        err = mean_rel_err(g.astype(np.float64) * k, c[..., :3].astype(np.float64) * k)
        print(f"C3 view {i} ({cam.w}x{cam.h}, seed {seed}) vs oracle: mean rel err {err:.3e}, identical {same:.4f}")
        assert err <= TOL_IMAGE
