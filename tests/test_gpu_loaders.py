"""GPU: a scene read from OBJ + MTL files with PFM / PGM image maps (diffuse
texture, bump and transparency maps, a light group) renders on the GPU as the
oracle renders it."""
import numpy as np
import pytest

from noor_amd import abi
from noor_amd.loaders import obj_scene
from noor_amd.renderer import Renderer
from noor_amd.scene import make_camera, make_options, make_spec
from noor_amd.scenes import Workload
from oracle.binding import OracleScene
from tests.helpers import mean_rel_err
from tests.test_loaders import CUBE, MTL, _pfm, _pnm, _write

pytestmark = pytest.mark.gpu


def test_obj_scene_gpu_vs_oracle(tmp_path):
    _write(tmp_path / "cube.obj", CUBE)
    _write(tmp_path / "cube.mtl", MTL)
    _pfm(tmp_path / "tex.pfm", np.random.default_rng(1).random((4, 8, 3)).astype(np.float32))
    _pnm(tmp_path / "bump.pgm", (np.arange(16).reshape(4, 4) * 16).astype(np.uint8))
    _pnm(tmp_path / "alpha.pgm", np.array([[255, 0], [0, 255]], np.uint8))
    sb, m = obj_scene(str(tmp_path / "cube.obj"))
    sb.set_env(np.full((8, 16, 4), 0.5, np.float32))
    payload = sb.build()
    cam = make_camera((2.0, 1.5, 3.0), (0.0, 0.0, 0.0), (0, 1, 0), np.radians(50.0), 64, 48)
    spec = make_spec(payload, bounces=4, rr=3, skydome_type=abi.SKY_HDR)
    wl = Workload("obj", payload, cam, spec, make_options(), None, 1, info={})
    r = Renderer.from_workload(wl)
    r.render(spp=4)
    g = r.film()
    c, _ = OracleScene.from_workload(wl).render(1, 4)
    err = mean_rel_err(g, c)
    print(f"obj scene: megakernel vs oracle {err:.3e}, identical {np.mean(np.all(g == c, axis=-1)):.4f}", flush=True)
    assert err <= 1e-4 and float(g[..., :3].max()) > 0
    w = Renderer.from_workload(wl)
    w.render_wavefront(spp=4)
    np.testing.assert_array_equal(w.film(), g)
