"""GPU: the device builder of bottom-level trees (noor_scene_build_device)
against the host builder.  Both make the same split decisions with the same
binary32 arithmetic and partition stably, so the payloads must be equal byte
for byte (nodes, triangle order, Wald records, attributes), and with them the
images.  The host builder itself is pinned by the oracle's BVH tests (the
traversal equals brute force; node invariants)."""
import ctypes as C

import numpy as np
import pytest

from noor_amd import abi, scenes
from noor_amd._lib import lib
from noor_amd.renderer import Renderer

pytestmark = pytest.mark.gpu


def _stats(payload):
    s = abi.BuildStats()
    assert lib().noor_payload_stats(payload.handle, C.byref(s)) == abi.OK
    return s


CASES = [
    ("terrain", dict(grid=320, n_spheres=30, env_w=64, env_h=32)),       # one 203k-triangle mesh + instances
    ("procedural", dict(target_tris=240_000, n_meshes=10)),              # several meshes above 16k triangles
]


@pytest.mark.parametrize("name,kw", CASES)
def test_device_build_equals_host_build(name, kw):
    host = getattr(scenes, name)(width=64, height=48, spp=2, **kw)
    dev = getattr(scenes, name)(width=64, height=48, spp=2, build_device=0, **kw)
    a, b = host.payload.arrays(), dev.payload.arrays()
    print(f"{name}: {host.payload.view.num_triangles} triangles, {host.payload.view.num_bvh_nodes} nodes")
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    sh, sd = _stats(host.payload), _stats(dev.payload)
    for f in ("height", "stack_depth", "num_nodes", "num_inner", "num_leaves", "max_leaf_tris"):
        assert getattr(sh, f) == getattr(sd, f), f
    print(f"{name}: host build {sh.build_seconds * 1e3:.1f} ms, device build {sd.build_seconds * 1e3:.1f} ms")
    r1, r2 = Renderer.from_workload(host), Renderer.from_workload(dev)
    r1.render(spp=2)
    r2.render(spp=2)
    np.testing.assert_array_equal(r1.film(), r2.film())


def test_device_build_rejects_too_many_bins():
    from noor_amd.scene import Mesh, SceneBuilder, material
    sb = SceneBuilder()
    m = sb.add_material(material(abi.MAT_DIFFUSE))
    pos, tri, _ = scenes.grid_heightfield(120, 10.0, lambda x, z: 0.1 * np.sin(x) * np.cos(z))
    sb.add_mesh(Mesh(pos, tri, material=m))
    with pytest.raises(abi.NoorError):
        sb.build(num_bins=32, device=0)
