"""GPU: noor_trace_noalloc (traversal stack in the library's fixed scratch,
device arrays in and out, no allocation) against noor_trace on the same rays,
bit for bit: two calls in a row, the second with fewer rays, closest and any
hit, with a 4-entry LDS stack so that the scratch holds most of the stack."""
import ctypes as C

import numpy as np
import pytest
import torch

from noor_amd import scenes
from noor_amd._lib import lib
from noor_amd.renderer import Renderer, hits_to_numpy, rays_from_numpy

pytestmark = pytest.mark.gpu


def _device_rays(rays):
    a = np.ctypeslib.as_array(C.cast(rays, C.POINTER(C.c_float)), shape=(len(rays), 8)).copy()
    return torch.from_numpy(a).cuda()


@pytest.mark.parametrize("any_hit", [0, 1])
def test_noalloc_trace_equals_trace(any_hit):
    wl = scenes.procedural(width=16, height=16, n_meshes=24, target_tris=60000)
    r = Renderer.from_workload(wl)
    r.set_stack_lds(4)  # most of the traversal stack in device memory: the scratch
    # rays aimed near the instances (the scene is sparse)
    v = wl.payload.view
    T = np.stack([np.ctypeslib.as_array(C.cast(v.object_to_world_row[k], C.POINTER(C.c_float)),
                                        shape=(v.num_instances, 4))[:, 3] for k in range(3)], axis=1)
    rng = np.random.default_rng(31)
    tgt = T[rng.integers(0, v.num_instances, 40000)] + rng.normal(size=(40000, 3)) * 0.7
    org = tgt + rng.normal(size=(40000, 3)) * 20
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = rays_from_numpy(org.astype(np.float32), d.astype(np.float32))
    ref = hits_to_numpy(r.trace(rays, any_hit=bool(any_hit)))
    assert (ref[:, 6] > 0).sum() > 10000
    d_rays = _device_rays(rays)
    for n in (len(rays), len(rays) // 3):
        d_hits = torch.full((len(rays), 8), -7, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        st_ = lib().noor_trace_noalloc(r._h, C.c_void_p(d_rays.data_ptr()), n, any_hit,
                                       C.c_void_p(d_hits.data_ptr()))
        assert st_ == 0
        r.synchronize()
        got = d_hits.cpu().numpy().view(np.uint32)
        np.testing.assert_array_equal(got[:n], ref[:n])
        assert np.all(got[n:].view(np.int32) == -7)  # nothing written past n


def test_noalloc_trace_rejects_bad_arguments():
    wl = scenes.cornell(subdiv=2, width=16, height=16)
    r = Renderer.from_workload(wl)
    assert lib().noor_trace_noalloc(None, None, 1, 0, None) != 0
    assert lib().noor_trace_noalloc(r._h, None, 1, 0, None) != 0
    assert lib().noor_trace_noalloc(r._h, None, 0, 0, None) == 0
