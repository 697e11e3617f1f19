"""Pins of the oracle's traversal, environment sampling and integrator.

  - triangle tests against double-precision Moller-Trumbore away from edges;
  - watertightness: rays through shared vertices of a closed mesh never miss;
  - BVH traversal == brute force over every (instance, triangle) pair;
  - BVH invariants (child boxes inside parents, leaves partition the ids);
  - camera: each primary ray passes through its own pixel footprint;
  - environment distribution: CDF rows monotone and ending at 1, sampling
    pdf equals the light's pdf_Li, and a sample histogram follows the pdf;
  - white furnace: a convex Lambert object under a uniform environment
    reflects exactly albedo x radiance (corrected sampling mode);
  - film: the running mean equals the arithmetic mean; a render split into
    two calls equals one call (resume);
  - FP64 vs FP32 integrators converge to the same image.
"""
import ctypes as C
import math

import numpy as np
import pytest

from noor_amd import abi, scenes
from noor_amd.renderer import hits_to_numpy, rays_from_numpy
from noor_amd.scene import Mesh, SceneBuilder, make_camera, make_options, make_spec, material
from oracle.binding import OracleScene
from oracle import binding as ob
from tests.helpers import random_rays, vertex_rays


def _single_tri_scene(tri_test):
    sb = SceneBuilder()
    m = sb.add_material(material(abi.MAT_DIFFUSE))
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.2], [0.2, 1.0, -0.3]], np.float32)
    sb.add_mesh(Mesh(pos, np.array([[0, 1, 2]]), material=m))
    p = sb.build()
    cam = make_camera((0, 0, 3), (0, 0, 0), w=8, h=8)
    return p, OracleScene(p, cam, make_spec(p), make_options(tri_test=tri_test)), pos


def _moller(o, d, v0, v1, v2):
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-14:
        return None
    inv = 1.0 / det
    s = o - v0
    u = (s @ p) * inv
    q = np.cross(s, e1)
    v = (d @ q) * inv
    t = (e2 @ q) * inv
    if u < 0 or v < 0 or u + v > 1 or t < 0:
        return None
    return t, u, v


@pytest.mark.parametrize("tri", [abi.TRI_WATERTIGHT, abi.TRI_WALD])
def test_triangle_tests_match_fp64_moller(tri):
    p, o, pos = _single_tri_scene(tri)
    rng = np.random.default_rng(5)
    n = 4000
    tgt = rng.uniform(-0.2, 1.1, (n, 3)) * np.array([1, 1, 0.2])
    org = tgt + rng.normal(size=(n, 3)) * 2
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = rays_from_numpy(org.astype(np.float32), d.astype(np.float32))
    hits = hits_to_numpy(o.trace(rays)[0])
    v = pos.astype(np.float64)
    checked = 0
    for i in range(n):
        oo = np.array([rays[i].o.x, rays[i].o.y, rays[i].o.z], np.float64)
        dd = np.array([rays[i].d.x, rays[i].d.y, rays[i].d.z], np.float64)
        ref = _moller(oo, dd, v[0], v[1], v[2])
        hit = hits[i, 6] == 1
        if ref is None:
            # near-edge rays may legitimately differ; require a clear miss otherwise
            continue
        t, u, w = ref
        if min(u, w, 1 - u - w) < 1e-4:
            continue
        assert hit, i
        got = hits[i, :3].view(np.float32)
        assert got[0] == pytest.approx(t, rel=2e-5)
        assert got[1] == pytest.approx(u, abs=2e-5)
        assert got[2] == pytest.approx(w, abs=2e-5)
        checked += 1
    assert checked > 1000


def test_watertight_closed_mesh_never_leaks():
    sb = SceneBuilder()
    m = sb.add_material(material(abi.MAT_DIFFUSE))
    v, f = scenes.icosphere(3)
    sb.add_mesh(Mesh(v * 1.7 + 0.3, f, material=m))
    p = sb.build()
    cam = make_camera((0, 0, 5), (0, 0, 0), w=8, h=8)
    o = OracleScene(p, cam, make_spec(p), make_options(tri_test=abi.TRI_WATERTIGHT))
    verts = (v * 1.7 + 0.3).astype(np.float32)
    # rays through vertices (shared by 5-6 triangles) and edge midpoints, arriving
    # from outside within 30 degrees of the outward normal so that they enter the
    # solid there: a watertight test may not let any of them through
    e = np.concatenate([verts, ((verts[f[:, 0]] + verts[f[:, 1]]) * 0.5).astype(np.float32)])
    rng = np.random.default_rng(3)
    idx = rng.integers(0, len(e), 6000)
    tgt = e[idx].astype(np.float64)
    nrm = tgt - 0.3
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    jit = rng.normal(size=tgt.shape) * 0.3
    org = tgt + 3.0 * (nrm + jit - nrm * np.sum(jit * nrm, axis=1, keepdims=True))
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = rays_from_numpy(org.astype(np.float32), d.astype(np.float32))
    hits = hits_to_numpy(o.trace(rays)[0])
    assert (hits[:, 6] == 1).all()


@pytest.mark.parametrize("tri", [abi.TRI_WATERTIGHT, abi.TRI_WALD])
def test_bvh_closest_equals_brute_force(tri):
    wl = scenes.cornell(subdiv=3, width=16, height=16, tri_test=tri)
    o = OracleScene.from_workload(wl)
    st = wl.payload.stats
    c = (np.array(st.scene_min) + np.array(st.scene_max)) / 2
    rays = random_rays(3000, c, 2.0 * st.scene_radius, seed=9)
    a = hits_to_numpy(o.trace(rays)[0])
    b = hits_to_numpy(o.trace(rays, brute=True)[0])
    np.testing.assert_array_equal(a, b)
    inside = random_rays(3000, c, 0.3 * st.scene_radius, seed=10)
    np.testing.assert_array_equal(hits_to_numpy(o.trace(inside)[0]), hits_to_numpy(o.trace(inside, brute=True)[0]))


def test_bvh_instanced_equals_brute_force():
    wl = scenes.procedural(width=16, height=16, n_meshes=24, target_tris=30000)
    o = OracleScene.from_workload(wl)
    st = wl.payload.stats
    c = (np.array(st.scene_min) + np.array(st.scene_max)) / 2
    rays = random_rays(1500, c, 1.5 * st.scene_radius, seed=4)
    np.testing.assert_array_equal(hits_to_numpy(o.trace(rays)[0]), hits_to_numpy(o.trace(rays, brute=True)[0]))
    # any-hit agrees with closest-hit existence
    anyh = hits_to_numpy(o.trace(rays, any_hit=True)[0])
    np.testing.assert_array_equal(anyh[:, 6] > 0, hits_to_numpy(o.trace(rays)[0])[:, 6] > 0)


def test_bvh_invariants():
    wl = scenes.procedural(width=16, height=16, n_meshes=12, target_tris=20000)
    arr = wl.payload.arrays()
    A, B = arr["bvh_a"], arr["bvh_b"]
    lo = A[:, :3].view(np.float32)
    hi = B[:, :3].view(np.float32)
    w = B[:, 3]
    n = len(A)
    nt = wl.payload.view.num_triangles
    seen = np.zeros(nt, np.int32)
    blas_roots = set()
    for i in range(n):
        flags = int(w[i])
        if flags & abi.c_u32(0x80000000).value:  # triangle leaf
            s, cnt = int(A[i, 3]), flags & 0x7FFFFFFF
            seen[s:s + cnt] += 1
            continue
        if flags & 0x20000000:  # instance entry
            blas_roots.add(int(A[i, 3]))
            continue
        if flags & (0x40000000 | 0x08000000):  # mesh leaf / light node
            continue
        for ch in (i + 1, int(A[i, 3])):
            assert (lo[ch] >= lo[i]).all() and (hi[ch] <= hi[i]).all(), (i, ch)
    assert (seen == 1).all()  # leaves partition the triangle ids
    assert sorted(set(arr["prim_ids"].tolist())) == list(range(nt))


def test_camera_rays_pass_through_their_pixel():
    wl = scenes.cornell(subdiv=2, width=40, height=30)
    o = OracleScene.from_workload(wl)
    cam = wl.camera
    M = np.array([[cam.world_to_raster.row[r].x, cam.world_to_raster.row[r].y, cam.world_to_raster.row[r].z,
                   [cam.world_to_raster.t.x, cam.world_to_raster.t.y, cam.world_to_raster.t.z][r]] for r in range(3)])
    for (x, y) in [(0, 0), (39, 29), (20, 15), (7, 22)]:
        for frame in (1, 2, 3):
            r = o.camera_ray(x, y, frame)
            d = np.array([r.d.x, r.d.y, r.d.z], np.float64)
            assert np.linalg.norm(d) == pytest.approx(1.0, abs=1e-6)
            # a point along the ray projected back onto the raster grid (affine part,
            # then the perspective divide by -z_camera)
            c2w = np.array([[cam.camera_to_world.row[k].x, cam.camera_to_world.row[k].y, cam.camera_to_world.row[k].z]
                            for k in range(3)])
            dc = np.linalg.solve(c2w, d)  # camera-space direction (the frame need not be orthonormal)
            r2c = np.array([[cam.raster_to_camera.row[k].x, cam.raster_to_camera.row[k].y,
                             cam.raster_to_camera.row[k].z] for k in range(3)])
            t = np.array([cam.raster_to_camera.t.x, cam.raster_to_camera.t.y, cam.raster_to_camera.t.z])
            # solve r2c @ (rx, ry, 0) + t = s * dc for rx, ry, s
            A = np.stack([r2c[:, 0], r2c[:, 1], -dc], 1)
            rx, ry, s = np.linalg.solve(A, -t)
            assert x - 0.5 - 1e-3 <= rx <= x + 0.5 + 1e-3
            assert y - 0.5 - 1e-3 <= ry <= y + 0.5 + 1e-3
    del M


def _env_scene(compat, w=64, h=32, seed=3):
    sb = SceneBuilder()
    m = sb.add_material(material(abi.MAT_DIFFUSE, diffuse=(1.0, 1.0, 1.0)))
    v, f = scenes.icosphere(3)
    sb.add_mesh(Mesh(v, f, normals=v, material=m))
    sb.set_env(scenes.env_map(w, h, seed, lobes=2, sigma_deg=10.0, peak=20.0))
    p = sb.build()
    cam = make_camera((0, 0, 4), (0, 0, 0), w=24, h=24, fov=np.radians(40))
    spec = make_spec(p, bounces=2, rr=5, skydome_type=abi.SKY_HDR)
    return p, OracleScene(p, cam, spec, make_options(compat=compat))


def test_env_cdf_invariants():
    p, o = _env_scene(abi.COMPAT_ALL)
    W, H, func, cdf = o.dist()
    assert (W, H) == (64, 32)  # nearest powers of two, clamped to [32, 1024]
    for r in list(range(H)):
        row = cdf[r * (W + 1): r * (W + 1) + W + 1]
        assert row[0] == 0.0 and row[-1] == pytest.approx(1.0, abs=1e-6)
        assert (np.diff(row) >= 0).all()
        # the exclusive scan matches a double-precision prefix sum of func / W
        fr = func[r * (W + 1): r * (W + 1) + W].astype(np.float64)
        ref = np.concatenate([[0.0], np.cumsum(fr / W)]) / (fr / W).sum()
        np.testing.assert_allclose(row, ref, atol=2e-6)
    marg = cdf[H * (W + 1): H * (W + 1) + H + 1]
    assert marg[0] == 0.0 and marg[-1] == pytest.approx(1.0, abs=1e-6) and (np.diff(marg) >= 0).all()
    # row integrals were copied into the marginal function
    np.testing.assert_array_equal(func[H * (W + 1): H * (W + 1) + H], func[[r * (W + 1) + W for r in range(H)]])


def test_env_sampling_pdf_matches_pdf_li_and_histogram():
    p, o = _env_scene(0)  # corrected mode: consistent sampling and pdf
    rng = np.random.default_rng(8)
    n = 40000
    u = rng.random((n, 2)).astype(np.float32)
    dirs = np.zeros((n, 3))
    rel = []
    for i in range(n):
        d, pdf, pdf_li = ob.env_sample(o, float(u[i, 0]), float(u[i, 1]))
        if i < 4000:
            rel.append(abs(pdf - pdf_li) / pdf_li)
        dirs[i] = d
    # the sampler's pdf (from theta = pi*v) and pdf_Li (from acos of the FP32
    # direction) differ by FP32 round-off in sin(theta); a direction that rounds
    # into the neighbouring texel (about 1 in 1000) takes that texel's value
    rel = np.array(rel)
    assert np.mean(rel < 1e-3) > 0.995 and np.median(rel) < 1e-5
    # The sampler is piecewise constant in (u, v) = (phi / 2pi, theta / pi):
    # texel (iu, iv) is chosen with probability func / sum(func), uniformly inside.
    # Histogram the samples over blocks of 4x4 texels against that exact mass.
    W, H, func, _ = o.dist()
    f2 = np.stack([func[r * (W + 1): r * (W + 1) + W] for r in range(H)]).astype(np.float64)
    mass = f2 / f2.sum()
    uu = np.mod(np.arctan2(dirs[:, 2], dirs[:, 0]), 2 * np.pi) / (2 * np.pi)
    vv = np.arccos(np.clip(dirs[:, 1], -1, 1)) / np.pi
    bu, bv = W // 4, H // 4
    hist, _, _ = np.histogram2d(vv, uu, bins=[bv, bu], range=[[0, 1], [0, 1]])
    expect = mass.reshape(bv, 4, bu, 4).sum(axis=(1, 3)) * n
    mask = expect > 20
    chi2 = np.sum((hist[mask] - expect[mask]) ** 2 / expect[mask])
    dof = mask.sum() - 1
    assert dof > 40
    assert chi2 < dof + 6 * math.sqrt(2 * dof), (chi2, dof)
    # a planted error in the marginal (rows drawn uniformly) is detected
    bad_exp = (mass / mass.sum(axis=1, keepdims=True) * (1.0 / H)).reshape(bv, 4, bu, 4).sum(axis=(1, 3)) * n
    m2 = bad_exp > 20
    assert np.sum((hist[m2] - bad_exp[m2]) ** 2 / bad_exp[m2]) > 10 * (m2.sum() + 6 * math.sqrt(2 * m2.sum()))


def _pdf_dir(o, d):
    W, H, func, _ = o.__dict__.setdefault("_dist", o.dist())
    theta = math.acos(max(-1.0, min(1.0, float(d[1]))))
    phi = math.atan2(float(d[2]), float(d[0]))
    if phi < 0:
        phi += 2 * math.pi
    iu = min(int(phi / (2 * math.pi) * W), W - 1)
    iv = min(int(theta / math.pi * H), H - 1)
    st = math.sin(theta)
    if st == 0:
        return 0.0
    return func[iv * (W + 1) + iu] / func[H * (W + 1) + H] / (2 * math.pi * math.pi * st)


def test_white_furnace_env_lambert_sphere():
    """A convex albedo-1 Lambert sphere under a constant environment E
    reflects E everywhere (direct lighting only, corrected sampling)."""
    sb = SceneBuilder()
    m = sb.add_material(material(abi.MAT_DIFFUSE, diffuse=(1.0, 1.0, 1.0)))
    v, f = scenes.icosphere(4)
    sb.add_mesh(Mesh(v, f, normals=v, material=m))
    env = np.ones((32, 64, 4), np.float32)
    env[..., :3] = 0.7
    sb.set_env(env)
    p = sb.build()
    cam = make_camera((0, 0, 3), (0, 0, 0), w=32, h=32, fov=np.radians(30))
    spec = make_spec(p, bounces=2, rr=5, skydome_type=abi.SKY_HDR, mis=True)
    o = OracleScene(p, cam, spec, make_options(compat=0))
    film, _ = o.render(1, 256, fp64=True, threads=8)
    rays_center = film[12:20, 12:20, :3]
    # with MIS the per-pixel std at 256 spp is about 0.01, so the 64-pixel mean is
    # good to ~0.0013 (without MIS the estimator is unbiased too but 4x noisier)
    assert rays_center.mean() == pytest.approx(0.7, abs=0.006)
    assert np.abs(rays_center - 0.7).max() < 0.06
    # a pixel that sees only the environment returns it exactly
    np.testing.assert_allclose(film[0, 0, :3], 0.7, rtol=1e-6)


def test_film_running_mean_and_resume():
    wl = scenes.cornell(subdiv=3, width=24, height=20, spp=6)
    o = OracleScene.from_workload(wl)
    full, _ = o.render(1, 6)
    part = np.zeros_like(full)
    o.render(1, 2, film=part)
    o.render(3, 4, film=part)
    np.testing.assert_array_equal(part, full)
    # running mean vs arithmetic mean of the 6 single-frame samples (FP64)
    frames = [o.render(k, 1, fp64=True)[0] for k in range(1, 7)]
    # a single-frame render at frame k > 1 lerps from a zero film; rebuild the samples
    samples = [frames[0]]
    for k in range(2, 7):
        samples.append(frames[k - 1] * k)
    mean = np.mean(samples, axis=0)
    run = o.render(1, 6, fp64=True)[0]
    np.testing.assert_allclose(run[..., :3], mean[..., :3], rtol=1e-9, atol=1e-9)


def test_fp64_and_fp32_converge():
    wl = scenes.cornell(subdiv=3, width=32, height=32, spp=64)
    o = OracleScene.from_workload(wl)
    a, _ = o.render(1, 64)
    b, _ = o.render(1, 64, fp64=True)
    rmse = float(np.sqrt(np.mean((a[..., :3] - b[..., :3]) ** 2)))
    assert rmse < 0.02 * float(np.mean(b[..., :3])) + 1e-3
