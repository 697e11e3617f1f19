"""Seeded synthetic scenes for the five benchmark / parity workloads.

There are no scene assets: every workload is generated here from a seed.
  cornell(...)      512^2 Cornell box (12 quads) + 6x subdivided icosphere with
                    radial value-noise displacement (amplitude 0.05), one quad
                    area light of radiance 15, Lambert albedo U[0.2, 0.8].
  procedural(...)   1M triangles in 256 meshes (spheres, tori R=1 r=0.3,
                    256x256 heightfield patches), random rotation / scale
                    U[0.5, 3] / position in a 100-unit cube, materials 40/30/30
                    Lambert / GGX copper / glass (IOR 1.5), Hosek sky.
  terrain(...)      1582^2 fBm heightfield (8 octaves, amplitude 20, ~5M
                    triangles) + 1000 instanced GGX spheres, lit by a 2048x1024
                    log-normal HDR environment with three 2-degree hot spots.
  adjoint_scene(...) the Cornell geometry without ceiling, 64 glossy materials
                    (albedo + roughness), lit by a 512x256 environment.
"""

import numpy as np

from . import abi
from .scene import Mesh, SceneBuilder, cook_hosek, make_camera, make_options, make_spec, material

# Copper refractive index and absorption (RGB), as used by PBRT-v3 for "Cu".
CU_ETA = (0.200438, 0.924033, 1.10221)
CU_K = (3.91295, 2.45285, 2.14219)


# ---------------------------------------------------------------------------------------------
# geometry primitives

def icosphere(subdiv):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                  [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        es = np.sort(e, axis=1)
        key = es[:, 0] * (len(v) + 1) + es[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a = uk // (len(v) + 1)
        b = uk % (len(v) + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv.reshape(3, -1).T + len(v)  # midpoints of edges 01, 12, 20
        v = np.concatenate([v, mid])
        f0, f1, f2 = f[:, 0], f[:, 1], f[:, 2]
        m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
        f = np.concatenate([np.stack([f0, m01, m20], 1), np.stack([f1, m12, m01], 1),
                            np.stack([f2, m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(73856093)) ^ (iy.astype(np.uint64) * np.uint64(19349663)) ^ \
        (iz.astype(np.uint64) * np.uint64(83492791)) ^ np.uint64(seed * 2654435761 & 0xFFFFFFFF)
    h ^= h >> np.uint64(13)
    h *= np.uint64(0x5bd1e995)
    h ^= h >> np.uint64(15)
    return (h & np.uint64(0xFFFFFF)).astype(np.float64) / float(0xFFFFFF) * 2.0 - 1.0


def value_noise3(p, seed):
    """Trilinear value noise in [-1, 1] on a hashed integer lattice, smoothstep weights."""
    p = np.asarray(p, np.float64)
    i = np.floor(p).astype(np.int64)
    f = p - i
    w = f * f * (3.0 - 2.0 * f)
    out = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                val = _hash3(i[:, 0] + dx + 1000, i[:, 1] + dy + 1000, i[:, 2] + dz + 1000, seed)
                wt = (w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) * \
                     (w[:, 2] if dz else 1 - w[:, 2])
                out += wt * val
    return out


def value_noise2(x, z, seed):
    p = np.stack([x.ravel(), z.ravel(), np.zeros(x.size)], 1)
    return value_noise3(p, seed).reshape(x.shape)


def displaced_sphere(subdiv, radius, amplitude, seed, freq=4.0):
    v, f = icosphere(subdiv)
    n = value_noise3(v * freq, seed)
    r = radius + amplitude * n
    return v * r[:, None], f


def torus(R, r, nu, nv):
    u = np.arange(nu) * 2 * np.pi / nu
    w = np.arange(nv) * 2 * np.pi / nv
    U, W = np.meshgrid(u, w, indexing="ij")
    x = (R + r * np.cos(W)) * np.cos(U)
    y = r * np.sin(W)
    z = (R + r * np.cos(W)) * np.sin(U)
    pos = np.stack([x, y, z], -1).reshape(-1, 3)
    nx = np.cos(W) * np.cos(U)
    ny = np.sin(W)
    nz = np.cos(W) * np.sin(U)
    nrm = np.stack([nx, ny, nz], -1).reshape(-1, 3)
    uv = np.stack([U / (2 * np.pi), W / (2 * np.pi)], -1).reshape(-1, 2)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    tris = np.concatenate([np.stack([a, d, b], -1).reshape(-1, 3), np.stack([b, d, c], -1).reshape(-1, 3)])
    return pos, tris, nrm, uv


def grid_heightfield(n, extent, height_fn):
    """n x n vertex grid over [-extent/2, extent/2]^2 in xz, y = height_fn(x, z)."""
    s = np.linspace(-extent / 2, extent / 2, n)
    X, Z = np.meshgrid(s, s, indexing="ij")
    Y = height_fn(X, Z)
    pos = np.stack([X, Y, Z], -1).reshape(-1, 3)
    uv = np.stack([(X - X.min()) / extent, (Z - Z.min()) / extent], -1).reshape(-1, 2)
    i = np.arange(n - 1)[:, None]
    j = np.arange(n - 1)[None, :]
    a = i * n + j
    b = (i + 1) * n + j
    c = (i + 1) * n + j + 1
    d = i * n + j + 1
    tris = np.concatenate([np.stack([a, d, b], -1).reshape(-1, 3), np.stack([b, d, c], -1).reshape(-1, 3)])
    return pos, tris.astype(np.uint32), uv


def fbm2(x, z, octaves, seed, base_freq):
    h = np.zeros_like(x)
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        h += amp * value_noise2(x * freq, z * freq, seed + 17 * o)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return h / norm


def quad(corner, u, v):
    """Two triangles over corner + [0,1]u + [0,1]v, wound so that the geometric
    normal (and the vertex normals derived from it) is v x u.  The path's
    sidedness test (noor/cuda/accumulate.cuh:41-45) ends a path scattered to
    the side the geometric normal points away from, so walls must face the
    room."""
    c, u, v = (np.asarray(a, np.float64) for a in (corner, u, v))
    pos = np.stack([c, c + u, c + u + v, c + v])
    return pos, np.array([[0, 2, 1], [0, 3, 2]], np.uint32)


def random_rotation(rng):
    """Uniform rotation on SO(3) from a random unit quaternion (Shoemake)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def hosek_sky(turbidity=3.0, albedo=0.3, elevation_deg=30.0, sun_lon=np.pi / 4):
    """The 1M-triangle scene's sky: Hosek-Wilkie, turbidity 3, grey ground of
    albedo 0.3, sun 30 degrees above the horizon, cooked as HosekSky::update
    does (noor/header/hosek.h:61-83)."""
    lat = np.float32(np.radians(90.0 - elevation_deg))
    return cook_hosek(turbidity, (albedo, albedo, albedo), lat, np.float32(sun_lon))


def env_map(width, height, seed, lobes=3, sigma_deg=2.0, peak=1.0e4):
    """Log-normal (mu 0, sigma 1) background + Gaussian hot-spot lobes, RGBA32F."""
    rng = np.random.default_rng(seed)
    bg = np.exp(rng.normal(0.0, 1.0, (height, width))).astype(np.float64)
    tint = np.array([1.0, 0.95, 0.9])
    img = bg[..., None] * tint[None, None, :]
    u = (np.arange(width) + 0.5) / width
    v = (np.arange(height) + 0.5) / height
    phi = 2 * np.pi * u[None, :]
    theta = np.pi * v[:, None]
    d = np.stack([np.sin(theta) * np.cos(phi), np.cos(theta) * np.ones_like(phi), np.sin(theta) * np.sin(phi)], -1)
    sig = np.radians(sigma_deg)
    for _ in range(lobes):
        th = np.arccos(rng.uniform(0.2, 0.95))
        ph = rng.uniform(0, 2 * np.pi)
        c = np.array([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)])
        ang = np.arccos(np.clip(d @ c, -1.0, 1.0))
        img += peak * np.exp(-0.5 * (ang / sig) ** 2)[..., None]
    out = np.ones((height, width, 4), np.float32)
    out[..., :3] = img
    return out


class Workload:
    """A built scene plus the camera / spec / options / sky to render it."""

    def __init__(self, name, payload, camera, spec, options, hosek, spp, info=None):
        self.name = name
        self.payload = payload
        self.camera = camera
        self.spec = spec
        self.options = options
        self.hosek = hosek
        self.spp = spp
        self.info = info or {}


# ---------------------------------------------------------------------------------------------
# workloads

def _cornell_builder(seed, subdiv, with_ceiling=True, with_light=True, materials=None):
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()

    def lam():
        return sb.add_material(material(abi.MAT_DIFFUSE, diffuse=tuple(rng.uniform(0.2, 0.8, 3))))

    L = 1.0
    walls = [
        ((-L, -L, -L), (2 * L, 0, 0), (0, 0, 2 * L)),     # floor (normal +y)
        ((-L, L, -L), (0, 0, 2 * L), (2 * L, 0, 0)),      # ceiling (normal -y)
        ((-L, -L, -L), (0, 2 * L, 0), (2 * L, 0, 0)),     # back wall (normal +z)
        ((-L, -L, L), (2 * L, 0, 0), (0, 2 * L, 0)),      # front wall (normal -z)
        ((-L, -L, -L), (0, 0, 2 * L), (0, 2 * L, 0)),     # left wall (normal +x)
        ((L, -L, -L), (0, 2 * L, 0), (0, 0, 2 * L)),      # right wall (normal -x)
    ]
    meshes = []
    for k, (c, u, v) in enumerate(walls):
        if k == 1 and not with_ceiling:
            continue
        if k == 3 and not with_ceiling:
            continue
        pos, tri = quad(c, u, v)
        meshes.append((pos, tri))
    # a block: 6 quads of an axis-aligned box
    b0, b1 = np.array([-0.7, -1.0, -0.6]), np.array([-0.2, -0.2, -0.1])
    dx, dy, dz = b1 - b0
    block = [
        (b0, (0, 0, dz), (dx, 0, 0)), ((b0[0], b1[1], b0[2]), (dx, 0, 0), (0, 0, dz)),
        (b0, (dx, 0, 0), (0, dy, 0)), ((b0[0], b0[1], b1[2]), (0, dy, 0), (dx, 0, 0)),
        (b0, (0, dy, 0), (0, 0, dz)), ((b1[0], b0[1], b0[2]), (0, 0, dz), (0, dy, 0)),
    ]
    bp, bt = [], []
    for c, u, v in block:
        pos, tri = quad(c, u, v)
        bt.append(tri + 4 * len(bp))
        bp.append(pos)
    meshes.append((np.concatenate(bp), np.concatenate(bt)))
    sp, sf = displaced_sphere(subdiv, 0.4, 0.05, seed + 1)
    sp = sp + np.array([0.35, -0.55, 0.15])
    return sb, rng, meshes, (sp, sf), lam


def cornell(seed=1, subdiv=6, width=512, height=512, spp=16, bounces=5, rr=3,
            rng_kind=abi.RNG_PCG32, tri_test=abi.TRI_WATERTIGHT, compat=abi.COMPAT_ALL):
    sb, rng, meshes, (sp, sf), lam = _cornell_builder(seed, subdiv)
    for pos, tri in meshes:
        sb.add_mesh(Mesh(pos, tri, material=lam()))
    sb.add_mesh(Mesh(sp, sf, material=lam()))
    # area light just below the ceiling, facing down (one-sided)
    sb.add_quad_light((-0.25, 0.99, -0.25), (0.5, 0, 0), (0, 0, 0.5), (15.0, 15.0, 15.0))
    payload = sb.build()
    cam = make_camera((0.0, 0.0, 0.95), (0.0, -0.1, -1.0), (0, 1, 0), np.radians(60.0), width, height)
    spec = make_spec(payload, bounces=bounces, rr=rr, skydome_type=abi.SKY_CONSTANT)
    opt = make_options(rng=rng_kind, tri_test=tri_test, compat=compat)
    return Workload("cornell", payload, cam, spec, opt, None, spp,
                    info=dict(triangles=payload.view.num_triangles, seed=seed))


def procedural(seed=2, width=1920, height=1080, spp=64, bounces=8, rr=3, target_tris=1_000_000,
               rng_kind=abi.RNG_PCG32, tri_test=abi.TRI_WATERTIGHT, compat=abi.COMPAT_ALL, n_meshes=256,
               build_device=None):
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    n_hf = max(1, int(round(4 * n_meshes / 256)))
    n_rest = n_meshes - n_hf
    n_sph = n_rest // 2
    n_tor = n_rest - n_sph
    hf_n = 256
    hf_tris = 2 * (hf_n - 1) ** 2
    rest_budget = max(target_tris - n_hf * hf_tris, 1000)
    sph_sub = 3 if rest_budget / max(n_rest, 1) > 1500 else 2
    sph_tris = 20 * 4 ** sph_sub
    tor_budget = max(rest_budget - n_sph * sph_tris, 200 * n_tor)
    seg = max(8, int(np.sqrt(tor_budget / max(n_tor, 1) / 4)))
    sph = icosphere(sph_sub)
    tor = torus(1.0, 0.3, 2 * seg, seg)
    kinds = ["hf"] * n_hf + ["sph"] * n_sph + ["tor"] * n_tor
    rng.shuffle(kinds)
    for k, kind in enumerate(kinds):
        u = rng.random()
        if u < 0.4:
            m = material(abi.MAT_DIFFUSE, diffuse=tuple(rng.uniform(0.2, 0.8, 3)))
        elif u < 0.7:
            r = rng.uniform(0.05, 0.5)
            m = material(abi.MAT_METAL, eta=CU_ETA, k=CU_K, roughness=(r, r), specular=(1, 1, 1))
        else:
            m = material(abi.MAT_GLASS, coat_ior=1.5)
        mi = sb.add_material(m)
        if kind == "sph":
            mesh = Mesh(sph[0], sph[1], normals=sph[0], material=mi)
        elif kind == "tor":
            mesh = Mesh(tor[0], tor[1], normals=tor[2], uvs=tor[3], material=mi)
        else:
            s = int(rng.integers(1 << 30))
            pos, tri, uv = grid_heightfield(hf_n, 2.0, lambda x, z: 0.15 * fbm2(x, z, 5, s, 2.0))
            mesh = Mesh(pos, tri, uvs=uv, material=mi)
        idx = sb.add_mesh(mesh)
        M = np.eye(4)
        M[:3, :3] = random_rotation(rng) * rng.uniform(0.5, 3.0)
        M[:3, 3] = rng.uniform(0.0, 100.0, 3)
        sb.add_instance(idx, M)
    payload = sb.build(device=build_device)
    cam = make_camera((50.0, 75.0, 210.0), (50.0, 45.0, 50.0), (0, 1, 0), np.radians(45.0), width, height)
    spec = make_spec(payload, bounces=bounces, rr=rr, skydome_type=abi.SKY_PHYSICAL)
    opt = make_options(rng=rng_kind, tri_test=tri_test, compat=compat)
    return Workload("procedural", payload, cam, spec, opt, hosek_sky(), spp,
                    info=dict(triangles=payload.view.num_triangles, meshes=n_meshes, seed=seed))


def terrain(seed=3, grid=1582, n_spheres=1000, width=1920, height=1080, spp=32, bounces=8, rr=3,
            env_w=2048, env_h=1024, rng_kind=abi.RNG_PCG32, tri_test=abi.TRI_WATERTIGHT,
            compat=abi.COMPAT_ALL, extent=1000.0, build_device=None):
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    ter = sb.add_material(material(abi.MAT_DIFFUSE, diffuse=(0.45, 0.42, 0.38)))
    s = int(rng.integers(1 << 30))
    hfn = lambda x, z: 20.0 * fbm2(x, z, 8, s, 4.0 / extent)
    pos, tri, uv = grid_heightfield(grid, extent, hfn)
    sb.add_mesh(Mesh(pos, tri, uvs=uv, material=ter))
    sv, sf = icosphere(3)
    # per-sphere GGX conductors: a few shared materials with random roughness
    mats = []
    for _ in range(16):
        r = rng.uniform(0.05, 0.5)
        mats.append(sb.add_material(material(abi.MAT_METAL, eta=CU_ETA, k=CU_K, roughness=(r, r),
                                             specular=tuple(rng.uniform(0.5, 1.0, 3)))))
    sphere_meshes = [sb.add_mesh(Mesh(sv, sf, normals=sv, material=m)) for m in mats]
    xs = rng.uniform(-0.45 * extent, 0.45 * extent, n_spheres)
    zs = rng.uniform(-0.45 * extent, 0.45 * extent, n_spheres)
    rad = rng.uniform(2.0, 8.0, n_spheres)
    ys = hfn(xs, zs) + rad * 0.6
    for i in range(n_spheres):
        M = np.eye(4)
        M[:3, :3] *= rad[i]
        M[:3, 3] = (xs[i], ys[i], zs[i])
        sb.add_instance(sphere_meshes[i % len(sphere_meshes)], M)
    sb.add_instance(0, np.eye(4))
    # mesh 0 (terrain) must be instanced explicitly once; sphere meshes only via instances
    sb.set_env(env_map(env_w, env_h, seed + 7))
    payload = sb.build(device=build_device)
    cam = make_camera((0.0, 90.0, 0.48 * extent), (0.0, 0.0, 0.0), (0, 1, 0), np.radians(50.0), width, height)
    spec = make_spec(payload, bounces=bounces, rr=rr, skydome_type=abi.SKY_HDR)
    opt = make_options(rng=rng_kind, tri_test=tri_test, compat=compat)
    return Workload("terrain", payload, cam, spec, opt, None, spp,
                    info=dict(triangles=payload.view.num_triangles, spheres=n_spheres, seed=seed,
                              grid=grid))


def adjoint_scene(seed=4, subdiv=6, width=512, height=512, bounces=4, rr=3, n_materials=64,
                  env_w=512, env_h=256, params=None, rng_kind=abi.RNG_PCG32):
    """Cornell geometry without ceiling and front wall, 64 GLOSSY materials,
    lit by an environment map.  `params` = dict(albedo (M,3), roughness (M,))
    overrides the seeded ground truth."""
    sb, rng, meshes, (sp, sf), _ = _cornell_builder(seed, subdiv, with_ceiling=False)
    M = n_materials
    if params is None:
        albedo = rng.uniform(0.2, 0.8, (M, 3))
        rough = rng.uniform(0.1, 0.6, M)
    else:
        albedo = np.asarray(params["albedo"], np.float64)
        rough = np.asarray(params["roughness"], np.float64)
    for i in range(M):
        sb.add_material(material(abi.MAT_GLOSSY, diffuse=tuple(albedo[i]), specular=(0.25, 0.25, 0.25),
                                 roughness=(rough[i], rough[i]), coat_ior=1.5))
    # walls and block: materials 0..len(meshes)-1
    for k, (pos, tri) in enumerate(meshes):
        sb.add_mesh(Mesh(pos, tri, material=k % M))
    # sphere triangles by lat-long sector over the remaining materials
    c = sp[sf].mean(axis=1) - np.array([0.35, -0.55, 0.15])
    th = np.arccos(np.clip(c[:, 1] / np.linalg.norm(c, axis=1), -1, 1))
    ph = np.arctan2(c[:, 2], c[:, 0]) + np.pi
    rest = M - len(meshes)
    nth = 6
    nph = max(1, rest // nth)
    sec = (np.minimum((th / np.pi * nth).astype(int), nth - 1) * nph +
           np.minimum((ph / (2 * np.pi) * nph).astype(int), nph - 1))
    sb.add_mesh(Mesh(sp, sf, material=(len(meshes) + sec % rest).astype(np.uint32)))
    env = env_map(env_w, env_h, seed + 11, lobes=3, sigma_deg=6.0, peak=30.0)
    sb.set_env(env)
    payload = sb.build()
    cam = make_camera((0.0, 0.0, 0.95), (0.0, -0.1, -1.0), (0, 1, 0), np.radians(60.0), width, height)
    spec = make_spec(payload, bounces=bounces, rr=rr, skydome_type=abi.SKY_HDR)
    opt = make_options(rng=rng_kind, compat=0)
    return Workload("adjoint", payload, cam, spec, opt, None, 1,
                    info=dict(albedo=albedo, roughness=rough, env=env, n_materials=M))


def batch_views(n=48, seed=7, center=(50.0, 45.0, 50.0), sizes=((256, 256), (512, 512), (1280, 720), (1920, 1080)),
                spps=(1, 4, 16, 64)):
    """`n` seeded views of the procedural scene for batched rendering: eye on a
    shell around the scene, resolution and spp drawn uniformly from `sizes`
    and `spps`, a per-view RNG seed.  Returns [(camera, frame, spp, seed)]."""
    rng = np.random.default_rng(seed)
    c = np.asarray(center, np.float64)
    out = []
    for i in range(n):
        d = rng.normal(size=3)
        d[1] = abs(d[1]) + 0.3
        d /= np.linalg.norm(d)
        eye = c + d * rng.uniform(120.0, 220.0)
        look = c + rng.normal(size=3) * 10.0
        w, h = sizes[int(rng.integers(len(sizes)))]
        spp = int(spps[int(rng.integers(len(spps)))])
        cam = make_camera(tuple(eye), tuple(look), (0, 1, 0), np.radians(rng.uniform(35.0, 60.0)), w, h)
        out.append((cam, 1, spp, int(rng.integers(1 << 20))))
    return out


def lights(seed=6, subdiv=4, width=96, height=80, spp=4, bounces=6, rr=3, mis=False, sky=True,
           rng_kind=abi.RNG_PCG32, compat=abi.COMPAT_ALL):
    """Every light type of the reference beside the HDR environment: a
    spherical and a disk area light, a point, a spot and a distant light, over
    diffuse walls, a glossy block and a copper sphere in the open Cornell
    geometry.  With lights present the reference turns the sky off
    (noor/source/scene.cpp:51-91); `sky` turns it back on so that the infinite
    light is among the sampled ones."""
    sb, rng, meshes, (sp, sf), lam = _cornell_builder(seed, subdiv, with_ceiling=False)
    walls = [lam() for _ in range(4)]
    block = sb.add_material(material(abi.MAT_GLOSSY, diffuse=(0.3, 0.3, 0.6), specular=(0.4, 0.4, 0.4),
                                     roughness=(0.2, 0.2)))
    metal = sb.add_material(material(abi.MAT_METAL, eta=CU_ETA, k=CU_K, roughness=(0.15, 0.15)))
    for k, (pos, tri) in enumerate(meshes):
        sb.add_mesh(Mesh(pos, tri, material=walls[k] if k < 4 else block))
    sb.add_mesh(Mesh(sp, sf, material=metal))
    sb.add_sphere_light((0.3, 0.55, -0.35), 0.12, (9.0, 8.0, 6.0))
    sb.add_disk_light((-0.45, 0.9, 0.2), (0.0, -1.0, 0.0), 0.2, (4.0, 4.5, 6.0))
    sb.add_point_light((0.05, 0.45, 0.35), (0.6, 0.55, 0.45))
    sb.add_spot_light((0.55, 0.75, 0.55), (-0.35, 1.0, -0.35), 0.96, 0.85, (3.0, 2.2, 1.5))
    sb.add_distant_light((0.3, 1.0, 0.6), (0.8, 0.75, 0.7))
    sb.set_env(env_map(128, 64, seed + 3, lobes=2, sigma_deg=8.0, peak=20.0))
    payload = sb.build()
    cam = make_camera((0.0, 0.0, 0.95), (0.0, -0.1, -1.0), (0, 1, 0), np.radians(60.0), width, height)
    spec = make_spec(payload, bounces=bounces, rr=rr, skydome_type=abi.SKY_HDR, mis=mis)
    if sky:
        spec.lighting_type |= abi.LIGHT_SKY_MASK
    opt = make_options(rng=rng_kind, compat=compat)
    return Workload("lights", payload, cam, spec, opt, None, spp, info=dict(triangles=payload.view.num_triangles))


def showcase(seed=5, subdiv=4, width=96, height=80, spp=4, bounces=6, rr=3, mis=False,
             rng_kind=abi.RNG_PCG32, compat=abi.COMPAT_ALL):
    """The secondary materials of the reference's factory (shadow catcher,
    rough glass, substrate, clear coat) beside glossy and diffuse ones, in the
    open Cornell geometry under an HDR environment."""
    sb, rng, meshes, (sp, sf), _ = _cornell_builder(seed, subdiv, with_ceiling=False)
    mats = [
        sb.add_material(material(abi.MAT_SHADOWCATCHER)),                                      # floor
        sb.add_material(material(abi.MAT_SUBSTRATE, diffuse=(0.6, 0.3, 0.2), specular=(0.1, 0.1, 0.1),
                                 roughness=(0.2, 0.2))),                                        # back wall
        sb.add_material(material(abi.MAT_CLEARCOAT, diffuse=(1.0, 1.0, 1.0), specular=(1.0, 1.0, 1.0),
                                 eta=CU_ETA, k=CU_K, roughness=(0.3, 0.3), coat_ior=1.5,
                                 coat_weight=0.4, coat_thickness=0.5, coat_sigma=0.3)),         # left wall
        sb.add_material(material(abi.MAT_GLOSSY, diffuse=(0.2, 0.5, 0.3), specular=(0.3, 0.3, 0.3),
                                 roughness=(0.25, 0.25))),                                      # right wall
        sb.add_material(material(abi.MAT_ROUGHGLASS, specular=(1.0, 1.0, 1.0), transmission=(0.9, 0.95, 1.0),
                                 roughness=(0.2, 0.2), coat_ior=1.5)),                          # block
        sb.add_material(material(abi.MAT_CLEARCOAT, diffuse=(1.0, 1.0, 1.0), specular=(1.0, 1.0, 1.0),
                                 eta=CU_ETA, k=CU_K, roughness=(0.1, 0.1), coat_ior=1.6,
                                 coat_weight=0.6)),                                             # sphere
    ]
    for k, (pos, tri) in enumerate(meshes):
        sb.add_mesh(Mesh(pos, tri, material=mats[min(k, 4)]))
    sb.add_mesh(Mesh(sp, sf, material=mats[5]))
    sb.set_env(env_map(128, 64, seed + 3, lobes=2, sigma_deg=8.0, peak=20.0))
    payload = sb.build()
    cam = make_camera((0.0, 0.0, 0.95), (0.0, -0.1, -1.0), (0, 1, 0), np.radians(60.0), width, height)
    spec = make_spec(payload, bounces=bounces, rr=rr, skydome_type=abi.SKY_HDR, mis=mis)
    opt = make_options(rng=rng_kind, compat=compat)
    return Workload("showcase", payload, cam, spec, opt, None, spp, info=dict(triangles=payload.view.num_triangles))


# ---------------------------------------------------------------------------------------------
# image textures

def checker(n, checks, c0=(0.8, 0.8, 0.8), c1=(0.1, 0.1, 0.1)):
    """n x n RGB checkerboard of checks x checks squares."""
    i = (np.arange(n) * checks // n)
    m = ((i[:, None] + i[None, :]) & 1).astype(np.float32)[:, :, None]
    return np.asarray(c0, np.float32) * (1 - m) + np.asarray(c1, np.float32) * m


def noise_image(n, seed, freq=8.0):
    """n x n value-noise image in [0, 1], periodic in both directions."""
    t = np.arange(n) / n
    X, Z = np.meshgrid(t, t, indexing="ij")
    # sum of periodic value-noise octaves: sample on a torus so the image tiles
    a = 2 * np.pi * X
    b = 2 * np.pi * Z
    p = np.stack([np.cos(a), np.sin(a), np.cos(b) + 2 * np.sin(b)], -1).reshape(-1, 3)
    v = value_noise3(p * freq / (2 * np.pi), seed).reshape(n, n)
    v = (v - v.min()) / max(v.max() - v.min(), 1e-12)
    return v.astype(np.float32)


def quad_uv(corner, u, v):
    """quad() with texture coordinates (0,0), (1,0), (1,1), (0,1) at corner,
    corner + u, corner + u + v, corner + v."""
    pos, tri = quad(corner, u, v)
    return pos, tri, np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float64)


def textured(seed=8, width=96, height=80, spp=4, bounces=6, rr=3, mis=False, lens_radius=0.0,
             rng_kind=abi.RNG_PCG32, compat=abi.COMPAT_ALL, tex=64, flat=False, env_size=(128, 64),
             camera_type=abi.CAMERA_PERSP):
    """The image-texture pipeline of the reference: mip-mapped colour,
    roughness and emittance textures read at the pixel footprint (ray
    differentials from the camera, carried through mirror, metal and glass
    bounces), bump maps and alpha-cut surfaces, in the open Cornell geometry
    under an HDR environment.  `flat` replaces every image by its mean and
    drops the bump and alpha maps: the same scene without textures."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    mat = material
    if flat:
        def mat(mtype, bump=None, alpha=None, bumpfactor=1.0, uvscale=None, **kw):
            for k, v in kw.items():
                if isinstance(v, np.ndarray):
                    a = v.reshape(-1, v.shape[-1]) if v.ndim == 3 else v.reshape(-1, 1)
                    m = a.mean(0)
                    kw[k] = tuple(float(x) for x in (m if m.size > 1 else np.repeat(m, 3)))
            return material(mtype, **kw)
    ck = checker(tex, 8)
    bumpimg = noise_image(tex, seed + 1)
    stripes = np.zeros((tex // 2, tex // 2, 3), np.float32)
    stripes[:, :, 0] = 0.2 + 0.6 * (np.arange(tex // 2)[None, :] % 8 < 4)
    stripes[:, :, 1] = 0.3
    stripes[:, :, 2] = 0.2 + 0.5 * noise_image(tex // 2, seed + 2)
    rough = 0.05 + 0.5 * noise_image(tex // 2, seed + 3)
    yy, xx = np.mgrid[0:tex, 0:tex] / tex
    holes = (((xx * 4) % 1 - 0.5) ** 2 + ((yy * 4) % 1 - 0.5) ** 2 > 0.09).astype(np.float32)
    glow = np.zeros((16, 16, 3), np.float32)
    glow[:, :, :] = (np.arange(16)[:, None, None] % 4 < 2) * np.array([6.0, 5.0, 3.0], np.float32)
    mats = dict(
        floor=sb.add_material(mat(abi.MAT_DIFFUSE, diffuse=ck, bump=bumpimg, bumpfactor=0.02,
                                       uvscale={"diffuse": (4, 4), "bump": (2, 2)})),
        back=sb.add_material(mat(abi.MAT_GLOSSY, diffuse=stripes, specular=(0.3, 0.3, 0.3), roughness=rough,
                                      uvscale={"diffuse": (2, 1), "roughness": (1, 2)})),
        left=sb.add_material(mat(abi.MAT_METAL, eta=CU_ETA, k=CU_K, roughness=rough * 0.5)),
        right=sb.add_material(mat(abi.MAT_MIRROR, specular=(0.9, 0.9, 0.9))),
        cut=sb.add_material(mat(abi.MAT_DIFFUSE, diffuse=ck[:, :, ::-1].copy(), alpha=holes,
                                     uvscale={"transparency": (1, 1)})),
        glass=sb.add_material(mat(abi.MAT_GLASS, specular=(1, 1, 1), transmission=(1, 1, 1), coat_ior=1.5)),
        torus=sb.add_material(mat(abi.MAT_SUBSTRATE, diffuse=stripes, specular=(0.05, 0.05, 0.05),
                                       roughness=(0.15, 0.15), bump=bumpimg, bumpfactor=0.01,
                                       uvscale={"diffuse": (4, 1), "bump": (8, 2)})),
        glow=sb.add_material(mat(abi.MAT_EMITTER, emittance=glow, emitter_scale=1.5)),
    )
    walls = [
        (quad_uv((-1.0, -0.5, -1.5), (2.0, 0.0, 0.0), (0.0, 0.0, 2.0)), mats["floor"]),    # floor, normal +y
        (quad_uv((-1.0, -0.5, -1.0), (0.0, 1.5, 0.0), (2.0, 0.0, 0.0)), mats["back"]),     # back, normal +z
        (quad_uv((-1.0, -0.5, -1.5), (0.0, 0.0, 2.0), (0.0, 1.5, 0.0)), mats["left"]),     # left, normal +x
        (quad_uv((1.0, -0.5, -1.5), (0.0, 1.5, 0.0), (0.0, 0.0, 2.0)), mats["right"]),     # right, normal -x
        (quad_uv((-0.55, -0.45, -0.6), (0.0, 0.6, 0.0), (0.5, 0.0, 0.0)), mats["cut"]),    # alpha-cut panel
        (quad_uv((0.2, 0.55, -0.95), (0.5, 0.0, 0.0), (0.0, 0.25, 0.0)), mats["glow"]),   # textured emitter
    ]
    for (pos, tri, uv), m in walls:
        sb.add_mesh(Mesh(pos, tri, uvs=uv, material=m))
    sv, sf = icosphere(3)
    sb.add_mesh(Mesh(sv * 0.2 + np.array([0.35, -0.28, -0.45]), sf, material=mats["glass"]))
    tp, tt, tn, tu = torus(0.22, 0.08, 48, 24)
    R = random_rotation(rng)
    sb.add_mesh(Mesh(tp @ R.T + np.array([-0.1, -0.3, -0.75]), tt, normals=tn @ R.T, uvs=tu, material=mats["torus"]))
    sb.set_env(env_map(env_size[0], env_size[1], seed + 3, lobes=2, sigma_deg=8.0, peak=20.0))
    payload = sb.build()
    cam = make_camera((0.0, 0.0, 0.95), (0.0, -0.15, -1.0), (0, 1, 0), np.radians(60.0), width, height,
                      lens_radius=lens_radius, camera_type=camera_type, ortho_zoom=1.2)
    spec = make_spec(payload, bounces=bounces, rr=rr, skydome_type=abi.SKY_HDR, mis=mis)
    opt = make_options(rng=rng_kind, compat=compat)
    return Workload("textured", payload, cam, spec, opt, None, spp, info=dict(triangles=payload.view.num_triangles))

