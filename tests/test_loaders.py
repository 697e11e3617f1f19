"""Payload producers for files (include/noor_loaders.h), on the CPU.

OBJ + MTL (noor/source/objLoader.cpp:28-147, noor/header/objLoader.h:90-169;
glm_obj itself is not shipped with the reference, its published behaviour is
pinned here by closed forms): fan triangulation and negative indices, glmUnitize
(centre and 2 / max(|max| + |min|) per axis), facet and crease-angle vertex
normals (a cube keeps its face normals, a sphere gets radial ones), MTL fields
and map options, light groups as quad lights.  Images
(noor/source/texture.cpp:49-76): Radiance RGBE written by an independent
encoder here (flat and run-length scanlines), PFM in both byte orders, PNM
8/16-bit with stb_image's 2.2 gamma.
"""
import os
import struct

import numpy as np
import pytest

from noor_amd import abi, scenes
from noor_amd.loaders import ObjModel, load_image, obj_scene


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


CUBE = """# unit test cube 2 x 4 x 2 at the origin corner
mtllib cube.mtl
v 0 0 0
v 2 0 0
v 2 4 0
v 0 4 0
v 0 0 2
v 2 0 2
v 2 4 2
v 0 4 2
vt 0 0
vt 1 0
vt 1 1
vt 0 1
g box
usemtl red
f 1/1 4/4 3/3 2/2
f 5/1 6/2 7/3 8/4
f 1/1 2/2 6/3 5/4
usemtl shiny
f 4 8 7 3
f 1 5 8 4
f -7 -6 -2 -1
g ceiling_light
usemtl lamp
f 3 4 8 7
"""

MTL = """newmtl red
Kd 0.8 0.1 0.1
Ks 0 0 0
Ns 10
illum 2
map_Kd -s 2 3 tex.pfm
newmtl shiny
Kd 0.1 0.1 0.1
Ks 0.9 0.9 0.9
Ni 1.7
Ns 50
illum 6
map_bump -bm 0.25 bump.pgm
map_d alpha.pgm
newmtl lamp
Kd 0 0 0
Ke 5 4 3
illum 5
"""


def _pfm(path, img, little=True):
    h, w, c = img.shape
    hdr = (b"PF\n" if c == 3 else b"Pf\n") + f"{w} {h}\n{-1.0 if little else 1.0}\n".encode()
    data = img[::-1].astype("<f4" if little else ">f4").tobytes()
    with open(path, "wb") as f:
        f.write(hdr + data)


def _pnm(path, img, maxval=255):
    h, w = img.shape[:2]
    c = 1 if img.ndim == 2 else 3
    dt = ">u2" if maxval > 255 else "u1"
    with open(path, "wb") as f:
        f.write(f"P{5 if c == 1 else 6}\n# comment\n{w} {h}\n{maxval}\n".encode() + img.astype(dt).tobytes())


@pytest.fixture
def cube_dir(tmp_path):
    _write(tmp_path / "cube.obj", CUBE)
    _write(tmp_path / "cube.mtl", MTL)
    _pfm(tmp_path / "tex.pfm", np.random.default_rng(1).random((4, 8, 3)).astype(np.float32))
    _pnm(tmp_path / "bump.pgm", (np.arange(16).reshape(4, 4) * 16).astype(np.uint8))
    _pnm(tmp_path / "alpha.pgm", np.full((2, 2), 255, np.uint8))
    return tmp_path


def test_obj_geometry_unitize_and_fans(cube_dir):
    m = ObjModel(str(cube_dir / "cube.obj"))
    assert m.indices.shape == (14, 3)                      # 7 quads as fans
    np.testing.assert_allclose(m.bbox_min, [0, 0, 0])
    np.testing.assert_allclose(m.bbox_max, [2, 4, 2])
    assert m.unitize_scale == pytest.approx(0.5)           # 2 / max(2, 4, 2)
    p = m.positions
    np.testing.assert_allclose(p.min(0), [-0.5, -1.0, -0.5], atol=1e-7)
    np.testing.assert_allclose(p.max(0), [0.5, 1.0, 0.5], atol=1e-7)
    # first quad 1 4 3 2 -> triangles (1, 4, 3), (1, 3, 2) in file positions
    src = (np.array([[0, 0, 0], [2, 0, 0], [2, 4, 0], [0, 4, 0], [0, 0, 2], [2, 0, 2], [2, 4, 2], [0, 4, 2]],
                    np.float64) - [1, 2, 1]) * 0.5
    tri = p[m.indices]
    np.testing.assert_allclose(tri[0], src[[0, 3, 2]], atol=1e-7)
    np.testing.assert_allclose(tri[1], src[[0, 2, 1]], atol=1e-7)
    # negative indices: f -7 -6 -2 -1 = f 2 3 7 8 ... of 8 vertices: 2 3 7 8
    np.testing.assert_allclose(tri[10], src[[1, 2, 6]], atol=1e-7)
    np.testing.assert_allclose(tri[11], src[[1, 6, 7]], atol=1e-7)
    # texture coordinates: given for the first three quads, (0, 0) elsewhere
    np.testing.assert_allclose(m.uvs[m.indices[0]], [[0, 0], [0, 1], [1, 1]])
    np.testing.assert_allclose(m.uvs[m.indices[6]], 0)


def test_obj_crease_normals_cube_and_sphere(tmp_path):
    # cube faces meet at 90 degrees: every corner keeps its facet normal
    _write(tmp_path / "cube.obj", CUBE.replace("mtllib cube.mtl\n", ""))
    m = ObjModel(str(tmp_path / "cube.obj"))
    tri = m.positions[m.indices].astype(np.float64)
    fn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    for k in range(3):
        np.testing.assert_allclose(m.normals[m.indices[:, k]], fn, atol=1e-6)
    # a subdivided icosahedron (neighbouring faces within 30 degrees): radial normals
    v, f = scenes.icosphere(2)
    lines = [f"v {a} {b} {c}" for a, b, c in v] + [f"f {a + 1} {b + 1} {c + 1}" for a, b, c in f]
    _write(tmp_path / "ico.obj", "\n".join(lines) + "\n")
    m = ObjModel(str(tmp_path / "ico.obj"), unitize=False)
    radial = m.positions / np.linalg.norm(m.positions, axis=1, keepdims=True)
    assert np.min(np.sum(radial * m.normals, axis=1)) > 0.99
    assert m.positions.shape[0] == v.shape[0]               # one normal per vertex: no splits
    # at a 5 degree crease no face is within another's: facet normals again
    m5 = ObjModel(str(tmp_path / "ico.obj"), unitize=False, crease_deg=5.0)
    tri = m5.positions[m5.indices].astype(np.float64)
    fn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    np.testing.assert_allclose(m5.normals[m5.indices[:, 0]], fn, atol=1e-6)


def test_mtl_fields_and_lights(cube_dir):
    m = ObjModel(str(cube_dir / "cube.obj"))
    mats = {d["name"]: d for d in m.materials}
    red, shiny, lamp = mats["red"], mats["shiny"], mats["lamp"]
    assert red["type"] == abi.MAT_DIFFUSE and shiny["type"] == abi.MAT_METAL and lamp["type"] == abi.MAT_EMITTER
    np.testing.assert_allclose(red["diffuse"], [0.8, 0.1, 0.1], rtol=1e-6)
    assert red["roughness"] == pytest.approx(0.1) and shiny["roughness"] == pytest.approx(0.02)
    assert shiny["ior"] == pytest.approx(1.7)
    assert red["map_kd"] == "tex.pfm" and red["uvscale"] == pytest.approx((2.0, 3.0))
    assert shiny["map_bump"] == "bump.pgm" and shiny["bumpfactor"] == pytest.approx(0.25)
    assert shiny["map_d"] == "alpha.pgm"
    np.testing.assert_allclose(lamp["emittance"], [5, 4, 3])
    names = [d["name"] for d in m.materials]
    assert [names[i] for i in m.material] == ["red"] * 6 + ["shiny"] * 6 + ["lamp"] * 2
    # light group: quad from its first triangle (3 4 8): base = vertex 4, u = v8 - v4, v = v3 - v4
    assert m.lights.shape == (1, 15)
    s = 0.5
    v3, v4, v8 = [(np.array(a, np.float64) - [1, 2, 1]) * s for a in ([2, 4, 0], [0, 4, 0], [0, 4, 2])]
    L = m.lights[0]
    np.testing.assert_allclose(L[0:3], v4, atol=1e-7)
    np.testing.assert_allclose(L[3:6], v8 - v4, atol=1e-7)
    np.testing.assert_allclose(L[6:9], v3 - v4, atol=1e-7)
    n = np.cross(v8 - v4, v3 - v4)
    np.testing.assert_allclose(L[9:12], n / np.linalg.norm(n), atol=1e-7)
    np.testing.assert_allclose(L[12:15], [5, 4, 3])


def _rgbe_encode(img):
    """Ward's RGBE: shared exponent e = frexp(max) so that max = m 2^e, m in
    [0.5, 1); bytes = floor(c * 256 / 2^e), exponent byte e + 128."""
    mx = img[..., :3].max(-1)
    m, e = np.frexp(mx)
    out = np.zeros(img.shape[:2] + (4,), np.uint8)
    ok = mx > 1e-32
    scale = np.where(ok, 256.0 / np.ldexp(1.0, e), 0)
    out[..., :3] = np.floor(img[..., :3] * scale[..., None]).clip(0, 255)
    out[..., 3] = np.where(ok, e + 128, 0)
    return out


def _rgbe_decode(b):
    f = np.ldexp(1.0, b[..., 3].astype(np.int32) - 136)
    rgb = (b[..., :3] + 0.5) * f[..., None]
    return np.where(b[..., 3:4] == 0, 0.0, rgb)


def _rle_scanline(row):
    """Run-length scanline: 2 2 hi lo, then each channel as runs (count > 128:
    repeat) or literals."""
    w = row.shape[0]
    out = bytearray([2, 2, w >> 8, w & 255])
    for c in range(4):
        ch = row[:, c]
        i = 0
        while i < w:
            j = i
            while j < w and j - i < 127 and ch[j] == ch[i]:
                j += 1
            if j - i >= 3:
                out += bytes([128 + (j - i), ch[i]])
                i = j
            else:
                k = i
                while k < w and k - i < 128 and not (k + 2 < w and ch[k] == ch[k + 1] == ch[k + 2]):
                    k += 1
                out += bytes([k - i]) + bytes(ch[i:k])
                i = k
    return bytes(out)


@pytest.mark.parametrize("rle", [False, True])
def test_radiance_hdr(tmp_path, rle):
    rng = np.random.default_rng(3)
    h, w = 5, 40
    img = (rng.random((h, w, 3)) * 10 ** rng.uniform(-3, 3, (h, w, 1))).astype(np.float32)
    img[1, 3:20] = [0.5, 0.25, 4.0]   # runs
    img[2, 0] = 0                     # zero exponent
    b = _rgbe_encode(img)
    body = b"".join(_rle_scanline(b[y]) for y in range(h)) if rle else b.tobytes()
    p = tmp_path / "x.hdr"
    with open(p, "wb") as f:
        f.write(b"#?RADIANCE\n# test\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode() + body)
    got = load_image(str(p))
    assert got.shape == (h, w, 4)
    np.testing.assert_allclose(got[..., :3], _rgbe_decode(b), rtol=1e-6)
    # RGBE keeps 8 bits relative to the brightest channel of a texel
    err = np.abs(got[..., :3] - img).max(-1)
    assert np.all(err <= img.max(-1) / 128 + 1e-30)
    assert np.all(got[..., 3] == 1)


@pytest.mark.parametrize("little", [True, False])
def test_pfm_and_pnm(tmp_path, little):
    rng = np.random.default_rng(4)
    img = rng.standard_normal((3, 5, 3)).astype(np.float32)
    _pfm(tmp_path / "c.pfm", img, little)
    np.testing.assert_array_equal(load_image(str(tmp_path / "c.pfm"))[..., :3], img)
    grey = rng.random((4, 2, 1)).astype(np.float32)
    _pfm(tmp_path / "g.pfm", grey, little)
    np.testing.assert_array_equal(load_image(str(tmp_path / "g.pfm"))[..., :3], np.repeat(grey, 3, -1))
    px = rng.integers(0, 256, (3, 4, 3)).astype(np.uint8)
    _pnm(tmp_path / "c.ppm", px)
    np.testing.assert_allclose(load_image(str(tmp_path / "c.ppm"))[..., :3], (px / 255.0) ** 2.2, rtol=1e-5, atol=1e-7)
    g16 = rng.integers(0, 4096, (2, 3)).astype(np.uint16)
    _pnm(tmp_path / "g.pgm", g16, maxval=4095)
    np.testing.assert_allclose(load_image(str(tmp_path / "g.pgm"))[..., 0], (g16 / 4095.0) ** 2.2, rtol=1e-5,
                               atol=1e-7)


def _png(path, arr, ctype, depth=8, palette=None, trns=None, seed=0, interlace=False):
    """A PNG written here: rows packed at `depth`, each with a random filter
    type (the filter arithmetic of ISO/IEC 15948 applied forward); with
    `interlace`, the seven Adam7 sub-images one after the other."""
    import zlib
    rng = np.random.default_rng(seed)
    h, w = arr.shape[:2]
    chans = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    arr = arr.reshape(h, w, chans)
    if interlace:
        subs = [arr[y0::dy, x0::dx] for x0, y0, dx, dy in
                ((0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2))]
    else:
        subs = [arr]
    raw = b"".join(_png_rows(sub, chans, depth, rng) for sub in subs if sub.size)
    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xffffffff)
    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0,
                                                                 1 if interlace else 0))
    if palette is not None:
        png += chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        png += chunk(b"tRNS", bytes(np.asarray(trns, np.uint8)))
    png += chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b"")
    with open(path, "wb") as fh:
        fh.write(png)


def _png_rows(arr, chans, depth, rng):
    h, w = arr.shape[:2]
    a = arr.reshape(h, w * chans).astype(np.int64)
    if depth == 16:
        rows = [np.stack([r >> 8, r & 255], -1).reshape(-1) for r in a]
    elif depth == 8:
        rows = list(a)
    else:
        per = 8 // depth
        rows = []
        for r in a:
            pad = np.zeros(-(-len(r) // per) * per, np.int64)
            pad[:len(r)] = r
            rows.append(np.array([sum(int(v) << (8 - depth * (k + 1)) for k, v in enumerate(pad[i:i + per]))
                                  for i in range(0, len(pad), per)]))
    bpp = max(1, chans * depth // 8)
    raw, prev = b"", np.zeros(len(rows[0]), np.int64)
    for r in rows:
        f = int(rng.integers(0, 5))
        left = np.concatenate([np.zeros(bpp, np.int64), r[:-bpp]])
        ul = np.concatenate([np.zeros(bpp, np.int64), prev[:-bpp]])
        if f == 0:
            e = r
        elif f == 1:
            e = r - left
        elif f == 2:
            e = r - prev
        elif f == 3:
            e = r - (left + prev) // 2
        else:
            p_ = left + prev - ul
            pa, pb, pc = np.abs(p_ - left), np.abs(p_ - prev), np.abs(p_ - ul)
            e = r - np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, ul))
        raw += bytes([f]) + bytes((e % 256).astype(np.uint8))
        prev = r
    return raw


@pytest.mark.parametrize("ctype,depth,interlace", [(2, 8, False), (6, 8, False), (0, 8, False), (4, 8, False),
                                                   (2, 16, False), (0, 4, False), (0, 1, False), (3, 8, False),
                                                   (3, 2, False), (6, 8, True), (0, 2, True), (3, 4, True)])
def test_png(tmp_path, ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 100 + depth)
    h, w = (11, 14) if interlace else (7, 13)
    chans = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    top = (1 << depth) - 1
    if ctype == 3:
        pal = rng.integers(0, 256, (1 << depth, 3))
        trns = rng.integers(0, 256, 1 << depth)
        idx = rng.integers(0, 1 << depth, (h, w))
        _png(tmp_path / "p.png", idx, ctype, depth, pal, trns, interlace=interlace)
        got = load_image(str(tmp_path / "p.png"))
        np.testing.assert_allclose(got[..., :3], (pal[idx] / 255.0) ** 2.2, rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(got[..., 3], trns[idx] / 255.0, rtol=1e-6)
        return
    arr = rng.integers(0, top + 1, (h, w, chans))
    _png(tmp_path / "p.png", arr, ctype, depth, interlace=interlace)
    got = load_image(str(tmp_path / "p.png"))
    v8 = arr >> 8 if depth == 16 else arr * (255 // top)
    col = (v8 / 255.0) ** 2.2
    if chans <= 2:
        np.testing.assert_allclose(got[..., :3], np.repeat(col[..., :1], 3, -1), rtol=1e-5, atol=1e-7)
    else:
        np.testing.assert_allclose(got[..., :3], col[..., :3], rtol=1e-5, atol=1e-7)
    alpha = v8[..., -1] / 255.0 if chans in (2, 4) else np.ones((h, w))
    np.testing.assert_allclose(got[..., 3], alpha, rtol=1e-6)


def _exr(path, img, comp, half):
    """A scanline OpenEXR file written here: channels in name order, each
    line channel by channel; ZIP (16 lines) and ZIPS (1 line) chunks split
    into even and odd bytes and delta-coded before zlib."""
    import zlib
    h, w, c = img.shape
    names = ["R", "G", "B", "A"][:c] if c > 1 else ["Y"]
    order = sorted(range(len(names)), key=lambda k: names[k])
    ptype = 1 if half else 2
    def attr(name, typ, data):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data
    chl = b"".join(names[k].encode() + b"\0" + struct.pack("<iB3xii", ptype, 0, 1, 1) for k in order) + b"\0"
    hdr = (b"\x76\x2f\x31\x01" + struct.pack("<BBBB", 2, 0, 0, 0) + attr("channels", "chlist", chl)
           + attr("compression", "compression", bytes([comp]))
           + attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
           + attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
           + attr("lineOrder", "lineOrder", b"\0") + attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
           + attr("screenWindowCenter", "v2f", struct.pack("<2f", 0, 0))
           + attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")
    lines = 16 if comp == 3 else 1
    dt = "<f2" if half else "<f4"
    chunks = []
    for y0 in range(0, h, lines):
        raw = b"".join(img[y, :, k].astype(dt).tobytes() for y in range(y0, min(h, y0 + lines)) for k in order)
        data = raw
        if comp in (1, 2, 3):
            a = np.frombuffer(raw, np.uint8).astype(np.int64)
            t = np.concatenate([a[0::2], a[1::2]])
            d = t.copy()
            d[1:] = (t[1:] - t[:-1] + 128) % 256
            if comp == 1:  # runs of >= 3 equal bytes: (n - 1, value); literals: (-n, bytes)
                z, i, n = bytearray(), 0, len(d)
                while i < n:
                    j = i
                    while j < n and j - i < 128 and d[j] == d[i]:
                        j += 1
                    if j - i >= 3:
                        z += bytes([j - i - 1, int(d[i])])
                        i = j
                        continue
                    k = i
                    while k < n and k - i < 127 and not (k + 2 < n and d[k] == d[k + 1] == d[k + 2]):
                        k += 1
                    z += bytes([(256 - (k - i)) % 256]) + bytes(d[i:k].astype(np.uint8))
                    i = k
                z = bytes(z)
            else:
                z = zlib.compress(bytes(d.astype(np.uint8)))
            data = z if len(z) < len(raw) else raw
        chunks.append(struct.pack("<ii", y0, len(data)) + data)
    table_end = len(hdr) + 8 * len(chunks)
    offs, pos = [], table_end
    for ch in chunks:
        offs.append(pos)
        pos += len(ch)
    with open(path, "wb") as f:
        f.write(hdr + struct.pack(f"<{len(offs)}Q", *offs) + b"".join(chunks))


@pytest.mark.parametrize("comp,half,chans", [(0, True, 3), (0, False, 4), (1, True, 3), (2, True, 4), (3, True, 3),
                                             (3, False, 3), (3, True, 1)])
def test_openexr(tmp_path, comp, half, chans):
    rng = np.random.default_rng(comp * 10 + chans)
    h, w = 37, 23
    img = (rng.standard_normal((h, w, chans)) * 10 ** rng.uniform(-3, 3, (h, w, 1))).astype(np.float32)
    img[0, :4] = 0.0
    img[2, :] = 1.5                          # runs
    img[1, :3, 0] = [1e-7, -6e-8, 65504.0]   # half subnormals and the largest half
    if half:
        img = img.astype(np.float16).astype(np.float32)
    _exr(tmp_path / "x.exr", img, comp, half)
    got = load_image(str(tmp_path / "x.exr"))
    want = np.ones((h, w, 4), np.float32)
    if chans == 1:
        want[..., :3] = img
    else:
        want[..., :chans] = img
    np.testing.assert_array_equal(got, want)


def _tga(path, img, kind, rle=False, top=False, cmap=None):
    """A TGA written here: kind 2 true colour (BGR or BGRA), 3 grey, 1
    colour-mapped (8-bit indices); run-length packets of up to 128 texels."""
    h, w = img.shape[:2]
    if kind == 2:
        bpp = img.shape[2]
        px = img[..., [2, 1, 0, 3][:bpp]]
    else:
        bpp = 1
        px = img.reshape(h, w, 1)
    rows = px if top else px[::-1]
    flat = rows.reshape(-1, bpp).astype(np.uint8)
    if rle:
        body, i, n = bytearray(), 0, len(flat)
        while i < n:
            j = i
            while j < n and j - i < 128 and np.array_equal(flat[j], flat[i]):
                j += 1
            if j - i >= 2:
                body += bytes([127 + (j - i)]) + bytes(flat[i])
                i = j
            else:
                k = i
                while k < n and k - i < 128 and not (k + 1 < n and np.array_equal(flat[k], flat[k + 1])):
                    k += 1
                body += bytes([k - i - 1]) + bytes(flat[i:k].reshape(-1))
                i = k
        body = bytes(body)
    else:
        body = bytes(flat.reshape(-1))
    cm = b""
    if kind == 1:
        cm = bytes(np.asarray(cmap, np.uint8)[:, [2, 1, 0]].reshape(-1))
    hdr = struct.pack("<BBBHHBHHHHBB", 0, 1 if kind == 1 else 0, kind + (8 if rle else 0), 0,
                      len(cmap) if kind == 1 else 0, 24 if kind == 1 else 0, 0, 0, w, h, bpp * 8,
                      (0x20 if top else 0) | (8 if bpp == 4 else 0))
    with open(path, "wb") as f:
        f.write(hdr + cm + body)


@pytest.mark.parametrize("kind,chans,rle,top", [(2, 3, False, False), (2, 4, True, True), (3, 1, True, False),
                                                (1, 1, False, True), (1, 1, True, False)])
def test_tga(tmp_path, kind, chans, rle, top):
    rng = np.random.default_rng(kind * 10 + chans)
    h, w = 6, 9
    if kind == 1:
        pal = rng.integers(0, 256, (16, 3))
        idx = rng.integers(0, 16, (h, w))
        idx[2, :] = 5
        _tga(tmp_path / "t.tga", idx, 1, rle, top, pal)
        np.testing.assert_allclose(load_image(str(tmp_path / "t.tga"))[..., :3], (pal[idx] / 255.0) ** 2.2,
                                   rtol=1e-5, atol=1e-7)
        return
    img = rng.integers(0, 256, (h, w, chans))
    img[1, :] = img[1, 0]  # runs
    _tga(tmp_path / "t.tga", img if chans > 1 else img[..., 0], kind, rle, top)
    got = load_image(str(tmp_path / "t.tga"))
    col = (img[..., :3] / 255.0) ** 2.2 if chans > 1 else np.repeat((img / 255.0) ** 2.2, 3, -1)
    np.testing.assert_allclose(got[..., :3], col, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(got[..., 3], img[..., 3] / 255.0 if chans == 4 else 1.0, rtol=1e-6)


# JPEG: the Huffman tables of ITU T.81 Annex K.3 (luminance DC and AC)
_DC_BITS = [0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0]
_DC_VALS = list(range(12))
_AC_BITS = [0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d]
_AC_VALS = [
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07,
    0x22, 0x71, 0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0,
    0x24, 0x33, 0x62, 0x72, 0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28,
    0x29, 0x2a, 0x34, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49,
    0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69,
    0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89,
    0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7,
    0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5,
    0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
    0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8,
    0xf9, 0xfa]
_ZZ = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14,
       21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53,
       60, 61, 54, 47, 55, 62, 63]


def _dct_matrix():
    c = np.array([[(np.sqrt(0.5) if u == 0 else 1.0) * np.cos((2 * x + 1) * u * np.pi / 16) for x in range(8)]
                  for u in range(8)]) / 2
    return c  # F = c f c^T, f = c^T F c


def _huff_codes(bits, vals):
    codes, code, k = {}, 0, 0
    for l in range(1, 17):
        for _ in range(bits[l - 1]):
            codes[vals[k]] = (code, l)
            code += 1
            k += 1
        code <<= 1
    return codes


def _jpeg(path, comps, sampling, q=2, restart=0, grey_factors=None):
    """A baseline JPEG written here: `comps` = planes of 8-bit samples (Y, or
    Y Cb Cr with chroma at 1/sampling resolution), uniform quantiser q, the
    Annex K luminance tables for every component.  Returns the dequantised
    coefficient planes (what a decoder must invert)."""
    C = _dct_matrix()
    dc, ac = _huff_codes(_DC_BITS, _DC_VALS), _huff_codes(_AC_BITS, _AC_VALS)
    H, W = comps[0].shape
    hs = [sampling, 1, 1][:len(comps)]
    mcu = 8 * sampling
    mx, my = -(-W // mcu), -(-H // mcu)
    out_bits, deq = [], [np.zeros((my * mcu // (sampling // h), mx * mcu // (sampling // h))) for h in hs]
    def put(v, n):
        for i in range(n - 1, -1, -1):
            out_bits.append((v >> i) & 1)
    def size(v):
        return 0 if v == 0 else int(abs(v)).bit_length()
    pred = [0] * len(comps)
    data = bytearray()
    def flush_bits():
        while len(out_bits) % 8:
            out_bits.append(1)
        for i in range(0, len(out_bits), 8):
            byte = int("".join(map(str, out_bits[i:i + 8])), 2)
            data.append(byte)
            if byte == 0xFF:
                data.append(0)
        out_bits.clear()
    n_mcu = 0
    for by in range(my):
        for bx in range(mx):
            if restart and n_mcu and n_mcu % restart == 0:
                flush_bits()
                data.extend([0xFF, 0xD0 + ((n_mcu // restart - 1) % 8)])
                pred = [0] * len(comps)
            n_mcu += 1
            for ci, (plane, h) in enumerate(zip(comps, hs)):
                for sy in range(h):
                    for sx in range(h):
                        y0, x0 = (by * h + sy) * 8, (bx * h + sx) * 8
                        blk = np.zeros((8, 8))
                        ph, pw = plane.shape
                        for y in range(8):
                            for x in range(8):
                                blk[y, x] = plane[min(y0 + y, ph - 1), min(x0 + x, pw - 1)]
                        F = np.rint(C @ (blk - 128.0) @ C.T / q).astype(int)
                        deq[ci][y0:y0 + 8, x0:x0 + 8] = F * q
                        zz = F.reshape(-1)[_ZZ]
                        d = int(zz[0]) - pred[ci]
                        pred[ci] = int(zz[0])
                        code, l = dc[size(d)]
                        put(code, l)
                        put(d if d >= 0 else d + (1 << size(d)) - 1, size(d))
                        run = 0
                        for k in range(1, 64):
                            v = int(zz[k])
                            if v == 0:
                                run += 1
                                continue
                            while run > 15:
                                put(*ac[0xF0])
                                run -= 16
                            code, l = ac[(run << 4) | size(v)]
                            put(code, l)
                            put(v if v >= 0 else v + (1 << size(v)) - 1, size(v))
                            run = 0
                        if run:
                            put(*ac[0x00])
    flush_bits()
    def seg(m, payload):
        return bytes([0xFF, m]) + struct.pack(">H", len(payload) + 2) + payload
    hdr = b"\xff\xd8" + seg(0xDB, bytes([0]) + bytes([q] * 64))
    nc = len(comps)
    sof = struct.pack(">BHHB", 8, H, W, nc) + b"".join(
        bytes([i + 1, (h << 4) | h, 0]) for i, h in enumerate(hs))
    if grey_factors is not None:  # a one-component frame stating other factors (the scan is unchanged)
        assert nc == 1
        sof = sof[:-2] + bytes([grey_factors, 0])
    hdr += seg(0xC0, sof)
    hdr += seg(0xC4, bytes([0x00]) + bytes(_DC_BITS) + bytes(_DC_VALS))
    hdr += seg(0xC4, bytes([0x10]) + bytes(_AC_BITS) + bytes(_AC_VALS))
    if restart:
        hdr += seg(0xDD, struct.pack(">H", restart))
    hdr += seg(0xDA, bytes([nc]) + b"".join(bytes([i + 1, 0x00]) for i in range(nc)) + bytes([0, 63, 0]))
    with open(path, "wb") as f:
        f.write(hdr + bytes(data) + b"\xff\xd9")
    return deq, C


@pytest.mark.parametrize("kind,sampling,restart", [("grey", 1, 0), ("colour", 1, 0), ("colour", 2, 0),
                                                   ("colour", 2, 3), ("grey", 1, 5)])
def test_baseline_jpeg(tmp_path, kind, sampling, restart):
    """Decoded texels against the inverse DCT of the coefficients written
    (ITU T.81 A.3.3, level shift, JFIF YCbCr), within one 8-bit step, and
    against the source image within the quantiser's error."""
    rng = np.random.default_rng(sampling * 10 + restart)
    H, W = 21, 30
    yy, xx = np.mgrid[0:H, 0:W]
    smooth = lambda ph: 128 + 90 * np.sin(xx / 5.0 + ph) * np.cos(yy / 7.0 - ph)  # noqa: E731
    Y = np.clip(smooth(0.3) + rng.normal(0, 3, (H, W)), 0, 255).round()
    if kind == "grey":
        comps = [Y]
    else:
        cb = np.clip(smooth(1.1), 0, 255).round()[::sampling, ::sampling]
        cr = np.clip(smooth(2.0), 0, 255).round()[::sampling, ::sampling]
        comps = [Y, cb, cr]
    deq, C = _jpeg(tmp_path / "x.jpg", comps, sampling, q=2, restart=restart)
    got = load_image(str(tmp_path / "x.jpg"))
    assert got.shape == (H, W, 4)
    def plane(ci):
        P = deq[ci]
        out = np.zeros_like(P)
        for y0 in range(0, P.shape[0], 8):
            for x0 in range(0, P.shape[1], 8):
                out[y0:y0 + 8, x0:x0 + 8] = C.T @ P[y0:y0 + 8, x0:x0 + 8] @ C
        return np.clip(np.rint(out + 128.0), 0, 255)
    Yd = plane(0)[:H, :W]
    if kind == "grey":
        want = np.repeat(Yd[..., None], 3, -1)
    else:
        s = sampling
        cbd = plane(1).repeat(s, 0).repeat(s, 1)[:H, :W] - 128.0
        crd = plane(2).repeat(s, 0).repeat(s, 1)[:H, :W] - 128.0
        want = np.clip(np.rint(np.stack([Yd + 1.402 * crd, Yd - 0.344136 * cbd - 0.714136 * crd, Yd + 1.772 * cbd],
                                        -1)), 0, 255)
    v8 = got[..., :3] ** (1 / 2.2) * 255.0
    assert np.max(np.abs(v8 - want)) <= 1.01
    assert np.max(np.abs(v8[..., 0] - (Y if kind == "grey" else want[..., 0]))) <= 8.0
    assert np.all(got[..., 3] == 1.0)


@pytest.mark.parametrize("factors", [0x22, 0x21, 0x12])
def test_grey_jpeg_ignores_sampling_factors(tmp_path, factors):
    """A one-component scan is not interleaved: one 8x8 block per MCU in
    raster order whatever the frame's sampling factors (T.81 A.2.2)."""
    rng = np.random.default_rng(factors)
    Y = np.clip(120 + 60 * rng.standard_normal((27, 35)), 0, 255).round()
    _jpeg(tmp_path / "a.jpg", [Y], 1)
    _jpeg(tmp_path / "b.jpg", [Y], 1, grey_factors=factors)
    np.testing.assert_array_equal(load_image(str(tmp_path / "b.jpg")), load_image(str(tmp_path / "a.jpg")))


def test_exr_chunk_offset_near_2_64(tmp_path):
    """An offset table entry close to 2^64 must not wrap past the size check."""
    img = np.ones((4, 5, 3), np.float32)
    _exr(tmp_path / "ok.exr", img, 0, True)
    data = bytearray(open(tmp_path / "ok.exr", "rb").read())
    load_image(str(tmp_path / "ok.exr"))
    # the offset table follows the header (4 chunks of one line, no compression)
    first = struct.unpack_from("<Q", data, len(data) - 4 * (8 + 5 * 3 * 2) - 4 * 8)[0]
    assert first == len(data) - 4 * (8 + 5 * 3 * 2)
    struct.pack_into("<Q", data, len(data) - 4 * (8 + 5 * 3 * 2) - 4 * 8, 0xFFFFFFFFFFFFFFF8)
    with open(tmp_path / "bad.exr", "wb") as f:
        f.write(bytes(data))
    with pytest.raises(abi.NoorError) as e:
        load_image(str(tmp_path / "bad.exr"))
    assert e.value.status == abi.ERR_INVALID_PAYLOAD and "chunk offset" in str(e.value)


def test_header_sizes_are_capped(tmp_path):
    """Headers stating more than 2^28 pixels are refused before any allocation."""
    cases = {
        "big.hdr": b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 65536 +X 65536\n",
        "big.pfm": b"PF\n65536 65536\n-1.0\n",
        "big.ppm": b"P6\n65536 65536\n255\n",
        "big.tga": bytes([0, 0, 2]) + bytes(9) + struct.pack("<HH", 65535, 65535) + bytes([24, 0]),
        "big.png": b"\x89PNG\r\n\x1a\n" + struct.pack(">I", 13) + b"IHDR" +
                   struct.pack(">IIBBBBB", 65536, 65536, 8, 2, 0, 0, 0) + b"\0\0\0\0",
    }
    for name, data in cases.items():
        with open(tmp_path / name, "wb") as f:
            f.write(data)
        with pytest.raises(abi.NoorError) as e:
            load_image(str(tmp_path / name))
        assert e.value.status == abi.ERR_INVALID_PAYLOAD and "too large" in str(e.value), name


def test_bad_files_fail_cleanly(tmp_path):
    _write(tmp_path / "bad.obj", "v 0 0 0\nf 1 2 3\n")
    with pytest.raises(abi.NoorError):
        ObjModel(str(tmp_path / "bad.obj"))
    with pytest.raises(abi.NoorError):
        ObjModel(str(tmp_path / "missing.obj"))
    with open(tmp_path / "t.hdr", "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 4 +X 4\n\x01\x02")
    with pytest.raises(abi.NoorError):
        load_image(str(tmp_path / "t.hdr"))
    with open(tmp_path / "x.gif", "wb") as f:
        f.write(b"GIF89a")
    with pytest.raises(abi.NoorError) as e:
        load_image(str(tmp_path / "x.gif"))
    assert e.value.status == abi.ERR_UNSUPPORTED
    _jpeg(tmp_path / "t.jpg", [np.full((64, 64), 77.0)], 1)
    data = open(tmp_path / "t.jpg", "rb").read()
    with open(tmp_path / "t.jpg", "wb") as f:  # entropy-coded data cut short
        f.write(data[:len(data) // 2])
    with pytest.raises(abi.NoorError) as e:
        load_image(str(tmp_path / "t.jpg"))
    assert e.value.status == abi.ERR_INVALID_PAYLOAD
    for seg in (b"\xff\xc0\x00\x02", b"\xff\xdd\x00\x02", b"\xff\xda\x00\x02"):  # segments cut short
        with open(tmp_path / "s.jpg", "wb") as f:
            f.write(b"\xff\xd8" + seg)
        with pytest.raises(abi.NoorError) as e:
            load_image(str(tmp_path / "s.jpg"))
        assert e.value.status == abi.ERR_INVALID_PAYLOAD
    # a valid frame header, then scan headers shorter than they claim
    sof = b"\xff\xc0\x00\x0b\x08\x00\x10\x00\x10\x01\x01\x11\x00"
    for sos in (b"\xff\xda\x00\x02", b"\xff\xda\x00\x04\x03\x01"):  # empty; Ns = 3 in 2 bytes
        with open(tmp_path / "q.jpg", "wb") as f:
            f.write(b"\xff\xd8" + sof + sos)
        with pytest.raises(abi.NoorError) as e:
            load_image(str(tmp_path / "q.jpg"))
        assert e.value.status == abi.ERR_INVALID_PAYLOAD and "bad scan header" in str(e.value)
    with open(tmp_path / "p.jpg", "wb") as f:  # progressive: refused
        f.write(b"\xff\xd8\xff\xc2\x00\x0b\x08\x00\x10\x00\x10\x01\x01\x11\x00")
    with pytest.raises(abi.NoorError) as e:
        load_image(str(tmp_path / "p.jpg"))
    assert e.value.status == abi.ERR_UNSUPPORTED
    with open(tmp_path / "x.png", "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n\x00\x00")
    with pytest.raises(abi.NoorError):
        load_image(str(tmp_path / "x.png"))


def test_obj_scene_renders_with_the_oracle(cube_dir):
    from noor_amd.scene import make_camera, make_options, make_spec
    from noor_amd.scenes import Workload
    from oracle.binding import OracleScene
    sb, m = obj_scene(str(cube_dir / "cube.obj"))
    payload = sb.build()
    assert payload.view.num_triangles == 14 and payload.view.num_area_lights == 1
    cam = make_camera((2.0, 1.5, 3.0), (0.0, 0.0, 0.0), (0, 1, 0), np.radians(50.0), 24, 20)
    spec = make_spec(payload, bounces=3, rr=3, skydome_type=abi.SKY_HDR)
    wl = Workload("obj", payload, cam, spec, make_options(), None, 1, info={})
    f, _ = OracleScene.from_workload(wl).render(1, 2)
    assert np.isfinite(f).all() and f[..., :3].max() > 0
