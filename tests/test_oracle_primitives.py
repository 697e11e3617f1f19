"""Pins of the CPU oracle's primitives against known answers and closed forms.

Each check is chosen so that a plausible mistake (a wrong constant, a swapped
operand, a dropped term) fails it:
  - TEA: the published 32-cycle test vector of the cipher; the 16-cycle seed
    of the reference uses the same round function with its own key.
  - LCG: Numerical Recipes constants, stepped independently here.
  - PCG32: the published pcg32-demo output for seed 42, stream 54.
  - portable transcendentals: ulp error against libm double precision.
  - Hosek-Wilkie: samples recorded from the reference's own Hosek code.
  - Fresnel: normal-incidence closed forms and total internal reflection.
  - GGX: the normalisation integral of D(wh) cos(theta_h) over the hemisphere.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from noor_amd import abi
from oracle import binding as ob

HERE = os.path.dirname(os.path.abspath(__file__))


def L():
    return ob.lib()


# ---- RNG --------------------------------------------------------------------------------------

def test_tea_published_vector():
    key = (C.c_uint32 * 4)(0, 0, 0, 0)
    v1 = C.c_uint32()
    v0 = L().oracle_tea(0, 0, key, 32, C.byref(v1))
    assert (v0, v1.value) == (0x41EA3A0A, 0x94BAA940)


def _tea_py(v0, v1, key, rounds):
    """TEA encryption written from Wheeler & Needham (1994), 32-bit words."""
    m = 0xFFFFFFFF
    s = 0
    for _ in range(rounds):
        s = (s + 0x9E3779B9) & m
        v0 = (v0 + ((((v1 << 4) + key[0]) & m) ^ ((v1 + s) & m) ^ (((v1 >> 5) + key[1]) & m))) & m
        v1 = (v1 + ((((v0 << 4) + key[2]) & m) ^ ((v0 + s) & m) ^ (((v0 >> 5) + key[3]) & m))) & m
    return v0, v1


def test_tea16_seed_matches_independent_tea():
    key = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)
    rng = np.random.default_rng(0)
    for _ in range(200):
        a, b = (int(x) for x in rng.integers(0, 2**32, 2))
        assert L().oracle_tea16(a, b) == _tea_py(a, b, key, 16)[0]
    # the Python TEA itself reproduces the published vector
    assert _tea_py(0, 0, (0, 0, 0, 0), 32) == (0x41EA3A0A, 0x94BAA940)


def test_lcg_stream():
    key = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)
    gid, frame = 12345, 7
    s = _tea_py(gid, frame, key, 16)[0]
    expect = []
    for _ in range(16):
        s = (1664525 * s + 1013904223) & 0xFFFFFFFF
        expect.append((s & 0xFFFFFF) / float(1 << 24))
    got = ob.rng_sequence(abi.RNG_TEA_LCG, gid, frame, 16)
    np.testing.assert_array_equal(got, np.array(expect, np.float32))


def test_pcg32_published_demo_sequence():
    st = C.c_uint64()
    inc = C.c_uint64()
    L().oracle_pcg32_seed(C.byref(st), C.byref(inc), 42, 54)
    out = [L().oracle_pcg32_next(C.byref(st), inc.value) for _ in range(6)]
    assert out == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_pcg32_stream_floats():
    # the float stream is the top 24 bits of pcg32 outputs for (initstate = gid, initseq = frame)
    gid, frame = 99, 3
    st = C.c_uint64()
    inc = C.c_uint64()
    L().oracle_pcg32_seed(C.byref(st), C.byref(inc), gid, frame)
    expect = [(L().oracle_pcg32_next(C.byref(st), inc.value) >> 8) / float(1 << 24) for _ in range(8)]
    np.testing.assert_array_equal(ob.rng_sequence(abi.RNG_PCG32, gid, frame, 8), np.array(expect, np.float32))
    u = ob.rng_sequence(abi.RNG_PCG32, 1, 1, 200000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.005


# ---- portable math --------------------------------------------------------------------------------

def _ulp_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    sp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    sp = np.maximum(sp, np.float64(np.finfo(np.float32).tiny))
    return np.abs(got - ref) / sp


@pytest.mark.parametrize("name,ref,lo,hi,bound", [
    ("oracle_sinf", np.sin, -8.0, 8.0, 2.0),
    ("oracle_cosf", np.cos, -8.0, 8.0, 2.0),
    ("oracle_acosf", np.arccos, -1.0, 1.0, 2.0),
    ("oracle_expf", np.exp, -87.0, 88.0, 2.0),
])
def test_portable_math_ulp(name, ref, lo, hi, bound):
    f = getattr(L(), name)
    xs = np.random.default_rng(1).uniform(lo, hi, 40000).astype(np.float32)
    got = np.array([f(float(x)) for x in xs], np.float32)
    err = _ulp_err(got, ref(xs.astype(np.float64)))
    assert err.max() <= bound, err.max()


def test_portable_log_pow_atan2():
    rng = np.random.default_rng(2)
    xs = np.exp(rng.uniform(-80, 80, 20000)).astype(np.float32)
    got = np.array([L().oracle_logf(float(x)) for x in xs], np.float32)
    assert _ulp_err(got, np.log(xs.astype(np.float64))).max() <= 2.0
    ys, xs2 = rng.uniform(-5, 5, (2, 20000)).astype(np.float32)
    got = np.array([L().oracle_atan2f(float(y), float(x)) for y, x in zip(ys, xs2)], np.float32)
    assert _ulp_err(got, np.arctan2(ys.astype(np.float64), xs2.astype(np.float64))).max() <= 3.0
    b = rng.uniform(0.01, 2.0, 5000).astype(np.float32)
    got = np.array([L().oracle_powf(float(x), 2.4) for x in b], np.float32)
    rel = np.abs(got / np.power(b.astype(np.float64), 2.4) - 1.0)
    # exp(y log x) amplifies the rounding of y log x by |y log x| (< 2 here)
    assert rel.max() < 3e-6
    # special cases
    assert L().oracle_sinf(0.0) == 0.0 and L().oracle_cosf(0.0) == 1.0
    assert L().oracle_expf(0.0) == 1.0 and L().oracle_logf(1.0) == 0.0
    assert L().oracle_acosf(1.0) == 0.0
    assert math.isnan(L().oracle_acosf(1.5))
    assert L().oracle_powf(0.0, 2.4) == 0.0


# ---- Hosek-Wilkie ----------------------------------------------------------------------------------

def _hosek_state():
    fx = json.load(open(os.path.join(os.path.dirname(HERE), "noor_amd", "data", "hosek_t3_a03_e30.json")))
    st = abi.HosekState()
    for c in range(3):
        for j in range(9):
            st.configs[c][j] = fx["configs"][c][j]
        st.radiances[c] = fx["radiances"][c]
    return st


def test_hosek_radiance_matches_reference_samples():
    st = _hosek_state()
    g = json.load(open(os.path.join(HERE, "golden", "hosek_reference.json")))
    for s in g["samples"]:
        got = [L().oracle_hosek_radiance(C.byref(st), s["theta"], s["gamma"], c) for c in range(3)]
        np.testing.assert_allclose(got, s["rgb"], rtol=2e-6)


def test_hosek_query_properties():
    from noor_amd import scenes
    hk = scenes.hosek_sky()
    out = (C.c_float * 3)()
    # below the horizon the sky is black (noor/cuda/hosek.cuh:88)
    L().oracle_hosek_query(C.byref(hk), (C.c_float * 3)(0.3, -0.2, 0.93), out)
    assert list(out) == [0.0, 0.0, 0.0]
    # towards the sun: solar_scale * radiance; just outside: inverse-sRGB of 0.025 * radiance
    d = (C.c_float * 3)(hk.solar_dir.x, hk.solar_dir.y, hk.solar_dir.z)
    L().oracle_hosek_query(C.byref(hk), d, out)
    sun = np.array(out)
    theta = math.acos(hk.solar_dir.y)
    rad = np.array([L().oracle_hosek_radiance(C.byref(hk.state_r), theta, 0.0, c) for c in range(3)])
    np.testing.assert_allclose(sun, 250.0 * rad, rtol=1e-5)


# ---- Fresnel and GGX -------------------------------------------------------------------------------

def test_fresnel_closed_forms():
    assert ob.fresnel_dielectric(1.0, 1.0, 1.5) == pytest.approx(0.04, abs=1e-7)
    assert ob.fresnel_dielectric(-1.0, 1.0, 1.5) == pytest.approx(0.04, abs=1e-7)
    # total internal reflection from inside glass beyond the critical angle
    assert ob.fresnel_dielectric(-math.cos(math.radians(50.0)), 1.0, 1.5) == 1.0
    # grazing incidence reflects everything
    assert ob.fresnel_dielectric(1e-6, 1.0, 1.5) == pytest.approx(1.0, abs=1e-4)
    eta = np.array([0.2, 0.92, 1.1])
    k = np.array([3.9, 2.45, 2.14])
    expect = ((eta - 1) ** 2 + k ** 2) / ((eta + 1) ** 2 + k ** 2)
    np.testing.assert_allclose(ob.fresnel_conductor(1.0, eta, k), expect, rtol=1e-5)


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.6])
def test_ggx_normalisation(alpha):
    # integral over the hemisphere of D(wh) cos(theta_h) d(omega) = 1
    n = 400
    th = (np.arange(n) + 0.5) / n * (math.pi / 2)
    ph = (np.arange(64) + 0.5) / 64 * 2 * math.pi
    total = 0.0
    for t in th:
        wh = np.array([math.sin(t) * math.cos(ph[0]), math.sin(t) * math.sin(ph[0]), math.cos(t)], np.float32)
        total += ob.ggx_D(alpha, alpha, wh) * math.cos(t) * math.sin(t) * (math.pi / 2 / n) * 2 * math.pi
    assert total == pytest.approx(1.0, rel=2e-2 if alpha < 0.1 else 5e-3)


def test_lambert_closed_form_and_reciprocity():
    p = ob.material_params(diffuse=(0.3, 0.5, 0.7))
    wo = np.array([0.3, 0.2, 0.932738], np.float32)
    wi = np.array([-0.5, 0.1, 0.860233], np.float32)
    f, pdf = ob.bsdf_eval(abi.MAT_DIFFUSE, p, wo, wi)
    np.testing.assert_allclose(f, np.array([0.3, 0.5, 0.7]) / np.float32(3.141592654), rtol=1e-6)
    assert pdf == pytest.approx(wi[2] / math.pi, rel=1e-6)
    g, _ = ob.bsdf_eval(abi.MAT_DIFFUSE, p, wi, wo)
    np.testing.assert_array_equal(f, g)
    # below the surface: no reflection
    f2, pdf2 = ob.bsdf_eval(abi.MAT_DIFFUSE, p, wo, np.array([0.1, 0.1, -0.99], np.float32))
    assert not f2.any() and pdf2 == 0.0


@pytest.mark.parametrize("mtype,params", [
    (abi.MAT_DIFFUSE, dict()),
    (abi.MAT_METAL, dict(eta=(0.2, 0.92, 1.1), k=(3.9, 2.45, 2.14), rough=(0.3, 0.3))),
    (abi.MAT_GLOSSY, dict(rough=(0.25, 0.25), specular=(0.5, 0.5, 0.5))),
    (abi.MAT_SUBSTRATE, dict(rough=(0.3, 0.3), specular=(0.2, 0.3, 0.4), diffuse=(0.6, 0.5, 0.4))),
    (abi.MAT_ROUGHGLASS, dict(rough=(0.35, 0.35), specular=(0.9, 0.9, 0.9), transmission=(0.8, 0.9, 1.0))),
])
def test_bsdf_sampling_pdf_consistency(mtype, params):
    """Sampled directions follow the reported pdf: a chi-square test on a
    (cos theta, phi) grid, and the sampled f / pdf agree with eval."""
    p = ob.material_params(**params)
    wo = np.array([0.4, -0.2, 0.894427], np.float32)
    nt, nph = 8, 8
    counts = np.zeros((nt, nph))
    n = 30000
    for i in range(n):
        wi, f, pdf, flags = ob.bsdf_sample(mtype, p, wo, i, 1)
        if pdf == 0.0:
            continue
        fe, pe = ob.bsdf_eval(mtype, p, wo, wi)
        assert pe == pytest.approx(pdf, rel=2e-3, abs=1e-6)
        np.testing.assert_allclose(fe, f, rtol=2e-3, atol=1e-6)
        if wi[2] <= 0:
            continue
        ti = min(int(wi[2] * nt), nt - 1)
        pi_ = min(int((math.atan2(wi[1], wi[0]) + math.pi) / (2 * math.pi) * nph), nph - 1)
        counts[ti, pi_] += 1
    # expected counts by integrating the pdf over each cell (midpoint rule on a finer grid)
    expect = np.zeros((nt, nph))
    sub = 6
    for ti in range(nt):
        for pj in range(nph):
            acc = 0.0
            for a in range(sub):
                for b in range(sub):
                    c = (ti + (a + 0.5) / sub) / nt
                    ph = ((pj + (b + 0.5) / sub) / nph) * 2 * math.pi - math.pi
                    s = math.sqrt(max(0.0, 1 - c * c))
                    wi = np.array([s * math.cos(ph), s * math.sin(ph), c], np.float32)
                    acc += ob.bsdf_eval(mtype, p, wo, wi)[1]
            expect[ti, pj] = acc / sub / sub * (1.0 / nt) * (2 * math.pi / nph) * n
    mask = expect > 20
    chi2 = np.sum((counts[mask] - expect[mask]) ** 2 / expect[mask])
    dof = mask.sum() - 1
    assert chi2 < dof + 6 * math.sqrt(2 * dof), (chi2, dof)


def test_white_furnace_lambert():
    # albedo-1 Lambert: E[f cos / pdf] = 1 for every sample
    p = ob.material_params(diffuse=(1.0, 1.0, 1.0))
    wo = np.array([0.0, 0.6, 0.8], np.float32)
    for i in range(200):
        wi, f, pdf, _ = ob.bsdf_sample(abi.MAT_DIFFUSE, p, wo, i, 2)
        assert f[0] * abs(wi[2]) / pdf == pytest.approx(1.0, rel=1e-5)


def _sphere_quadrature(fn, n_t=96, n_p=96):
    """Integral of fn(wi) over the unit sphere (midpoint rule in cos(theta), phi)."""
    acc = 0.0
    for i in range(n_t):
        c = -1 + 2 * (i + 0.5) / n_t
        s_ = math.sqrt(max(0.0, 1 - c * c))
        for j in range(n_p):
            ph = 2 * math.pi * (j + 0.5) / n_p
            acc += fn(np.array([s_ * math.cos(ph), s_ * math.sin(ph), c], np.float32))
    return acc * (2.0 / n_t) * (2 * math.pi / n_p)


def test_fresnel_blend_pdf_integrates_to_one():
    # half cosine, half visible-normal reflection (a little of it below the horizon)
    p = ob.material_params(rough=(0.3, 0.3))
    wo = np.array([0.3, 0.2, 0.932738], np.float32)
    total = _sphere_quadrature(lambda wi: ob.bsdf_eval(abi.MAT_SUBSTRATE, p, wo, wi)[1])
    assert 0.95 <= total <= 1.005, total


def test_rough_glass_lobe_pdfs():
    """The reflection lobe integrates to at most one; the transmission lobe
    integrates to one over the directions refraction can produce (half vector
    between wo and -wi).  Outside that set the reference's (PBRT v3) pdf is not
    zero, so its integral over the whole lower hemisphere exceeds one."""
    p = ob.material_params(rough=(0.3, 0.3), coat_ior=1.5)
    wo = np.array([0.3, 0.2, 0.932738], np.float32)

    def refl(wi):
        return 2 * ob.bsdf_eval(abi.MAT_ROUGHGLASS, p, wo, wi)[1] if wi[2] > 0 else 0.0

    def trans(wi):
        if wi[2] >= 0:
            return 0.0
        wh = wo.astype(np.float64) + 1.5 * wi
        wh /= np.linalg.norm(wh)
        wh *= np.sign(wh[2])
        if not (wo @ wh > 0 and wi @ wh < 0):
            return 0.0
        return 2 * ob.bsdf_eval(abi.MAT_ROUGHGLASS, p, wo, wi)[1]

    r = _sphere_quadrature(refl, 64, 64)
    t = _sphere_quadrature(trans, 192, 96)
    assert 0.9 <= r <= 1.005, r
    assert t == pytest.approx(1.0, abs=0.02), t


def test_fresnel_blend_closed_forms():
    """FresnelBlend (noor/cuda/bxdf.cuh:669-752): at normal incidence and
    reflection the diffuse term is 28/(23 pi) R (1 - S) (1 - 0.5^5)^2 and
    the Schlick term reduces to S."""
    R, S = (0.6, 0.5, 0.4), (0.2, 0.3, 0.4)
    p = ob.material_params(diffuse=R, specular=S, rough=(0.4, 0.4))
    wo = np.array([0.0, 0.0, 1.0], np.float32)
    wi = np.array([0.0, 0.0, 1.0], np.float32)
    f, _ = ob.bsdf_eval(abi.MAT_SUBSTRATE, p, wo, wi)
    a = 0.4 * 0.4
    D = 1.0 / (math.pi * a * a)  # GGX at the half vector = normal
    spec = D / (4 * 1 * 1) * np.array(S)
    diff = 28 / (23 * math.pi) * np.array(R) * (1 - np.array(S)) * (1 - 0.5 ** 5) ** 2
    np.testing.assert_allclose(f, diff + spec, rtol=2e-5)


def test_rough_glass_reciprocity_of_transmission():
    """Microfacet transmission (noor/cuda/bxdf.cuh:559-667) satisfies
    f(wo, wi) / eta_i^2 = f(wi, wo) / eta_o^2 (PBRT's generalised reciprocity)."""
    p = ob.material_params(rough=(0.3, 0.3), transmission=(1, 1, 1), coat_ior=1.5)
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(200):
        a = rng.normal(size=3)
        a[2] = abs(a[2]) + 0.2
        b = rng.normal(size=3)
        b[2] = -abs(b[2]) - 0.2
        wo = (a / np.linalg.norm(a)).astype(np.float32)
        wi = (b / np.linalg.norm(b)).astype(np.float32)
        wh = wo + 1.5 * wi
        wh = wh / np.linalg.norm(wh) * np.sign(wh[2])
        if not (wo @ wh > 0 and wi @ wh < 0):
            continue  # no refraction through this half vector (see the pdf test)
        f1, _ = ob.bsdf_eval(abi.MAT_ROUGHGLASS, p, wo, wi)  # outside -> inside
        f2, _ = ob.bsdf_eval(abi.MAT_ROUGHGLASS, p, wi, wo)  # inside -> outside
        checked += 1
        # the (1 - F) factors match for a shared half vector; the eta^2 factor flips
        assert f1[0] / 1.5 ** 2 == pytest.approx(f2[0] * 1.0, rel=1e-3)  # FP32 at grazing half vectors
    assert checked > 40


def test_clear_coat_energy_and_mirror_coat():
    """ClearCoat (noor/cuda/bxdf.cuh:842-940): coat samples are mirror
    reflections; with white colours and a unit-reflectance conductor the
    sampling weight f |cos| / pdf averages at most one."""
    p = ob.material_params(diffuse=(1, 1, 1), specular=(1, 1, 1), eta=(0.0, 0.0, 0.0), k=(1e4, 1e4, 1e4),
                           rough=(0.3, 0.3), coat_ior=1.5, coat_weight=0.5)
    wo = np.array([0.2, -0.3, 0.932738], np.float32)
    wo /= np.linalg.norm(wo)
    acc, n, mirror = 0.0, 20000, 0
    for i in range(n):
        wi, f, pdf, flags = ob.bsdf_sample(abi.MAT_CLEARCOAT, p, wo, i, 3)
        if pdf == 0.0:
            continue
        if flags & 16:  # specular: the coat
            mirror += 1
            np.testing.assert_allclose(wi, [-wo[0], -wo[1], wo[2]], atol=1e-7)
        acc += f[0] * abs(wi[2]) / pdf
    assert 0.3 < acc / n <= 1.02, acc / n
    assert 0 < mirror < n
