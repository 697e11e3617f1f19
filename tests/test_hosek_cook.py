"""The Hosek-Wilkie cook (noor_hosek_cook_state, noor_hosek_cook) against
what pins it:

- the state recorded from the reference's own cook for turbidity 3, albedo
  0.3, elevation 30 degrees (noor_amd/data/hosek_t3_a03_e30.json), bit for bit;
- the published model's structure, in float64 from the coefficient tables:
  a quintic Bezier curve in s = (elevation / (pi/2))^(1/3) per parameter
  (so elevation 0 and pi/2 return the first and last control points exactly),
  linear in ground albedo between the two albedo tables and linear in
  turbidity between neighbouring integer rows (Hosek & Wilkie 2012, sec. 5;
  noor/external/ArHosekSkyModel.cpp:147-304);
- the sky assembly of HosekSky::update (noor/header/hosek.h:61-83).
"""
import json
import math
import os

import numpy as np
import pytest

from noor_amd import abi
from noor_amd._lib import lib
from noor_amd.scene import cook_hosek, cook_hosek_state

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(os.path.dirname(HERE), "noor_amd", "data")


@pytest.fixture(scope="module")
def tables():
    d = json.load(open(os.path.join(DATA, "hosek_rgb_dataset.json")))
    cfg = np.array([[np.float32(x) for x in t] for t in d["configs"]], np.float64).reshape(3, 2, 10, 6, 9)
    rad = np.array([[np.float32(x) for x in t] for t in d["radiances"]], np.float64).reshape(3, 2, 10, 6)
    return cfg, rad


def state_arrays(st):
    return (np.array([[st.configs[c][i] for i in range(9)] for c in range(3)], np.float64),
            np.array([st.radiances[c] for c in range(3)], np.float64))


def model64(tables, T, a, elev):
    """The blended state in float64: Bernstein weights C(5,k) s^k (1-s)^(5-k)."""
    cfg, rad = tables
    s = (elev / (math.pi / 2)) ** (1.0 / 3.0)
    w = np.array([math.comb(5, k) * s ** k * (1 - s) ** (5 - k) for k in range(6)])
    ti = int(T)
    tr = T - ti
    rows = [(ti - 1, 1 - tr)] + ([(ti, tr)] if ti < 10 else [])
    C = np.zeros((3, 9))
    R = np.zeros(3)
    for r, wt in rows:
        for ai, wa in ((0, 1 - a), (1, a)):
            C += wt * wa * np.einsum("k,ckp->cp", w, cfg[:, ai, r])
            R += wt * wa * (w @ rad[:, ai, r].T)
    return C, R


def test_recorded_state_bit_exact():
    fx = json.load(open(os.path.join(DATA, "hosek_t3_a03_e30.json")))
    st = cook_hosek_state(3.0, 0.3, np.float32(np.radians(30.0)))
    C, R = state_arrays(st)
    np.testing.assert_array_equal(C.astype(np.float32), np.array(fx["configs"], np.float32))
    np.testing.assert_array_equal(R.astype(np.float32), np.array(fx["radiances"], np.float32))
    assert np.float32(st.solar_radius) == np.float32(fx["solar_radius"])
    assert st.turbidity == 3.0 and np.float32(st.albedo) == np.float32(0.3)


@pytest.mark.parametrize("T", [1, 2, 5, 9, 10])
@pytest.mark.parametrize("a", [0, 1])
def test_control_point_endpoints(tables, T, a):
    """Elevation 0 gives control point 0 and pi/2 control point 5 of the
    (albedo, turbidity) row, exactly."""
    cfg, rad = tables
    for elev, k in ((0.0, 0), (np.float32(np.pi / 2), 5)):
        C, R = state_arrays(cook_hosek_state(T, a, elev))
        np.testing.assert_array_equal(C, cfg[:, a, T - 1, k])
        np.testing.assert_array_equal(R, rad[:, a, T - 1, k])


@pytest.mark.parametrize("T,a,elev_deg", [(3.0, 0.3, 30.0), (2.5, 0.0, 10.0), (7.25, 0.8, 55.0),
                                           (9.5, 1.0, 80.0), (10.0, 0.5, 5.0), (1.0, 0.1, 89.0)])
def test_against_float64_model(tables, T, a, elev_deg):
    elev = float(np.float32(np.radians(elev_deg)))
    C, R = state_arrays(cook_hosek_state(T, a, elev))
    C64, R64 = model64(tables, T, float(np.float32(a)), elev)
    scale = np.abs(C64).max(axis=1, keepdims=True)
    assert np.max(np.abs(C - C64) / scale) < 2e-6
    np.testing.assert_allclose(R, R64, rtol=2e-6)
    # planted mistakes: the albedo tables swapped, the turbidity row off by
    # one, s without the cube root
    if 0.0 < a < 1.0 and a != 0.5:
        Cb, _ = model64(tables, T, 1.0 - a, elev)
        assert np.max(np.abs(C - Cb) / scale) > 1e-3
    if T < 9:
        Cb, _ = model64(tables, T + 1, a, elev)
        assert np.max(np.abs(C - Cb) / scale) > 1e-3
    Cb, _ = model64(tables, T, a, (elev / (math.pi / 2)) ** 3 * (math.pi / 2))
    assert np.max(np.abs(C - Cb) / scale) > 1e-3


def test_linear_in_turbidity_and_albedo():
    e = float(np.float32(np.radians(40.0)))
    a0, a1, am = (state_arrays(cook_hosek_state(4.0, a, e)) for a in (0.0, 1.0, 0.25))
    np.testing.assert_allclose(am[0], 0.75 * a0[0] + 0.25 * a1[0], rtol=1e-5, atol=1e-6)
    t0, t1, tm = (state_arrays(cook_hosek_state(T, 0.4, e)) for T in (4.0, 5.0, 4.5))
    np.testing.assert_allclose(tm[1], 0.5 * (t0[1] + t1[1]), rtol=1e-6)


def test_sky_assembly():
    lat, lon = np.float32(np.radians(60.0)), np.float32(np.radians(45.0))
    hk = cook_hosek(3.0, (0.3, 0.3, 0.3), lat, lon)
    d = np.array([hk.solar_dir.x, hk.solar_dir.y, hk.solar_dir.z])
    np.testing.assert_allclose(d, [math.sin(lat) * math.cos(lon), math.cos(lat), math.sin(lat) * math.sin(lon)],
                               atol=1e-6)
    assert abs(np.linalg.norm(d) - 1) < 1e-6
    assert hk.solar_scale == 250.0
    assert np.float32(hk.solar_radius) == np.float32(2.0) * np.float32(hk.state_r.solar_radius)
    # elevation = pi/2 - sun_lat in binary32; grey ground: three equal states
    e = np.float32(np.float32(3.141592654) / np.float32(2.0)) - lat
    assert np.float32(hk.state_r.elevation) == e
    for st in (hk.state_g, hk.state_b):
        np.testing.assert_array_equal(state_arrays(st)[0], state_arrays(hk.state_r)[0])
    # a coloured ground cooks each channel state with its own albedo
    hk2 = cook_hosek(3.0, (0.1, 0.5, 0.9), lat, lon)
    for st, a in ((hk2.state_r, 0.1), (hk2.state_g, 0.5), (hk2.state_b, 0.9)):
        ref = cook_hosek_state(3.0, a, e)
        np.testing.assert_array_equal(state_arrays(st)[0], state_arrays(ref)[0])
        assert np.float32(st.albedo) == np.float32(a)
    assert (hk2.ground_albedo.x, hk2.ground_albedo.z) == (np.float32(0.1), np.float32(0.9))


def test_rejects_out_of_range():
    st = abi.HosekState()
    import ctypes as C
    for args in ((0.5, 0.3, 0.5), (10.5, 0.3, 0.5), (3.0, -0.1, 0.5), (3.0, 1.1, 0.5), (3.0, 0.3, 1.6),
                 (3.0, 0.3, -0.1), (float("nan"), 0.3, 0.5)):
        assert lib().noor_hosek_cook_state(*args, C.byref(st)) == abi.ERR_INVALID_ARG
    hk = abi.Hosek()
    alb = (C.c_float * 3)(0.3, 0.3, 0.3)
    assert lib().noor_hosek_cook(3.0, alb, 1.7, 0.0, 250.0, 2.0, C.byref(hk)) == abi.ERR_INVALID_ARG
    assert lib().noor_hosek_cook(3.0, None, 0.5, 0.0, 250.0, 2.0, C.byref(hk)) == abi.ERR_INVALID_ARG
