"""CPU-side checks of the C-ABI library: it loads, exports every declared entry
point, and its host geometry generator satisfies the invariants that pin it
(the reference's generator library is absent, so its lists are this library's
own definition, documented in tdlbcpp_amd/csrc/rushton.h)."""
import math
import os
import subprocess

import numpy as np
import pytest

import tdlbcpp_amd as td

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    names = td.exported_symbols()
    assert len(names) >= 30
    L = td.lib()
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", td.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(names) <= exported


def test_error_path_reports_message():
    bad = td.CuParams()
    bad.x = bad.y = bad.z = 4
    bad.ghost = 2
    h = td.ctypes.c_void_p()
    rc = td.lib().tdlb_cu_create(td.ctypes.byref(bad), td.ctypes.byref(td.make_flow()), td.ctypes.byref(h))
    assert rc == -6
    assert b"ghost" in td.lib().tdlb_last_error()


@pytest.mark.parametrize("T", [60, 96, 128])
def test_rushton_default_dimensions(T):
    p = td.rushton_defaults(T)
    assert p.blade_outer_radius == int(T / 6)          # integer outerRadius (main.cpp:275)
    assert p.tank_radius == pytest.approx((T - 1) / 2)
    assert p.disk_plane == int(T / 3)                    # clearance T/3
    assert p.blade_outer_radius - p.blade_inner_radius == pytest.approx(T / 3 / 4)  # blade length D/4
    assert p.n_blades == 6 and p.n_baffles == 4


def _xz(pts):
    return pts["i"] + pts["i_frac"], pts["k"] + pts["k_frac"]


@pytest.mark.parametrize("T", [48, 60, 96])
def test_rushton_points_inside_interior_and_fractions_centered(T):
    p = td.rushton_defaults(T)
    for g, (a, d) in [(td.GEOM_FIXED, (0, 0)), (td.GEOM_ROTATING_NON_UPDATING, (0, 0.01)),
                      (td.GEOM_ROTATING, (0.37, 0.01))]:
        pts = td.rushton_generate(p, g, a, d)
        assert len(pts) > 0
        assert pts["i"].min() >= 0 and pts["i"].max() <= T - 1
        assert pts["k"].min() >= 0 and pts["k"].max() <= T - 1
        assert pts["j"].min() >= 0 and pts["j"].max() <= T - 1
        assert np.all(pts["i_frac"] >= -0.5) and np.all(pts["i_frac"] < 0.5)
        assert np.all(pts["k_frac"] >= -0.5) and np.all(pts["k_frac"] < 0.5)
        x, z = _xz(pts)
        r = np.hypot(x - p.center_x, z - p.center_z)
        assert r.max() <= p.tank_radius + 1e-9


def test_rushton_wall_is_on_the_tank_radius_and_baffles_radial():
    p = td.rushton_defaults(60)
    pts = td.rushton_generate(p, td.GEOM_FIXED)
    x, z = _xz(pts)
    r = np.hypot(x - p.center_x, z - p.center_z)
    wall = np.isclose(r, p.tank_radius, atol=1e-9)
    per_plane = math.ceil(2 * math.pi * p.tank_radius)
    assert wall.sum() == per_plane * 60
    th = np.mod(np.arctan2(z[~wall] - p.center_z, x[~wall] - p.center_x), 2 * math.pi)
    baffle_angles = np.mod(p.baffle_angle0 + np.arange(4) * math.pi / 2, 2 * math.pi)
    assert np.all(np.min(np.abs(th[:, None] - baffle_angles[None, :]), axis=1) < 1e-9)


def test_rushton_blades_rotate_rigidly_and_have_sixfold_symmetry():
    p = td.rushton_defaults(96)
    a, d = 0.4, 0.01
    b0 = td.rushton_generate(p, td.GEOM_ROTATING, a, d)
    b1 = td.rushton_generate(p, td.GEOM_ROTATING, a + math.pi / 3, d)
    x0, z0 = _xz(b0)
    x1, z1 = _xz(b1)
    # rotating by one blade pitch maps the set of points onto itself
    s0 = np.round(np.stack([x0, z0, b0["j"]], 1), 6)
    s1 = np.round(np.stack([x1, z1, b1["j"]], 1), 6)
    assert {tuple(v) for v in s0} == {tuple(v) for v in s1}
    # radii within [inner, outer], heights within the blade planes
    r = np.hypot(x0 - p.center_x, z0 - p.center_z)
    assert r.min() == pytest.approx(p.blade_inner_radius)
    assert r.max() <= p.blade_outer_radius + 1e-9
    assert b0["j"].min() == p.blade_plane_lo and b0["j"].max() == p.blade_plane_lo + p.blade_planes - 1
    # rigid rotation velocity: u = -omega (z - zc), w = omega (x - xc), v = 0
    np.testing.assert_allclose(b0["u_delta"], -d * (z0 - p.center_z), atol=1e-12)
    np.testing.assert_allclose(b0["w_delta"], d * (x0 - p.center_x), atol=1e-12)
    assert np.all(b0["v_delta"] == 0)


def test_blade_colour_classes_are_conflict_free():
    """Blades are emitted colour-major; inside a colour no two 3x3 (i,k) stencils on
    one plane share a cell, so each colour is one parallel level of the forcing
    (T = 30 and T >= 36)."""
    for T in (30, 60, 128):
        p = td.rushton_defaults(T)
        n_radial = int(math.floor(p.blade_outer_radius - p.blade_inner_radius + 1e-9)) + 1
        sizes = []
        for c in range(10):
            nb = len(range(c // 5, 6, 2))
            nr = (n_radial - c % 5 + 4) // 5 if n_radial > c % 5 else 0
            sizes.append(nb * p.blade_planes * nr)
        off = np.concatenate([[0], np.cumsum(sizes)])
        for a in np.linspace(0.0, math.pi / 3, 13):
            pts = td.rushton_generate(p, td.GEOM_ROTATING, a, 0.01)
            assert off[-1] == len(pts)
            for c in range(10):
                seg = pts[off[c]:off[c + 1]]
                seen = set()
                for q in seg:
                    cells = {(q["i"] + i1, q["j"], q["k"] + k1) for k1 in (-1, 0, 1) for i1 in (-1, 0, 1)}
                    assert not (cells & seen), (T, a, c)
                    seen |= cells


def test_step_increment_ramp():
    p = td.rushton_defaults(60, tip_speed=0.1, startup_steps=30)
    full = 0.1 / p.blade_outer_radius
    assert td.rushton_step_increment(p, 30) == pytest.approx(full)
    assert td.rushton_step_increment(p, 1000) == pytest.approx(full)
    assert td.rushton_step_increment(p, 15) == pytest.approx(0.5 * full)
    assert td.rushton_step_increment(p, 0) == 0.0
    incs = [td.rushton_step_increment(p, s) for s in range(31)]
    assert all(b >= a for a, b in zip(incs, incs[1:]))


def test_solid_mask_closed_form():
    p = td.rushton_defaults(20)
    m = td.rushton_solid_mask(p, 22, 22, 22, 1).reshape(22, 22, 22)
    i, k = np.meshgrid(np.arange(22) - 1, np.arange(22) - 1, indexing="ij")
    outside = (i - p.center_x) ** 2 + (k - p.center_z) ** 2 > p.tank_radius ** 2
    for j in (0, 5, 21):
        np.testing.assert_array_equal(m[:, j, :], outside.astype(np.uint8))
