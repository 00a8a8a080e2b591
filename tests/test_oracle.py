"""Pins of the CPU oracle against the reference's own golden vectors and against
what the mathematics fixes (closed forms, invariants, limits).

Nothing here runs on a GPU.  Each test names the oracle function it pins.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

E = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1],
              [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1], [0, -1, 1]])


def tag_cu(layout, streaming=O.SIMPLE, x=3, y=3, z=3):
    cu = O.OracleCU(x, y, z, ghost=1, dtype=O.TAGS, layout=layout, streaming=streaming)
    q, f, nu = O.fill_for_test(cu)
    cu.set_q_ijkl(q)
    cu.set(O.FIELD_F, f.reshape(-1))
    return cu, q


# ---------------------------------------------------------------- layouts (a1)
@pytest.mark.parametrize("layout", [O.IJKL, O.LIJK])
def test_memory_layout_addresses(layout):
    """ComputeUnitMemoryTest.cpp:36-58: IJKL slot = index*18+l, LIJK slot = l*size+index."""
    cu, q = tag_cu(layout)
    raw = cu.get(O.FIELD_Q)
    xg, yg, zg = cu.xg, cu.yg, cu.zg
    size = xg * yg * zg
    for (i, j, k, l) in [(0, 0, 0, 0), (1, 2, 3, 7), (4, 4, 4, 17), (2, 0, 1, 9)]:
        index = i * yg * zg + j * zg + k
        addr = index * 18 + l if layout == O.IJKL else l * size + index
        assert raw[addr] == q[i, j, k, l]


# ---------------------------------------------------------------- streaming (a5, a6)
@pytest.mark.parametrize("layout", [O.IJKL, O.LIJK])
def test_push_matches_reference_fixture(layout, golden_dir):
    """StreamingSimpleTest.cpp:92-126 with tests/StreamingSimplePushTest.hpp."""
    cu, _ = tag_cu(layout)
    cu.streaming_push()
    expected = np.load(f"{golden_dir}/streaming_push_Q.npy")
    np.testing.assert_array_equal(cu.q_ijkl(), expected)
    np.testing.assert_array_equal(cu.get(O.FIELD_F).reshape(5, 5, 5, 3),
                                  np.load(f"{golden_dir}/streaming_push_F.npy"))


@pytest.mark.parametrize("layout", [O.IJKL, O.LIJK])
def test_pull_matches_reference_fixture(layout, golden_dir):
    """StreamingSimpleTest.cpp:56-90 with tests/StreamingSimpleTest.hpp."""
    cu, _ = tag_cu(layout)
    cu.streaming_pull()
    np.testing.assert_array_equal(cu.q_ijkl(), np.load(f"{golden_dir}/streaming_pull_Q.npy"))


def _interior(shape, g=1):
    xg, yg, zg = shape
    return lambda i, j, k: (g <= i < xg - g) and (g <= j < yg - g) and (g <= k < zg - g)


def test_push_closed_form_nonsquare():
    """Push == two-lattice push whose destination starts as a copy of the source:
    dst[c][d] = src[c-e_d][d] if c-e_d is interior else src[c][d], on every cell."""
    cu, q = tag_cu(O.LIJK, x=4, y=3, z=5)
    cu.streaming_push()
    got = cu.q_ijkl()
    inside = _interior(q.shape[:3])
    exp = q.copy()
    for i in range(q.shape[0]):
        for j in range(q.shape[1]):
            for k in range(q.shape[2]):
                for d in range(18):
                    s = (i - E[d, 0], j - E[d, 1], k - E[d, 2])
                    if inside(*s):
                        exp[i, j, k, d] = q[s + (d,)]
    np.testing.assert_array_equal(got, exp)


# ---------------------------------------------------------------- bounce-back (a8)
def test_bounce_back_face_and_edge_rules():
    """Boundary.hpp:136-250: ghost slot d of a face cell g = slot opp(d) of g+e_d;
    only inward slots of face/edge ghosts change; corners untouched; interior untouched."""
    cu, q = tag_cu(O.IJKL, x=3, y=4, z=5)
    cu.bounce_back()
    got = cu.q_ijkl()
    xg, yg, zg = q.shape[:3]
    inside = _interior((xg, yg, zg))
    changed = np.argwhere(got != q)
    assert len(changed) > 0
    for i, j, k, d in changed:
        assert not inside(i, j, k)
        # inward: moving along e_d from the ghost reaches a cell whose ghost-ness decreases
        s = (i + E[d, 0], j + E[d, 1], k + E[d, 2])
        assert 0 <= s[0] < xg and 0 <= s[1] < yg and 0 <= s[2] < zg
    # every face cell's inward slots: value is the opposite slot of the neighbour (taken
    # before the later faces overwrite it, which only matters on face rims)
    for j in range(1, yg - 1):
        for k in range(1, zg - 1):
            for d in (0, 6, 12, 8, 14):  # Q01, Q07, Q13, Q09, Q15 on the i=0 face
                s = (1, j + E[d, 1], k + E[d, 2])
                assert got[0, j, k, d] == q[s + (d ^ 1,)]
    # corners are untouched
    for c in [(0, 0, 0), (xg - 1, yg - 1, zg - 1), (0, yg - 1, 0)]:
        np.testing.assert_array_equal(got[c], q[c])
    # the 12 edge lines carry exactly one changed slot each
    assert got[0, 0, 2, 6] == q[1, 1, 2, 7]          # Q07 <- Q08 of (1,1,k)
    assert got[1, yg - 1, zg - 1, 11] == q[1, yg - 2, zg - 2, 10]   # Q12 <- Q11


# ---------------------------------------------------------------- esoteric twist (a7)
@pytest.mark.parametrize("layout", [O.IJKL, O.LIJK])
@pytest.mark.parametrize("even", [True, False])
def test_esotwist_read_home_cells(layout, even):
    """StreamingEsoteric.hpp:29-81: direction d of cell c lives at c + max(-e_d, 0)
    (periodic on the interior), in slot d at even parity and opp(d) at odd."""
    cu, q = tag_cu(layout, O.ESOTWIST, x=3, y=4, z=5)
    cu.even_step = even
    r = cu.read_cells()
    nx, ny, nz = r.shape[:3]
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for d in range(18):
                    hi = (i + max(-E[d, 0], 0)) % nx
                    hj = (j + max(-E[d, 1], 0)) % ny
                    hk = (k + max(-E[d, 2], 0)) % nz
                    slot = d if even else d ^ 1
                    assert r[i, j, k, d] == q[hi + 1, hj + 1, hk + 1, slot]


@pytest.mark.parametrize("layout", [O.IJKL, O.LIJK])
@pytest.mark.parametrize("even", [True, False])
def test_esotwist_write_then_toggle_is_periodic_pull(layout, even):
    """write(P) then reading at the toggled parity streams: read'(c)[d] = v(c-e_d)[d]
    with periodic wrap; writeMoments(P) is the inverse of read(P)."""
    cu, q = tag_cu(layout, O.ESOTWIST, x=3, y=4, z=5)
    cu.even_step = even
    vals = cu.read_cells()
    # writeMoments is the inverse of read at the same parity
    cu.write_cells(vals, as_moments=True)
    np.testing.assert_array_equal(cu.q_ijkl(), q)
    # a distinct value per (cell, direction), written as post-collision populations
    v = (np.arange(vals.size, dtype=np.uint64) + 7).reshape(vals.shape)
    cu.write_cells(v, as_moments=False)
    cu.streaming_push()  # toggles parity
    assert cu.even_step == (not even)
    r = cu.read_cells()
    nx, ny, nz = r.shape[:3]
    for d in range(18):
        exp = np.roll(v[..., d], shift=(E[d, 0], E[d, 1], E[d, 2]), axis=(0, 1, 2))
        np.testing.assert_array_equal(r[..., d], exp)


# ---------------------------------------------------------------- moments (a3)
def test_moment_rows_match_lattice_vectors():
    """M01 = sum f, M02..M04 = sum f e, M06/M08/M09 = sum f e_a e_b (exact, integer data)."""
    rng = np.random.default_rng(1)
    for _ in range(20):
        f = rng.integers(-1000, 1000, 18).astype(np.float64)
        m = O.moments_of(f)
        assert m[0] == f.sum()
        np.testing.assert_array_equal(m[1:4], E.T @ f)
        assert m[5] == (E[:, 0] * E[:, 1]) @ f
        assert m[7] == (E[:, 0] * E[:, 2]) @ f
        assert m[8] == (E[:, 1] * E[:, 2]) @ f


def _moment_matrix():
    return np.stack([O.moments_of(np.eye(18)[d]) for d in range(18)], axis=1)


def test_moment_matrix_full_rank_and_symmetric_under_parity():
    """The 18 moments are independent, and odd-order rows (M02-M04, M11-M16) flip sign
    under e -> -e while even-order rows are unchanged."""
    M = _moment_matrix()
    assert np.linalg.matrix_rank(M) == 18
    flip = np.array([d ^ 1 for d in range(18)])
    odd = [1, 2, 3, 10, 11, 12, 13, 14, 15]
    for r in range(18):
        sign = -1 if r in odd else 1
        np.testing.assert_array_equal(M[r, flip], sign * M[r])


# ---------------------------------------------------------------- collision (a4)
def one_cell(dtype=O.F64, use_les=False, **flow):
    return O.OracleCU(1, 1, 1, ghost=1, dtype=dtype, layout=O.LIJK, streaming=O.SIMPLE,
                      use_les=use_les, **flow)


def collide(cu, m, f=(0.0, 0.0, 0.0)):
    cu.write_cells(np.asarray(m).reshape(1, 1, 1, 18))
    F = np.zeros((3, 3, 3, 3), dtype=cu.np_dtype)
    F[1, 1, 1] = f
    cu.set(O.FIELD_F, F.reshape(-1))
    cu.collision()
    return cu.read_cells().reshape(18)


def test_collision_rest_state_populations():
    """Rest moments (rho,0..0) collide to rho/12 on faces and rho/24 on edges (exact)."""
    for dt in (O.F32, O.F64):
        cu = one_cell(dt, nu=0.01)
        rho = 1.0
        p = collide(cu, [rho] + [0.0] * 17)
        npdt = np.float32 if dt == O.F32 else np.float64
        assert np.all(p[:6] == npdt(rho) / npdt(12)) or np.allclose(p[:6], rho / 12, rtol=1e-7)
        assert np.allclose(p[6:], rho / 24, rtol=1e-7)


def test_collision_relaxes_moments_as_eggels_somers():
    """moments(collision(m)) is the relaxed moment vector, which holds only if the
    inverse transform E satisfies M.E = 24 I:
      rho kept; j -> j + F; Pi_ab -> b (2 (j_a + F_a/2) u_b - c Pi_ab) with
      u = (j + F/2)/rho, b = 1/(1+6nu), c = 1-6nu; 3rd order -> -g3 m; 4th order -> 0."""
    rng = np.random.default_rng(3)
    nu, g3 = 0.02, 0.8
    pairs = {4: (0, 0), 5: (0, 1), 6: (1, 1), 7: (0, 2), 8: (1, 2), 9: (2, 2)}
    for _ in range(10):
        m = rng.normal(0, 0.01, 18)
        m[0] = 1.0 + rng.normal(0, 0.01)
        f = rng.normal(0, 1e-3, 3)
        cu = one_cell(O.F64, nu=nu, g3=g3)
        p = collide(cu, m, f)
        a = O.moments_of(p)
        rho, j = m[0], m[1:4]
        u = (j + 0.5 * f) / rho
        b, c = 1 / (1 + 6 * nu), 1 - 6 * nu
        exp = np.zeros(18)
        exp[0] = rho
        exp[1:4] = j + f
        for s, (x, y) in pairs.items():
            exp[s] = b * (2 * (j[x] + 0.5 * f[x]) * u[y] - c * m[s])
        exp[10:16] = -g3 * m[10:16]
        np.testing.assert_allclose(a, exp, rtol=1e-12, atol=1e-15)


def test_les_with_zero_cs0_is_plain_collision():
    """With cs0 = 0 the LES branch gives Nu = 0 and exactly the non-LES populations."""
    rng = np.random.default_rng(5)
    m = rng.normal(0, 0.01, 18)
    m[0] = 1.0
    for dt in (O.F32, O.F64):
        p0 = collide(one_cell(dt, nu=0.003), m)
        cu = one_cell(dt, use_les=True, nu=0.003, cs0=0.0)
        p1 = collide(cu, m)
        np.testing.assert_array_equal(p0, p1)
        assert np.all(cu.get(O.FIELD_NU) == 0)


def test_les_eddy_viscosity_scales_with_cs0_squared():
    """Nu = cs0^2 sqrt|S^2| with S independent of cs0 on the first step (Nu_old = 0)."""
    rng = np.random.default_rng(7)
    m = rng.normal(0, 0.01, 18)
    m[0] = 1.0
    nus = []
    for cs0 in (0.1, 0.2):
        cu = one_cell(O.F64, use_les=True, nu=0.003, cs0=cs0)
        collide(cu, m)
        nus.append(cu.get(O.FIELD_NU).reshape(3, 3, 3)[1, 1, 1])
    assert nus[0] > 0
    assert nus[1] / nus[0] == pytest.approx(4.0, rel=1e-12)


def test_shear_wave_decays_at_lattice_viscosity():
    """Physics pin of collision + streaming: a periodic shear wave u_x(y) decays as
    exp(-nu k^2 t) with nu the input viscosity (tau = 3 nu + 1/2)."""
    ny, nu, U = 32, 0.05, 1e-5
    cu = O.OracleCU(2, ny, 2, dtype=O.F64, layout=O.LIJK, streaming=O.ESOTWIST, nu=nu, g3=0.0)
    y = np.arange(ny)
    k = 2 * np.pi / ny
    m = np.zeros((2, ny, 2, 18))
    m[..., 0] = 1.0
    m[..., 1] = (U * np.sin(k * y))[None, :, None]
    cu.write_cells(m)

    def amp():
        jx = cu.read_cells()[..., 1].mean(axis=(0, 2))
        return 2 * np.dot(jx, np.sin(k * y)) / ny

    def step(n):
        for _ in range(n):
            cu.collision()
            cu.streaming_push()
            cu.moments()

    step(20)
    a0 = amp()
    step(100)
    a1 = amp()
    assert a1 / a0 == pytest.approx(math.exp(-nu * k * k * 100), rel=5e-3)


# ---------------------------------------------------------------- forcing (a9)
def test_weight_partition_of_unity_and_values():
    """calcWeight (Forcing.hpp:18-35) is the 3-point Roma kernel: w(0)=2/3, w(1)=1/6,
    w(1.5)=0, symmetric, and sum_i w(i - s) = 1 for |s| <= 1/2."""
    assert O.weight(0.0, O.F64) == pytest.approx(2 / 3, abs=1e-15)
    assert O.weight(1.0, O.F64) == pytest.approx(1 / 6, abs=1e-15)
    assert O.weight(1.5, O.F64) == pytest.approx(0.0, abs=1e-15)
    assert O.weight(2.0, O.F64) == 0.0
    for s in np.linspace(-0.5, 0.5, 41):
        tot = sum(O.weight(i - s, O.F64) for i in (-1, 0, 1))
        assert tot == pytest.approx(1.0, abs=1e-14)
        assert O.weight(s, O.F64) == O.weight(-s, O.F64)
    # continuity at |x| = 1/2
    assert O.weight(0.5, O.F64) == pytest.approx(O.weight(0.5 + 1e-12, O.F64), abs=1e-9)


def pts(*rows):
    a = np.zeros(len(rows), dtype=O.POINT_DTYPE)
    for n, r in enumerate(rows):
        a[n] = r
    return a


def test_forcing_single_point_closed_form():
    """Rest fluid, one point with wall velocity uD: Sum = -rho*uD*sum(w) = -rho*uD, so
    F(c) = beta * w(c) * rho * uD on the 3x3 stencil (i,k) and 0 elsewhere."""
    cu = O.OracleCU(6, 4, 6, dtype=O.F64, layout=O.IJKL, streaming=O.SIMPLE)
    cu.initialise(1.0)
    beta, uD, wD = 1.9, 0.03, -0.02
    cu.forcing(pts((2, 1, 3, 0.2, -0.4, uD, 0.0, wD)), 0.97, beta)
    F = cu.get(O.FIELD_F).reshape(cu.xg, cu.yg, cu.zg, 3)
    exp = np.zeros_like(F)
    for i1 in (-1, 0, 1):
        for k1 in (-1, 0, 1):
            w = O.weight(i1 - 0.2, O.F64) * O.weight(k1 + 0.4, O.F64)
            exp[3 + i1, 2, 4 + k1] = [beta * w * uD, 0.0, beta * w * wD]
    np.testing.assert_allclose(F, exp, rtol=1e-13, atol=1e-18)
    assert np.all(cu.get(O.FIELD_O) == 0)


def test_forcing_is_sequential_over_the_list():
    """Two overlapping points in one call == the first alone followed by the second
    alone, on the second stencil (later points see F already updated by earlier ones);
    the order of the list matters."""
    p1 = (2, 1, 2, 0.1, 0.0, 0.02, 0.0, 0.0)
    p2 = (3, 1, 2, -0.3, 0.2, 0.0, 0.0, 0.01)

    def run(*calls):
        cu = O.OracleCU(6, 3, 6, dtype=O.F64, layout=O.LIJK, streaming=O.SIMPLE)
        cu.initialise(1.0)
        for c in calls:
            cu.forcing(pts(*c), 0.9, 1.5)
        return cu.get(O.FIELD_F).reshape(cu.xg, cu.yg, cu.zg, 3)

    both = run([p1, p2])
    seq = run([p1], [p2])
    rev = run([p2, p1])
    st2 = (slice(3, 6), 2, slice(2, 5))
    np.testing.assert_array_equal(both[st2], seq[st2])
    assert not np.allclose(both, rev)


def test_forcing_wraps_stencil_periodically_in_i_and_k():
    """A point on the first interior cell spreads onto the last interior cell (i and k)."""
    cu = O.OracleCU(5, 3, 5, dtype=O.F64, layout=O.LIJK, streaming=O.SIMPLE)
    cu.initialise(1.0)
    cu.forcing(pts((0, 1, 0, 0.0, 0.0, 0.05, 0.0, 0.0)), 0.97, 1.9)
    F = cu.get(O.FIELD_F).reshape(cu.xg, cu.yg, cu.zg, 3)
    assert F[5, 2, 5, 0] != 0 and F[1, 2, 1, 0] != 0 and F[0, 2, 0, 0] == 0
    np.testing.assert_allclose(F[..., 0].sum(), 1.9 * 0.05, rtol=1e-13)


# ---------------------------------------------------------------- flow params (a11)
def test_calc_nu_closed_form():
    """FlowParams::calcNuAndRe_m: Re_m = reM*pi/2, nu = uav*R/Re_m (R an int)."""
    nu, rem = O.calc_nu(7300.0, 0.1, 10, O.F64)
    assert rem == pytest.approx(7300 * math.pi / 2, rel=1e-15)
    assert nu == pytest.approx(0.1 * 10 / (7300 * math.pi / 2), rel=1e-15)
    nu32, _ = O.calc_nu(7300.0, 0.1, 10, O.F32)
    assert nu32 == pytest.approx(8.72e-5, rel=1e-3)


# ---------------------------------------------------------------- whole step (a12)
def test_rest_state_is_a_fixed_point_over_steps():
    for streaming in (O.SIMPLE, O.ESOTWIST):
        cu = O.OracleCU(6, 5, 4, dtype=O.F32, layout=O.LIJK, streaming=streaming, nu=1e-3)
        cu.initialise(1.0)
        q0 = cu.read_cells()
        for _ in range(5):
            cu.collision()
            cu.bounce_back()
            cu.streaming_push()
            cu.moments()
        # rho/12 and rho/24 are inexact in float: a few ulps of drift over 5 steps
        np.testing.assert_allclose(cu.read_cells(), q0, atol=1e-6)


def test_esotwist_conserves_mass_without_forcing():
    rng = np.random.default_rng(11)
    cu = O.OracleCU(6, 5, 7, dtype=O.F64, layout=O.IJKL, streaming=O.ESOTWIST, nu=0.01)
    m = np.zeros((6, 5, 7, 18))
    m[..., 0] = 1 + 0.01 * rng.standard_normal((6, 5, 7))
    m[..., 1:4] = 0.001 * rng.standard_normal((6, 5, 7, 3))
    cu.write_cells(m)
    mass0 = m[..., 0].sum()
    mom0 = m[..., 1:4].sum(axis=(0, 1, 2))
    for _ in range(10):
        cu.collision()
        cu.streaming_push()
        cu.moments()
    r = cu.read_cells()
    assert r[..., 0].sum() == pytest.approx(mass0, rel=1e-13)
    np.testing.assert_allclose(r[..., 1:4].sum(axis=(0, 1, 2)), mom0, atol=1e-12)


# ---------------------------------------------------------------- regrid, output planes, vorticity
@pytest.mark.parametrize("layout", [O.IJKL, O.LIJK], ids=["IJKL", "LIJK"])
def test_double_resolution_replicates_each_cell_into_a_2x2x2_block(layout):
    """doubleResolutionFullCU (ComputeUnit.hpp:229-316): the doubled grid is the
    old one upsampled by replication -- interior cell a of the old grid fills
    cells 2a-1 and 2a on every axis, the ghost layers stay ghost layers."""
    o = O.OracleCU(3, 4, 2, dtype=O.TAGS, layout=layout, streaming=O.SIMPLE)
    q, f, nu = O.fill_for_test(o)
    o.set_q_ijkl(q)
    o.set(O.FIELD_NU, nu)
    o.set(O.FIELD_F, f)
    old_q = o.q_ijkl().copy()
    old_nu = o.get(O.FIELD_NU).reshape(o.xg, o.yg, o.zg).copy()
    o.double_resolution()
    assert (o.xg, o.yg, o.zg) == (8, 10, 6)
    new_q = o.q_ijkl()
    new_nu = o.get(O.FIELD_NU).reshape(o.xg, o.yg, o.zg)

    def src(a, n_old_interior):  # new index -> old index along one axis
        if a == 0:
            return 0
        if a == 2 * n_old_interior + 1:
            return n_old_interior + 1
        return (a + 1) // 2

    for i in range(o.xg):
        for j in range(o.yg):
            for k in range(o.zg):
                s = (src(i, 3), src(j, 4), src(k, 2))
                np.testing.assert_array_equal(new_q[i, j, k], old_q[s])
                assert new_nu[i, j, k] == old_nu[s]
    assert not o.get(O.FIELD_F).any() and not o.get(O.FIELD_O).any()


def test_plane_xz_records_hold_raw_slots_and_nonzero_forces():
    """savePlaneXZ (ComputeUnitOutput.hpp:354-447): one record per plane cell not
    excluded, the first 4 raw slots of the cell; an F record where all three
    components of F are non-zero (Force::isNotZero, QVec.hpp:54-56)."""
    o = O.OracleCU(4, 3, 5, dtype=O.F32, layout=O.LIJK, streaming=O.ESOTWIST)
    rng = np.random.default_rng(3)
    q4 = rng.standard_normal((o.xg, o.yg, o.zg, 18)).astype(np.float32)
    o.set_q_ijkl(q4)
    F = np.zeros((o.xg, o.yg, o.zg, 3), np.float32)
    F[2, 2, 3] = (1.0, -2.0, 0.5)
    F[1, 2, 1] = (0.0, 0.0, 3.0)
    o.set(O.FIELD_F, F.ravel())
    ex = np.zeros((o.xg, o.yg, o.zg), np.uint8)
    ex[3, 2, 2] = 1
    qr, fr = o.plane_xz(2, ex.ravel(), i0=10, k0=20)
    want = [(i, k) for i in range(1, 5) for k in range(1, 6) if not ex[i, 2, k]]
    assert len(qr) == len(want) == 19
    for r, (i, k) in zip(qr, want):
        assert (r["i"], r["j"], r["k"]) == (10 + i - 1, 1, 20 + k - 1)
        np.testing.assert_array_equal(r["q"], q4[i, 2, k, :4])
    assert [(r["i"], r["j"], r["k"]) for r in fr] == [(11, 2, 22)]  # (0, 0, 3) is "zero"
    np.testing.assert_array_equal(fr[0]["q"], F[2, 2, 3])


def test_vorticity_of_a_shear_flow_is_log_of_shear_squared():
    """calcVorticityXZ (ComputeUnitOutput.hpp:113-160) on u_x = a k: the only
    velocity gradient is du_x/dk = a, so the field is log(a^2) in the interior."""
    o = O.OracleCU(5, 5, 7, dtype=O.F64, layout=O.IJKL, streaming=O.SIMPLE)
    a = 0.01
    m = np.zeros(o.interior_shape())
    m[..., 0] = 1.0
    m[..., 1] = a * np.arange(1, 8)[None, None, :]
    o.write_cells(m, as_moments=True)
    v = o.vorticity(0, 3)
    # cells next to the ghost layer read zero-density ghosts (0/0): interior of the plane only
    np.testing.assert_allclose(v[1:-1, 1:-1], np.log(a * a), rtol=1e-12)
    ex = np.zeros(o.size, np.uint8)
    ex[o.q_ijkl().shape[1] * o.zg * 2 + o.zg * 3 + 4] = 1  # cell (2, 3, 4)
    assert np.isnan(o.vorticity(0, 3, ex)[1, 3])
    # plane k = 4: same shear, same value away from the ghost layers
    np.testing.assert_allclose(o.vorticity(1, 4)[1:-1, 1:-1], np.log(a * a), rtol=1e-12)


@pytest.mark.parametrize("streaming", [O.SIMPLE, O.ESOTWIST], ids=["Simple", "Esotwist"])
def test_threaded_oracle_is_identical(streaming):
    """The threaded CPU baseline (OpenMP collision and moments, cells split over
    i-planes) gives the single-threaded oracle's fields bit for bit."""
    import tdlbcpp_amd as td
    from tests.helpers import ReferenceLoop, flow_for_tank
    T = 36
    out = []
    try:
        for threads in (1, 4):
            assert O.set_threads(threads) == threads
            oflow, _ = flow_for_tank(T, re=7300.0)
            o = O.OracleCU(T, dtype=O.F32, layout=O.IJKL, streaming=streaming, use_les=True, **oflow)
            o.initialise(1.0)
            loop = ReferenceLoop(o, td.rushton_defaults(T))
            loop.run(6)
            out.append([o.get(f).copy() for f in (O.FIELD_Q, O.FIELD_F, O.FIELD_NU)])
    finally:
        O.set_threads(1)
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
    assert np.abs(out[0][1]).max() > 0  # the forcing ran
