"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same
seeded inputs, and with the reference's golden streaming vectors.

Bar: bit-exact for data movement, index lists and masks; for the float path
the reference bar is 1e-5 relative after 20 steps (|a-b| <= 1e-5 max(|b|, rho0/24)),
and the tests also record whether the result is bit-identical.
"""
import sys

import numpy as np
import pytest

import tdlbcpp_amd as td
from oracle import oracle as O
from tests.helpers import ReferenceLoop, assert_close_moments, flow_for_tank, random_moments

pytestmark = pytest.mark.gpu

# IJKL_SOA keeps the lattice LIJK in HBM and IJKL at the boundary: it must match the IJKL oracle
LAYOUTS = [(td.LAYOUT_IJKL, O.IJKL), (td.LAYOUT_LIJK, O.LIJK), (td.LAYOUT_IJKL_SOA, O.IJKL)]
LAYOUT_IDS = ["IJKL", "LIJK", "IJKL_SOA"]
STREAMS = [(td.STREAM_SIMPLE, O.SIMPLE), (td.STREAM_ESOTWIST, O.ESOTWIST)]


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def tag_pair(layout, x=3, y=3, z=3):
    ocu = O.OracleCU(x, y, z, dtype=O.TAGS, layout=layout[1], streaming=O.SIMPLE)
    q, f, nu = O.fill_for_test(ocu)
    ocu.set_q_ijkl(q)
    gcu = td.ComputeUnit(x, y, z, layout=layout[0], streaming=td.STREAM_SIMPLE, precision=td.PREC_F64)
    gcu.set(td.FIELD_Q, ocu.get(O.FIELD_Q))
    return ocu, gcu


# ---------------------------------------------------------------- data movement
@pytest.mark.parametrize("layout", LAYOUTS, ids=LAYOUT_IDS)
def test_gpu_push_matches_reference_fixture(layout, golden_dir):
    case(f"push golden {layout}")
    _, g = tag_pair(layout)
    g.streaming_push()
    got = g.q_ijkl(dtype=np.uint64)
    np.testing.assert_array_equal(got, np.load(f"{golden_dir}/streaming_push_Q.npy"))


@pytest.mark.parametrize("layout", LAYOUTS, ids=LAYOUT_IDS)
def test_gpu_pull_matches_reference_fixture(layout, golden_dir):
    case(f"pull golden {layout}")
    _, g = tag_pair(layout)
    g.streaming_pull()
    np.testing.assert_array_equal(g.q_ijkl(dtype=np.uint64), np.load(f"{golden_dir}/streaming_pull_Q.npy"))


@pytest.mark.parametrize("layout", LAYOUTS, ids=LAYOUT_IDS)
@pytest.mark.parametrize("op", ["push", "pull", "bounce_back", "bb_then_push"])
def test_gpu_data_movement_matches_oracle_nonsquare(layout, op):
    case(f"data movement {op} {layout}")
    o, g = tag_pair(layout, 4, 3, 5)
    steps = {"push": ["streaming_push"], "pull": ["streaming_pull"], "bounce_back": ["bounce_back"],
             "bb_then_push": ["bounce_back", "streaming_push", "bounce_back", "streaming_push"]}[op]
    for s in steps:
        getattr(o, s)()
        getattr(g, s)()
    np.testing.assert_array_equal(g.get(td.FIELD_Q, np.uint64), o.get(O.FIELD_Q))


# ---------------------------------------------------------------- collision / moments
def seeded_pair(layout, stream, prec, les, shape=(7, 6, 9), seed=1, even=True):
    rng = np.random.default_rng(seed)
    odt = O.F32 if prec == td.PREC_F32 else O.F64
    nu = 0.003
    o = O.OracleCU(*shape, dtype=odt, layout=layout[1], streaming=stream[1], use_les=les, nu=nu, cs0=0.12,
                   g3=0.8)
    g = td.ComputeUnit(*shape, layout=layout[0], streaming=stream[0], precision=prec, nu=nu, cs0=0.12, g3=0.8,
                       use_les=1 if les else 0)
    xg, yg, zg = o.xg, o.yg, o.zg
    m = random_moments((xg, yg, zg), rng)
    o.set_q_ijkl(m)
    F = np.zeros((xg, yg, zg, 3))
    sel = rng.random((xg, yg, zg)) < 0.2
    F[sel] = 1e-4 * rng.standard_normal((sel.sum(), 3))
    o.set(O.FIELD_F, F.reshape(-1))
    nuv = 1e-4 * rng.random((xg, yg, zg))
    o.set(O.FIELD_NU, nuv.reshape(-1))
    for f in (O.FIELD_Q, O.FIELD_F, O.FIELD_NU):
        g.set({O.FIELD_Q: td.FIELD_Q, O.FIELD_F: td.FIELD_F, O.FIELD_NU: td.FIELD_NU}[f], o.get(f))
    o.even_step = even
    g.even_step = even
    return o, g


@pytest.mark.parametrize("layout", LAYOUTS, ids=LAYOUT_IDS)
@pytest.mark.parametrize("stream", STREAMS, ids=["Simple", "Esotwist"])
@pytest.mark.parametrize("prec", [td.PREC_F32, td.PREC_F64], ids=["f32", "f64"])
@pytest.mark.parametrize("les", [False, True], ids=["noLES", "LES"])
def test_gpu_collision_and_moments_bit_exact(layout, stream, prec, les):
    for even in (True, False) if stream[0] == td.STREAM_ESOTWIST else (True,):
        case(f"collision {layout} {stream} prec={prec} les={les} even={even}")
        o, g = seeded_pair(layout, stream, prec, les, even=even)
        o.collision()
        g.collision()
        np.testing.assert_array_equal(g.get(td.FIELD_Q), o.get(O.FIELD_Q))
        np.testing.assert_array_equal(g.get(td.FIELD_NU), o.get(O.FIELD_NU))
        o.moments()
        g.moments()
        np.testing.assert_array_equal(g.get(td.FIELD_Q), o.get(O.FIELD_Q))


# ---------------------------------------------------------------- forcing
def random_points(rng, shape, n):
    pts = np.zeros(n, dtype=td.POINT_DTYPE)
    pts["i"] = rng.integers(0, shape[0], n)
    pts["j"] = rng.integers(0, shape[1], n)
    pts["k"] = rng.integers(0, shape[2], n)
    pts["i_frac"] = rng.uniform(-0.5, 0.5, n)
    pts["k_frac"] = rng.uniform(-0.5, 0.5, n)
    pts["u_delta"] = rng.uniform(-0.05, 0.05, n)
    pts["v_delta"] = rng.uniform(-0.05, 0.05, n)
    pts["w_delta"] = rng.uniform(-0.05, 0.05, n)
    return pts


@pytest.mark.parametrize("layout", LAYOUTS, ids=LAYOUT_IDS)
@pytest.mark.parametrize("stream", STREAMS, ids=["Simple", "Esotwist"])
def test_gpu_forcing_bit_exact_with_overlaps_and_wrap(layout, stream):
    case(f"forcing {layout} {stream}")
    shape = (8, 5, 7)
    o, g = seeded_pair(layout, stream, td.PREC_F32, False, shape=shape, seed=3)
    rng = np.random.default_rng(4)
    # dense overlapping list incl. points on the first/last interior rows (periodic wrap)
    pts = random_points(rng, shape, 120)
    for call in range(3):
        o.forcing(pts, 0.97, 1.9)
        g.forcing(pts, 0.97, 1.9)
        np.testing.assert_array_equal(g.get(td.FIELD_F), o.get(O.FIELD_F))
        np.testing.assert_array_equal(g.get(td.FIELD_O), o.get(O.FIELD_O))
        pts = random_points(rng, shape, 40 + 30 * call)
    g.forcing(pts[:0], 0.97, 1.9)
    o.forcing(pts[:0], 0.97, 1.9)
    np.testing.assert_array_equal(g.get(td.FIELD_F), o.get(O.FIELD_F))


# ---------------------------------------------------------------- main loop
@pytest.mark.parametrize("layout", LAYOUTS, ids=LAYOUT_IDS)
@pytest.mark.parametrize("stream", STREAMS, ids=["Simple", "Esotwist"])
def test_gpu_main_loop_20_steps_matches_oracle(layout, stream):
    T = 36
    case(f"main loop T={T} 20 steps {layout} {stream}")
    oflow, gflow = flow_for_tank(T)
    o = O.OracleCU(T, dtype=O.F32, layout=layout[1], streaming=stream[1], use_les=True, **oflow)
    g = td.ComputeUnit(T, layout=layout[0], streaming=stream[0], precision=td.PREC_F32, **gflow)
    rp = td.rushton_defaults(T)
    o.initialise(1.0)
    g.initialise(1.0)
    ref = ReferenceLoop(o, rp)
    g.bind_rushton(rp, step=1, angle=0.0)
    np.testing.assert_array_equal(g.get(td.FIELD_F), o.get(O.FIELD_F))
    ref.run(20)
    g.run(20)
    assert g.step == ref.step and g.angle == ref.angle
    assert_close_moments(g.get(td.FIELD_Q), o.get(O.FIELD_Q), what="Q")
    assert_close_moments(g.get(td.FIELD_F), o.get(O.FIELD_F), what="F", rho0=1e-3)
    assert_close_moments(g.get(td.FIELD_NU), o.get(O.FIELD_NU), what="Nu", rho0=1e-3)
    exact = np.array_equal(g.get(td.FIELD_Q), o.get(O.FIELD_Q))
    print(f"[gpu case] bit-identical Q after 20 steps: {exact}", flush=True)
    assert exact


@pytest.mark.parametrize("layout,stream", [(LAYOUTS[0], STREAMS[1]), (LAYOUTS[0], STREAMS[0]),
                                           (LAYOUTS[1], STREAMS[1])], ids=["IJKL-Esotwist", "IJKL-Simple",
                                                                           "LIJK-Esotwist"])
def test_gpu_non_cubic_grid_matches_oracle(layout, stream):
    """A tank of diameter 42 and height 29 in a 42 x 29 x 53 grid: the IJKL
    kernels' tiles are partial in i and k (and wrap in k), the last chunk of
    planes holds one plane; 6 main-loop steps bit-identical to the oracle."""
    X, Y, Z = 42, 29, 53
    case(f"non-cubic {X}x{Y}x{Z} {layout} {stream}")
    oflow, gflow = flow_for_tank(X)
    o = O.OracleCU(X, Y, Z, dtype=O.F32, layout=layout[1], streaming=stream[1], use_les=True, **oflow)
    g = td.ComputeUnit(X, Y, Z, layout=layout[0], streaming=stream[0], precision=td.PREC_F32, **gflow)
    rp = td.rushton_defaults(X, height=Y)
    o.initialise(1.0)
    g.initialise(1.0)
    ref = ReferenceLoop(o, rp)
    g.bind_rushton(rp, step=1, angle=0.0)
    ref.run(6)
    g.run(6)
    np.testing.assert_array_equal(g.get(td.FIELD_NU), o.get(O.FIELD_NU))
    np.testing.assert_array_equal(g.get(td.FIELD_Q), o.get(O.FIELD_Q))


@pytest.mark.parametrize("T", [12, 20, 31, 35])
def test_gpu_small_tanks_match_oracle(T):
    """Tanks below 36, where blades of one colour class would come within 3 sqrt(2)
    cells of each other at the inner radius, give every blade a colour class of
    its own: they bind and run the fused loop bit-identical to the oracle."""
    case(f"small tank T={T} Esotwist/LIJK and Simple/IJKL")
    for layout, stream in ((LAYOUTS[1], STREAMS[1]), (LAYOUTS[0], STREAMS[0])):
        oflow, gflow = flow_for_tank(T)
        o = O.OracleCU(T, dtype=O.F32, layout=layout[1], streaming=stream[1], use_les=True, **oflow)
        g = td.ComputeUnit(T, layout=layout[0], streaming=stream[0], precision=td.PREC_F32, **gflow)
        rp = td.rushton_defaults(T)
        o.initialise(1.0)
        g.initialise(1.0)
        ref = ReferenceLoop(o, rp)
        g.bind_rushton(rp, step=1, angle=0.0)
        ref.run(8)
        g.run(8)
        np.testing.assert_array_equal(g.rotating_points()["i"], ref.rot["i"])
        np.testing.assert_array_equal(g.get(td.FIELD_F), o.get(O.FIELD_F))
        np.testing.assert_array_equal(g.get(td.FIELD_Q), o.get(O.FIELD_Q))
        g.close()


@pytest.mark.parametrize("T,steps,team_cost", [(64, 20, None), (64, 10, "8"), (128, 6, None)],
                         ids=["64", "64-teams", "128"])
def test_gpu_one_launch_step_matches_oracle(T, steps, team_cost, monkeypatch):
    """Esotwist/LIJK tanks of 64^3 and more can take the one-launch step (the fused
    stream-collide and the IB forcing gather in one persistent kernel, XCD by XCD,
    then the forcing levels): bit-identical to the oracle, also with every plane's
    levels run by a team of workgroups (a low TDLB_FORCE_TEAM_COST)."""
    case(f"one-launch step T={T} {steps} steps Esotwist/LIJK fp32 team cost {team_cost}")
    monkeypatch.setenv("TDLB_PIPE", "1")
    if team_cost:
        monkeypatch.setenv("TDLB_FORCE_TEAM_COST", team_cost)
    oflow, gflow = flow_for_tank(T)
    o = O.OracleCU(T, dtype=O.F32, layout=O.LIJK, streaming=O.ESOTWIST, use_les=True, **oflow)
    g = td.ComputeUnit(T, layout=td.LAYOUT_LIJK, streaming=td.STREAM_ESOTWIST, precision=td.PREC_F32, **gflow)
    rp = td.rushton_defaults(T)
    o.initialise(1.0)
    g.initialise(1.0)
    ref = ReferenceLoop(o, rp)
    g.bind_rushton(rp, step=1, angle=0.0)
    assert g.one_launch
    ref.run(steps)
    g.run(steps)
    g.synchronize()  # reports a wait of the step kernel that timed out
    np.testing.assert_array_equal(g.get(td.FIELD_F), o.get(O.FIELD_F))
    np.testing.assert_array_equal(g.get(td.FIELD_NU), o.get(O.FIELD_NU))
    np.testing.assert_array_equal(g.get(td.FIELD_Q), o.get(O.FIELD_Q))


@pytest.mark.parametrize("prec", [td.PREC_F16, td.PREC_BF16, td.PREC_E3M12], ids=["fp16", "bf16", "e3m12"])
def test_gpu_one_launch_step_equals_separate_launches(prec, monkeypatch):
    """16-bit storage (no oracle): the one-launch step leaves the state of the
    separate stream-collide and forcing launches, bit for bit."""
    T, steps = 96, 12
    case(f"one-launch vs separate launches T={T} {steps} steps prec={prec}")
    _, gflow = flow_for_tank(T)
    rp = td.rushton_defaults(T)
    out = []
    for pipe in ("1", "0"):
        monkeypatch.setenv("TDLB_PIPE", pipe)
        g = td.ComputeUnit(T, layout=td.LAYOUT_LIJK, streaming=td.STREAM_ESOTWIST, precision=prec, **gflow)
        g.initialise(1.0)
        g.bind_rushton(rp, step=1, angle=0.0)
        assert g.one_launch == (pipe == "1")
        g.run(steps)
        g.synchronize()
        out.append([g.get(f) for f in (td.FIELD_Q, td.FIELD_F, td.FIELD_NU)])
        g.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


def test_gpu_disk_plane_team_at_295_matches_oracle():
    """A 295^3 tank: the impeller disk's plane runs on a team whose levels
    overflow one pass of a workgroup, so team members write F lines that
    another member has cached; the forcing at bind and after one step is
    bit-identical to the oracle (the members' L1 is dropped at the barrier)."""
    T = 295
    case(f"disk plane team T={T} bind + 1 step Esotwist/LIJK fp32")
    oflow, gflow = flow_for_tank(T)
    o = O.OracleCU(T, dtype=O.F32, layout=O.LIJK, streaming=O.ESOTWIST, use_les=True, **oflow)
    g = td.ComputeUnit(T, layout=td.LAYOUT_LIJK, streaming=td.STREAM_ESOTWIST, precision=td.PREC_F32, **gflow)
    rp = td.rushton_defaults(T)
    o.initialise(1.0)
    g.initialise(1.0)
    ref = ReferenceLoop(o, rp)
    g.bind_rushton(rp, step=1, angle=0.0)
    np.testing.assert_array_equal(g.get(td.FIELD_F), o.get(O.FIELD_F))
    ref.run(1)
    g.run(1)
    g.synchronize()
    np.testing.assert_array_equal(g.get(td.FIELD_F), o.get(O.FIELD_F))
    g.close()


@pytest.mark.parametrize("stream", STREAMS, ids=["Simple", "Esotwist"])
def test_gpu_forcing_plane_teams_match_oracle(stream, monkeypatch):
    """Planes forced by teams of workgroups meeting at a barrier in global memory
    (the impeller disk's plane of a large tank) give the sequential result: a low
    TDLB_FORCE_TEAM_COST gives every plane of a 48^3 tank a team of about four
    workgroups (about 190 in all, every one resident at once)."""
    T = 48
    case(f"forcing teams T={T} 12 steps {stream}")
    monkeypatch.setenv("TDLB_FORCE_TEAM_COST", "8")
    oflow, gflow = flow_for_tank(T)
    o = O.OracleCU(T, dtype=O.F32, layout=O.LIJK, streaming=stream[1], use_les=True, **oflow)
    g = td.ComputeUnit(T, layout=td.LAYOUT_LIJK, streaming=stream[0], precision=td.PREC_F32, **gflow)
    rp = td.rushton_defaults(T)
    o.initialise(1.0)
    g.initialise(1.0)
    ref = ReferenceLoop(o, rp)
    g.bind_rushton(rp, step=1, angle=0.0)
    monkeypatch.delenv("TDLB_FORCE_TEAM_COST")
    ref.run(12)
    g.run(12)
    g.synchronize()  # reports a team barrier that timed out
    np.testing.assert_array_equal(g.get(td.FIELD_F), o.get(O.FIELD_F))
    np.testing.assert_array_equal(g.get(td.FIELD_Q), o.get(O.FIELD_Q))


@pytest.mark.parametrize("stream", STREAMS, ids=["Simple", "Esotwist"])
def test_gpu_eager_calls_equal_fused_run(stream):
    """The five reference calls per step and the fused fast path leave identical state."""
    T = 30
    case(f"eager vs fused T={T} {stream}")
    _, gflow = flow_for_tank(T)
    rp = td.rushton_defaults(T)
    a = td.ComputeUnit(T, layout=td.LAYOUT_LIJK, streaming=stream[0], **gflow)
    b = td.ComputeUnit(T, layout=td.LAYOUT_LIJK, streaming=stream[0], **gflow)
    a.initialise(1.0)
    b.initialise(1.0)
    a.bind_rushton(rp)
    fixed = td.rushton_generate(rp, td.GEOM_FIXED)
    d0 = float(np.float32(td.rushton_step_increment(rp, 1)))
    nonupd = td.rushton_generate(rp, td.GEOM_ROTATING_NON_UPDATING, 0.0, d0)
    b.forcing(np.concatenate([fixed, nonupd, td.rushton_generate(rp, td.GEOM_ROTATING, 0.0, d0)]), 0.97, 1.9)
    step, angle = 1, 0.0
    for _ in range(6):
        step += 1
        d = float(np.float32(td.rushton_step_increment(rp, step)))
        angle += d
        b.collision()
        b.bounce_back()
        b.streaming_push()
        b.moments()
        b.forcing(np.concatenate([fixed, nonupd, td.rushton_generate(rp, td.GEOM_ROTATING, angle, d)]), 0.97, 1.9)
    a.run(6)
    for f in (td.FIELD_Q, td.FIELD_F, td.FIELD_NU, td.FIELD_O):
        np.testing.assert_array_equal(a.get(f), b.get(f))
    assert a.even_step == b.even_step


def test_gpu_rotating_points_bit_exact_with_host_generator():
    T = 60
    case(f"blade list T={T}")
    _, gflow = flow_for_tank(T)
    rp = td.rushton_defaults(T)
    g = td.ComputeUnit(T, layout=td.LAYOUT_LIJK, streaming=td.STREAM_ESOTWIST, **gflow)
    g.initialise(1.0)
    g.bind_rushton(rp)
    for n in (1, 7, 31):
        g.run(n)
        d = float(np.float32(td.rushton_step_increment(rp, g.step)))
        host = td.rushton_generate(rp, td.GEOM_ROTATING, g.angle, d)
        dev = g.rotating_points()
        for name in ("i", "j", "k"):
            np.testing.assert_array_equal(dev[name], host[name])
        for name in ("i_frac", "k_frac", "u_delta", "v_delta", "w_delta"):
            np.testing.assert_array_equal(dev[name], host[name].astype(np.float32).astype(np.float64))


def test_gpu_solid_mask_bit_exact():
    T = 48
    case(f"solid mask T={T}")
    rp = td.rushton_defaults(T)
    g = td.ComputeUnit(T, layout=td.LAYOUT_LIJK, streaming=td.STREAM_ESOTWIST)
    g.initialise(1.0)
    g.bind_rushton(rp)
    np.testing.assert_array_equal(g.solid_mask(), td.rushton_solid_mask(rp, g.xg, g.yg, g.zg, 1))
