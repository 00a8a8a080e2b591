"""Regrid (doubleResolutionFullCU), output exclusion, XZ plane records and the
vorticity field, each against the oracle's restatement of the reference."""
import sys

import numpy as np
import pytest

import tdlbcpp_amd as td
from oracle import oracle as O
from tests.helpers import ReferenceLoop, assert_close_moments, flow_for_tank

pytestmark = pytest.mark.gpu

CASES = [((td.LAYOUT_IJKL, O.IJKL), (td.STREAM_SIMPLE, O.SIMPLE)),
         ((td.LAYOUT_LIJK, O.LIJK), (td.STREAM_ESOTWIST, O.ESOTWIST)),
         ((td.LAYOUT_IJKL_SOA, O.IJKL), (td.STREAM_ESOTWIST, O.ESOTWIST))]
IDS = ["IJKL-Simple", "LIJK-Esotwist", "IJKL_SOA-Esotwist"]


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def pair(T, layout, stream):
    oflow, gflow = flow_for_tank(T)
    o = O.OracleCU(T, dtype=O.F32, layout=layout[1], streaming=stream[1], use_les=True, **oflow)
    g = td.ComputeUnit(T, layout=layout[0], streaming=stream[0], precision=td.PREC_F32, **gflow)
    o.initialise(1.0)
    g.initialise(1.0)
    return o, g


@pytest.mark.parametrize("layout,stream", CASES, ids=IDS)
def test_double_resolution_then_run_matches_oracle(layout, stream):
    case(f"regrid {layout} {stream}")
    o, g = pair(36, layout, stream)
    rp = td.rushton_defaults(36)
    ref = ReferenceLoop(o, rp)
    g.bind_rushton(rp)
    ref.run(5)
    g.run(5)
    o.double_resolution()
    g.double_resolution()
    assert (g.xg, g.yg, g.zg) == (o.xg, o.yg, o.zg) == (74, 74, 74)
    for f in (td.FIELD_Q, td.FIELD_F, td.FIELD_NU, td.FIELD_O):
        np.testing.assert_array_equal(g.get(f), o.get(f))
    # main.cpp:360-399: the tank of the new size at the same step and angle
    rp2 = td.rushton_defaults(72)
    ref2 = ReferenceLoop(o, rp2, step=ref.step, angle=ref.angle)
    g.bind_rushton(rp2, step=ref.step, angle=ref.angle)
    ref2.run(3)
    g.run(3)
    assert_close_moments(g.get(td.FIELD_Q), o.get(O.FIELD_Q), what="Q")
    assert np.array_equal(g.get(td.FIELD_Q), o.get(O.FIELD_Q))


@pytest.mark.parametrize("layout,stream", CASES, ids=IDS)
def test_plane_and_vorticity_match_oracle(layout, stream):
    case(f"plane/vorticity {layout} {stream}")
    T = 36
    o, g = pair(T, layout, stream)
    rp = td.rushton_defaults(T)
    ref = ReferenceLoop(o, rp)
    g.bind_rushton(rp)
    ref.run(12)
    g.run(12)
    ex = g.output_exclude()
    # exclusion: outside the tank plus every cell holding a point of the step's lists
    want = td.rushton_solid_mask(rp, g.xg, g.yg, g.zg).copy()
    d = td.rushton_step_increment(rp, ref.step)
    lists = [td.rushton_generate(rp, td.GEOM_FIXED), ref.nonupd, g.rotating_points()]
    for pts in lists:
        want.reshape(g.xg, g.yg, g.zg)[pts["i"] + 1, pts["j"] + 1, pts["k"] + 1] = 1
    np.testing.assert_array_equal(ex, want)
    del d
    for cut in (5, T // 3 + 1, T - 2):
        gq, gf = g.plane_xz(cut, ex)
        oq, of = o.plane_xz(cut, ex)
        for a, b in ((gq, oq), (gf, of)):  # fields (the alignment bytes are unset in the reference)
            assert len(a) == len(b)
            for name in ("i", "j", "k", "q"):
                np.testing.assert_array_equal(a[name], b[name])
        for axis in (0, 1):
            gv, ov = g.vorticity(axis, cut, ex), o.vorticity(axis, cut, ex)
            np.testing.assert_array_equal(np.isnan(gv), np.isnan(ov))
            fin = np.isfinite(ov)
            # the device log and the host's may differ in the last place
            np.testing.assert_allclose(gv[fin], ov[fin], rtol=2e-6)
            np.testing.assert_array_equal(np.isinf(gv), np.isinf(ov))
