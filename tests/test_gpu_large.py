"""A tank past 2^32 / 18 cells, the size where an IJKL slot address (cell * 18 +
slot) no longer fits 32 bits and the LIJK slot planes lie more than 4 GiB apart
in total.  The IJKL and LIJK kernels address the same state differently; they
agree bit for bit at small sizes (test_gpu_parity.py), so they must agree here
too, through the deterministic FP64 sums of the whole state, and Esotwist must
hold the mass.
"""
import sys

import pytest

import tdlbcpp_amd as td


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
def test_large_tank_ijkl_and_lijk_agree():
    T, steps = 650, 3  # 652^3 = 277 M cells incl. ghosts > 2^32 / 18 = 238.6 M
    assert (T + 2) ** 3 * 18 > 2 ** 32
    flow = td.flow_for_tank(T, re=1.0e5, uav=0.1, cs0=0.1, use_les=1)
    sums = []
    for layout, name in ((td.LAYOUT_LIJK, "LIJK"), (td.LAYOUT_IJKL, "IJKL")):
        case(f"large tank {T}^3 Esotwist {name} fp32, {steps} steps")
        cu = td.ComputeUnit(T, layout=layout, streaming=td.STREAM_ESOTWIST, precision=td.PREC_F32, **flow)
        try:
            cu.initialise(1.0)
            cu.bind_rushton(td.rushton_defaults(T, tip_speed=0.1, startup_steps=10))
            m0 = cu.reductions()["mass"]
            cu.run(steps)
            r = cu.reductions()
        finally:
            cu.close()
        print(f"[gpu case] {name}: mass {r['mass']:.6f} (start {m0:.6f}) momentum {r['momentum']} "
              f"torque {r['torque']:.6e}", flush=True)
        assert abs(r["mass"] / m0 - 1.0) < 1e-6
        assert r["torque"] != 0.0
        sums.append(r)
    assert sums[0] == sums[1]


@pytest.mark.gpu
def test_access_pattern_probe():
    """tdlb_probe_planes: a pass over 19 planes of 128^3 values takes a positive
    time below what 1 TB/s would allow, the non-temporal schedule too, and bad
    arguments are refused before anything is allocated."""
    print("[gpu case] access-pattern probe 19 x 128^3", flush=True)
    cells = 128 ** 3
    for nt in (True, False):
        ms = td.probe_planes(4, cells, 19, nontemporal=nt, passes=5, stride_bytes=td.lijk_plane_bytes(128, 4))
        assert 0.0 < ms < 2 * 19 * cells * 4 / 1e12 * 1e3
    for args in ((3, cells, 0, 19), (4, 0, 0, 19), (4, cells, 0, 21), (4, cells, 4 * cells - 128, 19),
                 (4, cells, 4 * cells + 4, 19)):
        elem, n, stride, planes = args
        ms = td.ctypes.c_double()
        assert td.lib().tdlb_probe_planes(0, elem, n, stride, planes, 1, 3, td.ctypes.byref(ms)) == -1
