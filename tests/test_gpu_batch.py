"""Batches of tanks: a member advances bit for bit as the same tank run alone.

The batch has no counterpart in the reference; its contract is that each
member's Q, F, Nu and O after n steps equal those of tdlb_cu_run on an
identical handle, whatever the other members, budgets, tiles or graph use.
"""
import sys

import numpy as np
import pytest

import tdlbcpp_amd as td

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def make_tank(T, prec, streaming, layout, tip=0.1, re=7300.0, startup=20):
    flow = td.flow_for_tank(T, re=re, uav=tip, precision=td.PREC_F64)
    cu = td.ComputeUnit(T, layout=layout, streaming=streaming, precision=prec, **flow)
    cu.initialise(1.0)
    cu.bind_rushton(td.rushton_defaults(T, tip_speed=tip, startup_steps=startup))
    return cu


def fields(cu):
    return [cu.get(f) for f in (td.FIELD_Q, td.FIELD_F, td.FIELD_NU, td.FIELD_O)]


SPECS = [(36, 0.10), (48, 0.07), (40, 0.12), (36, 0.05)]


@pytest.mark.parametrize("cfg", [
    (td.PREC_F16, td.STREAM_ESOTWIST, td.LAYOUT_LIJK, True, True),
    (td.PREC_F32, td.STREAM_SIMPLE, td.LAYOUT_IJKL, False, True),
    (td.PREC_BF16, td.STREAM_SIMPLE, td.LAYOUT_LIJK, True, False),
    (td.PREC_F32, td.STREAM_ESOTWIST, td.LAYOUT_IJKL, True, True),
    (td.PREC_F16, td.STREAM_SIMPLE, td.LAYOUT_IJKL_SOA, True, True),
    (td.PREC_E3M12, td.STREAM_ESOTWIST, td.LAYOUT_LIJK, True, True),
], ids=["fp16-Eso-LIJK-graph", "fp32-Simple-IJKL-nograph", "bf16-Simple-LIJK-notune", "fp32-Eso-IJKL-graph",
        "fp16-Simple-IJKL_SOA-graph", "e3m12-Eso-LIJK-graph"])
def test_batch_members_equal_solo_runs(cfg):
    prec, st, ly, graph, tune = cfg
    case(f"batch vs solo {cfg}")
    solo = [make_tank(T, prec, st, ly, tip) for T, tip in SPECS]
    mem = [make_tank(T, prec, st, ly, tip) for T, tip in SPECS]
    budgets = [25, 12, 25, 5]
    b = td.Batch(prec, ly, st, use_graph=graph, autotune=tune)
    for cu, n in zip(mem, budgets):
        b.add(cu, n)
    assert b.run(10) == 10
    assert b.run(30) == 15
    info = b.info()
    assert info["members"] == 4 and all(t in (0, 1, 2) for t in info["tiles"])
    assert info["block_size"] in (256, 512)
    for cu, n in zip(solo, budgets):
        cu.run(n)
    for a, s, n in zip(mem, solo, budgets):
        assert (a.step, a.angle) == (s.step, s.angle)
        assert b.member(a)[2] == 0
        for x, y in zip(fields(a), fields(s)):
            np.testing.assert_array_equal(x, y)
    b.close()


def test_batch_resumes_after_member_is_read():
    """Reading a member materialises its moments; the next batch run settles it
    with one step on its own and continues identically to a solo run."""
    case("batch resume after download")
    prec, st, ly = td.PREC_F16, td.STREAM_ESOTWIST, td.LAYOUT_LIJK
    a, s = make_tank(40, prec, st, ly), make_tank(40, prec, st, ly)
    other = make_tank(36, prec, st, ly, tip=0.09)
    b = td.Batch(prec, ly, st)
    b.add(a, 20)
    b.add(other, 20)
    b.run(7)
    s.run(7)
    np.testing.assert_array_equal(a.get(td.FIELD_Q), s.get(td.FIELD_Q))
    b.run(13)
    s.run(13)
    for x, y in zip(fields(a), fields(s)):
        np.testing.assert_array_equal(x, y)
    red_a, red_s = a.reductions(), s.reductions()
    assert red_a == red_s
    b.close()


def test_batch_rejects_mismatched_members():
    case("batch rejects mismatched members")
    b = td.Batch(td.PREC_F16, td.LAYOUT_LIJK, td.STREAM_ESOTWIST)
    wrong = make_tank(36, td.PREC_F32, td.STREAM_ESOTWIST, td.LAYOUT_LIJK)
    with pytest.raises(td.TdlbError):
        b.add(wrong, 5)
    unbound = td.ComputeUnit(36, layout=td.LAYOUT_LIJK, streaming=td.STREAM_ESOTWIST, precision=td.PREC_F16,
                             **td.flow_for_tank(36))
    with pytest.raises(td.TdlbError):
        b.add(unbound, 5)
    with pytest.raises(td.TdlbError):
        td.Batch(td.PREC_F64, td.LAYOUT_LIJK, td.STREAM_ESOTWIST)
    b.close()


@pytest.mark.parametrize("prec", [td.PREC_F32, td.PREC_F16, td.PREC_E3M12], ids=["fp32", "fp16", "e3m12"])
def test_cache_hinted_solo_run_equals_unhinted_batch_member(prec):
    """At 64^3 every k-row of the LIJK planes is a whole number of 128-byte lines,
    so tdlb_cu_run launches the fused kernel built with the non-temporal cache
    hint; batch members never take it.  The hint only changes cache policy: the
    two runs of one tank must agree bit for bit."""
    T, steps = 64, 12
    case(f"cache-hinted solo vs batch member T={T} prec={prec}")
    solo = make_tank(T, prec, td.STREAM_ESOTWIST, td.LAYOUT_LIJK)
    mem = make_tank(T, prec, td.STREAM_ESOTWIST, td.LAYOUT_LIJK)
    b = td.Batch(prec, td.LAYOUT_LIJK, td.STREAM_ESOTWIST)
    b.add(mem, steps)
    assert b.run(steps) == steps
    solo.run(steps)
    assert (mem.step, mem.angle) == (solo.step, solo.angle)
    for x, y in zip(fields(mem), fields(solo)):
        np.testing.assert_array_equal(x, y)
    b.close()


def test_batch_of_64_uneven_fp16_tanks_equals_solo_runs():
    """The batched-tank workload at its shape: 64 FP16 Esotwist/LIJK tanks with
    edges drawn from {48, 64, 96, 128, 160}, tip speeds and Reynolds numbers as
    the bench draws them (numpy PCG64, seed 1), graph replay and autotuning on,
    10 steps: every member equals its solo run bit for bit."""
    import bench
    specs = bench.batch_tanks(64, steps=10, seed=1)
    case(f"batch of {len(specs)} fp16 tanks, edges {sorted(set(s['edge'] for s in specs))}")
    prec, st, ly = td.PREC_F16, td.STREAM_ESOTWIST, td.LAYOUT_LIJK
    mem = [make_tank(s["edge"], prec, st, ly, tip=s["tip"], re=s["re"], startup=100) for s in specs]
    b = td.Batch(prec, ly, st, use_graph=True, autotune=True)
    for cu in mem:
        b.add(cu, 10)
    assert b.run(10) == 10
    info = b.info()
    print(f"[gpu case] block size {info['block_size']}, tiles {info['tiles']}", flush=True)
    b.close()
    for cu, s in zip(mem, specs):
        solo = make_tank(s["edge"], prec, st, ly, tip=s["tip"], re=s["re"], startup=100)
        solo.run(10)
        assert (cu.step, cu.angle) == (solo.step, solo.angle)
        for x, y in zip(fields(cu), fields(solo)):
            np.testing.assert_array_equal(x, y)
        solo.close()
        cu.close()
