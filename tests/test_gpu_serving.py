"""Serving loop: requests admitted by HBM footprint, continuous batching, and
each served tank's result identical to the same tank run alone."""
import sys

import numpy as np
import pytest

import tdlbcpp_amd as td

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def solo(T, tip, re, steps, prec, streaming, layout, startup=100):
    flow = td.flow_for_tank(T, re=re, uav=tip, cs0=0.1, use_les=1, precision=prec)
    cu = td.ComputeUnit(T, layout=layout, streaming=streaming, precision=prec, **flow)
    cu.initialise(1.0)
    cu.bind_rushton(td.rushton_defaults(T, tip_speed=tip, startup_steps=startup))
    cu.run(steps)
    return cu


def drain(s, limit=10000):
    out = []
    for _ in range(limit):
        resident, queued, _ = s.tick()
        out += s.poll()
        if resident == 0 and queued == 0:
            return out
    raise AssertionError("server did not drain")


@pytest.mark.parametrize("cfg", [(td.PREC_F16, td.STREAM_ESOTWIST, td.LAYOUT_LIJK),
                                 (td.PREC_F32, td.STREAM_SIMPLE, td.LAYOUT_IJKL)], ids=["fp16-Eso", "fp32-Simple"])
def test_served_tanks_equal_solo_runs(cfg):
    prec, st, ly = cfg
    case(f"serving vs solo {cfg}")
    reqs = [(36, 0.10, 5000.0, 30), (40, 0.07, 12000.0, 20), (48, 0.12, 3000.0, 9), (20, 0.05, 20000.0, 16)]
    s = td.Server(prec, ly, st, tick_steps=7, keep_handles=True)
    tickets = [s.submit(*reqs[0])]
    results = []
    for _ in range(2):
        s.tick()
        results += s.poll()
    tickets += [s.submit(*r) for r in reqs[1:]]  # arrive while the first one runs
    results += drain(s)
    assert sorted(r["ticket"] for r in results) == sorted(tickets)
    by_ticket = {r["ticket"]: r for r in results}
    for t, (T, tip, re, steps) in zip(tickets, reqs):
        r = by_ticket[t]
        assert r["error"] == 0 and r["steps"] == steps and r["edge"] == T
        assert r["submit_s"] <= r["admit_s"] <= r["start_s"] <= r["finish_s"]
        assert np.isfinite(r["power_number"]) and abs(r["mass_drift"]) < 1e-2
        a = s.take(t, T)
        b = solo(T, tip, re, steps, prec, st, ly)
        for f in (td.FIELD_Q, td.FIELD_F, td.FIELD_NU, td.FIELD_O):
            np.testing.assert_array_equal(a.get(f), b.get(f))
        red = b.reductions()
        assert red["mass"] == r["mass"] and red["torque"] == r["torque"]
        a.close()
        b.close()
    s.close()


def test_admission_respects_the_hbm_budget():
    case("serving admission by footprint")
    # room for one 48^3 fp16 tank (about 10 MB) but not two
    s = td.Server(td.PREC_F16, td.LAYOUT_LIJK, td.STREAM_ESOTWIST, hbm_budget=12e6, tick_steps=5)
    for _ in range(3):
        s.submit(48, 0.1, 5000.0, 12)
    most = 0
    results = []
    for _ in range(100):
        resident, queued, _ = s.tick()
        most = max(most, resident)
        results += s.poll()
        if resident == 0 and queued == 0:
            break
    assert len(results) == 3 and all(r["error"] == 0 for r in results)
    assert most == 1
    assert s.stats()["reserved_bytes"] == 0
    with pytest.raises(td.TdlbError):
        s.submit(160, 0.1, 5000.0, 10)  # larger than the whole budget
    s.close()


def test_served_burst_of_uneven_tanks_up_to_320():
    """The served workload at its shape: a burst of FP16 Esotwist/LIJK requests
    with edges log-uniform in [64, 320] as the bench draws them, plus one 320^3
    tank, admitted by footprint against the HBM budget and run with continuous
    batching; sampled results (the 320^3 tank among them) equal solo runs bit
    for bit, and every request completes."""
    import bench
    prec, st, ly = td.PREC_F16, td.STREAM_ESOTWIST, td.LAYOUT_LIJK
    draws = bench.serve_requests(4.0, rate=4.0, seed=1)[:12]
    reqs = [(r["edge"], r["tip"], r["re"], 8 + 3 * k) for k, r in enumerate(draws)] + [(320, 0.1, 1.0e4, 12)]
    case(f"served burst of {len(reqs)} tanks, edges {[r[0] for r in reqs]}")
    s = td.Server(prec, ly, st, tick_steps=5, keep_handles=True)
    tickets = [s.submit(*r) for r in reqs[:6]]
    results = []
    for _ in range(2):
        s.tick()
        results += s.poll()
    tickets += [s.submit(*r) for r in reqs[6:]]  # arrive while the first ones run
    results += drain(s)
    assert sorted(r["ticket"] for r in results) == sorted(tickets)
    assert all(r["error"] == 0 for r in results)
    by_ticket = {r["ticket"]: r for r in results}
    sample = [0, len(reqs) // 2, len(reqs) - 1]
    for k in range(len(reqs)):
        a = s.take(tickets[k], reqs[k][0])
        if k in sample:
            T, tip, re, steps = reqs[k]
            assert by_ticket[tickets[k]]["steps"] == steps
            b = solo(T, tip, re, steps, prec, st, ly)
            for f in (td.FIELD_Q, td.FIELD_F, td.FIELD_NU, td.FIELD_O):
                np.testing.assert_array_equal(a.get(f), b.get(f))
            b.close()
        a.close()
    s.close()
