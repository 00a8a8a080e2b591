"""Macroscopic fields, FP64 device sums and the 16-bit population storage.

rho/u follow QVec::velocity(F) (QVec.hpp:125-131) and must equal the oracle's
moments pushed through the same formula bit for bit, in both handle phases.
The FP64 sums are checked against numpy sums of the same data.  The 16-bit
storage has no reference counterpart: it is checked against the fp64 build of
the same kernels (velocity L2 over the fluid, power number, mass drift).
"""
import sys

import numpy as np
import pytest

import tdlbcpp_amd as td
from oracle import oracle as O
from tdlbcpp_amd import precision as P
from tests.helpers import ReferenceLoop, flow_for_tank

pytestmark = pytest.mark.gpu

LAYOUTS = [(td.LAYOUT_IJKL, O.IJKL), (td.LAYOUT_LIJK, O.LIJK)]
STREAMS = [(td.STREAM_SIMPLE, O.SIMPLE), (td.STREAM_ESOTWIST, O.ESOTWIST)]


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def host_macroscopic(o):
    """{rho, ux, uy, uz} from the oracle's moments and F, QVec::velocity(F) in float."""
    m = o.read_cells().astype(np.float32)
    F = o.get(O.FIELD_F).reshape(o.xg, o.yg, o.zg, 3)[1:-1, 1:-1, 1:-1].astype(np.float32)
    rho = m[..., 0].astype(np.float64)
    out = np.empty(m.shape[:3] + (4,), dtype=np.float32)
    out[..., 0] = m[..., 0]
    for a in range(3):
        out[..., 1 + a] = ((1.0 / rho) * (m[..., 1 + a].astype(np.float64) + 0.5 * F[..., a].astype(np.float64))
                           ).astype(np.float32)
    return out, m, F


@pytest.mark.parametrize("stream", STREAMS, ids=["Simple", "Esotwist"])
@pytest.mark.parametrize("layout", LAYOUTS, ids=["IJKL", "LIJK"])
def test_macroscopic_and_sums_match_oracle(layout, stream):
    T = 36
    case(f"macroscopic T={T} {layout} {stream}")
    oflow, gflow = flow_for_tank(T)
    o = O.OracleCU(T, dtype=O.F32, layout=layout[1], streaming=stream[1], use_les=True, **oflow)
    g = td.ComputeUnit(T, layout=layout[0], streaming=stream[0], precision=td.PREC_F32, **gflow)
    rp = td.rushton_defaults(T)
    o.initialise(1.0)
    g.initialise(1.0)
    ref = ReferenceLoop(o, rp)
    g.bind_rushton(rp)
    ref.run(9)
    g.run(9)
    want, m, F = host_macroscopic(o)
    lazy = g.macroscopic()        # computed from the fused path's populations
    red_lazy = g.reductions()
    g.get(td.FIELD_Q)             # materialise
    explicit = g.macroscopic()    # computed from the stored moments
    red = g.reductions()
    np.testing.assert_array_equal(lazy, want)
    np.testing.assert_array_equal(explicit, want)
    assert red == red_lazy
    m64, F64 = m.astype(np.float64), F.astype(np.float64)
    np.testing.assert_allclose(red["mass"], m64[..., 0].sum(), rtol=1e-13)
    for a in range(3):
        np.testing.assert_allclose(red["momentum"][a], m64[..., 1 + a].sum(), rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(red["force"][a], F64[..., a].sum(), rtol=1e-9, atol=1e-14)
    i = np.arange(T)[:, None] - rp.center_x
    k = np.arange(T)[None, :] - rp.center_z
    inside = (i * i + k * k) <= (rp.blade_outer_radius + 2.0) ** 2
    tq = (i[:, None, :] * F64[..., 2] - k[:, None, :] * F64[..., 0])[np.broadcast_to(inside[:, None, :], F64.shape[:3])]
    np.testing.assert_allclose(red["torque"], tq.sum(), rtol=1e-9, atol=1e-14)
    assert red["torque"] != 0.0


PREC16 = [td.PREC_F16, td.PREC_BF16, td.PREC_E3M12]
PREC16_IDS = ["fp16", "bf16", "e3m12"]


@pytest.mark.parametrize("prec", PREC16, ids=PREC16_IDS)
def test_16bit_storage_rejects_per_call_operations(prec):
    case(f"16-bit per-call rejection {prec}")
    _, gflow = flow_for_tank(36)
    g = td.ComputeUnit(36, layout=td.LAYOUT_LIJK, streaming=td.STREAM_ESOTWIST, precision=prec, **gflow)
    g.initialise(1.0)
    for call in (g.collision, g.moments, g.streaming_push, g.streaming_pull, g.bounce_back):
        with pytest.raises(td.TdlbError):
            call()
    q = g.get(td.FIELD_Q)
    assert q.dtype == np.float32 and q.size == g.size * 18


@pytest.mark.parametrize("stream", STREAMS, ids=["Simple", "Esotwist"])
@pytest.mark.parametrize("prec", PREC16, ids=PREC16_IDS)
def test_16bit_storage_tracks_fp64(prec, stream):
    """A short tank run in 16-bit storage stays close to fp64 and conserves mass."""
    T, steps = 48, 300
    case(f"16-bit storage T={T} {steps} steps prec={prec} {stream}")
    kw = dict(streaming=stream[0], startup_steps=50, samples=5)
    ref = P.tank_run(T, td.PREC_F64, steps, **kw)
    r = P.tank_run(T, prec, steps, **kw)
    f32 = P.tank_run(T, td.PREC_F32, steps, **kw)
    l2 = P.velocity_l2(r["mac"], ref["mac"], ref["fluid"])
    l2_32 = P.velocity_l2(f32["mac"], ref["mac"], ref["fluid"])
    # Simple streaming is not mass-conserving at the walls (a wall cell keeps its
    # own value in the directions nobody pushes into), so compare with fp64's drift
    drift = r["mass"][-1] / r["mass0"] - 1.0
    drift64 = ref["mass"][-1] / ref["mass0"] - 1.0
    po_err = abs(r["po"] - ref["po"]) / abs(ref["po"])
    print(f"[gpu case] velocity L2 vs fp64: {l2:.3e} (fp32 {l2_32:.3e}); Po {r['po']:.5f} vs {ref['po']:.5f} "
          f"(rel {po_err:.2e}); mass drift {drift:.3e} (fp64 {drift64:.3e})", flush=True)
    assert np.isfinite(r["mac"]).all()
    # Simple streaming is unstable on this tank (fp32 already differs from fp64
    # by ~7e-4 after 300 steps, Esotwist by ~2e-5), so its L2 measures how much
    # each format seeds the instability: E3M12's linear step below 2^-9 (5e-7)
    # is coarser than fp16's at the tiny deviations of the fluid near rest.
    esotwist = stream[0] == td.STREAM_ESOTWIST
    l2_max, po_max, drift_max = {td.PREC_F16: (2e-3, 1e-2, 1e-4), td.PREC_BF16: (5e-2, 5e-2, 5e-3),
                                 td.PREC_E3M12: (1e-3 if esotwist else 1e-2, 5e-3 if esotwist else 1e-2,
                                                 1e-4)}[prec]
    assert l2 <= l2_max
    assert po_err <= po_max
    assert abs(drift - drift64) <= drift_max
    if esotwist:
        # the deviation-form collision conserves the mass of the 16-bit formats
        assert abs(drift) <= 1e-6


# Measured at the target's workload (96^3, 5000 steps from rest, Esotwist, LIJK,
# LES, tip speed 0.1), the runs being deterministic for a given build: velocity
# L2 at the last step / of the mean field over the last 1250 steps, against fp64,
# Re 1e5: fp32 7.1e-5 / 3.9e-5, fp16 5.6e-3 / 3.0e-3, bf16 3.7e-2 / 1.9e-2,
# e3m12 2.7e-3 / 8.4e-4.  Builds that change the rounding of the 16-bit path
# move the instantaneous values by up to 2-3x (the turbulent impeller flow
# amplifies any storage rounding by step 5000); the mean field moves less.
# The bar is a velocity L2 <= 1e-3.  E3M12 meets it on the mean field; no
# 16-bit format meets it on the instantaneous field, and fp16 and bf16 miss it
# on both, so those are held to bands around their measurements only, as
# regression guards.
TARGET_LIMITS = {  # (instantaneous L2, mean-field L2, |Po rel. error|)
    "fp32": (2e-4, 1e-4, 1e-4),
    "fp16": (2e-2, 6e-3, 1e-3),
    "bf16": (1.5e-1, 6e-2, 1e-2),
    "e3m12": (6e-3, 1e-3, 5e-4),
}


@pytest.mark.parametrize("re_", [1.0e5, 7300.0], ids=["Re1e5", "Re7300"])
def test_storage_accuracy_at_the_target_workload(re_):
    """96^3 x 5000 steps of every storage format against the fp64 build: E3M12
    meets the 1e-3 velocity-L2 bar on the mean field; every format stays within
    its measured band; the 16-bit formats conserve mass."""
    case(f"accuracy 96^3 x 5000 Re={re_:g}")
    r = P.precision_report(T=96, steps=5000, precisions=(td.PREC_F32, td.PREC_F16, td.PREC_BF16, td.PREC_E3M12),
                           re=re_, samples=20)
    for name, v in r["variants"].items():
        print(f"[gpu case] {name}: L2 {v['velocity_l2']:.3e} mean field {v['velocity_l2_mean']:.3e} "
              f"Po err {v['po_rel_err']:.2e} mass drift {v['mass_drift']:.2e} ({v['seconds']:.1f} s)", flush=True)
    for name, (l2_max, mean_max, po_max) in TARGET_LIMITS.items():
        v = r["variants"][name]
        assert v["velocity_l2"] <= l2_max, name
        assert v["velocity_l2_mean"] <= mean_max, name
        assert v["po_rel_err"] <= po_max, name
        assert abs(v["mass_drift"]) <= 1e-6, name
    assert r["variants"]["e3m12"]["velocity_l2_mean"] <= 1e-3  # the bar


def test_fp64_rest_tank_conserves_mass_exactly():
    """The fp64 build with the impeller at rest keeps the fluid at rest."""
    T = 36
    case("fp64 rest tank")
    _, gflow = flow_for_tank(T, dtype=O.F64)
    g = td.ComputeUnit(T, layout=td.LAYOUT_LIJK, streaming=td.STREAM_ESOTWIST, precision=td.PREC_F64, **gflow)
    g.initialise(1.0)
    rp = td.rushton_defaults(T, tip_speed=0.0)
    g.bind_rushton(rp)
    g.run(20)
    mac = g.macroscopic()
    assert np.abs(mac[..., 1:]).max() < 1e-14
    red = g.reductions()
    assert abs(red["mass"] - T ** 3) < 1e-9
