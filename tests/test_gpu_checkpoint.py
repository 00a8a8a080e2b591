"""Checkpoint files and the C++ driver on the GPU.

The reference's own checkpoint test (ComputeUnitCheckpointTest.cpp:46-76) is a
write-then-read round trip of every field; this file does that through the C
ABI, checks the files hold the raw arrays in the reference's layout, and runs
the driver end to end on configuration 1 (60^3 Rushton tank, Simple streaming,
IJKL, FP32, LES) against the oracle.
"""
import glob
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import tdlbcpp_amd as td
from oracle import oracle as O
from tests.helpers import ReferenceLoop, assert_close_moments

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("streaming", [td.STREAM_SIMPLE, td.STREAM_ESOTWIST], ids=["Simple", "Esotwist"])
@pytest.mark.parametrize("layout", [td.LAYOUT_IJKL, td.LAYOUT_LIJK, td.LAYOUT_IJKL_SOA],
                         ids=["IJKL", "LIJK", "IJKL_SOA"])
def test_checkpoint_write_read_round_trip(tmp_path, streaming, layout):
    case(f"checkpoint round trip {streaming} {layout}")
    T = 36
    flow = td.flow_for_tank(T)
    a = td.ComputeUnit(T, layout=layout, streaming=streaming, **flow)
    a.initialise(1.0)
    a.bind_rushton(td.rushton_defaults(T))
    a.run(7)  # lazy state: the write must materialise the reference's moments
    rc = td.lib().tdlb_cu_checkpoint_write(a._h, str(tmp_path).encode(), b"device", b'{"x": 1}')
    assert rc == 0, td.lib().tdlb_last_error()
    Q, F, Nu = a.get(td.FIELD_Q), a.get(td.FIELD_F), a.get(td.FIELD_NU)
    # raw files hold the arrays exactly (names DiskOutputTree.h:560-590)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "checkpoint_grid.0.0.0.device.Q", dtype=np.float32), Q)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "checkpoint_grid.0.0.0.device.F", dtype=np.float32), F)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "checkpoint_grid.0.0.0.device.Nu", dtype=np.float32), Nu)
    assert json.loads((tmp_path / "AllParams.0.0.0.device.json").read_text()) == {"x": 1}
    b = td.ComputeUnit(T, layout=layout, streaming=streaming, **flow)
    missing = td.ctypes.c_int(-1)
    assert td.lib().tdlb_cu_checkpoint_read(b._h, str(tmp_path).encode(), b"device", td.ctypes.byref(missing)) == 0
    assert missing.value == 0
    b.even_step = a.even_step
    for f in (td.FIELD_Q, td.FIELD_F, td.FIELD_NU):
        np.testing.assert_array_equal(b.get(f), a.get(f))
    # both continue identically from the restored state
    for cu in (a, b):
        cu.collision()
        cu.streaming_push()
        cu.moments()
    np.testing.assert_array_equal(b.get(td.FIELD_Q), a.get(td.FIELD_Q))


@pytest.mark.parametrize("prec", [td.PREC_F16, td.PREC_BF16, td.PREC_E3M12], ids=["fp16", "bf16", "e3m12"])
def test_16bit_checkpoint_round_trip(tmp_path, prec):
    """A 16-bit handle writes the reference's float files (moments decoded from the
    16-bit populations, Nu widened), reads them back exactly, and its moments
    agree with an fp32 tank run alongside to the 16-bit rounding."""
    case(f"16-bit checkpoint round trip prec={prec}")
    T, steps = 36, 7
    flow = td.flow_for_tank(T)
    rp = td.rushton_defaults(T)
    ref = td.ComputeUnit(T, layout=td.LAYOUT_LIJK, streaming=td.STREAM_ESOTWIST, precision=td.PREC_F32, **flow)
    a = td.ComputeUnit(T, layout=td.LAYOUT_LIJK, streaming=td.STREAM_ESOTWIST, precision=prec, **flow)
    for cu in (ref, a):
        cu.initialise(1.0)
        cu.bind_rushton(rp)
        cu.run(steps)
    rc = td.lib().tdlb_cu_checkpoint_write(a._h, str(tmp_path).encode(), b"device", b'{"x": 1}')
    assert rc == 0, td.lib().tdlb_last_error()
    Q, F, Nu = a.get(td.FIELD_Q), a.get(td.FIELD_F), a.get(td.FIELD_NU)
    assert Q.dtype == np.float32 and Nu.dtype == np.float32
    np.testing.assert_array_equal(np.fromfile(tmp_path / "checkpoint_grid.0.0.0.device.Q", dtype=np.float32), Q)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "checkpoint_grid.0.0.0.device.Nu", dtype=np.float32), Nu)
    b = td.ComputeUnit(T, layout=td.LAYOUT_LIJK, streaming=td.STREAM_ESOTWIST, precision=prec, **flow)
    missing = td.ctypes.c_int(-1)
    assert td.lib().tdlb_cu_checkpoint_read(b._h, str(tmp_path).encode(), b"device", td.ctypes.byref(missing)) == 0
    assert missing.value == 0
    for f, want in ((td.FIELD_Q, Q), (td.FIELD_F, F), (td.FIELD_NU, Nu)):
        np.testing.assert_array_equal(b.get(f), want)
    # moments (rho, j, and the rest) of the 16-bit run against fp32: the 16-bit
    # rounding of the deviations (relative step 2^-11 fp16, 2^-8 bf16, 2^-12 e3m12)
    q16 = Q.reshape(18, -1)
    q32 = ref.get(td.FIELD_Q).reshape(18, -1)
    tol = {td.PREC_F16: 2e-5, td.PREC_BF16: 2e-4, td.PREC_E3M12: 1e-5}[prec]
    np.testing.assert_allclose(q16, q32, rtol=0, atol=tol)
    assert np.abs(q32[1:4]).max() > 1e-4  # the impeller has set the fluid moving


def test_checkpoint_read_reports_missing_and_wrong_size(tmp_path):
    case("checkpoint missing / wrong size")
    a = td.ComputeUnit(8, layout=td.LAYOUT_LIJK, streaming=td.STREAM_SIMPLE)
    a.initialise(1.0)
    missing = td.ctypes.c_int(0)
    assert td.lib().tdlb_cu_checkpoint_read(a._h, str(tmp_path).encode(), b"device", td.ctypes.byref(missing)) == 0
    assert missing.value == 7
    (tmp_path / "checkpoint_grid.0.0.0.device.Q").write_bytes(b"\0" * 100)
    assert td.lib().tdlb_cu_checkpoint_read(a._h, str(tmp_path).encode(), b"device", None) == -5


def _config1_input(tmp_path):
    with open(os.path.join(ROOT, "tests", "golden", "input_debug_gridx60_numSteps20.json")) as fh:
        j = json.load(fh)
    j["FlowParams"].update(initialRho=1.0, cs0=0.1, useLES=True, streaming="Simple", uav=0.1,
                           reMNonDimensional=7300.0, alpha=0.97, beta=1.9, g3=0.8)
    j["RunningParams"].update(num_steps=20, step=1, angle=0.0, doubleResolutionAtStep=1000)
    j["CheckpointParams"].update(checkpointRepeat=10, checkpointWriteRootDir=str(tmp_path),
                                 checkpointWriteDirAppendTime=False, checkpointWriteDirPrefix="cfg1")
    j["OutputParams"]["outputRootDir"] = str(tmp_path / "out")
    p = tmp_path / "input.json"
    p.write_text(json.dumps(j))
    return p, j


def _picture(v):
    """The 8-bit picture of a vorticity plane as the driver forms it, in float:
    floor(255 (v - min) / (max - min)) over the finite values, 0 elsewhere."""
    v = np.asarray(v, dtype=np.float32)
    fin = np.isfinite(v)
    lo, hi = v[fin].min(), v[fin].max()
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.floor(np.float32(255) * ((v - lo) / (hi - lo)))
    return np.where(fin & np.isfinite(s), np.clip(s, 0, 255), 0).astype(np.int32)


def _check_outputs(tmp_path, o, ref, T):
    """The step-20 outputs of the 60^3 input (main.cpp:462-487 and the XZ_planes
    records of --write-planes) against the oracle's records and vorticity."""
    from PIL import Image
    outs = glob.glob(str(tmp_path / "out_*"))
    assert len(outs) == 1, outs
    od = outs[0]
    names = sorted(os.listdir(od))
    for step in ("00000010", "00000020"):
        assert f"plot.xyPlane.V5.step_{step}.cut_30.jpeg" in names
        assert f"plot.xzPlane.V5.step_{step}.cut_40.jpeg" in names
        assert f"plot.xzPlane.V5.step_{step}.cut_40" in names
    rp = td.rushton_defaults(T, tip_speed=0.1, startup_steps=30)
    xg = T + 2
    ex = td.rushton_solid_mask(rp, xg, xg, xg).copy().reshape(xg, xg, xg)
    for pts in (ref.fixed, ref.nonupd, ref.rot):
        ex[pts["i"] + 1, pts["j"] + 1, pts["k"] + 1] = 1
    ex = ex.reshape(-1)
    pdir = os.path.join(od, "plot.xzPlane.V5.step_00000020.cut_40")
    oq, of = o.plane_xz(40, ex)
    for fname, want, qlen in (("CUid.0.0.0.V5.QVec.bin", oq, 4), ("CUid.0.0.0.V5.F3.bin", of, 3)):
        got = np.fromfile(os.path.join(pdir, fname), dtype=want.dtype)
        assert len(got) == len(want) > 0
        for name in ("i", "j", "k", "q"):
            np.testing.assert_array_equal(got[name], want[name])
        meta = json.loads(open(os.path.join(pdir, fname + ".json")).read())
        assert meta["BinFileParams"]["binFileSizeInStructs"] == len(want)
        assert meta["BinFileParams"]["QOutputLength"] == qlen
        assert meta["BinFileParams"]["filePath"] == os.path.join(pdir, fname)
        assert meta["RunningParams"]["step"] == 20
    for axis, cut, tag in ((0, 40, "xzPlane"), (1, 30, "xyPlane")):
        want = _picture(o.vorticity(axis, cut, ex))
        if axis == 1:
            want = want.T  # pict[(i-1) + x (j-1)]: rows are j
        with Image.open(os.path.join(od, f"plot.{tag}.V5.step_00000020.cut_{cut}.jpeg")) as im:
            assert im.format == "JPEG" and im.mode == "L" and im.size == (T, T)
            got = np.asarray(im, dtype=np.int32)
        # JPEG at quality 100 (at most 1 level) plus a last-place difference of
        # the device's log that can move a floor by one level
        err = np.abs(got - want)
        assert err.max() <= 3 and err.mean() < 0.5, (err.max(), err.mean())


@pytest.mark.parametrize("mode", ["fused", "per-call"])
def test_driver_config1_matches_oracle(tmp_path, mode):
    """tdlb_run on configuration 1: the checkpoint written at running.step 20
    (after 19 main-loop iterations) equals the oracle's Q, F and Nu to the bar,
    and the outputs of step 20 (vorticity pictures, XZ plane records) match the
    oracle's."""
    case(f"driver config 1 {mode}")
    inp, j = _config1_input(tmp_path)
    args = [os.path.join(ROOT, "tdlbcpp_amd", "tdlb_run"), "-i", str(inp), "--write-planes"]
    if mode == "per-call":
        args.append("--per-call")
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    d = tmp_path / "cfg1_checkpoint" / "cfg1_checkpoint_step_20"
    allp = json.loads((d / "AllParams.0.0.0.device.json").read_text())
    assert allp["RunningParams"]["step"] == 20
    assert allp["ComputeUnitParams"]["x"] == 60 and allp["BinFileParams"]["structName"] == "checkpoint"
    nu = allp["FlowParams"]["nu"]
    T = 60
    o = O.OracleCU(T, dtype=O.F32, layout=O.IJKL, streaming=O.SIMPLE, use_les=True, initialRho=1.0, cs0=0.1,
                   g3=0.8, nu=nu, alpha=0.97, beta=1.9)
    o.initialise(1.0)
    ref = ReferenceLoop(o, td.rushton_defaults(T, tip_speed=0.1, startup_steps=30))
    ref.run(19)
    q = np.fromfile(d / "checkpoint_grid.0.0.0.device.Q", dtype=np.float32)
    f = np.fromfile(d / "checkpoint_grid.0.0.0.device.F", dtype=np.float32)
    n = np.fromfile(d / "checkpoint_grid.0.0.0.device.Nu", dtype=np.float32)
    assert_close_moments(q, o.get(O.FIELD_Q), what="Q")
    assert_close_moments(f, o.get(O.FIELD_F), what="F", rho0=1e-3)
    assert_close_moments(n, o.get(O.FIELD_NU), what="Nu", rho0=1e-3)
    print(f"[gpu case] config 1 bit-identical Q: {np.array_equal(q, o.get(O.FIELD_Q))}", flush=True)
    assert len(glob.glob(str(tmp_path / "cfg1_checkpoint" / "*"))) == 2
    _check_outputs(tmp_path, o, ref, T)


@pytest.mark.parametrize("mode", ["fused", "per-call"])
def test_driver_regrid_inside_the_run_matches_oracle(tmp_path, mode):
    """doubleResolutionAtStep inside the run (main.cpp:349-399): the grid doubles
    at the top of that step, the tank of the new size is bound with angle +=
    its increment, nu is re-derived for the new blade radius, and the run goes
    on; the checkpoint at the last step equals the oracle's."""
    case(f"driver regrid {mode}")
    inp, j = _config1_input(tmp_path)
    T, steps, at = 36, 8, 5
    j["GridParams"].update(x=T, y=T, z=T)
    j["RunningParams"].update(num_steps=steps, doubleResolutionAtStep=at)
    j["CheckpointParams"].update(checkpointRepeat=steps)
    inp.write_text(json.dumps(j))
    args = [os.path.join(ROOT, "tdlbcpp_amd", "tdlb_run"), "-i", str(inp)]
    if mode == "per-call":
        args.append("--per-call")
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "INCREASING RESOLUTION TO grid.x 72" in r.stdout
    d = tmp_path / "cfg1_checkpoint" / f"cfg1_checkpoint_step_{steps}"
    allp = json.loads((d / "AllParams.0.0.0.device.json").read_text())
    assert allp["ComputeUnitParams"]["x"] == 2 * T and allp["ComputeUnitParams"]["resolution"] == 3
    flow = dict(initialRho=1.0, cs0=0.1, g3=0.8, alpha=0.97, beta=1.9)
    nu, _ = O.calc_nu(7300.0, 0.1, int(T / 6), O.F32)
    o = O.OracleCU(T, dtype=O.F32, layout=O.IJKL, streaming=O.SIMPLE, use_les=True, nu=nu, **flow)
    o.initialise(1.0)
    rp = td.rushton_defaults(T, tip_speed=0.1, startup_steps=30)
    ref = ReferenceLoop(o, rp)
    ref.run(at - 2)  # steps 2 .. at-1
    o.double_resolution()
    rp2 = td.rushton_defaults(2 * T, tip_speed=0.1, startup_steps=30)
    nu2, _ = O.calc_nu(7300.0, 0.1, int(2 * T / 6), O.F32)
    o.set_flow(nu=nu2)
    angle = ref.angle + ref.rnd(td.rushton_step_increment(rp2, at))
    ref2 = ReferenceLoop(o, rp2, step=at, angle=angle)
    ref2.step = at - 1  # the regular body of step `at` follows
    ref2.run(steps - at + 1)
    q = np.fromfile(d / "checkpoint_grid.0.0.0.device.Q", dtype=np.float32)
    f = np.fromfile(d / "checkpoint_grid.0.0.0.device.F", dtype=np.float32)
    assert_close_moments(q, o.get(O.FIELD_Q), what="Q")
    assert_close_moments(f, o.get(O.FIELD_F), what="F", rho0=1e-3)
    assert np.array_equal(q, o.get(O.FIELD_Q))
