"""The C++ driver's reading of the reference's input JSON (Params/*.hpp formats,
main.cpp:135-164) — CPU only (--print-params does not touch the GPU)."""
import json
import os
import subprocess

import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "tdlbcpp_amd", "tdlb_run")
GOLD = os.path.join(ROOT, "tests", "golden")


def run(*args):
    return subprocess.run([RUN, *args], capture_output=True, text=True, timeout=60)


@pytest.fixture(scope="module")
def sample():
    with open(os.path.join(GOLD, "input_debug_gridx60_numSteps20.json")) as fh:
        return json.load(fh)


def test_print_params_round_trips_the_reference_sample(sample):
    r = run("--print-params", "-i", os.path.join(GOLD, "input_debug_gridx60_numSteps20.json"))
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    for section in ("GridParams", "CheckpointParams", "BinFileParams", "OutputParams"):
        assert got[section] == sample[section], section
    run_exp = dict(sample["RunningParams"])
    run_exp["doubleResolutionAtStep"] = 10          # RunningParams.hpp:68 default when absent
    assert got["RunningParams"] == run_exp
    flow = dict(sample["FlowParams"])
    nu, rem = O.calc_nu(flow["reMNonDimensional"], flow["uav"], 10, O.F32)  # int(60/6)
    assert got["FlowParams"]["nu"] == pytest.approx(nu, rel=1e-7)
    assert got["FlowParams"]["Re_m"] == pytest.approx(rem, rel=1e-7)
    for k, v in flow.items():
        if k not in ("nu", "Re_m"):
            assert got["FlowParams"][k] == v, k
    cu = got["ComputeUnitParams"]
    assert (cu["x"], cu["y"], cu["z"], cu["ghost"]) == (60, 60, 60, 1)   # ghost = multiStep (main.cpp:225)


def test_nu_from_json_keeps_the_reference_main_behaviour(sample):
    r = run("--print-params", "--nu-from-json", "-i", os.path.join(GOLD, "input_debug_gridx60_numSteps20.json"))
    assert r.returncode == 0
    assert json.loads(r.stdout)["FlowParams"]["nu"] == sample["FlowParams"]["nu"]


def test_missing_key_fails_like_the_reference(tmp_path, sample):
    bad = json.loads(json.dumps(sample))
    del bad["FlowParams"]["g3"]
    p = tmp_path / "bad.json"
    p.write_text(json.dumps(bad))
    r = run("--print-params", "-i", str(p))
    assert r.returncode != 0
    assert "Failed to parse FlowParams" in r.stderr


def test_mistyped_value_fails(tmp_path, sample):
    bad = json.loads(json.dumps(sample))
    bad["FlowParams"]["useLES"] = "yes"
    p = tmp_path / "bad.json"
    p.write_text(json.dumps(bad))
    r = run("--print-params", "-i", str(p))
    assert r.returncode != 0 and "FlowParams" in r.stderr


def test_malformed_json_fails(tmp_path):
    p = tmp_path / "broken.json"
    p.write_text('{"GridParams": {"x": 60,,}')
    r = run("--print-params", "-i", str(p))
    assert r.returncode != 0 and "parsing input json" in r.stderr


def test_second_sample_input_parses():
    r = run("--print-params", "-i", os.path.join(GOLD, "input_debug_gridx72_IncreaseResolution.json"))
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    assert got["GridParams"]["x"] == 72 and got["FlowParams"]["streaming"] == "Simple"
