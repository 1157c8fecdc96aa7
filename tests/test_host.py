"""Host-side tests (no GPU): C++ unit tests, the CLI's drop-in behaviour on the
host device, and the C ABI library's exported symbols."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import onesolver_amd
import oracle
from conftest import GOLDEN, ROOT
from qubo_io import load_qubo

CLI = onesolver_amd.CLI_PATH


def test_cpp_unit_tests():
    r = subprocess.run([onesolver_amd.TEST_BIN_PATH], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout


def _declared_symbols():
    syms = []
    inc = os.path.join(ROOT, "include")
    for name in os.listdir(inc):
        if name.endswith(".h"):
            with open(os.path.join(inc, name)) as f:
                text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
            syms += re.findall(r"^\s*(?:const\s+)?[\w]+\s*\**\s*(os_\w+)\s*\(", text, flags=re.M)
    return sorted(set(syms))


def test_library_exports_every_declared_symbol():
    syms = _declared_symbols()
    assert len(syms) >= 8
    lib = ctypes.CDLL(onesolver_amd.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert lib.os_version() >= 10000


def test_null_and_invalid_arguments_are_rejected_without_a_gpu():
    L = onesolver_amd.lib()
    assert L.os_device_count(None) == onesolver_amd.OS_ERR_INVALID
    assert b"null" in L.os_last_error()
    assert L.os_ctx_create(0, None) == onesolver_amd.OS_ERR_INVALID
    assert L.os_anneal_reference(None, 4, None, None, 1, 1, 1, 0, None, None, None, None) \
        == onesolver_amd.OS_ERR_INVALID


def _run(args, **kw):
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120, **kw)


def test_cli_help_lists_reference_options():
    r = _run(["--help"])
    assert r.returncode == 0
    for opt in ["--input arg", "--output arg", "--num-iter arg (=100)", "--num-tries arg (=100)",
                "--schedule-type arg (=geometric)", "--beta-min arg (=0.10000000000000001)",
                "--beta-max arg (=1)", "--device-type arg (=host)"]:
        assert opt in r.stdout


@pytest.mark.parametrize("args,msg", [
    ([], "No input file provided."),
    (["--input", "x"], "No output file provided."),
    (["--input", "x", "--output", "y", "--device-type", "tpu"], "Unknown device type: tpu"),
    (["--input", "x", "--output", "y", "--schedule-type", "cosine"], "Unknown beta schedule: cosine"),
    (["--input", "x", "--output", "y", "--beta-min", "-1"],
     "Invalid schedule, both ends of beta range need to be positive"),
    (["--input", "x", "--output", "y", "--beta-min", "2", "--beta-max", "1"],
     "Invalid schedule, initial beta is not lesser than final beta"),
    (["--input", "/nonexistent.qubo", "--output", "y"], "can not open input file: /nonexistent.qubo"),
])
def test_cli_validation_returns_minus_one(args, msg):
    r = _run(args)
    assert r.returncode == 255
    assert msg in r.stderr


@pytest.mark.parametrize("args,msg", [
    (["--path", "dense"], "--beta-kind, --path, --refresh-interval and --warm-start need --mode sweep"),
    (["--warm-start", "w.bin"], "need --mode sweep"),
    (["--mode", "sweeps"], "Unknown mode: sweeps"),
    (["--mode", "sweep"], "Sweep mode runs on the gpu device"),
    (["--mode", "sweep", "--device-type", "cpu"], "Sweep mode runs on the gpu device"),
    (["--mode", "sweep", "--device-type", "gpu", "--path", "csr"], "Unknown path: csr"),
    (["--mode", "sweep", "--device-type", "gpu", "--beta-kind", "x"], "Unknown beta kind: x"),
    (["--mode", "sweep", "--device-type", "gpu", "--refresh-interval", "-2"],
     "Invalid refresh interval, need 0 or more"),
])
def test_cli_sweep_mode_validation(args, msg):
    r = _run(["--input", "x", "--output", "y"] + args)
    assert r.returncode == 255
    assert msg in r.stderr


def test_cli_help_lists_sweep_mode_options():
    r = _run(["--help"])
    for opt in ["--mode arg (=reference)", "--beta-kind arg (=temperature)", "--path arg (=auto)",
                "--refresh-interval arg (=0)", "--warm-start arg"]:
        assert opt in r.stdout


def test_cli_parse_errors_exit_one(tmp_path):
    r = _run(["--bogus", "1"])
    assert r.returncode == 1 and "error: unrecognised option '--bogus'" in r.stderr
    r = _run(["--input", os.path.join(GOLDEN, "examples", "dwave_doc.qubo"), "--output",
              str(tmp_path / "o.csv")])
    assert r.returncode == 1
    assert "error: Parsing failed. Incorrect file format." in r.stderr
    r = _run(["--num-iter", "abc"])
    assert r.returncode == 1 and "is invalid" in r.stderr


def test_cli_host_device_solves_test1(tmp_path):
    out = tmp_path / "result.csv"
    r = _run(["--input", os.path.join(GOLDEN, "examples", "test1.qubo"), "--output", str(out),
              "--num-iter", "500", "--num-tries", "32", "--beta-max", "5"])
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].startswith("Reading input from: ")
    assert lines[2] == "Schedule type: geometric"
    assert lines[3] == "Beta range: [0.1, 5]"
    assert lines[4] == "Number of iterations: 500"
    assert lines[5] == "Number of tries: 32"
    assert lines[6].startswith("Using device: ")
    text = out.read_text()
    head, row = text.splitlines()
    assert head == "0,1,2,3,energy"
    vals = row.split(",")
    x = np.array([int(v) for v in vals[:4]], dtype=np.uint8)
    q = load_qubo(os.path.join(GOLDEN, "examples", "test1.qubo"))
    assert float(vals[4]) == oracle.energy(q, x) == -12.0


def test_cli_prefix_and_equals_forms(tmp_path):
    out = tmp_path / "r.csv"
    r = _run(["--inp=" + os.path.join(GOLDEN, "examples", "simple.qubo"), "--out", str(out),
              "--num-i=50", "--num-t", "4", "--schedule=linear"])
    assert r.returncode == 0, r.stderr
    assert "Schedule type: linear" in r.stdout
    r = _run(["--beta", "1"])
    assert r.returncode == 1 and "ambiguous" in r.stderr


def test_bench_rejects_more_than_one_gpu():
    """bench.py runs on one GPU; --gpus N > 1 exits with status 2 and says so."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def _run_exh(args):
    return subprocess.run([onesolver_amd.EXHAUSTIVE_CLI_PATH] + args, capture_output=True,
                          text=True, timeout=120)


def test_exhaustive_cli_help_and_validation():
    r = _run_exh(["--help"])
    assert r.returncode == 0
    for opt in ["--input arg", "--output arg", "--device-type arg (=host)"]:
        assert opt in r.stdout
    for args, msg in [([], "No input file provided."),
                      (["--input", "x"], "No output file provided."),
                      (["--input", "x", "--output", "y", "--device-type", "tpu"],
                       "Unknown device type: tpu"),
                      (["--input", "/nonexistent.qubo", "--output", "y"],
                       "can not open input file: /nonexistent.qubo")]:
        r = _run_exh(args)
        assert r.returncode == 255 and msg in r.stderr, (args, r.stderr)
    r = _run_exh(["--bogus", "1"])
    assert r.returncode == 1 and "error: unrecognised option '--bogus'" in r.stderr


@pytest.mark.parametrize("device", ["host", "cpu"])
def test_exhaustive_cli_host_devices_solve_test1(tmp_path, device):
    out = tmp_path / "result.csv"
    r = _run_exh(["--input", os.path.join(GOLDEN, "examples", "test1.qubo"), "--output",
                  str(out), "--device-type", device])
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].startswith("Reading input from: ")
    assert lines[1].startswith("Output will be saved to: ")
    assert lines[2].startswith("Using device: ")
    assert out.read_text().splitlines() == ["0,1,2,3,energy", "1,1,0,1,-12"]


def _batch(args):
    return subprocess.run([onesolver_amd.BATCH_CLI_PATH] + args, capture_output=True, text=True,
                          timeout=60)


def test_batch_cli_help_and_validation(tmp_path):
    r = _batch(["--help"])
    assert r.returncode == 0 and "--beta-kind" in r.stdout and "--path" in r.stdout
    d = str(tmp_path)
    for args, msg in [([], "No input file provided."),
                      (["--input", d], "No output directory provided."),
                      (["--input", d, "--output", d, "--schedule-type", "x"],
                       "Unknown beta schedule: x"),
                      (["--input", d, "--output", d, "--beta-min", "2"],
                       "Invalid schedule, initial beta is not lesser than final beta"),
                      (["--input", d, "--output", d, "--beta-kind", "x"], "Unknown beta kind: x"),
                      (["--input", d, "--output", d, "--path", "x"], "Unknown path: x"),
                      (["--input", d, "--output", d + "/missing"],
                       "can not open output directory: " + d + "/missing")]:
        r = _batch(args)
        assert r.returncode == 255 and msg in r.stderr, (args, r.stderr)
    r = _batch(["--input", d + "/nothing.txt", "--output", d])
    assert r.returncode == 1 and "can not open input file" in r.stderr
    r = _batch(["--input", d, "--output", d])
    assert r.returncode == 1 and "no .qubo or .osqb files" in r.stderr


def test_batch_cli_without_gpu_reports_no_device(tmp_path):
    import shutil
    if onesolver_amd.device_count() > 0:
        pytest.skip("a GPU is present")
    shutil.copy(os.path.join(GOLDEN, "examples", "test1.qubo"), tmp_path / "a.qubo")
    r = _batch(["--input", str(tmp_path), "--output", str(tmp_path)])
    assert r.returncode == 1
    assert "No devices of given type could be initialized." in r.stderr
