"""CPU tests of the metrosim command line (mcgpu_amd/bin/metrosim): option
handling of CommandParsing.cpp:36-255 and the input-file errors the
reference's own tests look for (TestUtil.cpp:132-144 reads the first line
holding "Error" and expects "loadBoxData()" at columns 7-19).  None of these
reach the device."""
import subprocess
from pathlib import Path

import pytest

from config_cases import write_config

METROSIM = Path(__file__).resolve().parents[1] / "mcgpu_amd" / "bin" / "metrosim"


def run(args, cwd):
    if not METROSIM.exists():
        pytest.fail(f"{METROSIM} is missing: build with python -m mcgpu_amd.build")
    return subprocess.run([str(METROSIM), *args], cwd=cwd, capture_output=True,
                          text=True, timeout=60)


def error_token(text: str) -> str:
    """getErrorResult (TestUtil.cpp:132-144): the first line containing
    "Error", characters 7..19."""
    for line in text.splitlines():
        if "Error" in line:
            return line[7:20]
    return "ERROR: COULD NOT PARSE ERROR FILE!"


@pytest.mark.parametrize("args,msg", [
    ([], "Input file not specified"),
    (["a.config", "b.config"], "Too many input files specified"),
    (["a.config", "-s", "-p"], "Cannot specify both GPU and CPU modes"),
    (["a.txt"], "Must specify a config or state file"),
    (["a.config", "-n", "0"], "Step count must be greater than zero"),
    (["a.config", "-n", "x"], "Invalid step count"),
    (["a.config", "-i", "-3"], "Status Interval must be non-negative"),
    (["a.config", "-q"], "Unknown option -q"),
    (["a.config", "--steps"], "A required argument is missing"),
])
def test_usage_errors(args, msg, tmp_path):
    r = run(args, tmp_path)
    assert r.returncode != 0
    assert msg in r.stderr


def test_help_and_version_exit(tmp_path):
    r = run(["-h"], tmp_path)
    assert r.returncode != 0 and "Usage: metrosim <inputfile> [options]" in r.stdout
    r = run(["-V"], tmp_path)
    assert r.returncode != 0 and "MCGPU Monte Carlo Simulator" in r.stdout


@pytest.mark.parametrize("z,prim,edge,n", [
    ("meohtip3.z", "1", 23.7856, 256),          # meohtip3Test.cpp:48-57
    ("t3pdim.z", "[1,2]", 26.15, 256),          # t3pdimTest.cpp:75-84
    ("meoh500.z", "[1,2],[3,4]", 32.91, 500),   # Methanol500Test.cpp:182-190
])
def test_load_errors_report_loadboxdata(z, prim, edge, n, tmp_path):
    cfg = write_config(tmp_path / "bad.config", z, edge, n, 100, 11.0, 0.12, 12.0, prim)
    r = run([str(cfg), "-s", "-i", "10000"], tmp_path)
    assert r.returncode != 0
    assert error_token(r.stdout + r.stderr) == "loadBoxData()"


def test_unknown_strategy_is_rejected(tmp_path):
    cfg = write_config(tmp_path / "s.config", "meoh.z", 26.15, 256, 100, 11.0, 0.12,
                       12.0, "1")
    r = run([str(cfg), "-S", "octree"], tmp_path)
    assert r.returncode != 0 and "Unknown energy calculation strategy" in r.stderr
