"""The device FP64 division correction (RoundQuotient in fvs_physics.hpp),
compiled for the host: quotients one ulp off in either direction are
returned correctly rounded; zero, infinite and NaN quotients are kept."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
CSRC = HERE.parent / "finitevolumesolver_amd" / "csrc"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_round_quotient_on_host(tmp_path):
    exe = tmp_path / "round_quotient"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", f"-I{HERE / 'host_stub'}",
                    f"-I{CSRC}", str(HERE / "host_stub" / "round_quotient.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout
