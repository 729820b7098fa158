"""The X sweep's slot walker (finitevolumesolver_amd/csrc/fvs_xslots.hpp)
compiled for the host: over ragged 2D and 3D row counts and row lengths
(1 .. 256 cells, around the 60-slot chunk) and persistent wave counts from
4 to 4,096, every lane's loads stay inside the arrays, its row never lies
below the wave's base row (the unsigned offset cannot wrap), valid lanes sit
on their true slot after each stride, and every cell is stored exactly
once (tests/host_stub/xslots.cpp)."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
CSRC = HERE.parent / "finitevolumesolver_amd" / "csrc"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_x_sweep_slot_walker_on_host(tmp_path):
    exe = tmp_path / "xslots"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{HERE / 'host_stub'}", f"-I{CSRC}",
                    str(HERE / "host_stub" / "xslots.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    bad, cases = r.stdout.split()[-2:]
    assert r.returncode == 0 and bad == "0", r.stdout[-2000:]
    assert int(cases) > 500
