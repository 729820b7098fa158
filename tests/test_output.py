"""HDF5 output in the reference's layout (finitevolumesolver_amd/output.py,
io/fvs_hdf5.cpp) checked with the HDF5 tools of the image (h5dump) as an
independent reader: the composite onto the finest level follows
WriteHdf5UnRestricted (src/AMReX/output/WriteHdf5Impl.cpp:26-98), the file
layout CreateHdf5Database / OpenHdf5Database
(src/AMReX/PatchHierarchy.cpp:442-600)."""
from __future__ import annotations

import re
import shutil
import subprocess

import numpy as np
import pytest

import amr_cases as A
from finitevolumesolver_amd import output

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
needs_lib = pytest.mark.skipif(not output.LIB_PATH.exists(), reason="libfvs_hdf5.so not built")


def _interiors(state):
    return [[state(l, p) for p in range(len(lay))] for l, lay in enumerate(A.LAYOUTS)]


def test_composite_takes_the_finest_level_covering_each_cell():
    ints = _interiors(A.moving_state)
    fab = output.composite(A.DOMAIN, A.LAYOUTS, ints)
    assert fab.shape == (6, 128, 128)
    rng = np.random.default_rng(3)
    for _ in range(400):
        I, J = rng.integers(0, 128, 2)
        want = None
        for l in (2, 1, 0):
            r = 2 ** (2 - l)
            for (i0, j0, w, h), u in zip(A.LAYOUTS[l], ints[l]):
                if i0 <= I // r < i0 + w and j0 <= J // r < j0 + h:
                    want = u[:, J // r - j0, I // r - i0]
                    break
            if want is not None:
                break
        np.testing.assert_array_equal(fab[:4, J, I], want)
    comp = output.complete_2d(ints[0][0])
    e_kin = 0.5 * (ints[0][0][1] ** 2 + ints[0][0][2] ** 2) / ints[0][0][0]
    np.testing.assert_allclose(comp[4], (ints[0][0][3] - e_kin) * 0.4, rtol=1e-14)


def _dump(args):
    return subprocess.run([H5DUMP, *args], check=True, capture_output=True, text=True).stdout


@needs_lib
def test_hdf5_layout_and_appended_frames(tmp_path):
    if not shutil.which(H5DUMP):
        pytest.skip("no h5dump")
    f1 = output.composite(A.DOMAIN, A.LAYOUTS, _interiors(A.moving_state))
    f2 = output.composite(A.DOMAIN, A.LAYOUTS, _interiors(A.resting_blob))
    path = tmp_path / "grid.h5"
    output.write_hdf5(path, f1, (-1.0, -1.0), (A.DX0 / 4, A.DX0 / 4), 0.5, 7)
    output.write_hdf5(path, f2, (-1.0, -1.0), (A.DX0 / 4, A.DX0 / 4), 0.75, 9)
    head = _dump(["-H", str(path)])
    assert re.search(r'DATASET "data".*?\( 2, 6, 128, 128 \) / \( H5S_UNLIMITED, 6, 128, 128 \)',
                     head, re.S)
    assert 'ATTRIBUTE "xlower"' in head and 'ATTRIBUTE "cell_size"' in head
    fields = _dump(["-d", "/fields", str(path)])
    for name in output.FIELDS_2D:
        assert f'"{name}"' in fields
    def values(text):
        body = re.sub(r"\([^)]*\)", "", text[text.index("DATA {") + 6:])
        body = body[:body.index("}")]
        return [float(v) for v in re.findall(r"-?[0-9.]+(?:e[-+]?[0-9]+)?", body)]
    assert values(_dump(["-d", "/times", str(path)])) == [0.5, 0.75]
    assert values(_dump(["-d", "/cycles", str(path)])) == [7, 9]
    # frame 2, component 0 (density), the first 6 values in memory order
    got = values(_dump(["-d", "/data", "-s", "1,0,0,0", "-c", "1,1,1,6", "-m", "%.17g",
                        str(path)]))
    np.testing.assert_array_equal(np.array(got), f2[0].ravel()[:6])
    # a frame of another shape is refused
    with pytest.raises(OSError):
        output.write_hdf5(path, f1[:, :64], (-1.0, -1.0), (1.0, 1.0), 1.0, 10)


class _FakeHierarchy:
    """Stand-in with the attributes write_checkpoint reads (no GPU)."""

    class _FM:
        class equation:
            gamma = 1.4

    flux_method = _FM()
    ratio = 2
    cycle = 3
    time = 0.1 + 0.2
    layouts = [[(0, 0, 16, 32), (16, 0, 16, 32)], [(16, 16, 16, 24), (48, 0, 16, 16)]]
    nlevels = 2

    def interiors(self, level):
        rng = np.random.default_rng(level)
        out = []
        for (_, _, nx, ny) in self.layouts[level]:
            rho = 1.0 + rng.random((ny, nx))
            m0, m1 = rng.random((ny, nx)) - 0.5, rng.random((ny, nx)) - 0.5
            out.append(np.stack([rho, m0, m1, 3.0 + rng.random((ny, nx))]))
        return out


def test_checkpoint_layout_and_round_trip(tmp_path):
    """The checkpoint directory of WriteCheckpointFile
    (src/AMReX/PatchHierarchy.cpp:320-371): Header lines (title, finest
    level, per-level cycles and time points, BoxArrays), per level a VisMF
    header and data file whose FABs hold the complete state; reading it back
    returns the layouts, cycles, time (bit for bit: 17 digits) and the
    conservative components unchanged, with p and a from CompleteFromCons."""
    from finitevolumesolver_amd import amrex_files as ck
    from finitevolumesolver_amd.output import complete_2d
    h = _FakeHierarchy()
    ck.write_checkpoint(tmp_path / "chk00003", h)
    lines = (tmp_path / "chk00003" / "Header").read_text().split("\n")
    assert lines[:4] == ["Checkpoint File", "1", "3 6 ", "0.30000000000000004 0.30000000000000004 "]
    assert lines[4:8] == ["(2 0", "((0,0) (15,31) (0,0))", "((16,0) (31,31) (0,0))", ")"]
    finest, cycles, times, layouts = ck.read_header(tmp_path / "chk00003")
    assert finest == 1 and cycles == [3, 6] and times == [h.time, h.time]
    assert layouts == h.layouts
    head = (tmp_path / "chk00003" / "Level_0" / "data_H").read_text().split("\n")
    assert head[:4] == ["1", "1", "6", "0"] and head[8] == "2"
    for l in range(2):
        fabs = ck.read_level(tmp_path / "chk00003" / f"Level_{l}")
        for got, u in zip(fabs, h.interiors(l)):
            np.testing.assert_array_equal(got, complete_2d(u))
    raw = (tmp_path / "chk00003" / "Level_0" / "data_D_00000").read_bytes()
    assert raw.startswith(b"FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"
                          b"((0,0) (15,31) (0,0)) 6\n")


def test_plotfile_layout_and_round_trip(tmp_path):
    """WritePlotFile's HyperCLaw-V1.1 directory: variable names, time,
    problem domain, ratios, level domains and steps in the Header, each
    level's boxes in physical coordinates; the Cell multifabs read back to
    the complete states."""
    from finitevolumesolver_amd import amrex_files as ck
    from finitevolumesolver_amd.output import complete_2d
    h = _FakeHierarchy()
    h.domain, h.dx = (32, 32), [2.0 / 32, 1.0 / 32]
    ck.write_plotfile(tmp_path / "plt00003", h, prob_lo=(-1.0, -1.0))
    head = (tmp_path / "plt00003" / "Header").read_text().split("\n")
    assert head[:9] == ["HyperCLaw-V1.1", "6", "Density", "Momentum_0", "Momentum_1", "Energy",
                        "Pressure", "SpeedOfSound", "2"]
    assert head[9] == "0.30000000000000004" and head[10] == "1"
    assert head[11:14] == ["-1 -1 ", "1 1 ", "2 "]
    assert head[14] == "((0,0) (31,31) (0,0)) ((0,0) (63,63) (0,0)) " and head[15] == "3 6 "
    assert "Level_1/Cell" in head
    i = head.index("1 2 0.30000000000000004")
    assert head[i + 1] == "6" and head[i + 2] == "-0.5 0" and head[i + 3] == "-0.5 0.25"
    names, time, layouts, fabs = ck.read_plotfile(tmp_path / "plt00003")
    assert names[0] == "Density" and time == h.time and layouts == h.layouts
    for l in range(2):
        for got, u in zip(fabs[l], h.interiors(l)):
            np.testing.assert_array_equal(got, complete_2d(u))
