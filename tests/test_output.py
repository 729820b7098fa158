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
