"""The oracle against its committed golden vectors (tests/golden): every
output recomputed from the stored inputs must equal the stored one bit for
bit.  A mismatch means the oracle's arithmetic changed; the stored SHA-256
of the oracle sources tells whether the sources did."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import golden as G  # noqa: E402


@pytest.mark.parametrize("case", list(G.CASES))
def test_oracle_reproduces_golden_vectors(case):
    inp, want, digest = G.load_case(case)
    got = G.CASES[case][1](inp)
    assert set(got) == set(want)
    changed = digest != G.oracle_hash()
    for k in want:
        assert np.array_equal(np.asarray(got[k]), want[k]), (
            f"{case}:{k} differs from the golden vector"
            + (" (oracle sources changed since the vectors were written)" if changed else ""))


def test_golden_vectors_are_finite_and_sized():
    """Fixture sanity: finite outputs, ~10^3 faces per flux case."""
    for case in G.CASES:
        _, want, _ = G.load_case(case)
        for k, v in want.items():
            if v.dtype.kind == "f":
                assert np.all(np.isfinite(v)), (case, k)
    _, f2, _ = G.load_case("fluxes_2d")
    _, f3, _ = G.load_case("fluxes_3d")
    faces = sum(v.size // 4 for k, v in f2.items() if "/flux" in k) + \
        sum(v.size // 5 for k, v in f3.items() if "/flux" in k)
    assert faces >= 1000
