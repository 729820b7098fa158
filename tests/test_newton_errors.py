"""Newton iteration cap and vacuum on the CPU oracle: the outcomes the GPU
error paths are checked against (tests/test_gpu_newton_errors.py).

NewtonIteration (include/fub/NewtonIteration.hpp:44-50) throws when it
stops at its iteration cap with f(x) > tolerance; an iterate stopped below
the root (f(x) < -tolerance) is returned as if converged.  With one
iteration from the mean pressure, Toro's test 2 (the 123 problem) stops
above its root and is reported, the other four stop below theirs.  A pair
that generates vacuum (2 aL / (g - 1) + 2 aR / (g - 1) <= uR - uL) has a
pressure function positive everywhere, so Newton never reaches the
tolerance and every cap is reported."""
from __future__ import annotations

import numpy as np

import oracle_lib as O
from test_oracle import TORO, _pressure_function, _prim_pair

G = 1.4


def test_one_iteration_reports_only_positive_residuals():
    L_, R_ = _prim_pair(TORO)
    pM, _, it, conv = O.exact_middle_state(L_, R_, 2, 0, max_it=1)
    assert np.all(it == 1)
    for i in range(len(TORO)):
        f = _pressure_function(pM[i], L_, R_, i)
        assert conv[i] == (0 if f > 1e-7 else 1), (i, f)
    assert list(conv) == [1, 0, 1, 1, 1]
    # with the reference's cap every pair converges
    _, _, it, conv = O.exact_middle_state(L_, R_, 2, 0)
    assert np.all(conv == 1) and np.all(it < 20)


def test_vacuum_pair_hits_the_cap():
    rows = [(1.0, -5.0, 0.4, 1.0, 5.0, 0.4), (1.0, -4.0, 0.2, 0.5, 6.0, 0.1)]
    L_, R_ = _prim_pair([r + (0.0, 0.0) for r in rows])
    for i, r in enumerate(rows):
        aL, aR = np.sqrt(G * r[2] / r[0]), np.sqrt(G * r[5] / r[3])
        assert 2 * aL / (G - 1) + 2 * aR / (G - 1) <= r[4] - r[1]
        assert _pressure_function(1e-300, L_, R_, i) > 0.0
    for cap in (50, 5000):
        _, _, it, conv = O.exact_middle_state(L_, R_, 2, 0, max_it=cap)
        assert np.all(conv == 0) and np.all(it == cap)
