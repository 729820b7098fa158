"""The facts the exact-Godunov CFL pass's speed floor rests on, checked on
the CPU oracle's exact solver (ExactRiemannSolver.hpp:34-353).

The level's CFL pass skips the Newton solve of a face whose certified speed
bound stays below the floor.  The bound applies to pairs with positive
densities and pressures, a pressure ratio of at most 4, sound speeds of at
most 1e6 and |uR - uL| no larger than the smaller sound speed, and it is
used only with a Newton iteration cap of at least 16 (a skipped face then
is one the reference would have solved without reaching its cap).  Checked
here over that class, its corners included:
  - Newton from the mean pressure converges within 16 iterations;
  - Toro's pressure function is increasing and concave, so for uR >= uL
    (f(max p) >= 0) the Newton iterates from the mean never pass max(pL, pR);
  - the signal speeds then stay within the shock speeds of a star pressure
    max(pL, pR): |uK| + aK sqrt(1 + (g + 1) / (2 g) (max p / pK - 1))."""
from __future__ import annotations

import numpy as np

import oracle_lib as O

G = 1.4
CAP = 16


def _class_pairs(n, seed):
    rng = np.random.default_rng(seed)
    rhoL = 10.0 ** rng.uniform(-3, 3, n)
    rhoR = 10.0 ** rng.uniform(-3, 3, n)
    pmin = 10.0 ** rng.uniform(-3, 3, n)
    ratio = rng.uniform(1.0, 4.0, n)
    ratio[: n // 8] = 4.0  # the class's corner
    left_high = rng.random(n) < 0.5
    pL = np.where(left_high, pmin * ratio, pmin)
    pR = np.where(left_high, pmin, pmin * ratio)
    aL, aR = np.sqrt(G * pL / rhoL), np.sqrt(G * pR / rhoR)
    amin = np.minimum(aL, aR)
    du = rng.uniform(-1.0, 1.0, n) * amin
    du[n // 8: n // 4] = -amin[n // 8: n // 4]
    du[n // 4: 3 * n // 8] = amin[n // 4: 3 * n // 8]
    uL = rng.uniform(-2.0, 2.0, n) * np.maximum(aL, aR)
    uR = uL + du

    def complete(rho, u, p):
        cons = np.zeros((4, n))
        cons[0] = rho
        cons[1] = rho * u
        cons[3] = p / (G - 1.0) + 0.5 * rho * u * u
        return O.complete_from_cons(cons, 2, G)

    return complete(rhoL, uL, pL), complete(rhoR, uR, pR)


def test_newton_from_the_mean_converges_within_the_cap_on_the_class():
    Lq, Rq = _class_pairs(400_000, seed=11)
    _, _, it, conv = O.exact_middle_state(Lq, Rq, 2, 0, max_it=CAP)
    assert np.all(conv == 1), int(np.sum(conv == 0))
    assert int(it.max()) <= 8, int(it.max())


def test_star_pressure_and_signals_within_the_pmax_shock_bound():
    Lq, Rq = _class_pairs(200_000, seed=12)
    uL, uR = Lq[1] / Lq[0], Rq[1] / Rq[0]
    pL, pR, aL, aR = Lq[4], Rq[4], Lq[5], Rq[5]
    keep = uR >= uL
    pM, _, _, conv = O.exact_middle_state(Lq, Rq, 2, 0)
    pmax = np.maximum(pL, pR)
    assert np.all(conv[keep] == 1)
    assert np.all(pM[keep] <= pmax[keep])
    sL, sR = O.exact_signals(Lq, Rq, 2, 0, G)
    c = (G + 1.0) / (2.0 * G)
    wL = np.sqrt(1.0 + c * (pmax / pL - 1.0))
    wR = np.sqrt(1.0 + c * (pmax / pR - 1.0))
    bound = np.maximum(np.abs(uL) + aL * wL, np.abs(uR) + aR * wR)
    speed = np.maximum(np.abs(sL), np.abs(sR))
    assert np.all(speed[keep] <= bound[keep] * (1.0 + 1e-15)), \
        float(np.max(speed[keep] / bound[keep]))
