"""MusclHancockPrim in the oracle against what fixes it
(src/equations/perfect_gas/MusclHancockPrimMethod.hpp:59-176).

  * constant stencil: zero slopes, the face states are the cell state and
    the HLLEM flux is the physical flux (to the ulps of rho * (m / rho));
  * supersonic flow to the right with uniform velocity and pressure: HLLEM
    upwinds to the left face state, whose density is the textbook
    MUSCL-Hancock value rho_1 + s (1 - lambda u) / 2 with the limited slope
    s = sigma(r) (rho_2 - rho_0) / 2 worked out by hand (linear data: the
    central slope; r = 1/2: sigma = 2/3; an extremum: zero), so a wrong sign
    of lambda, a wrong cell or a wrong limiter branch fails;
  * the stable dt is the HLLEM dt of the middle two cells."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as O
from finitevolumesolver_amd.patch import PerfectGas, make_desc

GAMMA = 1.4


def _complete(rho, u, p):
    D = len(u)
    a = np.sqrt(GAMMA * p / rho)
    E = p / (GAMMA - 1.0) + 0.5 * rho * sum(v * v for v in u)
    return np.array([rho, *[rho * v for v in u], E, p, a])


def _stencil(states):
    return np.stack(states)[:, :, None]  # (4, D+4, 1)


@pytest.mark.parametrize("D", [2, 3])
def test_constant_state_gives_physical_flux(D):
    desc = make_desc(PerfectGas(dim=D), reconstruction="MusclHancockPrim")
    q = _complete(1.3, [0.4, -0.2, 0.1][:D], 2.1)
    for d in range(D):
        f, rc = O.face_flux(desc, _stencil([q] * 4), d, 0.01, 0.1)
        assert rc == 0
        ref = O.physical_flux(np.ascontiguousarray(q[:, None]), D, d)[:, 0]
        np.testing.assert_allclose(f[:, 0], ref, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("rhos,slope", [
    ((1.0, 1.1, 1.2, 1.3), 0.1),                  # linear: r = 1, central slope
    ((1.0, 1.1, 1.4, 1.0), (2.0 / 3.0) * 0.2),    # r = 1/2: sigma = 2/3 of (1.4 - 1) / 2
    ((1.0, 1.2, 1.1, 5.0), 0.0),                  # extremum: zero slope
])
def test_supersonic_face_density_closed_form(rhos, slope):
    D = 2
    desc = make_desc(PerfectGas(dim=D), reconstruction="MusclHancockPrim")
    u, p, dt, dx = 3.0, 1.0, 0.01, 0.1
    lam = dt / dx
    st = [_complete(r, [u, 0.0], p) for r in rhos]
    f, rc = O.face_flux(desc, _stencil(st), 0, dt, dx)
    assert rc == 0
    rho_face = rhos[1] + 0.5 * slope * (1.0 - lam * u)
    np.testing.assert_allclose(f[0, 0], rho_face * u, rtol=1e-14)
    # momentum flux of the upwinded state: rho u^2 + p (uniform u, p)
    np.testing.assert_allclose(f[1, 0], rho_face * u * u + p, rtol=1e-14)


def test_stable_dt_is_hllem_of_middle_cells():
    import ctypes as C
    D = 2
    prim = make_desc(PerfectGas(dim=D), reconstruction="MusclHancockPrim")
    st = [_complete(r, [0.3 * r, -0.1], 1.0 + 0.2 * r) for r in (1.0, 1.4, 0.7, 2.0)]
    row = np.ascontiguousarray(np.stack(st, axis=1)[:, None, None, :])  # (6, 1, 1, 4)
    out = C.c_double()
    rc = O.load().oracle_stable_dt(C.byref(prim), C.byref(O.host_view(row)), 0.1, 0,
                                   C.byref(out))
    assert rc == 0
    L, R = (np.ascontiguousarray(s[:, None]) for s in st[1:3])
    assert out.value == O.hllem_stable_dt(L, R, D, 0, 0.1)[0]


# MusclHancockCharacteristic (src/equations/perfect_gas/
# MusclHancockCharacteristicMethod.hpp:26-320): with uniform velocity and
# pressure only the entropy wave carries a jump, so the face density is
# rho_1 + s (1 - lambda u) / 2 with s the VanLeer mean 2 dl dr / (dl + dr)
# of the two density differences (linear data: the difference itself).

@pytest.mark.parametrize("D", [2, 3])
def test_characteristic_constant_state_gives_physical_flux(D):
    desc = make_desc(PerfectGas(dim=D), reconstruction="MusclHancockCharacteristic")
    q = _complete(1.3, [0.4, -0.2, 0.1][:D], 2.1)
    for d in range(D):
        f, rc = O.face_flux(desc, _stencil([q] * 4), d, 0.01, 0.1)
        assert rc == 0
        ref = O.physical_flux(np.ascontiguousarray(q[:, None]), D, d)[:, 0]
        np.testing.assert_allclose(f[:, 0], ref, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("rhos,slope", [
    ((1.0, 1.1, 1.2, 1.3), 0.1),                  # linear
    ((1.0, 1.1, 1.4, 1.0), 2 * 0.1 * 0.3 / 0.4),  # VanLeer mean of 0.1 and 0.3
    ((1.0, 1.2, 1.1, 5.0), 0.0),                  # extremum
])
def test_characteristic_supersonic_entropy_wave(rhos, slope):
    desc = make_desc(PerfectGas(dim=2), reconstruction="MusclHancockCharacteristic")
    u, p, dt, dx = 3.0, 1.0, 0.01, 0.1
    st = [_complete(r, [u, 0.0], p) for r in rhos]
    f, rc = O.face_flux(desc, _stencil(st), 0, dt, dx)
    assert rc == 0
    rho_face = rhos[1] + 0.5 * slope * (1.0 - dt / dx * u)
    np.testing.assert_allclose(f[0, 0], rho_face * u, rtol=1e-14)
    np.testing.assert_allclose(f[1, 0], rho_face * u * u + p, rtol=1e-14)
