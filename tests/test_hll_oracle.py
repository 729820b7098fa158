"""Pins the oracle's HLL flux with Einfeldt signal speeds
(include/fub/flux_method/HllMethod.hpp:159-347,
src/equations/perfect_gas/EinfeldtSignalVelocities.hpp:24-97) by closed
forms and symmetries (CPU only).  The reference's own HLL test
(tests/test.simd_FluxMethod.cpp:54-186) asserts only that its scalar and
array paths agree; its single-state input has the closed form below, and
its 16 x 16 inputs are the golden case hll_simd_2d."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle_lib as O
from finitevolumesolver_amd.initial_data import random_states
from finitevolumesolver_amd.patch import PerfectGas, make_desc

GAMMA = 1.4


def _hll(qL, qR, dim, d):
    desc = make_desc(PerfectGas(dim), base="HLL", reconstruction="None")
    f, rc = O.face_flux(desc, np.stack([qL, qR]), d, 0.0, 1.0)
    assert rc == 0
    return f


def _hll_dt(qL, qR, dim, d, dx):
    """Stable dt of each face (qL[:, i], qR[:, i]) through the view driver:
    a two-cell patch per face."""
    import ctypes as C
    desc = make_desc(PerfectGas(dim), base="HLL", reconstruction="None")
    out = []
    for i in range(qL.shape[1]):
        shape = [dim + 4, 1, 1, 1]
        shape[3 - d] = 2
        st = np.zeros(shape)
        idx0 = [slice(None), 0, 0, 0]
        idx1 = [slice(None), 0, 0, 0]
        idx1[3 - d] = 1
        st[tuple(idx0)] = qL[:, i]
        st[tuple(idx1)] = qR[:, i]
        v = O.host_view(st)
        dt = C.c_double()
        O.load().oracle_stable_dt(C.byref(desc), C.byref(v), dx, d, C.byref(dt))
        out.append(dt.value)
    return np.array(out)


def from_prim(rho, vel, p, dim):
    n = len(rho)
    out = np.empty((dim + 4, n))
    O.load().oracle_complete_from_prim(dim, GAMMA, n, O.ptr(np.ascontiguousarray(rho, dtype=float)),
                                       O.ptr(np.ascontiguousarray(vel, dtype=float)),
                                       O.ptr(np.ascontiguousarray(p, dtype=float)), O.ptr(out))
    return out


def test_reference_single_state_closed_form():
    """MUSCL-Hancock(HLL) between gases at rest, rho = 1, E = 2 | E = 1
    (test.simd_FluxMethod.cpp:54-106): MinMod slopes vanish, so the flux is
    HLL(A, B) with Roe speeds sqrt(0.84), sL = -sqrt(1.12), sR = sqrt(0.84)."""
    cons = np.array([[1.0, 0.0, 0.0, 2.0]] * 2 + [[1.0, 0.0, 0.0, 1.0]] * 2).T
    comp = O.complete_from_cons(np.ascontiguousarray(cons), 2)
    desc = make_desc(PerfectGas(2), base="HLL", reconstruction="MusclHancock")
    f, rc = O.face_flux(desc, np.ascontiguousarray(comp.T[:, :, None]), 0, 1.0, 1.0)
    assert rc == 0
    sL, sR = -math.sqrt(1.4 * 0.8), math.sqrt(0.4 * 2.1)
    expect = [0.0, (sR * 0.8 - sL * 0.4) / (sR - sL), 0.0, sL * sR * (1.0 - 2.0) / (sR - sL)]
    assert f[0, 0] == 0.0 and f[2, 0] == 0.0
    assert f[1, 0] == pytest.approx(expect[1], rel=1e-15)
    assert f[3, 0] == pytest.approx(expect[3], rel=1e-15)


@pytest.mark.parametrize("dim", [2, 3])
def test_hll_consistency(dim):
    q = O.complete_from_cons(random_states(200, dim, seed=21), dim)
    for d in range(dim):
        f = _hll(q, q, dim, d)
        ref = O.physical_flux(q, dim, d)
        assert np.max(np.abs(f - ref) / np.maximum(np.abs(ref), 1.0)) < 1e-13


@pytest.mark.parametrize("dim", [2, 3])
def test_hll_supersonic_upwinding(dim):
    rng = np.random.default_rng(22)
    n = 100
    for d in range(dim):
        rhoL, pL = rng.uniform(0.5, 2, n), rng.uniform(0.5, 2, n)
        rhoR, pR = rhoL * rng.uniform(0.9, 1.1, n), pL * rng.uniform(0.9, 1.1, n)
        a = np.sqrt(GAMMA * np.maximum(pL / rhoL, pR / rhoR))
        vL, vR = rng.uniform(-0.3, 0.3, (dim, n)), rng.uniform(-0.3, 0.3, (dim, n))
        vL[d] = vR[d] = 4.0 * a
        qL, qR = from_prim(rhoL, vL, pL, dim), from_prim(rhoR, vR, pR, dim)
        f = _hll(qL, qR, dim, d)
        ref = O.physical_flux(qL, dim, d)
        assert np.max(np.abs(f - ref) / np.maximum(np.abs(ref), 1.0)) < 1e-13
        # and leftwards: the right state's flux
        vL[d] = vR[d] = -4.0 * a
        qL, qR = from_prim(rhoL, vL, pL, dim), from_prim(rhoR, vR, pR, dim)
        f = _hll(qL, qR, dim, d)
        ref = O.physical_flux(qR, dim, d)
        assert np.max(np.abs(f - ref) / np.maximum(np.abs(ref), 1.0)) < 1e-13


@pytest.mark.parametrize("dim", [2, 3])
def test_hll_mirror_symmetry(dim):
    qL = O.complete_from_cons(random_states(200, dim, seed=23), dim)
    qR = O.complete_from_cons(random_states(200, dim, seed=24), dim)
    for d in range(dim):
        f = _hll(qL, qR, dim, d)
        mL, mR = qL.copy(), qR.copy()
        mL[1 + d] *= -1
        mR[1 + d] *= -1
        g = _hll(mR, mL, dim, d)
        expect = -f.copy()
        expect[1 + d] = f[1 + d]
        assert np.max(np.abs(g - expect) / np.maximum(np.abs(f), 1.0)) < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_hll_stable_dt_uniform_state(dim):
    """L = R = q: the Einfeldt Roe sound speed uses only the normal velocity,
    roeA^2 = a^2 + (gamma - 1) / 2 |u_t|^2, and dt = dx / (|u_n| + roeA)."""
    q = O.complete_from_cons(random_states(40, dim, seed=25), dim)
    dx = 0.01
    for d in range(dim):
        dt = _hll_dt(q, q, dim, d, dx)
        u = q[1:1 + dim] / q[0]
        ut2 = np.sum(u ** 2, axis=0) - u[d] ** 2
        roeA = np.sqrt(q[dim + 3] ** 2 + 0.5 * (GAMMA - 1.0) * ut2)
        expect = dx / (np.abs(u[d]) + roeA)
        assert np.max(np.abs(dt - expect) / expect) < 1e-13
