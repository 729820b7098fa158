"""Oracle-side composition of one split sweep on a ghost-free transmissive
patch, built from the oracle's patch-level operators exactly as one split
step of DimensionalSplitLevelIntegrator runs them: fill ghosts, CFL check,
ComputeNumericFluxes, UpdateConservatively."""
from __future__ import annotations

import ctypes as C

import numpy as np

import oracle_lib as O
from finitevolumesolver_amd import _lib as L


def padded_complete(cons: np.ndarray, D: int, gw: int, gamma: float = 1.4):
    nz, ny, nx = cons.shape[1:]
    g = [gw, gw, gw if D == 3 else 0]
    comp = np.zeros((D + 4, nz + 2 * g[2], ny + 2 * g[1], nx + 2 * g[0]))
    lo = (-g[0], -g[1], -g[2])
    inner = comp[:, g[2]:g[2] + nz, g[1]:g[1] + ny, g[0]:g[0] + nx]
    inner[:D + 2] = cons
    cv = O.host_view(comp, lo)
    iv = O.host_view(inner, (0, 0, 0))
    # complete the interior from its own conservative variables
    O.load().oracle_complete_view(D, gamma, C.byref(iv), C.byref(iv))
    dlo = L.i64x3((0, 0, 0))
    dhi = L.i64x3((nx - 1, ny - 1, nz - 1))
    for d in range(D):
        O.load().oracle_transmissive_fill(C.byref(cv), dlo, dhi, d)
    return comp, g


def oracle_sweep(desc: L.MethodDesc, cons: np.ndarray, dt: float, dx: float, direction: int):
    D = desc.dim
    # Two stencil widths of ghosts, so that the sequential driver's CFL face
    # range (which skips the box's last face) still covers every face.
    gw = 2 * (2 if desc.reconstruction == L.RECON_MUSCL_HANCOCK else 1)
    comp, g = padded_complete(cons, D, gw, desc.gamma)
    nz, ny, nx = cons.shape[1:]
    n = [nx, ny, nz]
    # states: interior grown by gw in the sweep direction
    sl = [slice(g[d], g[d] + n[d]) for d in range(3)]
    sl[direction] = slice(g[direction] - gw, g[direction] + n[direction] + gw)
    lo = [0, 0, 0]
    lo[direction] = -gw
    states = comp[:, sl[2], sl[1], sl[0]]
    sv = O.host_view(states, lo)
    fshape = [nx, ny, nz]
    fshape[direction] += 1
    flux = np.zeros((D + 2, fshape[2], fshape[1], fshape[0]))
    fv = O.host_view(flux, (0, 0, 0))
    rc = O.load().oracle_numeric_fluxes(C.byref(desc), C.byref(fv), C.byref(sv), dt, dx, direction)
    assert rc == 0
    dtc = C.c_double()
    O.load().oracle_stable_dt(C.byref(desc), C.byref(sv), dx, direction, C.byref(dtc))
    out = np.ascontiguousarray(cons.copy())
    ov = O.host_view(out, (0, 0, 0))
    O.load().oracle_update_conservatively(C.byref(ov), C.byref(ov), C.byref(fv), dt, dx, direction)
    return out, dtc.value, flux
