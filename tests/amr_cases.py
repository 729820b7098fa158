"""Shared cases of the AMR hierarchy tests: a small three-level hierarchy
and the per-patch oracle sweep the AMR oracle composes."""
from __future__ import annotations

import ctypes as C

import numpy as np

import oracle_lib as O

# Level 0: two patches covering a 32 x 32 domain on [-1, 1]^2.  Level 1
# (64 x 64): patches touching along x = 32, one isolated, one at the
# domain corner.  Level 2 (128 x 128): properly nested patches, one pair
# sharing a face (x = 60), one spanning two level-1 patches, one on the
# domain edge y = 0.
DOMAIN = (32, 32)
DX0 = 2.0 / 32
LAYOUTS = [
    [(0, 0, 16, 32), (16, 0, 16, 32)],
    [(16, 16, 16, 24), (32, 16, 16, 16), (8, 44, 24, 16), (48, 0, 16, 16)],
    [(36, 36, 24, 40), (60, 36, 16, 24), (100, 0, 24, 16)],
]


def oracle_sweep(desc):
    """sweep(cons, dt, dx, d, grow=0) of one 2D patch with two ghost layers:
    the oracle's FluxMethod + UpdateConservatively over the interior rows
    grown by `grow` ghost rows on each side across d
    (src/HyperbolicPatchIntegrator.cpp:35-154), the face fluxes of the
    interior rows and the minimum HLLEM stable dt of their faces 0..n."""
    lib = O.load()

    def sweep(cons, dt, dx, d, grow=0):
        ny, nx = cons.shape[1] - 4, cons.shape[2] - 4
        comp = O.complete_from_cons(np.ascontiguousarray(cons).reshape(4, -1), 2)
        comp = comp.reshape(6, 1, ny + 4, nx + 4)
        sv = O.host_view(np.ascontiguousarray(comp), (-2, -2, 0))
        gy, gx = (grow, 0) if d == 0 else (0, grow)
        fg = np.zeros((4, 1, ny + 2 * gy + (d == 1), nx + 2 * gx + (d == 0)))
        lib.oracle_numeric_fluxes(C.byref(desc), C.byref(O.host_view(fg, (-gx, -gy, 0))),
                                  C.byref(sv), dt, dx, d)
        prev = np.ascontiguousarray(
            cons[:, 2 - gy:2 + ny + gy, 2 - gx:2 + nx + gx].reshape(4, 1, ny + 2 * gy, nx + 2 * gx))
        nxt = prev.copy()
        lib.oracle_update_conservatively(C.byref(O.host_view(nxt, (-gx, -gy, 0))),
                                         C.byref(O.host_view(prev, (-gx, -gy, 0))),
                                         C.byref(O.host_view(fg, (-gx, -gy, 0))), dt, dx, d)
        f = fg[:, :, gy:gy + ny + (d == 1), gx:gx + nx + (d == 0)]
        c = comp[:, 0, 2:-2, :] if d == 0 else comp[:, 0, :, 2:-2]
        if d == 0:
            L, R = c[:, :, 1:nx + 2], c[:, :, 2:nx + 3]
        else:
            L, R = c[:, 1:ny + 2, :], c[:, 2:ny + 3, :]
        dts = O.hllem_stable_dt(np.ascontiguousarray(L.reshape(6, -1)),
                                np.ascontiguousarray(R.reshape(6, -1)), 2, d, dx)
        return nxt[:, 0], f[:, 0], float(dts.min())

    return sweep


def centres(layout_patch, level, ratio=2):
    i0, j0, nx, ny = layout_patch
    dx = DX0 / ratio ** level
    x = -1.0 + (i0 + np.arange(nx) + 0.5) * dx
    y = -1.0 + (j0 + np.arange(ny) + 0.5) * dx
    return np.meshgrid(x, y)


def cons_from(rho, u, v, p, gamma=1.4):
    return np.stack([rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)])


def moving_state(level, p, layouts=LAYOUTS):
    """Smooth blob with a swirling velocity field."""
    x, y = centres(layouts[level][p], level)
    r2 = (x - 0.1) ** 2 + (y + 0.05) ** 2
    rho = 1.0 + 0.4 * np.exp(-r2 / 0.08)
    u = 0.3 * np.sin(np.pi * x) * np.cos(np.pi * y)
    v = -0.2 * np.cos(np.pi * x) * np.sin(np.pi * y)
    p_ = 1.0 + 0.5 * np.exp(-r2 / 0.05)
    return cons_from(rho, u, v, p_)


def resting_blob(level, p, layouts=LAYOUTS):
    """Compactly supported density and pressure bump at rest (radius 0.5):
    no flux reaches the domain edges within a few steps, so the composite
    integrals of mass and energy are invariants of the hierarchy step."""
    x, y = centres(layouts[level][p], level)
    s = np.maximum(0.0, 1.0 - (x * x + y * y) / 0.25) ** 2
    rho = 1.0 + 0.5 * s
    p_ = 1.0 + 0.8 * s
    z = np.zeros_like(x)
    return cons_from(rho, z, z, p_)


def shock_bubble(level, p, layouts=LAYOUTS):
    """The shock-bubble state of the batched-AMR workload on the small
    hierarchy: a Mach 1.5 shock at x = -0.5 (on the left face of the first
    level-1 patch, so it crosses coarse-fine boundaries in the first steps)
    moving into rho = 1, p = 1, and a light bubble (r = 0.2, rho = 0.138)
    at the origin, inside level 2."""
    from finitevolumesolver_amd.initial_data import shock_bubble_state
    x, y = centres(layouts[level][p], level)
    rho, u, v, p_ = shock_bubble_state(x, y)
    return cons_from(rho, u, v, p_)


def oracle_run(o, steps, cfl, max_restarts=64, restore_cycle=True, final_time=float("inf"),
               output_times=(), eps=1e-12):
    """RunSimulation's loop (include/fub/RunSimulation.hpp:86-191) over the
    AMR oracle: dt = cfl * stable dt, one coarse step, each level's sweep
    minima checked against the level's own time step; a failed attempt is
    undone (state and, as the reference restores its whole solver from the
    backup at :151, the cycle count) and retried with cfl times the stable dt
    the sweeps found.  Returns (accepted dts, attempts per step)."""
    import copy
    dts, attempts = [], []
    t = 0.0
    outs = sorted(output_times)

    def limited(dt):
        cap = final_time - t
        if outs:
            cap = min(cap, outs[0] - t)
        return min(cap, dt)

    for _ in range(steps):
        if not t + eps < final_time:
            break
        dt = limited(cfl * o.stable_dt())
        saved, cycle = copy.deepcopy(o.U), o.cycle
        for attempt in range(max_restarts + 1):
            o.advance(dt)
            got = o.level_cfl
            if all(got[l] >= dt / o.r ** l for l in range(len(o.U))):
                break
            limit = min(got[l] * o.r ** l for l in range(len(o.U)))
            o.U = copy.deepcopy(saved)
            if restore_cycle:
                o.cycle = cycle
            dt = limited(cfl * limit)
        else:
            raise RuntimeError("too many restarts")
        dts.append(dt)
        attempts.append(attempt + 1)
        t += dt
        while outs and t + eps >= outs[0]:
            outs.pop(0)
    return dts, attempts
