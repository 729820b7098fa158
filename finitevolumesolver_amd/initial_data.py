"""Initial data of the benchmark workloads (host, numpy FP64).

Random numbers come from SplitMix64 (Steele, Lea & Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014), normals from Box-Muller, so
the same seed gives the same field on every machine.  States are returned as
conservative SoA arrays of shape (D+2, nz, ny, nx); pressure enters through
E = p / (gamma - 1) + 0.5 |rho u|^2 / rho (PerfectGas::CompleteFromPrim,
src/equations/PerfectGas.cpp:146-158).
"""
from __future__ import annotations

import math

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.uint64)
        s = self.state
        for i in range(n):
            s = (s + 0x9E3779B97F4A7C15) & MASK64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
            out[i] = z ^ (z >> 31)
        self.state = s
        return out

    def next_u64_fast(self, n: int) -> np.ndarray:
        """Vectorised stream: element i is the (i+1)-th output of next_u64."""
        with np.errstate(over="ignore"):
            idx = np.arange(1, n + 1, dtype=np.uint64)
            s = np.uint64(self.state) + idx * np.uint64(0x9E3779B97F4A7C15)
            z = s
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = int(s[-1]) if n else self.state
        return z

    def uniform(self, n: int) -> np.ndarray:
        """Uniform doubles in [0, 1) from the top 53 bits."""
        return (self.next_u64_fast(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(1.0 - u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


def cons_from_prim(rho, vel, p, gamma=1.4) -> np.ndarray:
    """rho, p: arrays of the grid shape; vel: list of D arrays."""
    D = len(vel)
    gm1inv = 1.0 / (gamma - 1.0)
    mom = [rho * v for v in vel]
    sq = mom[0] * mom[0]
    if D == 2:
        sq = mom[0] * mom[0] + mom[1] * mom[1]
    elif D == 3:
        sq = (mom[0] * mom[0] + mom[1] * mom[1]) + mom[2] * mom[2]
    e_kin = 0.5 * sq / rho
    E = e_kin + p * gm1inv
    out = np.stack([rho, *mom, E])
    if D == 2:
        out = out[:, None, :, :]
    return np.ascontiguousarray(out)


def cell_centers(n: int, lo: float, hi: float) -> np.ndarray:
    dx = (hi - lo) / n
    return lo + (np.arange(n) + 0.5) * dx


def circular_riemann_2d(n: int = 256, gamma: float = 1.4):
    """2D circular Riemann problem on [-1, 1]^2: r < 0.4 rho=1, p=1, else
    rho=0.125, p=0.1, fluid at rest.  Returns (cons, dx)."""
    x = cell_centers(n, -1.0, 1.0)
    X, Y = np.meshgrid(x, x, indexing="xy")  # Y rows, X columns -> [ny, nx]
    r = np.sqrt(X * X + Y * Y)
    inside = r < 0.4
    rho = np.where(inside, 1.0, 0.125)
    p = np.where(inside, 1.0, 0.1)
    zero = np.zeros_like(rho)
    dx = 2.0 / n
    return cons_from_prim(rho, [zero, zero], p, gamma), (dx, dx)


def _fourier_field(coords, seed: int, n_modes: int = 8, kmax: int = 4) -> np.ndarray:
    """sum_m a_m cos(2 pi k_m . x + phi_m), sum |a_m| = 1, 1 <= |k|_inf <= kmax."""
    rng = SplitMix64(seed)
    D = len(coords)
    field = np.zeros(np.broadcast_shapes(*[c.shape for c in coords]))
    ks, amps, phis = [], [], []
    while len(ks) < n_modes:
        k = (rng.uniform(D) * (2 * kmax + 1)).astype(np.int64) - kmax
        if np.max(np.abs(k)) < 1:
            continue
        ks.append(k)
        amps.append(rng.uniform(1)[0] * 2.0 - 1.0)
        phis.append(rng.uniform(1)[0] * 2.0 * np.pi)
    amps = np.asarray(amps)
    amps = amps / np.sum(np.abs(amps))
    for k, a, ph in zip(ks, amps, phis):
        arg = ph + 2.0 * np.pi * sum(float(k[d]) * coords[d] for d in range(D))
        field = field + a * np.cos(arg)
    return field


def _box_smooth(f: np.ndarray, passes: int = 3) -> np.ndarray:
    """5-point box filter per axis (periodic wrap), `passes` times."""
    for _ in range(passes):
        for ax in range(f.ndim):
            f = (np.roll(f, -2, ax) + np.roll(f, -1, ax) + f + np.roll(f, 1, ax)
                 + np.roll(f, 2, ax)) / 5.0
    return f


def fourier_modes_3d(n: int = 256, gamma: float = 1.4, seed0: int = 1):
    """3D smooth random field on [0, 1]^3: rho = 1 + 0.2 F_1, p = 1 + 0.1 F_2,
    velocities N(0, 0.1) smoothed by 3 box-filter passes per axis.
    Returns (cons (5, n, n, n), dx)."""
    x = cell_centers(n, 0.0, 1.0)
    X = x[None, None, :]
    Y = x[None, :, None]
    Z = x[:, None, None]
    rho = 1.0 + 0.2 * _fourier_field([X, Y, Z], seed0)
    p = 1.0 + 0.1 * _fourier_field([X, Y, Z], seed0 + 1)
    rho = np.broadcast_to(rho, (n, n, n)).copy()
    p = np.broadcast_to(p, (n, n, n)).copy()
    vel = []
    for s in range(3):
        v = SplitMix64(seed0 + 2 + s).normal(n ** 3).reshape(n, n, n) * 0.1
        vel.append(_box_smooth(v))
    dx = 1.0 / n
    return cons_from_prim(rho, vel, p, gamma), (dx, dx, dx)


def random_states(n: int, dim: int, seed: int, gamma: float = 1.4,
                  rho_range=(0.1, 10.0), p_range=(0.1, 10.0), u_scale=1.0):
    """n random conservative states (D+2, n) with log-uniform rho, p."""
    rng = SplitMix64(seed)
    lr = np.log(rho_range)
    lp = np.log(p_range)
    rho = np.exp(lr[0] + (lr[1] - lr[0]) * rng.uniform(n))
    p = np.exp(lp[0] + (lp[1] - lp[0]) * rng.uniform(n))
    vel = [u_scale * (2.0 * rng.uniform(n) - 1.0) for _ in range(dim)]
    gm1inv = 1.0 / (gamma - 1.0)
    mom = [rho * v for v in vel]
    if dim == 2:
        sq = mom[0] * mom[0] + mom[1] * mom[1]
    else:
        sq = (mom[0] * mom[0] + mom[1] * mom[1]) + mom[2] * mom[2]
    E = 0.5 * sq / rho + p * gm1inv
    return np.stack([rho, *mom, E])


def shock_channel(nx: int, ny: int, mach: float = 2.0, x_shock: float = -1.0,
                  gamma: float = 1.4):
    """Channel [-1.5, 2.5] x [-1, 1]: a Mach `mach` shock at x_shock moving
    into gas at rest with rho = 1.4, p = 1 (a = 1); post-shock state from
    the Rankine-Hugoniot relations (the reference's ShockMachnumber,
    include/fub/initial_data/ShockMachnumber.hpp:41-65).  Returns cons
    (4, ny, nx) and dx."""
    dx = 4.0 / nx
    rho1, p1, u1 = 1.4, 1.0, 0.0
    a1 = math.sqrt(gamma * p1 / rho1)
    m2 = mach * mach
    p2 = p1 * (1.0 + 2.0 * gamma / (gamma + 1.0) * (m2 - 1.0))
    rho2 = rho1 * (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    u2 = mach * a1 * (1.0 - rho1 / rho2)
    x = -1.5 + (np.arange(nx) + 0.5) * dx
    post = np.broadcast_to(x[None, :] < x_shock, (ny, nx))
    rho = np.where(post, rho2, rho1)
    u = np.where(post, u2, u1)
    p = np.where(post, p2, p1)
    cons = cons_from_prim(rho, [u, np.zeros_like(u)], p, gamma)
    return cons, dx


def _splitmix_uniform_device(seed: int, n: int, torch, device):
    """The SplitMix64 stream of `seed` (as SplitMix64(seed).uniform(n)),
    generated on the device with wrapping int64 arithmetic."""
    def srl(z, s):  # logical right shift of int64 bit patterns
        return (z >> s) & ((1 << (64 - s)) - 1)

    def wrap(v):  # a uint64 constant as its int64 bit pattern
        return v - (1 << 64) if v >= (1 << 63) else v

    golden = wrap(0x9E3779B97F4A7C15)
    idx = torch.arange(1, n + 1, dtype=torch.int64, device=device)
    z = wrap(seed & MASK64) + idx * golden
    z = (z ^ srl(z, 30)) * wrap(0xBF58476D1CE4E5B9)
    z = (z ^ srl(z, 27)) * wrap(0x94D049BB133111EB)
    z = z ^ srl(z, 31)
    return srl(z, 11).to(torch.float64) * (1.0 / (1 << 53))


def fourier_modes_3d_device(n: int, seed0: int = 1, gamma: float = 1.4, device="cuda"):
    """fourier_modes_3d generated on the device (torch FP64) for large
    benchmark grids: the same mode draws and SplitMix64 streams; values
    agree with the host version to rounding.  Returns (cons (5, n, n, n)
    device tensor, dx)."""
    import torch
    x = torch.from_numpy(cell_centers(n, 0.0, 1.0)).to(device)
    X, Y, Z = x.view(1, 1, n), x.view(1, n, 1), x.view(n, 1, 1)

    def field(seed):
        rng = SplitMix64(seed)
        ks, amps, phis = [], [], []
        while len(ks) < 8:
            k = (rng.uniform(3) * 9).astype(np.int64) - 4
            if np.max(np.abs(k)) < 1:
                continue
            ks.append(k)
            amps.append(rng.uniform(1)[0] * 2.0 - 1.0)
            phis.append(rng.uniform(1)[0] * 2.0 * np.pi)
        amps = np.asarray(amps)
        amps = amps / np.sum(np.abs(amps))
        f = torch.zeros((n, n, n), dtype=torch.float64, device=device)
        for k, a, ph in zip(ks, amps, phis):
            arg = ph + 2.0 * np.pi * (float(k[0]) * X + float(k[1]) * Y + float(k[2]) * Z)
            f += float(a) * torch.cos(arg)
        return f

    rho = 1.0 + 0.2 * field(seed0)
    p = 1.0 + 0.1 * field(seed0 + 1)
    vel = []
    m = (n ** 3 + 1) // 2
    for s in range(3):
        u = _splitmix_uniform_device(seed0 + 2 + s, 2 * m, torch, device)
        u1, u2 = u[:m], u[m:]
        r = torch.sqrt(-2.0 * torch.log(1.0 - u1))
        v = torch.cat([r * torch.cos(2 * np.pi * u2), r * torch.sin(2 * np.pi * u2)])[:n ** 3]
        v = v.view(n, n, n) * 0.1
        for _ in range(3):
            for ax in range(3):
                v = (torch.roll(v, -2, ax) + torch.roll(v, -1, ax) + v + torch.roll(v, 1, ax)
                     + torch.roll(v, 2, ax)) / 5.0
        vel.append(v)
    gm1inv = 1.0 / (gamma - 1.0)
    mom = [rho * v for v in vel]
    sq = (mom[0] * mom[0] + mom[1] * mom[1]) + mom[2] * mom[2]
    E = 0.5 * sq / rho + p * gm1inv
    cons = torch.stack([rho, *mom, E]).contiguous()
    dx = 1.0 / n
    return cons, (dx, dx, dx)


def blocky_field_3d_device(n: int, seed: int = 5, block: int = 4, gamma: float = 1.4,
                           umax: float = 0.5, device="cuda"):
    """Strong-wave 3D field on the device: piecewise-constant blocks of
    block^3 cells with density and pressure log-uniform in [0.1, 10] and
    velocities uniform in [-umax, umax] (SplitMix64 streams per variable),
    so nearly every block face is a strong Riemann problem.  With umax = 0.5
    no face generates vacuum (|du| <= 1 < 2 (aL + aR) / (gamma - 1), the
    sound speeds being at least sqrt(1.4 * 0.1 / 10)).  Returns (cons
    (5, n, n, n) device tensor, dx)."""
    import torch
    nb = (n + block - 1) // block

    def field(s, lo, hi, log):
        v = _splitmix_uniform_device(seed * 16 + s, nb ** 3, torch, device).view(nb, nb, nb)
        v = torch.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * v) if log else lo + (hi - lo) * v
        v = v.repeat_interleave(block, 0).repeat_interleave(block, 1).repeat_interleave(block, 2)
        return v[:n, :n, :n].contiguous()

    rho, p = field(0, 0.1, 10.0, True), field(1, 0.1, 10.0, True)
    vel = [field(2 + d, -umax, umax, False) for d in range(3)]
    cons = torch.empty((5, n, n, n), dtype=torch.float64, device=device)
    cons[0] = rho
    for d in range(3):
        cons[1 + d] = rho * vel[d]
    cons[4] = p / (gamma - 1.0) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    return cons, (1.0 / n, 1.0 / n, 1.0 / n)


def amr_patch_layout(level: int, count: int = 300, base: int = 512, ratio: int = 2,
                     sides=(16, 24, 32, 48, 64), band=(-0.75, 0.35)):
    """Patches of one refined level of the 2D shock-bubble hierarchy: `count`
    patches with sides drawn independently per axis from `sides` (SplitMix64
    seed 10 + level), placed without overlap on an 8-aligned lattice by
    shelf packing inside the x band `band` of [-1, 1]^2 (continuing to the
    right of it when the band is full).  Returns a list of (i0, j0, nx, ny)
    in the level's cell indices and the level's cell width."""
    n = base * ratio ** level
    dx = 2.0 / n
    rng = SplitMix64(10 + level)
    u = rng.uniform(2 * count)
    sizes = [(sides[int(u[2 * p] * len(sides))], sides[int(u[2 * p + 1] * len(sides))])
             for p in range(count)]
    x0 = int((band[0] + 1.0) / dx) // 8 * 8
    x1 = int((band[1] + 1.0) / dx) // 8 * 8
    out = []
    cx, cy, shelf = x0, 0, 0
    for nx, ny in sizes:
        if cy + ny > n:
            raise ValueError("level too small for the patches")
        if cx + nx > x1 and cx > x0:
            cx = x0
            cy += shelf
            shelf = 0
        if cy + ny > n:  # band full: continue in the next band to the right
            x0, x1 = x1, min(n, x1 + (x1 - x0))
            cx, cy, shelf = x0, 0, 0
        out.append((cx, cy, nx, ny))
        cx += (nx + 7) // 8 * 8
        shelf = max(shelf, (ny + 7) // 8 * 8)
    return out, dx


def amr_nested_levels(base: int = 512, ratio: int = 2, count: int = 300, shrink: int = 4):
    """Properly nested three-level layout of the 2D shock-bubble hierarchy:
    level 0 one base^2 patch, level 1 the `count` seeded patches of
    amr_patch_layout(1), level 2 every level-1 patch shrunk by `shrink`
    level-1 cells per side and refined (kept when both sides stay >= 8), so
    that each level-2 patch's ghost layers and the level-1 cells beside it
    lie inside its parent.  Returns per level a list of (i0, j0, nx, ny)."""
    l1, _ = amr_patch_layout(1, count, base, ratio)
    l2 = []
    for i0, j0, nx, ny in l1:
        a, b = nx - 2 * shrink, ny - 2 * shrink
        if a >= 8 and b >= 8:
            l2.append(((i0 + shrink) * ratio, (j0 + shrink) * ratio, a * ratio, b * ratio))
    return [[(0, 0, base, base)], l1, l2]


def shock_bubble_state(x, y, mach: float = 1.5, x_shock: float = -0.5, r_bubble: float = 0.2,
                       rho_bubble: float = 0.138, gamma: float = 1.4):
    """Primitive state (rho, u, v, p) of the 2D shock-bubble problem at points
    x, y: a Mach `mach` shock at x_shock moving right into rho = 1, p = 1
    (Rankine-Hugoniot post-shock state behind it), a light bubble of radius
    r_bubble at the origin."""
    m2 = mach * mach
    a1 = math.sqrt(gamma)
    p2 = 1.0 + 2.0 * gamma / (gamma + 1.0) * (m2 - 1.0)
    rho2 = (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    u2 = mach * a1 * (1.0 - 1.0 / rho2)
    post = x < x_shock
    bubble = (x * x + y * y) < r_bubble * r_bubble
    rho = np.where(post, rho2, np.where(bubble, rho_bubble, 1.0))
    u = np.where(post, u2, 0.0)
    p = np.where(post, p2, 1.0)
    return rho, u, np.zeros_like(u), p
