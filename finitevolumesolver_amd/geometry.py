"""Analytic embedded-boundary geometry for cut-cell patches (host, numpy).

Produces the arrays of CutCellData<2> (include/fub/CutCellData.hpp:48-75)
for bodies made of disks and convex polygons, with the fluid outside the
bodies: volume fractions alpha (exact areas), face fractions beta (exact
intercepts), boundary normals and boundary centroids.  The reference takes
these moments from AMReX's EB2 database (src/AMReX/cutcell/PatchHierarchy.cpp:507-545);
here they are computed in closed form:

  * disk / rectangle area: integral of the clipped chord height with the
    antiderivative of sqrt(r^2 - x^2), split at the clip breakpoints;
  * polygon / rectangle area: Sutherland-Hodgman clipping and the shoelace
    formula;
  * boundary centroid: midpoint of the boundary segment (chord) inside the
    cell, relative to the cell centre in units of the cell width;
  * boundary normal: AMReX orientation, pointing into the body (the cut-cell
    method negates it, src/CutCellData.cpp:90-105).

Bodies must not share a cell.  Arrays are indexed [j, i] (y, x) over the
cell box grown by `ghosts`; face arrays have one more entry along their
direction.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Disk:
    cx: float
    cy: float
    r: float


@dataclass
class Polygon:
    """Convex polygon, vertices counter-clockwise."""
    vertices: list


@dataclass
class Geometry2D:
    """CutCellData<2> arrays of a patch with `ghosts` ghost cells.

    lo: index of the first (ghost) cell in each direction, (ix, iy).
    alpha [ny, nx], beta_x [ny, nx + 1], beta_y [ny + 1, nx],
    normals [2, ny, nx], centroids [2, ny, nx].
    """
    lo: tuple
    alpha: np.ndarray
    beta_x: np.ndarray
    beta_y: np.ndarray
    normals: np.ndarray
    centroids: np.ndarray
    meta: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.alpha.shape


# ---- disk -------------------------------------------------------------------------

def _F(x, r):
    """Antiderivative of sqrt(r^2 - x^2) on [-r, r]."""
    x = min(max(x, -r), r)
    return 0.5 * (x * math.sqrt(max(r * r - x * x, 0.0)) + r * r * math.asin(x / r))


def _disk_rect_area(x0, x1, y0, y1, r):
    """Area of the disk of radius r at the origin inside [x0, x1] x [y0, y1]."""
    a, b = max(x0, -r), min(x1, r)
    if a >= b:
        return 0.0
    # breakpoints where the upper/lower half circle crosses y0 or y1
    pts = {a, b}
    for y in (y0, y1):
        if abs(y) < r:
            h = math.sqrt(r * r - y * y)
            for x in (-h, h):
                if a < x < b:
                    pts.add(x)
    pts = sorted(pts)
    area = 0.0
    for u, v in zip(pts[:-1], pts[1:]):
        m = 0.5 * (u + v)
        hm = math.sqrt(max(r * r - m * m, 0.0))
        # upper curve clip(h, y0, y1), lower curve clip(-h, y0, y1) on (u, v)
        def integral(sign):
            val = sign * hm
            if val >= y1:
                return y1 * (v - u)
            if val <= y0:
                return y0 * (v - u)
            return sign * (_F(v, r) - _F(u, r))
        area += integral(+1.0) - integral(-1.0)
    return max(area, 0.0)


def _disk_segment_cover(a, b, r, c):
    """Length of the segment {c} x [a, b] (coordinate c across) inside the disk."""
    if abs(c) >= r:
        return 0.0
    h = math.sqrt(r * r - c * c)
    return max(0.0, min(b, h) - max(a, -h))


def _disk_chord(x0, x1, y0, y1, r):
    """Intersection points of the circle with the boundary of the cell."""
    pts = []
    for x in (x0, x1):
        if abs(x) < r:
            h = math.sqrt(r * r - x * x)
            for y in (-h, h):
                if y0 <= y <= y1:
                    pts.append((x, y))
    for y in (y0, y1):
        if abs(y) < r:
            h = math.sqrt(r * r - y * y)
            for x in (-h, h):
                if x0 < x < x1:
                    pts.append((x, y))
    return pts


# ---- convex polygon ------------------------------------------------------------------

def _clip(poly, a, b, c):
    """Sutherland-Hodgman: keep the part of poly with a x + b y <= c."""
    out = []
    n = len(poly)
    for k in range(n):
        p, q = poly[k], poly[(k + 1) % n]
        fp = a * p[0] + b * p[1] - c
        fq = a * q[0] + b * q[1] - c
        if fp <= 0:
            out.append(p)
        if (fp < 0 < fq) or (fq < 0 < fp):
            t = fp / (fp - fq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    return out


def _area(poly):
    if len(poly) < 3:
        return 0.0
    s = 0.0
    for k in range(len(poly)):
        p, q = poly[k], poly[(k + 1) % len(poly)]
        s += p[0] * q[1] - q[0] * p[1]
    return abs(0.5 * s)


def _edges(vertices):
    n = len(vertices)
    for k in range(n):
        p, q = vertices[k], vertices[(k + 1) % n]
        ex, ey = q[0] - p[0], q[1] - p[1]
        ln = math.hypot(ex, ey)
        # outward normal of a counter-clockwise polygon
        yield p, q, (ey / ln, -ex / ln)


def _poly_cell(vertices, x0, x1, y0, y1):
    """Cell rectangle clipped to the polygon."""
    rect = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
    for p, q, (nx, ny) in _edges(vertices):
        rect = _clip(rect, nx, ny, nx * p[0] + ny * p[1])
        if not rect:
            break
    return rect


def _poly_segment_cover(vertices, p0, p1):
    """Length of segment p0-p1 inside the convex polygon."""
    t0, t1 = 0.0, 1.0
    dx, dy = p1[0] - p0[0], p1[1] - p0[1]
    for p, q, (nx, ny) in _edges(vertices):
        num = nx * p[0] + ny * p[1] - (nx * p0[0] + ny * p0[1])
        den = nx * dx + ny * dy
        if den == 0.0:
            if num < 0:
                return 0.0
            continue
        t = num / den
        if den > 0:
            t1 = min(t1, t)
        else:
            t0 = max(t0, t)
        if t0 >= t1:
            return 0.0
    return (t1 - t0) * math.hypot(dx, dy)


def _poly_boundary(vertices, x0, x1, y0, y1):
    """Boundary segment of the polygon inside the cell: (midpoint, outward normal)."""
    best = None
    for p, q, n in _edges(vertices):
        seg = [p, q]
        for a, b, c in ((-1, 0, -x0), (1, 0, x1), (0, -1, -y0), (0, 1, y1)):
            seg = _clip_segment(seg, a, b, c)
            if seg is None:
                break
        if seg is None:
            continue
        ln = math.hypot(seg[1][0] - seg[0][0], seg[1][1] - seg[0][1])
        if best is None or ln > best[0]:
            mid = (0.5 * (seg[0][0] + seg[1][0]), 0.5 * (seg[0][1] + seg[1][1]))
            best = (ln, mid, n)
    return best


def _clip_segment(seg, a, b, c):
    p, q = seg
    fp = a * p[0] + b * p[1] - c
    fq = a * q[0] + b * q[1] - c
    if fp > 0 and fq > 0:
        return None
    if fp <= 0 and fq <= 0:
        return [p, q]
    t = fp / (fp - fq)
    m = (p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1]))
    return [p, m] if fp <= 0 else [m, q]


# ---- patch geometry -------------------------------------------------------------------

def make_geometry(bodies, n, xlo, dx, ghosts=2):
    """Geometry of the cells [0, n) grown by `ghosts`, cell (i, j) spanning
    [xlo[0] + i dx, ...) x [xlo[1] + j dx, ...) (square cells)."""
    nx, ny = n
    gx, gy = nx + 2 * ghosts, ny + 2 * ghosts
    alpha = np.ones((gy, gx))
    beta_x = np.ones((gy, gx + 1))
    beta_y = np.ones((gy + 1, gx))
    normals = np.zeros((2, gy, gx))
    centroids = np.zeros((2, gy, gx))
    xs = xlo[0] + (np.arange(gx + 1) - ghosts) * dx
    ys = xlo[1] + (np.arange(gy + 1) - ghosts) * dx
    A = dx * dx
    for body in bodies:
        if isinstance(body, Disk):
            bx0, bx1 = body.cx - body.r, body.cx + body.r
            by0, by1 = body.cy - body.r, body.cy + body.r
        else:
            vx = [v[0] for v in body.vertices]
            vy = [v[1] for v in body.vertices]
            bx0, bx1, by0, by1 = min(vx), max(vx), min(vy), max(vy)
        i0 = max(int(math.floor((bx0 - xs[0]) / dx)) - 1, 0)
        i1 = min(int(math.ceil((bx1 - xs[0]) / dx)) + 1, gx)
        j0 = max(int(math.floor((by0 - ys[0]) / dx)) - 1, 0)
        j1 = min(int(math.ceil((by1 - ys[0]) / dx)) + 1, gy)
        for j in range(j0, j1):
            for i in range(i0, i1):
                x0, x1, y0, y1 = xs[i], xs[i + 1], ys[j], ys[j + 1]
                if isinstance(body, Disk):
                    cov = _disk_rect_area(x0 - body.cx, x1 - body.cx, y0 - body.cy,
                                          y1 - body.cy, body.r)
                else:
                    cov = _area(_poly_cell(body.vertices, x0, x1, y0, y1))
                frac = cov / A
                if frac <= 0.0:
                    continue
                alpha[j, i] -= frac
                if alpha[j, i] < 1e-14:
                    alpha[j, i] = 0.0
                xc, yc = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
                if 0.0 < alpha[j, i] < 1.0:
                    if isinstance(body, Disk):
                        pts = _disk_chord(x0 - body.cx, x1 - body.cx, y0 - body.cy,
                                          y1 - body.cy, body.r)
                        if len(pts) >= 2:
                            mx = 0.5 * (pts[0][0] + pts[-1][0]) + body.cx
                            my = 0.5 * (pts[0][1] + pts[-1][1]) + body.cy
                        else:
                            mx, my = xc, yc
                        nrm = math.hypot(mx - body.cx, my - body.cy) or 1.0
                        # into the body
                        normals[0, j, i] = -(mx - body.cx) / nrm
                        normals[1, j, i] = -(my - body.cy) / nrm
                    else:
                        hit = _poly_boundary(body.vertices, x0, x1, y0, y1)
                        if hit is None:
                            mx, my, out = xc, yc, (0.0, 0.0)
                        else:
                            (_, (mx, my), out) = hit
                        normals[0, j, i] = -out[0]
                        normals[1, j, i] = -out[1]
                    centroids[0, j, i] = (mx - xc) / dx
                    centroids[1, j, i] = (my - yc) / dx
        # faces
        for j in range(j0, j1):
            for i in range(i0, min(i1 + 1, gx + 1)):
                x, y0, y1 = xs[i], ys[j], ys[j + 1]
                if isinstance(body, Disk):
                    cov = _disk_segment_cover(y0 - body.cy, y1 - body.cy, body.r, x - body.cx)
                else:
                    cov = _poly_segment_cover(body.vertices, (x, y0), (x, y1))
                beta_x[j, i] = max(beta_x[j, i] - cov / dx, 0.0)
        for j in range(j0, min(j1 + 1, gy + 1)):
            for i in range(i0, i1):
                y, x0, x1 = ys[j], xs[i], xs[i + 1]
                if isinstance(body, Disk):
                    cov = _disk_segment_cover(x0 - body.cx, x1 - body.cx, body.r, y - body.cy)
                else:
                    cov = _poly_segment_cover(body.vertices, (x0, y), (x1, y))
                beta_y[j, i] = max(beta_y[j, i] - cov / dx, 0.0)
    for b in (beta_x, beta_y):
        b[b < 1e-14] = 0.0
    return Geometry2D((-ghosts, -ghosts), alpha, beta_x, beta_y, normals, centroids,
                      {"dx": dx, "xlo": tuple(xlo), "n": (nx, ny), "ghosts": ghosts})


def perturb_small_cells(geom: Geometry2D, fraction=0.01, seed=7, lo=1e-4, hi=1e-2):
    """Shrinks a seeded `fraction` of the cut cells to volume fractions
    log-uniform in [lo, hi]; their face fractions are capped at
    sqrt(2 alpha) (a corner cut of that area), keeping alpha > 0 cells
    connected only through proportionally small faces."""
    from .initial_data import SplitMix64
    cut = np.argwhere((geom.alpha > 0.0) & (geom.alpha < 1.0))
    rng = SplitMix64(seed)
    u = rng.uniform(len(cut))
    v = rng.uniform(len(cut))
    pick = u < fraction
    a = np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * v)
    for (j, i), p, an in zip(cut, pick, a):
        if not p:
            continue
        geom.alpha[j, i] = an
        cap = math.sqrt(2.0 * an)
        geom.beta_x[j, i] = min(geom.beta_x[j, i], cap)
        geom.beta_x[j, i + 1] = min(geom.beta_x[j, i + 1], cap)
        geom.beta_y[j, i] = min(geom.beta_y[j, i], cap)
        geom.beta_y[j + 1, i] = min(geom.beta_y[j + 1, i], cap)
    geom.meta["perturbed"] = int(pick.sum())
    return geom


def channel_geometry(nx=2048, ny=1024, ghosts=2, perturb=True):
    """Cylinder (centre (0, 0), r = 0.25) and a 30-degree wedge in the
    channel [-1.5, 2.5] x [-1, 1]."""
    dx = 4.0 / nx
    t30 = math.tan(math.radians(30.0))
    wedge = Polygon([(0.9, -0.7), (1.6, -0.7), (1.6, -0.7 + 0.7 * t30)])
    geom = make_geometry([Disk(0.0, 0.0, 0.25), wedge], (nx, ny), (-1.5, -1.0), dx, ghosts)
    if perturb:
        perturb_small_cells(geom)
    return geom
