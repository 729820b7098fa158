"""Pins the CPU oracle of the KBN cut-cell method (no GPU).

Closed forms with exact binary fractions, invariants (fraction sums,
rotation and reflection identities, conservation of TimeIntegrator2),
special cases that reduce to the regular method (alpha = beta = 1), and
textbook Riemann facts (a gas at rest against a wall keeps its pressure).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O
from kbn_cases import circle_case, uncut_case
from finitevolumesolver_amd import _lib as L
from finitevolumesolver_amd.geometry import Disk, Polygon, make_geometry
from finitevolumesolver_amd.patch import PerfectGas, make_desc


def _desc(**kw):
    return make_desc(PerfectGas(dim=2), **kw)


# ---- geometry generator -------------------------------------------------------------

def test_disk_areas_sum_to_pi_r2():
    dx = 1.0 / 64
    g = make_geometry([Disk(0.503, 0.49, 0.3)], (64, 64), (0.0, 0.0), dx, 2)
    covered = np.sum(1.0 - g.alpha) * dx * dx
    assert covered == pytest.approx(math.pi * 0.09, rel=1e-12)
    # the covered length of every vertical grid line equals its chord
    for i in range(g.beta_x.shape[1]):
        x = (i - 2) * dx
        h = math.sqrt(max(0.09 - (x - 0.503) ** 2, 0.0))
        assert np.sum(1.0 - g.beta_x[:, i]) * dx == pytest.approx(2 * h, abs=1e-12)


def test_polygon_areas_and_normals():
    dx = 1.0 / 16
    tri = Polygon([(0.25, 0.25), (0.75, 0.25), (0.75, 0.5)])
    g = make_geometry([tri], (16, 16), (0.0, 0.0), dx, 2)
    assert np.sum(1.0 - g.alpha) * dx * dx == pytest.approx(0.5 * 0.5 * 0.25, rel=1e-12)
    cut = (g.alpha > 0) & (g.alpha < 1)
    # normals are unit vectors pointing into the body (AMReX orientation)
    nrm = np.hypot(g.normals[0][cut], g.normals[1][cut])
    np.testing.assert_allclose(nrm, 1.0, rtol=1e-12)


# ---- FillCutCellData ------------------------------------------------------------------

def _fill_one(beta_row):
    """FillCutCellData on a 1-row x geometry with the given face fractions."""
    nf = len(beta_row)
    geom = make_geometry([], (nf - 1 - 4, 1), (0.0, 0.0), 1.0, 2)
    geom.beta_x[:] = np.asarray(beta_row)[None, :]
    h = O.HostCutCellData(geom).fill()
    return h.fractions


def test_fill_cutcell_closed_forms():
    fr = _fill_one([1.0, 1.0, 0.5, 0.25, 0.0, 0.75, 1.0, 1.0])
    us, sl, sr, ds = (fr[k][0][:, 0, 0, :] for k in
                      ("unshielded", "shielded_left", "shielded_right", "doubly_shielded"))
    # face 2: betaL 1, beta 0.5, betaR 0.25 -> dL 0, dR 0.25
    assert (us[0, 2], sl[0, 2], sr[0, 2], ds[0, 2]) == (0.25, 0.0, 0.25, 0.0)
    assert (us[1, 2], sl[1, 2], sr[1, 2], ds[1, 2]) == (0.5, 0.0, 0.5, 0.0)
    # face 3: 0.5, 0.25, 0.0 -> dL 0, dR 0.25: fully shielded from the right
    assert (us[0, 3], sr[0, 3], sr[1, 3]) == (0.0, 0.25, 1.0)
    # face 5: 0.0, 0.75, 1.0 -> dL 0.75, dR 0
    assert (us[0, 5], sl[0, 5], sl[1, 5]) == (0.0, 0.75, 1.0)
    # face 4 (beta 0): relative fractions divide by 1
    assert us[1, 4] == 0.0 and ds[1, 4] == 0.0


def test_fill_cutcell_fractions_sum_to_beta():
    rng = np.random.default_rng(2)
    row = np.concatenate([[1.0], rng.integers(0, 5, 40) / 4.0, [1.0]])
    fr = _fill_one(row)
    tot = sum(fr[k][0][0, 0, 0, :] for k in fr)
    inner = slice(1, len(row) - 1)
    np.testing.assert_array_equal(tot[inner], row[inner])


# ---- IsCutCell -----------------------------------------------------------------------

def test_is_cut_cell_classification():
    geom = make_geometry([], (6, 3), (0.0, 0.0), 1.0, 2)
    a = geom.alpha
    a[2, 2], a[2, 3], a[2, 4] = 0.0, 0.5, 1.0
    geom.beta_x[2, 5] = 0.5      # right face of cell (i=4, j=2) partly blocked
    h = O.HostCutCellData(geom)
    s = h.c_struct()
    lo = np.array([-2, -2, 0], np.int64)
    hi = np.array([5, 4, 0], np.int64)
    flags = np.zeros((7, 8), np.int32)
    O.load().oracle_kbn_is_cut_cell(2, C.byref(s), O.ptr(lo), O.ptr(hi), O.ptr(flags))
    assert flags[2, 2] == 0   # covered
    assert flags[2, 3] == 1   # 0 < alpha < 1
    assert flags[2, 4] == 1   # alpha = 1 with a face fraction below 1
    assert flags[2, 5] == 1   # its right neighbour shares that face
    assert flags.sum() == 3


# ---- rotations and reflections --------------------------------------------------------

def test_rotation_maps_normal_onto_x():
    rng = np.random.default_rng(3)
    for _ in range(50):
        t = rng.uniform(0, 2 * np.pi)
        n = np.array([math.cos(t), math.sin(t)])
        np.testing.assert_allclose(O.kbn_rotate(n, [1.0, 0.0], n), [1.0, 0.0], atol=1e-15)
        m = rng.normal(size=2)
        r = O.kbn_rotate(n, [1.0, 0.0], m)
        assert np.linalg.norm(r) == pytest.approx(np.linalg.norm(m), rel=1e-15)
        back = O.kbn_rotate([1.0, 0.0], n, r)
        np.testing.assert_allclose(back, m, atol=1e-15)
    # opposite vectors: minus the identity
    np.testing.assert_array_equal(O.kbn_rotate([-1.0, 0.0], [1.0, 0.0], [0.3, -0.7]), [-0.3, 0.7])


def test_rotation_3d():
    rng = np.random.default_rng(4)
    for _ in range(30):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        np.testing.assert_allclose(O.kbn_rotate(n, [1.0, 0.0, 0.0], n), [1, 0, 0], atol=1e-14)
        m = rng.normal(size=3)
        r = O.kbn_rotate(n, [1.0, 0.0, 0.0], m)
        assert np.linalg.norm(r) == pytest.approx(np.linalg.norm(m), rel=1e-14)


# ---- HLLEM Riemann state (the boundary solver alternative) ----------------------------

def _complete(rho, u, v, p, gamma=1.4):
    q = np.array([[rho], [rho * u], [rho * v], [0.5 * rho * (u * u + v * v) + p / (gamma - 1)],
                  [p], [math.sqrt(gamma * p / rho)]])
    return q


def test_hllem_state_consistency_and_upwinding():
    q = _complete(1.3, 0.2, 0.0, 0.9)
    for larr in (False, True):
        s = O.hllem_riemann_state(q, q, 2, 0, larrouturou=larr)
        np.testing.assert_allclose(s, q, rtol=1e-14, atol=1e-15)
    # supersonic to the right: the left state
    L_ = _complete(1.0, 3.0, 0.4, 1.0)
    R_ = _complete(0.5, 3.2, -0.1, 0.8)
    s = O.hllem_riemann_state(L_, R_, 2, 0, larrouturou=False)
    np.testing.assert_array_equal(s, L_)
    # Larrouturou: the left state without its transverse kinetic energy
    s = O.hllem_riemann_state(L_, R_, 2, 0, larrouturou=True)
    assert s[2, 0] == 0.0
    assert s[3, 0] == pytest.approx(L_[3, 0] - 0.5 * 1.0 * 0.4 ** 2, rel=1e-15)


# ---- boundary problems -----------------------------------------------------------------

@pytest.mark.parametrize("solver", ["exact", "hllem"])
def test_gas_at_rest_against_the_wall(solver):
    """A cut cell at rest: the reference state is the state itself and the
    boundary flux is the pure wall pressure p e_dir."""
    geom, comp, lo = circle_case(perturb=False)
    comp[1:3] = 0.0
    comp[3] = comp[4] / 0.4  # E = p / (gamma - 1) at rest
    desc = _desc(boundary_solver=solver)
    h = O.HostCutCellData(geom)
    s = h.c_struct()
    refs = np.zeros_like(comp)
    O.load().oracle_kbn_reference_states(C.byref(desc), C.byref(O.host_view(refs, lo)),
                                         C.byref(O.host_view(comp, lo)), C.byref(s))
    cut = (geom.alpha > 0) & (geom.alpha < 1)
    np.testing.assert_allclose(refs[0, 0][cut], comp[0, 0][cut], rtol=1e-14)
    np.testing.assert_allclose(refs[4, 0][cut], comp[4, 0][cut], rtol=1e-14)
    assert np.max(np.abs(refs[1:3, 0][:, cut])) <= 1e-15
    for d in range(2):
        fb = np.zeros((4,) + comp.shape[1:])
        O.load().oracle_kbn_boundary_fluxes(C.byref(desc), C.byref(O.host_view(fb, lo)),
                                            C.byref(O.host_view(comp, lo)),
                                            C.byref(O.host_view(refs, lo)), C.byref(s), 0.0, 1.0, d)
        np.testing.assert_allclose(fb[1 + d, 0][cut], comp[4, 0][cut], rtol=1e-13)
        others = [k for k in range(4) if k != 1 + d]
        assert np.max(np.abs(fb[others][:, 0][:, cut])) <= 1e-14


def test_moving_gas_pushes_on_the_wall():
    """Flow into the wall raises the boundary pressure above the cell's
    (two-shock), flow away lowers it (two-rarefaction)."""
    geom, comp, lo = circle_case(perturb=False, wedge=False)
    h = O.HostCutCellData(geom)
    s = h.c_struct()
    desc = _desc()
    cut = (geom.alpha > 0) & (geom.alpha < 1)
    jj, ii = np.nonzero(cut)
    n = -geom.normals[:, jj, ii]  # into the fluid
    for sign in (+1.0, -1.0):
        c = comp.copy()
        u = -sign * 0.3 * n       # sign +1: towards the wall
        rho, p = c[0, 0, jj, ii], c[4, 0, jj, ii]
        c[1, 0, jj, ii] = rho * u[0]
        c[2, 0, jj, ii] = rho * u[1]
        c[3, 0, jj, ii] = 0.5 * rho * 0.09 + p / 0.4
        refs = np.zeros_like(c)
        O.load().oracle_kbn_reference_states(C.byref(desc), C.byref(O.host_view(refs, lo)),
                                             C.byref(O.host_view(c, lo)), C.byref(s))
        pr = refs[4, 0, jj, ii]
        if sign > 0:
            assert np.all(pr > p)
        else:
            assert np.all(pr < p)
        # the wall state has zero normal velocity
        un = (refs[1, 0, jj, ii] * n[0] + refs[2, 0, jj, ii] * n[1]) / refs[0, 0, jj, ii]
        assert np.max(np.abs(un)) <= 1e-12


# ---- reduction to the regular method ---------------------------------------------------

def _regular_pipeline(desc, comp, lo, dt, dx, d):
    fshape = list(comp.shape)
    fshape[0] = 4
    fshape[3 - d] -= 3  # faces with a full MUSCL stencil
    flo = list(lo)
    flo[d] += 2
    f = np.zeros(fshape)
    O.load().oracle_numeric_fluxes(C.byref(desc), C.byref(O.host_view(f, flo)),
                                   C.byref(O.host_view(comp, lo)), dt, dx, d)
    return f, flo


def test_uncut_geometry_reduces_to_the_regular_method():
    geom, comp, lo = uncut_case()
    desc = _desc()
    h = O.HostCutCellData(geom).fill()
    s = h.c_struct()
    dt, dx = 0.004, 1.0 / 32
    for d in range(2):
        f_reg, flo = _regular_pipeline(desc, comp, lo, dt, dx, d)
        f_kbn = np.zeros_like(f_reg)
        O.load().oracle_kbn_regular_fluxes(C.byref(desc), C.byref(O.host_view(f_kbn, flo)),
                                           C.byref(O.host_view(comp, lo)), C.byref(s), dt, dx, d)
        np.testing.assert_array_equal(f_kbn, f_reg)
        dt_reg = C.c_double()
        dt_kbn = C.c_double()
        O.load().oracle_stable_dt(C.byref(desc), C.byref(O.host_view(comp, lo)), dx, d,
                                  C.byref(dt_reg))
        O.load().oracle_kbn_stable_dt(C.byref(desc), C.byref(O.host_view(comp, lo)), C.byref(s),
                                      dx, d, C.byref(dt_kbn))
        assert dt_kbn.value == dt_reg.value
        # stabilised fluxes equal the regular ones; the h-box update is the
        # regular update
        fb = np.zeros((4,) + comp.shape[1:])
        stab, fsl, fsr = (np.zeros_like(f_reg) for _ in range(3))
        O.load().oracle_kbn_cutcell_fluxes(2, C.byref(O.host_view(stab, flo)),
                                           C.byref(O.host_view(fsl, flo)),
                                           C.byref(O.host_view(fsr, flo)),
                                           C.byref(O.host_view(fb, lo)),
                                           C.byref(O.host_view(f_reg, flo)), C.byref(s), d)
        np.testing.assert_array_equal(stab, f_reg)
        cells_lo = list(flo)
        cshape = list(f_reg.shape)
        cshape[3 - d] -= 1
        cells = np.ascontiguousarray(comp[:4][tuple(slice(c - l, c - l + n) for c, l, n in
                                                       zip([0, 0, cells_lo[1], cells_lo[0]],
                                                           [0, 0, lo[1], lo[0]], cshape))])
        reg = cells.copy()
        O.load().oracle_update_conservatively(C.byref(O.host_view(reg, cells_lo)),
                                              C.byref(O.host_view(cells, cells_lo)),
                                              C.byref(O.host_view(f_reg, flo)), dt, dx, d)
        kbn = cells.copy()
        O.load().oracle_kbn_update(C.byref(O.host_view(kbn, cells_lo)),
                                   C.byref(O.host_view(f_reg, flo)), C.byref(O.host_view(stab, flo)),
                                   C.byref(O.host_view(fsl, flo)), C.byref(O.host_view(fsr, flo)),
                                   C.byref(O.host_view(fb, lo)), C.byref(s), dt, dx, d)
        np.testing.assert_array_equal(kbn, reg)


# ---- conservation of TimeIntegrator2 and the h-box branches ----------------------------

def test_update2_telescopes():
    """sum(alpha (u_new - u_old)) = dt/dx [beta F at the two ends of each row
    + sum over cells of (beta_R - beta_L) F_B]: every interior face flux
    cancels between its two cells."""
    geom, comp, lo = circle_case()
    h = O.HostCutCellData(geom)
    s = h.c_struct()
    rng = np.random.default_rng(9)
    dt, dx = 0.01, 1.0 / 32
    for d in range(2):
        gy, gx = geom.alpha.shape
        cshape = (4, 1, gy - 2 * (d == 1), gx - 2 * (d == 0))
        clo = [lo[0] + (d == 0), lo[1] + (d == 1), 0]
        src = rng.normal(size=cshape)
        dest = src.copy()
        fshape = list(cshape)
        fshape[3 - d] += 1
        fst = rng.normal(size=fshape)
        fb = rng.normal(size=(4,) + comp.shape[1:])
        O.load().oracle_kbn_update2(C.byref(O.host_view(dest, clo)), C.byref(O.host_view(src, clo)),
                                    C.byref(O.host_view(fst, clo)), C.byref(O.host_view(fb, lo)),
                                    C.byref(s), dt, dx, d)
        sl = (slice(clo[1] - lo[1], clo[1] - lo[1] + cshape[2]),
              slice(clo[0] - lo[0], clo[0] - lo[0] + cshape[3]))
        alpha = geom.alpha[sl]
        beta = (geom.beta_x if d == 0 else geom.beta_y)
        bsl = (slice(sl[0].start, sl[0].stop + (d == 1)), slice(sl[1].start, sl[1].stop + (d == 0)))
        beta = beta[bsl]
        live = alpha > 0
        lhs = np.sum(alpha * (dest - src)[:, 0] * live, axis=(1, 2))
        if d == 0:
            bL, bR = beta[:, :-1], beta[:, 1:]
            fL, fR = fst[:, 0, :, :-1], fst[:, 0, :, 1:]
        else:
            bL, bR = beta[:-1, :], beta[1:, :]
            fL, fR = fst[:, 0, :-1, :], fst[:, 0, 1:, :]
        fbc = fb[:, 0][:, sl[0], sl[1]]
        same = bL == bR
        # cells with equal betas use the regular form: alpha * (F_L - F_R)
        rhs = np.where(live & same, alpha * (fL - fR), 0.0) + \
            np.where(live & ~same, bL * fL - bR * fR - (bL - bR) * fbc, 0.0)
        rhs = dt / dx * np.sum(rhs, axis=(1, 2))
        np.testing.assert_allclose(lhs, rhs, rtol=1e-12, atol=1e-12)


def test_hbox_branches_closed_form():
    """One cell per branch of TimeIntegrator with binary fractions."""
    geom = make_geometry([], (3, 1), (0.0, 0.0), 1.0, 2)
    # cell i = 0: beta_L 1 > beta_R 0.5, alpha 0.75, centroid 0.25
    geom.alpha[2, 2] = 0.75
    geom.beta_x[2, 3] = 0.5
    geom.centroids[0, 2, 2] = 0.25
    h = O.HostCutCellData(geom).fill()
    s = h.c_struct()
    lo = [0, 0, 0]
    flo = [0, 0, 0]
    u = np.full((1, 1, 1, 1), 2.0)
    fus = np.array([[[[1.0, 0.5]]]])
    fst = np.array([[[[0.75, 0.25]]]])
    fsl = np.array([[[[0.5, 0.125]]]])
    fsr = np.array([[[[0.25, 0.375]]]])
    fb = np.full((1, 1, 5, 7), 0.0625)
    fb_lo = [-2, -2, 0]
    dt, dx = 0.5, 1.0
    O.load().oracle_kbn_update(C.byref(O.host_view(u, lo)), C.byref(O.host_view(fus, flo)),
                               C.byref(O.host_view(fst, flo)), C.byref(O.host_view(fsl, flo)),
                               C.byref(O.host_view(fsr, flo)), C.byref(O.host_view(fb, fb_lo)),
                               C.byref(s), dt, dx, 0)
    # fractions of face 0 (betas 1 | 1, 1, 0.5): dL 0, dR 0.5 -> us 0.5, sR 0.5
    us0, sl0, sr0 = 0.5, 0.0, 0.5
    dlg = min(max(sr0 * (0.5 + 0.25) / 0.75, 0.0), 1.0)
    expect = (2.0 + dt / dx * (us0 / 0.75) * (1.0 - 0.25) + dt / dx * (sl0 / 0.75) * (0.5 - 0.25)
              + dt / dx * dlg * (1.0 - 0.0625))
    assert u[0, 0, 0, 0] == expect


def test_channel_geometry_full_size_solid_area():
    """The 2048 x 1024 channel geometry before small cells are shrunk: the
    covered area sum (1 - alpha) dx^2 over the interior equals the
    cylinder's pi r^2 plus the wedge triangle's 0.7 * 0.7 tan 30 / 2 (both
    bodies lie inside the channel, apart from each other)."""
    import math
    from finitevolumesolver_amd.geometry import channel_geometry
    g = channel_geometry(2048, 1024, perturb=False)
    dx = 4.0 / 2048
    solid = float((1.0 - g.alpha[2:-2, 2:-2]).sum()) * dx * dx
    exact = math.pi * 0.25 ** 2 + 0.5 * 0.7 * 0.7 * math.tan(math.radians(30.0))
    assert abs(solid / exact - 1.0) < 1e-12
