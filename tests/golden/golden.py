"""Golden vectors of the FP64 oracle: inputs and expected outputs, written by
`python tests/golden/golden.py` and compared by tests/test_golden.py (oracle
against the vectors, CPU) and tests/test_gpu_golden.py (GPU against the
vectors).  Every file records the SHA-256 of the oracle sources it came from.

Cases (all seeded, generated here from the oracle; nothing is read from the
reference):
  fluxes_2d / fluxes_3d  per-face numeric fluxes and stable dt of five flux
                         methods (MUSCL-Hancock MinMod + HLLEM Larrouturou,
                         MUSCL-Hancock VanLeer + plain HLLEM, first-order
                         HLLEM, first-order exact Godunov, MUSCL-Hancock +
                         exact Godunov) in every direction, on a rough field
                         with adversarial cells (a density drop of 100 at equal
                         pressure, a pressure ratio of 900, Mach 5 flow, exact
                         zero momenta);
  sweep_2d               one fused split sweep per direction of a 16 x 16
                         patch (transmissive);
  cutcell_2d             FillCutCellData, IsCutCell, masked regular fluxes,
                         boundary fluxes (HLLEM boundary solver), stabilised
                         fluxes and the h-box update (TimeIntegrator) on a
                         12 x 12 patch cut by a disk and a wedge;
  hll_simd_2d            HLL (Einfeldt signals) fluxes and stable dt on the
                         inputs of the reference's own HLL test
                         (tests/test.simd_FluxMethod.cpp:54-186): one
                         MUSCL-Hancock(HLL) face and a 16 x 16 patch in X and Y.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
for p in (str(ROOT), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle_lib as O  # noqa: E402
from fields import rough_field  # noqa: E402
from sweep_ref import oracle_sweep, padded_complete  # noqa: E402

import finitevolumesolver_amd as fvs  # noqa: E402
from finitevolumesolver_amd import _lib as L  # noqa: E402

METHODS = {
    # name: (base, reconstruction, limiter)
    "muscl_minmod_hllem_larr": ("HLLEM", "MusclHancock", "MinMod"),
    "muscl_vanleer_hllem": ("HLLEM_plain", "MusclHancock", "VanLeer"),
    "first_order_hllem_larr": ("HLLEM", "None", "MinMod"),
    "first_order_godunov": ("Godunov", "None", "MinMod"),
    "muscl_minmod_godunov": ("Godunov", "MusclHancock", "MinMod"),
}
EXACT = ("first_order_godunov", "muscl_minmod_godunov")


def oracle_hash() -> str:
    h = hashlib.sha256()
    for f in sorted((ROOT / "oracle").glob("*.[ch]pp")):
        h.update(f.name.encode())
        h.update(f.read_bytes())
    return h.hexdigest()


def desc_of(D: int, name: str) -> L.MethodDesc:
    base, recon, lim = METHODS[name]
    return fvs.make_desc(fvs.PerfectGas(dim=D), base=base, reconstruction=recon, limiter=lim)


def flux_inputs(D: int) -> dict:
    """A rough field with adversarial cells; conservative (D+2, nz, ny, nx)."""
    shape = (1, 12, 12) if D == 2 else (6, 6, 6)
    cons = rough_field(shape, seed=77 + D)
    rho, mom, E = cons[0], cons[1:1 + D], cons[D + 1]
    ke = 0.5 * sum(m * m for m in mom) / rho
    p = (E - ke) * 0.4
    # a light cell: density / 100 at unchanged pressure (sound speed x 10)
    rho.flat[60] *= 1e-2
    # pressure ratio 900 between neighbours (100 already makes the 2D
    # VanLeer half-step states negative: MUSCL-Hancock's own limit)
    p.flat[17] *= 30.0
    p.flat[18] /= 30.0
    # supersonic flow: velocity 5 sound speeds
    a = np.sqrt(1.4 * p.flat[30] / rho.flat[30])
    mom[0].flat[30] = 5.0 * a * rho.flat[30]
    # exact zero momenta (signed zeros)
    for m in mom:
        m.flat[41] = 0.0
    mom[0].flat[42] = -0.0
    ke = 0.5 * sum(m * m for m in mom) / rho
    E[...] = p / 0.4 + ke
    cons = np.ascontiguousarray(cons)
    inp = {"cons": cons, "dx": 1.0 / shape[2], "dt": 0.1 / shape[2]}
    # complete states with transmissive ghosts, one and two layers
    for gw in (1, 2):
        inp[f"comp_gw{gw}"], _ = padded_complete(cons, D, gw)
    return inp


def flux_outputs(D: int, inp: dict) -> dict:
    """Oracle fluxes and stable dt of every method and direction."""
    cons, dx, dt = inp["cons"], inp["dx"], inp["dt"]
    out = {}
    for name in METHODS:
        desc = desc_of(D, name)
        gw = 2 if METHODS[name][1] == "MusclHancock" else 1
        comp, g = padded_complete(cons, D, gw)
        nz, ny, nx = cons.shape[1:]
        lo = (-g[0], -g[1], -g[2])
        assert np.array_equal(comp, inp[f"comp_gw{gw}"])
        sv = O.host_view(comp, lo)
        for d in range(D):
            fshape = [nx, ny, nz]
            fshape[d] += 1
            f = np.zeros((D + 2, fshape[2], fshape[1], fshape[0]))
            rc = O.load().oracle_numeric_fluxes(C.byref(desc), C.byref(O.host_view(f)),
                                                C.byref(sv), dt, dx, d)
            assert rc == 0, (name, d, rc)
            sdt = C.c_double()
            O.load().oracle_stable_dt(C.byref(desc), C.byref(sv), dx, d, C.byref(sdt))
            assert np.all(np.isfinite(f)), (name, d)
            out[f"{name}/flux{d}"] = f
            out[f"{name}/dt{d}"] = np.array([sdt.value])
    return out


def sweep_inputs() -> dict:
    cons = rough_field((1, 16, 16), seed=91)
    return {"cons": np.ascontiguousarray(cons), "dx": 1.0 / 16, "dt": 0.2 / 16}


def sweep_outputs(inp: dict) -> dict:
    desc = desc_of(2, "muscl_minmod_hllem_larr")
    out = {}
    for d in range(2):
        ref, ref_dt, _ = oracle_sweep(desc, inp["cons"], inp["dt"], inp["dx"], d)
        out[f"out{d}"] = ref
        out[f"dt{d}"] = np.array([ref_dt])
    return out


def cutcell_inputs() -> dict:
    from kbn_cases import circle_case
    geom, comp, lo = circle_case(nx=12, ny=12, seed=4)
    return {"alpha": geom.alpha, "beta_x": geom.beta_x, "beta_y": geom.beta_y,
            "normals": geom.normals, "centroids": geom.centroids, "comp": comp,
            "lo": np.array(lo, np.int64), "dx": np.array([1.0 / 32])}


class _Geom:
    """The geometry arrays of a cut-cell fixture in the shape HostCutCellData
    and CutCellData.from_geometry2d read."""

    def __init__(self, inp):
        self.alpha = inp["alpha"]
        self.beta_x = inp["beta_x"]
        self.beta_y = inp["beta_y"]
        self.normals = inp["normals"]
        self.centroids = inp["centroids"]
        self.lo = tuple(int(v) for v in inp["lo"])
        self.meta = {}


def cutcell_outputs(inp: dict) -> dict:
    """Oracle pipeline of one split step per direction (HLLEM boundary
    solver, so every stage is bit-exact on the GPU)."""
    geom = _Geom(inp)
    comp, lo, dx = inp["comp"], geom.lo, float(inp["dx"][0])
    desc = fvs.make_desc(fvs.PerfectGas(dim=2), base="HLLEM", reconstruction="MusclHancock",
                         limiter="MinMod", boundary_solver="hllem")
    h = O.HostCutCellData(geom).fill()
    s = h.c_struct()
    out = {}
    for k, arrs in h.fractions.items():
        for d in range(2):
            out[f"{k}{d}"] = arrs[d]
    gy, gx = geom.alpha.shape
    lo3 = np.array([lo[0], lo[1], 0], np.int64)
    hi3 = np.array([lo[0] + gx - 1, lo[1] + gy - 1, 0], np.int64)
    flags = np.zeros((gy, gx), np.int32)
    O.load().oracle_kbn_is_cut_cell(2, C.byref(s), O.ptr(lo3), O.ptr(hi3), O.ptr(flags))
    out["is_cut"] = flags
    refs = np.zeros_like(comp)
    O.load().oracle_kbn_reference_states(C.byref(desc), C.byref(O.host_view(refs, lo)),
                                         C.byref(O.host_view(comp, lo)), C.byref(s))
    out["refs"] = refs
    for d in range(2):
        dt_ref = C.c_double()
        O.load().oracle_kbn_stable_dt(C.byref(desc), C.byref(O.host_view(comp, lo)), C.byref(s),
                                      dx, d, C.byref(dt_ref))
        dt = 0.5 * dt_ref.value
        out[f"dt{d}"] = np.array([dt_ref.value])
        fshape = [4, 1, gy, gx]
        fshape[3 - d] -= 3
        flo = [lo[0], lo[1], 0]
        flo[d] += 2
        f = np.zeros(fshape)
        O.load().oracle_kbn_regular_fluxes(C.byref(desc), C.byref(O.host_view(f, flo)),
                                           C.byref(O.host_view(comp, lo)), C.byref(s), dt, dx, d)
        out[f"regular{d}"] = f
        fb = np.zeros((4,) + comp.shape[1:])
        O.load().oracle_kbn_boundary_fluxes(C.byref(desc), C.byref(O.host_view(fb, lo)),
                                            C.byref(O.host_view(comp, lo)),
                                            C.byref(O.host_view(refs, lo)), C.byref(s), dt, dx, d)
        out[f"boundary{d}"] = fb
        sshape = list(fshape)
        sshape[3 - d] -= 2
        slo = list(flo)
        slo[d] += 1
        sl_ = tuple([slice(None), slice(None)] +
                    [slice(1, -1) if (3 - a) == d else slice(None) for a in (2, 3)])
        f_in = np.ascontiguousarray(f[sl_])
        outs = [np.zeros(sshape) for _ in range(3)]
        O.load().oracle_kbn_cutcell_fluxes(2, *[C.byref(O.host_view(o, slo)) for o in outs],
                                           C.byref(O.host_view(fb, lo)),
                                           C.byref(O.host_view(f_in, slo)), C.byref(s), d)
        stab, fsl, fsr = outs
        for key, arr in (("stable", stab), ("shielded_left_flux", fsl),
                         ("shielded_right_flux", fsr)):
            out[f"{key}{d}"] = arr
        cshape = list(sshape)
        cshape[3 - d] -= 1
        clo = list(slo)
        off = [clo[1] - lo[1], clo[0] - lo[0]]
        cells = np.ascontiguousarray(comp[:4, :, off[0]:off[0] + cshape[2],
                                          off[1]:off[1] + cshape[3]])
        u_ref = cells.copy()
        O.load().oracle_kbn_update(C.byref(O.host_view(u_ref, clo)),
                                   C.byref(O.host_view(f_in, slo)), C.byref(O.host_view(stab, slo)),
                                   C.byref(O.host_view(fsl, slo)), C.byref(O.host_view(fsr, slo)),
                                   C.byref(O.host_view(fb, lo)), C.byref(s), dt, dx, d)
        out[f"update{d}"] = u_ref
        out[f"update_cells{d}"] = cells
        out[f"update_lo{d}"] = np.array(clo, np.int64)
        out[f"flux_lo{d}"] = np.array(slo, np.int64)
        out[f"regular_lo{d}"] = np.array(flo, np.int64)
        out[f"dt_used{d}"] = np.array([dt])
    return out


def hll_inputs() -> dict:
    """The inputs of the reference's tests/test.simd_FluxMethod.cpp:54-186:
    one MUSCL-Hancock(HLL) face between two pairs of gases at rest
    (rho = 1, E = 2 | E = 1; dt = dx = 1), and a 16 x 16 patch with
    rho = 1, u = 0, E = 1 + sin(x^2 + y^2) / 2 at x = i dx, y = j dx
    (dx = 0.0125), grown by one cell in the sweep direction (X and Y)."""
    import math
    inp = {"single": np.array([[1.0, 0.0, 0.0, 2.0], [1.0, 0.0, 0.0, 2.0],
                               [1.0, 0.0, 0.0, 1.0], [1.0, 0.0, 0.0, 1.0]]),
           "dx": 0.0125}
    for d in range(2):
        lo = [0, 0]
        hi = [15, 15]
        lo[d] -= 1
        hi[d] += 1
        ny, nx = hi[1] - lo[1] + 1, hi[0] - lo[0] + 1
        cons = np.zeros((4, 1, ny, nx))
        cons[0] = 1.0
        for j in range(ny):
            for i in range(nx):
                x, y = (lo[0] + i) * 0.0125, (lo[1] + j) * 0.0125
                cons[3, 0, j, i] = 1.0 + 0.5 * math.sin(x * x + y * y)
        inp[f"patch{d}"] = cons
        inp[f"patch_lo{d}"] = np.array([lo[0], lo[1], 0], np.int64)
    return inp


def hll_outputs(inp: dict) -> dict:
    """Oracle HLL (Einfeldt signals) fluxes and stable dt on those inputs."""
    out = {}
    desc = fvs.make_desc(fvs.PerfectGas(dim=2), base="HLL", reconstruction="MusclHancock")
    st = np.zeros((4, 6))
    for s_, q in enumerate(inp["single"]):
        st[s_, :4] = q
        st[s_, 4] = q[3] * 0.4  # p = (E - 0) / (1 / (gamma - 1)) is exact here
    comp = np.zeros((6, 1, 1, 4))
    comp[:4, 0, 0, :] = inp["single"].T
    cv = O.host_view(comp)
    O.load().oracle_complete_view(2, 1.4, C.byref(cv), C.byref(cv))
    stencil = np.ascontiguousarray(comp[:, 0, 0, :].T)  # (4 states, 6 components)
    f = np.zeros(4)
    rc = O.load().oracle_face_flux(C.byref(desc), 0, 1.0, 1.0, 1,
                                   stencil.ctypes.data_as(C.POINTER(C.c_double)),
                                   f.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0
    out["single_flux"] = f
    out["single_stencil"] = stencil
    first = fvs.make_desc(fvs.PerfectGas(dim=2), base="HLL", reconstruction="None")
    dx = float(inp["dx"])
    for d in range(2):
        cons, lo = inp[f"patch{d}"], tuple(int(v) for v in inp[f"patch_lo{d}"])
        comp = np.zeros((6,) + cons.shape[1:])
        comp[:4] = cons
        cv = O.host_view(comp, lo)
        O.load().oracle_complete_view(2, 1.4, C.byref(cv), C.byref(cv))
        out[f"complete{d}"] = comp
        fshape = [4, 1, 16, 16]
        fshape[3 - d] += 1
        flux = np.zeros(fshape)
        dtc = C.c_double()
        O.load().oracle_stable_dt(C.byref(first), C.byref(cv), dx, d, C.byref(dtc))
        rc = O.load().oracle_numeric_fluxes(C.byref(first), C.byref(O.host_view(flux)),
                                            C.byref(cv), dtc.value, dx, d)
        assert rc == 0
        out[f"flux{d}"] = flux
        out[f"dt{d}"] = np.array([dtc.value])
    return out


CASES = {
    "fluxes_2d": (lambda: flux_inputs(2), lambda i: flux_outputs(2, i)),
    "fluxes_3d": (lambda: flux_inputs(3), lambda i: flux_outputs(3, i)),
    "sweep_2d": (sweep_inputs, sweep_outputs),
    "cutcell_2d": (cutcell_inputs, cutcell_outputs),
    "hll_simd_2d": (hll_inputs, hll_outputs),
}


def path(case: str) -> Path:
    return HERE / f"{case}.npz"


def load_case(case: str):
    """(inputs, expected outputs, oracle hash) of a committed fixture."""
    z = np.load(path(case), allow_pickle=False)
    inp = {k[4:]: z[k] for k in z.files if k.startswith("in::")}
    out = {k[5:]: z[k] for k in z.files if k.startswith("out::")}
    if "dx" in inp and inp["dx"].ndim == 0:
        inp["dx"] = float(inp["dx"])
    if "dt" in inp and inp["dt"].ndim == 0:
        inp["dt"] = float(inp["dt"])
    return inp, out, str(z["oracle_sha256"])


def write_all() -> None:
    digest = oracle_hash()
    for case, (make_in, make_out) in CASES.items():
        inp = make_in()
        out = make_out(inp)
        arrays = {f"in::{k}": np.asarray(v) for k, v in inp.items()}
        arrays.update({f"out::{k}": np.asarray(v) for k, v in out.items()})
        arrays["oracle_sha256"] = np.array(digest)
        np.savez_compressed(path(case), **arrays)
        print(f"{path(case).name}: {len(inp)} inputs, {len(out)} outputs")


if __name__ == "__main__":
    write_all()
