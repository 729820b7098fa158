"""fvs_patch_stable_dt_noalloc (partial minima in the library's fixed device
array, no allocation) against fvs_patch_stable_dt (partials in the handle's
buffer), through the C ABI: bit for bit, on two calls in a row whose second
patch is smaller, so the second call finds the first call's keys in the
array and must not read them."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from fields import rough_field

import finitevolumesolver_amd as fvs
from finitevolumesolver_amd import _lib as L

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dts(entry, fm, cons, dx):
    import torch
    D = fm.equation.dim
    out = torch.full((3,), float("nan"), dtype=torch.float64, device=cons.device)
    n = (cons.shape[3], cons.shape[2], cons.shape[1])
    p = L.ptr_array([cons[c].data_ptr() for c in range(D + 2)])
    dxa = (C.c_double * 3)(*(list(dx) + [1.0] * (3 - len(dx))))
    L.check(getattr(L.load(), entry)(fm.handle, p, L.i64x3(n), dxa, out.data_ptr()))
    torch.cuda.synchronize()
    return out[:D].cpu().numpy()


@pytest.mark.parametrize("dim,base,precision,shapes", [
    (3, "HLLEM", "fp64", [(40, 36, 44), (9, 11, 13)]),
    (3, "HLLEM", "fp64_rcp", [(33, 30, 70), (12, 8, 10)]),
    (2, "HLLEM", "fp64", [(1, 300, 260), (1, 17, 23)]),
    (3, "Godunov", "fp64", [(24, 20, 30), (6, 7, 9)]),
])
def test_noalloc_stable_dt_matches_allocating_entry(cuda, dim, base, precision, shapes):
    import torch
    kw = {"reconstruction": "None"} if base == "Godunov" else {}
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=dim), base=base, precision=precision, **kw)
    for k, shape in enumerate(shapes):
        _note(f"noalloc stable dt: dim={dim} base={base} precision={precision} "
              f"call {k} shape={shape}")
        cons = torch.from_numpy(rough_field(shape, seed=5 + k)).to(cuda)
        if dim == 2:
            cons = cons.reshape(4, 1, shape[1], shape[2]).contiguous()
        dx = [1.0 / shape[2], 1.3 / shape[1], 0.9 / max(shape[0], 1)][:dim]
        ref = _dts("fvs_patch_stable_dt", fm, cons, dx)
        got = _dts("fvs_patch_stable_dt_noalloc", fm, cons, dx)
        assert np.all(np.isfinite(ref)) and np.all(ref > 0), ref
        assert np.array_equal(got, ref), (k, got, ref)


def test_noalloc_rejects_patch_beyond_its_array(cuda):
    """A patch whose reduction needs more keys than the fixed array holds is
    rejected before any launch (the sizes are checked on the host; the
    component pointers are never read)."""
    import torch
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=3))
    dummy = torch.zeros(8, dtype=torch.float64, device=cuda)
    p = L.ptr_array([dummy.data_ptr()] * 5)
    n = (4096, 4096, 4096)
    dxa = (C.c_double * 3)(1.0, 1.0, 1.0)
    out = torch.zeros(3, dtype=torch.float64, device=cuda)
    rc = L.load().fvs_patch_stable_dt_noalloc(fm.handle, p, L.i64x3(n), dxa, out.data_ptr())
    assert rc == L.FVS_BAD_ARGUMENT
    assert b"allocation-free" in L.load().fvs_last_error()
