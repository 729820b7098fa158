"""Checkpoints (with restart) and plotfiles of an AMR hierarchy in the
reference's AMReX layouts.

WriteCheckpointFile (src/AMReX/PatchHierarchy.cpp:320-371) writes a
directory with a text `Header` (title line, finest level, the cycle count of
every level, the time point of every level at 17 significant digits, the
BoxArray of every level) and, per level, `Level_<l>/data`: the level's
complete state (density, momentum, energy, pressure, speed of sound) as an
AMReX VisMF multifab, a text header `data_H` (version, layout, component and
ghost counts, BoxArray, the byte offset of each patch in `data_D_00000`,
per-patch minima and maxima) and the data file, where every patch is a FAB
header line followed by its components in Fortran order (x fastest, one
component after the other) as little-endian IEEE doubles.
ReadCheckpointFile (:373-440) reads it back into a hierarchy of the same
geometry.  Here the hierarchy's patch interiors are downloaded, completed on
the host (PerfectGas::CompleteFromCons, array form) and written by numpy;
restart uploads the conservative components and sets the cycle count and
time point, so a restarted run continues bit for bit.

WritePlotFile (include/fub/AMReX/PatchHierarchy.hpp, through
amrex::WriteMultiLevelPlotfile) writes `Header` in the HyperCLaw-V1.1 form
(variable names Density, Momentum_0, Momentum_1, Energy, Pressure,
SpeedOfSound; time; problem domain; refinement ratios; level domains, steps
and cell sizes; per level its boxes in physical coordinates) and
`Level_<l>/Cell` as a VisMF multifab like the checkpoint's.  AMReX itself is
not available here, so the byte layout follows AMReX's documented formats
and is checked by reading it back, not by AMReX's own readers.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from .output import complete_2d

# FAB header of IEEE little-endian doubles (the real descriptor AMReX writes
# on x86: bit layout, then byte order).
_REAL_DESCRIPTOR = "((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"


def _box(lo, hi) -> str:
    return f"(({lo[0]},{lo[1]}) ({hi[0]},{hi[1]}) (0,0))"


def _boxarray(layout) -> str:
    """BoxArray::writeOn: '(' size ' 0', one box per line, ')'."""
    lines = [f"({len(layout)} 0"]
    for (i0, j0, nx, ny) in layout:
        lines.append(_box((i0, j0), (i0 + nx - 1, j0 + ny - 1)))
    return "\n".join(lines) + "\n)"


def _real(v: float) -> str:
    return format(float(v), ".17g")


def _write_level(level_dir: Path, layout, fabs, name: str = "data") -> None:
    """VisMF::Write of one level: <name>_H and <name>_D_00000."""
    level_dir.mkdir(parents=True, exist_ok=True)
    offsets = []
    with open(level_dir / f"{name}_D_00000", "wb") as f:
        for (i0, j0, nx, ny), fab in zip(layout, fabs):
            offsets.append(f.tell())
            ncomp = fab.shape[0]
            head = f"FAB {_REAL_DESCRIPTOR}{_box((i0, j0), (i0 + nx - 1, j0 + ny - 1))} {ncomp}\n"
            f.write(head.encode())
            f.write(np.ascontiguousarray(fab, dtype="<f8").tobytes())
    ncomp = fabs[0].shape[0] if fabs else 0
    lines = ["1", "1", str(ncomp), "0", _boxarray(layout), str(len(layout))]
    lines += [f"FabOnDisk: {name}_D_00000 {o}" for o in offsets]
    lines.append("")
    for reduce in (np.min, np.max):
        lines.append(f"{len(fabs)},{ncomp}")
        for fab in fabs:
            lines.append("".join(_real(reduce(fab[c])) + "," for c in range(ncomp)))
    with open(level_dir / f"{name}_H", "w") as f:
        f.write("\n".join(lines) + "\n")


def write_checkpoint(path, hierarchy) -> None:
    """Writes the hierarchy's state (pool 0 of every level) as the
    reference's checkpoint directory `path`."""
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    h = hierarchy
    gamma = h.flux_method.equation.gamma
    n = h.nlevels
    cycles = [h.cycle * h.ratio ** l for l in range(n)]
    head = ["Checkpoint File", str(n - 1), "".join(f"{c} " for c in cycles),
            "".join(f"{_real(h.time)} " for _ in range(n))]
    for l in range(n):
        head.append(_boxarray(h.layouts[l]))
    with open(path / "Header", "w") as f:
        f.write("\n".join(head) + "\n")
    for l in range(n):
        fabs = [complete_2d(np.asarray(u), gamma) for u in h.interiors(l)]
        _write_level(path / f"Level_{l}", h.layouts[l], fabs)


def _parse_boxarray(lines, k):
    """BoxArray::readFrom starting at lines[k]; returns (layout, next k)."""
    n = int(lines[k].strip("(").split()[0])
    layout = []
    for i in range(n):
        txt = lines[k + 1 + i].replace("(", " ").replace(")", " ").replace(",", " ").split()
        lo0, lo1, hi0, hi1 = (int(v) for v in txt[:4])
        layout.append((lo0, lo1, hi0 - lo0 + 1, hi1 - lo1 + 1))
    if lines[k + 1 + n].strip() != ")":
        raise ValueError("malformed BoxArray in checkpoint header")
    return layout, k + 2 + n


def read_header(path):
    """(finest level, cycles, time points, layouts) of a checkpoint."""
    lines = (Path(path) / "Header").read_text().split("\n")
    if lines[0] != "Checkpoint File":
        raise ValueError(f"{path} is not a checkpoint directory")
    finest = int(lines[1])
    cycles = [int(v) for v in lines[2].split()]
    times = [float(v) for v in lines[3].split()]
    layouts, k = [], 4
    for _ in range(finest + 1):
        lay, k = _parse_boxarray(lines, k)
        layouts.append(lay)
    return finest, cycles, times, layouts


def read_level(level_dir, name: str = "data") -> list:
    """VisMF::Read of one level: the (ncomp, ny, nx) FAB of every patch."""
    level_dir = Path(level_dir)
    hl = (level_dir / f"{name}_H").read_text().split("\n")
    ncomp = int(hl[2])
    layout, k = _parse_boxarray(hl, 4)
    nfab = int(hl[k])
    entries = [hl[k + 1 + i].split() for i in range(nfab)]
    fabs = []
    with open(level_dir / entries[0][1], "rb") as f:
        data = f.read()
    for (i0, j0, nx, ny), e in zip(layout, entries):
        off = int(e[2])
        end = data.index(b"\n", off)
        head = data[off:end].decode()
        if not head.startswith("FAB ") or _REAL_DESCRIPTOR not in head:
            raise ValueError(f"unsupported FAB header {head!r}")
        count = ncomp * nx * ny
        arr = np.frombuffer(data, dtype="<f8", count=count, offset=end + 1)
        fabs.append(arr.reshape(ncomp, ny, nx).copy())
    return fabs


def read_checkpoint(path, flux_method, domain, dx0: float, ratio: int = 2):
    """ReadCheckpointFile: a new AmrHierarchy on the checkpoint's layouts
    with its conservative state, cycle count and time point."""
    from .amr import AmrHierarchy
    finest, cycles, times, layouts = read_header(path)
    h = AmrHierarchy(flux_method, domain, layouts, dx0, ratio)
    states = [read_level(Path(path) / f"Level_{l}") for l in range(finest + 1)]
    h.set_state(lambda l, p: states[l][p][:4])
    h.cycle = cycles[0]
    h.time = times[0]
    return h


PLOT_NAMES_2D = ("Density", "Momentum_0", "Momentum_1", "Energy", "Pressure", "SpeedOfSound")


def write_plotfile(path, hierarchy, prob_lo=(0.0, 0.0), varnames=PLOT_NAMES_2D) -> None:
    """WritePlotFile: every level's complete state as a HyperCLaw-V1.1
    plotfile directory `path` (problem domain [prob_lo, prob_lo + domain *
    dx0])."""
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    h = hierarchy
    gamma = h.flux_method.equation.gamma
    n = h.nlevels
    dx = [h.dx[l] for l in range(n)]
    prob_hi = [prob_lo[d] + h.domain[d] * dx[0] for d in range(2)]
    steps = [h.cycle * h.ratio ** l for l in range(n)]
    head = ["HyperCLaw-V1.1", str(len(varnames)), *varnames, "2", _real(h.time), str(n - 1),
            "".join(_real(v) + " " for v in prob_lo), "".join(_real(v) + " " for v in prob_hi),
            "".join(f"{h.ratio} " for _ in range(n - 1)),
            "".join(_box((0, 0), (h.domain[0] * h.ratio ** l - 1, h.domain[1] * h.ratio ** l - 1))
                    + " " for l in range(n)),
            "".join(f"{s_} " for s_ in steps)]
    for l in range(n):
        head.append(f"{_real(dx[l])} {_real(dx[l])} ")
    head += ["0", "0"]
    for l in range(n):
        head.append(f"{l} {len(h.layouts[l])} {_real(h.time)}")
        head.append(str(steps[l]))
        for (i0, j0, nx, ny) in h.layouts[l]:
            for d, (a, w) in enumerate(((i0, nx), (j0, ny))):
                head.append(f"{_real(prob_lo[d] + a * dx[l])} {_real(prob_lo[d] + (a + w) * dx[l])}")
        head.append(f"Level_{l}/Cell")
    with open(path / "Header", "w") as f:
        f.write("\n".join(head) + "\n")
    for l in range(n):
        fabs = [complete_2d(np.asarray(u), gamma) for u in h.interiors(l)]
        _write_level(path / f"Level_{l}", h.layouts[l], fabs, name="Cell")


def read_plotfile(path):
    """(variable names, time, per-level layouts, per-level FABs) of a
    plotfile written by write_plotfile."""
    lines = (Path(path) / "Header").read_text().split("\n")
    if lines[0] != "HyperCLaw-V1.1":
        raise ValueError(f"{path} is not a HyperCLaw-V1.1 plotfile")
    nvar = int(lines[1])
    names = lines[2:2 + nvar]
    time = float(lines[3 + nvar])
    finest = int(lines[4 + nvar])
    fabs, layouts = [], []
    for l in range(finest + 1):
        fabs.append(read_level(Path(path) / f"Level_{l}", name="Cell"))
        hl = (Path(path) / f"Level_{l}" / "Cell_H").read_text().split("\n")
        layouts.append(_parse_boxarray(hl, 4)[0])
    return names, time, layouts, fabs


__all__ = ["write_checkpoint", "read_checkpoint", "read_header", "read_level",
           "write_plotfile", "read_plotfile"]
