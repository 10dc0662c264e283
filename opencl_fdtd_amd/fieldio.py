"""Field vectors and simulation output files.

``vec``/``unvec`` convert between the flat 3N layout the simulation keeps
(component-major, C order inside a component: ``fdfd_tools.vec`` as used
at ``opencl_fdtd/simulation.py:347, 357``) and a list of three (sx, sy, sz)
arrays.  ``save_fields`` writes what the reference example saves after its
run (``fdtd.py:181-191``: the grid and the unvec'd E, H and, with Poynting,
S) plus F/G with Bloch boundaries.  The reference pickles a dict with dill
and lzma; here the file is a compressed ``.npz`` (arrays only, nothing
executable), read back by ``load_fields``.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence

import numpy


def vec(fields: Sequence) -> numpy.ndarray:
    """Concatenation of the C-order flattenings of three same-shape arrays."""
    parts = [numpy.asarray(f) for f in fields]
    if len(parts) != 3 or any(p.shape != parts[0].shape for p in parts):
        raise ValueError("vec needs three arrays of one shape")
    return numpy.concatenate([p.ravel(order="C") for p in parts])


def unvec(v, shape: Sequence[int]) -> List[numpy.ndarray]:
    """The three (sx, sy, sz) components of a flat 3N vector (host array,
    or anything with ``.get()`` / ``.cpu()`` such as ``sim.E``)."""
    if hasattr(v, "get"):
        v = v.get()
    elif hasattr(v, "cpu"):
        v = v.cpu().numpy()
    v = numpy.asarray(v)
    n = int(numpy.prod(shape))
    if v.size != 3 * n:
        raise ValueError("vector of %d values does not hold 3 x %d" % (v.size, n))
    return [v[k * n:(k + 1) * n].reshape(tuple(shape), order="C") for k in range(3)]


def field_dict(sim) -> Dict[str, List[numpy.ndarray]]:
    """The unvec'd fields of a simulation: E and H, S when Poynting is on,
    F and G with Bloch boundaries (fdtd.py:183-191)."""
    out = {"E": unvec(sim.E, sim.shape), "H": unvec(sim.H, sim.shape)}
    for name in ("S", "F", "G"):
        v = getattr(sim, name, None)
        if v is not None:
            out[name] = unvec(v, sim.shape)
    return out


def save_fields(path: str, sim, grid: Optional[Dict[str, numpy.ndarray]] = None, step: Optional[int] = None) -> str:
    """Write ``field_dict(sim)`` as ``<name>_<x|y|z>`` arrays of a compressed
    ``.npz``, with the grid shape, dt and, when given, the step count and the
    caller's grid description (arrays, e.g. cell edge coordinates, saved
    as ``grid_<key>``).  Returns the path written."""
    arrays = {"shape": numpy.asarray(sim.shape, dtype=numpy.int64), "dt": numpy.float64(sim.dt)}
    if step is not None:
        arrays["step"] = numpy.int64(step)
    for name, comps in field_dict(sim).items():
        for axis, c in zip("xyz", comps):
            arrays["%s_%s" % (name, axis)] = c
    for k, v in (grid or {}).items():
        arrays["grid_" + k] = numpy.asarray(v)
    if not path.endswith(".npz"):
        path += ".npz"
    numpy.savez_compressed(path, **arrays)
    return path


def load_fields(path: str) -> Dict:
    """The dict ``save_fields`` wrote: 'shape', 'dt', ('step'), each field as
    a list of three arrays, and 'grid' (a dict) when one was saved."""
    with numpy.load(path, allow_pickle=False) as z:
        out = {"shape": tuple(int(s) for s in z["shape"]), "dt": float(z["dt"])}
        if "step" in z.files:
            out["step"] = int(z["step"])
        names = sorted({k[:-2] for k in z.files if k[-2:] in ("_x", "_y", "_z") and not k.startswith("grid_")})
        for name in names:
            out[name] = [z["%s_%s" % (name, a)] for a in "xyz"]
        grid = {k[5:]: z[k] for k in z.files if k.startswith("grid_")}
        if grid:
            out["grid"] = grid
    return out
