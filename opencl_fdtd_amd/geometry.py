"""Photonic-crystal geometry: lattice sites, cavity defects and their
rasterisation onto the Yee grid.

Regenerates the helpers of the reference's example layer (``pcgen.py``
and the L3 pattern of ``fdtd.py:28-81``) without masque or gridlock: hole
lists are ``(n, 2)`` arrays of lattice sites ``[x, y]`` or ``(n, 3)`` arrays
``[x, y, r]`` in units of the lattice constant, and ``draw_holes`` /
``slab_epsilon`` turn them into the per-axis permittivity grids a
``Simulation`` takes.

Two defects of ``pcgen.py`` are documented, not reproduced:

* ``triangular_lattice`` takes its row count from ``dims[0]``
  (``pcgen.py:34``); here rows come from ``dims[1]``, as its docstring
  says.  The two agree for square ``dims``, which is how every caller in the
  reference (``ln_defect``, ``r6_defect``) uses it.
* ``l3_shift_perturbed_defect`` tests ``numpy.fabs(row) in
  perturbed_holes`` against a generator of 2-tuples (``pcgen.py:209-212``):
  a 3-vector never equals a 2-tuple, so the reference perturbs no hole
  (under NumPy >= 2 the comparison raises).  Here the intended rule is
  applied: a hole whose (|x|, |y|) is one of the four listed sites gets the
  perturbed radius.

Sorting: the reference sorts sites by x with ``argsort``'s default
(unstable) kind, so the order of sites sharing an x is unspecified; here the
sort is stable on (x, y).
"""
from __future__ import annotations

from typing import Sequence

import numpy

SQRT3_2 = numpy.sqrt(3.0) / 2.0


def _sort_xy(xy: numpy.ndarray) -> numpy.ndarray:
    return xy[numpy.lexsort((xy[:, 1], xy[:, 0]))]


def triangular_lattice(dims: Sequence[int], asymmetrical: bool = False) -> numpy.ndarray:
    """Sites ``[[x0, y0], ...]`` of a triangular lattice of unit pitch centred
    on (0, 0) (``pcgen.py:11-48``): ``dims[1]`` rows at y = j * sqrt(3)/2 for
    j = -(dims[1]-1)/2 .. (dims[1]-1)/2; a row holds ``dims[0]`` sites at
    x = -(dims[0]-1)/2 .. (dims[0]-1)/2, shifted by 1/2 on odd rows, where
    the symmetric lattice drops the last site so that it stays symmetric
    about x = 0 (``asymmetrical=True`` keeps it)."""
    nx, ny = int(dims[0]), int(dims[1])
    drop = 0 if asymmetrical else 1
    xmax = (nx - 1) / 2.0
    ymax = (ny - 1) / 2.0
    rows = []
    for j in numpy.linspace(-ymax, ymax, ny):
        odd = numpy.floor(j) % 2
        n_row = nx - int(drop * odd)
        xs = numpy.linspace(-xmax, xmax - drop * odd, n_row) + 0.5 * odd
        rows.append(numpy.stack([xs, numpy.full_like(xs, j * SQRT3_2)], axis=1))
    return _sort_xy(numpy.concatenate(rows))


def square_lattice(dims: Sequence[int]) -> numpy.ndarray:
    """Sites of a unit square lattice of ``dims[0] x dims[1]`` points,
    x = i - dims[0]/2 and y = j - dims[1]/2 (``pcgen.py:51-63``)."""
    xs, ys = numpy.meshgrid(numpy.arange(dims[0]) - dims[0] / 2.0, numpy.arange(dims[1]) - dims[1] / 2.0,
                            indexing="xy")
    return _sort_xy(numpy.stack([xs.ravel(), ys.ravel()], axis=1))


def nanobeam_holes(a_defect: float, num_defect_holes: int, num_mirror_holes: int) -> numpy.ndarray:
    """``[[x, r], ...]`` of a nanobeam cavity (``pcgen.py:68-90``): per side,
    ``num_defect_holes`` holes whose spacing and radius grow linearly from
    ``a_defect`` towards 1 (the spacing of the first is halved, since the
    mirror image doubles it), then ``num_mirror_holes`` holes of spacing and
    radius 1; mirrored about x = 0."""
    a = numpy.linspace(a_defect, 1.0, num_defect_holes, endpoint=False)
    xs = numpy.cumsum(a) - a[0] / 2.0
    mx = xs[-1] + numpy.arange(1, num_mirror_holes + 1)
    mr = numpy.ones_like(mx)
    x = numpy.concatenate([-mx[::-1], -xs[::-1], xs, mx])
    r = numpy.concatenate([mr[::-1], a[::-1], a, mr])
    return numpy.stack([x, r], axis=1)


def ln_defect(mirror_dims: Sequence[int], defect_length: int) -> numpy.ndarray:
    """Ln cavity (``pcgen.py:93-108``): a symmetric triangular lattice of
    (2 m + 1) sites per direction with the ``defect_length`` (odd) sites of
    the centre row nearest x = 0 removed."""
    if defect_length % 2 != 1:
        raise Exception("defect_length must be odd!")
    xy = triangular_lattice([2 * d + 1 for d in mirror_dims])
    h = defect_length // 2
    gone = (xy[:, 1] == 0) & (numpy.abs(xy[:, 0]) <= h) & (xy[:, 0] == numpy.round(xy[:, 0]))
    return xy[~gone]


def _expand(v, n):
    out = numpy.ones(n)
    v = numpy.atleast_1d(numpy.asarray(v, dtype=float))
    out[:v.size] = v
    return out


def ln_shift_defect(mirror_dims: Sequence[int], defect_length: int,
                    shifts_a: Sequence[float] = (0.15, 0, 0.075),
                    shifts_r: Sequence[float] = (1, 1, 1)) -> numpy.ndarray:
    """Ln cavity with the first holes beside the defect moved outwards by
    ``shifts_a[k]`` (in lattice constants) and their radius scaled by
    ``shifts_r[k]`` (``pcgen.py:111-158``).  Returns ``[[x, y, r], ...]``
    with r = 1 elsewhere; the shorter list is padded with ones."""
    if shifts_a is not None and not hasattr(shifts_a, "__len__"):
        shifts_a = [shifts_a]
    if shifts_r is not None and not hasattr(shifts_r, "__len__"):
        shifts_r = [shifts_r]
    xy = ln_defect(mirror_dims, defect_length)
    xyr = numpy.concatenate([xy, numpy.ones((xy.shape[0], 1))], axis=1)
    n = max(len(shifts_a), len(shifts_r))
    sa, sr = _expand(shifts_a, n), _expand(shifts_r, n)
    h = defect_length // 2
    for k in range(n):
        for sign in (-1.0, 1.0):
            x0 = sign * (h + k + 1)
            hit = (xyr[:, 0] == x0) & (xyr[:, 1] == 0)
            xyr[hit] = (x0 + sign * sa[k], 0.0, sr[k])
    return xyr


def r6_defect(mirror_dims: Sequence[int]) -> numpy.ndarray:
    """R6 (ring of six removed neighbours of the centre) cavity
    (``pcgen.py:161-179``)."""
    xy = triangular_lattice([2 * d + 1 for d in mirror_dims])
    ring = numpy.array([[1.0, 0.0], [0.5, SQRT3_2], [0.5, -SQRT3_2]])
    ring = numpy.concatenate([ring, -ring])
    keep = numpy.ones(xy.shape[0], bool)
    for p in ring:
        keep &= numpy.any(xy != p, axis=1)
    return xy[keep]


def l3_shift_perturbed_defect(mirror_dims: Sequence[int], perturbed_radius: float = 1.1,
                              shifts_a: Sequence[float] = (), shifts_r: Sequence[float] = ()) -> numpy.ndarray:
    """L3 cavity with shifted holes and four beam-forming holes of radius
    ``perturbed_radius`` (``pcgen.py:182-213``): with xs the distinct
    nonzero |x| and ys the distinct |y| of all holes, the holes at
    (|x|, |y|) = (xs[3], ys[1]), (xs[7], ys[1]), (xs[2], ys[2]), (xs[6],
    ys[2]) (all four sign combinations) are perturbed."""
    xyr = ln_shift_defect(mirror_dims, 3, shifts_a, shifts_r)
    ax, ay = numpy.abs(xyr[:, 0]), numpy.abs(xyr[:, 1])
    xs = numpy.unique(ax[ax != 0])
    ys = numpy.unique(ay)
    for a, b in ((3, 1), (7, 1), (2, 2), (6, 2)):
        if a < xs.size and b < ys.size:
            xyr[(ax == xs[a]) & (ay == ys[b]), 2] = perturbed_radius
    return xyr


def perturbed_l3(a: float, radius: float, xy_size: Sequence[int] = (10, 10), perturbed_radius: float = 1.1,
                 shifts_a: Sequence[float] = (0.15, 0, 0.075), shifts_r: Sequence[float] = (1.0, 1.0, 1.0)
                 ) -> numpy.ndarray:
    """The hole list of the example's perturbed L3 cavity (``fdtd.py:28-81``)
    as ``[[x, y, r], ...]`` in the units of ``a``: positions scaled by the
    lattice constant ``a``, radii by ``a * radius``.  (The example also draws
    two undercut trenches through masque; they lie outside the simulated
    slab's hole pattern and are left out.)"""
    xyr = l3_shift_perturbed_defect(xy_size, perturbed_radius, shifts_a, shifts_r).copy()
    xyr *= a
    xyr[:, 2] *= radius
    return xyr


def draw_holes(shape_xy: Sequence[int], holes: numpy.ndarray, center_xy: Sequence[float] = None,
               pitch: float = 1.0, radius: float = None) -> numpy.ndarray:
    """Boolean (sx, sy) mask of the cells whose centres lie inside a hole.
    ``holes`` is ``[[x, y], ...]`` (radius ``radius``) or ``[[x, y, r],
    ...]``, in lattice units scaled by ``pitch`` cells; cell (i, j) has its
    centre at (i - cx, j - cy) cells from ``center_xy`` (default: the grid
    centre cell, sx // 2, sy // 2).  A cell is inside when its squared
    distance to the hole centre is <= r^2."""
    sx, sy = int(shape_xy[0]), int(shape_xy[1])
    cx, cy = (sx // 2, sy // 2) if center_xy is None else center_xy
    xs = numpy.arange(sx, dtype=float)[:, None] - cx
    ys = numpy.arange(sy, dtype=float)[None, :] - cy
    mask = numpy.zeros((sx, sy), bool)
    holes = numpy.asarray(holes, dtype=float)
    for h in holes:
        hx, hy = h[0] * pitch, h[1] * pitch
        r = (h[2] * pitch) if holes.shape[1] > 2 else radius
        if r is None:
            raise Exception("radius is required for [x, y] hole lists")
        x0, x1 = int(numpy.floor(hx + cx - r)), int(numpy.ceil(hx + cx + r))
        y0, y1 = int(numpy.floor(hy + cy - r)), int(numpy.ceil(hy + cy + r))
        if x1 < 0 or y1 < 0 or x0 >= sx or y0 >= sy:
            continue
        x0, y0 = max(x0, 0), max(y0, 0)
        x1, y1 = min(x1 + 1, sx), min(y1 + 1, sy)
        sub = (xs[x0:x1] - hx) ** 2 + (ys[:, y0:y1] - hy) ** 2 <= r * r
        mask[x0:x1, y0:y1] |= sub
    return mask


def slab_epsilon(shape: Sequence[int], hole_mask: numpy.ndarray, z0: int, thickness: int,
                 eps_slab: Sequence[float], eps_background: float = 1.0):
    """Per-axis permittivity grids ``[eps_xx, eps_yy, eps_zz]`` of a slab of
    ``thickness`` cells starting at z = ``z0``, permittivity ``eps_slab``
    (one value per axis) outside the holes and ``eps_background`` in them and
    around the slab."""
    out = []
    for val in eps_slab:
        e = numpy.full(tuple(shape), float(eps_background))
        layer = numpy.where(hole_mask, float(eps_background), float(val))
        e[:, :, z0:z0 + thickness] = layer[:, :, None]
        out.append(e)
    return out
