"""Synthetic workloads of the benchmark configurations (host-side builders).

Each builder returns a ``Workload``: the constructor arguments of
``Simulation`` plus the soft source of the caller loop (``fdtd.py:136-143,
162-175``).  All randomness comes from ``numpy.random.default_rng(seed)``.

  small      64^3, 20 random eps boxes, CPML-8 on 6 faces, Ez point source
  phc_slab   395x345x73 photonic-crystal slab with an L3 defect, CPML-8,
             Ey dipole (the README-scale run, fdtd.py:84-175)
  bloch      64x64x256 Bloch-periodic supercell, CPML-16 on z, PEC mirror
  adjoint    128^3, CPML-8, 32^3 design region
  learned    256x256x128 interior with absorbing padding of a given width
  large      512^3, the small configuration's construction (HBM-scale grid)
"""
from __future__ import annotations

import dataclasses
from typing import Callable, Dict, List, Optional, Tuple

import numpy


@dataclasses.dataclass
class Workload:
    name: str
    shape: Tuple[int, int, int]
    epsilon: List[numpy.ndarray]
    pmls: List[Dict]
    steps: int
    dt: float
    bloch_boundaries: List[Dict] = dataclasses.field(default_factory=list)
    do_poynting: bool = True
    src_index: Optional[int] = None          # flat index into E (component offset included)
    src_wave: Optional[numpy.ndarray] = None  # float64 per-step source values
    plane_source: Optional[Dict] = None       # plane-wave source (bloch config)

    @property
    def cells(self) -> int:
        return int(numpy.prod(self.shape))


def pml_faces(thickness: int, axes: str = "xyz") -> List[Dict]:
    return [{"axis": a, "polarity": p, "thickness": thickness} for a in axes for p in "np"]


def gaussian_pulse(steps: int, dt: float, wl: float, dwl: float, phase: float = 0.0) -> numpy.ndarray:
    """The source of fdtd.py:136-143 with dx = 1: sin(w t0 + phase) exp(-alpha t0^2),
    t0 = n dt - delay."""
    w = 2 * numpy.pi / wl
    fwhm = dwl * w * w / (2 * numpy.pi)
    alpha = (fwhm ** 2) / 8 * numpy.log(2)
    delay = 7 / numpy.sqrt(2 * alpha)
    t0 = numpy.arange(steps) * dt - delay
    return numpy.sin(w * t0 + phase) * numpy.exp(-alpha * t0 ** 2)


def modulated_sine(steps: int, dt: float, wl: float, width: float, n0: float) -> numpy.ndarray:
    """sin(2 pi / wl * (n - n0) * dt) * exp(-((n - n0) / width)^2 / 2)."""
    n = numpy.arange(steps, dtype=numpy.float64)
    return numpy.sin(2 * numpy.pi / wl * (n - n0) * dt) * numpy.exp(-0.5 * ((n - n0) / width) ** 2)


def _index(shape, comp, x, y, z) -> int:
    return int(numpy.ravel_multi_index((x, y, z), shape, order="C")) + comp * int(numpy.prod(shape))


def small(seed: int = 1, steps: int = 1000, n: int = 64) -> Workload:
    """64^3, eps=1 background with 20 axis-aligned boxes (side 4-16, eps~U[2,12]),
    CPML-8 on all faces, dt = 0.99/sqrt(3), soft Ez source at the centre."""
    rng = numpy.random.default_rng(seed)
    shape = (n, n, n)
    eps = numpy.ones(shape)
    for _ in range(20):
        side = rng.integers(4, 17, size=3)
        lo = [int(rng.integers(0, n - s + 1)) for s in side]
        val = rng.uniform(2, 12)
        eps[lo[0]:lo[0] + side[0], lo[1]:lo[1] + side[1], lo[2]:lo[2] + side[2]] = val
    dt = 0.99 / numpy.sqrt(3)
    return Workload("small_64", shape, [eps.copy(), eps.copy(), eps.copy()], pml_faces(8), steps, dt,
                    src_index=_index(shape, 2, n // 2, n // 2, n // 2),
                    src_wave=modulated_sine(steps, dt, 20.0, 30.0, 120.0))


def phc_slab(seed: int = 2, steps: int = 8000, shape=(395, 345, 73), pitch: float = 20.0,
             radius: float = 7.0, thickness: int = 11) -> Workload:
    """Photonic-crystal slab: 11-cell slab, (eps_xx, eps_yy, eps_zz) =
    (12.25, 12.25, 11.9), air holes of radius 7 on a triangular lattice of
    pitch 20 with an L3 defect (3 holes removed on the centre row); CPML-8;
    Ey dipole at the defect centre with the fdtd.py pulse at wavelength 45."""
    from .geometry import draw_holes, ln_defect, slab_epsilon
    rng = numpy.random.default_rng(seed)
    sx, sy, sz = shape
    cx, cy, cz = sx // 2, sy // 2, sz // 2
    z0 = cz - thickness // 2
    # lattice sites (in pitches) reaching past every cell of the grid
    mx = int(numpy.ceil((sx / 2 + radius) / pitch)) + 1
    my = int(numpy.ceil((sy / 2 + radius) / (pitch * numpy.sqrt(3) / 2))) + 1
    holes = draw_holes((sx, sy), ln_defect((mx, my), 3), (cx, cy), pitch=pitch, radius=radius)
    eps = slab_epsilon(shape, holes, z0, thickness, (12.25, 12.25, 11.9))
    dt = 0.99 / numpy.sqrt(3)
    wl = 45.0
    phase = float(rng.uniform(0, 2 * numpy.pi))
    return Workload("phc_slab_395x345x73", shape, eps, pml_faces(8), steps, dt,
                    src_index=_index(shape, 1, cx, cy, cz),
                    src_wave=gaussian_pulse(steps, dt, wl, wl * 200 / 1550, phase))


def bloch(seed: int = 3, steps: int = 4000, shape=(64, 64, 256)) -> Workload:
    """Bloch-periodic supercell: phases on x and y with k ~ U[0, pi/64] per
    cell, CPML-16 on the z faces, eps ~ U[1, 9] per 8^3 block, a 5-plane PEC
    mirror (eps = inf) at z = 200..204, plane-wave soft source at z = 24."""
    rng = numpy.random.default_rng(seed)
    sx, sy, sz = shape
    blocks = rng.uniform(1, 9, size=(sx // 8, sy // 8, sz // 8))
    eps = numpy.kron(blocks, numpy.ones((8, 8, 8)))
    eps[:, :, 200:205] = numpy.inf
    kx, ky = rng.uniform(0, numpy.pi / 64, size=2)
    bl = [{"axis": "x", "real": float(numpy.cos(kx * sx)), "imag": float(numpy.sin(kx * sx))},
          {"axis": "y", "real": float(numpy.cos(ky * sy)), "imag": float(numpy.sin(ky * sy))}]
    dt = 0.99 / numpy.sqrt(3)
    phase = numpy.exp(1j * (kx * numpy.arange(sx)[:, None] + ky * numpy.arange(sy)[None, :]))
    wave = modulated_sine(steps, dt, 20.0, 30.0, 120.0)
    return Workload("bloch_64x64x256", shape, [eps.copy(), eps.copy(), eps.copy()],
                    pml_faces(16, "z"), steps, dt, bloch_boundaries=bl,
                    plane_source={"comp": 0, "z": 24, "phase": phase, "wave": wave, "k": (kx, ky)})


def adjoint(seed: int = 4, steps: int = 2000, n: int = 128) -> Workload:
    """128^3, CPML-8, eps = 1 outside a 32^3 design region with eps ~ U[1, 4]."""
    rng = numpy.random.default_rng(seed)
    shape = (n, n, n)
    eps = numpy.ones(shape)
    c = n // 2
    eps[c - 16:c + 16, c - 16:c + 16, 40:72] = rng.uniform(1, 4, size=(32, 32, 32))
    dt = 0.99 / numpy.sqrt(3)
    return Workload("adjoint_128", shape, [eps.copy(), eps.copy(), eps.copy()], pml_faces(8), steps, dt,
                    src_index=_index(shape, 0, c, c, 20),
                    src_wave=modulated_sine(steps, dt, 20.0, 30.0, 120.0))


def learned(pad: int, seed: int = 5, steps: int = 500, interior=(256, 256, 128),
            poynting: bool = False) -> Workload:
    """256x256x128 interior (eps ~ U[1,4] in 8^3 blocks near the faces),
    padded by ``pad`` absorbing cells on every face (CPML of that thickness)."""
    rng = numpy.random.default_rng(seed)
    ix, iy, iz = interior
    shape = (ix + 2 * pad, iy + 2 * pad, iz + 2 * pad)
    blocks = rng.uniform(1, 4, size=(ix // 8, iy // 8, iz // 8))
    inner = numpy.kron(blocks, numpy.ones((8, 8, 8)))
    core = numpy.ones(interior, bool)
    core[16:-16, 16:-16, 16:-16] = False
    inner = numpy.where(core, inner, 1.0)
    eps = numpy.ones(shape)
    eps[pad:pad + ix, pad:pad + iy, pad:pad + iz] = inner
    dt = 0.99 / numpy.sqrt(3)
    return Workload("learned_pad%d" % pad, shape, [eps.copy(), eps.copy(), eps.copy()],
                    pml_faces(pad) if pad > 0 else [], steps, dt, do_poynting=poynting,
                    src_index=_index(shape, 2, pad + ix // 2, pad + iy // 2, pad + iz // 2),
                    src_wave=modulated_sine(steps, dt, 20.0, 30.0, 120.0))


def large(seed: int = 6, steps: int = 200, n: int = 512) -> Workload:
    """Synthetic 512^3 grid (134 M cells, 1.6 GB per field): the small
    configuration's construction (eps = 1 with 20 boxes of eps ~ U[2, 12],
    CPML-8 on all faces, Ez point source at the centre) at a size far beyond
    the 256 MB Infinity Cache, for an HBM figure free of cache effects."""
    w = small(seed=seed, steps=steps, n=n)
    w.name = "large_%d" % n
    return w


BUILDERS: Dict[str, Callable[..., Workload]] = {
    "small": small, "phc_slab": phc_slab, "bloch": bloch, "adjoint": adjoint, "large": large,
}


def algorithmic_bytes_per_cell(shape, pmls, dtype_bytes=4, poynting=True, bloch=False) -> float:
    """Compulsory HBM bytes per cell-update of one fused step (E kernel + H/S kernel):
    E r+w, H r, eps r (E step, 12 words); E r, H r, H w (H step, 9 words);
    S w (3 words, when Poynting is on); per CPML slab cell and face, psi_E and
    psi_H read and written (8 words).  Bloch pairs double the E/H terms."""
    words = 12 + 9
    if bloch:
        words += 9 + 9
    if poynting:
        words += 3
    n = float(numpy.prod(shape))
    slab = 0.0
    for p in pmls:
        r = "xyz".find(p["axis"])
        T = min(int(p.get("thickness", 8)), shape[r])
        slab += T * n / shape[r]
    per_face = 8 * (2 if bloch else 1)
    return dtype_bytes * (words + per_face * slab / n)
