"""CPU oracle for the FDTD time-stepping hot path (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module; it is the checker, never the product.

``OracleSimulation`` mirrors the reference host class
``opencl_fdtd/simulation.py:25-357`` on host numpy arrays, and steps them with
``liboracle_fdtd.so`` (``oracle/fdtd_oracle.cpp``), which restates the
per-cell template bodies in the reference's operation order.

Coefficients follow ``simulation.py:274-319`` with the arithmetic done in the
field dtype, as NumPy 1.x value-based casting did for the reference
(float32 arrays combined with float64 scalars stay float32).  The reference's
own behaviour under NumPy >= 2 (float64 coefficient arrays handed to ``float*``
kernel arguments) is a defect, not a behaviour to reproduce.
"""
from __future__ import annotations

import ctypes
import os
import warnings
from typing import Dict, List, Optional, Sequence

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_fdtd.so")
MAX_FACES = 12


class OracleFace(ctypes.Structure):
    _fields_ = [
        ("axis", ctypes.c_int32), ("polarity", ctypes.c_int32),
        ("thickness", ctypes.c_int32), ("pad_", ctypes.c_int32),
        ("p0e", ctypes.c_void_p), ("p1e", ctypes.c_void_p), ("p2e", ctypes.c_void_p),
        ("p0h", ctypes.c_void_p), ("p1h", ctypes.c_void_p), ("p2h", ctypes.c_void_p),
        ("psi_eu", ctypes.c_void_p), ("psi_ev", ctypes.c_void_p),
        ("psi_hu", ctypes.c_void_p), ("psi_hv", ctypes.c_void_p),
    ]


class OracleGrid(ctypes.Structure):
    _fields_ = [
        ("sx", ctypes.c_int64), ("sy", ctypes.c_int64), ("sz", ctypes.c_int64),
        ("dtype", ctypes.c_int32), ("n_face", ctypes.c_int32),
        ("face", OracleFace * MAX_FACES),
        ("inv_d_uniform", ctypes.c_double),
        ("inv_de", ctypes.c_void_p * 3), ("inv_dh", ctypes.c_void_p * 3),
        ("nthreads", ctypes.c_int32), ("pad2_", ctypes.c_int32),
    ]


_lib = None


def load_library() -> ctypes.CDLL:
    """Load (building it first if needed) the oracle shared library."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    G = ctypes.POINTER(OracleGrid)
    D = ctypes.POINTER(ctypes.c_double)
    lib.oracle_update_e.argtypes = [G, P, P, P, P, ctypes.c_double, ctypes.c_int, D]
    lib.oracle_update_h.argtypes = [G, P, P, P, P, ctypes.c_double, ctypes.c_int, D]
    lib.oracle_update_s.argtypes = [G, P, P, ctypes.c_double]
    lib.oracle_update_j.argtypes = [G, P, P, P, ctypes.c_double, ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_int64)]
    for f in ("oracle_update_e", "oracle_update_h", "oracle_update_s", "oracle_update_j"):
        getattr(lib, f).restype = ctypes.c_int
    _lib = lib
    return lib


def build() -> None:
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _ptr(a: Optional[numpy.ndarray]) -> Optional[int]:
    return None if a is None else a.ctypes.data


def vec(fields: Sequence[numpy.ndarray]) -> numpy.ndarray:
    """Concatenation of C-order flattenings (the fdfd_tools.vec layout used at
    simulation.py:347, 357)."""
    return numpy.hstack([numpy.asarray(f).ravel(order="C") for f in fields])


def unvec(v: numpy.ndarray, shape) -> List[numpy.ndarray]:
    n = int(numpy.prod(shape))
    return [v[k * n:(k + 1) * n].reshape(shape, order="C") for k in range(3)]


def pml_coefficients(pml: Dict, dt, dtype) -> tuple:
    """CPML coefficients of one face, simulation.py:279-304.

    Returns (p0e, p1e, p2e, p0h, p1h, p2h), each of length thickness, index
    ir = 0 being the outermost layer.  All arithmetic happens in ``dtype``.
    """
    dtype = numpy.dtype(dtype).type
    T = pml["thickness"]
    sigma_max = dtype(-pml["ln_R_per_layer"] / 2 * (pml["m"] + 1))
    # simulation.py:285,291: kappa_max is a float64 scalar, kappa_max - 1 is
    # a scalar-scalar operation (float64), rounded to the array dtype once.
    km1 = dtype(numpy.float64(numpy.sqrt(pml["mu_eff"] * pml["epsilon_eff"])) - 1)
    alpha_max = dtype(pml["cfs_alpha"])
    m = dtype(pml["m"])
    ma = dtype(pml["ma"])
    Tt = dtype(T)
    dtt = dtype(dt)
    one = dtype(1)

    def par(x):
        scaling = (x / Tt) ** m
        sigma = scaling * sigma_max
        kappa = one + scaling * km1
        alpha = ((one - x / Tt) ** ma) * alpha_max
        with numpy.errstate(invalid="ignore", divide="ignore"):
            p0 = numpy.exp(-(sigma / kappa + alpha) * dtt)
            p1 = sigma / (sigma + kappa * alpha) * (p0 - one)
            p2 = one / kappa
        return [numpy.ascontiguousarray(a, dtype=dtype) for a in (p0, p1, p2)]

    xe = numpy.arange(1, T + 1, dtype=dtype)[::-1].copy()
    xh = numpy.arange(1, T + 1, dtype=dtype)[::-1].copy()
    if pml["polarity"] == "p":
        xe -= dtype(0.5)
    elif pml["polarity"] == "n":
        xh -= dtype(0.5)
    return tuple(par(xe)) + tuple(par(xh))


def pml_defaults(pml: Dict) -> Dict:
    """Defaults of simulation.py:161-168."""
    pml = dict(pml)
    pml.setdefault("thickness", 8)
    pml.setdefault("epsilon_eff", 1.0)
    pml.setdefault("mu_eff", 1.0)
    pml.setdefault("ln_R_per_layer", -1.6)
    pml.setdefault("m", 3.5)
    pml.setdefault("cfs_alpha", 0)
    pml.setdefault("ma", 1)
    return pml


class _PmlSet:
    """One system's CPML state (coefficients + psi), simulation.py:274-319."""

    def __init__(self, pmls, shape, dt, dtype):
        self.faces = []
        for pml in pmls:
            a = "xyz".find(pml["axis"])
            coefs = pml_coefficients(pml, dt, dtype)
            psi_shape = list(shape)
            psi_shape[a] = pml["thickness"]
            psi = [numpy.zeros(int(numpy.prod(psi_shape)), dtype=dtype) for _ in range(4)]
            self.faces.append((a, 0 if pml["polarity"] == "n" else 1, pml["thickness"], coefs, psi))

    def fill(self, grid: OracleGrid, e_or_h: str):
        grid.n_face = len(self.faces)
        for k, (a, pol, T, coefs, psi) in enumerate(self.faces):
            f = grid.face[k]
            f.axis, f.polarity, f.thickness = a, pol, T
            f.p0e, f.p1e, f.p2e = [_ptr(c) for c in coefs[:3]]
            f.p0h, f.p1h, f.p2h = [_ptr(c) for c in coefs[3:]]
            f.psi_eu, f.psi_ev, f.psi_hu, f.psi_hv = [_ptr(p) for p in psi]


class OracleSimulation:
    """Host mirror of ``opencl_fdtd.Simulation`` (simulation.py:77-271).

    Fields are flat numpy arrays of length 3N (component-major, C order inside
    each component).  ``bloch_mode`` selects the Bloch semantics:
    ``'phase'`` applies the phase factor at the seam (E/F and H/G coupled),
    ``'decoupled'`` reproduces the reference as shipped (two independent
    periodic systems; simulation.py:187-224 with the Bloch block unrendered).
    """

    def __init__(self, epsilon, pmls, bloch_boundaries=(), dxes=None, dt=None,
                 initial_fields=None, float_type=numpy.float32, do_poynting=True,
                 do_fieldsrc=False, bloch_mode="phase", nthreads=0):
        if initial_fields is None:
            initial_fields = {}
        self.shape = tuple(int(s) for s in epsilon[0].shape)
        self.arg_type = numpy.dtype(float_type).type
        if len(epsilon) != 3:
            raise Exception("Epsilon must be a list with length of 3")
        if not all(e.shape == epsilon[0].shape for e in epsilon[1:]):
            raise Exception("All epsilon grids must have the same shape")
        self.N = int(numpy.prod(self.shape))
        self.eps = vec(epsilon).astype(self.arg_type)
        if dxes is None:
            dxes = 1.0
        self.uniform = isinstance(dxes, (float, int))
        if self.uniform:
            min_dx = dxes
            self.inv_dxes = None
            inv_uniform = 1.0 / dxes
        else:
            self.inv_dxes = [(1 / numpy.asarray(d, dtype=numpy.float64)).astype(self.arg_type)
                             for d in list(dxes[0]) + list(dxes[1])]
            min_dx = min(min(d) for d in list(dxes[0]) + list(dxes[1]))
            inv_uniform = 1.0
        max_dt = min_dx * .99 / numpy.sqrt(3)
        if dt is None:
            self.dt = max_dt
        elif dt <= 0:
            raise Exception("Invalid dt: {}".format(dt))
        else:
            if dt > max_dt:
                warnings.warn("Warning: unstable dt: {}".format(dt))
            self.dt = dt

        def field(key):
            v = initial_fields.get(key)
            if v is None:
                return numpy.zeros(3 * self.N, dtype=self.arg_type)
            return vec(v).astype(self.arg_type)

        self.E = field("E")
        self.H = field("H")
        self.bloch = list(bloch_boundaries)
        self.bloch_mode = bloch_mode
        if self.bloch:
            self.F = field("F")
            self.G = field("G")
        self.pmls = [pml_defaults(p) for p in pmls]
        self.do_poynting = do_poynting
        if do_poynting:
            self.oS = numpy.zeros(6 * self.N, dtype=self.arg_type)
            self.S = numpy.zeros(3 * self.N, dtype=self.arg_type)
        else:
            self.oS = None
            self.S = None
        if do_fieldsrc:
            self.Jr = numpy.zeros(3 * self.N, dtype=self.arg_type)
            self.Ji = numpy.zeros(3 * self.N, dtype=self.arg_type)
        self._pml_main = _PmlSet(self.pmls, self.shape, self.dt, self.arg_type)
        self._pml_bloch = _PmlSet(self.pmls, self.shape, self.dt, self.arg_type) if self.bloch else None

        self._lib = load_library()
        self._grid = OracleGrid()
        g = self._grid
        g.sx, g.sy, g.sz = self.shape
        g.dtype = 0 if self.arg_type == numpy.float32 else 1
        g.inv_d_uniform = inv_uniform
        g.nthreads = nthreads
        if not self.uniform:
            for r in range(3):
                g.inv_de[r] = _ptr(self.inv_dxes[r])
                g.inv_dh[r] = _ptr(self.inv_dxes[3 + r])
        # Bloch seams: phase factor c + i s per axis
        self._seam_mask = 0
        c = [1.0, 1.0, 1.0]
        s = [0.0, 0.0, 0.0]
        for b in self.bloch:
            r = "xyz".find(b["axis"])
            self._seam_mask |= 1 << r
            c[r] = float(b["real"])
            s[r] = float(b["imag"])
        # E-step neighbour across the seam: (H + iG)(c - is); H-step: (E + iF)(c + is)
        arr = ctypes.c_double * 6
        self._seam_E = arr(*c, *s)                        # real system, E-step: Hc + Gs
        self._seam_F = arr(*c, *[-v for v in s])          # imag system, F-step: Gc - Hs
        self._seam_H = arr(*c, *[-v for v in s])          # real system, H-step: Ec - Fs
        self._seam_G = arr(*c, *s)                        # imag system, G-step: Fc + Es

    def _grid_for(self, pml_set):
        pml_set.fill(self._grid, "")
        return ctypes.byref(self._grid)

    def _mask(self):
        return self._seam_mask if (self.bloch and self.bloch_mode == "phase") else 0

    # --- the update callables -------------------------------------------------
    def update_E(self):
        rc = self._lib.oracle_update_e(self._grid_for(self._pml_main), _ptr(self.E), _ptr(self.H),
                                       _ptr(self.G) if self.bloch else None, _ptr(self.eps),
                                       self.dt, self._mask(), self._seam_E)
        assert rc == 0, rc

    def update_F(self):
        rc = self._lib.oracle_update_e(self._grid_for(self._pml_bloch), _ptr(self.F), _ptr(self.G),
                                       _ptr(self.H), _ptr(self.eps), self.dt, self._mask(), self._seam_F)
        assert rc == 0, rc

    def update_H(self):
        rc = self._lib.oracle_update_h(self._grid_for(self._pml_main), _ptr(self.E), _ptr(self.H),
                                       _ptr(self.F) if self.bloch else None,
                                       _ptr(self.oS) if self.do_poynting else None,
                                       self.dt, self._mask(), self._seam_H)
        assert rc == 0, rc

    def update_G(self):
        rc = self._lib.oracle_update_h(self._grid_for(self._pml_bloch), _ptr(self.F), _ptr(self.G),
                                       _ptr(self.E), None, self.dt, self._mask(), self._seam_G)
        assert rc == 0, rc

    def update_S(self):
        rc = self._lib.oracle_update_s(ctypes.byref(self._grid), _ptr(self.oS), _ptr(self.S), self.dt)
        assert rc == 0, rc

    def update_J(self, c, s, xmin, xmax, ymin, ymax, zmin, zmax):
        box = (ctypes.c_int64 * 6)(xmin, xmax, ymin, ymax, zmin, zmax)
        rc = self._lib.oracle_update_j(ctypes.byref(self._grid), _ptr(self.E), _ptr(self.Jr),
                                       _ptr(self.Ji), c, s, box)
        assert rc == 0, rc

    def step(self, source=None, n: int = 0):
        """One step in the order of the reference caller loop (fdtd.py:162-175):
        E (then F), the host soft source, H (then G), S."""
        self.update_E()
        if self.bloch:
            self.update_F()
        if source is not None:
            source(self, n)
        self.update_H()
        if self.bloch:
            self.update_G()
        if self.do_poynting:
            self.update_S()
