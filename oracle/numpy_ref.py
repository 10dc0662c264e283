"""Vectorised brute-force restatement of the FDTD step (TEST INFRASTRUCTURE ONLY).

An independent second restatement of the reference step, written with whole
array operations (``numpy.roll`` for the periodic wrap of ``common.cl:75-85``,
boolean slab masks for the CPML faces of ``update_e.cl:64-100`` /
``update_h.cl:80-116``) instead of the per-cell index arithmetic of
``fdtd_oracle.cpp``.  Each elementwise operation is one IEEE operation in the
array dtype, so in float32 it reproduces the per-cell restatement bit for bit,
and in float64 it is the brute-force reference for the invariant tests.

Fields are arrays of shape (3, sx, sy, sz).  Only for small grids.
"""
from __future__ import annotations

import numpy

from .oracle import pml_coefficients, pml_defaults


def _uv(r):
    return [a for a in range(3) if a != r]


class NumpyFDTD:
    def __init__(self, eps, pmls=(), dt=None, dx=1.0, dtype=numpy.float64,
                 bloch=(), do_poynting=True):
        self.dtype = numpy.dtype(dtype).type
        self.eps = numpy.asarray(eps, dtype=self.dtype)
        self.shape = self.eps.shape[1:]
        self.dt = dt if dt is not None else dx * .99 / numpy.sqrt(3)
        self.inv = self.dtype(1.0 / dx)
        self.E = numpy.zeros((3,) + self.shape, self.dtype)
        self.H = numpy.zeros_like(self.E)
        self.F = numpy.zeros_like(self.E)
        self.G = numpy.zeros_like(self.E)
        self.S = numpy.zeros_like(self.E)
        self.do_poynting = do_poynting
        # Bloch phase factors per axis (c, s); the seam neighbour is
        # c*own + b*partner with b chosen per step type.
        self.bloch = {"xyz".find(b["axis"]): (self.dtype(b["real"]), self.dtype(b["imag"])) for b in bloch}
        self.faces = []
        self.faces_b = []
        for p in pmls:
            p = pml_defaults(p)
            r = "xyz".find(p["axis"])
            T = p["thickness"]
            s = self.shape[r]
            coords = numpy.arange(s)
            if p["polarity"] == "n":
                mask1 = coords < T
                ir = numpy.where(mask1, coords, 0)
            else:
                mask1 = (coords >= s - T) if T <= s else numpy.zeros(s, bool)
                ir = numpy.where(mask1, s - 1 - coords, 0)
            coefs = pml_coefficients(p, self.dt, self.dtype)
            bshape = [1, 1, 1]
            bshape[r] = s
            mask = mask1.reshape(bshape)
            full = [numpy.where(mask1, c[ir], 0).astype(self.dtype).reshape(bshape) for c in coefs]
            for lst in (self.faces, self.faces_b):
                lst.append(dict(r=r, mask=numpy.broadcast_to(mask, self.shape), c=full,
                                psi=[numpy.zeros(self.shape, self.dtype) for _ in range(4)]))

    def _back(self, A, P, b, c, s):
        """A at i - e_b with periodic wrap; Bloch seam mix at coordinate 0."""
        out = numpy.roll(A, 1, axis=b)
        if b in self.bloch and P is not None:
            sl = [slice(None)] * 3
            sl[b] = 0
            sl = tuple(sl)
            out[sl] = c * out[sl] + s * numpy.roll(P, 1, axis=b)[sl]
        return out

    def _fwd(self, A, P, b, c, s):
        out = numpy.roll(A, -1, axis=b)
        if b in self.bloch and P is not None:
            sl = [slice(None)] * 3
            sl[b] = -1
            sl = tuple(sl)
            out[sl] = c * out[sl] + s * numpy.roll(P, -1, axis=b)[sl]
        return out

    def _e(self, E, H, P, faces, sign):
        dt = self.dtype(self.dt)
        dH = [[None] * 3 for _ in range(3)]
        for a in range(3):
            for b in range(3):
                if a == b:
                    continue
                cb, sb = self.bloch.get(b, (1, 0))
                Hb = self._back(H[a], P[a] if P is not None else None, b, cb, sign * sb)
                dH[a][b] = (H[a] - Hb) * self.inv
        pE = [numpy.zeros(self.shape, self.dtype) for _ in range(3)]
        for f in faces:
            r, m = f["r"], f["mask"]
            u, v = _uv(r)
            p0, p1, p2 = f["c"][:3]
            dH[v][r] = numpy.where(m, dH[v][r] * p2, dH[v][r])
            dH[u][r] = numpy.where(m, dH[u][r] * p2, dH[u][r])
            f["psi"][0] = numpy.where(m, p0 * f["psi"][0] + p1 * dH[v][r], f["psi"][0])
            f["psi"][1] = numpy.where(m, p0 * f["psi"][1] + p1 * dH[u][r], f["psi"][1])
            zu = numpy.where(m, f["psi"][0], 0)
            zv = numpy.where(m, f["psi"][1], 0)
            if r != 1:
                pE[u] = numpy.where(m, pE[u] - zu, pE[u]); pE[v] = numpy.where(m, pE[v] + zv, pE[v])
            else:
                pE[u] = numpy.where(m, pE[u] + zu, pE[u]); pE[v] = numpy.where(m, pE[v] - zv, pE[v])
        E[0] = E[0] + dt / self.eps[0] * (dH[2][1] - dH[1][2] + pE[0])
        E[1] = E[1] + dt / self.eps[1] * (dH[0][2] - dH[2][0] + pE[1])
        E[2] = E[2] + dt / self.eps[2] * (dH[1][0] - dH[0][1] + pE[2])

    def _h(self, E, H, P, faces, sign, want_s):
        dt = self.dtype(self.dt)
        dE = [[None] * 3 for _ in range(3)]
        Ef = [[None] * 3 for _ in range(3)]
        for a in range(3):
            for b in range(3):
                if a == b:
                    continue
                cb, sb = self.bloch.get(b, (1, 0))
                Ef[a][b] = self._fwd(E[a], P[a] if P is not None else None, b, cb, sign * sb)
                dE[a][b] = (Ef[a][b] - E[a]) * self.inv
        pH = [numpy.zeros(self.shape, self.dtype) for _ in range(3)]
        for f in faces:
            r, m = f["r"], f["mask"]
            u, v = _uv(r)
            p0, p1, p2 = f["c"][3:]
            dE[v][r] = numpy.where(m, dE[v][r] * p2, dE[v][r])
            dE[u][r] = numpy.where(m, dE[u][r] * p2, dE[u][r])
            f["psi"][2] = numpy.where(m, p0 * f["psi"][2] + p1 * dE[v][r], f["psi"][2])
            f["psi"][3] = numpy.where(m, p0 * f["psi"][3] + p1 * dE[u][r], f["psi"][3])
            zu = numpy.where(m, f["psi"][2], 0)
            zv = numpy.where(m, f["psi"][3], 0)
            if r != 1:
                pH[u] = numpy.where(m, pH[u] + zu, pH[u]); pH[v] = numpy.where(m, pH[v] - zv, pH[v])
            else:
                pH[u] = numpy.where(m, pH[u] - zu, pH[u]); pH[v] = numpy.where(m, pH[v] + zv, pH[v])
        H0 = H.copy()
        H[0] = H[0] - dt * (dE[2][1] - dE[1][2] - pH[0])
        H[1] = H[1] - dt * (dE[0][2] - dE[2][0] - pH[1])
        H[2] = H[2] - dt * (dE[1][0] - dE[0][1] - pH[2])
        if not want_s:
            return
        aE = [[None if a == b else Ef[a][b] + E[a] for b in range(3)] for a in range(3)]
        aH = H + H0
        oSxy = aE[1][0] * aH[2]
        oSxz = -aE[2][0] * aH[1]
        oSyz = aE[2][1] * aH[0]
        oSyx = -aE[0][1] * aH[2]
        oSzx = aE[0][2] * aH[1]
        oSzy = -aE[1][2] * aH[0]
        sf = dt * self.dtype(0.125)
        R = lambda A, b: numpy.roll(A, 1, axis=b)
        self.S[0] = (oSxy + oSxz + R(oSxy, 1) + R(oSxz, 2)) * sf
        self.S[1] = (oSyz + oSyx + R(oSyz, 2) + R(oSyx, 0)) * sf
        self.S[2] = (oSzx + oSzy + R(oSzx, 0) + R(oSzy, 1)) * sf

    def step(self, source=None, n=0):
        bl = bool(self.bloch)
        self._e(self.E, self.H, self.G if bl else None, self.faces, +1)
        if bl:
            self._e(self.F, self.G, self.H, self.faces_b, -1)
        if source is not None:
            source(self, n)
        self._h(self.E, self.H, self.F if bl else None, self.faces, -1, self.do_poynting)
        if bl:
            self._h(self.F, self.G, self.E, self.faces_b, +1, False)
