"""CPML face parameters and coefficients.

Restates ``opencl_fdtd/simulation.py:161-168`` (defaults) and ``:279-304``
(coefficients).  The arithmetic is done in the field dtype: the reference was
written for NumPy 1.x, whose value-based casting kept a float32 array
combined with float64 scalars in float32.  (Under NumPy >= 2 the reference
would build float64 coefficient arrays and hand them to ``float*`` kernel
arguments, a defect that is not reproduced.)
"""
from __future__ import annotations

from typing import Dict, List

import numpy

PML_DEFAULTS = (
    ("thickness", 8),
    ("epsilon_eff", 1.0),
    ("mu_eff", 1.0),
    ("ln_R_per_layer", -1.6),
    ("m", 3.5),
    ("cfs_alpha", 0),
    ("ma", 1),
)


def with_defaults(pml: Dict) -> Dict:
    """Fill the face dict in place with the defaults of simulation.py:161-168
    (the reference mutates the caller's dicts the same way) and validate it."""
    for k, v in PML_DEFAULTS:
        pml.setdefault(k, v)
    if pml.get("axis") not in ("x", "y", "z"):
        raise Exception("PML axis must be one of 'x', 'y', 'z', got {!r}".format(pml.get("axis")))
    if pml.get("polarity") not in ("n", "p"):
        raise Exception("PML polarity must be 'n' or 'p', got {!r}".format(pml.get("polarity")))
    if int(pml["thickness"]) < 1:
        raise Exception("PML thickness must be >= 1, got {!r}".format(pml["thickness"]))
    return pml


def face_coefficients(pml: Dict, dt: float, dtype) -> List[numpy.ndarray]:
    """[p0e, p1e, p2e, p0h, p1h, p2h] for one face; index 0 = outermost layer.

    sigma_max = -ln_R/2 * (m + 1); kappa_max = sqrt(mu_eff * eps_eff);
    depth x = T, T-1, ..., 1 (minus 0.5 for E on 'p' faces and for H on 'n'
    faces); s = (x/T)^m; sigma = s*sigma_max; kappa = 1 + s*(kappa_max - 1);
    alpha = (1 - x/T)^ma * cfs_alpha; p0 = exp(-(sigma/kappa + alpha)*dt);
    p1 = sigma/(sigma + kappa*alpha) * (p0 - 1); p2 = 1/kappa.
    """
    t = numpy.dtype(dtype).type
    T = int(pml["thickness"])
    one = t(1)
    sigma_max = t(-pml["ln_R_per_layer"] / 2 * (pml["m"] + 1))
    # kappa_max is a float64 scalar in the reference (numpy.sqrt of Python
    # floats, simulation.py:285); kappa_max - 1 is formed between scalars, in
    # float64, and rounded once when it meets the float32 array (:291).
    km1 = t(numpy.float64(numpy.sqrt(pml["mu_eff"] * pml["epsilon_eff"])) - 1)
    alpha_max = t(pml["cfs_alpha"])
    m, ma, Tt, dtt = t(pml["m"]), t(pml["ma"]), t(T), t(dt)

    def par(x):
        scaling = (x / Tt) ** m
        sigma = scaling * sigma_max
        kappa = one + scaling * km1
        alpha = ((one - x / Tt) ** ma) * alpha_max
        with numpy.errstate(invalid="ignore", divide="ignore"):
            p0 = numpy.exp(-(sigma / kappa + alpha) * dtt)
            p1 = sigma / (sigma + kappa * alpha) * (p0 - one)
            p2 = one / kappa
        return [numpy.ascontiguousarray(a, dtype=t) for a in (p0, p1, p2)]

    xe = numpy.arange(1, T + 1, dtype=t)[::-1].copy()
    xh = xe.copy()
    if pml["polarity"] == "p":
        xe -= t(0.5)
    else:
        xh -= t(0.5)
    return par(xe) + par(xh)
