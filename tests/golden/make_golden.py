"""Writes the golden fixtures of the FDTD step from the CPU oracle.

The reference ships no tests or golden vectors, so these fixtures are the
oracle's outputs on small seeded inputs that exercise every feature (CPML on
all faces, PEC, non-uniform spacing, Bloch pairs, Poynting, the point
source).  The oracle itself is pinned by tests/test_oracle_kat.py.  Each
fixture stores its inputs and the fields after the given number of steps.

    python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.oracle import OracleSimulation  # noqa: E402


def cases():
    """name -> (constructor kwargs, steps, source (index, wave) or None)."""
    out = {}
    rng = numpy.random.default_rng(101)
    shape = (8, 7, 6)
    eps = [rng.uniform(1, 4, shape) for _ in range(3)]
    init = {"E": [rng.standard_normal(shape) for _ in range(3)],
            "H": [rng.standard_normal(shape) for _ in range(3)]}
    out["cpml6_f32"] = (dict(epsilon=eps, pmls=[{"axis": a, "polarity": p, "thickness": 2}
                                                for a in "xyz" for p in "np"],
                             initial_fields=init, float_type=numpy.float32), 30,
                        (2 * 336 + 3 * 42 + 3 * 6 + 2, 0.01 * numpy.sin(numpy.arange(30) * 0.4)))

    rng = numpy.random.default_rng(102)
    eps = [rng.uniform(1, 3, shape) for _ in range(3)]
    init = {k: [rng.standard_normal(shape) for _ in range(3)] for k in "EHFG"}
    bl = [{"axis": "x", "real": math.cos(0.8), "imag": math.sin(0.8)},
          {"axis": "y", "real": math.cos(-1.3), "imag": math.sin(-1.3)}]
    out["bloch_f32"] = (dict(epsilon=eps, pmls=[{"axis": "z", "polarity": p, "thickness": 2} for p in "np"],
                             bloch_boundaries=bl, initial_fields=init, float_type=numpy.float32), 20, None)

    rng = numpy.random.default_rng(103)
    eps = [rng.uniform(1, 4, shape) for _ in range(3)]
    mask = rng.random(shape) < 0.15
    for e in eps:
        e[mask] = numpy.inf
    init = {"E": [numpy.where(mask, 0.0, rng.standard_normal(shape)) for _ in range(3)],
            "H": [rng.standard_normal(shape) for _ in range(3)]}
    dxes = [[rng.uniform(0.8, 1.2, s) for s in shape], [rng.uniform(0.8, 1.2, s) for s in shape]]
    out["pec_nonuniform_f64"] = (dict(epsilon=eps, pmls=[{"axis": "y", "polarity": p, "thickness": 2,
                                                          "cfs_alpha": 0.05} for p in "np"],
                                      dxes=dxes, initial_fields=init, float_type=numpy.float64), 25, None)
    return out


def run_case(kw, steps, src):
    o = OracleSimulation(**kw)
    for t in range(steps):
        o.update_E()
        if o.bloch:
            o.update_F()
        if src is not None:
            o.E[src[0]] += o.arg_type(src[1][t])
        o.update_H()
        if o.bloch:
            o.update_G()
        o.update_S()
    out = {"E": o.E, "H": o.H, "S": o.S}
    if o.bloch:
        out["F"] = o.F
        out["G"] = o.G
    return out


def main():
    for name, (kw, steps, src) in cases().items():
        res = run_case(kw, steps, src)
        path = os.path.join(HERE, name + ".npz")
        numpy.savez_compressed(path, steps=steps, **{"out_" + k: v for k, v in res.items()})
        print("wrote", path, {k: float(numpy.abs(v).max()) for k, v in res.items()})


if __name__ == "__main__":
    main()
