"""Writes tests/golden/adjoint_fd_f64.npz: the Poynting-flux objective of a
small problem and its central finite-difference eps-gradient at a few design
cells, computed with the FP64 CPU oracle (test infrastructure only).

    python tests/golden/make_adjoint_golden.py
"""
import os
import sys

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.oracle import OracleSimulation, unvec  # noqa: E402


def problem():
    shape = (12, 10, 14)
    rng = numpy.random.default_rng(201)
    eps = numpy.ones((3,) + shape)
    design = (slice(4, 8), slice(3, 7), slice(5, 9))
    block = rng.uniform(1, 4, size=(4, 4, 4))
    for a in range(3):
        eps[(a,) + design] = block
    pmls = [{"axis": a, "polarity": p, "thickness": 3} for a in "xyz" for p in "np"]
    steps = 80
    dt = 0.99 / numpy.sqrt(3)
    n = numpy.arange(steps)
    wave = numpy.sin(2 * numpy.pi / 10 * n * dt) * numpy.exp(-0.5 * ((n - 20) / 6.0) ** 2)
    N = int(numpy.prod(shape))
    src = 2 * N + numpy.ravel_multi_index((6, 5, 3), shape)
    cells = [(4, 3, 5), (5, 4, 6), (7, 6, 8), (6, 5, 7), (4, 6, 8)]
    return dict(shape=shape, eps=eps, design=design, pmls=pmls, steps=steps, dt=dt, wave=wave, src=src,
                axis=2, index=10, cells=cells)


def objective(pr, eps):
    o = OracleSimulation([eps[0], eps[1], eps[2]], [dict(p) for p in pr["pmls"]], dt=pr["dt"],
                         float_type=numpy.float64)
    J = 0.0
    for t in range(pr["steps"]):
        o.update_E()
        o.E[pr["src"]] += pr["wave"][t]
        o.update_H()
        o.update_S()
        Sz = unvec(o.S, pr["shape"])[2]
        J += float(Sz[:, :, pr["index"]].sum())
    return J


def main():
    pr = problem()
    J0 = objective(pr, pr["eps"])
    grads = []
    for c in pr["cells"]:
        h = 1e-5 * pr["eps"][(0,) + c]
        ep, em = pr["eps"].copy(), pr["eps"].copy()
        for a in range(3):
            ep[(a,) + c] += h
            em[(a,) + c] -= h
        grads.append((objective(pr, ep) - objective(pr, em)) / (2 * h))
    path = os.path.join(HERE, "adjoint_fd_f64.npz")
    numpy.savez_compressed(path, J=J0, grad=numpy.array(grads), cells=numpy.array(pr["cells"]))
    print("wrote", path, J0, grads)


if __name__ == "__main__":
    main()
