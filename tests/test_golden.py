"""Golden fixtures (tests/golden/*.npz, written by make_golden.py from the
oracle): the oracle reproduces them (CPU), and the HIP path reproduces them
bit for bit (GPU)."""
import os

import numpy
import pytest

from tests.golden.make_golden import cases, run_case

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(cases().keys())


def _fixture(name):
    with numpy.load(os.path.join(HERE, name + ".npz")) as z:
        return {k[4:]: z[k] for k in z.files if k.startswith("out_")}, int(z["steps"])


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    kw, steps, src = cases()[name]
    want, n = _fixture(name)
    assert n == steps
    got = run_case(kw, steps, src)
    for k, v in want.items():
        assert numpy.array_equal(got[k], v), k


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_reproduces_fixture(name):
    from opencl_fdtd_amd import Simulation
    kw, steps, src = cases()[name]
    want, _ = _fixture(name)
    print("golden case", name, flush=True)
    sim = Simulation(**kw)
    for t in range(steps):
        sim.update_E([])
        if sim.update_F:
            sim.update_F([])
        if src is not None:
            sim.E[src[0]] += float(src[1][t])
        e = sim.update_H([])
        if sim.update_S:
            e = sim.update_S([e])
    e.wait()
    for k, v in want.items():
        got = getattr(sim, k).get()
        bad = numpy.flatnonzero(got != v)
        assert bad.size == 0, (k, bad.size, bad[:5])
