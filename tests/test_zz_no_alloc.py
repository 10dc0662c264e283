"""fdtd_energy_noalloc (the entry point that allocates nothing: its partial
sums live in the library's static device scratch) against fdtd_energy (the
same reduction with a caller-supplied partials buffer), through the C-ABI,
on two calls in a row, the second on a smaller grid."""
import ctypes

import numpy
import pytest

pytestmark = pytest.mark.gpu


def _energy_pair(shape, seed):
    import torch
    from opencl_fdtd_amd import Simulation
    rng = numpy.random.default_rng(seed)
    eps = [rng.uniform(1, 4, shape) for _ in range(3)]
    eps[0][rng.random(shape) < 0.05] = numpy.inf
    init = {"E": [rng.standard_normal(shape) for _ in range(3)], "H": [rng.standard_normal(shape) for _ in range(3)]}
    sim = Simulation(eps, [], initial_fields=init)
    hp = sim.H.clone()
    sim.step_n(3)
    E, H = sim.E, sim.H
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    parts = torch.zeros(int(sim._lib.fdtd_reduce_partials(sim.plan)), dtype=torch.float64, device="cuda")
    ref = torch.zeros(1, dtype=torch.float64, device="cuda")
    got = torch.full((1,), numpy.nan, dtype=torch.float64, device="cuda")
    assert sim._lib.fdtd_energy(sim.plan, ptr(E), ptr(H), ptr(hp), ptr(sim.eps), ptr(parts), ptr(ref), st) == 0
    assert sim._lib.fdtd_energy_noalloc(sim.plan, ptr(E), ptr(H), ptr(hp), ptr(sim.eps), ptr(got), st) == 0
    torch.cuda.synchronize()
    return float(ref.item()), float(got.item())


def test_energy_noalloc_equals_energy_on_two_calls():
    for shape, seed in (((40, 36, 50), 1), ((9, 7, 5), 2)):   # the second call on a smaller grid
        print("case energy_noalloc", shape, flush=True)
        ref, got = _energy_pair(shape, seed)
        assert numpy.isfinite(ref) and ref != 0.0
        assert got == ref, (shape, got, ref)


def test_energy_noalloc_rejects_a_null_plan():
    from opencl_fdtd_amd import _lib
    lib = _lib.load()
    rc = lib.fdtd_energy_noalloc(None, None, None, None, None, None, None)
    assert rc != 0 and b"plan" in lib.fdtd_last_error()
