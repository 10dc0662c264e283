"""The hand-written adjoint and the FP64 reductions (GPU).

* transpose test of one step: <A x, w> = <x, A^T w> over the whole state
  (E, H, psi_E, psi_H) in FP64, with CPML and PEC cells;
* the FP64 adjoint gradient of the Poynting-flux objective equals the FP64
  central finite differences computed with the CPU oracle
  (tests/golden/adjoint_fd_f64.npz), and the objective equals the oracle's;
* FP32 adjoint vs FP32 and FP64 central differences on the GPU (with
  checkpointed recompute);
* plane-flux and energy reductions vs numpy in FP64.
"""
import ctypes
import os

import numpy
import pytest

from tests.golden.make_adjoint_golden import problem as oracle_problem

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _torch():
    import torch
    return torch


def test_adjoint_step_is_the_transpose():
    torch = _torch()
    from opencl_fdtd_amd import Simulation, _lib
    rng = numpy.random.default_rng(301)
    shape = (9, 8, 10)
    eps = [rng.uniform(1, 4, shape) for _ in range(3)]
    mask = rng.random(shape) < 0.05
    for e in eps:
        e[mask] = numpy.inf
    pmls = [{"axis": a, "polarity": p, "thickness": 3, "cfs_alpha": 0.03} for a in "xyz" for p in "np"]
    sim = Simulation(eps, pmls, float_type=numpy.float64)
    dev = sim.device
    n3, npsi = 3 * sim.N, sim.psi_elems
    r = lambda n: torch.as_tensor(rng.standard_normal(n), device=dev)
    x = {"E": r(n3), "H": r(n3), "psi_e": r(npsi), "psi_h": r(npsi)}
    x["E"].view(3, -1)[:, torch.as_tensor(mask.ravel(), device=dev)] = 0.0
    sim.restore(x)
    E0, H0 = sim._E.clone(), sim.H.clone()
    sim.step_n(1)
    y = sim.checkpoint()
    w = {"E": r(n3), "H": r(n3), "psi_e": r(npsi), "psi_h": r(npsi)}
    Eb, Hb = w["E"].clone(), w["H"].clone()
    pe_out, ph_out = torch.zeros(npsi, dtype=torch.float64, device=dev), torch.zeros(npsi, dtype=torch.float64, device=dev)
    grad = torch.zeros(n3, dtype=torch.float64, device=dev)
    b = _lib.AdjointBufs()
    b.E_prev, b.E_next, b.H_prev, b.H_next = E0.data_ptr(), y["E"].data_ptr(), H0.data_ptr(), y["H"].data_ptr()
    b.eps, b.Eb, b.Hb = sim.eps.data_ptr(), Eb.data_ptr(), Hb.data_ptr()
    b.psib_e, b.psib_e_out = w["psi_e"].data_ptr(), pe_out.data_ptr()
    b.psib_h, b.psib_h_out = w["psi_h"].data_ptr(), ph_out.data_ptr()
    b.coef, b.eps_grad = sim.coef.data_ptr(), grad.data_ptr()
    b.src_index, b.obj_axis = -1, -1
    _lib.check(sim._lib.fdtd_adjoint_step(sim.plan, ctypes.byref(b),
                                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "adjoint")
    torch.cuda.synchronize()
    lhs = sum(float(torch.dot(y[k], w[k])) for k in ("E", "H", "psi_e", "psi_h"))
    rhs = float(torch.dot(x["E"], Eb) + torch.dot(x["H"], Hb) + torch.dot(x["psi_e"], pe_out)
                + torch.dot(x["psi_h"], ph_out))
    print("dot product", lhs, rhs, flush=True)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def _oracle_problem_gpu(float_type):
    torch = _torch()
    from opencl_fdtd_amd.adjoint import FluxProblem
    pr = oracle_problem()
    p = FluxProblem(shape=pr["shape"], pmls=pr["pmls"], steps=pr["steps"], src_index=int(pr["src"]),
                    src_wave=pr["wave"], axis=pr["axis"], index=pr["index"], dt=pr["dt"], checkpoint_every=25,
                    float_type=float_type, base_eps=pr["eps"], design=pr["design"])
    return pr, p, torch.as_tensor(pr["eps"], device="cuda")


def test_gradient_matches_oracle_finite_differences():
    from opencl_fdtd_amd.adjoint import FluxAdjoint
    pr, p, eps = _oracle_problem_gpu(numpy.float64)
    fix = numpy.load(os.path.join(HERE, "adjoint_fd_f64.npz"))
    J, grad = FluxAdjoint(p, eps).gradient()
    g = grad.sum(dim=0).cpu().numpy()
    print("J gpu %.15g oracle %.15g" % (J, float(fix["J"])), flush=True)
    assert abs(J - float(fix["J"])) <= 1e-12 * abs(float(fix["J"]))
    for c, fd in zip(fix["cells"], fix["grad"]):
        a = g[tuple(c)]
        print("cell", tuple(c), "adjoint %.10e oracle FD %.10e" % (a, fd), flush=True)
        assert abs(a - fd) <= 1e-5 * numpy.abs(fix["grad"]).max()


def test_fp32_adjoint_vs_central_differences():
    """FP32 fields, checkpoints every 25 of 80 steps: the adjoint gradient
    against FP32 central differences (relative step 1e-3) and FP64 ones."""
    torch = _torch()
    from opencl_fdtd_amd.adjoint import FluxAdjoint, flux
    pr, p32, eps = _oracle_problem_gpu(numpy.float32)
    _, p64, _ = _oracle_problem_gpu(numpy.float64)
    J, grad = FluxAdjoint(p32, eps).gradient()
    g = grad.sum(dim=0).cpu().numpy()
    scale = None
    for c in pr["cells"]:
        fd = {}
        for name, p, rel in (("f32", p32, 1e-3), ("f64", p64, 1e-6)):
            h = rel * float(eps[(0,) + c])
            ep, em = eps.clone(), eps.clone()
            for a in range(3):
                ep[(a,) + c] += h
                em[(a,) + c] -= h
            fd[name] = (flux(p, ep) - flux(p, em)) / (2 * h)
        scale = scale or max(abs(fd["f64"]), 1e-30)
        print("cell", c, "adjoint f32 %.6e  FD f32 %.6e  FD f64 %.6e" % (g[c], fd["f32"], fd["f64"]), flush=True)
        assert abs(g[c] - fd["f64"]) <= 1e-2 * numpy.abs(numpy.load(os.path.join(HERE, "adjoint_fd_f64.npz"))["grad"]).max()


def test_gradient_box_equals_whole_grid_inside():
    """With a design region the adjoint accumulates the eps gradient only
    there: bit-identical to the whole-grid gradient inside, zero outside."""
    torch = _torch()
    import dataclasses
    from opencl_fdtd_amd.adjoint import FluxAdjoint
    pr, p, eps = _oracle_problem_gpu(numpy.float32)
    J1, boxed = FluxAdjoint(p, eps).gradient()
    J2, whole = FluxAdjoint(dataclasses.replace(p, design=None), eps).gradient()
    assert J1 == J2
    sel = (slice(None),) + tuple(pr["design"])
    assert torch.equal(boxed[sel], whole[sel])
    assert float(whole[sel].abs().max()) > 0
    inside = torch.zeros_like(boxed, dtype=torch.bool)
    inside[sel] = True
    assert float(boxed[~inside].abs().max()) == 0.0
    assert float(whole[~inside].abs().max()) > 0


def test_autograd_function():
    torch = _torch()
    from opencl_fdtd_amd.adjoint import PoyntingFluxObjective
    pr, p, eps = _oracle_problem_gpu(numpy.float64)
    d = torch.as_tensor(pr["eps"][(0,) + pr["design"]], device="cuda").clone().requires_grad_(True)
    J = PoyntingFluxObjective.apply(d, p)
    J.backward()
    fix = numpy.load(os.path.join(HERE, "adjoint_fd_f64.npz"))
    g = d.grad.cpu().numpy()
    lo = [s.start for s in pr["design"]]
    for c, fd in zip(fix["cells"], fix["grad"]):
        a = g[tuple(int(c[k]) - lo[k] for k in range(3))]
        assert abs(a - fd) <= 1e-5 * numpy.abs(fix["grad"]).max()


def test_fp64_reductions():
    torch = _torch()
    from opencl_fdtd_amd import Simulation
    rng = numpy.random.default_rng(302)
    shape = (11, 13, 9)
    eps = [rng.uniform(1, 3, shape) for _ in range(3)]
    init = {"E": [rng.standard_normal(shape) for _ in range(3)], "H": [rng.standard_normal(shape) for _ in range(3)]}
    sim = Simulation(eps, [], initial_fields=init)
    hp = sim.H.clone()
    sim.step_n(1)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    acc = torch.zeros(1, dtype=torch.float64, device="cuda")
    for axis, idx in ((0, 3), (1, 12), (2, 0)):
        acc.zero_()
        assert sim._lib.fdtd_plane_sum(sim.plan, ctypes.c_void_p(sim.S.data_ptr()), axis, axis, idx,
                                       ctypes.c_void_p(acc.data_ptr()), st) == 0
        S = sim.S.get().astype(numpy.float64).reshape(3, *shape)
        want = numpy.take(S[axis], idx, axis=axis).sum()
        assert abs(float(acc.item()) - want) <= 1e-12 * max(abs(want), 1)
    parts = torch.zeros(int(sim._lib.fdtd_reduce_partials(sim.plan)), dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    assert sim._lib.fdtd_energy(sim.plan, ctypes.c_void_p(sim._E.data_ptr()), ctypes.c_void_p(sim.H.data_ptr()),
                                ctypes.c_void_p(hp.data_ptr()), ctypes.c_void_p(sim.eps.data_ptr()),
                                ctypes.c_void_p(parts.data_ptr()), ctypes.c_void_p(out.data_ptr()), st) == 0
    E = sim.E.get().astype(numpy.float64)
    epsv = sim.eps.cpu().numpy().astype(numpy.float64)     # eps as the FP32 simulation holds it
    want = float(numpy.sum(epsv * E * E)
                 + numpy.sum(sim.H.get().astype(numpy.float64) * hp.cpu().numpy().astype(numpy.float64)))
    assert abs(float(out.item()) - want) <= 1e-12 * abs(want)


def test_adjoint_config_128():
    """128^3, CPML-8, 2000 steps, checkpoints every 100, flux through z = 100,
    eps ~ U[1,4] in a 32^3 design region: the FP32 adjoint gradient on 16
    seeded design cells against FP32 central differences (relative step
    1e-3) and FP64 central differences (relative step 1e-6)."""
    torch = _torch()
    from opencl_fdtd_amd import workloads
    from opencl_fdtd_amd.adjoint import FluxAdjoint, FluxProblem, flux
    w = workloads.adjoint()
    eps = torch.as_tensor(numpy.array(w.epsilon), device="cuda")
    c = w.shape[0] // 2
    design = (slice(c - 16, c + 16), slice(c - 16, c + 16), slice(40, 72))
    mk = lambda ft: FluxProblem(shape=w.shape, pmls=w.pmls, steps=w.steps, src_index=w.src_index,
                                src_wave=w.src_wave, axis=2, index=100, dt=w.dt, checkpoint_every=100,
                                float_type=ft, design=design)
    p32, p64 = mk(numpy.float32), mk(numpy.float64)
    print("case adjoint_128 forward+adjoint", flush=True)
    J, grad = FluxAdjoint(p32, eps).gradient()
    g = grad.sum(dim=0).cpu().numpy()
    rng = numpy.random.default_rng(4)
    cells = [tuple(int(rng.integers(s.start, s.stop)) for s in design) for _ in range(16)]
    rows = []
    for cell in cells:
        fd = {}
        for name, p, rel in (("f32", p32, 1e-3), ("f64", p64, 1e-6)):
            h = rel * float(eps[(0,) + cell])
            ep, em = eps.clone(), eps.clone()
            for a in range(3):
                ep[(a,) + cell] += h
                em[(a,) + cell] -= h
            fd[name] = (flux(p, ep) - flux(p, em)) / (2 * h)
        rows.append((cell, g[cell], fd["f32"], fd["f64"]))
        print("cell", cell, "adjoint f32 %.6e  FD f32 %.6e  FD f64 %.6e" % rows[-1][1:], flush=True)
    gmax = max(abs(r[3]) for r in rows)
    err_adj = max(abs(r[1] - r[3]) for r in rows) / gmax
    err_fd32 = max(abs(r[2] - r[3]) for r in rows) / gmax
    print("J %.9e  max|adjoint - FD64|/max|g| %.3e  max|FD32 - FD64|/max|g| %.3e" % (J, err_adj, err_fd32),
          flush=True)
    assert err_adj < 1e-3
