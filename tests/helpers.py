"""Shared helpers for the parity tests: run the HIP Simulation and the CPU
oracle side by side on the same inputs."""
from __future__ import annotations

import numpy


def rel_l2(a, b) -> float:
    a = numpy.asarray(a, dtype=numpy.float64)
    b = numpy.asarray(b, dtype=numpy.float64)
    den = numpy.linalg.norm(b)
    num = numpy.linalg.norm(a - b)
    return float(num / den) if den > 0 else float(num)


def comp_rel_l2(a, b, n) -> list:
    """Relative L2 per component (x, y, z) of two flat 3N arrays."""
    return [rel_l2(a[k * n:(k + 1) * n], b[k * n:(k + 1) * n]) for k in range(3)]


def random_fields(rng, shape, scale=1.0):
    return [scale * rng.standard_normal(shape) for _ in range(3)]


def build_pair(epsilon, pmls, dtype=numpy.float32, bloch=(), init=None, dxes=None, dt=None,
               poynting_mode="fused", bloch_mode="phase", do_poynting=True, do_fieldsrc=False):
    from opencl_fdtd_amd import Simulation
    from oracle.oracle import OracleSimulation
    sim = Simulation([numpy.asarray(e) for e in epsilon], [dict(p) for p in pmls],
                     bloch_boundaries=[dict(b) for b in bloch], dxes=dxes, dt=dt,
                     initial_fields=init, float_type=dtype, do_poynting=do_poynting,
                     do_fieldsrc=do_fieldsrc, poynting_mode=poynting_mode, bloch_mode=bloch_mode)
    orc = OracleSimulation([numpy.asarray(e) for e in epsilon], [dict(p) for p in pmls],
                           bloch_boundaries=[dict(b) for b in bloch], dxes=dxes, dt=dt,
                           initial_fields=init, float_type=dtype, do_poynting=do_poynting,
                           do_fieldsrc=do_fieldsrc, bloch_mode=bloch_mode)
    return sim, orc


def step_gpu(sim, steps, src_index=None, wave=None):
    for t in range(steps):
        sim.update_E([])
        if sim.update_F:
            sim.update_F([])
        if wave is not None:
            sim.E[src_index] += float(wave[t])
        e = sim.update_H([])
        if sim.update_G:
            e = sim.update_G([e])
        if sim.update_S:
            e = sim.update_S([e])
    e.wait()


def step_oracle(orc, steps, src_index=None, wave=None):
    def src(o, t):
        # the source value is converted to the field type, then added
        o.E[src_index] += o.arg_type(wave[t])
    for t in range(steps):
        orc.step(src if wave is not None else None, t)


def fetch(sim, name):
    v = getattr(sim, name)
    return None if v is None else v.get()


def run_workload(w, steps, backend, dtype=numpy.float32, **sim_kwargs):
    """Runs a ``workloads.Workload`` for ``steps`` steps in the reference
    caller-loop order (E, F, soft source, H, G, S; fdtd.py:162-175) on the HIP
    library ('gpu') or the CPU oracle ('oracle').  A plane source adds
    Re(phase) * f(t) to E and Im(phase) * f(t) to F on its z plane, each
    product rounded to the field type.  Returns the simulation object."""
    import os
    pmls = [dict(p) for p in w.pmls]
    if backend == "gpu":
        import torch
        from opencl_fdtd_amd import Simulation
        sim = Simulation(w.epsilon, pmls, dt=w.dt, float_type=dtype, bloch_boundaries=w.bloch_boundaries,
                         do_poynting=w.do_poynting, **sim_kwargs)
    else:
        from oracle.oracle import OracleSimulation
        sim = OracleSimulation(w.epsilon, pmls, dt=w.dt, float_type=dtype, bloch_boundaries=w.bloch_boundaries,
                               do_poynting=w.do_poynting, nthreads=min(len(os.sched_getaffinity(0)), 16))
    ps = w.plane_source
    if ps is not None:
        z, comp = ps["z"], ps["comp"]
        pr = numpy.real(ps["phase"]).astype(dtype)
        pi = numpy.imag(ps["phase"]).astype(dtype)
        if backend == "gpu":
            pr_t, pi_t = torch.as_tensor(pr, device="cuda"), torch.as_tensor(pi, device="cuda")
    for t in range(steps):
        if backend == "gpu":
            sim.update_E([])
        else:
            sim.update_E()
            if sim.bloch:
                sim.update_F()
        if ps is not None:
            a = dtype(ps["wave"][t])
            if backend == "gpu":
                sim.E.view(3, *w.shape)[comp, :, :, z] += pr_t * float(a)
                sim.F.view(3, *w.shape)[comp, :, :, z] += pi_t * float(a)
            else:
                sim.E.reshape(3, *w.shape)[comp, :, :, z] += pr * a
                sim.F.reshape(3, *w.shape)[comp, :, :, z] += pi * a
        elif w.src_wave is not None:
            sim.E[w.src_index] += float(dtype(w.src_wave[t])) if backend == "gpu" else dtype(w.src_wave[t])
        if backend == "gpu":
            e = sim.update_H([])
            if sim.update_G:
                e = sim.update_G([e])
            if sim.update_S:
                e = sim.update_S([e])
        else:
            sim.update_H()
            if sim.bloch:
                sim.update_G()
            if sim.do_poynting:
                sim.update_S()
    if backend == "gpu":
        e.wait()
    return sim
