"""Benchmark of the FDTD time-stepping hot path on one MI355X.

A step is one iteration of the reference caller loop (fdtd.py:162-175):
E update (CPML psi_E), soft source, H update (CPML psi_H) with the Poynting
vector S.  The default workload is the README-scale photonic-crystal slab,
395 x 345 x 73 cells, FP32, 8-cell CPML on all faces, Poynting on.

Prints one JSON line.  value = Gcell-updates/s over the whole grid (PML
included) with all fields resident in HBM.  roofline = the dominant kernel's
algorithmic HBM bytes per launch / its average launch duration (HIP events on
the launch stream).  cpu_baseline = the oracle (C++/OpenMP restatement of the
reference's three-pass dataflow, FP32) on a bounded sample of the same grid.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def kernel_bytes(shape, pmls, kind, dtype_bytes=4, poynting=True):
    """Algorithmic bytes of one launch: E kernel = E r+w, H r, eps r (12
    words/cell); H kernel = E r, H r, H w, S w (12 words/cell with Poynting,
    9 without); plus psi r+w (4 words) per CPML slab cell and face.  The
    fused H+E pass ("HE") = E r, H r, eps r, H w, S w, E w (18 words/cell;
    15 without Poynting) plus both psi blocks (8 words per slab cell and
    face).  The two-step pass ("HE2") moves the same per launch, for two
    steps."""
    n = float(numpy.prod(shape))
    if kind in ("HE", "HE2"):
        words, psi_words = (18 if poynting else 15), 8
    else:
        words, psi_words = (12 if (kind == "E" or poynting) else 9), 4
    slab = 0.0
    for p in pmls:
        r = "xyz".find(p["axis"])
        slab += min(int(p.get("thickness", 8)), shape[r]) * n / shape[r]
    return dtype_bytes * (words * n + psi_words * slab)


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC
    summary (profiles/traffic_<workload>.md): FETCH_SIZE and WRITE_SIZE per
    dispatch, each scaled by the factor the same counter reads for a copy
    kernel of known size.  None when no summary exists."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.md" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.strip().strip("|").split("|")
        if len(parts) >= 2 and parts[0].strip() == "hbm_bytes_per_launch " + kernel:
            return float(parts[1])
    return None


def cpu_baseline(w, seconds: float):
    from oracle.oracle import OracleSimulation
    cores = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        cores = min(cores, int(env))
    orc = OracleSimulation(w.epsilon, [dict(p) for p in w.pmls], dt=w.dt, nthreads=cores)
    orc.step(None, 0)   # warm-up: page in all buffers
    steps = 0
    t0 = time.perf_counter()
    while True:
        orc.step(None, steps)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 400:
            break
    rate = w.cells * steps / el / 1e9
    return {"value": rate, "unit": "Gcell-updates/s", "cores": cores, "kind": "port",
            "sample": "%d steps of the full %s grid (FP32, three-pass reference dataflow: E, H+oS, S), "
                      "%.1f s on %s" % (steps, "x".join(map(str, w.shape)), el, platform.processor() or "host")}


def _run_fields(w, dtype, steps, backend):
    """Runs the workload for ``steps`` steps in the reference caller-loop
    order; backend 'gpu' (this library) or 'oracle' (CPU, test
    infrastructure).  Returns the final fields as float64 numpy arrays."""
    pmls = [dict(p) for p in w.pmls]
    if backend == "gpu":
        import torch
        from opencl_fdtd_amd import Simulation
        sim = Simulation(w.epsilon, pmls, dt=w.dt, float_type=dtype, bloch_boundaries=w.bloch_boundaries,
                         do_poynting=w.do_poynting)
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
                Ev = sim.E.view(3, *w.shape)
                Fv = sim.F.view(3, *w.shape)
                Ev[comp, :, :, z] += pr_t * float(a)
                Fv[comp, :, :, z] += pi_t * float(a)
            else:
                Ev = sim.E.reshape(3, *w.shape)
                Fv = sim.F.reshape(3, *w.shape)
                Ev[comp, :, :, z] += pr * a
                Fv[comp, :, :, z] += pi * a
        elif w.src_wave is not None:
            if backend == "gpu":
                sim.E[w.src_index] += float(dtype(w.src_wave[t]))
            else:
                sim.E[w.src_index] += dtype(w.src_wave[t])
        if backend == "gpu":
            sim.update_H([])
        else:
            sim.update_H()
            if sim.bloch:
                sim.update_G()
            if sim.do_poynting:
                sim.update_S()
    out = {}
    for k in ("E", "H", "S", "F", "G"):
        v = getattr(sim, k, None)
        if v is None:
            continue
        out[k] = (v.get() if backend == "gpu" else v).astype(numpy.float64)
    return out


def _rel(a, b, n):
    res = []
    for k in range(3):
        x, y = a[k * n:(k + 1) * n], b[k * n:(k + 1) * n]
        d = numpy.linalg.norm(y)
        res.append(float(numpy.linalg.norm(x - y) / d) if d > 0 else float(numpy.linalg.norm(x - y)))
    return res


def precision_report():
    """Per configuration: GPU FP32 vs GPU FP64 after the full step count
    (error of the FP32 path against an FP64 run), and GPU FP32 vs the FP32
    oracle (bit-exactness) after ``oracle_steps``."""
    from opencl_fdtd_amd import workloads
    cases = [("small", workloads.small(), 1000), ("phc_slab", workloads.phc_slab(), 300),
             ("bloch", workloads.bloch(), 4000), ("adjoint", workloads.adjoint(), 2000),
             ("learned_pad16", workloads.learned(16, steps=500), 100),
             ("learned_pad8", workloads.learned(8, steps=500), 100)]
    rows = []
    for name, w, oracle_steps in cases:
        t0 = time.perf_counter()
        f32 = _run_fields(w, numpy.float32, w.steps, "gpu")
        f64 = _run_fields(w, numpy.float64, w.steps, "gpu")
        vs64 = {k: _rel(f32[k], f64[k], w.cells) for k in f32}
        g32 = _run_fields(w, numpy.float32, oracle_steps, "gpu")
        o32 = _run_fields(w, numpy.float32, oracle_steps, "oracle")
        vso = {k: _rel(g32[k], o32[k], w.cells) for k in g32}
        exact = all(numpy.array_equal(g32[k], o32[k]) for k in g32)
        row = {"config": name, "shape": list(w.shape), "steps": w.steps, "fp32_vs_fp64_relL2": vs64,
               "oracle_steps": oracle_steps, "fp32_vs_oracle_fp32_relL2": vso, "bit_exact_vs_oracle": exact,
               "seconds": round(time.perf_counter() - t0, 1)}
        rows.append(row)
        print(json.dumps(row), flush=True)
    return rows


def learned_report(steps: int = 300, warmup: int = 20, seeds=(1000, 1001, 1002, 1003), profile_only: bool = False,
                   poynting: bool = False):
    """Interior 256x256x128: interior Gcell-updates/s of CPML-16, CPML-8, the
    learned 4-cell shell and its CPML-4 base (Poynting off by default; with
    poynting=True every run steps through the fused H+E pass), and
    reflection (dB vs a 48-cell padded run) of each on unseen scenarios."""
    import torch
    from opencl_fdtd_amd import Simulation, learned as L
    net = L.ShellNet.load(L.WEIGHTS)
    interior = (256, 256, 128)
    n_int = int(numpy.prod(interior))
    sc = L.make_scenario(seeds[0], interior=interior, steps=steps + warmup)
    rows = {}

    def timed(name, pad, faces, runner_net=None):
        shape, eps, idx, sl = L.embed(sc, pad)
        sim = Simulation(eps, faces, dt=sc.dt, do_poynting=poynting)
        wave = torch.as_tensor(sc.wave(), dtype=torch.float32, device="cuda")
        runner = L.ShellRunner(sim, runner_net) if runner_net is not None else None
        st = torch.cuda.current_stream()

        def run(n0, n):
            if runner is None:
                sim.step_n(n, idx, wave[n0:n0 + n])
            else:
                runner.step_n(n, idx, wave[n0:n0 + n])
        run(0, warmup)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        run(warmup, steps)
        b.record(st)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b)
        rows[name] = {"shape": list(shape), "cells": int(numpy.prod(shape)), "ms_per_step": ms / steps,
                      "poynting": poynting, "fused_pass": bool(sim.fused_steps),
                      "interior_gcell_s": n_int * steps / ms / 1e6,
                      "total_gcell_s": int(numpy.prod(shape)) * steps / ms / 1e6}
        print(json.dumps({name: rows[name]}), file=sys.stderr, flush=True)

    if profile_only:                           # the learned-4 loop alone (for kernel profiles)
        timed("learned4", 4, net.faces(), net)
        return rows
    timed("cpml16", 16, L.cpml_faces(16))
    timed("cpml8", 8, L.cpml_faces(8))
    timed("learned4", 4, net.faces(), net)
    timed("cpml4_base", 4, net.faces())
    refl = L.evaluate_main(net, seeds, interior=interior, steps=600, log=lambda s: print(s, file=sys.stderr))
    out = {"speed": rows, "reflection_db": refl, "correction_period": net.period,
           "learned4_vs_cpml16_interior_speedup": rows["learned4"]["interior_gcell_s"] / rows["cpml16"]["interior_gcell_s"]}
    print(json.dumps(out), flush=True)
    return out


def learned_summary(steps: int = 300, warmup: int = 20, seeds=(1000, 1001, 1002, 1003), refl_steps: int = 600):
    """The metric's second half, sized for the default run: interior
    Gcell-updates/s of the learned 4-cell shell and of CPML-16 on the same
    256 x 256 x 128 interior, with Poynting off and on, and the reflection
    (dB vs the 48-cell padded run) of the learned shell, its CPML-4 base and
    CPML-16 on unseen scenarios (seeds never used in training)."""
    import torch
    from opencl_fdtd_amd import Simulation, learned as L
    t_start = time.perf_counter()
    net = L.ShellNet.load(L.WEIGHTS)
    interior = (256, 256, 128)
    n_int = int(numpy.prod(interior))
    sc = L.make_scenario(seeds[0], interior=interior, steps=steps + warmup)
    speed = {}
    for poynting in (False, True):
        for name, pad, faces, shell in (("cpml16", 16, L.cpml_faces(16), None), ("learned4", 4, net.faces(), net)):
            shape, eps, idx, sl = L.embed(sc, pad)
            sim = Simulation(eps, faces, dt=sc.dt, do_poynting=poynting)
            wave = torch.as_tensor(sc.wave(), dtype=torch.float32, device="cuda")
            runner = L.ShellRunner(sim, shell) if shell is not None else None
            run = (lambda n0, n: runner.step_n(n, idx, wave[n0:n0 + n])) if runner else \
                  (lambda n0, n: sim.step_n(n, idx, wave[n0:n0 + n]))
            run(0, warmup)
            torch.cuda.synchronize()
            st = torch.cuda.current_stream()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(st)
            run(warmup, steps)
            b.record(st)
            torch.cuda.synchronize()
            ms = a.elapsed_time(b)
            speed["%s_poynting_%s" % (name, "on" if poynting else "off")] = {
                "grid": list(shape), "ms_per_step": ms / steps, "interior_gcell_s": n_int * steps / ms / 1e6}
            del sim, runner
    ratio = {k: speed["learned4_poynting_" + k]["interior_gcell_s"] / speed["cpml16_poynting_" + k]["interior_gcell_s"]
             for k in ("off", "on")}
    base = L.ShellNet(net.T, weights=torch.zeros_like(net.master), base_pml=net.base_pml)
    refl = []
    for sd in seeds:
        scr = L.make_scenario(sd, interior=interior, steps=refl_steps)
        wide = L.run_interior(scr, L.WIDE_PAD, L.cpml_faces(L.WIDE_PAD), sample_every=4)
        refl.append({"seed": sd, "wavelength": round(scr.wavelength, 1),
                     "learned4": L.reflection_db(L.run_learned(scr, net, sample_every=4), wide),
                     "cpml4_base": L.reflection_db(L.run_learned(scr, base, sample_every=4), wide),
                     "cpml16": L.reflection_db(L.run_interior(scr, 16, L.cpml_faces(16), sample_every=4), wide)})
        del wide
        torch.cuda.empty_cache()
    out = {"interior": list(interior), "timed_steps": steps, "speed": speed,
           "correction_period": net.period,
           "schedule": ("the correction every %d steps between two-step passes (he2_march, the first after a "
                        "correction adding it), shell_conv (composed 5x5x5 MFMA convolution) after the pass that "
                        "reaches a corrected step" % net.period) if net.period > 1 else
                       "one fused pass per step and the correction after every step",
           "learned4_vs_cpml16_interior_speedup": ratio,
           "reflection_db": refl, "reflection_steps": refl_steps,
           "worst_learned4_db": max(r["learned4"] for r in refl),
           "mean_gain_over_base_db": float(numpy.mean([r["cpml4_base"] - r["learned4"] for r in refl])),
           "seconds": round(time.perf_counter() - t_start, 1)}
    print("bench: learned %s" % json.dumps(out), file=sys.stderr, flush=True)
    return out


def adjoint_report(repeats: int = 2):
    """The adjoint configuration (128^3, CPML-8, 2000 steps, checkpoints
    every 100, flux through z = 100, a 32^3 design region; FP32 fields):
    wall time of the objective alone and of the objective with its eps
    gradient (forward with checkpoints, segment recompute, reverse steps),
    best of ``repeats`` after one warm-up run.  One JSON line."""
    import torch
    from opencl_fdtd_amd import workloads
    from opencl_fdtd_amd.adjoint import FluxAdjoint, FluxProblem, flux
    w = workloads.adjoint()
    eps = torch.as_tensor(numpy.array(w.epsilon), device="cuda")
    c = w.shape[0] // 2
    design = (slice(c - 16, c + 16), slice(c - 16, c + 16), slice(40, 72))
    p = FluxProblem(shape=w.shape, pmls=w.pmls, steps=w.steps, src_index=w.src_index, src_wave=w.src_wave,
                    axis=2, index=100, dt=w.dt, checkpoint_every=100, float_type=numpy.float32, design=design)
    flux(p, eps)
    FluxAdjoint(p, eps).gradient()
    torch.cuda.synchronize()
    t_f, t_g = [], []
    for _ in range(repeats):
        t = time.perf_counter()
        J = flux(p, eps)
        torch.cuda.synchronize()
        t_f.append(time.perf_counter() - t)
        t = time.perf_counter()
        J2, g = FluxAdjoint(p, eps).gradient()
        torch.cuda.synchronize()
        t_g.append(time.perf_counter() - t)
    out = {"workload": w.name, "shape": list(w.shape), "steps": w.steps, "checkpoint_every": 100,
           "objective_s": min(t_f), "objective_us_per_step": min(t_f) / w.steps * 1e6,
           "gradient_s": min(t_g), "gradient_us_per_step": min(t_g) / w.steps * 1e6,
           "J": J, "J_with_gradient": J2, "timing": "host wall clock, synchronised, FP32 fields"}
    print(json.dumps(out), flush=True)


def configs_report(seconds: float = 1.0):
    """Throughput of every workload (C1 small 64^3, C2 README slab, C3 Bloch
    supercell, C4 the adjoint's 128^3 forward grid, a synthetic 512^3 grid
    with CPML-8): step_n timed with HIP
    events, Gcell-updates/s and the fraction of 8 TB/s that the step's
    algorithmic bytes per cell-update represent.  One JSON line per config."""
    import torch
    from opencl_fdtd_amd import Simulation
    from opencl_fdtd_amd import workloads
    out = []
    for name in ("small", "phc_slab", "bloch", "adjoint", "large"):
        w = workloads.BUILDERS[name](steps=10)
        sim = Simulation(w.epsilon, [dict(p) for p in w.pmls], dt=w.dt, bloch_boundaries=w.bloch_boundaries,
                         do_poynting=w.do_poynting)
        st = torch.cuda.current_stream()
        sim.step_n(20)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        sim.step_n(20)
        b.record(st)
        torch.cuda.synchronize()
        n = max(20, int(seconds * 1e3 / (a.elapsed_time(b) / 20)))
        a.record(st)
        sim.step_n(n)
        b.record(st)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b)
        gcell = w.cells * n / (ms / 1e3) / 1e9
        fused = bool(sim.fused_steps)
        two = bool(sim.two_step_passes)
        if fused:
            # one he_march per step (he2_march per two steps): 18 words per
            # cell, 8 per slab cell and face, per pass; the Bloch-pair pass
            # (heb_march) moves both systems: 30 words per cell, 16 per slab
            # cell and face
            nc = float(w.cells)
            slab = sum(min(int(p.get("thickness", 8)), w.shape["xyz".find(p["axis"])]) * nc /
                       w.shape["xyz".find(p["axis"])] for p in w.pmls)
            one = 4.0 * ((30 + 16 * slab / nc) if w.bloch_boundaries else (18 + 8 * slab / nc))
            bpc = one / 2 if two else one
        else:
            bpc = one = workloads.algorithmic_bytes_per_cell(w.shape, w.pmls, poynting=w.do_poynting,
                                                             bloch=bool(w.bloch_boundaries))
        row = {"workload": w.name, "shape": list(w.shape), "steps": n, "ms_per_step": ms / n,
               "gcell_updates_s": gcell, "fused_pass": fused, "steps_per_pass": 2 if two else 1,
               "bytes_per_cell_update": bpc, "hbm_frac": gcell * bpc / HBM_PEAK_GBS,
               "hbm_frac_one_step_pass_equivalent": gcell * one / HBM_PEAK_GBS}
        print(json.dumps(row), flush=True)
        out.append(row)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="phc_slab")
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--kernel-steps", type=int, default=200, help="steps timed kernel by kernel")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="leave out the learned-shell summary")
    ap.add_argument("--precision", action="store_true",
                    help="per-config FP32 vs FP64 and FP32 vs oracle report instead of the benchmark")
    ap.add_argument("--learned", action="store_true",
                    help="learned 4-cell shell vs CPML-8/16: interior speed and reflection")
    ap.add_argument("--learned-profile", action="store_true",
                    help="with --learned: only the timed learned-4 loop (for rocprofv3)")
    ap.add_argument("--poynting", action="store_true",
                    help="with --learned: Poynting on, so every run uses the fused H+E pass")
    ap.add_argument("--all-configs", action="store_true",
                    help="throughput of every workload (one JSON line each) instead of the benchmark")
    ap.add_argument("--adjoint", action="store_true",
                    help="time the adjoint configuration's objective and gradient instead of the benchmark")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the FDTD path is measured on a single MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.precision:
        precision_report()
        return
    if args.learned:
        learned_report(profile_only=args.learned_profile, poynting=args.poynting)
        return
    if args.all_configs:
        configs_report()
        return
    if args.adjoint:
        adjoint_report()
        return

    import torch
    from opencl_fdtd_amd import Simulation
    from opencl_fdtd_amd import workloads

    total = args.warmup + args.steps
    w = workloads.BUILDERS[args.workload](steps=max(total, args.kernel_steps))
    dtype = numpy.float32 if args.dtype == "f32" else numpy.float64
    sim = Simulation(w.epsilon, [dict(p) for p in w.pmls], dt=w.dt, float_type=dtype,
                     bloch_boundaries=w.bloch_boundaries, do_poynting=w.do_poynting)
    stream = torch.cuda.current_stream()
    wave = torch.as_tensor(w.src_wave, dtype=torch.float32 if dtype == numpy.float32 else torch.float64,
                           device="cuda")
    print("bench: %s %s cells=%d steps=%d warmup=%d" % (w.name, args.dtype, w.cells, args.steps, args.warmup),
          file=sys.stderr, flush=True)

    sim.step_n(args.warmup, w.src_index, wave[:args.warmup])
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    sim.step_n(args.steps, w.src_index, wave[args.warmup:args.warmup + args.steps])
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    gcell = w.cells * args.steps / (ms / 1e3) / 1e9
    print("bench: timed %.1f ms (wall %.3f s) -> %.2f Gcell/s" % (ms, wall, gcell), file=sys.stderr, flush=True)

    # kernel-by-kernel timing (HIP events on the launch stream).  An event
    # pair around a launch also times the marker packets themselves: that
    # overhead is measured with empty pairs and subtracted.
    empty = []
    for _ in range(50):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        b.record(stream)
        empty.append((a, b))
    dur = {"E": [], "H": [], "HE": [], "HE2": []}
    evs = []
    for _ in range(args.kernel_steps):
        a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        a.record(stream)
        sim.update_E([])
        b.record(stream)
        sim.update_H([])
        sim.flush()                  # the H half-step itself, not deferred into the next update_E
        c.record(stream)
        evs.append((a, b, c))
    # the fused H+E pass step_n runs for all but the first E and last H
    # half-steps (fdtd_update_he), launched alone between event pairs
    fused = sim.fused_steps
    hevs = []
    if fused:
        amp = wave[:1]
        for _ in range(args.kernel_steps):
            a, b = (torch.cuda.Event(enable_timing=True) for _ in range(2))
            a.record(stream)
            sim.update_HE([], w.src_index, amp)
            b.record(stream)
            hevs.append((a, b))
    # the two-step pass step_n runs for most steps (fdtd_update_he2)
    two = sim.two_step_passes
    h2evs = []
    if two:
        amp2 = wave[:2]
        for _ in range(args.kernel_steps):
            a, b = (torch.cuda.Event(enable_timing=True) for _ in range(2))
            a.record(stream)
            sim.update_HE2([], w.src_index, amp2)
            b.record(stream)
            h2evs.append((a, b))
    torch.cuda.synchronize()
    marker_ms = float(numpy.median([a.elapsed_time(b) for a, b in empty]))
    for a, b, c in evs:
        dur["E"].append(a.elapsed_time(b) - marker_ms)
        dur["H"].append(b.elapsed_time(c) - marker_ms)
    for a, b in hevs:
        dur["HE"].append(a.elapsed_time(b) - marker_ms)
    for a, b in h2evs:
        dur["HE2"].append(a.elapsed_time(b) - marker_ms)
    dbytes = 4 if dtype == numpy.float32 else 8
    kern = {}
    for k in ("E", "H", "HE", "HE2"):
        if not dur[k]:
            continue
        avg_ms = float(numpy.mean(dur[k]))
        byts = kernel_bytes(w.shape, w.pmls, k, dbytes, w.do_poynting)
        kern[k] = {"avg_us": avg_ms * 1e3, "bytes": byts, "GBs": byts / (avg_ms / 1e3) / 1e9,
                   "marker_us_subtracted": marker_ms * 1e3,
                   "us_percentiles_5_50_95": [float(numpy.percentile(dur[k], q)) * 1e3 for q in (5, 50, 95)]}
    # the dominant kernel: the two-step pass when step_n uses it (one launch
    # per two steps), else the fused pass (one launch per step), else the
    # slower of the two half-steps
    dom = "HE2" if two else "HE" if fused else max(kern, key=lambda k: kern[k]["avg_us"])
    print("bench: kernels %s" % json.dumps(kern), file=sys.stderr, flush=True)

    # the drop-in path: the reference caller loop call by call (fdtd.py:162-175,
    # with its two host waits per step); update_H is deferred into the next
    # update_E, which then runs the fused H+E pass.  Untimed warm-up steps
    # first: the first source add loads torch's indexing kernels.
    loop_steps = min(args.steps, 200)
    loop_warmup = max(args.warmup, 5)

    def call_loop(t0, n):
        for t in range(t0, t0 + n):
            sim.update_E([]).wait()
            sim.E[w.src_index] += wave[t % wave.numel()]
            e = sim.update_H([])
            if sim.update_S:
                e = sim.update_S([e])
            e.wait()
        sim.flush()
        torch.cuda.synchronize()

    call_loop(0, loop_warmup)
    t1 = time.perf_counter()
    call_loop(loop_warmup, loop_steps)
    loop_s = time.perf_counter() - t1
    per_call = {"steps": loop_steps, "warmup": loop_warmup, "ms_per_step": loop_s * 1e3 / loop_steps,
                "gcell_updates_s": w.cells * loop_steps / loop_s / 1e9,
                "timing": "host wall clock around the loop, two host waits per step as in fdtd.py",
                "h_deferred_into_fused_pass": bool(sim._defer)}
    print("bench: per-call loop %s" % json.dumps(per_call), file=sys.stderr, flush=True)

    # algorithmic bytes per cell update of the pass step_n runs, and of the
    # one-step fused pass (what a pass per step would have to move)
    step_bytes = (kern["HE2"]["bytes"] / 2 if two else kern["HE"]["bytes"] if fused
                  else kern["E"]["bytes"] + kern["H"]["bytes"])
    one_pass_bytes = kern["HE"]["bytes"] if fused else kern["E"]["bytes"] + kern["H"]["bytes"]
    names = {"E": ("e_march (E)", "e_march"), "H": ("h_march (H+S)", "h_march"),
             "HE": ("he_march (H+S, then E)", "he_march"),
             "HE2": ("he2_march (two steps: H+S, E, H+S, E)", "he2_march")}
    result = {
        "metric": "Gcell-updates/s + %HBM roofline; learned vs 16-cell CPML at -40 dB reflection",
        "value": gcell,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": w.name, "shape": list(w.shape), "pml": "CPML-8 on 6 faces",
                   "poynting": w.do_poynting, "cells": w.cells,
                   "steps_per_pass": 2 if two else 1,
                   "schedule": ("step_n calls chain (fdtd_step_n_chain): the warm-up call leaves its last H "
                                "half-step pending and the timed call runs it in its first pass, leaving its own "
                                "last one pending, as update_H defers into update_E; the timed region holds the "
                                "E and H half-steps of exactly `steps` steps") if sim._defer else
                               "each step_n call runs E first and H last",
                   "s_stored": ("only for the second H half-step of each two-step pass (the first is not "
                                "computed); the S after step_n is bit-identical to the per-step loop's")
                               if two else "every H half-step",
                   "bytes_per_cell_update": step_bytes / w.cells,
                   "hbm_frac_whole_step": gcell * step_bytes / w.cells / HBM_PEAK_GBS,
                   "hbm_frac_one_step_pass_equivalent": gcell * one_pass_bytes / w.cells / HBM_PEAK_GBS},
        "per_call_loop": per_call,
        "roofline": {"bound": "hbm", "kernel": names[dom][0],
                     "limiter": ("per-plane latency: one 1024-thread workgroup per CU, two barriers and one "
                                 "memory round trip per plane" if dom == "HE2" else "HBM streaming"),
                     "achieved": kern[dom]["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kern[dom]["GBs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(w.name, names[dom][1]),
                     "traffic_source": "profiles/traffic_%s.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE, "
                                       "calibrated on a copy kernel; recorded, not measured by this run)" % w.name,
                     "avg_us": kern[dom]["avg_us"], "bytes_per_launch": kern[dom]["bytes"],
                     "steps_per_launch": 2 if dom == "HE2" else 1,
                     "other_kernel": {k: v for k, v in kern.items() if k != dom}},
    }
    if not args.no_learned:
        del sim
        torch.cuda.empty_cache()
        result["learned"] = learned_summary()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
