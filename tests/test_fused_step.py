"""The fused H+E pass (fdtd_update_he) and the two-step pass
(fdtd_update_he2), both used by step_n, against the unfused kernels and the
CPU oracle.

The fused pass does, per cell, the same arithmetic in the same order as the
H kernel followed by the E kernel, so every comparison here is bit-exact:
fields, S and both CPML psi blocks.
"""
import math

import numpy
import pytest
import torch

from tests.helpers import build_pair, fetch, random_fields, step_oracle

pytestmark = pytest.mark.gpu


def _case(shape, T, seed, nonuniform=False, pec=False):
    rng = numpy.random.default_rng(seed)
    eps = [rng.uniform(1, 6, shape) for _ in range(3)]
    if pec:
        for e in eps:
            e[shape[0] // 2, :, :shape[2] // 3] = numpy.inf
    pmls = [{"axis": a, "polarity": p, "thickness": T} for a in "xyz" for p in "np"] if T else []
    init = {"E": random_fields(rng, shape), "H": random_fields(rng, shape)}
    dxes = None
    if nonuniform:
        dxes = [[rng.uniform(0.8, 1.25, s) for s in shape] for _ in range(2)]
    return eps, pmls, init, dxes


def _sim(eps, pmls, init, dxes, dtype, fuse):
    from opencl_fdtd_amd import Simulation
    return Simulation([numpy.asarray(e) for e in eps], [dict(p) for p in pmls], dxes=dxes,
                      initial_fields=init, float_type=dtype, fuse_steps=fuse)


def _state(sim):
    return {"E": fetch(sim, "E"), "H": fetch(sim, "H"), "S": fetch(sim, "S"),
            "psi_e": sim._psi_e.cpu().numpy(),
            "psi_h": sim._psi_h[sim._h].cpu().numpy()}


def _assert_same(a, b, what):
    for k in a:
        bad = numpy.flatnonzero(a[k] != b[k])
        assert bad.size == 0, "%s: %s differs at %d of %d values, first %s: %r vs %r" % (
            what, k, bad.size, a[k].size, bad[:4], a[k][bad[:4]], b[k][bad[:4]])


@pytest.mark.parametrize("shape, T, n, dtype, nonuniform", [
    ((11, 13, 9), 3, 2, numpy.float32, False),
    ((11, 13, 9), 3, 3, numpy.float32, False),
    ((11, 13, 9), 3, 8, numpy.float32, True),
    ((9, 40, 73), 4, 7, numpy.float32, False),       # rows of 73: 6 rows + halo per 512-thread tile
    ((6, 9, 200), 2, 5, numpy.float32, False),       # 1024-thread tiles (rows of 129..512 cells)
    ((5, 37, 3), 2, 6, numpy.float32, False),        # short rows, many per tile
    ((33, 8, 17), 3, 9, numpy.float64, True),
    ((5, 6, 160), 2, 4, numpy.float64, False),       # FP64, 1024-thread tiles
    ((6, 7, 5), 4, 5, numpy.float32, False),         # slabs of both faces overlap on every axis
    ((2, 9, 10), 1, 4, numpy.float32, False),        # two planes: every chunk is one plane
    ((3, 4, 341), 2, 3, numpy.float32, False),       # two owned rows per tile
    ((3, 4, 512), 2, 4, numpy.float32, False),       # whole rows of 512: one owned row per tile
    ((4, 3, 400), 3, 5, numpy.float64, True),        # one owned row per tile, FP64, nonuniform
    ((3, 4, 700), 2, 4, numpy.float32, False),       # rows too long for whole-row tiles: 64-cell z segments
    ((3, 2, 1500), 2, 3, numpy.float32, False),      # two rows: the halo rows repeat owned rows
    ((4, 3, 1100), 3, 3, numpy.float64, True),       # FP64, nonuniform, a short last segment
    ((5, 6, 1), 1, 6, numpy.float32, True),          # one-cell rows
])
def test_step_n_fused_equals_unfused(shape, T, n, dtype, nonuniform):
    print("case fused step_n", shape, T, n, dtype.__name__, nonuniform, flush=True)
    eps, pmls, init, dxes = _case(shape, T, seed=sum(shape) + n, nonuniform=nonuniform, pec=True)
    a = _sim(eps, pmls, init, dxes, dtype, True)
    b = _sim(eps, pmls, init, dxes, dtype, False)
    assert a._E_out is not None and b._E_out is None
    N = a.N
    idx = 2 * N + (shape[0] // 2) * shape[1] * shape[2] + (shape[1] // 3) * shape[2] + shape[2] // 2
    wave = numpy.sin(0.3 * numpy.arange(3 * n)).astype(dtype)
    for k in range(3):                           # three calls: E may change buffers between them
        a.step_n(n, idx, wave[k * n:(k + 1) * n])
        b.step_n(n, idx, wave[k * n:(k + 1) * n])
    torch.cuda.synchronize()
    _assert_same(_state(a), _state(b), "step_n fused vs unfused")


@pytest.mark.parametrize("shape, T, n", [
    ((160, 300, 250), 8, 6),     # 12 M cells, 1024-thread tiles of 3 rows, several rounds of workgroups
    ((40, 700, 100), 6, 5),      # 78 tile columns of 9 rows, short x
    ((24, 40, 512), 8, 5),       # rows of 512 in 64-cell z segments, two steps per pass
    ((40, 30, 600), 6, 6),       # the two-step pass on 64-cell z segments
])
def test_fused_large_grids_equal_unfused(shape, T, n):
    """Launch shapes the cost model gives larger grids (multiple rounds of
    workgroups), fused against unfused, bit for bit."""
    print("case fused large", shape, T, n, flush=True)
    rng = numpy.random.default_rng(sum(shape))
    eps = [rng.uniform(1, 4, shape).astype(numpy.float32) for _ in range(3)]
    pmls = [{"axis": a, "polarity": p, "thickness": T} for a in "xyz" for p in "np"]
    a = _sim(eps, pmls, None, None, numpy.float32, True)
    b = _sim(eps, pmls, None, None, numpy.float32, False)
    idx = 1 * a.N + (shape[0] // 2) * shape[1] * shape[2] + (shape[1] // 2) * shape[2] + shape[2] // 2
    wave = numpy.sin(0.25 * numpy.arange(n)).astype(numpy.float32)
    a.step_n(n, idx, wave)
    b.step_n(n, idx, wave)
    torch.cuda.synchronize()
    assert float(torch.abs(a.E).max()) > 0
    for name in ("E", "H", "S"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name
    del a, b
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", [numpy.float32, numpy.float64])
def test_update_he_equals_h_then_e(dtype):
    """One fdtd_update_he = update_H (+S), update_E, E[idx] += amp."""
    shape = (10, 12, 14)
    eps, pmls, init, _ = _case(shape, 3, seed=5)
    a = _sim(eps, pmls, init, None, dtype, True)
    b = _sim(eps, pmls, init, None, dtype, False)
    for s in (a, b):                             # non-zero psi first
        for _ in range(3):
            s.update_E([])
            s.update_H([])
    idx = 1 * a.N + 5 * shape[1] * shape[2] + 6 * shape[2] + 7
    tdt = torch.float32 if dtype == numpy.float32 else torch.float64
    for amp in (0.75, -1.5):
        t_amp = torch.tensor([amp], dtype=tdt, device="cuda")
        a.update_HE([], idx, t_amp)
        b.update_H([])
        b.update_E([])
        b.E[idx] += t_amp[0]
        torch.cuda.synchronize()
        _assert_same(_state(a), _state(b), "update_HE amp %g" % amp)


@pytest.mark.parametrize("shape, T, nonuniform", [
    ((10, 12, 14), 3, False),
    ((10, 12, 14), 3, True),
    ((7, 3, 256), 1, False),     # whole rows of 256: one owned row per two-step tile
    ((3, 2, 270), 1, False),     # two-step pass on 64-cell z segments, a short last one, two rows
    ((4, 5, 700), 2, True),      # nonuniform, z segments for both passes
    ((4, 2, 9), 1, False),       # two rows: the halo rows are owned rows of the same tile
    ((4, 25, 40), 2, False),     # four planes, several tiles per column
    ((2, 5, 9), 1, False),       # two planes: the chunk's extra planes wrap onto its own
    ((6, 20, 30), 0, False),     # no CPML: no psi blocks at all
])
def test_update_he2_equals_he_twice(shape, T, nonuniform):
    """One fdtd_update_he2 = fdtd_update_he twice (sources of both steps),
    bit for bit: fields, S and both psi blocks."""
    print("case update_HE2", shape, T, nonuniform, flush=True)
    eps, pmls, init, dxes = _case(shape, T, seed=11, nonuniform=nonuniform, pec=True)
    a = _sim(eps, pmls, init, dxes, numpy.float32, True)
    b = _sim(eps, pmls, init, dxes, numpy.float32, True)
    assert a._psi_e_out is not None
    for s in (a, b):                             # non-zero psi first
        for _ in range(3):
            s.update_E([])
            s.update_H([])
        s.update_E([])
    idx = 1 * a.N + (shape[0] // 2) * shape[1] * shape[2] + (shape[1] // 2) * shape[2] + shape[2] // 3
    for amps in ((0.75, -1.5), (2.0, 0.25)):
        t_amp = torch.tensor(amps, dtype=torch.float32, device="cuda")
        a.update_HE2([], idx, t_amp)
        b.update_HE([], idx, t_amp[0:1])
        b.update_HE([], idx, t_amp[1:2])
        torch.cuda.synchronize()
        _assert_same(_state(a), _state(b), "update_HE2 amps %s" % (amps,))
    a.update_HE2([])
    b.update_HE([])
    b.update_HE([])
    torch.cuda.synchronize()
    _assert_same(_state(a), _state(b), "update_HE2 without a source")


def test_update_he2_sources_on_tile_and_chunk_edges():
    """The source added inside the two-step pass on cells that other
    workgroups also evaluate: rows y0-1 and y1 of a tile (its halo rows),
    the periodic wraps, and, on this grid, x chunks of one plane, so every
    source plane is also an extra plane of the neighbouring chunks.  40 x 50
    x 30 cells: tiles of 31 rows (rows 0-30 and 31-49), 40 chunks."""
    shape = (40, 50, 30)
    eps, pmls, init, _ = _case(shape, 3, seed=21, pec=True)
    a = _sim(eps, pmls, init, None, numpy.float32, True)
    b = _sim(eps, pmls, init, None, numpy.float32, True)
    for s in (a, b):
        for _ in range(2):
            s.update_E([])
            s.update_H([])
        s.update_E([])
    N = a.N
    cells = [(0, 5, 31, 0), (1, 0, 30, 29), (2, 39, 0, 15), (0, 20, 49, 7), (2, 17, 32, 29), (1, 38, 1, 1)]
    for k, (comp, x, y, z) in enumerate(cells):
        idx = comp * N + (x * shape[1] + y) * shape[2] + z
        t_amp = torch.tensor([0.5 + k, -1.25 * (k + 1)], dtype=torch.float32, device="cuda")
        a.update_HE2([], idx, t_amp)
        b.update_HE([], idx, t_amp[0:1])
        b.update_HE([], idx, t_amp[1:2])
        torch.cuda.synchronize()
        _assert_same(_state(a), _state(b), "update_HE2 source at %s" % ((comp, x, y, z),))


def test_step_n_two_step_schedule_equals_unfused():
    """step_n on a grid where the two-step pass is chosen, for n = 1 .. 10:
    (n - 1) / 2 two-step passes and at most one one-step pass; E, H, S and
    both psi blocks end as the unfused loop leaves them, call after call,
    wherever the schedule leaves E and psi_E."""
    shape = (120, 120, 73)
    eps, pmls, init, _ = _case(shape, 8, seed=33, pec=True)
    a = _sim(eps, pmls, init, None, numpy.float32, True)
    b = _sim(eps, pmls, init, None, numpy.float32, False)
    assert a.two_step_passes
    idx = 1 * a.N + 60 * shape[1] * shape[2] + 61 * shape[2] + 36
    k0 = 0
    homes = set()
    for n in (1, 2, 3, 4, 5, 6, 7, 8, 9, 10):
        wave = numpy.sin(0.21 * numpy.arange(k0, k0 + n)).astype(numpy.float32)
        a.step_n(n, idx, wave)
        b.step_n(n, idx, wave)
        k0 += n
        torch.cuda.synchronize()
        homes.add((a._E.data_ptr(), a._psi_e.data_ptr()))
        _assert_same(_state(a), _state(b), "step_n(%d) two-step schedule vs unfused" % n)
    assert len(homes) > 1     # the schedule did move E / psi_E


@pytest.mark.parametrize("shape, T, dtype, poynting", [
    ((120, 120, 73), 8, numpy.float32, True),
    ((23, 30, 73), 8, numpy.float32, False),
    ((11, 13, 9), 3, numpy.float64, True),
])
def test_step_n_calls_chain(shape, T, dtype, poynting):
    """Consecutive step_n calls with nothing read between them chain
    (fdtd_step_n_chain): each call leaves its last H half-step pending and
    the next call (or update_E) runs it in its first pass.  Mixed with
    zero-step calls and the per-call loop, E, H, S and both psi blocks equal
    those of the same calls without deferral, bit for bit."""
    from opencl_fdtd_amd import Simulation
    print("case step_n chain", shape, T, dtype.__name__, poynting, flush=True)
    eps, pmls, init, _ = _case(shape, T, seed=43, pec=True)
    mk = lambda d: Simulation([numpy.asarray(e) for e in eps], [dict(p) for p in pmls], initial_fields=init,
                              float_type=dtype, do_poynting=poynting, defer_h=d)
    a, b = mk(True), mk(False)
    idx = 1 * a.N + (shape[0] // 2) * shape[1] * shape[2] + (shape[1] // 3) * shape[2] + shape[2] // 2
    wave = numpy.sin(0.17 * numpy.arange(64)).astype(dtype)
    for s in (a, b):
        k = 0
        for op, n in (("n", 5), ("n", 6), ("n", 1), ("n", 0), ("loop", 2), ("n", 4), ("n", 3), ("n", 8)):
            if op == "n":
                s.step_n(n, idx, wave[k:k + n])
                k += n
                if s is a:
                    assert s.pending_h                # left for the next call
                continue
            for _ in range(n):
                s.update_E([])
                s.E[idx] += float(wave[k])
                s.update_H([])
                if s.update_S:                        # None without Poynting
                    s.update_S([])
                k += 1
    torch.cuda.synchronize()
    assert a.pending_h and not b.pending_h
    _assert_same(_state(a), _state(b), "chained step_n vs unchained")
    assert not a.pending_h                            # the reads ran it


def test_step_n_c_abi_keeps_e_home():
    """fdtd_step_n (the C-ABI entry point without the move report) keeps the
    pass counts even: E and psi_E end in the caller's buffers, with the same
    results as the unfused loop."""
    import ctypes
    shape = (120, 120, 73)
    eps, pmls, init, _ = _case(shape, 8, seed=34, pec=True)
    a = _sim(eps, pmls, init, None, numpy.float32, True)
    b = _sim(eps, pmls, init, None, numpy.float32, False)
    assert a.two_step_passes
    idx = 1 * a.N + 60 * shape[1] * shape[2] + 61 * shape[2] + 36
    k0 = 0
    for n in (5, 6, 7, 8, 9, 10):
        wave = torch.as_tensor(numpy.sin(0.21 * numpy.arange(k0, k0 + n)).astype(numpy.float32), device="cuda")
        E0, psi0 = a._E.data_ptr(), a._psi_e.data_ptr()
        f = a._fields()
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        assert a._lib.fdtd_step_n(a.plan, ctypes.byref(f), n, idx, ctypes.c_void_p(wave.data_ptr()), st) == 0
        if n % 2 == 1:
            a._h = 1 - a._h
        b.step_n(n, idx, wave)
        k0 += n
        torch.cuda.synchronize()
        assert a._E.data_ptr() == E0 and a._psi_e.data_ptr() == psi0
        _assert_same(_state(a), _state(b), "fdtd_step_n(%d) vs unfused" % n)


@pytest.mark.parametrize("shape, T, dtype", [
    ((23, 30, 73), 8, numpy.float32),
    ((120, 120, 73), 8, numpy.float32),
    ((11, 13, 9), 3, numpy.float64),
])
def test_per_call_loop_runs_the_fused_pass(shape, T, dtype):
    """The reference caller loop (update_E, E[idx] += f(t), update_H,
    update_S, wait) with the H half-step deferred into the next update_E:
    every step after the first runs as one fused pass, and E, H, S match the
    oracle bit for bit after each of several reads between the steps."""
    print("case per-call loop", shape, T, dtype.__name__, flush=True)
    eps, pmls, init, _ = _case(shape, T, seed=41, pec=True)
    sim, orc = build_pair(eps, pmls, init=init, dtype=dtype)
    idx = 2 * sim.N + (shape[0] // 2) * shape[1] * shape[2] + (shape[1] // 2) * shape[2] + shape[2] // 2
    wave = numpy.cos(0.2 * numpy.arange(40))
    t = 0
    for block in (7, 1, 12):
        for _ in range(block):
            sim.update_E([]).wait()
            assert not sim.pending_h
            sim.E[idx] += float(dtype(wave[t]))
            e = sim.update_H([])
            assert sim.pending_h          # the H half-step waits for the next update_E
            e = sim.update_S([e])
            e.wait()
            t += 1
    step_oracle(orc, t, idx, wave)
    for name in ("E", "H", "S"):
        g, o = fetch(sim, name), getattr(orc, name)
        bad = numpy.flatnonzero(g != o)
        assert bad.size == 0, (name, bad.size, bad[:4])
    assert not sim.pending_h


def test_per_call_loop_mixed_with_other_calls():
    """Deferred H half-steps next to reads, step_n, update_J, update_HE and
    checkpoints: each of them runs the recorded half-step first, so the
    fields equal those of the loop without deferral."""
    from opencl_fdtd_amd import Simulation
    shape = (12, 14, 73)
    eps, pmls, init, _ = _case(shape, 3, seed=42, pec=True)
    mk = lambda d: Simulation([numpy.asarray(e) for e in eps], [dict(p) for p in pmls], initial_fields=init,
                              do_fieldsrc=True, defer_h=d)
    a, b = mk(True), mk(False)
    rng = numpy.random.default_rng(3)
    for s in (a, b):
        s.Jr.copy_(torch.as_tensor(rng.standard_normal(3 * s.N).astype(numpy.float32)))
        s.Ji.copy_(torch.as_tensor(rng.standard_normal(3 * s.N).astype(numpy.float32)))
        rng = numpy.random.default_rng(3)
    idx = 1 * a.N + 6 * shape[1] * shape[2] + 7 * shape[2] + 30
    for s in (a, b):
        s.update_E([])
        s.update_H([])
        h_mid = s.H.clone()                       # read: runs the recorded H
        s.update_E([])
        s.update_H([])
        s.update_J([], 0.5, -0.25, 2, 9, 3, 11, 0, 73)   # J after a recorded H
        s.update_E([])
        s.update_H([])
        s.step_n(5, idx, numpy.linspace(-1, 1, 5).astype(numpy.float32))
        s.update_E([])
        s.update_H([])
        cp = s.checkpoint()
        s.update_E([])
        s.update_H([])
        s.update_HE([])
        s.update_H([])
        s.restore(cp)
        s.update_E([])
        s.update_H([])
        s._h_mid = h_mid
    torch.cuda.synchronize()
    assert torch.equal(a._h_mid, b._h_mid)
    _assert_same(_state(a), _state(b), "deferred vs immediate H")



@pytest.mark.parametrize("shape, T, two", [
    ((14, 11, 150), 4, False),       # rows of 150 cells (1024-thread tiles)
    ((6, 7, 5), 4, False),           # overlapping slabs on every axis
    ((23, 30, 73), 8, False),        # README row length, CPML-8
    ((120, 120, 73), 8, True),       # the same rows, a grid the two-step pass is chosen for
    ((40, 30, 600), 8, True),        # the two-step pass on 64-cell z segments
    ((6, 7, 700), 3, False),         # the one-step pass on 64-cell z segments
])
def test_fused_step_n_matches_oracle(shape, T, two):
    """step_n through the fused pass (and the two-step pass) against the
    oracle, with a source."""
    print("case fused vs oracle", shape, T, two, flush=True)
    eps, pmls, init, _ = _case(shape, T, seed=9, pec=True)
    sim, orc = build_pair(eps, pmls, init=init)
    assert sim._E_out is not None
    assert sim.two_step_passes == two
    idx = 2 * sim.N + (shape[0] // 2) * shape[1] * shape[2] + (shape[1] // 2) * shape[2] + shape[2] // 2
    wave = numpy.cos(0.2 * numpy.arange(21))
    sim.step_n(21, idx, wave.astype(numpy.float32)).wait()
    step_oracle(orc, 21, idx, wave)
    for name in ("E", "H", "S"):
        g, o = fetch(sim, name), getattr(orc, name)
        bad = numpy.flatnonzero(g != o)
        assert bad.size == 0, (name, bad.size, bad[:4])


def test_fusion_is_off_where_unsupported():
    from opencl_fdtd_amd import Simulation
    shape = (6, 5, 7)
    eps = [numpy.ones(shape)] * 3
    assert Simulation(eps, [])._E_out is not None
    assert Simulation(eps, [], poynting_mode="two_pass")._E_out is None
    assert Simulation(eps, [], do_poynting=False)._E_out is not None            # Poynting off: fused too
    assert Simulation(eps, [], do_poynting=False, fuse_steps=False)._E_out is None
    assert len(Simulation(eps, [], do_poynting=False, fuse_steps=False)._H) == 1   # no second H then
    bloch = [{"axis": "x", "real": 1.0, "imag": 0.0}]
    assert Simulation(eps, [], do_poynting=False, bloch_boundaries=bloch)._E_out is None
    sb = Simulation(eps, [], bloch_boundaries=bloch)
    assert sb._E_out is not None and sb._F_out is not None and not sb.two_step_passes
    assert Simulation(eps, [], bloch_boundaries=bloch, float_type=numpy.float64)._E_out is None
    nu = [[numpy.ones(n) for n in shape], [numpy.ones(n) for n in shape]]
    assert Simulation(eps, [], bloch_boundaries=bloch, dxes=nu)._E_out is None
    s = Simulation(eps, [], poynting_mode="two_pass")
    with pytest.raises(Exception):
        s.update_HE([])
    assert Simulation(eps, [])._psi_e_out is not None
    assert Simulation(eps, [], float_type=numpy.float64)._psi_e_out is None
    thick = [{"axis": "z", "polarity": p, "thickness": 3} for p in "np"]   # 3 + 3 of 5 z cells overlap
    assert Simulation([numpy.ones((6, 7, 5))] * 3, thick)._psi_e_out is None
    assert not Simulation(eps, [], float_type=numpy.float64).two_step_passes
    with pytest.raises(Exception):
        Simulation(eps, [], float_type=numpy.float64).update_HE2([])


def _bloch_case(shape, T, seed, axes="xyz", face_axes="z"):
    rng = numpy.random.default_rng(seed)
    eps = [rng.uniform(1, 6, shape) for _ in range(3)]
    for e in eps:
        e[shape[0] // 2, :, :shape[2] // 3] = numpy.inf
    bl = [{"axis": a, "real": math.cos(ph), "imag": math.sin(ph)} for a, ph in zip(axes, (0.7, -1.9, 2.4))]
    pmls = [{"axis": fa, "polarity": p, "thickness": T} for fa in face_axes for p in "np"] if T else []
    init = {k: random_fields(rng, shape) for k in "EHFG"}
    return eps, pmls, bl, init


@pytest.mark.parametrize("mode", ["phase", "decoupled"])
@pytest.mark.parametrize("shape, T, axes, face_axes", [
    ((8, 10, 12), 3, "xy", "z"),      # one tile per plane
    ((6, 24, 40), 4, "xyz", "z"),     # seams on every axis, several tiles per plane
    ((5, 9, 256), 16, "xy", "z"),     # rows of 256 (the Bloch supercell's): three owned rows per tile
    ((4, 3, 512), 2, "xz", "z"),      # one owned row per tile, z seam inside the row
    ((9, 12, 14), 2, "xy", "xyz"),    # CPML on every axis (the general variant of the pass)
    ((7, 16, 20), 3, "z", "x"),       # CPML on x only, Bloch on z
    ((6, 8, 10), 0, "xyz", ""),       # no CPML at all
])
def test_bloch_pair_pass_equals_paired_kernels(mode, shape, T, axes, face_axes):
    """The fused Bloch-pair pass (update_HE on a Bloch plan) against the
    paired h_march + e_march, bit for bit: E, H, F, G, S and the psi blocks,
    with a source; with CPML on z only, on other axes, and none (the pass is
    specialised for the axes that carry faces)."""
    from opencl_fdtd_amd import Simulation
    print("case bloch pair pass", mode, shape, T, axes, face_axes, flush=True)
    eps, pmls, bl, init = _bloch_case(shape, T, seed=sum(shape), axes=axes, face_axes=face_axes)
    mk = lambda fuse: Simulation([numpy.asarray(e) for e in eps], [dict(p) for p in pmls], bloch_boundaries=bl,
                                 initial_fields=init, bloch_mode=mode, fuse_steps=fuse, defer_h=False)
    a, b = mk(True), mk(False)
    assert a._E_out is not None and b._E_out is None
    for s in (a, b):                  # non-zero psi first
        for _ in range(2):
            s.update_E([])
            s.update_H([])
    idx = 1 * a.N + (shape[0] // 2) * shape[1] * shape[2] + (shape[1] // 2) * shape[2] + 1
    for amp in (0.75, -1.5, 2.0):
        t_amp = torch.tensor([amp], dtype=torch.float32, device="cuda")
        a.update_HE([], idx, t_amp)
        b.update_H([])
        b.update_E([])
        b.E[idx] += t_amp[0]
        torch.cuda.synchronize()
        st = lambda s: dict(_state(s), F=fetch(s, "F"), G=fetch(s, "G"), psi_f=s._psi_f.cpu().numpy(),
                            psi_g=s._psi_g[s._h].cpu().numpy())
        _assert_same(st(a), st(b), "Bloch pair pass amp %g" % amp)


def test_bloch_step_n_and_loop_match_oracle():
    """step_n and the per-call loop (update_H deferred into the fused Bloch
    pair pass) against the oracle with Bloch phases on x and y, CPML on z
    and eps = inf cells, on rows of 256 cells."""
    print("case bloch step_n vs oracle", flush=True)
    eps, pmls, bl, init = _bloch_case((6, 12, 256), 8, seed=5, axes="xy")
    sim, orc = build_pair(eps, pmls, bloch=bl, init=init)
    assert sim._F_out is not None
    sim.step_n(7)
    for _ in range(5):
        sim.update_E([])
        sim.update_F([])
        sim.update_H([])
        sim.update_G([])
        sim.update_S([]).wait()
    sim.step_n(4)
    for _ in range(16):
        orc.step()
    for name in ("E", "H", "F", "G", "S"):
        g, o = fetch(sim, name), getattr(orc, name)
        bad = numpy.flatnonzero(g != o)
        assert bad.size == 0, (name, bad.size, bad[:4])


@pytest.mark.parametrize("shape, T, dtype, mode", [
    ((10, 12, 14), 3, numpy.float32, "fused"),
    ((30, 40, 73), 4, numpy.float32, "fused"),       # step_n would take the two-step pass here
    ((6, 7, 8), 2, numpy.float64, "fused"),
    ((6, 7, 8), 2, numpy.float32, "two_pass"),       # S from oS after every H half-step
    ((3, 4, 700), 2, numpy.float32, "fused"),        # 64-cell z segments
])
def test_step_n_observed_equals_single_steps(shape, T, dtype, mode):
    """fdtd_step_n_observed (one call of n steps) against one step per call
    with fdtd_plane_sum and clones after every step: every E and H copy and
    the final fields and psi blocks bit for bit; the flux, an FP64 sum of
    the same FP32 S values in another order, against the exactly rounded
    sum (math.fsum) of those values to 1e-13."""
    import ctypes
    from opencl_fdtd_amd import Simulation, _lib
    print("case step_n_observed", shape, T, dtype.__name__, mode, flush=True)
    eps, pmls, init, _ = _case(shape, T, seed=sum(shape), pec=True)
    mk = lambda: Simulation([numpy.asarray(e) for e in eps], [dict(p) for p in pmls], initial_fields=init,
                            float_type=dtype, poynting_mode=mode)
    a, b = mk(), mk()
    n = 9
    idx = 2 * a.N + (shape[0] // 2) * shape[1] * shape[2] + (shape[1] // 3) * shape[2] + shape[2] // 2
    wave = numpy.sin(0.4 * numpy.arange(n)).astype(dtype)
    comp, axis, index = 2, 1, shape[1] // 2
    acc_a = torch.zeros(1, dtype=torch.float64, device="cuda")
    acc_b = torch.zeros(1, dtype=torch.float64, device="cuda")
    keep_E = [torch.empty_like(a._E) for _ in range(n)]
    keep_H = [torch.empty_like(a._E) for _ in range(n)]
    a.step_n_observed(n, idx, wave, flux=(comp, axis, index, acc_a), keep_E=keep_E, keep_H=keep_H)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ref_E, ref_H, values = [], [], []
    for k in range(n):
        b.step_n(1, idx, wave[k:k + 1])
        b.flush()                 # S of the step (step_n leaves its last H half-step pending)
        _lib.check(b._lib.fdtd_plane_sum(b.plan, ctypes.c_void_p(b._S.data_ptr()), comp, axis, index,
                                         ctypes.c_void_p(acc_b.data_ptr()), st), "fdtd_plane_sum")
        Sk = fetch(b, "S").reshape(3, *shape)[comp]
        values.extend(numpy.take(Sk, index, axis=axis).astype(numpy.float64).ravel().tolist())
        ref_E.append(b.E.clone())
        ref_H.append(b.H.clone())
    torch.cuda.synchronize()
    exact = math.fsum(values)
    assert exact != 0.0
    for got in (float(acc_a.item()), float(acc_b.item())):
        assert abs(got - exact) <= 1e-13 * sum(abs(v) for v in values), (got, exact)
    for k in range(n):
        assert torch.equal(keep_E[k], ref_E[k]), ("E after step", k)
        assert torch.equal(keep_H[k], ref_H[k]), ("H after step", k)
    _assert_same(_state(a), _state(b), "observed vs single steps")


def test_step_n_observed_rejects_bad_observations():
    from opencl_fdtd_amd import Simulation
    shape = (6, 7, 8)
    eps = [numpy.ones(shape)] * 3
    s = Simulation(eps, [])
    acc = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(Exception):
        s.step_n_observed(3, flux=(2, 3, 0, acc))                 # axis out of range
    with pytest.raises(Exception):
        s.step_n_observed(3, flux=(2, 1, 7, acc))                 # plane beyond sy
    with pytest.raises(Exception):
        s.step_n_observed(3, flux=(2, 1, 0, acc.float()))         # not FP64
    with pytest.raises(Exception):
        s.step_n_observed(3, keep_E=[torch.empty_like(s._E)] * 2)   # n tensors needed
    with pytest.raises(Exception):
        Simulation(eps, [], do_poynting=False).step_n_observed(2, flux=(2, 1, 0, acc))   # no S


@pytest.mark.parametrize("shape, dtype", [((5, 600, 500), numpy.float32), ((4, 300, 200), numpy.float64)])
def test_step_n_observed_flux_on_every_axis(shape, dtype):
    """The observed flux summed over many workgroups (a 600 x 500 plane:
    586 workgroups of 512 values) on planes normal to x, y and z, against
    the exactly rounded sum (math.fsum) of the S values after each step."""
    from opencl_fdtd_amd import Simulation
    print("case observed flux", shape, dtype.__name__, flush=True)
    rng = numpy.random.default_rng(7)
    eps = [rng.uniform(1, 3, shape) for _ in range(3)]
    init = {"E": random_fields(rng, shape), "H": random_fields(rng, shape)}
    for axis in range(3):
        sim = Simulation(eps, [], initial_fields=init, float_type=dtype)
        acc = torch.zeros(1, dtype=torch.float64, device="cuda")
        index = shape[axis] // 3
        values = []
        for k in range(2):
            sim.step_n_observed(1, flux=(axis, axis, index, acc))
            S = fetch(sim, "S").reshape(3, *shape)[axis]
            values.extend(numpy.take(S, index, axis=axis).astype(numpy.float64).ravel().tolist())
        exact = math.fsum(values)
        got = float(acc.item())
        assert abs(got - exact) <= 1e-13 * sum(abs(v) for v in values), (axis, got, exact)
        del sim
        torch.cuda.empty_cache()


def test_random_shapes_fused_equals_unfused():
    """A seeded sweep of 24 grids (rows of 1 to 1500 cells, 1 to 40 rows,
    2 to 9 planes; CPML thickness 0 to 3): step_n through whichever passes
    the plan picks (whole-row or z-segment tiles, one or two steps per pass)
    against the unfused kernels, and update_HE2 against update_HE twice
    where the two-step pass is supported, bit for bit."""
    rng = numpy.random.default_rng(2024)
    for case in range(24):
        shape = (int(rng.integers(2, 10)), int(rng.integers(1, 41)),
                 int(rng.choice([int(rng.integers(1, 80)), int(rng.integers(80, 600)), int(rng.integers(600, 1500))])))
        T = int(rng.integers(0, 4))
        T = min(T, min(shape) // 2)
        n = int(rng.integers(2, 7))
        print("case random", case, shape, T, n, flush=True)
        eps, pmls, init, _ = _case(shape, T, seed=100 + case, pec=bool(case % 3 == 0))
        a = _sim(eps, pmls, init, None, numpy.float32, True)
        b = _sim(eps, pmls, init, None, numpy.float32, False)
        idx = (shape[0] // 2) * shape[1] * shape[2] + (shape[1] // 2) * shape[2] + shape[2] // 2
        wave = numpy.cos(0.3 * numpy.arange(n)).astype(numpy.float32)
        a.step_n(n, idx, wave)
        b.step_n(n, idx, wave)
        torch.cuda.synchronize()
        _assert_same(_state(a), _state(b), "random %s" % (shape,))
        if a._psi_e_out is not None:
            c = _sim(eps, pmls, init, None, numpy.float32, True)
            d = _sim(eps, pmls, init, None, numpy.float32, True)
            amps = torch.tensor([0.5, -0.25], dtype=torch.float32, device="cuda")
            c.update_HE2([], idx, amps)
            d.update_HE([], idx, amps[0:1])
            d.update_HE([], idx, amps[1:2])
            torch.cuda.synchronize()
            _assert_same(_state(c), _state(d), "random HE2 %s" % (shape,))


@pytest.mark.parametrize("shape, T, n, dtype, nonuniform", [
    ((11, 13, 9), 3, 5, numpy.float32, False),
    ((30, 40, 73), 4, 6, numpy.float32, False),      # the two-step pass
    ((3, 4, 700), 2, 4, numpy.float32, False),       # z segments
    ((33, 8, 17), 3, 5, numpy.float64, True),
    ((5, 6, 160), 2, 4, numpy.float32, True),
])
def test_poynting_off_fused_equals_unfused(shape, T, n, dtype, nonuniform):
    """do_poynting=False: step_n through the fused passes (without S) against
    the separate kernels, bit for bit (E, H, both psi blocks)."""
    from opencl_fdtd_amd import Simulation
    print("case poynting off", shape, T, n, dtype.__name__, nonuniform, flush=True)
    eps, pmls, init, dxes = _case(shape, T, seed=sum(shape) + 3, nonuniform=nonuniform, pec=True)
    mk = lambda fuse: Simulation([numpy.asarray(e) for e in eps], [dict(p) for p in pmls], dxes=dxes,
                                 initial_fields=init, float_type=dtype, do_poynting=False, fuse_steps=fuse)
    a, b = mk(True), mk(False)
    assert a._E_out is not None and b._E_out is None and a.S is None
    idx = 2 * a.N + (shape[0] // 2) * shape[1] * shape[2] + (shape[1] // 3) * shape[2] + shape[2] // 2
    wave = numpy.sin(0.3 * numpy.arange(3 * n)).astype(dtype)
    for k in range(3):
        a.step_n(n, idx, wave[k * n:(k + 1) * n])
        b.step_n(n, idx, wave[k * n:(k + 1) * n])
    torch.cuda.synchronize()
    for name in ("E", "H"):
        assert numpy.array_equal(fetch(a, name), fetch(b, name)), name
    assert torch.equal(a._psi_e, b._psi_e) and torch.equal(a._psi_h[a._h], b._psi_h[b._h])
    if a._psi_e_out is not None:                      # the two-step pass itself, without S
        c, d = mk(True), mk(True)
        amps = torch.tensor([0.5, -0.25], dtype=torch.float32, device="cuda")
        c.update_HE2([], idx, amps)
        d.update_HE([], idx, amps[0:1])
        d.update_HE([], idx, amps[1:2])
        torch.cuda.synchronize()
        for name in ("E", "H"):
            assert numpy.array_equal(fetch(c, name), fetch(d, name)), ("HE2", name)


def test_poynting_off_matches_oracle_and_loop():
    """do_poynting=False against the oracle: step_n (two-step pass where
    chosen) and the per-call loop, whose update_H is deferred into the next
    update_E; update_S is None as in the reference."""
    shape, T = (40, 30, 73), 4
    eps, pmls, init, _ = _case(shape, T, seed=21, pec=True)
    sim, orc = build_pair(eps, pmls, init=init, do_poynting=False)
    assert sim.update_S is None and sim._defer
    idx = 1 * sim.N + (shape[0] // 2) * shape[1] * shape[2] + (shape[1] // 2) * shape[2] + shape[2] // 2
    wave = numpy.cos(0.2 * numpy.arange(30))
    sim.step_n(17, idx, wave[:17].astype(numpy.float32))
    for t in range(17, 30):
        sim.update_E([]).wait()
        sim.E[idx] += float(numpy.float32(wave[t]))
        sim.update_H([]).wait()
    step_oracle(orc, 30, idx, wave)
    for name in ("E", "H"):
        g, o = fetch(sim, name), getattr(orc, name)
        bad = numpy.flatnonzero(g != o)
        assert bad.size == 0, (name, bad.size, bad[:4])
