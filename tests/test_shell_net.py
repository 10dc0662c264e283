"""The learned shell's convolution: the CPU oracle pinned by its own
transpose and finite-difference identities (CPU), and the HIP forward and
backward kernels against the oracle (GPU)."""
import ctypes

import numpy
import pytest

from oracle import shell_net as SN


def _rand_weights(rng, scale=0.05):
    return SN.bf16_round(rng.standard_normal(SN.N_WEIGHTS) * scale).astype(numpy.float64)


def _fields(rng, shape):
    return rng.standard_normal((3,) + shape), rng.standard_normal((3,) + shape)


def test_oracle_boxes_partition_the_shell():
    shape, T = (12, 11, 13), 3
    cover = numpy.zeros(shape, int)
    for box in SN.boxes(shape, T):
        cover += SN.box_mask(shape, box)
    inner = numpy.zeros(shape, bool)
    inner[T:-T, T:-T, T:-T] = True
    assert numpy.all(cover[inner] == 0) and numpy.all(cover[~inner] == 1)


def test_oracle_backward_is_the_transpose():
    rng = numpy.random.default_rng(401)
    shape, T = (10, 9, 11), 2
    w = _rand_weights(rng)
    E, H = _fields(rng, shape)
    Ebar = rng.standard_normal((3,) + shape)
    _, y = SN.forward(E, H, w, T, exact_f32=False)
    xbE, xbH, _ = SN.backward(E, H, w, T, Ebar)
    lhs = float(numpy.sum(Ebar * y))
    rhs = float(numpy.sum(xbE * E) + numpy.sum(xbH * H))
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), 1)


def test_oracle_weight_gradient_finite_differences():
    """<Ebar, y(W)> is bilinear in (W1, W2): its derivative along a weight
    direction equals the central difference exactly (up to rounding)."""
    rng = numpy.random.default_rng(402)
    shape, T = (9, 8, 10), 2
    w = _rand_weights(rng)
    E, H = _fields(rng, shape)
    Ebar = rng.standard_normal((3,) + shape)
    _, _, wbar = SN.backward(E, H, w, T, Ebar)
    f = lambda ww: float(numpy.sum(Ebar * SN.forward(E, H, ww, T, exact_f32=False)[1]))
    for _ in range(6):
        d = rng.standard_normal(w.size)
        h = 1e-3
        fd = (f(w + h * d) - f(w - h * d)) / (2 * h)
        assert abs(fd - float(wbar @ d)) <= 1e-8 * max(abs(fd), 1)
    # a few single weights in W1 and W2
    for idx in (5, SN.W1_SIZE * 3 + 77, 6 * SN.W1_SIZE + 10, 6 * SN.W1_SIZE + 21 * SN.W2_SIZE + 400,
                6 * SN.W1_SIZE + 9 * SN.W2_SIZE + 1000):
        d = numpy.zeros(w.size)
        d[idx] = 1.0
        fd = (f(w + 1e-3 * d) - f(w - 1e-3 * d)) / 2e-3
        assert abs(fd - wbar[idx]) <= 1e-8 * max(abs(fd), 1)


def test_oracle_depth_heads():
    """Layer-2 head d acts only on the cells at depth d of their box (index
    d along the box's thin axis, counted from the box's low side)."""
    rng = numpy.random.default_rng(406)
    shape, T = (11, 10, 12), 3
    E, H = _fields(rng, shape)
    w = _rand_weights(rng)
    n1 = 6 * SN.W1_SIZE
    for b in (0, 3, 5):
        for d in range(T):
            ww = w.copy()
            ww[n1:] = 0.0
            head = n1 + (b * SN.N_DEPTH + d) * SN.W2_SIZE
            ww[head:head + SN.W2_SIZE] = w[head:head + SN.W2_SIZE]
            _, y = SN.forward(E, H, ww, T, exact_f32=False)
            (lo, hi) = SN.boxes(shape, T)[b]
            want = SN.box_mask(shape, (lo, hi))
            r = b // 2
            sel = [slice(None)] * 3
            keep = numpy.zeros(shape, bool)
            sel[r] = lo[r] + d
            keep[tuple(sel)] = True
            want &= keep
            support = numpy.abs(y).sum(axis=0) > 0
            assert numpy.array_equal(support, want), (b, d)


def test_oracle_composed_kernel_is_the_two_layers():
    """The 5x5x5 kernels the inference uses are the composition of the two
    3x3x3 layers: the composed form equals the two-layer form in float64 on a
    grid where every box has cells of every depth and taps wrap, so a wrong
    tap offset, axis order or depth head fails."""
    rng = numpy.random.default_rng(406)
    w = _rand_weights(rng)
    for shape, T in (((12, 11, 13), 2), ((13, 12, 14), 4), ((9, 14, 10), 3)):
        E, H = _fields(rng, shape)
        _, y2 = SN.forward(E, H, w, T, exact_f32=False)
        yc = SN.forward_composed(E, H, w, T)
        assert numpy.abs(y2).max() > 0.1
        assert numpy.abs(yc - y2).max() <= 1e-12 * numpy.abs(y2).max()


def test_oracle_composed_weights_exact_f32():
    """The float32 composition (the GPU's order) rounds to BF16 values that
    lie within one BF16 step of the float64 composition."""
    rng = numpy.random.default_rng(407)
    w = _rand_weights(rng)
    c32 = SN.compose(w, True).astype(numpy.float64)
    c64 = SN.compose(w, False)
    assert numpy.array_equal(c32, SN.bf16_round(c32).astype(numpy.float64))
    assert numpy.all(numpy.abs(c32 - c64) <= 2.0 ** -8 * numpy.abs(c64) + 1e-30)


def test_oracle_zero_weights_identity():
    rng = numpy.random.default_rng(403)
    shape = (8, 8, 8)
    E, H = _fields(rng, shape)
    Eo, y = SN.forward(E, H, numpy.zeros(SN.N_WEIGHTS), 2)
    assert numpy.array_equal(Eo, E.astype(numpy.float32)) and not y.any()


def _gpu_run(shape, T, w, E, H, Ebar=None):
    import torch
    from opencl_fdtd_amd import Simulation
    sim = Simulation([numpy.ones(shape)] * 3, [], do_poynting=False)
    lib = sim._lib
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    wb = torch.as_tensor(w, dtype=torch.float32).to(torch.bfloat16).cuda()
    assert numpy.array_equal(wb.float().cpu().numpy(), w.astype(numpy.float32))
    nh = int(lib.fdtd_shell_hidden_elems(sim.plan, T))
    hid = torch.zeros(nh, dtype=torch.float32, device="cuda")
    Et = torch.as_tensor(E.ravel(), dtype=torch.float32, device="cuda").contiguous()
    Ht = torch.as_tensor(H.ravel(), dtype=torch.float32, device="cuda").contiguous()
    if Ebar is None:
        assert lib.fdtd_shell_apply(sim.plan, T, ctypes.c_void_p(wb.data_ptr()), ctypes.c_void_p(Et.data_ptr()),
                                    ctypes.c_void_p(Ht.data_ptr()), ctypes.c_void_p(hid.data_ptr()), st) == 0
        torch.cuda.synchronize()
        return Et.cpu().numpy().reshape((3,) + shape)
    hbar = torch.zeros(nh, dtype=torch.float32, device="cuda")
    Eb = torch.as_tensor(Ebar.ravel(), dtype=torch.float32, device="cuda").contiguous()
    Ebo = torch.zeros_like(Eb)
    Hbo = torch.zeros_like(Eb)
    parts = torch.zeros(int(lib.fdtd_shell_partials_elems()), dtype=torch.float32, device="cuda")
    wg = torch.zeros(w.size, dtype=torch.float32, device="cuda")
    rc = lib.fdtd_shell_backward(sim.plan, T, ctypes.c_void_p(wb.data_ptr()), ctypes.c_void_p(Et.data_ptr()),
                                 ctypes.c_void_p(Ht.data_ptr()), ctypes.c_void_p(hid.data_ptr()),
                                 ctypes.c_void_p(hbar.data_ptr()), ctypes.c_void_p(Eb.data_ptr()),
                                 ctypes.c_void_p(Ebo.data_ptr()), ctypes.c_void_p(Hbo.data_ptr()),
                                 ctypes.c_void_p(parts.data_ptr()), ctypes.c_void_p(wg.data_ptr()), st)
    assert rc == 0
    torch.cuda.synchronize()
    return (Ebo.cpu().numpy().reshape((3,) + shape), Hbo.cpu().numpy().reshape((3,) + shape), wg.cpu().numpy())


@pytest.mark.gpu
@pytest.mark.parametrize("shape, T", [((12, 11, 13), 2), ((16, 14, 12), 4), ((13, 12, 17), 3),
                                      ((20, 37, 25), 4), ((8, 9, 30), 2), ((31, 8, 8), 2),
                                      ((33, 29, 47), 3)])
def test_hip_forward_matches_oracle(shape, T):
    """The MFMA inference kernel (the composed 5x5x5 convolution): same BF16
    operands as the oracle's FP32 emulation (inputs and composed weights),
    FP32 accumulation in another order, so the correction agrees to
    accumulation rounding; cells outside the shell are untouched."""
    rng = numpy.random.default_rng(404)
    w = _rand_weights(rng)
    E, H = _fields(rng, shape)
    E, H = E.astype(numpy.float32), H.astype(numpy.float32)
    got = _gpu_run(shape, T, w, E, H)
    want, y = SN.forward(E, H, w, T, exact_f32=True)
    inner = numpy.zeros(shape, bool)
    inner[T:-T, T:-T, T:-T] = True
    assert numpy.array_equal(got[:, inner], E[:, inner])
    err = numpy.abs((got - E) - y).max()
    print("max |y_gpu - y_oracle| = %.3e of max |y| = %.3e" % (err, numpy.abs(y).max()))
    assert err <= 2e-3 * numpy.abs(y).max()


@pytest.mark.gpu
@pytest.mark.parametrize("shape, T", [((12, 11, 13), 2), ((16, 14, 12), 4), ((33, 29, 47), 3),
                                      ((20, 37, 25), 4), ((8, 9, 30), 2)])
def test_hip_backward_matches_oracle(shape, T):
    """The MFMA backward (FP32 matrix cores: hidden layer, both transposed
    layers, both weight gradients) against the FP64 oracle; the larger
    shapes give every box many 16-cell groups and every weight-gradient
    workgroup a slice of several hundred cells."""
    print("case shell backward", shape, T, flush=True)
    rng = numpy.random.default_rng(405)
    w = _rand_weights(rng)
    E, H = _fields(rng, shape)
    E = SN.bf16_round(E).astype(numpy.float64)   # the GPU rounds its inputs to BF16
    H = SN.bf16_round(H).astype(numpy.float64)
    Ebar = rng.standard_normal((3,) + shape).astype(numpy.float32).astype(numpy.float64)
    gE, gH, gw = _gpu_run(shape, T, w, E.astype(numpy.float32), H.astype(numpy.float32), Ebar)
    oE, oH, ow = SN.backward(E, H, w, T, Ebar)
    for g, o in ((gE, oE), (gH, oH)):
        assert numpy.abs(g - o).max() <= 2e-5 * numpy.abs(o).max()
    # the W2 gradient uses BF16-rounded hidden values on the GPU
    n1 = 6 * SN.W1_SIZE
    assert numpy.abs(gw[:n1] - ow[:n1]).max() <= 2e-5 * numpy.abs(ow[:n1]).max()
    assert numpy.abs(gw[n1:] - ow[n1:]).max() <= 1e-2 * numpy.abs(ow[n1:]).max()


@pytest.mark.gpu
@pytest.mark.parametrize("poynting, period", [(False, 1), (True, 1), (False, 2), (True, 2), (False, 4), (True, 4)])
def test_shell_step_n_matches_stepwise(poynting, period):
    """fdtd_shell_step_n (one native launch loop; each correction is added by
    the next step's E update, with period 1 the last one by a final pass;
    with period 2 or 4 each pair of steps is one two-step pass that adds a
    pending correction in its first E update) against the same steps made
    one call at a time: E step + source, H step from that E, then, on a
    corrected step, E += the correction of that E and the H before the
    step."""
    import torch
    from opencl_fdtd_amd import Simulation, learned as L
    steps, split = {1: (41, 17), 2: (42, 18), 4: (44, 20)}[period]
    print("case shell step_n poynting=%s period=%d" % (poynting, period), flush=True)
    sc = L.make_scenario(7, interior=(24, 20, 28), steps=steps, near_face=6, src_margin=4)
    w = torch.as_tensor(_rand_weights(numpy.random.default_rng(8), 0.02), dtype=torch.float32)
    net = L.ShellNet(4, weights=w, period=period)
    shape, eps, idx, sl = L.embed(sc, 4)
    wave = torch.as_tensor(sc.wave(), dtype=torch.float32, device="cuda")
    a = Simulation(eps, net.faces(), dt=sc.dt, do_poynting=poynting)
    b = Simulation(eps, net.faces(), dt=sc.dt, do_poynting=poynting)
    if period > 1:
        assert a._psi_e_out is not None     # the plan has the two-step pass
    ra, rb = L.ShellRunner(a, net), L.ShellRunner(b, net)
    ra.step_n(split, idx, wave[:split])
    ra.step_n(sc.steps - split, idx, wave[split:])
    for n in range(sc.steps):
        rb.step(idx, wave[n])
    torch.cuda.synchronize()
    assert float(torch.abs(a.E).max()) > 0
    assert torch.equal(a.E, b.E) and torch.equal(a.H, b.H)
    if poynting:
        assert torch.equal(a.S, b.S)


@pytest.mark.gpu
@pytest.mark.parametrize("period", [1, 2, 4])
def test_trainer_gradient_directional_derivative(period):
    """The training gradient (reverse pass through the solver adjoint and
    the correction's backward kernels, with checkpointed recompute; the
    correction on every step, or every second or fourth) against central
    differences of the training loss along random directions of each
    layer's weights (the realised BF16 weight change).  The difference
    converges to the adjoint as O(t^2) down to the rounding floor of the
    forward (inputs and composed 5x5x5 weights rounded to BF16, a fixed
    error per weight that the two points of the difference see differently).
    A small correction (W2 ~ 0.003) keeps the run nearly linear in the
    weights, so a step of t = 0.3 keeps both terms under the 2 % bar.  Each layer separately, since along
    a direction over both their contributions nearly cancel."""
    import torch
    from opencl_fdtd_amd import learned as L
    print("case trainer gradient period=%d" % period, flush=True)
    sc = L.make_scenario(11, interior=(20, 18, 22), steps=44, near_face=5, src_margin=4, width_wl=0.5)
    rng = numpy.random.default_rng(12)
    w = rng.standard_normal(L.N_WEIGHTS)
    n1 = 6 * L.W1_SIZE
    w[:n1] *= 0.05
    w[n1:] *= 0.003
    net = L.ShellNet(4, weights=torch.as_tensor(w, dtype=torch.float32), base_pml=dict(L.BASE_PML), period=period)
    tr = L.ShellTrainer(net, [sc], checkpoint_every=16)
    loss0, g = tr.loss_and_grad(tr.cases[0])
    d_all = torch.as_tensor(rng.standard_normal(L.N_WEIGHTS) * numpy.abs(w), dtype=torch.float32, device=g.device)
    base = net.master.clone()
    t = 0.3
    for name, part in (("W1", slice(0, n1)), ("W2", slice(n1, None))):
        d = torch.zeros_like(d_all)
        d[part] = d_all[part]
        losses, realised = [], []
        for sgn in (1.0, -1.0):
            net.master.copy_(base + sgn * t * d)
            net.sync()
            realised.append(net.bf16.float())
            losses.append(tr.loss_and_grad(tr.cases[0])[0])
        fd = losses[0] - losses[1]
        pred = float(torch.dot(g.double(), (realised[0] - realised[1]).double()))
        print("%s: loss %.6e  central difference %.6e  adjoint %.6e" % (name, loss0, fd, pred))
        assert abs(fd) > 1e-3 * loss0
        assert abs(fd - pred) <= 0.02 * abs(fd)
    # checkpoint spacing changes nothing: the recompute is bit-exact
    net.master.copy_(base)
    net.sync()
    tr.K = 7
    assert torch.equal(tr.loss_and_grad(tr.cases[0])[1], g)


@pytest.mark.gpu
def test_stability_detectors_agree():
    """The two detectors of a growing mode: the 12000-step trace after a
    pulse with no static part (is_stable, model selection) and the
    trainer's stability term (the growth of a 2000-step run started from
    the harvested end of that pulse run, whose gradient training follows).
    The base CPML-4 is stable to both (the term is ~0); a random net of
    sizeable weights feeds a growing mode and both flag it."""
    import torch
    from opencl_fdtd_amd import learned as L
    print("case stability detectors", flush=True)
    rng = numpy.random.default_rng(31)
    w = rng.standard_normal(L.N_WEIGHTS)
    n1 = 6 * L.W1_SIZE
    w[:n1] *= 0.05
    w[n1:] *= 0.001
    for name, wt, grows in (("base", numpy.zeros(L.N_WEIGHTS), False), ("random", w, True)):
        net = L.ShellNet(4, weights=torch.as_tensor(wt, dtype=torch.float32), period=4)
        trace = L.stability_trace(net, steps=8000)
        tr = L.ShellTrainer(net, [], checkpoint_every=60)
        tr.stability_setup()
        term, g = tr.stability_loss_and_grad()
        print(name, "term %.3f" % term, " ".join("%.1e" % v for v in trace))
        assert L.is_stable(trace) == (not grows), name
        if grows:
            assert term > 0.5 and float(torch.linalg.vector_norm(g)) > 0
        else:
            assert abs(term) < 0.05


@pytest.mark.gpu
def test_zero_net_is_the_base_cpml():
    """With zero weights the shell step is exactly its base CPML-4 (the
    fallback): fields equal to the plain solver's after many steps."""
    import torch
    from opencl_fdtd_amd import Simulation, learned as L
    sc = L.make_scenario(9, interior=(30, 26, 34), steps=60, near_face=6, src_margin=4)
    net = L.ShellNet(4, weights=torch.zeros(L.N_WEIGHTS))
    shape, eps, idx, sl = L.embed(sc, 4)
    wave = torch.as_tensor(sc.wave(), dtype=torch.float32, device="cuda")
    a = Simulation(eps, net.faces(), dt=sc.dt, do_poynting=False)
    b = Simulation(eps, net.faces(), dt=sc.dt, do_poynting=False)
    L.ShellRunner(a, net).step_n(sc.steps, idx, wave)
    b.step_n(sc.steps, idx, wave)
    torch.cuda.synchronize()
    assert float(torch.abs(b.E).max()) > 0
    assert torch.equal(a.E, b.E) and torch.equal(a.H, b.H)


@pytest.mark.gpu
def test_trained_shell_gain_over_its_base():
    """The shipped net against the method without it (the same CPML-4, zero
    weights, same correction schedule) on unseen 256x256x128 scenarios, 600
    steps: reflection (interior field error over the run, every 4th step, vs
    the scenario padded by 48 cells of CPML) improves in every scenario (by
    more than 0.5 dB), by at least 3 dB on average, and stays at or below
    -40 dB."""
    import torch
    from opencl_fdtd_amd import learned as L
    net = L.ShellNet.load(L.WEIGHTS)
    assert net.period == 4
    base = L.ShellNet(net.T, weights=torch.zeros(L.N_WEIGHTS), base_pml=net.base_pml, period=net.period)
    gains = []
    for seed in (1000, 1001, 1002, 1003):
        sc = L.make_scenario(seed, interior=(256, 256, 128), steps=600)
        wide = L.run_interior(sc, L.WIDE_PAD, L.cpml_faces(L.WIDE_PAD), sample_every=4)
        db_net = L.reflection_db(L.run_learned(sc, net, sample_every=4), wide)
        db_base = L.reflection_db(L.run_learned(sc, base, sample_every=4), wide)
        del wide
        torch.cuda.empty_cache()
        print("seed %d: base %.2f dB, learned %.2f dB" % (seed, db_base, db_net))
        assert db_net <= -40.0
        gains.append(db_base - db_net)
    assert min(gains) > 0.5
    assert float(numpy.mean(gains)) >= 3.0


@pytest.mark.gpu
def test_trained_shell_long_run_is_stable():
    """Long runs with the trained shell.  After a pulse with no static part
    (a derivative of a Gaussian) the interior of a stable absorber empties
    to a flat round-off floor: over 12000 steps the energy of the second
    half must not rise, as with the base CPML-4.  (A learned correction can
    feed a slowly growing mode that a run of a few thousand steps, or a
    pulse that leaves a static field behind, does not show.)  And 4000
    steps after the scenario pulse the energy decays like the base's."""
    import torch
    from opencl_fdtd_amd import learned as L
    net = L.ShellNet.load(L.WEIGHTS)
    base = L.ShellNet(net.T, weights=torch.zeros(L.N_WEIGHTS), base_pml=net.base_pml, period=net.period)
    for name, nn in (("base", base), ("learned", net)):
        tr = L.stability_trace(nn, steps=12000)
        print(name, " ".join("%.2e" % v for v in tr))
        assert L.is_stable(tr), name
    sc = L.make_scenario(21, interior=(64, 64, 48), steps=4000)
    energy = {}
    for name, nn in (("learned", net), ("base", base)):
        e = [float(torch.sum(t.double() ** 2)) for t in L.run_learned(sc, nn, sample_every=200)]
        energy[name] = e
        print(name, ["%.3e" % v for v in e])
    e = energy["learned"]
    peak = max(e)
    assert e[-1] < 1e-3 * peak
    assert e[-1] <= 1.05 * e[-6]                   # not rising over the last 1000 steps
    assert e[-1] <= 2.0 * energy["base"][-1]
