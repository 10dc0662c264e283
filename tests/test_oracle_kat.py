"""Pins of the CPU oracle (no GPU): closed forms, invariants, special cases.

Each test is chosen so that a plausible restatement mistake (a dropped term,
a flipped sign, a wrong neighbour offset or staggering, a wrong factor)
fails it:

* two independent restatements (per-cell C++, vectorised numpy) agree bit for
  bit in float32 and float64;
* a discrete plane-wave eigenmode evolves exactly as the closed-form Yee
  dispersion predicts (all curl terms, dt/eps, 1/dx);
* the backward divergence of eps*E and the forward divergence of H are
  constant without sources (curl signs and offsets);
* the leapfrog energy sum eps*E_k^2 + H_k.H_{k-1} is conserved in a lossless
  periodic box with random anisotropic eps;
* the Poynting vector of a travelling plane wave along each axis, each
  polarisation and each direction has the closed-form flux (staggering,
  signs, the dt/8 factor, all six partial products);
* CPML with sigma = 0 is the periodic step bit for bit; CPML absorbs, and
  its reflection falls with thickness;
* PEC cells keep E; a Bloch phase of 1 is the periodic step; a Bloch cell
  equals the corresponding cells of a periodic supercell.
"""
import math

import numpy
import pytest

from oracle.numpy_ref import NumpyFDTD
from oracle.oracle import OracleSimulation, pml_coefficients, pml_defaults, unvec, vec


def _run(o, steps):
    for t in range(steps):
        o.step(None, t)


def _as3(v, shape):
    return numpy.array(unvec(v, shape))


# ---------------------------------------------------------------- restatements
@pytest.mark.parametrize("dtype", [numpy.float32, numpy.float64])
@pytest.mark.parametrize("bloch", [False, True])
def test_two_restatements_bit_identical(dtype, bloch):
    rng = numpy.random.default_rng(11)
    shape = (9, 8, 10)
    eps = [rng.uniform(1, 4, shape) for _ in range(3)]
    eps[1][rng.random(shape) < 0.05] = numpy.inf
    if bloch:
        bl = [{"axis": "x", "real": math.cos(0.9), "imag": math.sin(0.9)},
              {"axis": "z", "real": math.cos(-2.1), "imag": math.sin(-2.1)}]
        pmls = [{"axis": "y", "polarity": p, "thickness": 3} for p in "np"]
    else:
        bl = []
        pmls = [{"axis": a, "polarity": p, "thickness": 3, "cfs_alpha": 0.02} for a in "xyz" for p in "np"]
    init = {k: [rng.standard_normal(shape) for _ in range(3)] for k in "EHFG"}
    o = OracleSimulation(eps, [dict(p) for p in pmls], bloch_boundaries=bl, initial_fields=init, float_type=dtype)
    n = NumpyFDTD(numpy.array(eps), pmls=[dict(p) for p in pmls], dtype=dtype, bloch=bl)
    for k in "EHFG":
        getattr(n, k)[:] = numpy.array(init[k])
    for t in range(12):
        o.step()
        n.step()
    names = "EHSFG" if bloch else "EHS"
    for name in names:
        a = getattr(o, name)
        b = vec(getattr(n, name))
        assert numpy.array_equal(a, b), name


# ---------------------------------------------------------------- dispersion
def _eigenmode(shape, kvec, eps, dt, a, dx=1.0):
    """Complex amplitudes (E_hat, H_hat, lambda) of the discrete Yee mode with
    wave vector kvec (radians per cell): E_{n} = E_hat lam^n, H after step n
    = H_hat lam^n; derived from E' = E + dt/eps Kb x H, H' = H - dt Kf x E'."""
    k = numpy.asarray(kvec, dtype=float)
    Kb = (1 - numpy.exp(-1j * k)) / dx
    Kf = (numpy.exp(1j * k) - 1) / dx
    s2 = numpy.sum(numpy.sin(k / 2) ** 2) / dx ** 2
    theta = 2 * math.asin(dt * math.sqrt(s2 / eps))
    lam = numpy.exp(-1j * theta)
    Eh = numpy.cross(numpy.asarray(a, dtype=complex), Kb)      # Kb . Eh = 0
    Hh = -dt * lam / (lam - 1) * numpy.cross(Kf, Eh)
    return Eh, Hh, lam


def _mode_fields(shape, kvec, amp, lam_pow):
    idx = numpy.indices(shape)
    phase = numpy.exp(1j * sum(kvec[r] * idx[r] for r in range(3)))
    return [numpy.real(amp[c] * lam_pow * phase) for c in range(3)]


@pytest.mark.parametrize("dx", [1.0, 0.7])
def test_yee_dispersion_closed_form(dx):
    shape = (8, 10, 12)
    eps_val = 2.25
    m = (1, -2, 3)
    kvec = [2 * math.pi * m[r] / shape[r] for r in range(3)]
    dt = dx * 0.99 / math.sqrt(3)
    Eh, Hh, lam = _eigenmode(shape, kvec, eps_val, dt, a=(0.3, 1.0, -0.5), dx=dx)
    init = {"E": _mode_fields(shape, kvec, Eh, 1.0), "H": _mode_fields(shape, kvec, Hh, 1.0)}
    eps = [numpy.full(shape, eps_val) for _ in range(3)]
    o = OracleSimulation(eps, [], initial_fields=init, float_type=numpy.float64, dt=dt, dxes=dx, do_poynting=False)
    steps = 60
    _run(o, steps)
    E_ref = vec(_mode_fields(shape, kvec, Eh, lam ** steps))
    H_ref = vec(_mode_fields(shape, kvec, Hh, lam ** steps))
    scale = numpy.abs(E_ref).max()
    assert numpy.abs(o.E - E_ref).max() < 1e-11 * scale
    assert numpy.abs(o.H - H_ref).max() < 1e-11 * numpy.abs(H_ref).max()
    # the mode really moved: a wrong dispersion would be off by O(1)
    assert numpy.abs(lam ** steps - 1) > 0.5


# ---------------------------------------------------------------- invariants
def _div_back(F, shape):
    return sum(F[r] - numpy.roll(F[r], 1, axis=r) for r in range(3))


def _div_fwd(F, shape):
    return sum(numpy.roll(F[r], -1, axis=r) - F[r] for r in range(3))


def test_divergence_preserved():
    rng = numpy.random.default_rng(12)
    shape = (10, 9, 11)
    eps = [rng.uniform(1, 6, shape) for _ in range(3)]
    init = {"E": [rng.standard_normal(shape) for _ in range(3)],
            "H": [rng.standard_normal(shape) for _ in range(3)]}
    o = OracleSimulation(eps, [], initial_fields=init, float_type=numpy.float64, do_poynting=False)
    epsv = numpy.array(eps)
    dE0 = _div_back(epsv * _as3(o.E, shape), shape)
    dH0 = _div_fwd(_as3(o.H, shape), shape)
    _run(o, 40)
    dE1 = _div_back(epsv * _as3(o.E, shape), shape)
    dH1 = _div_fwd(_as3(o.H, shape), shape)
    assert numpy.abs(dE1 - dE0).max() < 1e-11
    assert numpy.abs(dH1 - dH0).max() < 1e-11
    # and the fields did change
    assert numpy.abs(_as3(o.E, shape) - numpy.array(init["E"])).max() > 0.1


def test_leapfrog_energy_conserved():
    rng = numpy.random.default_rng(13)
    shape = (9, 10, 8)
    eps = [rng.uniform(1, 5, shape) for _ in range(3)]
    init = {"E": [rng.standard_normal(shape) for _ in range(3)],
            "H": [rng.standard_normal(shape) for _ in range(3)]}
    o = OracleSimulation(eps, [], initial_fields=init, float_type=numpy.float64, do_poynting=False)
    epsv = vec(eps)
    energies = []
    for t in range(50):
        h_prev = o.H.copy()
        o.step()
        energies.append(float(numpy.sum(epsv * o.E * o.E) + numpy.sum(o.H * h_prev)))
    energies = numpy.array(energies)
    assert numpy.abs(energies - energies[0]).max() < 1e-11 * abs(energies[0])
    # a plain sum of squares is not conserved: the invariant is the specific one
    assert abs(float(numpy.sum(o.H ** 2)) - float(numpy.sum(h_prev ** 2))) > 1e-6


# ---------------------------------------------------------------- Poynting
@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("pol_k", [0, 1])
@pytest.mark.parametrize("sign", [1, -1])
def test_poynting_plane_wave_flux(d, pol_k, sign):
    """Travelling 1-D plane wave along axis d: the sum of S_d over one period
    equals L dt E0^2 cos(k/2) cos(theta/2) / 2 (sign of the direction), every
    step; the transverse components of S vanish."""
    L = 16
    shape = [2, 2, 2]
    shape[d] = L
    shape = tuple(shape)
    a = [r for r in range(3) if r != d][pol_k]
    k = sign * 2 * math.pi * 2 / L
    kvec = [0.0, 0.0, 0.0]
    kvec[d] = k
    dt = 0.99 / math.sqrt(3)
    theta = 2 * math.asin(dt * abs(math.sin(k / 2)))
    lam = numpy.exp(-1j * theta)
    Eh = numpy.zeros(3, complex)
    Eh[a] = 1.0
    Kf = (numpy.exp(1j * numpy.array(kvec)) - 1)
    Hh = -dt * lam / (lam - 1) * numpy.cross(Kf, Eh)
    init = {"E": _mode_fields(shape, kvec, Eh, 1.0), "H": _mode_fields(shape, kvec, Hh, 1.0)}
    o = OracleSimulation([numpy.ones(shape)] * 3, [], initial_fields=init, float_type=numpy.float64, dt=dt)
    expect = sign * L * dt * math.cos(k / 2) * math.cos(theta / 2) / 2
    ncol = 4   # (x, y) columns transverse to d
    for t in range(10):
        o.step()
        S = _as3(o.S, shape)
        assert abs(S[d].sum() / ncol - expect) < 1e-12 * abs(expect), (t, S[d].sum() / ncol, expect)
        for r in range(3):
            if r != d:
                assert numpy.abs(S[r]).max() < 1e-14


# ---------------------------------------------------------------- CPML
def test_cpml_zero_sigma_is_periodic():
    rng = numpy.random.default_rng(14)
    shape = (10, 9, 8)
    eps = [rng.uniform(1, 3, shape) for _ in range(3)]
    init = {"E": [rng.standard_normal(shape) for _ in range(3)],
            "H": [rng.standard_normal(shape) for _ in range(3)]}
    # ma = 0 keeps alpha = cfs_alpha > 0 on every layer (with ma >= 1 alpha
    # vanishes on the outermost layer and sigma/(sigma + kappa*alpha) is 0/0)
    pmls = [{"axis": ax, "polarity": p, "thickness": 3, "ln_R_per_layer": 0.0, "cfs_alpha": 0.1, "ma": 0}
            for ax in "xyz" for p in "np"]
    a = OracleSimulation(eps, pmls, initial_fields=init)
    b = OracleSimulation(eps, [], initial_fields=init)
    _run(a, 20)
    _run(b, 20)
    for name in "EHS":
        assert numpy.array_equal(getattr(a, name), getattr(b, name)), name


def _pulse_run(pad, steps, interior=24):
    n = interior + 2 * pad
    shape = (n, n, n)
    eps = [numpy.ones(shape)] * 3
    pmls = [{"axis": ax, "polarity": p, "thickness": pad} for ax in "xyz" for p in "np"] if pad else []
    o = OracleSimulation(eps, pmls, float_type=numpy.float64, do_poynting=False)
    c = n // 2
    idx = 2 * o.N + (c * n + c) * n + c
    tt = numpy.arange(steps)
    wave = numpy.exp(-0.5 * ((tt - 30) / 8.0) ** 2) * numpy.sin(2 * math.pi / 12 * tt * o.dt)
    hist = []
    for t in range(steps):
        o.update_E()
        o.E[idx] += wave[t]
        o.update_H()
        E = _as3(o.E, shape)[:, pad:pad + interior, pad:pad + interior, pad:pad + interior]
        hist.append(E.copy())
    return numpy.array(hist)


@pytest.mark.slow
def test_cpml_reflection_falls_with_thickness():
    """Reflection = ||E_T - E_wide|| / ||E_wide|| in the interior, with a
    40-cell padded run as the reflection-free reference (its own reflections
    cannot return within the window)."""
    steps = 90
    wide = _pulse_run(40, steps)
    refl = {}
    for T in (2, 4, 8):
        E = _pulse_run(T, steps)
        refl[T] = numpy.linalg.norm(E - wide) / numpy.linalg.norm(wide)
    db = {T: 20 * math.log10(r) for T, r in refl.items()}
    assert db[2] > db[4] > db[8], db
    assert db[8] < -30, db


def test_coefficients_closed_form():
    dt = 0.99 / math.sqrt(3)
    for pol in "np":
        p = pml_defaults({"axis": "x", "polarity": pol})
        c = pml_coefficients(p, dt, numpy.float64)
        T, m, smax = 8, 3.5, 1.6 / 2 * 4.5
        for ir in range(T):
            for eh, (p0, p1, p2) in (("e", c[:3]), ("h", c[3:])):
                x = T - ir - (0.5 if (pol == "p") == (eh == "e") else 0.0)
                sig = (x / T) ** m * smax
                assert p0[ir] == pytest.approx(math.exp(-sig * dt), rel=1e-14)
                assert p1[ir] == pytest.approx(math.exp(-sig * dt) - 1, rel=1e-13)
                assert p2[ir] == 1.0
    # kappa and alpha terms
    p = pml_defaults({"axis": "y", "polarity": "n", "epsilon_eff": 4.0, "cfs_alpha": 0.2, "ma": 2, "thickness": 5})
    c = pml_coefficients(p, dt, numpy.float64)
    x = 5 - 0 - 0.5   # H on an 'n' face, outermost layer
    s = (x / 5) ** 3.5
    sig, kap, alp = s * 3.6, 1 + s * (2.0 - 1), (1 - x / 5) ** 2 * 0.2
    p0 = math.exp(-(sig / kap + alp) * dt)
    assert c[3][0] == pytest.approx(p0, rel=1e-14)
    assert c[4][0] == pytest.approx(sig / (sig + kap * alp) * (p0 - 1), rel=1e-13)
    assert c[5][0] == pytest.approx(1 / kap, rel=1e-15)


# ---------------------------------------------------------------- PEC, Bloch
def test_pec_keeps_e():
    rng = numpy.random.default_rng(15)
    shape = (8, 9, 10)
    eps = [rng.uniform(1, 3, shape) for _ in range(3)]
    mask = rng.random(shape) < 0.2
    for e in eps:
        e[mask] = numpy.inf
    init = {"E": [rng.standard_normal(shape) for _ in range(3)], "H": [rng.standard_normal(shape) for _ in range(3)]}
    o = OracleSimulation(eps, [], initial_fields=init)
    E0 = _as3(o.E, shape).copy()
    _run(o, 15)
    E1 = _as3(o.E, shape)
    assert numpy.array_equal(E1[:, mask], E0[:, mask])
    assert not numpy.array_equal(E1[:, ~mask], E0[:, ~mask])


def test_bloch_unit_phase_is_periodic():
    rng = numpy.random.default_rng(16)
    shape = (8, 7, 9)
    eps = [rng.uniform(1, 3, shape) for _ in range(3)]
    init = {k: [rng.standard_normal(shape) for _ in range(3)] for k in "EHFG"}
    pmls = [{"axis": "z", "polarity": p, "thickness": 2} for p in "np"]
    bl = [{"axis": "x", "real": 1.0, "imag": 0.0}, {"axis": "y", "real": 1.0, "imag": 0.0}]
    a = OracleSimulation(eps, [dict(p) for p in pmls], bloch_boundaries=bl, initial_fields=init)
    b = OracleSimulation(eps, [dict(p) for p in pmls], initial_fields={"E": init["E"], "H": init["H"]})
    _run(a, 15)
    _run(b, 15)
    for name in "EHS":
        assert numpy.array_equal(getattr(a, name), getattr(b, name)), name


def test_bloch_cell_equals_supercell():
    """An L-cell Bloch run with phase exp(i phi), phi = 2 pi m / n, equals the
    first L cells of an nL-cell periodic run started from fields that satisfy
    f(x + L) = f(x) exp(i phi) (real parts -> E/H, imaginary parts -> F/G)."""
    rng = numpy.random.default_rng(17)
    L, n, m = 6, 4, 1
    phi = 2 * math.pi * m / n
    shape = (L, 5, 10)
    big = (n * L, 5, 10)
    eps1 = rng.uniform(1, 3, shape)
    eps = [eps1] * 3
    epsb = [numpy.concatenate([eps1] * n, axis=0)] * 3
    pmls = [{"axis": "z", "polarity": p, "thickness": 3} for p in "np"]
    fc = {k: [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)] for k in "EH"}

    def extend(f):
        return numpy.concatenate([f * numpy.exp(1j * phi * j) for j in range(n)], axis=0)

    bl = [{"axis": "x", "real": math.cos(phi), "imag": math.sin(phi)}]
    a = OracleSimulation(eps, [dict(p) for p in pmls], bloch_boundaries=bl, float_type=numpy.float64,
                         initial_fields={"E": [f.real for f in fc["E"]], "H": [f.real for f in fc["H"]],
                                         "F": [f.imag for f in fc["E"]], "G": [f.imag for f in fc["H"]]})
    re = OracleSimulation(epsb, [dict(p) for p in pmls], float_type=numpy.float64,
                          initial_fields={k: [extend(f).real for f in fc[k]] for k in "EH"})
    im = OracleSimulation(epsb, [dict(p) for p in pmls], float_type=numpy.float64,
                          initial_fields={k: [extend(f).imag for f in fc[k]] for k in "EH"})
    _run(a, 25)
    _run(re, 25)
    _run(im, 25)
    for mine, sup, part in (("E", re, "E"), ("H", re, "H"), ("F", im, "E"), ("G", im, "H")):
        A = _as3(getattr(a, mine), shape)
        B = _as3(getattr(sup, part), big)[:, :L]
        assert numpy.abs(A - B).max() < 1e-12 * numpy.abs(B).max(), mine


# ---------------------------------------------------------------- S, cell by cell
def _cyc(d):
    return (d + 1) % 3, (d + 2) % 3


# physical offsets of the Yee components from their cell index (E_a on the
# primal edge along a, H_a on the dual edge along a)
_DE = numpy.eye(3) * 0.5
_DH = (numpy.ones((3, 3)) - numpy.eye(3)) * 0.5


def _s_closed_form(shape, kvec, Eh, Hh, lam, t, dt):
    """Poynting vector of the discrete plane wave E = Re(Eh lam^t e^{ik.idx}),
    H = Re(Hh lam^t e^{ik.idx}) after step t, sampled at the S_d points.

    S_d sits on the E_d point.  It is dt/8 times the sum of the four E x H
    products on the edges around that point: E_a H_b at the two points half a
    cell along a from it, E_b H_a (negated) at the two points half a cell
    along b, (d, a, b) cyclic; E is averaged over the two cells along d and H
    over the two half-steps around E's time.  For a plane wave these averages
    are closed forms:
      2 cos(k_d/2) Re(A e^{ik.r} lam^t),  Re(B e^{ik.r} lam^t (1 + 1/lam)),
    with A, B the amplitudes at the physical positions, and the pair of
    points +-1/2 along a turns e^{2ik.r} into 2 cos(k_a) e^{2ik.r}."""
    k = numpy.asarray(kvec, dtype=float)
    A = [Eh[a] * numpy.exp(-1j * (k @ _DE[a])) for a in range(3)]
    B = [Hh[a] * numpy.exp(-1j * (k @ _DH[a])) for a in range(3)]
    idx = numpy.indices(shape).astype(float)
    out = []
    for d in range(3):
        a, b = _cyc(d)
        r = idx + _DE[d][:, None, None, None]
        e2 = numpy.exp(2j * sum(k[q] * r[q] for q in range(3)))
        tm = lam ** t * (1 + 1 / lam)          # H^t + H^{t-1}, relative to lam^t
        steady = numpy.real((A[a] * numpy.conj(B[b]) - A[b] * numpy.conj(B[a])) * numpy.conj(tm) * lam ** t)
        osc = numpy.real((math.cos(k[a]) * A[a] * B[b] - math.cos(k[b]) * A[b] * B[a]) * e2 * lam ** t * tm)
        out.append(dt / 4 * math.cos(k[d] / 2) * (steady + osc))
    return numpy.array(out)


@pytest.mark.parametrize("m", [(1, -2, 3), (2, 1, -1), (0, 3, 2)])
def test_poynting_oblique_wave_cell_by_cell(m):
    """Every S component of an oblique travelling wave, at every cell and
    step, equals the closed form above (FP64, 1e-12 of max|S|).  With k along
    two or three axes the four products sit at distinct phases, so a
    neighbour taken on the wrong side (+e instead of -e), a dropped product,
    a wrong sign or a missing time average moves S by O(1)."""
    shape = (8, 10, 12)
    kvec = [2 * math.pi * m[r] / shape[r] for r in range(3)]
    dt = 0.99 / math.sqrt(3)
    Eh, Hh, lam = _eigenmode(shape, kvec, 1.0, dt, a=(0.4, -1.0, 0.7))
    init = {"E": _mode_fields(shape, kvec, Eh, 1.0), "H": _mode_fields(shape, kvec, Hh, 1.0)}
    o = OracleSimulation([numpy.ones(shape)] * 3, [], initial_fields=init, float_type=numpy.float64, dt=dt)
    for t in range(1, 9):
        o.step()
        S = _as3(o.S, shape)
        ref = _s_closed_form(shape, kvec, Eh, Hh, lam, t, dt)
        scale = numpy.abs(ref).max()
        assert scale > 1e-3
        for d in range(3):
            err = numpy.abs(S[d] - ref[d]).max()
            assert err < 1e-12 * scale, (t, d, err, scale)
        # the oscillating part is a large share of S, so the test sees offsets
        assert numpy.abs(ref - ref.mean(axis=(1, 2, 3), keepdims=True)).max() > 0.2 * scale


# ---------------------------------------------------------------- CPML, 1-D model
def _rg_coefficients(depth, T, sigma_max, kappa_max, alpha_max, m, ma, dt):
    """Roden & Gedney's CPML (Microw. Opt. Tech. Lett. 27, 2000) in units
    eps0 = mu0 = 1, with the graded profiles sampled at a node of the given
    depth: b = exp(-(sigma/kappa + alpha) dt),
    a = sigma (b - 1) / (sigma kappa + kappa^2 alpha)."""
    g = (depth / T) ** m
    sigma = sigma_max * g
    kappa = 1 + (kappa_max - 1) * g
    alpha = alpha_max * (1 - depth / T) ** ma
    b = math.exp(-(sigma / kappa + alpha) * dt)
    a = sigma * (b - 1) / (sigma * kappa + kappa * kappa * alpha)
    return a, b, kappa


def _cpml_1d(L, faces, eps_a, E0, H0, sgn, dt, steps):
    """Independent FP64 1-D Yee + CPML line along one axis: E at x = i,
    H at x = i + 1/2, periodic.  dE/dt = sgn/eps dH/dx, dH/dt = sgn dE/dx,
    with dx -> dx/kappa + psi inside a PML.  A face of thickness T is a layer
    T cells thick whose outer edge is the outermost node of the line on its
    side: E_0 (x = 0) on the low side, H_{L-1} (x = L - 1/2) on the high side,
    so a node's depth is T - x (low) or x - (L - 1/2 - T) (high), and it is
    inside when 0 < depth <= T."""
    xE = numpy.arange(L, dtype=float)
    xH = xE + 0.5

    def layer(xs):
        a = numpy.zeros(L)
        b = numpy.ones(L)
        kap = numpy.ones(L)
        inside = numpy.zeros(L, bool)
        for side, T, par in faces:
            depth = (T - xs) if side == "n" else (xs - (L - 0.5 - T))
            for i in range(L):
                if 0 < depth[i] <= T:
                    a[i], b[i], kap[i] = _rg_coefficients(depth[i], T, *par, dt)
                    inside[i] = True
        return a, b, kap, inside

    aE, bE, kE, inE = layer(xE)
    aH, bH, kH, inH = layer(xH)
    E, H = E0.copy(), H0.copy()
    psiE, psiH = numpy.zeros(L), numpy.zeros(L)
    for _ in range(steps):
        D = H - numpy.roll(H, 1)
        psiE = numpy.where(inE, bE * psiE + aE * D, 0.0)
        E = E + dt / eps_a * sgn * (D / kE + psiE)
        D = numpy.roll(E, -1) - E
        psiH = numpy.where(inH, bH * psiH + aH * D, 0.0)
        H = H + dt * sgn * (D / kH + psiH)
    return E, H


@pytest.mark.parametrize("d", [0, 1, 2])
def test_cpml_matches_independent_1d_model(d):
    """A 3-D run whose fields are uniform across axis d matches the
    independent 1-D CPML line above on both polarisations, to 1e-12: the
    E/H half-cell depths, the psi recursion and its coefficients (sigma,
    kappa, CFS alpha, both grading exponents), the derivative each psi
    takes and the sign it enters the curl with on every axis."""
    L, Tn, Tp = 36, 5, 7
    dt = 0.99 / math.sqrt(3)
    rng = numpy.random.default_rng(20 + d)
    shape = [3, 2, 2]
    shape[d] = L
    shape = tuple(shape)
    fn = {"axis": "xyz"[d], "polarity": "n", "thickness": Tn, "ln_R_per_layer": -1.9, "m": 3.0,
          "cfs_alpha": 0.08, "ma": 1, "epsilon_eff": 2.25}
    fp = {"axis": "xyz"[d], "polarity": "p", "thickness": Tp, "ln_R_per_layer": -1.3, "m": 2.5,
          "cfs_alpha": 0.03, "ma": 2, "epsilon_eff": 1.0, "mu_eff": 1.7}

    def par(f):
        kappa_max = math.sqrt(f.get("mu_eff", 1.0) * f["epsilon_eff"])
        return (-f["ln_R_per_layer"] / 2 * (f["m"] + 1), kappa_max, f["cfs_alpha"], f["m"], f["ma"])

    faces = [("n", Tn, par(fn)), ("p", Tp, par(fp))]
    line = [slice(None) if r == d else 0 for r in range(3)]

    def bcast(v):
        sh = [1, 1, 1]
        sh[d] = L
        return numpy.broadcast_to(v.reshape(sh), shape).copy()

    eps1 = [rng.uniform(1, 3, L) for _ in range(3)]
    e0 = [rng.standard_normal(L) for _ in range(3)]
    h0 = [rng.standard_normal(L) for _ in range(3)]
    o = OracleSimulation([bcast(e) for e in eps1], [dict(fn), dict(fp)], float_type=numpy.float64, dt=dt,
                         initial_fields={"E": [bcast(v) for v in e0], "H": [bcast(v) for v in h0]},
                         do_poynting=False)
    steps = 80
    _run(o, steps)
    E3, H3 = _as3(o.E, shape), _as3(o.H, shape)
    a, b = _cyc(d)
    # (E_a, H_b): sgn = eps_{d b a} = -1; (E_b, H_a): +1
    for ea, hb, sgn in ((a, b, -1.0), (b, a, 1.0)):
        E1, H1 = _cpml_1d(L, faces, eps1[ea], e0[ea], h0[hb], sgn, dt, steps)
        scale = max(numpy.abs(E1).max(), numpy.abs(H1).max())
        assert numpy.abs(E3[ea][tuple(line)] - E1).max() < 1e-12 * scale, ("E", ea)
        assert numpy.abs(H3[hb][tuple(line)] - H1).max() < 1e-12 * scale, ("H", hb)
        # the PML did act: the same line without it differs by O(1)
        E0_, _ = _cpml_1d(L, [], eps1[ea], e0[ea], h0[hb], sgn, dt, steps)
        assert numpy.abs(E0_ - E1).max() > 0.05 * scale
    # the component along d does not move
    assert numpy.array_equal(E3[d][tuple(line)], e0[d])
    assert numpy.array_equal(H3[d][tuple(line)], h0[d])


def test_kappa_max_minus_one_rounds_once():
    """simulation.py:285,291: kappa_max is a float64 scalar, so kappa_max - 1
    is formed in float64 and rounded to float32 once, where it meets the
    float32 profile.  For eps_eff = 6, rounding sqrt(6) to float32 first
    gives a kappa one ulp off on some layers; the coefficients must follow
    the single rounding."""
    f32 = numpy.float32
    p = pml_defaults({"axis": "x", "polarity": "n", "epsilon_eff": 6.0, "thickness": 8, "m": 3.5})
    c = pml_coefficients(p, 0.5, f32)
    km1 = f32(math.sqrt(6.0) - 1.0)
    km1_bad = f32(f32(math.sqrt(6.0)) - f32(1))
    assert km1 != km1_bad                      # the two orders differ for this value
    x = numpy.arange(8, 0, -1).astype(f32)     # E depths on an 'n' face
    s = (x / f32(8)) ** f32(3.5)
    good = f32(1) / (f32(1) + s * km1)
    bad = f32(1) / (f32(1) + s * km1_bad)
    assert not numpy.array_equal(good, bad)
    assert numpy.array_equal(c[2], good)


# ---------------------------------------------------------------- J source
@pytest.mark.parametrize("box", [(0, 8, 0, 9, 0, 10), (0, 3, 4, 9, 7, 10), (5, 8, 0, 1, 2, 3), (2, 2, 0, 9, 0, 10)])
def test_update_j_adds_c_jr_plus_s_ji_inside_the_box(box):
    """update_j.cl:23-32: E += c Jr + s Ji on cells with min <= coord < max on
    every axis, nothing elsewhere; boxes touching 0 and the far faces, a
    one-cell-thick box and an empty one."""
    rng = numpy.random.default_rng(30)
    shape = (8, 9, 10)
    o = OracleSimulation([numpy.ones(shape)] * 3, [], float_type=numpy.float64, do_fieldsrc=True,
                         initial_fields={"E": [rng.standard_normal(shape) for _ in range(3)]})
    o.Jr[:] = rng.standard_normal(o.Jr.size)
    o.Ji[:] = rng.standard_normal(o.Ji.size)
    before = _as3(o.E, shape).copy()
    c, s = 0.37, -1.21
    o.update_J(c, s, *box)
    after = _as3(o.E, shape)
    inside = numpy.zeros(shape, bool)
    inside[box[0]:box[1], box[2]:box[3], box[4]:box[5]] = True
    J = c * _as3(o.Jr, shape) + s * _as3(o.Ji, shape)
    assert numpy.array_equal(after[:, ~inside], before[:, ~inside])
    if inside.any():
        assert numpy.abs(after[:, inside] - (before[:, inside] + J[:, inside])).max() <= 1e-15 * 8
    assert inside.sum() == (box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4])


# ---------------------------------------------------------------- non-uniform dx
def test_nonuniform_energy_with_cell_volumes():
    """With per-index spacings, E_a lives on a primal edge (length dx_e along
    a) crossing a dual face (dx_h along the other two axes), H_a on a dual
    edge crossing a primal face.  Weighting the leapfrog energy with those
    volumes, sum eps E_a^2 V^E_a + H_a H'_a V^H_a is exactly conserved when
    the E step divides by dx_h and the H step by dx_e, each indexed by its
    own axis (update_e.cl:29-33, update_h.cl:29-33).  Alternating spacings
    make every pairing mistake visible."""
    rng = numpy.random.default_rng(31)
    shape = (8, 6, 10)
    alt = lambda n, p, q: numpy.where(numpy.arange(n) % 2 == 0, p, q)
    de = [alt(shape[0], 0.8, 1.3), alt(shape[1], 1.1, 0.7), rng.uniform(0.7, 1.3, shape[2])]
    dh = [alt(shape[0], 1.2, 0.9), alt(shape[1], 0.75, 1.25), rng.uniform(0.7, 1.3, shape[2])]
    eps = [rng.uniform(1, 4, shape) for _ in range(3)]
    init = {"E": [rng.standard_normal(shape) for _ in range(3)], "H": [rng.standard_normal(shape) for _ in range(3)]}
    o = OracleSimulation(eps, [], dxes=[de, dh], initial_fields=init, float_type=numpy.float64,
                         do_poynting=False)

    def vol(own, other, a):
        g = numpy.meshgrid(*[(own if r == a else other)[r] for r in range(3)], indexing="ij")
        return g[0] * g[1] * g[2]

    VE = [vol(de, dh, a) for a in range(3)]
    VH = [vol(dh, de, a) for a in range(3)]
    VE_bad = [vol(dh, de, a) for a in range(3)]
    VH_bad = [vol(de, dh, a) for a in range(3)]
    epsv = numpy.array(eps)

    def energy(E, H, Hp, ve, vh):
        return sum(float(numpy.sum(epsv[a] * E[a] ** 2 * ve[a]) + numpy.sum(H[a] * Hp[a] * vh[a])) for a in range(3))

    w, w_bad = [], []
    for _ in range(40):
        hp = _as3(o.H, shape).copy()
        o.step()
        E, H = _as3(o.E, shape), _as3(o.H, shape)
        w.append(energy(E, H, hp, VE, VH))
        w_bad.append(energy(E, H, hp, VE_bad, VH_bad))
    w, w_bad = numpy.array(w), numpy.array(w_bad)
    assert numpy.abs(w - w[0]).max() < 1e-11 * abs(w[0])
    assert numpy.abs(w_bad - w_bad[0]).max() > 1e-4 * abs(w_bad[0])
