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
