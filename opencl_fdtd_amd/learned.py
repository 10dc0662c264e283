"""Thin absorbing boundaries: scenarios, reflection measurement, and the
learned 4-cell absorber.

A *scenario* is an interior region (eps ~ U[1, 4] in 8^3 blocks within 16
cells of the faces, 1 elsewhere) driven by a soft dipole with a seeded
position, polarisation and wavelength U[12, 40] cells.  The same scenario
can be embedded in domains padded by different absorbers: CPML of thickness
T on every face, or the learned 4-cell shell.  Reflection of a padded run
is measured against the same scenario padded by a 48-cell CPML, which
absorbs far below the levels compared here:

    R = 20 log10( ||E_pad - E_wide|| / ||E_wide|| )

with the norms over the interior cells and over every ``sample_every``-th
step of the run.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy
import torch

from .simulation import Simulation

WIDE_PAD = 48


@dataclasses.dataclass
class Scenario:
    interior: Tuple[int, int, int]
    eps: numpy.ndarray             # interior permittivity (sx, sy, sz), isotropic
    src: Tuple[int, int, int, int]  # (component, x, y, z) in interior coordinates
    wavelength: float
    steps: int
    dt: float = 0.99 / math.sqrt(3)
    width_wl: float = 1.5           # pulse envelope width in wavelengths (of steps)

    def wave(self) -> numpy.ndarray:
        """sin(2 pi/wl n dt) exp(-((n - n0)/w)^2 / 2), w = width_wl * wl, n0 = 4 w."""
        n = numpy.arange(self.steps, dtype=numpy.float64)
        width = self.width_wl * self.wavelength
        n0 = 4 * width
        return numpy.sin(2 * numpy.pi / self.wavelength * n * self.dt) * numpy.exp(-0.5 * ((n - n0) / width) ** 2)


def make_scenario(seed: int, interior=(256, 256, 128), steps: int = 600, near_face: int = 16,
                  src_margin: int = 8, width_wl: float = 1.5) -> Scenario:
    """eps ~ U[1, 4] per 8^3 block within ``near_face`` cells of a face (1
    elsewhere); dipole of random polarisation at a random position at least
    ``src_margin`` cells inside, wavelength U[12, 40]."""
    rng = numpy.random.default_rng(seed)
    ix, iy, iz = interior
    blocks = rng.uniform(1, 4, size=(-(-ix // 8), -(-iy // 8), -(-iz // 8)))
    eps = numpy.kron(blocks, numpy.ones((8, 8, 8)))[:ix, :iy, :iz]
    core = numpy.zeros(interior, bool)
    core[near_face:ix - near_face, near_face:iy - near_face, near_face:iz - near_face] = True
    eps = numpy.where(core, 1.0, eps)
    pos = tuple(int(rng.integers(src_margin, s - src_margin)) for s in interior)
    comp = int(rng.integers(0, 3))
    wl = float(rng.uniform(12, 40))
    return Scenario(interior, eps, (comp,) + pos, wl, steps, width_wl=width_wl)


def embed(sc: Scenario, pad: int):
    """(shape, eps list, source flat index, interior slices) of the scenario
    padded by ``pad`` cells of eps = 1 on every face."""
    ix, iy, iz = sc.interior
    shape = (ix + 2 * pad, iy + 2 * pad, iz + 2 * pad)
    eps = numpy.ones(shape)
    sl = (slice(pad, pad + ix), slice(pad, pad + iy), slice(pad, pad + iz))
    eps[sl] = sc.eps
    c, x, y, z = sc.src
    idx = int(numpy.ravel_multi_index((x + pad, y + pad, z + pad), shape)) + c * int(numpy.prod(shape))
    return shape, [eps, eps, eps], idx, sl


def cpml_faces(T: int) -> List[Dict]:
    return [{"axis": a, "polarity": p, "thickness": T} for a in "xyz" for p in "np"]


def run_interior(sc: Scenario, pad: int, pmls: List[Dict], sample_every: int = 4,
                 correction=None, float_type=numpy.float32) -> torch.Tensor:
    """Runs the padded scenario; returns the interior E every
    ``sample_every`` steps as a (samples, 3, ix, iy, iz) tensor.
    ``correction(sim, n)``, if given, is applied after each E update and
    source (the learned shell)."""
    shape, eps, idx, sl = embed(sc, pad)
    sim = Simulation(eps, pmls, dt=sc.dt, do_poynting=False, float_type=float_type)
    wave = torch.as_tensor(sc.wave(), dtype=sim._tdtype, device=sim.device)
    out = []
    if correction is None:
        for n0 in range(0, sc.steps - sample_every + 1, sample_every):
            sim.step_n(sample_every, idx, wave[n0:n0 + sample_every])
            out.append(sim.E.view(3, *shape)[(slice(None),) + sl].clone())
        return torch.stack(out)
    for n in range(sc.steps):
        sim.update_E([])
        sim.E[idx] += wave[n]
        correction(sim, n)
        sim.update_H([])
        if (n + 1) % sample_every == 0:
            out.append(sim.E.view(3, *shape)[(slice(None),) + sl].clone())
    return torch.stack(out)


def reflection_db(test: torch.Tensor, wide: torch.Tensor) -> float:
    num = torch.linalg.vector_norm((test - wide).double())
    den = torch.linalg.vector_norm(wide.double())
    return float(20 * torch.log10(num / den))


# ---------------------------------------------------------------------------
# The learned 4-cell shell
# ---------------------------------------------------------------------------

from . import _lib  # noqa: E402

N_HID, N_IN, N_OUT, N_TAP = 16, 6, 3, 27
N_DEPTH = 4                        # layer-2 heads per box: one per depth of the shell (T <= 4)
W1_SIZE = N_HID * N_IN * N_TAP
W2_SIZE = N_OUT * N_HID * N_TAP    # one head
N_WEIGHTS = 6 * (W1_SIZE + N_DEPTH * W2_SIZE)


# The CPML the shell runs under the correction: a 4-cell CPML tuned for the
# interior-field error against a 48-cell padded run (graded sigma of order 2,
# kappa_max = sqrt(6), CFS alpha 0.15), far better at 4 cells than the
# defaults of simulation.py:161-168.
BASE_PML = {"ln_R_per_layer": -2.2, "m": 2.0, "cfs_alpha": 0.15, "epsilon_eff": 6.0, "ma": 1}
# Thinner shells' bases, tuned the same way (mean reflection on the held-out
# 128 x 128 x 96 scenarios over a grid of ln R, m, CFS alpha and epsilon_eff;
# for 3 cells the runner-up, whose worst case is better).
THIN_BASE_PML = {
    2: {"ln_R_per_layer": -2.2, "m": 1.5, "cfs_alpha": 0.05, "epsilon_eff": 12.0, "ma": 1},
    3: {"ln_R_per_layer": -2.2, "m": 2.0, "cfs_alpha": 0.15, "epsilon_eff": 12.0, "ma": 1},
    4: BASE_PML,
}


class ShellNet:
    """Weights of the shell correction: an FP32 master copy (for training)
    and the BF16 copy the kernels read.  ``base_pml`` are the CPML parameters
    of the T-cell shell (defaults of simulation.py:161-168 otherwise).
    ``period``: the correction is applied after every step (1) or after
    every second or fourth step (2, 4: steps 0, period, 2 period, ...),
    which lets the solver run two steps per pass between corrections."""

    def __init__(self, thickness: int = 4, weights: Optional[torch.Tensor] = None, seed: int = 0,
                 w1_scale: float = 0.05, device="cuda", base_pml: Optional[Dict] = None, period: int = 2):
        if period not in (1, 2, 4):
            raise ValueError("correction period %d is not 1, 2 or 4" % period)
        self.T = thickness
        self.period = period
        self.base_pml = dict(BASE_PML if base_pml is None else base_pml)
        if weights is None:
            g = torch.Generator().manual_seed(seed)
            w = torch.zeros(N_WEIGHTS, dtype=torch.float32)
            w[:6 * W1_SIZE] = torch.randn(6 * W1_SIZE, generator=g) * w1_scale   # W2 = 0: starts as CPML-T
            weights = w
        self.master = weights.to(device=device, dtype=torch.float32).clone()
        self.sync()

    def sync(self) -> None:
        self.bf16 = self.master.to(torch.bfloat16).contiguous()

    def zero_(self) -> "ShellNet":
        self.master.zero_()
        self.sync()
        return self

    def faces(self) -> List[Dict]:
        return [dict(f, **self.base_pml) for f in cpml_faces(self.T)]

    def save(self, path: str) -> None:
        numpy.savez_compressed(path, thickness=self.T, period=self.period, weights=self.master.detach().cpu().numpy(),
                               base_pml_keys=numpy.array(list(self.base_pml.keys())),
                               base_pml_values=numpy.array(list(self.base_pml.values()), dtype=numpy.float64))

    @classmethod
    def load(cls, path: str, device="cuda") -> "ShellNet":
        with numpy.load(path, allow_pickle=False) as z:
            base = {str(k): float(v) for k, v in zip(z["base_pml_keys"], z["base_pml_values"])}
            if z["weights"].size != N_WEIGHTS:
                raise ValueError("%s holds %d weights, this net has %d" % (path, z["weights"].size, N_WEIGHTS))
            period = int(z["period"]) if "period" in z.files else 1
            return cls(int(z["thickness"]), weights=torch.as_tensor(z["weights"]), device=device, base_pml=base,
                       period=period)


class ShellRunner:
    """Applies a ShellNet to one Simulation (the thin domain with CPML-T).

    One step of the shell solver: E update and source (E_s); H update from
    E_s; on a corrected step (every step, or steps 0, period, 2 period, ...
    counted from the runner's creation) E = E_s + correction(E_s, H before
    the update).  ``step`` makes one such step call by call, ``step_n`` makes
    n natively (the correction is then added by the next step's E update;
    with period 2 or 4 each pair of steps runs as one two-step pass where the
    plan has it)."""

    def __init__(self, sim: Simulation, net: ShellNet):
        self.sim, self.net = sim, net
        self.k = 0                     # steps made (the correction's phase)
        self.lib = sim._lib
        nh = int(self.lib.fdtd_shell_hidden_elems(sim.plan, net.T))
        if nh <= 0:
            raise _lib.FdtdError("shell thickness %d does not fit the grid %s" % (net.T, sim.shape))
        self.hidden = torch.zeros(nh, dtype=torch.float32, device=sim.device)
        self.hbar = None
        self.partials = None

    def _st(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.sim.device).cuda_stream)

    def apply(self, E: torch.Tensor, H: torch.Tensor) -> None:
        """E += correction(E, H) on the shell cells (in place)."""
        _lib.check(self.lib.fdtd_shell_apply(self.sim.plan, self.net.T, ctypes.c_void_p(self.net.bf16.data_ptr()),
                                             ctypes.c_void_p(E.data_ptr()), ctypes.c_void_p(H.data_ptr()),
                                             ctypes.c_void_p(self.hidden.data_ptr()), self._st()),
                   "fdtd_shell_apply")

    def corrected(self, k: int) -> bool:
        """Whether step k (counted from the start) gets the correction."""
        return k % self.net.period == 0

    def step(self, src_index: Optional[int] = None, amp=None, keep: Optional[List[torch.Tensor]] = None) -> None:
        """One shell step call by call; ``keep`` receives a copy of E after
        the E update and source (the correction's input)."""
        sim = self.sim
        sim.update_E([])
        if src_index is not None:
            sim._E[src_index] += amp
        if keep is not None:
            keep.append(sim._E.clone())
        corr = self.corrected(self.k)
        self.k += 1
        if not corr:
            sim.update_H([])
            sim.flush()
            return
        Ec = sim._E.clone()
        self.apply(Ec, sim._H[sim._h])
        sim.update_H([])
        sim.flush()
        sim._E.copy_(Ec)

    def step_n(self, n: int, src_index: Optional[int] = None, src_wave: Optional[torch.Tensor] = None) -> None:
        """``n`` shell steps launched natively (fdtd_shell_step_n); source as
        Simulation.step_n.  n and the steps made so far must be multiples
        of the correction period."""
        P = self.net.period
        if n % P or self.k % P:
            raise ValueError("with correction period %d, step_n takes a multiple of %d steps from a multiple of %d"
                             % (P, P, P))
        sim = self.sim
        sim.flush()
        wave_ptr = None
        if src_wave is not None:
            wave = src_wave.to(device=sim.device, dtype=sim._tdtype).contiguous()
            if wave.numel() < n:
                raise ValueError("src_wave has %d values, need %d" % (wave.numel(), n))
            self._wave_keepalive = wave
            wave_ptr = ctypes.c_void_p(wave.data_ptr())
        f = sim._fields()
        moved = ctypes.c_int32(0)
        _lib.check(self.lib.fdtd_shell_step_n(sim.plan, ctypes.byref(f), self.net.T,
                                              ctypes.c_void_p(self.net.bf16.data_ptr()),
                                              ctypes.c_void_p(self.hidden.data_ptr()), int(n), self.net.period,
                                              int(src_index if src_index is not None else -1), wave_ptr,
                                              self._st(), ctypes.byref(moved)), "fdtd_shell_step_n")
        self.k += n
        if len(sim._H) == 2 and moved.value & 4:
            sim._h = 1 - sim._h
        if moved.value & 1:
            sim._swap_e()
        if moved.value & 2:
            sim._psi_e, sim._psi_e_out = sim._psi_e_out, sim._psi_e

    def backward(self, E, H, Ebar, Ebar_out, Hbar_out, wgrad) -> None:
        if self.hbar is None:
            self.hbar = torch.zeros_like(self.hidden)
            self.partials = torch.zeros(int(self.lib.fdtd_shell_partials_elems()), dtype=torch.float32,
                                        device=self.sim.device)
        p = lambda t: ctypes.c_void_p(t.data_ptr())
        _lib.check(self.lib.fdtd_shell_backward(self.sim.plan, self.net.T, p(self.net.bf16), p(E), p(H),
                                                p(self.hidden), p(self.hbar), p(Ebar), p(Ebar_out), p(Hbar_out),
                                                p(self.partials), p(wgrad), self._st()), "fdtd_shell_backward")


import ctypes  # noqa: E402


def run_learned(sc: Scenario, net: ShellNet, sample_every: int = 4, float_type=numpy.float32) -> torch.Tensor:
    """The scenario in the thin domain (CPML-T shell + the net's correction);
    interior E every ``sample_every`` steps."""
    pad = net.T
    shape, eps, idx, sl = embed(sc, pad)
    sim = Simulation(eps, net.faces(), dt=sc.dt, do_poynting=False, float_type=float_type)
    runner = ShellRunner(sim, net)
    wave = torch.as_tensor(sc.wave(), dtype=sim._tdtype, device=sim.device)
    out = []
    for n0 in range(0, sc.steps - sample_every + 1, sample_every):
        runner.step_n(sample_every, idx, wave[n0:n0 + sample_every])
        out.append(sim._E.view(3, *shape)[(slice(None),) + sl].clone())
    return torch.stack(out)


def stability_trace(net: ShellNet, steps: int = 12000, every: int = 400, interior=(48, 48, 48),
                    seed: int = 7) -> List[float]:
    """Interior energy every ``every`` steps of the thin domain (CPML-T shell
    and the net) after a pulse with no static part (a derivative of a
    Gaussian in a random component at the centre), so the energy of a stable
    absorber keeps falling towards round-off; a correction that feeds a
    growing mode shows as energy that rises again."""
    rng = numpy.random.default_rng(seed)
    sc = Scenario(tuple(interior), numpy.ones(interior), (int(rng.integers(0, 3)),) + tuple(s // 2 for s in interior),
                  16.0, steps)
    shape, eps, idx, sl = embed(sc, net.T)
    sim = Simulation(eps, net.faces(), dt=sc.dt, do_poynting=False)
    runner = ShellRunner(sim, net)
    n = numpy.arange(steps, dtype=numpy.float64)
    w, n0 = 6.0, 30.0
    wave = torch.as_tensor(-(n - n0) / w * numpy.exp(-0.5 * ((n - n0) / w) ** 2), dtype=sim._tdtype,
                           device=sim.device)
    out = []
    for k0 in range(0, steps - every + 1, every):
        runner.step_n(every, idx, wave[k0:k0 + every])
        out.append(float(torch.sum(sim._E.view(3, *shape)[(slice(None),) + sl].double() ** 2)))
    return out


def is_stable(trace: Sequence[float]) -> bool:
    """A stability trace whose last value is no higher than (5 % above) the
    lowest of its first half: over the second half nothing grows."""
    half = len(trace) // 2
    return trace[-1] <= 1.05 * min(trace[:half + 1])


class ShellTrainer:
    """End-to-end training of a ShellNet through the solver.

    Loss of one scenario: sum over steps of ||E_thin - E_wide||^2 on the
    interior, divided by sum of ||E_wide||^2 (10 log10 of it is the
    reflection in dB over every step).  Its gradient with respect to the
    weights is exact: the thin run is reversed step by step with the adjoint
    of the solver (fdtd_adjoint_step, split around the shell) and of the
    shell correction (fdtd_shell_backward), with segments of
    ``checkpoint_every`` steps recomputed from checkpoints."""

    def __init__(self, net: ShellNet, scenarios: Sequence[Scenario], checkpoint_every: int = 60,
                 lr: float = 1e-3, betas=(0.9, 0.999), loss_every: int = 1):
        self.net = net
        self.K = checkpoint_every
        self.loss_every = loss_every           # the loss compares E after every loss_every-th step
        self.lr, self.betas = lr, betas
        self.m = torch.zeros_like(net.master)
        self.v = torch.zeros_like(net.master)
        self.t = 0
        self.next_case = 0                     # scenarios are taken in turn across train() calls
        self.stab = None                       # the source-free run of the stability term (stability_setup)
        self.last_stab = None
        self.cases = []
        for sc in scenarios:
            wide = run_interior(sc, WIDE_PAD, cpml_faces(WIDE_PAD), sample_every=loss_every)
            self.cases.append((sc, wide, float(torch.sum(wide.double() ** 2))))

    def _thin(self, sc):
        shape, eps, idx, sl = embed(sc, self.net.T)
        sim = Simulation(eps, self.net.faces(), dt=sc.dt, do_poynting=False)
        return sim, idx, sl

    def loss_and_grad(self, case) -> Tuple[float, torch.Tensor]:
        sc, wide, Z = case
        sim, idx, sl = self._thin(sc)
        runner = ShellRunner(sim, self.net)
        shape = sim.shape
        wave = torch.as_tensor(sc.wave(), dtype=torch.float32, device=sim.device)
        Ev = lambda t: t.view(3, *shape)[(slice(None),) + sl]

        le = self.loss_every

        def per_step(n, E, Eb):
            # loss term of step n (on E after the step, every le-th step) and its adjoint into Eb
            if (n + 1) % le:
                return None if Eb is not None else torch.zeros((), dtype=torch.float64, device=sim.device)
            w = wide[(n + 1) // le - 1]
            if Eb is not None:
                Ev(Eb).add_((2.0 / Z) * (Ev(E) - w))
                return None
            return torch.sum((Ev(E) - w).double() ** 2)

        loss, wgrad = self._trajectory(sim, runner, sc.steps, idx, wave, per_step=per_step)
        return loss / Z, wgrad

    def _trajectory(self, sim, runner, n_steps, idx, wave, per_step=None, final_w: float = 0.0):
        """Runs ``n_steps`` shell steps from the simulation's state (source
        idx, wave if given), with checkpoints every K steps, and reverses
        them: the gradient of sum_n per_step(n, E after step n) + final_w
        (|E_n|^2 + |H_n|^2) with respect to the weights."""
        E = sim._E
        cps, loss = [], torch.zeros((), dtype=torch.float64, device=sim.device)
        runner.k = 0
        for n in range(n_steps):
            if n % self.K == 0:
                cps.append(sim.checkpoint())
            runner.step(idx, wave[n] if wave is not None else None)
            if per_step is not None:
                loss += per_step(n, E, None)
        N3 = 3 * sim.N
        Eb = torch.zeros(N3, dtype=torch.float32, device=sim.device)
        Hb = torch.zeros_like(Eb)
        if final_w:
            Hn = sim._H[sim._h]
            loss += final_w * (torch.sum(E.double() ** 2) + torch.sum(Hn.double() ** 2))
            Eb.add_(E, alpha=2.0 * final_w)
            Hb.add_(Hn, alpha=2.0 * final_w)
        Eb_mid = torch.zeros_like(Eb)
        npsi = max(sim.psi_elems, 1)
        pe = [torch.zeros(npsi, dtype=torch.float32, device=sim.device) for _ in range(2)]
        ph = [torch.zeros(npsi, dtype=torch.float32, device=sim.device) for _ in range(2)]
        wgrad = torch.zeros(N_WEIGHTS, dtype=torch.float32, device=sim.device)
        b = _lib.AdjointBufs()
        b.eps, b.coef = sim.eps.data_ptr(), sim.coef.data_ptr()
        b.eps_grad = None
        b.obj_axis = -1
        b.src_index = -1
        st = ctypes.c_void_p(torch.cuda.current_stream(sim.device).cuda_stream)
        cur = 0
        for s in reversed(range(len(cps))):
            n0, n1 = s * self.K, min(s * self.K + self.K, n_steps)
            sim.restore(cps[s])
            Es, Hs, Em = [E.clone()], [sim._H[sim._h].clone()], []
            runner.k = n0
            for n in range(n0, n1):
                runner.step(idx, wave[n] if wave is not None else None, keep=Em)
                Es.append(E.clone())
                Hs.append(sim._H[sim._h].clone())
            for n in reversed(range(n0, n1)):
                k = n - n0
                # step n: Em = Estep(E_k, H_k) + source; H_k+1 = Hstep(H_k, Em);
                # E_k+1 = Em + y(Em, H_k) on a corrected step, Em otherwise.
                # Eb: adjoint of E_k+1 (+ the loss), Hb: adjoint of H_k+1.
                if per_step is not None:
                    per_step(n, Es[k + 1], Eb)
                corr = runner.corrected(n)
                if corr:
                    Eb_mid.copy_(Eb)             # adjoint of E_k+1, the correction's output
                b.E_prev, b.E_next = Es[k].data_ptr(), Em[k].data_ptr()
                b.H_prev, b.H_next = Hs[k].data_ptr(), Hs[k + 1].data_ptr()
                b.Eb, b.Hb = Eb.data_ptr(), Hb.data_ptr()
                b.psib_e, b.psib_e_out = pe[cur].data_ptr(), pe[1 - cur].data_ptr()
                b.psib_h, b.psib_h_out = ph[cur].data_ptr(), ph[1 - cur].data_ptr()
                b.parts = 1                      # the H step: Hb -> adjoint of H_k, Eb += its Em part
                _lib.check(sim._lib.fdtd_adjoint_step(sim.plan, ctypes.byref(b), st), "adjoint H")
                # the correction: Eb += its Em part, Hb += its H_k part (the H step's
                # Jacobian in H is the identity, so adding after its adjoint is exact)
                if corr:
                    runner.backward(Em[k], Hs[k], Eb_mid, Eb, Hb, wgrad)
                b.parts = 2                      # the E step: Eb -> adjoint of E_k, Hb += its part
                _lib.check(sim._lib.fdtd_adjoint_step(sim.plan, ctypes.byref(b), st), "adjoint E")
                cur = 1 - cur
            del Es, Hs, Em
        return float(loss), wgrad

    def stability_setup(self, interior=(48, 48, 48), steps: int = 2000, harvest: int = 6000, seed: int = 7) -> None:
        """The runs the stability term uses, on the thin domain of an eps = 1
        interior: a pulse with no static part (a derivative of a Gaussian,
        over the first ~60 steps, in a seeded component at the centre) run
        for ``harvest`` steps, whose last state (rescaled to unit energy)
        starts a source-free run of ``steps`` steps.  After a stable
        absorber that state is the round-off floor, which stays flat; when
        the correction feeds a growing mode the harvest has let it dominate,
        and the short run shows its growth."""
        rng = numpy.random.default_rng(seed)
        sc = Scenario(tuple(interior), numpy.ones(interior), (int(rng.integers(0, 3)),) + tuple(s // 2 for s in interior),
                      16.0, harvest)
        shape, eps, idx, sl = embed(sc, self.net.T)
        sim = Simulation(eps, self.net.faces(), dt=sc.dt, do_poynting=False)
        n = numpy.arange(harvest, dtype=numpy.float64)
        w, n0 = 6.0, 30.0
        wave = torch.as_tensor(-(n - n0) / w * numpy.exp(-0.5 * ((n - n0) / w) ** 2), dtype=sim._tdtype,
                               device=sim.device)
        self.stab = {"sim": sim, "zero": sim.checkpoint(), "steps": steps, "harvest": harvest, "src": (idx, wave),
                     "start": None}
        self.stability_harvest()

    def stability_harvest(self) -> None:
        """Runs the pulse with the current weights and keeps its last state,
        rescaled to unit energy, as the stability term's start."""
        S = self.stab
        sim = S["sim"]
        sim.restore(S["zero"])
        runner = ShellRunner(sim, self.net)
        idx, wave = S["src"]
        P = self.net.period
        chunk = 400 - 400 % P
        for k0 in range(0, S["harvest"] - S["harvest"] % chunk, chunk):
            runner.step_n(chunk, idx, wave[k0:k0 + chunk])
        st = sim.checkpoint()
        e = float(torch.sum(st["E"].double() ** 2) + torch.sum(st["H"].double() ** 2))
        if not math.isfinite(e) or e <= 0.0:
            raise FloatingPointError("the stability run's fields are not finite (energy %r): the correction "
                                     "grows too fast for the harvest" % e)
        scale = 1.0 / math.sqrt(e)
        for v in st.values():
            v.mul_(scale)
        S["start"] = st

    def stability_loss_and_grad(self, window: int = 200) -> Tuple[float, torch.Tensor]:
        """log(A / B) for the source-free run of stability_setup, with A and
        B the sums of |E|^2 over its last ``window`` steps and over the
        ``window`` steps before its middle, and the gradient of that
        logarithm: the growth of the run's second half, whose gradient for
        a growing mode is its growth rate's, whatever the mode's start."""
        S = self.stab
        sim, steps = S["sim"], S["steps"]
        sim.restore(S["start"])
        runner = ShellRunner(sim, self.net)
        late0, mid0 = steps - window, steps // 2 - window
        acc = {"A": torch.zeros((), dtype=torch.float64, device=sim.device), "B": None}
        acc["B"] = torch.zeros_like(acc["A"])
        scale = {}

        def per_step(n, E, Eb):
            late, mid = n >= late0, mid0 <= n < mid0 + window
            if Eb is None:
                if late:
                    acc["A"] += torch.sum(E.double() ** 2)
                elif mid:
                    acc["B"] += torch.sum(E.double() ** 2)
                return torch.zeros((), dtype=torch.float64, device=sim.device)
            if not scale:
                scale["A"], scale["B"] = float(acc["A"]), float(acc["B"])
            if late:
                Eb.add_(E, alpha=2.0 / scale["A"])
            elif mid:
                Eb.add_(E, alpha=-2.0 / scale["B"])
            return None

        _, g = self._trajectory(sim, runner, steps, None, None, per_step=per_step)
        return math.log(scale["A"] / scale["B"]), g

    def adam(self, g: torch.Tensor) -> None:
        b1, b2 = self.betas
        self.t += 1
        self.m.mul_(b1).add_(g, alpha=1 - b1)
        self.v.mul_(b2).addcmul_(g, g, value=1 - b2)
        mh = self.m / (1 - b1 ** self.t)
        vh = self.v / (1 - b2 ** self.t)
        self.net.master.sub_(self.lr * mh / (vh.sqrt() + 1e-12))
        self.net.sync()

    def train(self, iterations: int, log: Optional[Callable[[int, float], None]] = None,
              decay_every: int = 0, decay: float = 0.5, batch: int = 1, stab_weight: float = 0.0,
              stab_every: int = 1, stab_hinge: float = math.log(1.02)) -> List[float]:
        """``iterations`` Adam steps, each on the mean loss and gradient of
        ``batch`` scenarios taken in turn, plus ``stab_weight`` times the
        stability term (stability_setup) when set; returns the mean
        losses."""
        hist = []
        lr0 = self.lr
        for it in range(iterations):
            if decay_every:
                self.lr = lr0 * decay ** (it // decay_every)
            Ls, gs = 0.0, None
            for _ in range(batch):
                L, g = self.loss_and_grad(self.cases[self.next_case % len(self.cases)])
                self.next_case += 1
                Ls += L / batch
                gs = g / batch if gs is None else gs.add_(g, alpha=1.0 / batch)
            if stab_weight and self.stab is not None and self.t % stab_every == 0:
                # hinge: only a second half that grows by more than exp(stab_hinge) pushes back
                Lst, gst = self.stability_loss_and_grad()
                self.last_stab = Lst
                if Lst > stab_hinge:
                    gs.add_(gst, alpha=stab_weight * stab_every)
            self.adam(gs)
            hist.append(Ls)
            if log:
                log(it, Ls)
        return hist


# ---------------------------------------------------------------------------
# Training and evaluation entry point
# ---------------------------------------------------------------------------

import json  # noqa: E402
import os  # noqa: E402
import time  # noqa: E402

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "shell4.npz")


def train_scenarios(n: int, seed0: int = 100, interior=(64, 64, 48), steps: int = 300) -> List[Scenario]:
    """Training trajectories: small interiors with eps ~ U[1, 4] near the
    faces and random dipoles (position anywhere 4 or more cells inside,
    polarisation, wavelength U[12, 40])."""
    return [make_scenario(seed0 + s, interior=interior, steps=steps, near_face=16, src_margin=4, width_wl=1.0)
            for s in range(n)]


def validation_scenarios(n: int = 4, seed0: int = 500, interior=(128, 128, 96), steps: int = 400):
    """Held-out scenarios for model selection, with their padded references."""
    scs = [make_scenario(seed0 + s, interior=interior, steps=steps) for s in range(n)]
    return [(sc, run_interior(sc, WIDE_PAD, cpml_faces(WIDE_PAD), sample_every=4)) for sc in scs]


def validation_db(net: ShellNet, val) -> float:
    """Mean reflection (dB vs the padded runs) of the net on ``val``."""
    return float(numpy.mean([reflection_db(run_learned(sc, net, sample_every=4), wide) for sc, wide in val]))


def train_main(iterations: int, lr: float, n_scen: int, out: str, log=print, batch: int = 4,
               val_every: int = 8, interior=(64, 64, 48), steps: int = 300, period: int = 2,
               stab_weight: float = 0.0, stab_every: int = 2, init: Optional[str] = None,
               harvest: int = 12000, check: int = 24000, stab_hinge: float = math.log(1.02),
               loss_every: int = 1, thickness: int = 4) -> ShellNet:
    """Train a shell net of ``thickness`` cells (correction every ``period``
    steps) from W2 = 0 (its base CPML, THIN_BASE_PML); every ``val_every`` Adam steps measure the
    held-out reflection and keep the best weights whose long source-free
    run does not grow (stability_trace), the base itself when nothing beats
    it.  ``stab_weight``: weight of the stability term (every ``stab_every``
    steps; its pulse run ``harvest`` steps long).  ``check``: length of the
    stability trace of model selection.  ``init``: start from these weights
    instead of W2 = 0."""
    net = ShellNet(thickness, seed=1, period=period, base_pml=THIN_BASE_PML[thickness])
    if init:
        net0 = ShellNet.load(init)
        if net0.period != period or net0.T != thickness:
            raise ValueError("%s has period %d and thickness %d, not %d and %d" % (
                init, net0.period, net0.T, period, thickness))
        net.master.copy_(net0.master)
        net.sync()
    tr = ShellTrainer(net, train_scenarios(n_scen, interior=interior, steps=steps), checkpoint_every=60, lr=lr,
                      loss_every=loss_every)
    if stab_weight:
        tr.stability_setup(harvest=harvest)
    val = validation_scenarios()
    best_db = validation_db(net, val)
    best = net.master.clone()
    if init and not is_stable(stability_trace(net, steps=check)):
        # an unstable start is no candidate: the record starts at its base (W2 = 0)
        keep = net.master.clone()
        best[6 * W1_SIZE:] = 0
        net.master.copy_(best)
        net.sync()
        best_db = validation_db(net, val)
        net.master.copy_(keep)
        net.sync()
    log("validation: start %.2f dB" % best_db)
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)

    def save_best():
        keep = net.master.clone()
        net.master.copy_(best)
        net.save(out)
        net.master.copy_(keep)
    t0 = time.time()
    decay_every = max(iterations // 3, 1)
    for it in range(0, iterations, val_every):
        n = min(val_every, iterations - it)
        lr_it = lr * 0.5 ** (it // decay_every)
        tr.lr = lr_it
        if stab_weight:
            tr.stability_harvest()
        tr.train(n, batch=batch, stab_weight=stab_weight, stab_every=stab_every, stab_hinge=stab_hinge,
                 log=lambda k, L: log("it %d loss %.2f dB stab %s lr %.1e (%.0f s)" % (
                     it + k, 10 * numpy.log10(L), "%.2e" % tr.last_stab if tr.last_stab is not None else "-",
                     tr.lr, time.time() - t0)))
        db = validation_db(net, val)
        trace = stability_trace(net, steps=check)
        stable = is_stable(trace)
        if db < best_db and stable:
            best_db, best = db, net.master.clone()
            save_best()                        # the best so far survives an interrupted run
        log("validation after %d: %.2f dB, long run %s (%.1e -> %.1e) (best %.2f)" % (
            it + n, db, "stable" if stable else "GROWS", trace[len(trace) // 2], trace[-1], best_db))
    net.master.copy_(best)
    net.sync()
    net.save(out)
    return net


def evaluate_main(net: ShellNet, seeds: Sequence[int], interior=(256, 256, 128), steps: int = 600,
                  log=print) -> List[Dict]:
    """Reflection (dB, vs the 48-cell padded run) of CPML-T with the shell's
    base parameters (the net zeroed: ``base``), the learned T-cell shell
    (``learned``), CPML-T with the defaults, CPML-8 and CPML-16 on unseen
    scenarios of the given interior."""
    rows = []
    base = ShellNet(net.T, weights=torch.zeros_like(net.master), base_pml=net.base_pml)
    for sd in seeds:
        sc = make_scenario(sd, interior=interior, steps=steps)
        wide = run_interior(sc, WIDE_PAD, cpml_faces(WIDE_PAD), sample_every=4)
        row = {"seed": sd, "wavelength": round(sc.wavelength, 1), "src": list(sc.src)}
        row["base"] = reflection_db(run_learned(sc, base, sample_every=4), wide)
        row["learned"] = reflection_db(run_learned(sc, net, sample_every=4), wide)
        for T in sorted({net.T, 8, 16}):
            row["cpml%d" % T] = reflection_db(run_interior(sc, T, cpml_faces(T), sample_every=4), wide)
        rows.append(row)
        log(json.dumps(row))
    return rows


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="train / evaluate a learned thin absorbing shell")
    ap.add_argument("--train", type=int, default=0, help="training iterations")
    ap.add_argument("--lr", type=float, default=5e-6)
    ap.add_argument("--scenarios", type=int, default=48)
    ap.add_argument("--batch", type=int, default=4, help="scenarios per Adam step")
    ap.add_argument("--interior", default="64,64,48", help="training interior (cells)")
    ap.add_argument("--steps", type=int, default=300, help="steps per training scenario")
    ap.add_argument("--period", type=int, default=2, help="correction every 1, 2 or 4 steps")
    ap.add_argument("--thickness", type=int, default=4, choices=sorted(THIN_BASE_PML), help="shell cells")
    ap.add_argument("--stab-weight", type=float, default=0.0, help="weight of the stability term")
    ap.add_argument("--init", default=None, help="start from these weights (same period)")
    ap.add_argument("--stab-harvest", type=int, default=12000, help="pulse run before the stability term's run")
    ap.add_argument("--stab-check", type=int, default=24000, help="stability trace of model selection (steps)")
    ap.add_argument("--loss-every", type=int, default=1, help="compare E after every n-th step (memory)")
    ap.add_argument("--stab-hinge", type=float, default=0.02,
                    help="growth of the stability run's second half (relative) above which the term acts")
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--evaluate", type=int, default=0, help="number of unseen 256x256x128 scenarios")
    a = ap.parse_args(argv)
    interior = tuple(int(v) for v in a.interior.split(","))
    net = (train_main(a.train, a.lr, a.scenarios, a.weights, batch=a.batch, interior=interior, steps=a.steps,
                      period=a.period, stab_weight=a.stab_weight, init=a.init, harvest=a.stab_harvest,
                      check=a.stab_check, stab_hinge=math.log1p(a.stab_hinge), loss_every=a.loss_every,
                      thickness=a.thickness)
           if a.train else ShellNet.load(a.weights))
    if a.evaluate:
        rows = evaluate_main(net, range(1000, 1000 + a.evaluate))
        gain = [r["base"] - r["learned"] for r in rows]
        print(json.dumps({"thickness": net.T, "scenarios": len(rows), "mean_gain_db": float(numpy.mean(gain)),
                          "improved": int(sum(g > 0 for g in gain)),
                          "worst_base_db": max(r["base"] for r in rows),
                          "worst_learned_db": max(r["learned"] for r in rows)}), flush=True)


if __name__ == "__main__":
    main()
