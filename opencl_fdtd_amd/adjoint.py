"""Gradients of a time-integrated Poynting flux with respect to permittivity.

Objective: J(eps) = sum over the steps n = 1..N of the sum, over the plane
{coordinate of ``axis`` == ``index``}, of S_axis after step n, accumulated in
FP64 on the GPU.  The fields are driven by a soft point source, as in the
reference caller loop (``fdtd.py:162-175``).

dJ/d eps comes from the hand-written discrete adjoint of the step
(``fdtd_adjoint_step`` in ``libfdtd_hip.so``).  It reverses one step at a
time and needs the forward E and H before and after that step.  Those are
regenerated segment by segment: the forward pass keeps a checkpoint (E, H,
psi) every ``checkpoint_every`` steps, and the backward pass recomputes each
segment from its checkpoint, storing the segment's fields, then walks it in
reverse.

``PoyntingFluxObjective.apply(eps_design, problem)`` wraps this as a
``torch.autograd.Function`` over an isotropic design region: one eps value
per cell, used for all three components.
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Dict, List, Optional, Sequence, Tuple

import numpy
import torch

from . import _lib
from .simulation import Simulation


@dataclasses.dataclass
class FluxProblem:
    shape: Tuple[int, int, int]
    pmls: List[Dict]
    steps: int
    src_index: int                 # flat index into E (component offset included)
    src_wave: numpy.ndarray        # per-step source values (float64)
    axis: int = 2                  # the flux plane's normal
    index: int = 0                 # the flux plane's coordinate along axis
    dt: Optional[float] = None
    checkpoint_every: int = 100
    float_type: type = numpy.float32
    base_eps: Optional[numpy.ndarray] = None   # (3, sx, sy, sz), default 1
    design: Optional[Tuple[slice, slice, slice]] = None   # the design region of the isotropic variable


class FluxAdjoint:
    """Forward objective and its eps-gradient for one ``FluxProblem``."""

    def __init__(self, problem: FluxProblem, eps: torch.Tensor):
        """eps: the full (3, sx, sy, sz) permittivity on the GPU."""
        self.p = problem
        self.sim = Simulation([eps[0], eps[1], eps[2]], [dict(x) for x in problem.pmls], dt=problem.dt,
                              float_type=problem.float_type, do_poynting=True)
        self.lib = self.sim._lib
        self.dev = self.sim.device
        tdt = self.sim._tdtype
        self.wave = torch.as_tensor(numpy.asarray(problem.src_wave[:problem.steps]), dtype=tdt, device=self.dev)
        # the values actually added: the source converted to the field type
        self.wave_host = self.wave.double().cpu().numpy()

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def forward(self, keep_checkpoints: bool = True) -> Tuple[float, List[Dict[str, torch.Tensor]]]:
        """Runs the N steps; returns J and the checkpoints (state before
        steps 0, K, 2K, ...).  Each segment is one call that steps on the
        device and adds the plane's flux after every step."""
        N = self.p.steps
        K = self.p.checkpoint_every if keep_checkpoints else max(N, 1)
        acc = torch.zeros(1, dtype=torch.float64, device=self.dev)
        cps = []
        for n0 in range(0, N, K):
            n1 = min(n0 + K, N)
            if keep_checkpoints:
                cps.append(self.sim.checkpoint())
            self.sim.step_n_observed(n1 - n0, self.p.src_index, self.wave[n0:n1],
                                     flux=(self.p.axis, self.p.axis, self.p.index, acc))
        return float(acc.item()), cps

    def gradient(self) -> Tuple[float, torch.Tensor]:
        """J and dJ/d eps (3, sx, sy, sz) in FP64: inside the problem's
        design region when it has one (zero outside, where the adjoint step
        then reads neither the forward E nor the gradient), else
        everywhere."""
        J, cps = self.forward()
        sim = self.sim
        N3 = 3 * sim.N
        tdt = sim._tdtype
        Eb = torch.zeros(N3, dtype=tdt, device=self.dev)
        Hb = torch.zeros(N3, dtype=tdt, device=self.dev)
        n_psi = max(sim.psi_elems, 1)
        pe = [torch.zeros(n_psi, dtype=tdt, device=self.dev) for _ in range(2)]
        ph = [torch.zeros(n_psi, dtype=tdt, device=self.dev) for _ in range(2)]
        grad = torch.zeros(N3, dtype=torch.float64, device=self.dev)
        K = self.p.checkpoint_every
        b = _lib.AdjointBufs()
        b.eps = sim.eps.data_ptr()
        b.Eb = Eb.data_ptr()
        b.Hb = Hb.data_ptr()
        b.coef = sim.coef.data_ptr()
        b.eps_grad = grad.data_ptr()
        b.obj_axis = self.p.axis
        b.obj_index = self.p.index
        if self.p.design is not None:
            for k, sl in enumerate(self.p.design):
                lo, hi, _ = sl.indices(sim.shape[k])
                b.grad_lo[k], b.grad_hi[k] = lo, hi
        cur = 0
        for s in reversed(range(len(cps))):
            n0 = s * K
            n1 = min(n0 + K, self.p.steps)
            sim.restore(cps[s])
            # the segment's fields: before its first step, then after each
            Es = [sim._E.clone()] + [torch.empty_like(sim._E) for _ in range(n1 - n0)]
            Hs = [sim.H.clone()] + [torch.empty_like(sim._E) for _ in range(n1 - n0)]
            sim.step_n_observed(n1 - n0, self.p.src_index, self.wave[n0:n1], keep_E=Es[1:], keep_H=Hs[1:])
            for n in reversed(range(n0, n1)):
                k = n - n0
                b.E_prev, b.E_next = Es[k].data_ptr(), Es[k + 1].data_ptr()
                b.H_prev, b.H_next = Hs[k].data_ptr(), Hs[k + 1].data_ptr()
                b.psib_e, b.psib_e_out = pe[cur].data_ptr(), pe[1 - cur].data_ptr()
                b.psib_h, b.psib_h_out = ph[cur].data_ptr(), ph[1 - cur].data_ptr()
                b.src_index = self.p.src_index
                b.src_amp = float(self.wave_host[n])
                _lib.check(self.lib.fdtd_adjoint_step(sim.plan, ctypes.byref(b), self._stream()),
                           "fdtd_adjoint_step")
                cur = 1 - cur
            del Es, Hs
        return J, grad.view(3, *sim.shape)


def full_eps(problem: FluxProblem, eps_design: torch.Tensor) -> torch.Tensor:
    """The (3, sx, sy, sz) permittivity with the design region set to the
    (isotropic) design values."""
    base = problem.base_eps if problem.base_eps is not None else numpy.ones((3,) + tuple(problem.shape))
    eps = torch.as_tensor(base, dtype=torch.float64, device=eps_design.device).clone()
    for a in range(3):
        eps[(a,) + tuple(problem.design)] = eps_design.to(torch.float64)
    return eps


class PoyntingFluxObjective(torch.autograd.Function):
    """J(eps_design) with the adjoint gradient.  eps_design: a tensor on the
    GPU with the shape of ``problem.design``."""

    @staticmethod
    def forward(ctx, eps_design: torch.Tensor, problem: FluxProblem) -> torch.Tensor:
        eps = full_eps(problem, eps_design.detach())
        ctx.problem = problem
        ctx.save_for_backward(eps_design.detach())
        J, _ = FluxAdjoint(problem, eps).forward(keep_checkpoints=False)
        return torch.tensor(J, dtype=torch.float64, device=eps_design.device)

    @staticmethod
    def backward(ctx, gJ: torch.Tensor):
        (eps_design,) = ctx.saved_tensors
        problem = ctx.problem
        eps = full_eps(problem, eps_design)
        _, grad = FluxAdjoint(problem, eps).gradient()
        g = grad[(slice(None),) + tuple(problem.design)].sum(dim=0)
        return (gJ * g).to(eps_design.dtype), None


def flux(problem: FluxProblem, eps: torch.Tensor) -> float:
    """J for a full (3, sx, sy, sz) permittivity (forward only)."""
    J, _ = FluxAdjoint(problem, eps).forward(keep_checkpoints=False)
    return J
