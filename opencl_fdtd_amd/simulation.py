"""The ``Simulation`` class: the reference host API over MI355X HIP kernels.

Mirrors ``opencl_fdtd/simulation.py:25-357``: the same constructor arguments,
field attributes (``E``, ``H``, ``S``, ``F``, ``G``, ``eps``, ``dt``,
``shape``, ``arg_type``) and ``update_*`` callables that take a list of events
to wait for and return an event.  Buffers are PyTorch-ROCm device tensors in
the reference layout (3N values, component-major, C order inside a
component), so ``sim.E[ind] += value`` works as in ``fdtd.py:168-169``.  The
updates run as hand-written gfx950 kernels through the C-ABI of
``libfdtd_hip.so``; nothing here computes field values on the host.

Differences from the reference, all documented behaviour:

* Poynting ``S`` is computed inside ``update_H`` (``poynting_mode='fused'``,
  the default); ``update_S`` is then an event-only no-op kept for call-shape
  parity.  In this mode ``H`` alternates between two device buffers, so read
  ``sim.H`` after each update rather than holding on to an old tensor.
  ``poynting_mode='two_pass'`` keeps the reference dataflow (6N ``oS`` buffer
  and a separate S kernel).
* With Bloch boundaries, ``update_E`` advances E and F together and
  ``update_H`` advances H and G together (the F/G steps read only H/G and
  E/F, which the paired step leaves untouched); ``update_F``/``update_G`` are
  event-only.  ``bloch_mode='phase'`` (default) applies the phase factor at
  the seam; ``'decoupled'`` reproduces the reference as shipped.
* ``step_n`` (n steps without host round trips) runs each H half-step and
  the next E half-step as one pass over the grid when the plan allows it
  (fused Poynting, or Poynting off without Bloch boundaries; with Bloch
  boundaries FP32, uniform dx and rows of at most 512 cells, and then the
  pass advances E/H and F/G together; ``fuse_steps=False`` turns it off and
  saves the second E buffer it needs, and with Poynting off the second H),
  and two steps per pass where that is faster.  E may end in either of its
  two buffers: read ``sim.E`` after each call.  Results are bit-identical.
  ``update_HE`` exposes that pass once; it moves E to the other buffer.
* The per-call loop gets the same pass (``defer_h=True``, the default, where
  the pass is available): ``update_H`` records the H half-step and returns
  at once, and the next ``update_E`` runs it together with the E half-step
  in one pass.  Reading ``sim.E``, ``sim.H``, ``sim.S`` (or F, G, oS), and
  every other call, runs a recorded H half-step first, so the fields read
  are always those of the reference loop; ``flush()`` runs it explicitly.
  An event returned by a deferred ``update_H`` completes when the work
  enqueued before it has; the deferred work itself runs with the next call.
* An unstable ``dt`` warns and is then used (the reference warns and leaves
  ``dt`` unset); malformed initial fields raise (the reference builds the
  exception without raising it).
"""
from __future__ import annotations

import ctypes
import warnings
from typing import Dict, List, Optional, Sequence

import numpy
import torch

from . import _lib
from .pml import face_coefficients, with_defaults

__all__ = ["Simulation", "Event", "DeviceField", "type_to_C"]


def type_to_C(float_type) -> str:
    """numpy float type -> C type name, simulation.py:360-376."""
    if float_type == numpy.float16:
        arg_type = "half"
    elif float_type == numpy.float32:
        arg_type = "float"
    elif float_type == numpy.float64:
        arg_type = "double"
    else:
        raise Exception("Unsupported type")
    return arg_type


class DeviceField(torch.Tensor):
    """A device tensor with the pyopencl-style ``get()`` used by callers
    (``fdtd.py:186-190``)."""

    def get(self) -> numpy.ndarray:
        with torch._C.DisableTorchFunctionSubclass():
            return self.detach().cpu().numpy()


class Event:
    """Completion marker of an enqueued update (pyopencl.Event stand-in).

    ``deferred``: the event of an ``update_H`` whose half-step is recorded
    for the next call's fused pass.  Its work is not enqueued yet, and every
    way to observe H or S runs it first on the simulation's stream, so
    ``wait()`` has nothing to wait for and returns at once."""

    def __init__(self, stream: torch.cuda.Stream, deferred: bool = False):
        self._event = torch.cuda.Event()
        self._event.record(stream)
        self._deferred = deferred

    def wait(self) -> None:
        if not self._deferred:
            self._event.synchronize()

    @property
    def torch_event(self) -> torch.cuda.Event:
        return self._event


def _vec(fields, dtype, device) -> torch.Tensor:
    parts = []
    for f in fields:
        t = f if isinstance(f, torch.Tensor) else torch.from_numpy(numpy.ascontiguousarray(f))
        parts.append(t.reshape(-1).to(dtype=dtype))
    return torch.cat(parts).to(device).contiguous()


class Simulation:
    """FDTD fields and update operations on one MI355X GPU.

    After constructing this object, call ``update_E``, ``update_H`` (and
    ``update_S``) each time step, as in the reference (simulation.py:28-51):

        for t in range(max_t):
            sim.update_E([]).wait()
            sim.E[ind] += numpy.sin(omega * t * sim.dt)
            event = sim.update_H([])
            if sim.update_S:
                event = sim.update_S([event])
            event.wait()
    """

    def __init__(self,
                 epsilon: List,
                 pmls: List[Dict],
                 bloch_boundaries: Sequence[Dict] = (),
                 dxes=None,
                 dt: Optional[float] = None,
                 initial_fields: Optional[Dict[str, List]] = None,
                 context=None,
                 queue: Optional[torch.cuda.Stream] = None,
                 float_type=numpy.float32,
                 do_poynting: bool = True,
                 do_fieldsrc: bool = False,
                 poynting_mode: str = "fused",
                 bloch_mode: str = "phase",
                 fuse_steps: bool = True,
                 defer_h: bool = True):
        if initial_fields is None:
            initial_fields = {}
        self._lib = _lib.load()
        if not torch.cuda.is_available():
            raise _lib.FdtdError("Simulation needs a ROCm GPU (torch.cuda.is_available() is False)")
        if context is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        else:
            self.device = torch.device(context) if not isinstance(context, int) else torch.device("cuda", context)
        if self.device.type != "cuda":
            raise _lib.FdtdError("context must name a GPU device, got {}".format(self.device))
        self.queue = queue
        self.context = self.device

        if len(epsilon) != 3:
            raise Exception("Epsilon must be a list with length of 3")
        shapes = [tuple(e.shape) for e in epsilon]
        if not all(s == shapes[0] for s in shapes[1:]):
            raise Exception("All epsilon grids must have the same shape. Shapes are {}".format(shapes))
        if len(shapes[0]) != 3:
            raise Exception("Epsilon grids must be 3-dimensional, got shape {}".format(shapes[0]))
        self.shape = shapes[0]
        self.arg_type = float_type
        ctype = type_to_C(float_type)
        if ctype == "half":
            raise Exception("float16 is not supported (the reference never enables cl_khr_fp16)")
        self._tdtype = torch.float32 if ctype == "float" else torch.float64
        self._ndtype = numpy.float32 if ctype == "float" else numpy.float64
        self.N = int(numpy.prod(self.shape))
        self.sources = {}

        # --- epsilon, dxes, dt (simulation.py:127-153, 340-347) -------------------
        self.eps = _vec(epsilon, self._tdtype, self.device)
        if dxes is None:
            dxes = 1.0
        if isinstance(dxes, (float, int)):
            uniform_dx = float(dxes)
            min_dx = dxes
            self.inv_dxes = None
        else:
            uniform_dx = None
            lists = list(dxes[0]) + list(dxes[1])
            for r, d in enumerate(lists):
                if len(d) != self.shape[r % 3]:
                    raise Exception("dxes[{}][{}] has length {}, expected {}".format(
                        r // 3, r % 3, len(d), self.shape[r % 3]))
            self.inv_dxes = [torch.from_numpy((1 / numpy.asarray(d, dtype=numpy.float64)).astype(self._ndtype))
                             .to(self.device) for d in lists]
            min_dx = min(min(d) for d in lists)
        max_dt = min_dx * .99 / numpy.sqrt(3)
        if dt is None:
            self.dt = max_dt
        elif dt <= 0:
            raise Exception("Invalid dt: {}".format(dt))
        else:
            if dt > max_dt:
                warnings.warn("Warning: unstable dt: {}".format(dt))
            self.dt = dt

        # --- fields (simulation.py:155-159, 349-357) ---------------------------
        self.bloch_boundaries = [dict(b) for b in bloch_boundaries]
        for b in self.bloch_boundaries:
            if b.get("axis") not in ("x", "y", "z"):
                raise Exception("Bloch axis must be one of 'x', 'y', 'z'")
        if poynting_mode not in ("fused", "two_pass"):
            raise Exception("poynting_mode must be 'fused' or 'two_pass'")
        if bloch_mode not in ("phase", "decoupled"):
            raise Exception("bloch_mode must be 'phase' or 'decoupled'")
        self.do_poynting = bool(do_poynting)
        self.poynting_mode = poynting_mode if do_poynting else "off"
        self.bloch_mode = bloch_mode
        # H (and psi_H) in two buffers: fused S reads the old H, and the fused
        # passes (also with Poynting off, without Bloch boundaries) read H
        # from one buffer while writing the other
        pingpong = self.poynting_mode == "fused" or (self.poynting_mode == "off" and bool(fuse_steps)
                                                      and not self.bloch_boundaries)
        self._E = self._create_field(initial_fields.get("E"))
        self._H = [self._create_field(initial_fields.get("H"))]
        if pingpong:
            self._H.append(torch.empty_like(self._H[0]))
        self._h = 0
        self._F = None
        self._G = None        # list: G, and with fused Poynting its ping-pong partner (indexed like H)
        if self.bloch_boundaries:
            self._F = self._create_field(initial_fields.get("F"))
            self._G = [self._create_field(initial_fields.get("G"))]
            if pingpong:
                self._G.append(torch.empty_like(self._G[0]))
        self._S = self._oS = None
        if self.do_poynting:
            self._S = torch.zeros_like(self._E)
            if self.poynting_mode == "two_pass":
                self._oS = torch.zeros(6 * self.N, dtype=self._tdtype, device=self.device)
        self.Jr = self.Ji = None
        if do_fieldsrc:
            self.Jr = torch.zeros_like(self._E)
            self.Ji = torch.zeros_like(self._E)

        # --- PML faces and coefficients (simulation.py:161-168, 274-319) -------
        self.pmls = [with_defaults(p) for p in pmls]
        desc = _lib.Desc()
        desc.shape[:] = list(self.shape)
        desc.dtype = _lib.F32 if ctype == "float" else _lib.F64
        if len(self.pmls) > _lib.MAX_FACES:
            raise Exception("at most {} PML faces are supported".format(_lib.MAX_FACES))
        desc.n_face = len(self.pmls)
        coefs = []
        for k, p in enumerate(self.pmls):
            desc.face[k].axis = "xyz".find(p["axis"])
            desc.face[k].polarity = 0 if p["polarity"] == "n" else 1
            desc.face[k].thickness = int(p["thickness"])
            coefs += face_coefficients(p, self.dt, self._ndtype)
        desc.dt = float(self.dt)
        desc.inv_dx = 1.0 / uniform_dx if uniform_dx is not None else 1.0
        desc.nonuniform = 0 if uniform_dx is not None else 1
        desc.poynting = {"off": _lib.POYNTING_OFF, "fused": _lib.POYNTING_FUSED,
                         "two_pass": _lib.POYNTING_TWO_PASS}[self.poynting_mode]
        mask = 0
        re, im = [1.0, 1.0, 1.0], [0.0, 0.0, 0.0]
        for b in self.bloch_boundaries:
            r = "xyz".find(b["axis"])
            mask |= 1 << r
            re[r], im[r] = float(b["real"]), float(b["imag"])
        desc.bloch_mask = mask
        desc.bloch_mode = _lib.BLOCH_PHASE if bloch_mode == "phase" else _lib.BLOCH_DECOUPLED
        desc.bloch_re[:] = re
        desc.bloch_im[:] = im
        self._desc = desc
        plan = ctypes.c_void_p()
        _lib.check(self._lib.fdtd_plan_create(ctypes.byref(desc), ctypes.byref(plan)), "fdtd_plan_create")
        self._plan = plan
        # second E buffer: step_n then runs each H half-step and the next E
        # half-step as one pass over the grid (fdtd_update_he)
        self._flux_partials = None      # step_n_observed's plane-sum partials, allocated on first use
        self._E_out = None
        if fuse_steps and self._lib.fdtd_update_he_supported(plan):
            self._E_out = torch.empty_like(self._E)
        # deferred H half-step of the per-call loop (run by the next update_E
        # as one pass, or by flush())
        self._defer = bool(defer_h) and self._E_out is not None
        self._pending_h = False
        n_psi = int(self._lib.fdtd_psi_elems(plan))
        self._coef = _vec([numpy.concatenate(coefs)] if coefs else [numpy.zeros(1)], self._tdtype, self.device)
        zeros = lambda: torch.zeros(max(n_psi, 1), dtype=self._tdtype, device=self.device)
        self._psi_e = zeros()
        # second psi_E block: step_n then runs two steps per pass where the
        # plan allows it (fdtd_update_he2)
        self._psi_e_out = None
        if self._E_out is not None and self._lib.fdtd_update_he2_supported(plan):
            self._psi_e_out = zeros()
        self._psi_h = [zeros()] + ([zeros()] if pingpong else [])
        self._psi_f = self._psi_g = None
        self._F_out = None
        if self.bloch_boundaries:
            self._psi_f = zeros()
            self._psi_g = [zeros()] + ([zeros()] if pingpong else [])
            if self._E_out is not None:
                self._F_out = torch.empty_like(self._F)   # the fused pass moves F with E
        self._inv_d = torch.cat(self.inv_dxes).contiguous() if self.inv_dxes is not None else None

        # sources (simulation.py:209-230 keeps the rendered template text of
        # each update): the specialisation the plan chose for each update
        self.sources["E"] = self._describe(0)
        self.sources["H"] = self._describe(1)
        if self.do_poynting:
            self.sources["S"] = self._describe(2)
        if self.bloch_boundaries:
            self.sources["F"] = self._describe(3)
            self.sources["G"] = self._describe(4)
        self.sources["step_n"] = self._describe(5)

        self.update_E = self._update_E
        self.update_H = self._update_H
        self.update_S = self._update_S if self.do_poynting else None
        self.update_F = self._noop if self.bloch_boundaries else None
        self.update_G = self._noop if self.bloch_boundaries else None
        self.update_J = self._update_J if do_fieldsrc else None

    def _describe(self, which: int) -> str:
        n = int(self._lib.fdtd_plan_describe(self._plan, which, None, 0))
        if n < 0:
            _lib.check(1, "fdtd_plan_describe")
        buf = ctypes.create_string_buffer(n + 1)
        self._lib.fdtd_plan_describe(self._plan, which, buf, n + 1)
        return buf.value.decode()

    # ------------------------------------------------------------------ fields
    def _create_field(self, initial_value):
        if initial_value is None:
            return torch.zeros(3 * self.N, dtype=self._tdtype, device=self.device)
        if len(initial_value) != 3:
            raise Exception("Initial field value must be a list of length 3")
        if not all(tuple(f.shape) == tuple(self.shape) for f in initial_value):
            raise Exception("Initial field list elements must have same shape as epsilon elements")
        return _vec(initial_value, self._tdtype, self.device)

    @property
    def E(self) -> DeviceField:
        self.flush()
        return self._E.as_subclass(DeviceField)

    @property
    def H(self) -> DeviceField:
        self.flush()
        return self._H[self._h].as_subclass(DeviceField)

    @property
    def F(self) -> Optional[DeviceField]:
        self.flush()
        return None if self._F is None else self._F.as_subclass(DeviceField)

    @property
    def G(self) -> Optional[DeviceField]:
        self.flush()
        return None if self._G is None else self._G[self._h if len(self._G) == 2 else 0].as_subclass(DeviceField)

    @property
    def S(self) -> Optional[DeviceField]:
        self.flush()
        return None if self._S is None else self._S.as_subclass(DeviceField)

    @property
    def oS(self) -> Optional[DeviceField]:
        self.flush()
        return None if self._oS is None else self._oS.as_subclass(DeviceField)

    @property
    def pending_h(self) -> bool:
        """True while an H half-step recorded by ``update_H`` has not run."""
        return self._pending_h

    def flush(self) -> None:
        """Runs the H half-step a deferred ``update_H`` recorded, if any."""
        if self._pending_h:
            self._pending_h = False
            self._launch(self._lib.fdtd_update_h, (), "fdtd_update_h")
            self._h = 1 - self._h

    # ------------------------------------------------------------------ plumbing
    def _stream(self) -> torch.cuda.Stream:
        return self.queue if self.queue is not None else torch.cuda.current_stream(self.device)

    def _fields(self, flush: bool = True) -> _lib.Fields:
        """The C-ABI view of the buffers; runs a deferred H half-step first
        unless ``flush`` is False."""
        if flush:
            self.flush()
        f = _lib.Fields()
        ptr = lambda t: None if t is None else t.data_ptr()
        f.E = ptr(self._E)
        f.H = ptr(self._H[self._h])
        f.H_out = ptr(self._H[1 - self._h]) if len(self._H) == 2 else f.H
        f.eps = ptr(self.eps)
        f.S = ptr(self._S)
        f.oS = ptr(self._oS)
        f.F = ptr(self._F)
        f.F_out = ptr(self._F_out)
        if self._G is not None:
            f.G = ptr(self._G[self._h if len(self._G) == 2 else 0])
            f.G_out = ptr(self._G[1 - self._h]) if len(self._G) == 2 else f.G
        else:
            f.G = f.G_out = None
        f.psi_e = ptr(self._psi_e)
        f.psi_h = ptr(self._psi_h[self._h if len(self._psi_h) == 2 else 0])
        f.psi_h_out = ptr(self._psi_h[1 - self._h]) if len(self._psi_h) == 2 else f.psi_h
        f.psi_f = ptr(self._psi_f)
        if self._psi_g is not None:
            f.psi_g = ptr(self._psi_g[self._h if len(self._psi_g) == 2 else 0])
            f.psi_g_out = ptr(self._psi_g[1 - self._h]) if len(self._psi_g) == 2 else f.psi_g
        else:
            f.psi_g = f.psi_g_out = None
        f.coef = ptr(self._coef)
        f.inv_d = ptr(self._inv_d)
        f.E_out = ptr(self._E_out)
        f.psi_e_out = ptr(self._psi_e_out)
        return f

    def _launch(self, fn, events, what, *extra):
        stream = self._stream()
        for e in events or ():
            stream.wait_event(e.torch_event if isinstance(e, Event) else e)
        f = self._fields(flush=False)
        with torch.cuda.device(self.device):
            _lib.check(fn(self._plan, ctypes.byref(f), *extra, ctypes.c_void_p(stream.cuda_stream)), what)
        return Event(stream)

    def _update_E(self, events: Sequence = ()) -> Event:
        if self._pending_h:
            # the recorded H half-step and this E half-step in one pass
            self._pending_h = False
            ev = self._launch(self._lib.fdtd_update_he, events, "fdtd_update_he", ctypes.c_int64(-1),
                              ctypes.c_void_p(None))
            self._h = 1 - self._h
            self._swap_e()
            return ev
        self.flush()
        return self._launch(self._lib.fdtd_update_e, events, "fdtd_update_e")

    def _swap_e(self) -> None:
        """E (and F with Bloch boundaries) moved to the other buffer."""
        self._E, self._E_out = self._E_out, self._E
        if self._F_out is not None:
            self._F, self._F_out = self._F_out, self._F

    def _update_H(self, events: Sequence = ()) -> Event:
        self.flush()
        if self._defer:
            stream = self._stream()
            for e in events or ():
                stream.wait_event(e.torch_event if isinstance(e, Event) else e)
            self._pending_h = True
            return Event(stream, deferred=True)
        ev = self._launch(self._lib.fdtd_update_h, events, "fdtd_update_h")
        if len(self._H) == 2:
            self._h = 1 - self._h
        return ev

    def _update_S(self, events: Sequence = ()) -> Event:
        if self.poynting_mode == "two_pass":
            return self._launch(self._lib.fdtd_update_s, events, "fdtd_update_s")
        ev = self._noop(events)
        if self._pending_h:
            ev._deferred = True          # S comes with the recorded H half-step
        return ev

    def _noop(self, events: Sequence = ()) -> Event:
        stream = self._stream()
        for e in events or ():
            stream.wait_event(e.torch_event if isinstance(e, Event) else e)
        return Event(stream)

    def _update_J(self, events, c, s, xmin, xmax, ymin, ymax, zmin, zmax) -> Event:
        self.flush()
        box = (ctypes.c_int64 * 6)(int(xmin), int(xmax), int(ymin), int(ymax), int(zmin), int(zmax))
        return self._launch(self._lib.fdtd_update_j, events, "fdtd_update_j",
                            ctypes.c_void_p(self.Jr.data_ptr()), ctypes.c_void_p(self.Ji.data_ptr()),
                            ctypes.c_double(c), ctypes.c_double(s), box)

    # ------------------------------------------------------------------ fast path
    def update_HE(self, events: Sequence = (), src_index: Optional[int] = None, src_amp=None) -> Event:
        """``update_H`` (with S) of one step and ``update_E`` of the next,
        then ``E[src_index] += src_amp[0]`` (a one-value device tensor), in
        one pass over the grid.  Needs ``fuse_steps`` and a plan that
        supports it (fused Poynting or Poynting off; with Bloch boundaries
        fused Poynting, FP32, uniform dx and sz <= 512, F moving with E).  The
        new E is in a second buffer afterwards: read ``sim.E`` again."""
        if self._E_out is None:
            raise _lib.FdtdError("update_HE is not available for this simulation (fdtd_update_he_supported)")
        self.flush()
        amp = None
        if src_amp is not None:
            if src_index is None:
                raise Exception("src_index is required with src_amp")
            amp = src_amp.to(device=self.device, dtype=self._tdtype).contiguous()
            self._amp_keepalive = amp
        ev = self._launch(self._lib.fdtd_update_he, events, "fdtd_update_he",
                          ctypes.c_int64(int(src_index) if amp is not None else -1),
                          ctypes.c_void_p(amp.data_ptr() if amp is not None else None))
        self._h = 1 - self._h
        self._swap_e()
        return ev

    def update_HE2(self, events: Sequence = (), src_index: Optional[int] = None, src_amp=None) -> Event:
        """``update_HE`` twice in one pass over the grid: H (with S), E,
        ``E[src_index] += src_amp[0]``, H (with S), E, ``E[src_index] +=
        src_amp[1]`` (``src_amp``: a two-value device tensor).  Needs a plan
        that supports it (FP32, fused Poynting or Poynting off, no Bloch
        boundaries, CPML slabs of one axis apart).  The new E is in a second buffer
        afterwards: read ``sim.E`` again."""
        if self._psi_e_out is None:
            raise _lib.FdtdError("update_HE2 is not available for this simulation (fdtd_update_he2_supported)")
        self.flush()
        amp = None
        if src_amp is not None:
            if src_index is None:
                raise Exception("src_index is required with src_amp")
            amp = src_amp.to(device=self.device, dtype=self._tdtype).contiguous()
            if amp.numel() < 2:
                raise Exception("src_amp needs two values")
            self._amp_keepalive = amp
        ev = self._launch(self._lib.fdtd_update_he2, events, "fdtd_update_he2",
                          ctypes.c_int64(int(src_index) if amp is not None else -1),
                          ctypes.c_void_p(amp.data_ptr() if amp is not None else None))
        self._h = 1 - self._h
        self._E, self._E_out = self._E_out, self._E
        self._psi_e, self._psi_e_out = self._psi_e_out, self._psi_e
        return ev

    def step_n(self, n: int, src_index: Optional[int] = None, src_wave=None, events: Sequence = ()) -> Event:
        """``n`` iterations of the reference caller loop (fdtd.py:162-175)
        enqueued without host synchronisation: E (+F), ``E[src_index] +=
        src_wave[k]``, H (+G) with S.  ``src_wave`` is a device tensor (or
        array) of ``n`` source values in the field dtype.  E and psi_E may end
        in either of their two buffers (``sim.E`` follows).  With ``defer_h``
        (the default where the fused pass exists) the last step's H half-step
        is left pending as ``update_H``'s is: the next ``step_n`` or
        ``update_E`` runs it in its first pass, and reading any field,
        ``flush`` or any other update runs it first, so consecutive calls
        run every H half-step fused with the next E half-step."""
        return self._step_n(n, src_index, src_wave, events, None)

    def step_n_observed(self, n: int, src_index: Optional[int] = None, src_wave=None, flux=None,
                        keep_E: Optional[Sequence[torch.Tensor]] = None,
                        keep_H: Optional[Sequence[torch.Tensor]] = None, events: Sequence = ()) -> Event:
        """``step_n`` with observations after every step k, on the device
        (fdtd_step_n_observed; no counterpart in the reference, whose caller
        reads fields on the host between steps):
        ``flux = (comp, axis, index, acc)``: ``acc[0] +=`` the FP64 sum of
        S_comp over the plane {axis coordinate == index} (``acc``: a float64
        device tensor), summed over many workgroups in a fixed order;
        ``keep_E`` / ``keep_H``: n device tensors of 3N values in the field
        dtype that receive E / H after step k.
        It runs one step per pass (the two-step pass does not keep the first
        step's S); the fields and S are those of ``step_n``."""
        obs = _lib.Observe()
        alive = []
        if flux is not None:
            comp, axis, index, acc = flux
            if not (isinstance(acc, torch.Tensor) and acc.dtype == torch.float64 and acc.device == self.device):
                raise Exception("the flux accumulator must be a float64 tensor on {}".format(self.device))
            obs.flux_acc = acc.data_ptr()
            obs.flux_comp, obs.flux_axis, obs.flux_index = int(comp), int(axis), int(index)
            if self._flux_partials is None:
                self._flux_partials = torch.empty(int(self._lib.fdtd_reduce_partials(self._plan)),
                                                  dtype=torch.float64, device=self.device)
            obs.flux_partials = self._flux_partials.data_ptr()
        for name, keep in (("E_keep", keep_E), ("H_keep", keep_H)):
            if keep is None:
                continue
            if len(keep) != n:
                raise Exception("{} holds {} tensors, need {}".format(name, len(keep), n))
            for t in keep:
                if (t.device != self.device or t.dtype != self._tdtype or t.numel() != 3 * self.N
                        or not t.is_contiguous()):
                    raise Exception("{}: each tensor must be a contiguous {} tensor of {} values on {}".format(
                        name, self._tdtype, 3 * self.N, self.device))
            arr = (ctypes.c_void_p * max(n, 1))(*[t.data_ptr() for t in keep])
            alive.append(arr)
            setattr(obs, name, ctypes.cast(arr, ctypes.c_void_p))
        return self._step_n(n, src_index, src_wave, events, obs)

    def _step_n(self, n, src_index, src_wave, events, obs) -> Event:
        # chained (fdtd_step_n_chain): a pending H half-step runs in the
        # call's first pass and the call's last one is left pending, as
        # update_H is, for the next call (or flush) to run
        chain = obs is None and self._defer
        if not chain:
            self.flush()
        stream = self._stream()
        for e in events or ():
            stream.wait_event(e.torch_event if isinstance(e, Event) else e)
        wave_ptr = None
        if src_wave is not None:
            if src_index is None:
                raise Exception("src_index is required with src_wave")
            wave = src_wave if isinstance(src_wave, torch.Tensor) else torch.as_tensor(numpy.asarray(src_wave))
            wave = wave.to(device=self.device, dtype=self._tdtype).contiguous()
            if wave.numel() < n:
                raise Exception("src_wave has {} values, need {}".format(wave.numel(), n))
            self._wave_keepalive = wave
            wave_ptr = wave.data_ptr()
        f = self._fields(flush=not chain)
        moved = ctypes.c_int32(0)
        idx = int(src_index if src_index is not None else -1)
        with torch.cuda.device(self.device):
            if chain:
                flags = _lib.CHAIN_DEFER_H | (_lib.CHAIN_H_PENDING if self._pending_h else 0)
                _lib.check(self._lib.fdtd_step_n_chain(self._plan, ctypes.byref(f), int(n), idx, wave_ptr, flags,
                                                       ctypes.c_void_p(stream.cuda_stream), ctypes.byref(moved)),
                           "fdtd_step_n_chain")
                self._pending_h = bool(moved.value & _lib.MOVED_H_PENDING)
            elif obs is None:
                _lib.check(self._lib.fdtd_step_n_moved(self._plan, ctypes.byref(f), int(n), idx, wave_ptr,
                                                       ctypes.c_void_p(stream.cuda_stream), ctypes.byref(moved)),
                           "fdtd_step_n_moved")
            else:
                _lib.check(self._lib.fdtd_step_n_observed(self._plan, ctypes.byref(f), int(n), idx, wave_ptr,
                                                          ctypes.byref(obs), ctypes.c_void_p(stream.cuda_stream),
                                                          ctypes.byref(moved)), "fdtd_step_n_observed")
        if len(self._H) == 2 and moved.value & 4:
            self._h = 1 - self._h
        if moved.value & 1:
            self._swap_e()
        if moved.value & 2:
            self._psi_e, self._psi_e_out = self._psi_e_out, self._psi_e
        return Event(stream)

    # ------------------------------------------------------------------ state
    def checkpoint(self) -> Dict[str, torch.Tensor]:
        """Copies of everything the step reads: E, H, psi (and F, G with
        Bloch).  ``restore`` puts them back, so stepping resumes bit for bit
        (``initial_fields`` restores only E/H/F/G, as in the reference)."""
        self.flush()
        st = {"E": self._E.clone(), "H": self._H[self._h].clone(), "psi_e": self._psi_e.clone(),
              "psi_h": self._psi_h[self._h if len(self._psi_h) == 2 else 0].clone()}
        if self._F is not None:
            g = self._h if len(self._G) == 2 else 0
            st.update(F=self._F.clone(), G=self._G[g].clone(), psi_f=self._psi_f.clone(),
                      psi_g=self._psi_g[g].clone())
        return st

    def restore(self, st: Dict[str, torch.Tensor]) -> None:
        self.flush()
        self._E.copy_(st["E"])
        self._H[self._h].copy_(st["H"])
        self._psi_e.copy_(st["psi_e"])
        self._psi_h[self._h if len(self._psi_h) == 2 else 0].copy_(st["psi_h"])
        if self._F is not None:
            g = self._h if len(self._G) == 2 else 0
            self._F.copy_(st["F"])
            self._G[g].copy_(st["G"])
            self._psi_f.copy_(st["psi_f"])
            self._psi_g[g].copy_(st["psi_g"])

    @property
    def fused_steps(self) -> bool:
        """True when step_n runs the fused H+E pass (see update_HE)."""
        return self._E_out is not None

    @property
    def two_step_passes(self) -> bool:
        """True when step_n runs two steps per pass (see update_HE2): the
        plan supports it and its time model beats the one-step pass."""
        return self._psi_e_out is not None and bool(self._lib.fdtd_update_he2_preferred(self._plan))

    @property
    def plan(self) -> ctypes.c_void_p:
        return self._plan

    @property
    def coef(self) -> torch.Tensor:
        return self._coef

    @property
    def psi_elems(self) -> int:
        return int(self._lib.fdtd_psi_elems(self._plan))

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                self._lib.fdtd_plan_destroy(plan)
            except Exception:
                pass
            self._plan = None
