"""CPU oracle of the learned absorbing shell (TEST INFRASTRUCTURE ONLY).

Restates the shell correction (no counterpart in the reference): on a shell
of T cells per face, split into six boxes (x faces first, then y, then z),
  h_k(j) = sum_{c,t} W1[b][k][c][t] x_c(j + t)        x = (E, H), 16 hidden
  y_o(i) = sum_{k,t} W2[b][d(i)][o][k][t] h_k(i + t)  3 outputs, i in box b
  E_out = E + y on the shell cells,
with t over the 27 offsets of a periodic 3x3x3 stencil, weights in BF16, and
d(i) the depth of i: its index along the box's thin axis (0..T-1; four
output heads are stored whatever T).  The net has no activation, so per box
and depth the two layers are one 5x5x5 convolution:
  Wc[b][d][o][c][s] = sum_k sum_u W2[b][d][o][k][u] W1[b][k][c][s - u]
  y_o(i) = sum_{c,s} Wc[b][d(i)][o][c][s] x_c(i + s).

``forward(..., exact_f32=True)`` reproduces the GPU arithmetic: the composed
weights summed in float32 in the GPU's order (k outer, u inner) and rounded
once to BF16, inputs rounded to BF16, products and sums in float32 (the
GPU's matrix cores sum in another order).  ``forward(..., exact_f32=False)``
and ``backward`` work in float64 on the two layers without rounding (the
backward treats the BF16 roundings as identity, as the GPU backward does);
``forward_composed`` is the float64 composed form, which the tests hold
against the two-layer form.
"""
from __future__ import annotations

import numpy

N_HID, N_IN, N_OUT, N_TAP = 16, 6, 3, 27
N_DEPTH = 4                                    # output heads per box (T <= 4)
W1_SIZE = N_HID * N_IN * N_TAP
W2_SIZE = N_OUT * N_HID * N_TAP                # one head
N_WEIGHTS = 6 * (W1_SIZE + N_DEPTH * W2_SIZE)


def offsets():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(N_TAP)]


def boxes(shape, T):
    """Six output boxes [o0, o1) per axis partitioning the shell."""
    out = []
    for r in range(3):
        for side in range(2):
            lo, hi = [], []
            for q in range(3):
                if q < r:
                    lo.append(T); hi.append(shape[q] - T)
                elif q > r:
                    lo.append(0); hi.append(shape[q])
                else:
                    lo.append(0 if side == 0 else shape[q] - T)
                    hi.append(T if side == 0 else shape[q])
            out.append((tuple(lo), tuple(hi)))
    return out


def box_mask(shape, box):
    m = numpy.zeros(shape, bool)
    (a0, b0, c0), (a1, b1, c1) = box
    m[a0:a1, b0:b1, c0:c1] = True
    return m


def bf16_round(a):
    a32 = numpy.asarray(a, dtype=numpy.float32)
    u = a32.view(numpy.uint32).astype(numpy.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(numpy.uint32).view(numpy.float32)
    return numpy.where(special, a32, r)


def split_weights(w, nbox=6):
    """Flat weight vector (float) -> W1 [box][16][6][27], W2 [box][depth][3][16][27]."""
    w = numpy.asarray(w)
    W1 = w[:nbox * W1_SIZE].reshape(nbox, N_HID, N_IN, N_TAP)
    W2 = w[nbox * W1_SIZE:].reshape(nbox, N_DEPTH, N_OUT, N_HID, N_TAP)
    return W1, W2


def depth_masks(shape, box, b):
    """Masks of the box's output cells at each depth along its thin axis."""
    r = b // 2
    (lo, hi) = box
    idx = numpy.arange(shape[r]).reshape([-1 if q == r else 1 for q in range(3)])
    m = box_mask(shape, box)
    return [m & ((idx - lo[r]) == d) for d in range(N_DEPTH)]


def _shift(a, d):
    """a[j + d] with periodic wrap, over the last three axes."""
    return numpy.roll(a, (-d[0], -d[1], -d[2]), axis=(-3, -2, -1))


def compose(w, exact_f32=True):
    """Wc[b][d][o][c][s0][s1][s2], offsets s - 2 along the grid axes: the
    two layers of each box and depth as one 5x5x5 kernel.  exact_f32: float32
    sums in the GPU's order (k outer, u inner; the products of BF16 values are
    exact), rounded to BF16 once; else float64."""
    W1, W2 = split_weights(w)
    dt = numpy.float32 if exact_f32 else numpy.float64
    W1 = W1.astype(dt)
    W2 = W2.astype(dt)
    Wc = numpy.zeros((6, N_DEPTH, N_OUT, N_IN, 5, 5, 5), dt)
    offs = offsets()
    for k in range(N_HID):
        for u, du in enumerate(offs):
            a = W2[:, :, :, k, u][:, :, :, None]             # (b, d, o, 1)
            for v, dv in enumerate(offs):
                s = (du[0] + dv[0] + 2, du[1] + dv[1] + 2, du[2] + dv[2] + 2)
                Wc[:, :, :, :, s[0], s[1], s[2]] += a * W1[:, None, None, k, :, v]
    return bf16_round(Wc) if exact_f32 else Wc


def _composed(x, Wc, shape, T, dt):
    y = numpy.zeros((3,) + shape, dt)
    masks = [(b, d, m) for b, box in enumerate(boxes(shape, T)) for d, m in enumerate(depth_masks(shape, box, b))
             if m.any()]
    for s0 in range(5):
        for s1 in range(5):
            for s2 in range(5):
                xs = _shift(x, (s0 - 2, s1 - 2, s2 - 2))
                for b, d, m in masks:
                    y[:, m] += (Wc[b, d, :, :, s0, s1, s2] @ xs[:, m]).astype(dt)
    return y


def forward(E, H, w, T, exact_f32=True):
    """E, H: (3, sx, sy, sz); w: flat weights already rounded to BF16 values
    (float).  Returns (E_out, y) with y the correction on the shell."""
    shape = E.shape[1:]
    if exact_f32:
        x = bf16_round(numpy.concatenate([E, H]))
        y = _composed(x, compose(w, True), shape, T, numpy.float32)
        return (E.astype(numpy.float32) + y).astype(numpy.float32), y
    W1, W2 = split_weights(w)
    dt = numpy.float64
    x = numpy.concatenate([E, H]).astype(dt)
    xs = [_shift(x, d) for d in offsets()]
    y = numpy.zeros((3,) + shape, dt)
    for b, box in enumerate(boxes(shape, T)):
        h = numpy.zeros((N_HID,) + shape, dt)
        for k in range(N_HID):
            acc = numpy.zeros(shape, dt)
            for t in range(N_TAP):
                for c in range(N_IN):
                    acc = acc + dt(W1[b, k, c, t]) * xs[t][c]
            h[k] = acc
        hs = [_shift(h, d) for d in offsets()]
        for dep, m in enumerate(depth_masks(shape, box, b)):
            if not m.any():
                continue
            yb = numpy.zeros((3,) + shape, dt)
            for t in range(N_TAP):
                for k in range(N_HID):
                    for o in range(N_OUT):
                        yb[o] = yb[o] + dt(W2[b, dep, o, k, t]) * hs[t][k]
            y[:, m] = yb[:, m]
    Eout = (E.astype(dt) + y).astype(dt)
    return Eout, y


def forward_composed(E, H, w, T):
    """float64 correction through the composed 5x5x5 kernels (no rounding)."""
    shape = E.shape[1:]
    x = numpy.concatenate([E, H]).astype(numpy.float64)
    return _composed(x, compose(w, False), shape, T, numpy.float64)


def backward(E, H, w, T, Ebar):
    """FP64 adjoint of E_out = E + y(E, H): returns (xbar_E, xbar_H, wbar)
    with wbar in the flat weight layout; BF16 roundings as identity."""
    shape = E.shape[1:]
    W1, W2 = split_weights(numpy.asarray(w, dtype=numpy.float64))
    x = numpy.concatenate([E, H]).astype(numpy.float64)
    xs = [_shift(x, d) for d in offsets()]
    xbar = numpy.zeros_like(x)
    W1b = numpy.zeros_like(W1)
    W2b = numpy.zeros_like(W2)
    n = int(numpy.prod(shape))
    X = numpy.array(xs).reshape(N_TAP, N_IN, n)                  # x_c(j + t)
    for b, box in enumerate(boxes(shape, T)):
        h = numpy.einsum("kct,tcn->kn", W1[b], X).reshape((N_HID,) + shape)
        Hs = numpy.array([_shift(h, d) for d in offsets()]).reshape(N_TAP, N_HID, n)
        hbar = numpy.zeros_like(h)
        for dep, m in enumerate(depth_masks(shape, box, b)):
            yb = numpy.where(m, Ebar, 0.0)
            # W2bar[d][o][k][t] = sum_{i at depth d} ybar_o(i) h_k(i + t)
            W2b[b, dep] = numpy.einsum("on,tkn->okt", yb.reshape(N_OUT, n), Hs)
            # hbar_k(j) = sum_{o,t} W2[d(j - t)][o][k][t] ybar_o(j - t)
            for t, d in enumerate(offsets()):
                hbar += numpy.einsum("ok,on->kn", W2[b, dep][:, :, t],
                                     _shift(yb, (-d[0], -d[1], -d[2])).reshape(N_OUT, n)).reshape(h.shape)
        # W1bar[k][c][t] = sum_j hbar_k(j) x_c(j + t)
        W1b[b] = numpy.einsum("kn,tcn->kct", hbar.reshape(N_HID, n), X)
        for t, d in enumerate(offsets()):
            xbar += numpy.einsum("kc,kn->cn", W1[b][:, :, t],
                                 _shift(hbar, (-d[0], -d[1], -d[2])).reshape(N_HID, n)).reshape(x.shape)
    wbar = numpy.concatenate([W1b.ravel(), W2b.ravel()])
    return xbar[:3], xbar[3:], wbar
