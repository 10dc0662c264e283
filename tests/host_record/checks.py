"""Checks of the library's host logic through the launch-recording build
(test infrastructure; no GPU): plan creation over many shapes, the launch
decomposition of every kernel (tiles of whole rows or of z segments with
their halo rows and columns, x chunks covering the grid), and the step_n schedule: the half-steps in
the order of the caller loop (fdtd.py:162-175), each pass reading the
buffers the previous one wrote, the source values in order, and the moved
bits reporting where E, psi_E and H end.

``run(lib_path)`` raises AssertionError on the first failure; run as a
script it exits 0 on success (used under AddressSanitizer)."""
from __future__ import annotations

import ctypes
import itertools
import re
import sys

import numpy

SLOTS = 16
(E_IN, E_OUT, H_IN, H_OUT, PSIE_IN, PSIE_OUT, PSIH_IN, PSIH_OUT, F_IN, F_OUT, G_IN, G_OUT, S, SRC, PTR0) = range(15)


class Record(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_char * 128), ("grid", ctypes.c_int32), ("block", ctypes.c_int32),
                ("has_geom", ctypes.c_int32)] + \
               [(n, ctypes.c_int32) for n in ("sx", "sy", "sz", "plane", "xchunk", "nchunk", "pblocks", "tile",
                                              "src_x", "src_p", "src_comp", "tyo", "tzo", "tzb", "tw")] + \
               [("buf", ctypes.c_void_p * SLOTS)]


def _load(path):
    from opencl_fdtd_amd import _lib as L
    lib = ctypes.CDLL(path)
    P, F = ctypes.c_void_p, ctypes.POINTER(L.Fields)
    lib.fdtd_plan_create.argtypes = [ctypes.POINTER(L.Desc), ctypes.POINTER(ctypes.c_void_p)]
    lib.fdtd_plan_destroy.argtypes = [P]
    lib.fdtd_plan_describe.argtypes = [P, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int64]
    lib.fdtd_plan_describe.restype = ctypes.c_int64
    lib.fdtd_psi_elems.argtypes = [P]
    lib.fdtd_psi_elems.restype = ctypes.c_int64
    lib.fdtd_step_n_moved.argtypes = [P, F, ctypes.c_int32, ctypes.c_int64, P, P, ctypes.POINTER(ctypes.c_int32)]
    lib.fdtd_step_n.argtypes = [P, F, ctypes.c_int32, ctypes.c_int64, P, P]
    lib.fdtd_step_n_chain.argtypes = [P, F, ctypes.c_int32, ctypes.c_int64, P, ctypes.c_int32, P,
                                      ctypes.POINTER(ctypes.c_int32)]
    lib.fdtd_step_n_observed.argtypes = [P, F, ctypes.c_int32, ctypes.c_int64, P, ctypes.POINTER(L.Observe), P,
                                         ctypes.POINTER(ctypes.c_int32)]
    lib.fdtd_shell_step_n.argtypes = [P, F, ctypes.c_int32, P, P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, P,
                                      P, ctypes.POINTER(ctypes.c_int32)]
    lib.fdtd_shell_hidden_elems.argtypes = [P, ctypes.c_int32]
    lib.fdtd_shell_hidden_elems.restype = ctypes.c_int64
    lib.fdtd_update_he2_preferred.argtypes = [P]
    lib.fdtd_update_he_supported.argtypes = [P]
    lib.fdtd_record_count.restype = ctypes.c_int64
    lib.fdtd_record_get.argtypes = [ctypes.c_int64, ctypes.POINTER(Record)]
    lib.fdtd_last_error.restype = ctypes.c_char_p
    return lib, L


def _records(lib):
    out = []
    for i in range(lib.fdtd_record_count()):
        r = Record()
        assert lib.fdtd_record_get(i, ctypes.byref(r)) == 0
        out.append(r)
    lib.fdtd_record_clear()
    return out


def _name(r):
    m = re.search(r"(\w+)<", r.kernel.decode())
    return m.group(1) if m else r.kernel.decode()


def _halo_rows(r):
    """Halo rows a whole-row marching launch's workgroup holds besides its
    tile."""
    name, text = _name(r), r.kernel.decode()
    if name == "heb_march":
        return 1
    if name == "h_march":
        mode = int(re.search(r"h_march<[^,]+,[^,]+,[^,]+,\s*(\d)", text).group(1))
        return 1 if mode == 2 else 0
    return 0


# he_march / he2_march: halo rows below and above the owned rows, halo
# columns before and after the owned cells of a z segment
TILE_HALO = {"he_march": (1, 0, 1, 0), "he2_march": (2, 1, 2, 1)}
_tiles_seen = set()


def check_tiles(r):
    """The 2-D tiles of he_march / he2_march (whole rows, or rows cut into z
    segments): every plane cell is owned by exactly one workgroup, and for
    each owned cell the workgroup holds its neighbours at every offset
    -yl..yh in y and -zl..zh in z, at the periodically wrapped cell, as the
    pass's stages need them.  The thread -> cell mapping restated from the
    kernel is enumerated for planes up to 70 000 cells."""
    name = _name(r)
    yl, yh, zl, zh = TILE_HALO[name]
    seg = r.tzb > 1
    if seg:
        assert r.tzo < r.sz and r.tzb == -(-r.sz // r.tzo) and r.tw == r.tzo + zl + zh, (name, r.tzo, r.tzb, r.tw)
    else:
        assert r.tzb == 1 and r.tzo == r.tw == r.sz, (name, r.tzo, r.tzb, r.tw, r.sz)
    assert r.tyo >= 1 and (r.tyo + yl + yh) * r.tw <= r.block, (name, r.tyo, r.tw, r.block)
    assert r.pblocks == -(-r.sy // r.tyo) * r.tzb, (name, r.pblocks, r.sy, r.tyo, r.tzb)
    key = (name, r.sy, r.sz, r.tyo, r.tzo, r.tzb, r.tw, r.block)
    if key in _tiles_seen or r.plane > 70000:
        return
    _tiles_seen.add(key)
    t = numpy.arange(r.block)
    rr, cz = t // r.tw, t % r.tw
    owned = numpy.zeros(r.plane, dtype=numpy.int64)
    for pb in range(r.pblocks):
        iy, iz = divmod(pb, r.tzb)
        y0 = iy * r.tyo
        nr = min(r.tyo, r.sy - y0) + yl + yh
        if seg:
            z0 = iz * r.tzo
            ncz = min(r.tzo, r.sz - z0) + zl + zh
            z = (z0 - zl + cz) % r.sz
        else:
            ncz = r.sz
            z = cz
        y = (y0 - yl + rr) % r.sy
        inn = (rr < nr) & (cz < ncz)
        own = inn & (rr >= yl) & (rr < nr - yh)
        if seg:
            own &= (cz >= zl) & (cz < ncz - zh)
        numpy.add.at(owned, (y * r.sz + z)[own], 1)
        Y = numpy.full((nr + 1, r.tw + 1), -1)
        Z = numpy.full((nr + 1, r.tw + 1), -1)
        Y[rr[inn], cz[inn]] = y[inn]
        Z[rr[inn], cz[inn]] = z[inn]
        ro, co = rr[own], cz[own]
        for dy in range(-yl, yh + 1):
            for dz in range(-zl, zh + 1):
                cn = co + dz if seg else (co + dz) % r.sz
                ok = (Y[ro + dy, cn] == (y[own] + dy) % r.sy) & (Z[ro + dy, cn] == (z[own] + dz) % r.sz)
                assert ok.all(), (name, r.sy, r.sz, r.tyo, r.tzo, pb, dy, dz)
    assert (owned == 1).all(), (name, r.sy, r.sz, r.tyo, r.tzo, int(owned.min()), int(owned.max()))


def check_decomposition(r):
    """The tiles cover the plane once and the x chunks cover every plane."""
    if not r.has_geom:
        return
    if _name(r) == "s_from_os":             # one thread per cell: plane blocks x planes
        assert r.grid == -(-r.plane // r.block) * r.sx, (r.grid, r.plane, r.block, r.sx)
        return
    assert r.pblocks * r.nchunk == r.grid, (_name(r), r.pblocks, r.nchunk, r.grid)
    assert r.xchunk > 0 and r.xchunk * r.nchunk >= r.sx > r.xchunk * (r.nchunk - 1), (_name(r), r.xchunk, r.sx)
    assert r.plane == r.sy * r.sz
    if _name(r) in TILE_HALO:
        check_tiles(r)
        return
    assert r.tile > 0 and r.tile * r.pblocks >= r.plane > r.tile * (r.pblocks - 1), (_name(r), r.tile, r.plane)
    halo = _halo_rows(r)
    if halo:
        assert r.tile % r.sz == 0, (_name(r), r.tile, r.sz)        # whole rows
        assert r.tile + halo * r.sz <= r.block, (_name(r), r.tile, halo, r.sz, r.block)


HALF = {"e_march": "E", "h_march": "H", "he_march": "HE", "heb_march": "HE", "he2_march": "HEHE"}


def _halves(r):
    """The half-steps a launch performs; a two-step pass launched without
    its last E (template argument E2 = false) does H, E, H."""
    name = _name(r)
    if name == "he2_march":
        args = [x.strip() for x in re.search(r"<(.*)>", r.kernel.decode()).group(1).split(",")]
        if len(args) >= 4 and args[3] == "false":
            return "HEH"
    return HALF[name]


def check_schedule(recs, f, n, moved, wave, esz, bloch, pingpong, obs=None, flags=0):
    """Half-steps E H E H ..., each launch reading what the previous wrote.
    flags (fdtd_step_n_chain): a pending H half-step first (1), the last H
    half-step left pending (2; *moved bit 3 reports a pending one).
    obs (fdtd_step_n_observed): dict with "flux" (S + comp * N * esz, the
    plane_sum input, or None) and "E_keep" / "H_keep" (destination lists or
    None); after the k-th E half-step E is copied to E_keep[k], after the
    k-th H half-step the flux is summed from S and H copied to H_keep[k]."""
    loc = {"E": f.E, "H": f.H, "psi_e": f.psi_e, "psi_h": f.psi_h, "F": f.F, "G": f.G}
    seq = []
    k_src = 0
    seen = {"flux": 0, "E": 0, "H": 0}
    prev = None
    for r in recs:
        name = _name(r)
        last, prev = prev, name
        if name in ("s_from_os", "add_point"):
            continue
        if name in ("plane_sum", "plane_partial"):
            assert obs is not None and obs["flux"] is not None, "plane sum without a flux"
            assert (name == "plane_partial") == (obs["partials"] is not None), name
            assert r.buf[PTR0] == obs["flux"], ("flux reads", r.buf[PTR0], obs["flux"])
            assert seq.count("H") == seen["flux"] + 1, ("flux after H half-step", seq, seen)
            seen["flux"] += 1
            continue
        if name == "reduce_accumulate":
            assert last == "plane_partial" and r.buf[PTR0] == obs["partials"], (last, r.buf[PTR0])
            continue
        if name == "copy":
            assert obs is not None
            dst, src = r.buf[E_OUT], r.buf[E_IN]
            which = "E" if obs["E_keep"] and dst in obs["E_keep"] else "H"
            k = seen[which]
            assert obs[which + "_keep"] and dst == obs[which + "_keep"][k], (which, k)
            assert src == loc[which], ("copy of the newest", which, k)
            assert seq.count(which) == k + 1, ("copy after the half-step", which, k, seq)
            assert r.grid == min(obs["bytes"], 2 ** 31 - 1), (r.grid, obs["bytes"])
            seen[which] += 1
            continue
        assert name in HALF, name
        b = r.buf
        if name == "e_march":
            assert b[E_IN] == loc["E"] and b[H_IN] == loc["H"], ("e_march reads", n)
            assert b[PSIE_IN] == loc["psi_e"]
            if bloch:
                assert b[F_IN] == loc["F"] and b[G_IN] == loc["G"]
        elif name == "h_march":
            assert b[E_IN] == loc["E"] and b[H_IN] == loc["H"] and b[PSIH_IN] == loc["psi_h"], ("h_march", n)
            loc["H"], loc["psi_h"] = b[H_OUT], b[PSIH_OUT]
            if pingpong:
                assert b[H_OUT] != b[H_IN]
            if bloch:
                assert b[F_IN] == loc["F"] and b[G_IN] == loc["G"]
                loc["G"] = b[G_OUT]
        else:
            assert b[E_IN] == loc["E"] and b[H_IN] == loc["H"], (name, n)
            assert b[PSIE_IN] == loc["psi_e"] and b[PSIH_IN] == loc["psi_h"], (name, n)
            assert b[E_OUT] != b[E_IN] and b[H_OUT] != b[H_IN]
            loc["E"], loc["H"] = b[E_OUT], b[H_OUT]
            loc["psi_e"], loc["psi_h"] = b[PSIE_OUT], b[PSIH_OUT]
            if name == "heb_march":
                assert b[F_IN] == loc["F"] and b[G_IN] == loc["G"]
                assert b[F_OUT] != b[F_IN] and b[G_OUT] != b[G_IN]
                loc["F"], loc["G"] = b[F_OUT], b[G_OUT]
        # the source of step k goes with the k-th E half-step, from wave + k
        if "E" in _halves(r) and wave:
            first_e = sum(1 for t in seq if t == "E")
            if name == "h_march":
                pass
            elif b[SRC]:
                assert b[SRC] == wave + first_e * esz, (name, first_e)
                k_src += _halves(r).count("E")
        seq += list(_halves(r))
    pend_out = bool(flags & 2 and n > 0) or bool(flags & 1 and n == 0)
    expect = ("H" if flags & 1 and n > 0 else "") + "EH" * n
    if flags & 2 and n > 0:
        expect = expect[:-1]
    assert "".join(seq) == expect, ("".join(seq), n, flags)
    if wave and n:
        assert k_src == n, (k_src, n)
    if obs is not None:
        assert seen["flux"] == (n if obs["flux"] is not None else 0), seen
        assert seen["E"] == (n if obs["E_keep"] else 0) and seen["H"] == (n if obs["H_keep"] else 0), seen
    if moved is not None:
        assert bool(moved & 1) == (loc["E"] == f.E_out) and (loc["E"] in (f.E, f.E_out)), moved
        assert bool(moved & 2) == (loc["psi_e"] == f.psi_e_out) or f.psi_e_out is None
        assert bool(moved & 4) == (loc["H"] == f.H_out and f.H_out != f.H), moved
        assert bool(moved & 8) == pend_out, (moved, n, flags)
    else:
        assert loc["E"] == f.E and loc["psi_e"] == f.psi_e


SHELL_KERNELS = ("shell_compose", "shell_conv", "shell_add")


def check_shell_schedule(recs, f, n, period, scratch, wave, esz, moved, pingpong):
    """fdtd_shell_step_n: the half-steps as step_n's (E H E H ..., each
    launch reading what the previous wrote, the moved bits), and between
    them the correction: shell_compose first; right after the E half-step
    of each corrected step k (k a multiple of the period) one shell_conv
    that reads the newest E and H and writes the scratch; the E update of
    step k + 1 (an e_march, a he_march, or the first E of a he2_march)
    carries the scratch as its pending correction, and no other E update
    does; with period 1 a final shell_add of the scratch into the newest
    E."""
    plain = [r for r in recs if _name(r) not in SHELL_KERNELS]
    check_schedule(plain, f, n, moved, wave, esz, False, pingpong)
    corrected = lambda k: k % period == 0
    if n == 0:
        assert not recs, [_name(r) for r in recs]
        return
    assert _name(recs[0]) == "shell_compose", _name(recs[0])
    loc = {"E": f.E, "H": f.H}
    n_e = 0                   # E half-steps made so far
    last = None               # the last half-step made ("E" or "H")
    convs, adds = [], 0
    for r in recs[1:]:
        name, b = _name(r), r.buf
        if name == "shell_conv":
            assert last == "E" and corrected(n_e - 1), ("correction after step", n_e - 1, last)
            assert b[E_IN] == loc["E"] and b[H_IN] == loc["H"], "the correction reads the newest E and H"
            convs.append(n_e - 1)
            continue
        if name == "shell_add":
            assert period == 1 and r is recs[-1], "shell_add ends a period-1 call"
            assert b[E_OUT] == loc["E"]
            adds += 1
            continue
        if name == "shell_compose":
            raise AssertionError("weights composed twice")
        halves = _halves(r)
        k = n_e                                  # step of the launch's first E half-step
        if "E" in halves:
            want = scratch if (k >= 1 and corrected(k - 1)) else None
            assert b[PTR0] == want, (name, "step", k, "pending", b[PTR0], want)
        if name == "h_march":
            loc["H"] = b[H_OUT]
        elif name != "e_march":
            loc["E"], loc["H"] = b[E_OUT], b[H_OUT]
        n_e += halves.count("E")
        last = halves[-1]
    assert convs == [k for k in range(n) if corrected(k)] if period == 1 else \
        convs == [k for k in range(0, n - 1) if corrected(k)], (convs, n, period)
    assert adds == (1 if period == 1 else 0), adds


def _shell(lib, L, plan, shape, esz, poynting):
    """fdtd_shell_step_n on fake device addresses: periods 1, 2 and 4, with
    and without the fused passes' buffers and the moved bits."""
    T = 4 if min(shape) >= 12 else 2
    nh = lib.fdtd_shell_hidden_elems(plan, T)
    assert nh > 0, lib.fdtd_last_error()
    checked = 0
    for period, with_out, use_moved in itertools.product((1, 2, 4), (False, True), (False, True)):
        for n in ((0, 1, 2, 3, 5) if period == 1 else (0, period, 2 * period, 3 * period)):
            f = L.Fields()
            addrs = iter(range(0x100000, 0x10000000, 0x100000))
            for name, _ in L.Fields._fields_:
                setattr(f, name, next(addrs))
            f.F = f.G = f.G_out = f.psi_f = f.psi_g = f.psi_g_out = f.F_out = None
            f.oS = None
            if not with_out:
                f.E_out = f.psi_e_out = None
            w, scratch, wave = 0x71000000, 0x72000000, 0x70000000
            moved = ctypes.c_int32(-1)
            rc = lib.fdtd_shell_step_n(plan, ctypes.byref(f), T, w, scratch, n, period, 1, wave, None,
                                       ctypes.byref(moved) if use_moved else None)
            assert rc == 0, (shape, n, period, lib.fdtd_last_error())
            recs = _records(lib)
            for r in recs:
                if _name(r) not in SHELL_KERNELS:
                    check_decomposition(r)
            check_shell_schedule(recs, f, n, period, scratch, wave, esz, moved.value if use_moved else None, True)
            checked += 1
    return checked


def _desc(L, shape, faces, dtype, poynting, bloch_mask=0, bloch_mode=0):
    d = L.Desc()
    d.shape[:] = list(shape)
    d.dtype = dtype
    d.n_face = len(faces)
    for k, (a, p, t) in enumerate(faces):
        d.face[k].axis, d.face[k].polarity, d.face[k].thickness = a, p, t
    d.dt, d.inv_dx, d.poynting = 0.5, 1.0, poynting
    d.bloch_mask, d.bloch_mode = bloch_mask, bloch_mode
    d.bloch_re[:] = [1.0, 1.0, 1.0]
    d.bloch_im[:] = [0.0, 0.0, 0.0]
    return d


SHAPES = [(1, 1, 1), (2, 3, 4), (3, 3, 3), (7, 5, 73), (5, 37, 3), (33, 4, 130), (2, 9, 10), (4, 25, 40),
          (120, 120, 73), (395, 345, 73), (64, 64, 256), (128, 128, 128), (24, 40, 512), (5, 6, 513),
          (1000, 10, 3), (3, 700, 400), (264, 264, 136), (17, 1, 1), (1, 17, 1), (1, 1, 300),
          (4, 7, 700), (6, 30, 300), (3, 2, 270), (2, 3, 1500)]


def _observed(lib, L, plan, shape, esz, poynting, bloch, pingpong):
    """fdtd_step_n_observed: the flux after every H half-step, the copies of
    E and H after every half-step, in the step_n schedule (one step per
    pass), on fake device addresses."""
    N = shape[0] * shape[1] * shape[2]
    checked = 0
    for n, flux, keep in itertools.product((0, 1, 2, 3, 5, 8), (False, "one", "many"), (False, True)):
        if flux and poynting == 0:
            continue
        f = L.Fields()
        addrs = iter(range(0x100000, 0x10000000, 0x100000))
        for name, _ in L.Fields._fields_:
            setattr(f, name, next(addrs))
        if not pingpong:
            f.H_out, f.psi_h_out, f.G_out, f.psi_g_out = f.H, f.psi_h, f.G, f.psi_g
        if not bloch:
            f.F = f.G = f.G_out = f.psi_f = f.psi_g = f.psi_g_out = f.F_out = None
        if poynting != 2:
            f.oS = None
        ob = L.Observe()
        info = {"flux": None, "partials": None, "E_keep": None, "H_keep": None, "bytes": 3 * N * esz}
        arrays = []
        if flux:
            comp, axis = 2, 1
            ob.flux_acc, ob.flux_comp, ob.flux_axis, ob.flux_index = 0x60000000, comp, axis, shape[1] // 2
            info["flux"] = f.S + comp * N * esz
            if flux == "many":
                ob.flux_partials = info["partials"] = 0x68000000
        if keep:
            for name, base in (("E_keep", 0x20000000), ("H_keep", 0x40000000)):
                ptrs = [base + k * 0x100000 for k in range(n)]
                arr = (ctypes.c_void_p * max(n, 1))(*ptrs)
                arrays.append(arr)
                setattr(ob, name, ctypes.cast(arr, ctypes.c_void_p))
                info[name] = ptrs
        moved = ctypes.c_int32(-1)
        wave = 0x70000000
        rc = lib.fdtd_step_n_observed(plan, ctypes.byref(f), n, 1, wave, ctypes.byref(ob), None, ctypes.byref(moved))
        assert rc == 0, (shape, n, flux, keep, lib.fdtd_last_error())
        recs = _records(lib)
        assert not any(_name(r) == "he2_march" for r in recs), "observed steps run one step per pass"
        for r in recs:
            check_decomposition(r)
        check_schedule(recs, f, n, moved.value, wave, esz, bloch, pingpong, obs=info)
        checked += 1
    return checked


def run(lib_path: str, quick: bool = False) -> int:
    lib, L = _load(lib_path)
    lib.fdtd_record_clear()
    checked = 0
    shapes = SHAPES[:8] if quick else SHAPES
    for shape, dtype, poynting, bloch in itertools.product(shapes, (0, 1), (0, 1, 2), (0, 3)):
        T = max(1, min(4, min(shape) // 3))
        faces = [(a, p, T) for a in range(3) for p in (0, 1)]
        d = _desc(L, shape, faces, dtype, poynting, bloch_mask=bloch)
        plan = ctypes.c_void_p()
        rc = lib.fdtd_plan_create(ctypes.byref(d), ctypes.byref(plan))
        assert rc == 0, (shape, dtype, poynting, bloch, lib.fdtd_last_error())
        try:
            for which in range(6):
                nch = lib.fdtd_plan_describe(plan, which, None, 0)
                buf = ctypes.create_string_buffer(nch + 1)
                assert lib.fdtd_plan_describe(plan, which, buf, nch + 1) == nch > 0
            esz = 4 if dtype == 0 else 8
            # H ping-pong: always with fused S; with Poynting off both ways
            # (in place, or two buffers for the fused passes without S)
            for n, with_out, homekeep, pingpong in itertools.product(
                    (0, 1, 2, 3, 4, 5, 8, 11), (False, True), (False, True),
                    (False, True) if poynting == 0 else (poynting == 1,)):
                f = L.Fields()
                addrs = iter(range(0x100000, 0x10000000, 0x100000))
                for name, _ in L.Fields._fields_:
                    setattr(f, name, next(addrs))
                if not pingpong:
                    f.H_out, f.psi_h_out, f.G_out, f.psi_g_out = f.H, f.psi_h, f.G, f.psi_g
                if not with_out:
                    f.E_out = f.psi_e_out = f.F_out = None
                if not bloch:
                    f.F = f.G = f.G_out = f.psi_f = f.psi_g = f.psi_g_out = f.F_out = None
                if poynting != 2:
                    f.oS = None
                wave = 0x70000000
                moved = ctypes.c_int32(-1)
                if homekeep:
                    rc = lib.fdtd_step_n(plan, ctypes.byref(f), n, 1, wave, None)
                else:
                    rc = lib.fdtd_step_n_moved(plan, ctypes.byref(f), n, 1, wave, None, ctypes.byref(moved))
                assert rc == 0, (shape, n, lib.fdtd_last_error())
                recs = _records(lib)
                for r in recs:
                    check_decomposition(r)
                check_schedule(recs, f, n, None if homekeep else moved.value, wave, esz, bool(bloch), pingpong)
                checked += 1
                if homekeep:
                    continue
                for flags in (1, 2, 3):
                    moved = ctypes.c_int32(-1)
                    rc = lib.fdtd_step_n_chain(plan, ctypes.byref(f), n, 1, wave, flags, None, ctypes.byref(moved))
                    assert rc == 0, (shape, n, flags, lib.fdtd_last_error())
                    recs = _records(lib)
                    for r in recs:
                        check_decomposition(r)
                    check_schedule(recs, f, n, moved.value, wave, esz, bool(bloch), pingpong, flags=flags)
                    checked += 1
            checked += _observed(lib, L, plan, shape, esz, poynting, bool(bloch), poynting in (0, 1))
            if not bloch and poynting != 2 and min(shape) >= 8:
                checked += _shell(lib, L, plan, shape, esz, poynting)
        finally:
            lib.fdtd_plan_destroy(plan)
    return checked


if __name__ == "__main__":
    sys.path.insert(0, sys.argv[2]) if len(sys.argv) > 2 else None
    n = run(sys.argv[1])
    print("host schedule checks: %d schedules" % n, flush=True)
