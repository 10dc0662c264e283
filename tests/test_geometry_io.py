"""Geometry generators (pcgen.py / fdtd.py:28-81 equivalents), field I/O
(fdtd.py:181-191) and the plan's specialisation text (sim.sources,
simulation.py:209-230).  CPU only."""
import ctypes
import math

import numpy
import pytest

from opencl_fdtd_amd import geometry as G
from opencl_fdtd_amd import fieldio

R3 = math.sqrt(3) / 2


def _set(xy):
    return sorted((round(float(p[0]), 9), round(float(p[1]), 9)) + tuple(round(float(q), 9) for q in p[2:])
                  for p in xy)


def test_triangular_lattice_by_hand():
    # 3 x 3 symmetric: rows y = -r3, 0, r3; odd rows one site shorter, shifted by 1/2
    want = [(-0.5, -R3), (0.5, -R3), (-1, 0), (0, 0), (1, 0), (-0.5, R3), (0.5, R3)]
    assert _set(G.triangular_lattice([3, 3])) == _set(want)
    # asymmetrical: odd rows keep all three sites
    want_a = want + [(1.5, -R3), (1.5, R3)]
    assert _set(G.triangular_lattice([3, 3], asymmetrical=True)) == _set(want_a)
    # sorted by x
    xy = G.triangular_lattice([5, 5])
    assert numpy.all(numpy.diff(xy[:, 0]) >= 0)
    # rows come from dims[1] (the reference takes dims[0], which agrees only for square dims)
    xy = G.triangular_lattice([5, 3])
    assert len(numpy.unique(numpy.round(xy[:, 1], 9))) == 3
    assert xy.shape[0] == 5 + 4 + 4
    # nearest-neighbour distance 1 everywhere
    xy = G.triangular_lattice([7, 7])
    d = numpy.sqrt(((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1))
    d[d == 0] = 9
    assert numpy.allclose(d.min(axis=1), 1.0)


def test_square_lattice_by_hand():
    xy = G.square_lattice([2, 3])
    assert _set(xy) == _set([(x, y) for x in (-1, 0) for y in (-1.5, -0.5, 0.5)])


def test_defects_by_hand():
    # L3 in a 3 x 3 lattice: the centre row (x = -1, 0, 1) removed, the odd rows stay
    assert _set(G.ln_defect([1, 1], 3)) == _set([(-0.5, -R3), (0.5, -R3), (-0.5, R3), (0.5, R3)])
    # L1 in a 5 x 5 lattice removes the centre only
    full, l1 = G.triangular_lattice([5, 5]), G.ln_defect([2, 2], 1)
    assert l1.shape[0] == full.shape[0] - 1 and (0.0, 0.0) not in _set(l1)
    with pytest.raises(Exception):
        G.ln_defect([2, 2], 2)
    # R6: the six neighbours of the centre removed, the centre kept
    r6 = _set(G.r6_defect([2, 2]))
    assert len(r6) == 23 - 6 and (0.0, 0.0) in r6
    for p in ((1, 0), (-1, 0), (0.5, R3), (-0.5, -R3)):
        assert (round(p[0], 9), round(p[1], 9)) not in r6
    # shifted L3: holes at x = +-2 move out by 0.15, x = +-3 stays, +-4 by 0.075, radii scaled
    xyr = G.ln_shift_defect([4, 4], 3, (0.15, 0, 0.075), (0.9, 1, 1.2))
    row = {round(float(x), 9): float(r) for x, y, r in xyr if y == 0}
    assert row[2.15] == 0.9 and row[-2.15] == 0.9 and row[3.0] == 1.0 and row[4.075] == 1.2
    assert 2.0 not in row and 4.0 not in row and 1.0 not in row and 0.0 not in row
    assert numpy.all(xyr[xyr[:, 1] != 0, 2] == 1.0)


def test_l3_perturbed_holes():
    """The four beam-forming sites of l3_shift_perturbed_defect, with xs the
    distinct nonzero |x| and ys the distinct |y|: (xs[3], ys[1]),
    (xs[7], ys[1]), (xs[2], ys[2]), (xs[6], ys[2]), every sign."""
    xyr = G.l3_shift_perturbed_defect([10, 10], perturbed_radius=1.3, shifts_a=(0.15, 0, 0.075),
                                      shifts_r=(1, 1, 1))
    # with the example's shifts, xs = 0.5, 1, 1.5, 2, 2.15, 2.5, 3, 3.5, 4, 4.075, ...
    # and ys = 0, r3, 2 r3: the sites are (2, r3), (3.5, r3), (1.5, 2 r3), (3, 2 r3);
    # odd rows (|y| = r3) hold half-integer x, even rows integer x, so the
    # holes that exist there are (+-3.5, +-r3) and (+-3, +-2 r3)
    pert = _set(xyr[xyr[:, 2] == 1.3][:, :2])
    want = [(sx * 3.5, sy * R3) for sx in (-1, 1) for sy in (-1, 1)] + \
           [(sx * 3.0, sy * 2 * R3) for sx in (-1, 1) for sy in (-1, 1)]
    assert pert == _set(want)
    assert numpy.all(xyr[xyr[:, 2] != 1.3, 2] == 1.0)
    holes = G.perturbed_l3(430.0, 0.285)
    assert holes.shape[1] == 3 and numpy.isclose(holes[:, 2].min(), 430 * 0.285)


def test_nanobeam_by_hand():
    h = G.nanobeam_holes(0.7, 3, 2)
    x = [0.35, 1.15, 2.05, 3.05, 4.05]
    r = [0.7, 0.8, 0.9, 1.0, 1.0]
    assert numpy.allclose(h[:, 0], [-v for v in x[::-1]] + x)
    assert numpy.allclose(h[:, 1], r[::-1] + r)


def test_draw_holes_and_slab():
    m = G.draw_holes((7, 7), numpy.array([[0.0, 0.0]]), radius=1.5)
    assert m.sum() == 9 and m[3, 3] and m[2, 2] and not m[1, 3]
    m = G.draw_holes((9, 9), numpy.array([[1.0, 0.0, 0.5]]), pitch=2.0)     # r = 1 cell at x = +2 cells
    assert m.sum() == 5 and m[6, 4] and m[5, 4] and m[7, 4] and m[6, 3]
    eps = G.slab_epsilon((9, 9, 10), m, 3, 4, (12.25, 12.25, 11.9))
    assert eps[0][0, 0, 3] == 12.25 and eps[2][0, 0, 6] == 11.9 and eps[0][0, 0, 7] == 1.0
    assert eps[1][6, 4, 4] == 1.0 and eps[1][0, 0, 2] == 1.0


def test_phc_slab_workload_uses_the_l3_lattice():
    """The README-scale workload: holes of radius 7 on the pitch-20 triangular
    lattice with three sites of the centre row missing; the slab is 11 cells
    thick around the centre plane."""
    from opencl_fdtd_amd.workloads import phc_slab
    w = phc_slab(steps=4, shape=(120, 100, 31))
    e = w.epsilon[0]
    cz = 31 // 2
    slab = e[:, :, cz]
    assert e[:, :, cz - 6].max() == 1.0 and e[:, :, cz + 6].max() == 1.0
    assert slab[60, 50] == 12.25                 # defect centre: no hole
    assert slab[60 + 40, 50] == 1.0               # lattice site at x = +2 pitches on the centre row
    assert slab[60 + 20, 50] == 12.25             # the removed site at x = +1 pitch
    assert slab[60 + 10, 50 + 17] == 1.0          # odd row, x = +1/2 pitch, y = 17.3 cells
    assert w.epsilon[2][60 + 40, 50 + 10, cz] == 11.9


def test_vec_unvec_round_trip():
    rng = numpy.random.default_rng(0)
    shape = (3, 4, 5)
    comps = [rng.standard_normal(shape) for _ in range(3)]
    v = fieldio.vec(comps)
    assert v.shape == (180,)
    # C order inside a component, components one after another
    assert v[1 * 60 + (2 * 4 + 3) * 5 + 4] == comps[1][2, 3, 4]
    back = fieldio.unvec(v, shape)
    assert all(numpy.array_equal(a, b) for a, b in zip(back, comps))
    with pytest.raises(ValueError):
        fieldio.unvec(v[:-1], shape)


class _FakeSim:
    def __init__(self, shape, rng, bloch=False, poynting=True):
        n = 3 * int(numpy.prod(shape))
        self.shape, self.dt = shape, 0.5
        self.E, self.H = rng.standard_normal(n), rng.standard_normal(n)
        self.S = rng.standard_normal(n) if poynting else None
        self.F = rng.standard_normal(n) if bloch else None
        self.G = rng.standard_normal(n) if bloch else None


@pytest.mark.parametrize("bloch, poynting", [(False, True), (True, False)])
def test_save_and_load_fields(tmp_path, bloch, poynting):
    rng = numpy.random.default_rng(1)
    sim = _FakeSim((4, 3, 5), rng, bloch, poynting)
    path = fieldio.save_fields(str(tmp_path / "out"), sim, grid={"edges_x": numpy.arange(5.0)}, step=42)
    d = fieldio.load_fields(path)
    assert d["shape"] == (4, 3, 5) and d["dt"] == 0.5 and d["step"] == 42
    assert numpy.array_equal(d["grid"]["edges_x"], numpy.arange(5.0))
    names = {"E", "H"} | ({"S"} if poynting else set()) | ({"F", "G"} if bloch else set())
    assert names == {k for k in d if k in "EHSFG"}
    for name in names:
        assert numpy.array_equal(fieldio.vec(d[name]), getattr(sim, name))


def test_plan_describe_names_each_update():
    """sim.sources text from the plan: kernel, template flags, faces; host only."""
    from opencl_fdtd_amd import _lib
    lib = _lib.load()
    d = _lib.Desc()
    d.shape[:] = [20, 30, 73]
    d.dtype = _lib.F32
    d.n_face = 2
    for k, (ax, pol) in enumerate(((0, 0), (2, 1))):
        d.face[k].axis, d.face[k].polarity, d.face[k].thickness = ax, pol, 4
    d.dt, d.inv_dx, d.poynting = 0.5, 1.0, _lib.POYNTING_FUSED
    d.bloch_mode = _lib.BLOCH_PHASE
    plan = ctypes.c_void_p()
    assert lib.fdtd_plan_create(ctypes.byref(d), ctypes.byref(plan)) == 0
    try:
        texts = []
        for which in range(6):
            n = lib.fdtd_plan_describe(plan, which, None, 0)
            buf = ctypes.create_string_buffer(n + 1)
            assert lib.fdtd_plan_describe(plan, which, buf, n + 1) == n
            texts.append(buf.value.decode())
        assert "e_march<float, PAIR=0" in texts[0]
        assert "h_march<float" in texts[1] and "Poynting fused" in texts[1]
        assert "enqueues nothing" in texts[2]
        assert "he_march" in texts[5]
        for t in texts:
            assert "axis x, polarity n, thickness 4" in t and "axis z, polarity p, thickness 4" in t
        small = ctypes.create_string_buffer(8)
        assert lib.fdtd_plan_describe(plan, 0, small, 8) == len(texts[0]) and len(small.value) == 7
        assert lib.fdtd_plan_describe(plan, 9, None, 0) == -1
    finally:
        lib.fdtd_plan_destroy(plan)
