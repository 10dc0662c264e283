"""Host-side checks that need no GPU: the C-ABI library loads and exports what
include/fdtd.h declares, plan validation, coefficient agreement with the
oracle, workload builders, and the loud failure without a GPU."""
import ctypes
import os
import re

import numpy
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "fdtd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fdtd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from opencl_fdtd_amd import _lib
    lib = _lib.load()
    declared = _header_functions()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.EXPORTS)
    assert lib.fdtd_abi_version() == _lib.ABI_VERSION


def test_struct_layouts_match_header():
    """ctypes mirrors of fdtd_desc / fdtd_fields have the C sizes."""
    from opencl_fdtd_amd import _lib
    # fdtd_face: 4 int32; fdtd_desc: 3 int64 + 2 int32 + 12 faces + 2 double + 4 int32 + 6 double
    assert ctypes.sizeof(_lib.Face) == 16
    assert ctypes.sizeof(_lib.Desc) == 24 + 8 + 12 * 16 + 16 + 16 + 48
    assert ctypes.sizeof(_lib.Fields) == 20 * 8
    assert ctypes.sizeof(_lib.AdjointBufs) == 13 * 8 + 8 + 8 + 4 * 4 + 6 * 4


def _desc(shape=(10, 11, 12), faces=(), **kw):
    from opencl_fdtd_amd import _lib
    d = _lib.Desc()
    d.shape[:] = list(shape)
    d.dtype = kw.get("dtype", 0)
    d.n_face = len(faces)
    for k, (a, p, t) in enumerate(faces):
        d.face[k].axis, d.face[k].polarity, d.face[k].thickness = a, p, t
    d.dt = kw.get("dt", 0.5)
    d.inv_dx = 1.0
    d.poynting = kw.get("poynting", 1)
    d.bloch_mask = kw.get("bloch_mask", 0)
    d.bloch_mode = kw.get("bloch_mode", 0)
    return d


def _create(d):
    from opencl_fdtd_amd import _lib
    lib = _lib.load()
    plan = ctypes.c_void_p()
    rc = lib.fdtd_plan_create(ctypes.byref(d), ctypes.byref(plan))
    msg = lib.fdtd_last_error().decode()
    if rc == 0:
        psi = lib.fdtd_psi_elems(plan)
        coef = lib.fdtd_coef_elems(plan)
        lib.fdtd_plan_destroy(plan)
        return rc, msg, psi, coef
    return rc, msg, None, None


def test_plan_sizes():
    faces = [(a, p, 8) for a in range(3) for p in range(2)]
    rc, _, psi, coef = _create(_desc((20, 30, 40), faces))
    assert rc == 0
    # psi: two arrays per face of T * (area of the face)
    assert psi == 2 * 8 * 2 * (30 * 40 + 20 * 40 + 20 * 30)
    assert coef == 6 * 8 * 6


@pytest.mark.parametrize("bad, needle", [
    (dict(dt=0.0), "dt"),
    (dict(dt=-1.0), "dt"),
    (dict(dtype=2), "dtype"),
    (dict(poynting=7), "poynting"),
    (dict(bloch_mask=9), "bloch_mask"),
    (dict(bloch_mode=5), "bloch_mode"),
])
def test_plan_rejects_bad_descriptions(bad, needle):
    rc, msg, _, _ = _create(_desc(**bad))
    assert rc != 0 and needle in msg


def test_plan_rejects_bad_faces():
    assert _create(_desc(faces=[(3, 0, 4)]))[0] != 0
    assert _create(_desc(faces=[(0, 2, 4)]))[0] != 0
    assert _create(_desc(faces=[(0, 0, 0)]))[0] != 0
    rc, msg, _, _ = _create(_desc(faces=[(0, 0, 2), (0, 1, 2), (0, 0, 3)]))
    assert rc != 0 and "more than two" in msg
    d = _desc((0, 4, 4))
    assert _create(d)[0] != 0


def test_plan_rejects_grids_beyond_32bit_offsets():
    rc, msg, _, _ = _create(_desc((1100, 1000, 1000)))
    assert rc != 0 and "4 GiB" in msg


def test_coefficients_match_oracle_bitwise():
    from opencl_fdtd_amd.pml import face_coefficients, with_defaults
    from oracle.oracle import pml_coefficients, pml_defaults
    for dtype in (numpy.float32, numpy.float64):
        for pml in ({"axis": "x", "polarity": "n"}, {"axis": "z", "polarity": "p", "thickness": 16},
                    {"axis": "y", "polarity": "n", "cfs_alpha": 0.05, "m": 3, "ma": 2, "epsilon_eff": 2.0}):
            a = face_coefficients(with_defaults(dict(pml)), 0.5715, dtype)
            b = pml_coefficients(pml_defaults(dict(pml)), 0.5715, dtype)
            for x, y in zip(a, b):
                assert x.dtype == dtype
                assert numpy.array_equal(x, y)


def test_simulation_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from opencl_fdtd_amd import Simulation, _lib
    eps = [numpy.ones((4, 4, 4))] * 3
    with pytest.raises(_lib.FdtdError):
        Simulation(eps, [])


def test_type_to_c():
    from opencl_fdtd_amd import type_to_C
    assert type_to_C(numpy.float32) == "float"
    assert type_to_C(numpy.float64) == "double"
    assert type_to_C(numpy.float16) == "half"
    with pytest.raises(Exception):
        type_to_C(numpy.int32)


def test_workloads():
    from opencl_fdtd_amd import workloads as W
    w = W.small(steps=10)
    assert w.shape == (64, 64, 64) and len(w.pmls) == 6
    e = w.epsilon[0]
    assert e.min() == 1.0 and e.max() <= 12 and (e > 1).mean() > 0.05
    assert 0 <= w.src_index < 3 * w.cells
    s = W.phc_slab(steps=10)
    assert s.shape == (395, 345, 73)
    exx = s.epsilon[0]
    slab = exx[:, :, 31:42]
    assert numpy.all(exx[:, :, :31] == 1) and numpy.all(exx[:, :, 42:] == 1)
    fill = (slab == 12.25).mean()
    assert 0.5 < fill < 0.9              # holes of r = 7 on a pitch-20 lattice
    # L3 defect: the centre cell of the slab is dielectric
    assert exx[197, 172, 36] == 12.25 and s.epsilon[2][197, 172, 36] == 11.9
    b = W.bloch(steps=10)
    assert numpy.isinf(b.epsilon[0][:, :, 200:205]).all()
    assert abs(numpy.isinf(b.epsilon[0]).mean() - 5 / 256) < 1e-9
    assert len(b.bloch_boundaries) == 2 and [p["axis"] for p in b.pmls] == ["z", "z"]
    bytes_c2 = W.algorithmic_bytes_per_cell(s.shape, s.pmls)
    assert abs(bytes_c2 - 105.8) < 0.1


def test_update_he_support_rules():
    """The fused H+E pass needs fused Poynting or Poynting off (long rows
    are cut into z segments), and with Bloch boundaries fused Poynting, FP32
    and rows of at most 512 cells (plan-level query, no GPU)."""
    from opencl_fdtd_amd import _lib
    lib = _lib.load()

    def supported(**kw):
        shape = kw.pop("shape", (10, 11, 12))
        d = _desc(shape, [(0, 0, 2)], **kw)
        plan = ctypes.c_void_p()
        assert lib.fdtd_plan_create(ctypes.byref(d), ctypes.byref(plan)) == 0
        r = lib.fdtd_update_he_supported(plan)
        lib.fdtd_plan_destroy(plan)
        return r
    assert supported(poynting=1) == 1
    assert supported(poynting=1, shape=(4, 4, 512)) == 1
    assert supported(poynting=1, shape=(4, 4, 513)) == 1
    assert supported(poynting=1, shape=(4, 4, 5000)) == 1
    assert supported(poynting=1, shape=(4, 4, 512), bloch_mask=1) == 1
    assert supported(poynting=1, shape=(4, 4, 513), bloch_mask=1) == 0
    assert supported(poynting=0) == 1        # Poynting off: the pass without S
    assert supported(poynting=2) == 0
    assert supported(poynting=0, bloch_mask=1) == 0
    # a Bloch pair runs the pass in FP32 on uniform grids
    assert supported(poynting=1, bloch_mask=1) == 1
    assert supported(poynting=1, bloch_mask=3, bloch_mode=1) == 1
    assert supported(poynting=1, bloch_mask=1, dtype=1) == 0
    assert supported(poynting=2, bloch_mask=1) == 0
    assert lib.fdtd_update_he_supported(None) == 0


def test_update_he_rejects_bad_calls():
    """fdtd_update_he validates before launching anything (fake device
    pointers: no kernel runs on these paths)."""
    from opencl_fdtd_amd import _lib
    lib = _lib.load()

    def call(poynting, e_out_same=False, h_out_same=False):
        d = _desc((10, 11, 12), [(0, 0, 2)], poynting=poynting)
        plan = ctypes.c_void_p()
        assert lib.fdtd_plan_create(ctypes.byref(d), ctypes.byref(plan)) == 0
        f = _lib.Fields()
        f.E, f.H, f.eps, f.S, f.oS = 0x1000, 0x2000, 0x3000, 0x4000, 0x5000
        f.H_out = f.H if h_out_same else 0x6000
        f.psi_e, f.psi_h, f.psi_h_out, f.coef = 0x7000, 0x8000, 0x9000, 0xa000
        f.E_out = f.E if e_out_same else 0xb000
        rc = lib.fdtd_update_he(plan, ctypes.byref(f), -1, None, None)
        msg = lib.fdtd_last_error().decode()
        lib.fdtd_plan_destroy(plan)
        return rc, msg
    rc, msg = call(poynting=2)
    assert rc != 0 and "Poynting" in msg
    rc, msg = call(poynting=1, e_out_same=True)
    assert rc != 0 and "E_out" in msg
    rc, msg = call(poynting=1, h_out_same=True)
    assert rc != 0 and "H_out" in msg
    assert lib.fdtd_update_he(None, None, -1, None, None) != 0


def test_update_he2_support_rules():
    """The two-step pass needs FP32, fused Poynting or Poynting off, no
    Bloch boundaries, a coefficient block that fits its LDS copy and the
    CPML slabs of one axis apart; rows of any length (plan-level query, no
    GPU)."""
    from opencl_fdtd_amd import _lib
    lib = _lib.load()

    def supported(shape=(10, 11, 12), faces=((0, 0, 2),), **kw):
        d = _desc(shape, list(faces), **kw)
        plan = ctypes.c_void_p()
        assert lib.fdtd_plan_create(ctypes.byref(d), ctypes.byref(plan)) == 0
        r = lib.fdtd_update_he2_supported(plan)
        lib.fdtd_plan_destroy(plan)
        return r
    six = [(a, p, 3) for a in range(3) for p in range(2)]
    assert supported() == 1
    assert supported(faces=six) == 1
    assert supported(faces=()) == 1
    assert supported(shape=(4, 4, 256)) == 1
    assert supported(shape=(4, 4, 257)) == 1
    assert supported(shape=(4, 4, 5000)) == 1
    assert supported(dtype=1) == 0
    assert supported(poynting=0) == 1
    assert supported(poynting=2) == 0
    assert supported(bloch_mask=2) == 0
    assert supported(shape=(6, 7, 5), faces=six) == 0          # both z slabs (3 of 5 cells) overlap
    assert supported(shape=(6, 7, 6), faces=six) == 1          # 3 + 3 of 6 cells: apart
    assert supported(shape=(8, 9, 200), faces=[(a, p, 120) for a in range(2, 3) for p in range(2)]) == 0
    assert lib.fdtd_update_he2_supported(None) == 0


def test_update_he2_preference():
    """step_n takes the two-step pass where its time model beats the
    one-step pass: rows of 73 cells (11 of a tile's 14 rows owned), rows of
    250 and 300 (12 of 15 rows of 64-cell z segments), but not a grid so
    small that every x chunk is one plane (6 x 7 x 700: the two-step pass
    would march three dry planes per chunk against one)."""
    from opencl_fdtd_amd import _lib
    lib = _lib.load()

    def pref(shape, t):
        d = _desc(shape, [(a, p, t) for a in range(3) for p in range(2)])
        plan = ctypes.c_void_p()
        assert lib.fdtd_plan_create(ctypes.byref(d), ctypes.byref(plan)) == 0
        r = (lib.fdtd_update_he2_supported(plan), lib.fdtd_update_he2_preferred(plan))
        lib.fdtd_plan_destroy(plan)
        return r
    assert pref((395, 345, 73), 8) == (1, 1)
    assert pref((64, 64, 250), 8) == (1, 1)
    assert pref((64, 64, 300), 8) == (1, 1)
    assert pref((6, 7, 700), 2) == (1, 0)
    assert lib.fdtd_update_he2_preferred(None) == 0


def test_update_he2_rejects_bad_calls():
    """fdtd_update_he2 validates before launching anything (fake device
    pointers: no kernel runs on these paths)."""
    from opencl_fdtd_amd import _lib
    lib = _lib.load()

    def call(dtype=0, psi_e_out_same=False, e_out_same=False):
        d = _desc((10, 11, 12), [(0, 0, 2)], dtype=dtype)
        plan = ctypes.c_void_p()
        assert lib.fdtd_plan_create(ctypes.byref(d), ctypes.byref(plan)) == 0
        f = _lib.Fields()
        f.E, f.H, f.eps, f.S, f.oS = 0x1000, 0x2000, 0x3000, 0x4000, 0x5000
        f.H_out = 0x6000
        f.psi_e, f.psi_h, f.psi_h_out, f.coef = 0x7000, 0x8000, 0x9000, 0xa000
        f.E_out = f.E if e_out_same else 0xb000
        f.psi_e_out = f.psi_e if psi_e_out_same else 0xc000
        rc = lib.fdtd_update_he2(plan, ctypes.byref(f), -1, None, None)
        msg = lib.fdtd_last_error().decode()
        lib.fdtd_plan_destroy(plan)
        return rc, msg
    rc, msg = call(dtype=1)
    assert rc != 0 and "FP32" in msg
    rc, msg = call(psi_e_out_same=True)
    assert rc != 0 and "psi_e_out" in msg
    rc, msg = call(e_out_same=True)
    assert rc != 0 and "fdtd_update_he2 needs E_out" in msg
    assert lib.fdtd_update_he2(None, None, -1, None, None) != 0


def test_adjoint_gradient_box_is_validated():
    """fdtd_adjoint_step rejects a gradient box that is empty or leaves the
    grid before launching anything (fake device pointers)."""
    from opencl_fdtd_amd import _lib
    lib = _lib.load()
    d = _desc((10, 11, 12), [(0, 0, 2)], poynting=1)
    plan = ctypes.c_void_p()
    assert lib.fdtd_plan_create(ctypes.byref(d), ctypes.byref(plan)) == 0
    try:
        for lo, hi in (((0, 0, 0), (11, 11, 12)), ((2, 3, 4), (2, 5, 6)), ((-1, 0, 0), (4, 4, 4)),
                       ((0, 0, 0), (4, 12, 4))):
            b = _lib.AdjointBufs()
            for name in ("E_prev", "E_next", "H_prev", "H_next", "eps", "Eb", "Hb", "psib_e", "psib_h", "coef"):
                setattr(b, name, 0x1000)
            b.psib_e_out, b.psib_h_out, b.eps_grad = 0x2000, 0x3000, 0x4000
            b.src_index, b.obj_axis = -1, -1
            b.grad_lo[:], b.grad_hi[:] = lo, hi
            assert lib.fdtd_adjoint_step(plan, ctypes.byref(b), None) != 0, (lo, hi)
            assert "gradient box" in lib.fdtd_last_error().decode()
    finally:
        lib.fdtd_plan_destroy(plan)
