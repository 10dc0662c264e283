"""Host-side checks of bench.py: the algorithmic bytes per launch of each
timed kernel, the traffic figure read from the committed rocprofv3 summary,
and the one-GPU guard.  No GPU is used."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

SHAPE = (395, 345, 73)
PMLS = [{"axis": a, "polarity": p, "thickness": 8} for a in "xyz" for p in "np"]


def _slab_faces(shape, pmls):
    """Slab cells summed over the faces: thickness (at most the axis length)
    times the cells of a plane across that axis."""
    n = 0
    for p in pmls:
        r = "xyz".find(p["axis"])
        other = shape[0] * shape[1] * shape[2] // shape[r]
        n += min(p["thickness"], shape[r]) * other
    return n


def test_kernel_bytes_readme_grid():
    n = SHAPE[0] * SHAPE[1] * SHAPE[2]
    faces = _slab_faces(SHAPE, PMLS)
    # E: E r+w, H r, eps r = 12 words; psi r+w 4 words per slab cell and face
    assert bench.kernel_bytes(SHAPE, PMLS, "E") == 4 * (12 * n + 4 * faces)
    # H with S: E r, H r, H w, S w = 12 words; without S 9
    assert bench.kernel_bytes(SHAPE, PMLS, "H") == 4 * (12 * n + 4 * faces)
    assert bench.kernel_bytes(SHAPE, PMLS, "H", poynting=False) == 4 * (9 * n + 4 * faces)
    # fused pass and two-step pass: 18 words per cell, both psi blocks (8 words)
    he = bench.kernel_bytes(SHAPE, PMLS, "HE")
    assert he == 4 * (18 * n + 8 * faces) == 813692440
    assert bench.kernel_bytes(SHAPE, PMLS, "HE2") == he
    assert bench.kernel_bytes(SHAPE, PMLS, "E", dtype_bytes=8) == 2 * bench.kernel_bytes(SHAPE, PMLS, "E")


def test_kernel_bytes_thin_axis():
    """A slab thicker than its axis counts the axis length, not the thickness."""
    shape = (4, 20, 30)
    pmls = [{"axis": "x", "polarity": "n", "thickness": 8}]
    n = 4 * 20 * 30
    assert bench.kernel_bytes(shape, pmls, "E") == 4 * (12 * n + 4 * 4 * 20 * 30)


def test_measured_traffic_reads_the_summary():
    v = bench.measured_traffic("phc_slab_395x345x73", "he2_march")
    assert v is not None and 8.0e8 < v < 9.0e8
    assert bench.measured_traffic("phc_slab_395x345x73", "no_such_kernel") is None
    assert bench.measured_traffic("no_such_workload", "he2_march") is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
