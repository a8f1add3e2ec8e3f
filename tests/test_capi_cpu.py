"""C ABI checks that need no GPU: exports, argument errors, host-side helpers,
and the derived constants of the input path."""
import ctypes
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from cupf_amd import capi, cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    syms = capi.exported_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    for s in syms:
        assert f" T {s}" in out, s


@pytest.mark.parametrize("field,value,msg", [
    ("K", 7, "K must be 5 or 8"),
    ("fnx", 3, "grid too small"),
    ("num_theta", 0, "num_theta"),
    ("bcx", 2, "bc digits"),
    ("thermal_type", 9, "thermal_type"),
])
def test_create_rejects_bad_parameters(field, value, msg):
    c = cases.make_line_case("bad", n=(8, 8, 8), grains=3, K=5, seed=1, z0_cells=4)
    p = dict(c.params)
    p[field] = value
    lib = capi.load()
    h = ctypes.c_void_p()
    rc = lib.pf_create(ctypes.byref(capi.PFParams.from_dict(p)), 0, ctypes.byref(h))
    assert rc == -1
    assert msg in lib.pf_last_error().decode()
    assert not h.value


def test_create_rejects_32bit_overflow():
    p = cases.make_line_case("big", n=(8, 8, 8), grains=3, K=8, seed=1, z0_cells=4).params
    p = dict(p, fnx=2000, fny=2000, fnz=200)
    lib = capi.load()
    h = ctypes.c_void_p()
    assert lib.pf_create(ctypes.byref(capi.PFParams.from_dict(p)), 0, ctypes.byref(h)) == -1
    assert "2^31" in lib.pf_last_error().decode()


def test_null_context_is_an_error():
    lib = capi.load()
    assert lib.pf_step(None, 1, 1, 0) == -1
    assert "null context" in lib.pf_last_error().decode()


def test_orientations_from_angles_equals_reference_angle_form():
    """Rows built on the host from theta reproduce kine_ani's angle form exactly."""
    th = cases.reference_angles(30, 4)
    R9 = capi.orientations_from_angles(th, 30)
    ng = 30
    lib = oracle.load("f32")
    P = oracle.OrcParams.from_dict(dict(kin_delta=np.float32(0.11), eps=np.float32(1e-8)))
    rng = np.random.default_rng(2)
    f32 = np.float32
    libm = ctypes.CDLL("libm.so.6")
    for fn in (libm.sinf, libm.cosf):
        fn.argtypes = [ctypes.c_float]
        fn.restype = ctypes.c_float
    for g in range(ng + 1):
        # host libm, as 3DgrainMPI/PhaseFieldCpu.cpp:383-384
        sa, ca = f32(libm.sinf(float(th[g]))), f32(libm.cosf(float(th[g])))
        sb, cb = f32(libm.sinf(float(th[g + ng]))), f32(libm.cosf(float(th[g + ng])))
        assert R9[g, 0] == f32(ca * cb) and R9[g, 2] == -sb and R9[g, 3] == -sa and R9[g, 7] == f32(sa * sb)
        u = rng.standard_normal(3)
        got = lib.orc_kine_ani(P, *u, R9[g].ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        a_s = 1 + 3 * 0.11
        assert 0.85 * a_s * (1 - 4 * 0.11 / a_s) < got < 1.2 * a_s


def test_derived_constants_match_line_py_values():
    """3DgrainMPI/line.py physics with R = 0.1 m/s: cfl doubled (R < 0.3)."""
    d = cases.derive(dict(cases.LINE_PY, R=0.1, G=1.0, UC=0.1, z0=2.0, Lx=5.12, asp_ratio_zx=1.0),
                     line_config=False)
    assert d["cfl"] == np.float32(3.0)
    assert d["dt"] == pytest.approx(0.0363668, rel=1e-5)
    assert d["tau0"] == pytest.approx(1.32804e-7, rel=1e-5)
    assert d["lamd"] == pytest.approx(58.3323, rel=1e-5)
    assert d["hi"] == np.float32(1.25)
    assert (d["nx"], d["ny"], d["nz"]) == (64, 64, 64)
    assert (d["fnx"], d["fny"], d["fnz"]) == (67, 67, 67)
    assert d["pts_cell"] == 4


def test_voronoi_substrate_is_seeded_and_complete():
    a = cases.voronoi_substrate(65, 65, 100, 0)
    b = cases.voronoi_substrate(65, 65, 100, 0)
    c = cases.voronoi_substrate(65, 65, 100, 1)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert set(np.unique(a)) == set(range(1, a.max() + 1))
    assert 95 <= a.max() <= 100


def test_quaternion_rotations_are_orthonormal():
    R = cases.quaternion_rotations(50, 3).astype(np.float64).reshape(-1, 3, 3)
    for r in R:
        assert np.allclose(r @ r.T, np.eye(3), atol=1e-6)
        assert np.linalg.det(r) == pytest.approx(1.0, abs=1e-6)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
