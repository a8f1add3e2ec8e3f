"""The phase_field drop-in's input path on the CPU (no device work: -P):
parsing of the input file and forcing folder, derived constants, grid and
initial fields."""
import os
import subprocess

import numpy as np
import pytest

from cupf_amd import cases, forcing

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cupf_amd", "bin", "phase_field")


def _params(path):
    out = {}
    for line in open(path):
        k, v = line.split()
        out[k] = float(v)
    return out


@pytest.fixture(scope="module")
def case_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("fc")
    forcing.write_line_case(str(d), 0, n=(16, 16, 16), grains=6, G=1.0, R=0.1, z0=0.6, t_end=2.5e-7)
    r = subprocess.run([CLI, "line.py", "-s", "0", "-b", "110", "-f", str(d / "forcing" / "case"),
                        "-o", str(d), "-D", str(d / "dump"), "-P"], cwd=d, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return d


def test_derived_constants_match_float32_derivation(case_dir):
    got = _params(case_dir / "dump" / "params.txt")
    inp = dict(cases.LINE_PY, Lx=1.28, asp_ratio_yx=1.0, asp_ratio_zx=1.0, moving_ratio=1.0, R=0.1, G=1.0)
    d = cases.derive(inp, line_config=False)
    for k in ("dt", "tau0", "lamd", "hi", "cfl"):
        assert np.float32(got[k]) == d[k], k
    for k in ("nx", "ny", "nz", "fnx", "fny", "fnz", "pts_cell"):
        assert int(got[k]) == d[k], k
    assert int(got["Mt"]) == 48 and int(got["thermalType"]) == 1
    assert (int(got["bcX"]), int(got["bcY"]), int(got["bcZ"])) == (1, 1, 0)


def test_initial_fields_match_initfield_restatement(case_dir):
    """psi trilinear from the macro grid, /W0, phi = tanh(psi/sqrt2), alpha from
    alpha.txt where phi > LS (3DgrainMPI/PhaseFieldCpu.cpp:394-449)."""
    fn = 19
    phi = np.fromfile(case_dir / "dump" / "init_phi.bin", np.float32).reshape(fn, fn, fn)
    alpha = np.fromfile(case_dir / "dump" / "init_alpha.bin", np.int32).reshape(fn, fn, fn)
    F = case_dir / "forcing" / "case0"
    f32 = np.float32
    X, Y, Z = (np.loadtxt(F / f"{a}.txt").astype(f32) for a in "xyz")
    Nx, Ny, Nz = X.size, Y.size, Z.size
    psi_mac = np.loadtxt(F / "psi.txt").astype(f32)
    c = (np.arange(fn) - 1).astype(f32) * f32(f32(0.8) * f32(0.1))
    Dx, Dy, Dz = X[-1] - X[-2], Y[-1] - Y[-2], Z[-1] - Z[-2]   # last spacing, as the reference
    want = np.zeros((fn, fn, fn), f32)
    for k in range(1, fn - 1):
        for j in (3, 8, 15):
            for i in (2, 8, 16):
                q = [(c[i] - X[0]) / Dx, (c[j] - Y[0]) / Dy, (c[k] - Z[0]) / Dz]
                kk = [int(v) for v in q]
                dd = [f32(v - f32(w)) for v, w in zip(q, kk)]
                for a, N in enumerate((Nx, Ny, Nz)):
                    if kk[a] == N - 1:
                        kk[a], dd[a] = N - 2, f32(1.0)
                o = kk[0] + kk[1] * Nx + kk[2] * Nx * Ny
                on = o + Nx * Ny
                one = f32(1)
                dx, dy, dz = dd
                u = psi_mac
                v = (((one - dx) * (one - dy) * u[o] + (one - dx) * dy * u[o + Nx]) + dx * (one - dy) * u[o + 1]
                     + dx * dy * u[o + Nx + 1]) * (one - dz) + \
                    (((one - dx) * (one - dy) * u[on] + (one - dx) * dy * u[on + Nx]) + dx * (one - dy) * u[on + 1]
                     + dx * dy * u[on + Nx + 1]) * dz
                want[k, j, i] = np.tanh(f32(f32(v / f32(0.1)) / f32(np.sqrt(2.0))))
                assert abs(float(phi[k, j, i]) - float(want[k, j, i])) < 4e-7, (k, j, i)
    assert (phi[0] == 0).all() and (phi[:, 0] == 0).all()       # halo zero
    img = np.loadtxt(F / "alpha.txt").astype(np.int32).reshape(17, 17)
    solid = phi[1:-1, 1:-1, 1:-1].astype(np.float64) > -0.995
    exp = np.where(solid, img[None], 0)
    assert np.array_equal(alpha[1:-1, 1:-1, 1:-1], exp)
    assert alpha.max() == 6
