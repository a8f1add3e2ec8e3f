"""End to end: the phase_field drop-in (C++ host over the HIP library) against
the CPU oracle driven in the reference's loop order from the same initial
fields (3DgrainMPI/APTPhaseField.cu:820-922): QoI volumes, junctions and the
saved alpha field must agree exactly."""
import glob
import os
import subprocess

import numpy as np
import pytest

import oracle
from cupf_amd import capi, forcing, h5
from oracle import qoi

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cupf_amd", "bin", "phase_field")


def _read_params(path):
    out = {}
    for line in open(path):
        k, v = line.split()
        out[k] = float(v)
    return out


def test_phase_field_cli_matches_oracle(tmp_path):
    d = tmp_path
    forcing.write_line_case(str(d), 0, n=(20, 20, 20), grains=8, G=1.0, R=0.1, z0=0.8, t_end=2.5e-7)
    out = d / "out"
    out.mkdir()
    r = subprocess.run([CLI, "line.py", "-s", "0", "-b", "110", "-f", str(d / "forcing" / "case"),
                        "-o", str(out), "-D", str(d / "dump"), "-k", "8"], cwd=d,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    P = _read_params(d / "dump" / "params.txt")
    fn = int(P["fnx"])
    L = fn ** 3
    phi0 = np.fromfile(d / "dump" / "init_phi.bin", np.float32)
    alpha0 = np.fromfile(d / "dump" / "init_alpha.bin", np.int32)
    coords = np.fromfile(d / "dump" / "coords.bin", np.float32)
    x, y, z = coords[:fn], coords[fn:2 * fn], coords[2 * fn:]
    F = d / "forcing" / "case0"
    ng = int(P["num_theta"])
    theta = np.loadtxt(F / "theta.txt").astype(np.float32)
    R9 = capi.orientations_from_angles(theta, ng)
    f32 = np.float32
    params = dict(fnx=fn, fny=fn, fnz=fn, K=8, num_theta=ng, bcx=1, bcy=1, bcz=0, thermal_type=1,
                  gather27=0, dt=f32(P["dt"]), hi=f32(P["hi"]), lamd=f32(P["lamd"]), L_cp=f32(P["L_cp"]),
                  kin_delta=f32(P["kin_delta"]), eps=f32(1e-8), G=f32(P["G"]), R=f32(P["R"]),
                  uc=f32(P["underCoolingRate"]), uc0=f32(P["underCoolingRate0"]), z0=f32(P["z0"]),
                  lyd=f32(P["lyd"]), lzd=f32(P["lzd"]), tau0=f32(P["tau0"]), sqrt2=f32(P["sqrt2"]),
                  dx=f32(P["dx"]))
    sim = oracle.OracleSim(params, x, y, z, R9, "f32")
    sim.init_from_phi_alpha(phi0, alpha0)
    sim.collect(clear_alpha=True, bc=False)
    Mt = int(P["Mt"])
    qoikts, fieldkts = Mt // int(P["nts"]), Mt
    vol = alpha_field = None
    for kt in range(Mt // 2):
        sim.step(1, 2 * kt + 1)
        if (2 * kt + 2) % qoikts == 0:
            _, a = sim.collect(clear_alpha=False, bc=True)
            vol = qoi.volume(a, fn, fn, fn)
        if (2 * kt + 2) % fieldkts == 0:
            _, alpha_field = sim.collect(clear_alpha=True, bc=True)
        sim.step(1, 2 * kt + 2)

    qfile = [f for f in glob.glob(str(out / "*_rank0.h5")) if "_time" not in f]
    ffile = glob.glob(str(out / f"*_time{Mt}.h5"))
    assert len(qfile) == 1 and len(ffile) == 1, os.listdir(out)
    assert np.array_equal(h5.read(ffile[0], "alpha", np.int32), alpha_field)
    assert h5.read(qfile[0], "grainToPF", np.int32).tolist() == vol[0]
    assert h5.read(qfile[0], "volume", np.int32).tolist() == vol[1]
    nodes = h5.read(qfile[0], "node_region", np.int32).reshape(-1, 9)
    nodes = [tuple(r) for r in nodes if r[0] != -1]
    want = qoi.junctions3d(alpha_field, fn, fn, fn)
    got = [(int(r[0]), int(r[1]), int(r[2]), int(r[3]), tuple(int(v) for v in r[4:8])) for r in nodes]
    assert got == want
    assert np.allclose(h5.read(qfile[0], "x_coordinates"), x)
