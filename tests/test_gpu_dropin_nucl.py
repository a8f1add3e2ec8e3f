"""End to end, nucleation (-n 1, liquid start with preMt > 0): the phase_field
drop-in against the CPU oracle driven in the reference's order
(3DgrainMPI/APTPhaseField.cu:779-818 and 852-922): the preMt phase (uniform
cooling at the initial undercooling rate, nucleation after every odd step),
the collect and the re-initialisation clipped to the initial phi, then the
main loop with nucleation from the macro temperature T_3D (interp4 at t and
t + dt) after every odd step, QoIs (per-grain volumes of nucleated ids, 3-D
junctions) and the saved field.

The case: 32^3 with the bottom 10 planes solid at the start (grain id 0: no
substrate with -n 1), nuc_rad 0.25 um (7^3 blocks, 64 of them), Nmax 10 so
that the preMt phase and the main loop both nucleate; the macro temperature
cools uniformly from -1.5 K to -3.5 K over the run.  Decisions draw from the
same counter-based generator on both sides, so events and fields agree
exactly (a draw within an ulp of the probability could differ between device
and host expf; the case has none).
"""
import glob
import os
import re
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


def test_nucleation_cli_matches_oracle(tmp_path):
    d = tmp_path
    nt = 11
    forcing.write_line_case(str(d), 0, n=(32, 32, 32), grains=12, G=1.0, R=0.1, z0=0.8, t_end=2.5e-7,
                            nts=2, inp_overrides=dict(nuc_rad=0.25, preMt=20), nmax=10.0,
                            temp=np.linspace(-1.5, -3.5, nt), macro=(13, 13, 23, nt))
    out = d / "out"
    out.mkdir()
    r = subprocess.run([CLI, "line.py", "-s", "0", "-b", "110", "-f", str(d / "forcing" / "case"),
                        "-o", str(out), "-D", str(d / "dump"), "-k", "8", "-n", "1", "-d", "1"],
                       cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    m = re.search(r"nucleation events: (\d+)", r.stdout)
    assert m, r.stdout[-2000:]
    gpu_events = int(m.group(1))

    P = _read_params(d / "dump" / "params.txt")
    fnx, fny, fnz = int(P["fnx"]), int(P["fny"]), int(P["fnz"])
    phi0 = np.fromfile(d / "dump" / "init_phi.bin", np.float32)
    alpha0 = np.fromfile(d / "dump" / "init_alpha.bin", np.int32)
    coords = np.fromfile(d / "dump" / "coords.bin", np.float32)
    x, y, z = coords[:fnx], coords[fnx:fnx + fny], coords[fnx + fny:]
    F = d / "forcing" / "case0"
    ng = int(P["num_theta"])
    R9 = capi.orientations_from_angles(np.loadtxt(F / "theta.txt").astype(np.float32), ng)
    T3d = h5.read(str(F / "Temp.h5"), "Temp").astype(np.float32)
    f32 = np.float32
    params = dict(fnx=fnx, fny=fny, fnz=fnz, K=8, num_theta=ng, bcx=1, bcy=1, bcz=0, thermal_type=1,
                  gather27=0, eps=f32(1e-8))
    for k in ("dt", "hi", "lamd", "L_cp", "kin_delta", "G", "R", "z0", "lyd", "lzd", "tau0", "sqrt2",
              "dx", "nuc_rad", "undcool_mean", "undcool_std", "nuc_Nmax", "nuc_Nmax0", "X0_mac",
              "Y0_mac", "Z0_mac", "t0_mac", "Dx_mac", "Dt_mac"):
        params[k] = f32(P[k])
    params.update(uc=f32(P["underCoolingRate"]), uc0=f32(P["underCoolingRate0"]),
                  pts_cell=int(P["pts_cell"]), Nx=int(P["Nx"]), Ny=int(P["Ny"]), Nz=int(P["Nz"]),
                  Nt=int(P["Nt"]))
    assert params["pts_cell"] == 3
    seed = int(P["seed_val"])
    sim = oracle.OracleSim(params, x, y, z, R9, "f32", T3d=T3d)
    sim.init_from_phi_alpha(phi0, alpha0)
    sim.nucl_reset(1)
    preMt = int(P["preMt"])
    sim.nucl_reset(0)
    for kt in range(preMt // 2):
        sim.step(2, 2 * kt + 1, use_initial_uc=True, nucleate=True, seed=seed)
    pre_events = sim.nucl_events
    sim.collect(clear_alpha=True, bc=False)
    sim.reinit_from_collect()
    sim.collect(clear_alpha=True, bc=False)
    Mt, nts = int(P["Mt"]), int(P["nts"])
    qoikts, fieldkts = Mt // nts, Mt
    vol = field = None
    for kt in range(Mt // 2):
        sim.step(1, 2 * kt + 1, nucleate=True, seed=seed)
        if (2 * kt + 2) % qoikts == 0:
            _, a = sim.collect(clear_alpha=False, bc=True)
            vol = qoi.volume(a, fnx, fny, fnz)
        if (2 * kt + 2) % fieldkts == 0:
            _, field = sim.collect(clear_alpha=True, bc=True)
        sim.step(1, 2 * kt + 2)
    print(f"[nucl] events: preMt {pre_events}, total {sim.nucl_events}, drop-in {gpu_events}; "
          f"grains at the end {len(vol[0])}", flush=True)
    # the case nucleates in both phases, and nucleated ids exceed num_theta
    assert pre_events > 0 and sim.nucl_events > pre_events
    assert max(vol[0]) > ng
    assert gpu_events == sim.nucl_events

    qfile = [f for f in glob.glob(str(out / "*_rank0.h5")) if "_time" not in f]
    ffile = glob.glob(str(out / f"*_time{Mt}.h5"))
    assert len(qfile) == 1 and len(ffile) == 1, os.listdir(out)
    assert np.array_equal(h5.read(ffile[0], "alpha", np.int32), field)
    assert h5.read(qfile[0], "grainToPF", np.int32).tolist() == vol[0]
    assert h5.read(qfile[0], "volume", np.int32).tolist() == vol[1]
    nodes = h5.read(qfile[0], "node_region", np.int32).reshape(-1, 9)
    got = [(int(r[0]), int(r[1]), int(r[2]), int(r[3]), tuple(int(v) for v in r[4:8]))
           for r in nodes if r[0] != -1]
    want = qoi.junctions3d(field, fnx, fny, fnz)
    assert got == want[:len(got)] and (len(got) == len(want) or len(got) == nodes.shape[0])
