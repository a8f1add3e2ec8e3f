"""End to end, line configuration (-l 1): the phase_field drop-in against the
CPU oracle driven in the reference's line-mode order
(3DgrainMPI/APTPhaseField.cu:655-697, 820-922): frame-0 line QoIs, odd step,
getLineQoIs sampling, the nts-frame stop, moveDomain every 20 steps (tip scan
of the plane sums, one-plane shifts with the alpha archive and loss areas,
then the tip height of the window), field saves of the full-height alpha, even
step; then QOILine::searchJunctionsOnImage on the archived alpha
(3DgrainMPI/QOI.cpp:103-160).

The case: a 24^3 window (fnz 27, shift threshold plane 21) of a 48-plane
column, the interface near plane 19.5, so the first moveDomain (step 20)
shifts the window.  top = -2 um puts every sampling threshold below the
window's height once it has moved, so frames 1..3 are taken at steps 21, 23
and 25 and the run stops there; -d 2 saves the field at step 24.
"""
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
TIPP = 20  # moving-domain cadence (3DgrainMPI/APTPhaseField.cu:29)


def _read_params(path):
    out = {}
    for line in open(path):
        k, v = line.split()
        out[k] = float(v)
    return out


def _oracle_line_run(P, phi0, alpha0, x, y, z, R9, num_nodes, save_fields):
    f32 = np.float32
    fnx, fny, fnz, fnz_f = (int(P[k]) for k in ("fnx", "fny", "fnz", "fnz_f"))
    ng, nts, Mt = int(P["num_theta"]), int(P["nts"]), int(P["Mt"])
    L = fnx * fny * fnz
    params = dict(fnx=fnx, fny=fny, fnz=fnz, K=8, num_theta=ng, bcx=1, bcy=1, bcz=0, thermal_type=1,
                  gather27=0, dt=f32(P["dt"]), hi=f32(P["hi"]), lamd=f32(P["lamd"]), L_cp=f32(P["L_cp"]),
                  kin_delta=f32(P["kin_delta"]), eps=f32(1e-8), G=f32(P["G"]), R=f32(P["R"]),
                  uc=f32(P["underCoolingRate"]), uc0=f32(P["underCoolingRate0"]), z0=f32(P["z0"]),
                  lyd=f32(P["lyd"]), lzd=f32(P["lzd"]), tau0=f32(P["tau0"]), sqrt2=f32(P["sqrt2"]),
                  dx=f32(P["dx"]))
    sim = oracle.OracleSim(params, x, y, z, R9, "f32")
    sim.init_from_phi_alpha(phi0, alpha0)
    tip_y = np.zeros(nts + 1, np.int32)
    total = np.zeros((nts + 1) * ng, np.int32)
    extra = np.zeros((nts + 1) * ng, np.int32)
    loss = np.zeros(ng, np.int32)

    def line_qois(alpha, kt, cur_tip, move_count):
        tip, t, e = qoi.line_qois(alpha, fnx, fny, fnz, cur_tip, ng)
        tip_y[kt] = tip + move_count
        total[kt * ng:(kt + 1) * ng] += t.astype(np.int32) + loss
        extra[kt * ng:(kt + 1) * ng] += e.astype(np.int32)
        return tip

    _, alpha = sim.collect(clear_alpha=True, bc=False)
    sim.md_enable(fnz_f)
    move_count, cur_tip, tip_front, samples, lowsl = 0, 1, 1, 0, 1
    tip_thres = int(0.8 * fnz)
    cur_tip = line_qois(alpha, 0, cur_tip, move_count)
    fieldkts = Mt // save_fields
    alpha_full = np.zeros(fnx * fny * fnz_f, np.int32)
    fields = {}
    W0, dx, z0, top = f32(P["W0"]), f32(P["dx"]), f32(P["z0"]), f32(P["top"])
    kt = 0
    for kt in range(Mt // 2 + 1000):
        sim.step(1, 2 * kt + 1)
        # getLineQoIs (float arithmetic of :657-658)
        loc = f32(f32(move_count + lowsl - 1) * W0) * dx
        threshold = z0 + f32(top * f32(samples + 1)) / f32(nts)
        if loc > threshold:
            _, alpha = sim.collect(clear_alpha=True, bc=True)
            loss[:] = sim.loss
            line_qois(alpha, samples + 1, 0, move_count)
            samples += 1
        if samples == nts:
            break
        if (2 * kt + 2) % TIPP == 0:  # moveDomain
            tip_front = sim.md_tip(tip_front)
            while tip_front >= tip_thres:
                sim.md_shift(move_count)
                move_count += 1
                tip_front -= 1
            _, alpha = sim.collect(clear_alpha=False, bc=True)
            lowsl = qoi.line_qois(alpha, fnx, fny, fnz, 1, ng)[0]
        if (2 * kt + 2) % fieldkts == 0:
            _, alpha = sim.collect(clear_alpha=True, bc=True)
            alpha_full = sim.alpha_full.copy()
            off = move_count * fnx * fny
            n = min(L, alpha_full.size - off)
            alpha_full[off:off + n] = alpha[:n]
            fields[2 * kt + 2] = alpha_full.copy()
        sim.step(1, 2 * kt + 2)
    cross, nodes = qoi.line_junctions(alpha_full, fnx, fny, tip_y, num_nodes)
    return dict(tip_y=tip_y, total=total, extra=extra, cross=cross, nodes=nodes, fields=fields,
                Mt=2 * kt, move_count=move_count, samples=samples, z=sim.z.copy())


def test_line_config_cli_matches_oracle(tmp_path):
    d = tmp_path
    forcing.write_line_case(str(d), 0, n=(24, 24, 24), grains=40, G=1.0, R=0.1, z0=1.48, t_end=2.5e-7,
                            nts=3, line_config=True, nz_full=48, top=-2.0)
    out = d / "out"
    out.mkdir()
    r = subprocess.run([CLI, "line.py", "-s", "0", "-b", "110", "-f", str(d / "forcing" / "case"),
                        "-o", str(out), "-D", str(d / "dump"), "-k", "8", "-l", "1", "-d", "2"],
                       cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    P = _read_params(d / "dump" / "params.txt")
    fnx, fny, fnz = int(P["fnx"]), int(P["fny"]), int(P["fnz"])
    phi0 = np.fromfile(d / "dump" / "init_phi.bin", np.float32)
    alpha0 = np.fromfile(d / "dump" / "init_alpha.bin", np.int32)
    coords = np.fromfile(d / "dump" / "coords.bin", np.float32)
    x, y, z = coords[:fnx], coords[fnx:fnx + fny], coords[fnx + fny:]
    F = d / "forcing" / "case0"
    ng = int(P["num_theta"])
    num_nodes = int(np.loadtxt(F / "NN.txt"))
    R9 = capi.orientations_from_angles(np.loadtxt(F / "theta.txt").astype(np.float32), ng)
    want = _oracle_line_run(P, phi0, alpha0, x, y, z, R9, num_nodes, save_fields=2)
    print(f"[line] oracle: Mt {want['Mt']}, shifts {want['move_count']}, frames {want['samples']}, "
          f"fields at {sorted(want['fields'])}", flush=True)
    # the case exercises what it is meant to
    assert want["move_count"] > 0 and want["samples"] == int(P["nts"]) and want["fields"]
    assert (want["total"][ng:] > 0).any() and (want["nodes"] != -1).any()

    qfile = [f for f in glob.glob(str(out / "*_rank0.h5")) if "_time" not in f]
    assert len(qfile) == 1, os.listdir(out)
    assert f"_Mt{want['Mt']}_" in os.path.basename(qfile[0])
    assert h5.read(qfile[0], "tip_y", np.int32).tolist() == want["tip_y"].tolist()
    assert np.array_equal(h5.read(qfile[0], "total_area", np.int32), want["total"])
    assert np.array_equal(h5.read(qfile[0], "extra_area", np.int32), want["extra"])
    assert np.array_equal(h5.read(qfile[0], "cross_sec", np.int32), want["cross"])
    assert np.array_equal(h5.read(qfile[0], "node_region", np.int32), want["nodes"])
    for step, alpha_full in want["fields"].items():
        ffile = glob.glob(str(out / f"*_time{step}.h5"))
        assert len(ffile) == 1, (step, os.listdir(out))
        assert np.array_equal(h5.read(ffile[0], "alpha", np.int32), alpha_full)
