"""Benchmark of the APT step on one MI355X.

Workload: the 256x256x512 line-configuration simulation (K=8 active slots,
~5000 seeded Voronoi grains, random quaternion orientations, linear thermal
field G=1e6 K/m, R=0.1 m/s, BC 110), state resident in HBM.  A step is one
APTrhs_psi time step over all (nx+1)(ny+1)(nz+1) updated points.

Prints one JSON line:
  value      Mcell-steps/s over the timed steps (--warmup W untimed, then
             --steps K timed; whole job, one GPU; every updated point of the
             domain counts, as every one is updated)
  roofline   per-kernel durations of the same window (the run is replayed
             from the initial state with HIP events around every launch on
             the context's stream): work-list plan, step kernel (in steps
             with heavy overflow together with the edge kernel that runs
             beside it on a second stream and joins before the fix-up),
             overflow fix-up.  The step kernel skips z planes of 64x4 column tiles
             whose stencil reaches only empty slots and whose destination is
             empty (their result is known without reading anything), so the
             algorithmic bytes of a step are those of the points it processes:
             processed points x K x 16 B (one read and one write of every
             slot's float32 phi and int32 id), counted exactly per step from
             the activity maps.  `kernel` names the kernel with the largest
             share of the window; `achieved` divides the step's algorithmic
             bytes by the summed per-launch duration of all three kernels
             (the step's points are finished only when the fix-up is), and the
             step kernel's own figure is given beside it.  `traffic` is the
             measured HBM bytes of one step of the same window (rocprof
             FETCH_SIZE + WRITE_SIZE of plan, step and fix-up kernels,
             profiles/traffic.json) or null if that window was not profiled.
  dense      the same window with plane skipping off (every point read and
             written): the step kernel's bytes (4.337 GB per step) over its
             duration, the kernel-quality figure that skipping cannot flatter
  full_run   the whole 10000-step simulation of this configuration (start-up
             transient included): ms per step and Mcell-steps/s, and per block
             of steps (ms/step, processed points, overflowed points)
  cpu_baseline  the oracle (C restatement, OpenMP on all host threads) on
             the state after the timed steps, float32 and float64 builds,
             for a bounded sample of steps (time-capped); the GPU's next steps
             from the same state are compared with both: arg mismatches and
             L-inf(phi) against float32, arg flips and L-inf(phi) against float64
  ensemble   64 seeded simulations, grids drawn from 32^3, 64^3, 96x96x128 and
             128x128x256, G log-uniform in [1e5, 1e7] K/m, R log-uniform in
             [0.01, 1] m/s, K=8, 2000 steps each: batched (one launch sequence
             per step for all of them, hipGraph replay, autotuned chunk depth)
             against back-to-back single runs on the same GPU; both stepping
             times, sims/h, and whether every simulation ended with the same bits
  dropin     the phase_field command (the reference's CLI surface) on the
             same configuration in line mode, 10000 steps: wall time, its
             evolve timer, output written
  serving    a seeded Poisson stream (4 requests/s) of the ensemble's
             distributions with 500..5000 steps each, continuously batched
             under an HBM budget: sims/h and p50/p99 latency from arrival to
             retirement
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(window: str):
    """Measured HBM bytes of one step of this window (rocprof FETCH_SIZE and
    WRITE_SIZE of the plan, step and fix-up kernels, scaled by the copy-kernel
    calibration) and the per-kernel split, from profiles/traffic.json; None
    for a window that was not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None, None
    try:
        with open(path) as f:
            w = json.load(f).get(window, {})
    except (OSError, ValueError):
        return None, None
    return w.get("hbm_bytes_per_step"), w.get("kernels")


def _bc_map(i: int, n: int, bc: int) -> int:
    if i == 0:
        return n - 2 if bc else 2
    if i == n - 1:
        return 1 if bc else n - 3
    return i


def processed_points(ctx, p) -> int:
    """Points in the tile planes the next step processes: those whose 3x3
    tile neighbourhood at k-1, k, k+1 (through the boundary maps) is active in
    the source buffer, or whose own tile plane is active in the destination."""
    fnx, fny, fnz = int(p["fnx"]), int(p["fny"]), int(p["fnz"])
    bcx, bcy, bcz = int(p["bcx"]), int(p["bcy"]), int(p["bcz"])
    src = ctx.activity(0, with_map=True)[2].astype(bool)
    dst = ctx.activity(1, with_map=True)[2].astype(bool)
    nz, nty, ntx = src.shape
    tx_of = lambda i: (i - 1) // 64
    ty_of = lambda j: (j - 1) // 4
    # tile-neighbour index lists through the boundary maps
    xs = [(tx_of(_bc_map(1 + 64 * t - 1, fnx, bcx)), t, tx_of(_bc_map(min(1 + 64 * t + 64, fnx - 1), fnx, bcx)))
          for t in range(ntx)]
    ys = [(ty_of(_bc_map(1 + 4 * t - 1, fny, bcy)), t, ty_of(_bc_map(min(1 + 4 * t + 4, fny - 1), fny, bcy)))
          for t in range(nty)]
    dil = np.zeros_like(src)
    for ty in range(nty):
        for tx in range(ntx):
            acc = np.zeros(nz, bool)
            for yy in ys[ty]:
                for xx in xs[tx]:
                    acc |= src[:, yy, xx]
            dil[:, ty, tx] = acc
    proc = np.zeros_like(src)
    for k in range(1, fnz - 1):
        proc[k] = dil[_bc_map(k - 1, fnz, bcz)] | dil[k] | dil[_bc_map(k + 1, fnz, bcz)] | dst[k]
    wx = np.array([min(64, fnx - 2 - 64 * t) for t in range(ntx)])
    wy = np.array([min(4, fny - 2 - 4 * t) for t in range(nty)])
    pts = np.outer(wy, wx)
    return int((proc[1:fnz - 1] * pts[None]).sum())


def ensemble_block(count: int, K: int) -> dict:
    from cupf_amd import ensemble
    specs = ensemble.ensemble_specs(count, seed=0, nsteps=2000)
    t0 = time.time()
    prepared = [ensemble.make_case(s, K) for s in specs]
    log(f"[bench] ensemble: {count} sims prepared in {time.time() - t0:.1f}s, "
        f"{sum(s.footprint(K) for s in specs) / 1e9:.1f} GB of device state")
    tb, kc, hb = ensemble.run_batched(specs, prepared, K=K, graph=True, autotune=True)
    log(f"[bench] ensemble batched: {tb:.2f} s (chunk depth {kc})")
    ta, each, ha = ensemble.run_back_to_back(specs, prepared, K=K)
    log(f"[bench] ensemble back-to-back: {ta:.2f} s; same bits: {ha == hb}")
    grids = {}
    for s in specs:
        grids["x".join(map(str, s.n))] = grids.get("x".join(map(str, s.n)), 0) + 1
    return {
        "workload": f"{count} sims x 2000 steps, K={K}",
        "grids": grids,
        "sims_per_h": round(count * 3600.0 / tb, 1),
        "unit": "sims/h",
        "batched_s": round(tb, 3),
        "back_to_back_s": round(ta, 3),
        "back_to_back_sims_per_h": round(count * 3600.0 / ta, 1),
        "speedup_vs_back_to_back": round(ta / tb, 3),
        "chunk_depth": kc,
        "identical_results": ha == hb,
        "note": "stepping time only (inputs prepared and contexts initialised beforehand, in both)",
    }


def serving_block(count: int, K: int) -> dict:
    from cupf_amd import ensemble
    specs = ensemble.request_stream(count, rate=4.0, seed=1)
    prepared = [ensemble.make_case(s, K) for s in specs]
    srv = ensemble.Server(hbm_budget=64 << 30, quantum=200, K=K)
    done, wall = srv.run(specs, prepared)
    lat = [d.latency for d in done]
    log(f"[bench] serving: {len(done)} requests in {wall:.2f} s, p50 {ensemble.percentile(lat, 50):.3f} s, "
        f"p99 {ensemble.percentile(lat, 99):.3f} s")
    return {
        "workload": f"{count} requests, Poisson 4/s, 500..5000 steps, ensemble grids, K={K}",
        "sims_per_h": round(len(done) * 3600.0 / wall, 1),
        "p50_latency_s": round(ensemble.percentile(lat, 50), 3),
        "p99_latency_s": round(ensemble.percentile(lat, 99), 3),
        "wall_s": round(wall, 3),
        "sim_steps": int(sum(s.nsteps for s in specs)),
        "hbm_budget_gb": 64,
        "note": "latency from arrival to retirement; arrivals span "
                f"{specs[-1].arrival:.1f} s, so sims/h is bounded by the arrival rate",
    }


def qoi_block(ctx, p) -> dict:
    """The device QoIs on the current state (median of 5 calls each, wall
    clock around the C-ABI call, which allocates its small result buffers,
    launches, copies the results back and synchronises), beside the download
    of the collected fields that a host-side QoI pass needs first."""
    from cupf_amd import capi

    def med(fn, reps=5):
        ts = []
        for _ in range(reps):
            ctx.sync()
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return sorted(ts)[reps // 2] * 1e3
    ms_collect = med(lambda: ctx.collect_on_device(0, capi.PF_COLLECT_CLEAR_ALPHA))
    ms_volume = med(lambda: ctx.qoi_volume(int(p["num_theta"])))
    ms_junc = med(ctx.qoi_junction_rows)
    ms_line = med(lambda: ctx.qoi_line(0))
    ms_host = med(lambda: ctx.collect(0, capi.PF_COLLECT_CLEAR_ALPHA), 3)
    ids, _, above = ctx.qoi_volume(int(p["num_theta"]))
    nj = len(ctx.qoi_junction_rows())
    field = int(p["fnx"]) * int(p["fny"]) * int(p["fnz"]) * 4
    log(f"[bench] device QoIs: collect {ms_collect:.3f} ms, volume {ms_volume:.3f} ms ({len(ids)} grains), "
        f"junctions {ms_junc:.3f} ms ({nj}), line {ms_line:.3f} ms; "
        f"collect + download of phi and alpha {ms_host:.1f} ms")
    return {
        "workload": "QoIs of the 256x256x512 state after the timed steps",
        "collect_on_device_ms": round(ms_collect, 3),
        "volume_ms": round(ms_volume, 3),
        "junctions3d_ms": round(ms_junc, 3),
        "line_ms": round(ms_line, 3),
        "collect_and_download_ms": round(ms_host, 2),
        "alpha_bytes": field,
        "grains": len(ids),
        "ids_above_num_theta": above,
        "junctions": nj,
        "note": "device calls each include their result allocation, launch, copy-back and sync; "
                "a host QoI pass would first need the download",
    }


def dropin_block(K: int) -> dict:
    """The drop-in `phase_field` command on the same configuration as the
    reference would run it: a forcing folder for a 256x256x512 window of a
    256x256x640 column (4999 grains, G 1e6 K/m, R 0.1 m/s, z0 2 um, t_end for
    10000 steps), line mode (-l 1: moving domain every 20 steps, 10 line-QoI
    samples, the alpha archive in the QoI file).  Reports the command's wall
    time, its own evolve timer and the files it wrote."""
    import re
    import shutil
    import subprocess
    import tempfile
    from cupf_amd import forcing
    cli = os.path.join(ROOT, "cupf_amd", "bin", "phase_field")
    d = tempfile.mkdtemp(prefix="cupf_dropin_")
    try:
        t0 = time.perf_counter()
        forcing.write_line_case(d, 0, n=(256, 256, 512), grains=5000, G=1.0, R=0.1, z0=2.0,
                                t_end=4.83e-5, nts=10, line_config=True, nz_full=640, top=48.0)
        t_case = time.perf_counter() - t0
        out = os.path.join(d, "out")
        os.makedirs(out)
        cmd = [cli, "line.py", "-s", "0", "-b", "110", "-f", os.path.join(d, "forcing", "case"),
               "-o", out, "-k", str(K), "-l", "1"]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError(f"phase_field failed ({r.returncode}): {r.stdout[-800:]} {r.stderr[-800:]}")
        m = re.search(r"time for (\d+) iterations: ([0-9.]+) s", r.stdout)
        steps, evolve_s = (int(m.group(1)), float(m.group(2))) if m else (None, None)
        files = {f: os.path.getsize(os.path.join(out, f)) for f in sorted(os.listdir(out))}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    res = {
        "workload": "phase_field -l 1 -k 8 on a 256x256x512 window of a 256x256x640 column, 4999 grains, "
                    "Mt 10000 (moving domain, 10 line-QoI samples)",
        "command": "phase_field line.py -s 0 -b 110 -f <case>/forcing/case -o <out> -k 8 -l 1",
        "case_write_s": round(t_case, 2),
        "wall_s": round(wall, 2),
        "evolve_s": evolve_s,
        "steps": steps,
        "evolve_ms_per_step": round(evolve_s * 1e3 / steps, 4) if steps else None,
        "setup_and_output_s": round(wall - evolve_s, 2) if evolve_s is not None else None,
        "output_bytes": sum(files.values()),
        "output_files": len(files),
    }
    log(f"[bench] drop-in phase_field: {wall:.2f} s wall, evolve {evolve_s} s for {steps} steps "
        f"({res['evolve_ms_per_step']} ms/step), {len(files)} files, {res['output_bytes'] / 1e6:.1f} MB")
    return res


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def new_ctx(case):
    from cupf_amd import capi
    ctx = capi.APTContext(case.params)
    ctx.set_coords(case.x, case.y, case.z)
    ctx.set_orientations(case.R9)
    ctx.init_from_phi_alpha(case.phi0, case.alpha0)
    return ctx


def profiled_window(ctx, p, first: int, steps: int):
    """Steps [first, first+steps) of ctx one at a time with HIP events around
    every launch; processed points of each step from the activity maps just
    before it.  Returns per-step lists (plan ms, step ms, fix-up ms, points)."""
    plan, main, fix, pts = [], [], [], []
    for q in range(steps):
        pts.append(processed_points(ctx, p))
        m, f, pl = ctx.step_profiled(1, first + q)
        plan.append(pl)
        main.append(m)
        fix.append(f)
    return plan, main, fix, pts


def cpu_block(ctx, case, step0: int, budget_s: float, threads: int) -> dict:
    """Oracle float32 and float64 from the GPU state after the timed steps, as
    many steps as fit the time budget (at least one); then the GPU steps on
    from the same state and is compared with both."""
    import oracle
    p = case.params
    n_upd = (p["fnx"] - 2) * (p["fny"] - 2) * (p["fnz"] - 2)
    ph, arg = ctx.download_state(0)
    res = {}
    nsteps = None
    for prec in ("f32", "f64"):
        sim = oracle.OracleSim(p, case.x, case.y, case.z, case.R9, prec)
        sim.set_state(ph, arg, 0)
        t0 = time.perf_counter()
        n = 0
        while True:
            sim.step(1, step0 + n)
            n += 1
            el = time.perf_counter() - t0
            if nsteps is not None:
                if n >= nsteps:
                    break
            elif el >= budget_s / 2:
                break
        nsteps = n
        sim.apply_bc(0)
        res[prec] = (sim.state(0), n, el)
        log(f"[bench] cpu oracle {prec}: {n} steps in {el:.1f} s, "
            f"{n_upd * n / el / 1e6:.3f} Mcell-steps/s on {threads} threads")
    ctx.step(nsteps, step0)
    ctx.apply_bc(0)
    gph, garg = ctx.download_state(0)
    o32, a32 = res["f32"][0]
    o64, a64 = res["f64"][0]
    cmp = {
        "gpu_vs_f32_arg_mismatches": int((garg != a32).sum()),
        "gpu_vs_f32_linf_phi": float(np.abs(gph.astype(np.float64) - o32.astype(np.float64)).max()),
        "gpu_vs_f64_arg_flips": int((garg != a64).sum()),
        "gpu_vs_f64_linf_phi": float(np.abs(gph.astype(np.float64) - o64).max()),
        "f32_vs_f64_arg_flips": int((a32 != a64).sum()),
        "f32_vs_f64_linf_phi": float(np.abs(o32.astype(np.float64) - o64).max()),
    }
    log(f"[bench] accuracy after {nsteps} steps: {cmp}")
    n32, e32 = res["f32"][1], res["f32"][2]
    n64, e64 = res["f64"][1], res["f64"][2]
    return {
        "value": round(n_upd * n32 / e32 / 1e6, 3),
        "unit": "Mcell-steps/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "f64_value": round(n_upd * n64 / e64 / 1e6, 3),
        "sample": f"{nsteps} steps of the 256x256x512 K=8 state after the timed GPU steps "
                  f"(steps {step0}..{step0 + nsteps - 1}), oracle float32 {e32:.1f} s and float64 "
                  f"{e64:.1f} s, OpenMP {threads} threads",
        "accuracy": cmp,
        "accuracy_steps": nsteps,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=400)
    ap.add_argument("--K", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=24.0,
                    help="time budget of the CPU oracle sample (float32 and float64 together)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-only", type=int, default=0,
                    help="only time a plain copy kernel over this many bytes (counter calibration)")
    ap.add_argument("--z0-cells", type=float, default=None,
                    help="initial interface height in cells (default: z0 = 2 um as in line.py)")
    ap.add_argument("--full-steps", type=int, default=10000,
                    help="length of the full-run measurement (0: skip)")
    ap.add_argument("--no-dense", action="store_true", help="skip the dense (no plane skipping) window")
    ap.add_argument("--no-qoi", action="store_true", help="skip the device QoI timings")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream measurement")
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in phase_field run")
    ap.add_argument("--ensemble-sims", type=int, default=64)
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only run the batched-ensemble measurement and print its JSON object")
    ap.add_argument("--serve-requests", type=int, default=48)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        return 2

    from cupf_amd import capi, cases

    if args.copy_only:
        gbs = capi.copy_bandwidth(args.copy_only, 5)
        print(json.dumps({"copy_bytes": args.copy_only, "copy_gbs": round(gbs, 1)}), flush=True)
        return 0

    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_block(args.ensemble_sims, args.K)}), flush=True)
        return 0

    t0 = time.time()
    case = cases.c2(K=args.K, z0_cells=args.z0_cells)
    p = case.params
    n_upd = (p["fnx"] - 2) * (p["fny"] - 2) * (p["fnz"] - 2)
    log(f"[bench] case {case.name}: fnx,fny,fnz={p['fnx']},{p['fny']},{p['fnz']} grains={case.n_grains} "
        f"K={args.K} updated points={n_upd} (setup {time.time() - t0:.1f}s)")

    ctx = new_ctx(case)
    step0 = 1
    if args.warmup:
        ms = ctx.step_timed(args.warmup, step0)
        log(f"[bench] warmup {args.warmup} steps: {ms / args.warmup:.3f} ms/step")
        step0 += args.warmup
    ovf0 = ctx.overflow_count()

    # timed region: stream synchronised on both sides, wall clock + HIP events
    ctx.sync()
    w0 = time.perf_counter()
    ev_ms = ctx.step_timed(args.steps, step0)
    ctx.sync()
    wall = time.perf_counter() - w0
    step_after = step0 + args.steps
    ovf = ctx.overflow_count() - ovf0
    mcell = n_upd * args.steps / wall / 1e6
    log(f"[bench] timed {args.steps} steps: wall {wall * 1e3 / args.steps:.3f} ms/step, "
        f"events {ev_ms / args.steps:.3f} ms/step, overflow points {ovf}")

    # the same window replayed with events around every launch
    rctx = new_ctx(case)
    if args.warmup:
        rctx.step(args.warmup, 1)
    rovf0 = rctx.overflow_count()
    plan, main_ms, fix, pts = profiled_window(rctx, p, step0, args.steps)
    rovf = rctx.overflow_count() - rovf0
    rctx.close()
    nprof = len(main_ms)
    avg_plan, avg_main, avg_fix = sum(plan) / nprof, sum(main_ms) / nprof, sum(fix) / nprof
    proc = sum(pts) / nprof
    log(f"[bench] per launch (same window): apt_plan {avg_plan:.4f} ms, apt_step {avg_main:.4f} ms, "
        f"apt_fixup {avg_fix:.4f} ms; processed points {min(pts)}..{max(pts)} (mean {proc:.0f}) "
        f"of {n_upd}; overflow points {rovf}")

    copy_gbs = capi.copy_bandwidth(1 << 30, 10)
    log(f"[bench] plain copy kernel: {copy_gbs:.0f} GB/s")
    bytes_step = proc * args.K * 16
    bytes_dense = n_upd * args.K * 16
    t_all = avg_plan + avg_main + avg_fix
    achieved = bytes_step / (t_all * 1e-3) / 1e9
    achieved_main = bytes_step / (avg_main * 1e-3) / 1e9
    dominant = "apt_step_tiled<%d,false>" % args.K if avg_main >= avg_fix else "apt_fixup<%d,false>" % args.K
    traffic, traffic_kernels = traffic_from_profiles(f"line-256x256x512:w{args.warmup}:s{args.steps}")

    dense = None
    if not args.no_dense:
        os.environ["CUPF_NO_ACTIVITY"] = "1"
        dctx = new_ctx(case)
        del os.environ["CUPF_NO_ACTIVITY"]
        if args.warmup:
            dctx.step(args.warmup, 1)
        dm, df, dp = dctx.step_profiled(args.steps, step0)
        g1 = ctx.download_state(0)
        g2 = dctx.download_state(0)
        same = bool(np.array_equal(g1[1], g2[1]) and np.array_equal(g1[0].view(np.uint32), g2[0].view(np.uint32)))
        del g1, g2
        dctx.close()
        d_main = dm / args.steps
        dense = {
            "workload": "same window, every plane processed (CUPF_NO_ACTIVITY=1)",
            "kernel_ms": round(d_main, 4),
            "plan_ms": round(dp / args.steps, 4),
            "fixup_ms": round(df / args.steps, 4),
            "algorithmic_bytes_per_launch": bytes_dense,
            "achieved_gbs": round(bytes_dense / (d_main * 1e-3) / 1e9, 1),
            "frac": round(bytes_dense / (d_main * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "same_bits_as_skipping": same,
        }
        log(f"[bench] dense window: step kernel {d_main:.4f} ms ({dense['achieved_gbs']} GB/s), "
            f"same bits as skipping: {same}")

    full = None
    if args.full_steps > 0:
        # the whole run in blocks (one stream sync between blocks): the
        # start-up transient, the thin solid slab, then the slab growing as the
        # front advances
        fctx = new_ctx(case)
        edges = [b for b in (5, 25, 100, 200, 300, 500, 1000, 2000, 3000, 5000, 7000) if b < args.full_steps]
        edges.append(args.full_steps)
        blocks = []
        fsec = 0.0
        s = 1
        for e in edges:
            n = e - s + 1
            pts = processed_points(fctx, p)
            o0 = fctx.overflow_count()
            fctx.sync()
            f0 = time.perf_counter()
            fctx.step(n, s)
            fctx.sync()
            dt = time.perf_counter() - f0
            fsec += dt
            blocks.append({"steps": [s, e], "ms_per_step": round(dt * 1e3 / n, 4),
                           "processed_points_first_step": pts,
                           "overflow_points_per_step": round((fctx.overflow_count() - o0) / n)})
            s = e + 1
        fovf = fctx.overflow_count()
        fctx.close()
        full = {
            "workload": f"the whole {args.full_steps}-step run from the initial state",
            "steps": args.full_steps,
            "ms_per_step": round(fsec * 1e3 / args.full_steps, 4),
            "mcell_steps_per_s": round(n_upd * args.full_steps / fsec / 1e6, 1),
            "wall_s": round(fsec, 3),
            "overflow_points": fovf,
            "blocks": blocks,
        }
        log(f"[bench] full run: {args.full_steps} steps in {fsec:.2f} s "
            f"({full['ms_per_step']} ms/step, {full['mcell_steps_per_s']} Mcell-steps/s)")
        for b in blocks:
            log(f"[bench]   steps {b['steps'][0]}..{b['steps'][1]}: {b['ms_per_step']} ms/step, "
                f"processed {b['processed_points_first_step']}, overflow/step {b['overflow_points_per_step']}")

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = cpu_block(ctx, case, step_after, args.cpu_seconds, threads)

    qoi = None if args.no_qoi else qoi_block(ctx, p)
    dropin = None
    if not args.no_dropin:
        try:
            dropin = dropin_block(args.K)
        except (RuntimeError, OSError, ValueError) as e:  # reported in the line, not fatal to it
            log(f"[bench] drop-in run failed: {e}")
            dropin = {"error": str(e)[:400]}

    ens = None
    if not args.no_ensemble:
        ens = ensemble_block(args.ensemble_sims, args.K)
    srv = None
    if not args.no_serve:
        srv = serving_block(args.serve_requests, args.K)

    out = {
        "metric": "Mcell-steps/s + %HBM roofline at 256^2x512 K=8; ensemble sims/h, p50/p99 latency",
        "value": round(mcell, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+i32",
        "data": "synthetic: seeded periodic Voronoi substrate, random quaternion orientations, "
                "line thermal field; state resident in HBM",
        "config": {
            "workload": "line-256x256x512",
            "grid": [p["fnx"], p["fny"], p["fnz"]],
            "K": args.K,
            "grains": case.n_grains,
            "G_K_per_m": 1e6,
            "R_m_per_s": 0.1,
            "bc": "110",
            "z0_um": float(p["z0"]),
            "window": f"steps {step0}..{step_after - 1} after {args.warmup} warm-up steps",
            "overflow_points_timed": ovf,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_by_kernel": traffic_kernels,
            "kernel": dominant,
            "kernels": "apt_plan + apt_step_tiled (+ apt_edge beside it) + apt_fixup (one step)",
            "step_ms": round(t_all, 4),
            "plan_ms": round(avg_plan, 4),
            "kernel_ms": round(avg_main, 4),
            "fixup_ms": round(avg_fix, 4),
            "step_kernel_achieved": round(achieved_main, 1),
            "step_kernel_frac": round(achieved_main / HBM_PEAK_GBS, 4),
            "copy_kernel_gbs": round(copy_gbs, 1),
            "frac_of_copy": round(achieved / copy_gbs, 4),
            "algorithmic_bytes_per_launch": int(bytes_step),
            "processed_points_per_launch": int(proc),
            "updated_points_per_launch": n_upd,
            "overflow_points_per_launch": int(rovf // max(1, args.steps)),
        },
        "dense": dense,
        "full_run": full,
        "cpu_baseline": cpu,
        "qoi": qoi,
        "dropin": dropin,
        "ensemble": ens,
        "serving": srv,
    }
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
