"""Ensembles and serving over the batched step (pf_batch_*).

The reference runs one simulation per process, one after another
(3Dgrain/createcmd:1, 3DgrainMPI/grains3D.py:380-399).  Here many
simulations share one GPU:

- ensemble: a fixed set of seeded simulations of uneven grids and thermal
  parameters stepped together, one batched launch sequence per time step
  (captured in a hipGraph, chunk depth picked by the autotuner);
- Server: continuous batching.  Requests arrive over time, are admitted while
  their HBM footprint fits the budget, share the batched step with whatever is
  already running, and retire at their own end step; the batch is re-formed at
  every admission or retirement.

Every member gets the same bits as when it runs alone (tests/test_gpu_batch.py,
tests/test_gpu_serve.py), so batching is a schedule, not an approximation.
"""
from __future__ import annotations

import dataclasses
import hashlib
import time

import numpy as np

from . import capi, cases

GRIDS = [(32, 32, 32), (64, 64, 64), (96, 96, 128), (128, 128, 256)]


@dataclasses.dataclass
class SimSpec:
    """One simulation of an ensemble or a served request."""
    n: tuple
    seed: int
    G: float             # K/um (1.0 = 1e6 K/m)
    R: float             # m/s
    nsteps: int
    grains: int
    arrival: float = 0.0  # s after the start of a served stream

    def footprint(self, K: int = 8) -> int:
        """Device bytes of a context: two K-slot buffers (phi + id), the single
        fields, the fix-up list, its records (7 x K id and phi entries for up
        to min(interior / 8, 2^20) points) and the coordinate tables."""
        fnx, fny, fnz = self.n[0] + 3, self.n[1] + 3, self.n[2] + 3
        L = fnx * fny * fnz
        interior = (fnx - 2) * (fny - 2) * (fnz - 2)
        rec = max(1, min(interior // 8, 1 << 20)) * 7 * K * 8
        return 16 * K * L + 12 * L + 4 * interior + rec


def _grains(n) -> int:
    # the 64^3 benchmark's density: ~100 grains on a 64x64 substrate
    return max(4, int(round(100.0 * n[0] * n[1] / (64 * 64))))


def ensemble_specs(count: int = 64, seed: int = 0, nsteps: int = 2000) -> list:
    """Grids uniform over GRIDS, G log-uniform in [1e5, 1e7] K/m, R log-uniform in
    [0.01, 1] m/s, per-simulation Voronoi seed."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        n = GRIDS[int(rng.integers(len(GRIDS)))]
        G = float(10.0 ** rng.uniform(-1.0, 1.0))     # K/um
        R = float(10.0 ** rng.uniform(-2.0, 0.0))     # m/s
        out.append(SimSpec(n=n, seed=1000 + q, G=G, R=R, nsteps=nsteps, grains=_grains(n)))
    return out


def request_stream(count: int, rate: float = 4.0, seed: int = 1) -> list:
    """Served requests: Poisson arrivals (mean `rate` per second), the ensemble's
    grid/G/R/seed distributions, step counts uniform in [500, 5000]."""
    rng = np.random.default_rng(seed)
    specs = ensemble_specs(count, seed=seed + 100)
    t = 0.0
    for s in specs:
        t += float(rng.exponential(1.0 / rate))
        s.arrival = t
        s.nsteps = int(rng.integers(500, 5001))
    return specs


def make_case(spec: SimSpec, K: int = 8) -> cases.Case:
    return cases.make_line_case(f"sim{spec.seed}", n=spec.n, grains=spec.grains, K=K, seed=spec.seed,
                                G=spec.G, R=spec.R)


def new_context(case: cases.Case) -> capi.APTContext:
    ctx = capi.APTContext(case.params)
    ctx.set_coords(case.x, case.y, case.z)
    ctx.set_orientations(case.R9)
    ctx.init_from_phi_alpha(case.phi0, case.alpha0)
    return ctx


def state_hash(ctx: capi.APTContext) -> str:
    """Digest of the current buffer (phi bits and ids) of a context."""
    ph, arg = ctx.download_state(0)
    h = hashlib.blake2b(digest_size=16)
    h.update(ph.view(np.uint32).tobytes())
    h.update(arg.tobytes())
    return h.hexdigest()


def run_back_to_back(specs, prepared=None, K: int = 8):
    """Each simulation alone with pf_step, one after another.  Returns
    (stepping seconds, per-simulation seconds, state hashes)."""
    total, each, hashes = 0.0, [], []
    for q, s in enumerate(specs):
        case = prepared[q] if prepared is not None else make_case(s, K)
        ctx = new_context(case)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(s.nsteps, 1)
        ctx.sync()
        dt = time.perf_counter() - t0
        total += dt
        each.append(dt)
        hashes.append(state_hash(ctx))
        ctx.close()
    return total, each, hashes


def run_batched(specs, prepared=None, K: int = 8, graph: bool = True, autotune: bool = True):
    """All simulations as one batch (equal step counts).  Returns (stepping
    seconds, chosen chunk depth, state hashes)."""
    ncommon = specs[0].nsteps
    if any(s.nsteps != ncommon for s in specs):
        raise ValueError("run_batched needs equal step counts; use Server for mixed ones")
    ctxs = [new_context(prepared[q] if prepared is not None else make_case(s, K))
            for q, s in enumerate(specs)]
    b = capi.Batch(ctxs)
    flags = (capi.PF_BATCH_GRAPH if graph else 0) | (capi.PF_BATCH_AUTOTUNE if autotune else 0)
    for c in ctxs:
        c.sync()
    t0 = time.perf_counter()
    b.step(ncommon, flags)
    dt = time.perf_counter() - t0
    kc = b.kchunk
    hashes = [state_hash(c) for c in ctxs]
    b.close()
    for c in ctxs:
        c.close()
    return dt, kc, hashes


@dataclasses.dataclass
class Done:
    spec: SimSpec
    admitted: float
    finished: float
    state_hash: str

    @property
    def latency(self) -> float:
        return self.finished - self.spec.arrival


class Server:
    """Continuous batching of simulation requests on one GPU.

    Requests (SimSpec with arrival times, in seconds from start()) are admitted
    in arrival order once they have arrived and their footprint fits
    `hbm_budget` bytes.  Running simulations advance together in quanta of at
    most `quantum` steps, cut short so that no simulation overshoots its end
    step; a finished simulation is hashed, freed and reported, and the batch is
    re-formed.  Inputs are prepared before the clock starts (they are the
    requests' payload); what is timed is admission, device set-up, stepping
    and retirement.
    """

    TUNE_MIN_STEPS = 400

    def __init__(self, hbm_budget: int = 64 << 30, quantum: int = 200, K: int = 8,
                 graph: bool = True, autotune: bool = True):
        self.budget = int(hbm_budget)
        self.quantum = int(quantum)
        self.K = K
        self.graph = graph
        self.autotune = autotune

    def run(self, specs, prepared, clock=time.perf_counter):
        pending = sorted(range(len(specs)), key=lambda q: specs[q].arrival)
        running = []   # [q, ctx, next_step, admitted]
        done = []
        used = 0
        t0 = clock()
        batch, batch_ids = None, None
        kchunk = None  # the last tuned chunk depth, carried into re-formed batches
        while pending or running:
            now = clock() - t0
            admitted = False
            while pending and specs[pending[0]].arrival <= now:
                q = pending[0]
                need = specs[q].footprint(self.K)
                if used + need > self.budget and running:
                    break
                pending.pop(0)
                ctx = new_context(prepared[q])
                running.append([q, ctx, 1, now])
                used += need
                admitted = True
            if not running:
                # idle until the next arrival
                wait = specs[pending[0]].arrival - (clock() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.05))
                continue
            ids = tuple(r[0] for r in running)
            if batch is None or admitted or ids != batch_ids:
                if batch is not None:
                    batch.close()
                batch = capi.Batch([r[1] for r in running], [r[2] for r in running])
                batch_ids = ids
                if kchunk is not None:
                    batch.kchunk = kchunk
            left = [specs[r[0]].nsteps - r[2] + 1 for r in running]
            n = min(self.quantum, min(left))
            batch.next_step = [r[2] for r in running]
            flags = capi.PF_BATCH_GRAPH if (self.graph and n >= 8) else 0
            # a tuning round spends 40 of the quantum's steps at trial depths:
            # worth it only for long quanta (with 200-step quanta the served
            # stream's p50 was 3-4 % lower untuned, at the default depth 4)
            if self.autotune and n >= self.TUNE_MIN_STEPS:
                flags |= capi.PF_BATCH_AUTOTUNE
            batch.step(n, flags)
            kchunk = batch.kchunk
            for r in running:
                r[2] += n
            still = []
            for r in running:
                q, ctx, nxt, adm = r
                if nxt > specs[q].nsteps:
                    done.append(Done(specs[q], adm, clock() - t0, state_hash(ctx)))
                    ctx.close()
                    used -= specs[q].footprint(self.K)
                else:
                    still.append(r)
            if len(still) != len(running):
                running = still
                if batch is not None:
                    batch.close()
                batch, batch_ids = None, None
        if batch is not None:
            batch.close()
        return done, clock() - t0


def percentile(xs, p: float) -> float:
    return float(np.percentile(np.asarray(xs, dtype=np.float64), p)) if xs else float("nan")
