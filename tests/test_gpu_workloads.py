"""The benchmark workloads themselves, checked at their real shapes.

- 64^3 plain run (K=8, ~100 grains, BC 110) for its full 2000 steps against
  the oracle: bit-exact ids and phi through the overflow transient and after.
- The batched ensemble's four grid classes (32^3, 64^3, 96x96x128,
  128x128x256, G log-uniform in [1e5, 1e7] K/m, R log-uniform in [0.01, 1]
  m/s) stepped as one batch (hipGraph replays, autotuned chunk depth): every
  member ends with the bits of its alone run.
- A served Poisson stream (4 requests/s) over the same grid classes with step
  counts of a few hundred, admitted under an HBM budget that holds only part
  of it: every request ends with the bits of its alone run.
"""
import numpy as np
import pytest

from cupf_amd import cases, ensemble
from tests.helpers import compare_states, first_mismatch, gpu_context, oracle_sim

pytestmark = pytest.mark.gpu


def test_plain_64cubed_full_2000_steps_against_the_oracle():
    c = cases.c1(K=8)
    ctx = gpu_context(c)
    sim = oracle_sim(c)
    s = 1
    for stop in (500, 1000, 2000):
        while s <= stop:
            n = min(250, stop - s + 1)
            ctx.step(n, s)
            sim.step(n, s)
            s += n
        print(f"[c1] step {stop}: overflow {sim.overflow}", flush=True)
        assert ctx.overflow_count() == sim.overflow
        bad, dphi = compare_states(ctx, sim)
        assert bad == 0, (stop, first_mismatch(ctx, sim, 8, 67, 67, 67))
        assert dphi == 0.0
    assert sim.overflow > 0


def _grid_class_specs(nsteps, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for q, n in enumerate(ensemble.GRIDS):
        G = float(10.0 ** rng.uniform(-1.0, 1.0))
        R = float(10.0 ** rng.uniform(-2.0, 0.0))
        out.append(ensemble.SimSpec(n=n, seed=3000 + q, G=G, R=R, nsteps=nsteps,
                                    grains=ensemble._grains(n)))
    return out


def test_ensemble_grid_classes_batched_equal_alone():
    specs = _grid_class_specs(300)
    prepared = [ensemble.make_case(s) for s in specs]
    _, _, h_alone = ensemble.run_back_to_back(specs, prepared)
    _, kc, h_batch = ensemble.run_batched(specs, prepared, graph=True, autotune=True)
    print(f"[c4] chunk depth {kc}", flush=True)
    for q, s in enumerate(specs):
        assert h_batch[q] == h_alone[q], f"grid {s.n} differs from its alone run"


def test_served_stream_over_grid_classes_equals_alone_runs():
    rng = np.random.default_rng(11)
    specs = []
    t = 0.0
    for q in range(8):
        n = ensemble.GRIDS[q % 4]
        t += float(rng.exponential(0.25))
        specs.append(ensemble.SimSpec(n=n, seed=4000 + q, G=float(10.0 ** rng.uniform(-1, 1)),
                                      R=float(10.0 ** rng.uniform(-2, 0)),
                                      nsteps=int(rng.integers(150, 450)), grains=ensemble._grains(n),
                                      arrival=t))
    prepared = [ensemble.make_case(s) for s in specs]
    budget = sum(sorted(s.footprint() for s in specs)[-3:])
    srv = ensemble.Server(hbm_budget=budget, quantum=100)
    done, wall = srv.run(specs, prepared)
    assert len(done) == len(specs)
    print(f"[c5] served {len(done)} in {wall:.2f} s", flush=True)
    for q, s in enumerate(specs):
        ctx = ensemble.new_context(prepared[q])
        ctx.step(s.nsteps, 1)
        h = ensemble.state_hash(ctx)
        ctx.close()
        d = [x for x in done if x.spec.seed == s.seed][0]
        assert d.state_hash == h, f"request {s.seed} ({s.n}, {s.nsteps} steps) differs from its alone run"
