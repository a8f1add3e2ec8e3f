"""Continuous batching (cupf_amd.ensemble.Server): each served simulation ends
with the same bits as when it runs alone, whatever it shared the GPU with.

The stream is small but exercises every path: requests arriving while others
run, admission held back by the HBM budget, quanta cut short at a member's
end step, the batch re-formed on every admission and retirement, graph
replays inside quanta."""
import pytest

from cupf_amd import ensemble

pytestmark = pytest.mark.gpu


def _stream():
    grids = [(16, 16, 16), (32, 32, 24), (20, 12, 30), (70, 9, 16), (24, 24, 24), (16, 40, 12)]
    steps = [37, 64, 23, 51, 90, 45]
    arrivals = [0.0, 0.0, 0.02, 0.05, 0.3, 0.31]
    return [ensemble.SimSpec(n=g, seed=500 + q, G=0.5 + q, R=0.05 * (q + 1), nsteps=s,
                             grains=6 + q, arrival=a)
            for q, (g, s, a) in enumerate(zip(grids, steps, arrivals))]


def test_served_results_equal_alone_runs():
    specs = _stream()
    prepared = [ensemble.make_case(s) for s in specs]
    # a budget that holds only part of the stream at once forces queueing
    budget = sum(s.footprint() for s in specs[:3])
    srv = ensemble.Server(hbm_budget=budget, quantum=16)
    done, wall = srv.run(specs, prepared)
    assert len(done) == len(specs)
    alone = {}
    for q, s in enumerate(specs):
        ctx = ensemble.new_context(prepared[q])
        ctx.step(s.nsteps, 1)
        alone[s.seed] = ensemble.state_hash(ctx)
        ctx.close()
    for d in done:
        assert d.state_hash == alone[d.spec.seed], f"request {d.spec.seed} differs from its alone run"
        assert d.finished >= d.spec.arrival
    assert wall > 0.0


def test_ensemble_batched_equals_back_to_back():
    specs = [ensemble.SimSpec(n=g, seed=700 + q, G=1.0 + q, R=0.1, nsteps=40, grains=8)
             for q, g in enumerate([(16, 16, 16), (32, 32, 24), (70, 9, 16), (24, 24, 24)])]
    prepared = [ensemble.make_case(s) for s in specs]
    _, _, h_alone = ensemble.run_back_to_back(specs, prepared)
    _, kc, h_batch = ensemble.run_batched(specs, prepared, graph=True, autotune=False)
    assert h_alone == h_batch
