"""Overflow handling (neighbourhoods holding more than K distinct grain ids,
3DgrainMPI/APTPhaseField.cu:222-231): the keep test in the tile, the fix-up
records, the streaming top-K fix-up and the edge tiles of a one-column strip.

The cases have the 256x256x512 benchmark's grain density (grains ~3.6 cells
wide, so most points of the start-up transient overflow), grain ids spread over
[1, 2^31 - 2] (the range nucleated ids take), and the fix-up grid cut to two
workgroups, so that every lane takes dozens of overflowed points in turn.  Bar:
ids and phi bit-exact against the oracle and equal overflow counts; batched
members equal to their single runs; both step kernel variants (with records,
keep test and the edge kernel beside it, and without) equal to each other and
to the oracle.
"""
import numpy as np
import pytest

from cupf_amd import capi, cases, ensemble
from tests.helpers import compare_states, first_mismatch, gpu_context, oracle_sim

pytestmark = pytest.mark.gpu


def _dense_case(name, n, seed, K=8):
    grains = int(round(n[0] * n[1] / 3.6 ** 2))
    c = cases.make_line_case(name, n=n, grains=grains, K=K, seed=seed)
    rng = np.random.default_rng(seed)
    ids = np.unique(c.alpha0[c.alpha0 > 0])
    while True:
        new = rng.integers(1, 2 ** 31 - 1, size=ids.size, dtype=np.int64)
        if np.unique(new).size == new.size:
            break
    lut = np.zeros(int(ids.max()) + 1, np.int64)
    lut[ids] = new
    a = c.alpha0.astype(np.int64)
    a[a > 0] = lut[a[a > 0]]
    c.alpha0[:] = a.astype(np.int32)
    return c


@pytest.mark.parametrize("K", [5, 8])
def test_fixup_tables_reused_with_large_ids(K, monkeypatch):
    monkeypatch.setenv("CUPF_FIX_BLOCKS", "2")
    c = _dense_case("ovf-large-ids", (48, 48, 24), seed=17, K=K)
    assert int(c.alpha0.max()) > 1 << 24
    ctx = gpu_context(c)
    sim = oracle_sim(c)
    s = 1
    for stop in (5, 12, 20):
        ctx.step(stop - s + 1, s)
        sim.step(stop - s + 1, s)
        s = stop + 1
        print(f"[ovf K={K}] step {stop}: overflow {sim.overflow}", flush=True)
        assert ctx.overflow_count() == sim.overflow
        bad, dphi = compare_states(ctx, sim)
        p = c.params
        assert bad == 0, first_mismatch(ctx, sim, K, p["fnx"], p["fny"], p["fnz"])
        assert dphi == 0.0
    # far more overflowed points than the two workgroups' 128 lanes
    assert sim.overflow > 20 * 128


def test_batched_fixup_tables_with_large_ids(monkeypatch):
    monkeypatch.setenv("CUPF_FIX_BLOCKS", "2")
    members = [_dense_case("ovf-b0", (40, 40, 20), seed=5), _dense_case("ovf-b1", (24, 56, 16), seed=6)]
    alone = []
    for c in members:
        ctx = gpu_context(c)
        ctx.step(15, 1)
        alone.append(ensemble.state_hash(ctx))
        assert ctx.overflow_count() > 0
        ctx.close()
    ctxs = [gpu_context(c) for c in members]
    b = capi.Batch(ctxs)
    b.step(15)
    b.close()
    for q, ctx in enumerate(ctxs):
        assert ensemble.state_hash(ctx) == alone[q], f"member {q} differs from its single run"


@pytest.mark.parametrize("K", [8, 5])
def test_step_variants_same_bits(K, monkeypatch):
    """The step kernel with fix-up records and keep test (its edge strip as
    the edge kernel beside it), the plain one, and the host's choice between
    them (the overflow variant after a new state and while the fix-up reports
    heavy overflow, then the plain one) end with the same bits, on a
    65-point-wide grid (a one-column +x strip), through the start-up overflow
    (150 steps, the overflow variant also against the oracle) and past it
    (400 steps), for K=8 and K=5."""
    c = _dense_case("ovf-variants", (64, 40, 24), seed=23, K=K)
    p = c.params
    assert p["fnx"] - 2 == 65  # one full 64-wide tile and a one-column strip
    ctxs = {}
    for v in ("0", "1", "auto"):
        if v == "auto":
            monkeypatch.delenv("CUPF_STEP_VARIANT", raising=False)
        else:
            monkeypatch.setenv("CUPF_STEP_VARIANT", v)
        ctxs[v] = gpu_context(c)
    monkeypatch.delenv("CUPF_STEP_VARIANT", raising=False)

    def state(ctx):
        ctx.apply_bc(0)
        ph, arg = ctx.download_state(0)
        return ph.view(np.int32).copy(), arg.copy(), ctx.overflow_count()

    sim = oracle_sim(c)
    sim.step(150, 1)
    for s0, n in ((1, 150), (151, 250)):
        got = {}
        for v, ctx in ctxs.items():
            ctx.step(n, s0)
            got[v] = state(ctx)
            print(f"[variants K={K}] step {s0 + n - 1} variant {v}: overflow {got[v][2]}", flush=True)
        if s0 == 1:
            assert got["1"][2] == sim.overflow and sim.overflow > 0
            bad, dphi = compare_states(ctxs["1"], sim)
            assert bad == 0 and dphi == 0.0, first_mismatch(ctxs["1"], sim, K, p["fnx"], p["fny"], p["fnz"])
        for v in ("0", "auto"):
            assert np.array_equal(got[v][1], got["1"][1]), (v, s0)
            assert np.array_equal(got[v][0], got["1"][0]), (v, s0)
            assert got[v][2] == got["1"][2]
    for ctx in ctxs.values():
        ctx.close()


@pytest.mark.parametrize("K", [8, 5])
def test_batch_variants_same_bits(K, monkeypatch):
    """batch_step with fix-up records and keep test, the one without, and the
    host's choice between them (per 8-step graph replay, then per step) give
    every member the bits of its single run, through the start-up overflow
    and past it; one member has a one-column +x strip (edge tiles with and
    without records)."""
    members = [_dense_case("bv0", (64, 40, 24), seed=29, K=K), _dense_case("bv1", (40, 40, 20), seed=31, K=K)]
    assert members[0].params["fnx"] - 2 == 65
    alone = []
    for c in members:
        ctx = gpu_context(c)
        ctx.step(200, 1)
        alone.append((ensemble.state_hash(ctx), ctx.overflow_count()))
        assert ctx.overflow_count() > 0
        ctx.close()
    for v in ("0", "1", "auto"):
        if v == "auto":
            monkeypatch.delenv("CUPF_STEP_VARIANT", raising=False)
        else:
            monkeypatch.setenv("CUPF_STEP_VARIANT", v)
        ctxs = [gpu_context(c) for c in members]
        b = capi.Batch(ctxs)
        b.step(96, capi.PF_BATCH_GRAPH)
        b.step(104)
        b.close()
        print(f"[batch variants K={K}] variant {v}: overflow {[x.overflow_count() for x in ctxs]}", flush=True)
        for q, ctx in enumerate(ctxs):
            assert ensemble.state_hash(ctx) == alone[q][0], f"variant {v}: member {q} differs from its single run"
            assert ctx.overflow_count() == alone[q][1]
            ctx.close()
    monkeypatch.delenv("CUPF_STEP_VARIANT", raising=False)


def test_fused_fixup_plan_same_bits():
    """pf_step sends each step's fix-up and the next step's plan as one launch
    within a call; one step per call, and pf_step_profiled, launch them
    separately.  All three end with the same bits and overflow counts,
    through the overflow start-up (one-column +x strip, edge kernel)."""
    c = _dense_case("fused", (64, 40, 24), seed=41)
    ctxs = [gpu_context(c) for _ in range(3)]
    n = 60
    ctxs[0].step(n, 1)
    for s in range(1, n + 1):
        ctxs[1].step(1, s)
    ctxs[2].step_profiled(n, 1)
    got = [(ensemble.state_hash(x), x.overflow_count()) for x in ctxs]
    print(f"[fused] overflow {[g[1] for g in got]}", flush=True)
    assert got[0][1] > 0
    assert got[1] == got[0] and got[2] == got[0]
    for x in ctxs:
        x.close()
