"""Batched ensembles (pf_batch_*): every member gets exactly the bits pf_step gives it.

Mixed members (grid shapes with partial tiles and edge strips, different
G/R, seeds, boundary conditions and thermal types) are stepped as one batch
and compared bit for bit with the same simulations stepped alone, through
hipGraph replays interleaved with eager single steps, after the autotuner
has switched chunk depths, and against the oracle for one member.
"""
import numpy as np
import pytest

from cupf_amd import capi, cases
from tests.helpers import compare_states, first_mismatch, gpu_context, oracle_sim

pytestmark = pytest.mark.gpu

SPECS = [
    # n, grains, seed, G (K/um), R (m/s), bc, thermal type, z0 cells
    ((32, 32, 32), 20, 11, 1.0, 0.1, (1, 1, 0), capi.THERMAL_LINE, 16),
    ((70, 9, 24), 12, 12, 0.3, 0.5, (0, 1, 0), capi.THERMAL_LINE, 12),
    ((95, 10, 20), 15, 13, 5.0, 0.02, (1, 1, 0), capi.THERMAL_LINE, 10),
    ((16, 16, 40), 6, 14, 2.0, 0.2, (1, 0, 1), capi.THERMAL_LINE_3DGRAIN, 20),
    ((130, 9, 12), 14, 15, 1.0, 0.1, (1, 1, 0), capi.THERMAL_LINE, 6),
]


def _case(spec, K=8, g27=0):
    n, grains, seed, G, R, bc, tt, z0c = spec
    return cases.make_line_case(f"b{seed}", n=n, grains=grains, K=K, seed=seed, G=G, R=R,
                                z0_cells=z0c, bc=bc, thermal_type=tt, gather27=g27)


def _same(c1, c2, which=(0, 1)):
    for w in which:
        p1, a1 = c1.download_state(w)
        p2, a2 = c2.download_state(w)
        if not (np.array_equal(a1, a2) and np.array_equal(p1.view(np.uint32), p2.view(np.uint32))):
            return False
    return True


def _pairs(K=8, g27=0):
    cs = [_case(sp, K, g27) for sp in SPECS]
    return cs, [gpu_context(c) for c in cs], [gpu_context(c) for c in cs]


def test_batch_matches_single_steps():
    cs, batched, single = _pairs()
    b = capi.Batch(batched)
    for s0 in (1, 11):
        b.step(10)
        for c in single:
            c.step(10, s0)
        for i, (x, y) in enumerate(zip(batched, single)):
            assert _same(x, y), f"member {i} differs after step {s0 + 9}"
        assert [m.overflow_count() for m in batched] == [m.overflow_count() for m in single]
    b.close()


def test_batch_graph_replays_interleaved_with_eager_work():
    """Graph replays (groups of 8 steps) across several calls, with single
    steps of a member and downloads in between: every replay must match."""
    cs, batched, single = _pairs()
    b = capi.Batch(batched)
    step = 1
    for rnd in range(3):
        b.step(16, capi.PF_BATCH_GRAPH)          # two replays of the 8-step graph
        for c in single:
            c.step(16, step)
        step += 16
        for i, (x, y) in enumerate(zip(batched, single)):
            assert _same(x, y), f"member {i} differs after graph round {rnd}"
        # eager work between replays: member 0 steps alone, both sides
        batched[0].step(3, step)
        single[0].step(3, step)
        b.next_step[0] += 3
        for c in single[1:]:
            c.step(3, step)
        for i in range(1, len(batched)):
            batched[i].step(3, step)
            b.next_step[i] += 3
        step += 3
    for i, (x, y) in enumerate(zip(batched, single)):
        assert _same(x, y), f"member {i} differs at the end"
    b.close()


def test_batch_autotune_and_chunk_depths_do_not_change_results():
    cs, batched, single = _pairs()
    b = capi.Batch(batched)
    b.step(80, capi.PF_BATCH_AUTOTUNE)
    assert b.kchunk in (2, 4, 8, 16)
    b.kchunk = 3
    b.step(7)
    for c in single:
        c.step(87, 1)
    for i, (x, y) in enumerate(zip(batched, single)):
        assert _same(x, y), f"member {i} differs"
    b.close()


@pytest.mark.parametrize("K,g27", [(5, 0), (5, 1), (8, 1)])
def test_batch_other_variants_match_single_steps(K, g27):
    """The K=5 slot set and the 27-point id gather (3Dgrain/) batched: eager
    steps, graph replays and a changed chunk depth, bit for bit against the
    same members stepped alone."""
    cs, batched, single = _pairs(K, g27)
    b = capi.Batch(batched)
    b.step(10)
    b.step(16, capi.PF_BATCH_GRAPH)
    b.kchunk = 8
    b.step(6)
    for c in single:
        c.step(32, 1)
    for i, (x, y) in enumerate(zip(batched, single)):
        assert _same(x, y), f"member {i} differs (K={K}, 27-point gather={g27})"
    assert [m.overflow_count() for m in batched] == [m.overflow_count() for m in single]
    b.close()


def test_batch_member_against_oracle():
    cs = [_case(sp) for sp in SPECS[:3]]
    ctxs = [gpu_context(c) for c in cs]
    b = capi.Batch(ctxs)
    b.step(12)
    sim = oracle_sim(cs[1])
    sim.step(12, 1)
    bad, dphi = compare_states(ctxs[1], sim)
    assert bad == 0, first_mismatch(ctxs[1], sim, 8, cs[1].params["fnx"], cs[1].params["fny"],
                                    cs[1].params["fnz"])
    assert dphi == 0.0
    b.close()


def test_batch_rejects_mixed_k_gather_and_nucleation():
    c8 = gpu_context(_case(SPECS[0], K=8))
    c5 = gpu_context(_case(SPECS[1], K=5))
    with pytest.raises(capi.PFError):
        capi.Batch([c8, c5])
    g27 = gpu_context(_case(SPECS[1], K=8, g27=1))
    with pytest.raises(capi.PFError):
        capi.Batch([c8, g27])
    b = capi.Batch([c8])
    with pytest.raises(capi.PFError):
        b.step(2, capi.PF_STEP_NUCLEATE)
    b.close()
