"""Plane skipping with the activity map: the same bits as processing every point.

The step skips z planes of 64x4 column tiles whose stencil reaches only empty
slot pairs (-1, -1.0) and whose destination holds only empty pairs.  These
tests run the skipping context beside one created with CUPF_NO_ACTIVITY=1
(every plane processed) and against the oracle, through the cases where the
map must be kept right: a front advancing into liquid, nucleation writing
into inactive planes, a moving-frame shift, uploads, and partial tiles in x
and y.  They also check that the map covers every tile plane that really
holds a non-empty slot.
"""
import os

import numpy as np
import pytest

from cupf_amd import capi, cases
from tests.helpers import compare_states, first_mismatch, gpu_context, oracle_sim

pytestmark = pytest.mark.gpu


def _dense_context(case, **kw):
    old = os.environ.get("CUPF_NO_ACTIVITY")
    os.environ["CUPF_NO_ACTIVITY"] = "1"
    try:
        return gpu_context(case, **kw)
    finally:
        if old is None:
            del os.environ["CUPF_NO_ACTIVITY"]
        else:
            os.environ["CUPF_NO_ACTIVITY"] = old


def _truth_map(ph, arg, K, fnx, fny, fnz):
    """Tile planes holding a slot other than (-1, -1.0), from a downloaded state."""
    a = arg.reshape(K, fnz, fny, fnx)
    p = ph.reshape(K, fnz, fny, fnx)
    ne = ((a != -1) | (p != np.float32(-1.0))).any(axis=0)[1:-1, 1:-1, 1:-1]
    ntx, nty = (fnx - 2 + 63) // 64, (fny - 2 + 3) // 4
    out = np.zeros((fnz, nty, ntx), np.uint8)
    kk, jj, ii = np.nonzero(ne)
    out[kk + 1, jj // 4, ii // 64] = 1
    return out


def _same_state(c1, c2):
    for which in (0, 1):
        p1, a1 = c1.download_state(which)
        p2, a2 = c2.download_state(which)
        assert np.array_equal(a1, a2), f"arg differs in buffer {which}"
        assert np.array_equal(p1.view(np.uint32), p2.view(np.uint32)), f"phi differs in buffer {which}"


def _check_cover(ctx, case):
    P = case.params
    for which in (0, 1):
        act, total, m = ctx.activity(which, with_map=True)
        ph, arg = ctx.download_state(which)
        t = _truth_map(ph, arg, int(P["K"]), int(P["fnx"]), int(P["fny"]), int(P["fnz"]))
        missing = np.argwhere((t == 1) & (m == 0))
        assert missing.size == 0, f"buffer {which}: non-empty tile planes not in the map: {missing[:5]}"
        assert act == int(m.sum()) and total == m.shape[1] * m.shape[2] * (m.shape[0] - 2)


@pytest.mark.parametrize("shape,K,g27,bc", [
    ((130, 9, 70), 8, 0, (1, 1, 0)),    # partial tiles in x and y, front in the lower planes
    ((127, 7, 66), 5, 0, (0, 0, 1)),    # full tiles only, periodic z
    ((70, 10, 50), 8, 1, (1, 0, 0)),    # 27-point gather
])
def test_skip_matches_dense_and_oracle(shape, K, g27, bc):
    c = cases.make_line_case("act", n=shape, grains=12, K=K, seed=51, z0_cells=10, bc=bc,
                             gather27=g27,
                             thermal_type=capi.THERMAL_LINE_3DGRAIN if g27 else capi.THERMAL_LINE)
    ctx = gpu_context(c)
    dense = _dense_context(c)
    sim = oracle_sim(c)
    for s in (1, 11, 21):
        ctx.step(10, s)
        dense.step(10, s)
        sim.step(10, s)
    a, total = ctx.activity(0)
    assert 0 < a < total // 2, (a, total)      # most tile planes are liquid and skipped
    _check_cover(ctx, c)
    _same_state(ctx, dense)
    bad, dphi = compare_states(ctx, sim)
    assert bad == 0, first_mismatch(ctx, sim, K, c.params["fnx"], c.params["fny"], c.params["fnz"])
    assert dphi == 0.0


def test_skip_with_nucleation_in_liquid():
    """All-liquid start: nucleation writes grains into tile planes the map
    holds as empty; they must be picked up on the next step."""
    c = cases.make_line_case("nucl-act", n=(70, 66, 34), grains=3, K=5, seed=37, z0_cells=-20)
    c.params.update(uc0=np.float32(8.0), nuc_Nmax0=np.float32(5.0))
    ctx = gpu_context(c)
    dense = _dense_context(c)
    sim = oracle_sim(c)
    for x in (ctx, dense, sim):
        x.nucl_reset(0)
    assert ctx.activity(0)[0] == 0
    for s in range(1, 41, 10):
        ctx.step(10, s, capi.PF_STEP_INITIAL_UC | capi.PF_STEP_NUCLEATE)
        dense.step(10, s, capi.PF_STEP_INITIAL_UC | capi.PF_STEP_NUCLEATE)
        sim.step(10, s, use_initial_uc=True, nucleate=True, seed=int(c.params["seed"]))
    assert sim.nucl_events > 0 and ctx.nucl_events() == sim.nucl_events
    _check_cover(ctx, c)
    _same_state(ctx, dense)
    assert compare_states(ctx, sim)[0] == 0


def test_skip_through_moving_frame_and_upload():
    c = cases.make_line_case("md-act", n=(66, 9, 40), grains=6, K=5, seed=31, z0_cells=20,
                             line_config=True, nz_full=80)
    ctx = gpu_context(c)
    dense = _dense_context(c)
    for x in (ctx, dense):
        x.step(6, 1)
        x.md_enable(c.d["fnz_f"])
        for mc in range(3):
            x.md_shift(mc)
        x.step(6, 7)
    _check_cover(ctx, c)
    _same_state(ctx, dense)
    # a state uploaded into the current buffer is mapped from its contents
    ph, arg = dense.download_state(0)
    ctx.upload_state(ph, arg, 0)
    _check_cover(ctx, c)
    for x in (ctx, dense):
        x.step(5, 13)
    _same_state(ctx, dense)


def test_c2_skip_matches_dense_full_size():
    """256x256x512, 5000 grains, 40 steps from the initial state: the benchmark
    workload, compared bit for bit with every plane processed."""
    c = cases.c2(K=8)
    ctx = gpu_context(c)
    dense = _dense_context(c)
    ctx.step(40, 1)
    dense.step(40, 1)
    _same_state(ctx, dense)
    a, total = ctx.activity(0)
    assert a < total // 4, (a, total)
