"""Device-side QoIs (pf_qoi_volume, pf_qoi_junctions3d) against the numpy
restatement of 3DgrainMPI/QOI.cpp in oracle/qoi.py, on the same collected
alpha field: a simulated state, and an uploaded block pattern with many
4-grain junctions and a liquid region."""
import ctypes

import numpy as np
import pytest

from cupf_amd import capi, cases
from oracle import qoi
from tests.helpers import gpu_context

pytestmark = pytest.mark.gpu


def _both(ctx, p, max_id=None):
    """Host alpha (for the restatement) and the device QoIs of the same field."""
    _, alpha = ctx.collect(0, capi.PF_COLLECT_CLEAR_ALPHA)
    ctx.collect_on_device(0, capi.PF_COLLECT_CLEAR_ALPHA)
    fnx, fny, fnz = int(p["fnx"]), int(p["fny"]), int(p["fnz"])
    if max_id is None:
        max_id = max(1, int(alpha.max()))
    return alpha, (fnx, fny, fnz), ctx.qoi_volume(max_id), ctx.qoi_junctions3d()


def test_device_qoi_on_simulated_state():
    c = cases.make_line_case("qoi", n=(48, 40, 32), grains=40, K=8, seed=31, G=1.0, R=0.1,
                             z0_cells=20)
    ctx = gpu_context(c)
    ctx.step(40, 1)
    alpha, dims, (ids, counts, above), rows = _both(ctx, c.params)
    ref_ids, ref_counts = qoi.volume(alpha, *dims)
    assert above == 0
    assert ids == ref_ids and counts == ref_counts
    assert rows == qoi.junctions3d(alpha, *dims)


def test_device_qoi_on_block_pattern():
    c = cases.make_line_case("qoi-b", n=(40, 36, 30), grains=8, K=8, seed=32)
    ctx = gpu_context(c)
    p = c.params
    fnx, fny, fnz, K = int(p["fnx"]), int(p["fny"]), int(p["fnz"]), 8
    L = fnx * fny * fnz
    k, j, i = np.meshgrid(np.arange(fnz), np.arange(fny), np.arange(fnx), indexing="ij")
    rng = np.random.default_rng(5)
    table = rng.integers(1, 40, size=(fnz // 4 + 1, fny // 3 + 1, fnx // 3 + 1))
    ids = table[k // 4, j // 3, i // 3].astype(np.int32)
    liquid = (k > 2 * fnz // 3) & (i < fnx // 2)
    ph = np.full((K, L), -1.0, dtype=np.float32)
    arg = np.full((K, L), -1, dtype=np.int32)
    ph[0] = np.where(liquid, -1.0, 1.0).ravel()
    arg[0] = np.where(liquid, -1, ids).ravel()
    ctx.upload_state(ph, arg, 0)
    # ids up to 39: count them all, then with a bound that leaves some above
    alpha, dims, (vid, vcnt, above), rows = _both(ctx, p, 64)
    ref_ids, ref_counts = qoi.volume(alpha, *dims)
    assert above == 0 and vid == ref_ids and vcnt == ref_counts
    ref_rows = qoi.junctions3d(alpha, *dims)
    assert len(ref_rows) > 0
    assert rows == ref_rows
    # the field spans several junction workgroups (2048 scan indices each);
    # a row buffer shorter than the count holds the first rows in scan order
    assert (fnx - 2) * (fny - 2) * (fnz - 2) > 4 * 2048
    cut = max(1, len(ref_rows) // 3)
    buf = np.full(8 * cut, -7, dtype=np.int32)
    cnt = ctypes.c_int()
    assert capi.load().pf_qoi_junctions3d(ctx._h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                          cut, ctypes.byref(cnt)) == 0
    assert cnt.value == len(ref_rows)
    got = [(int(x[0]), int(x[1]), int(x[2]), int(x[3]), tuple(int(v) for v in x[4:]))
           for x in buf.reshape(-1, 8)]
    assert got == ref_rows[:cut]
    ids20, cnt20, above20 = ctx.qoi_volume(20)
    a = alpha.reshape(fnz, fny, fnx)[1:-1, 1:-1, 1:-1]
    assert above20 == int((a > 20).sum())
    keep = [(g, n) for g, n in zip(ref_ids, ref_counts) if g <= 20]
    assert list(zip(ids20, cnt20)) == keep


def test_device_line_qois():
    c = cases.make_line_case("qoi-l", n=(48, 40, 40), grains=30, K=8, seed=33, G=1.0, R=0.1,
                             z0_cells=12)
    ctx = gpu_context(c)
    p = c.params
    dims = (int(p["fnx"]), int(p["fny"]), int(p["fnz"]))
    ng = int(p["num_theta"])
    for s0 in (1, 31):
        ctx.step(30, s0)
        _, alpha = ctx.collect(0, capi.PF_COLLECT_CLEAR_ALPHA)
        ctx.collect_on_device(0, capi.PF_COLLECT_CLEAR_ALPHA)
        for cur in (0, 5):
            tip, total, extra = ctx.qoi_line(cur)
            rt, rtot, rext = qoi.line_qois(alpha, *dims, cur, ng)
            assert tip == rt
            assert np.array_equal(total, rtot) and np.array_equal(extra, rext)
