"""pf_qoi_volume_noalloc, the per-grain volume QoI for callers that cannot
allocate device memory (no counterpart in the reference), against the
allocating pf_qoi_volume through the C ABI: two calls in a row, the second on
a smaller field with a smaller max_id (so the fixed device array still holds
the first call's larger contents), then the size limit.  Bar: counts, first
indices and the count of larger ids identical."""
import ctypes

import numpy as np
import pytest

from cupf_amd import capi, cases
from tests.helpers import gpu_context

pytestmark = pytest.mark.gpu


def _call(ctx, name, max_id):
    counts = np.full(max_id + 1, -7, np.int32)
    first = np.full(max_id + 1, -7, np.int32)
    above = ctypes.c_longlong(-7)
    capi._check(getattr(capi.load(), name)(ctx._h, max_id, capi._iptr(counts), capi._iptr(first),
                                           ctypes.byref(above)))
    return counts, first, above.value


def test_noalloc_volume_equals_allocating_volume_on_two_calls():
    big = cases.make_line_case("noalloc-big", n=(64, 64, 48), grains=120, K=8, seed=21, z0_cells=30)
    small = cases.make_line_case("noalloc-small", n=(24, 16, 20), grains=9, K=5, seed=22, z0_cells=10)
    for c, steps, max_id in ((big, 30, 120), (small, 10, 5)):
        ctx = gpu_context(c)
        ctx.step(steps, 1)
        ctx.collect_on_device(0, capi.PF_COLLECT_CLEAR_ALPHA)
        ref = _call(ctx, "pf_qoi_volume", max_id)
        got = _call(ctx, "pf_qoi_volume_noalloc", max_id)
        print(f"[noalloc] {c.name}: max_id {max_id}, grains {int((ref[0] > 0).sum())}, above {ref[2]}", flush=True)
        assert np.array_equal(got[0], ref[0])
        assert np.array_equal(got[1], ref[1])
        assert got[2] == ref[2]
        assert ref[0].sum() > 0
        if c is small:
            assert ref[2] > 0  # ids above max_id counted, not indexed
        ctx.close()


def test_noalloc_volume_rejects_ids_beyond_its_array():
    c = cases.make_line_case("noalloc-limit", n=(12, 12, 12), grains=3, K=5, seed=3, z0_cells=6)
    ctx = gpu_context(c)
    ctx.collect_on_device(0, capi.PF_COLLECT_CLEAR_ALPHA)
    limit = 1048575
    with pytest.raises(capi.PFError, match="PF_NOALLOC_MAX_ID"):
        _call(ctx, "pf_qoi_volume_noalloc", limit + 1)
    counts, _, above = _call(ctx, "pf_qoi_volume_noalloc", limit)
    assert counts.sum() > 0 and above == 0
    ctx.close()
