"""The timed and profiled stepping entry points (pf_step_timed,
pf_step_profiled) compute exactly what pf_step computes: the events only
bracket the launches.  A dense substrate (many small grains) so that the
start-up overflows K and the fix-up runs, as in the benchmark's warm-up."""
import numpy as np
import pytest

from cupf_amd import capi, cases
from tests.helpers import gpu_context

pytestmark = pytest.mark.gpu


def _same(c1, c2):
    for w in (0, 1):
        p1, a1 = c1.download_state(w)
        p2, a2 = c2.download_state(w)
        assert np.array_equal(a1, a2)
        assert np.array_equal(p1.view(np.uint32), p2.view(np.uint32))


def test_timed_and_profiled_steps_equal_plain_steps():
    c = cases.make_line_case("timing", n=(70, 40, 48), grains=200, K=8, seed=21, G=1.0, R=0.1,
                             z0_cells=12)
    plain, timed, prof = gpu_context(c), gpu_context(c), gpu_context(c)
    plain.step(60, 1)
    ms = timed.step_timed(60, 1)
    m, f, pl = prof.step_profiled(60, 1)
    assert ms > 0.0 and m > 0.0 and f > 0.0 and pl > 0.0
    # events on one stream: the three parts of every step are disjoint spans
    assert m + f + pl < 60 * 50.0
    _same(plain, timed)
    _same(plain, prof)
    assert plain.overflow_count() == timed.overflow_count() == prof.overflow_count()


def test_profiled_rejects_nucleation():
    c = cases.make_line_case("timing-n", n=(16, 16, 16), grains=6, K=8, seed=22)
    ctx = gpu_context(c)
    with pytest.raises(capi.PFError):
        ctx.step_profiled(2, 1, capi.PF_STEP_NUCLEATE)
