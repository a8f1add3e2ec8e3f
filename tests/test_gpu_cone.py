"""All-liquid start under the cone melt-pool field with nucleation (the 256^3
cone configuration): a liquid phase with uniform cooling and nucleation (the
reference's preMt phase, 3DgrainMPI/APTPhaseField.cu:779-816), then the cone
field (thermal type 3, whose `continue` leaves points ahead of the tip
untouched).  Against the oracle at 48^3; at 256^3 the skipping path against
every plane processed, bit for bit."""
import os

import numpy as np
import pytest

from cupf_amd import capi, cases
from tests.helpers import compare_states, gpu_context, oracle_sim
from tests.test_gpu_activity import _check_cover, _dense_context, _same_state

pytestmark = pytest.mark.gpu

PHASE1 = capi.PF_STEP_INITIAL_UC | capi.PF_STEP_NUCLEATE


def _liquid_cone(n):
    c = cases.make_cone_case("cone", n=n, seed=9)
    c.params.update(uc0=np.float32(8.0), nuc_Nmax0=np.float32(5.0))
    return c


def test_cone_nucleation_against_oracle():
    c = _liquid_cone((48, 48, 48))
    ctx = gpu_context(c)
    sim = oracle_sim(c)
    ctx.nucl_reset(0)
    sim.nucl_reset(0)
    for s in (1, 11):
        ctx.step(10, s, PHASE1)
        sim.step(10, s, use_initial_uc=True, nucleate=True, seed=int(c.params["seed"]))
    assert sim.nucl_events > 0 and ctx.nucl_events() == sim.nucl_events
    ctx.step(20, 21)
    sim.step(20, 21)
    bad, dphi = compare_states(ctx, sim)
    assert bad == 0 and dphi <= 1e-5


def test_cone_256_skip_matches_dense():
    c = _liquid_cone((256, 256, 256))
    ctx = gpu_context(c)
    dense = _dense_context(c)
    for x in (ctx, dense):
        x.nucl_reset(0)
        x.step(40, 1, PHASE1)
        x.step(20, 41)
    assert ctx.nucl_events() > 0 and ctx.nucl_events() == dense.nucl_events()
    _same_state(ctx, dense)
    _check_cover(ctx, c)
