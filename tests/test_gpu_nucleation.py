"""Main-loop nucleation driven by the macro temperature field T_3D
(3DgrainMPI/APTPhaseField.cu:486-496 with interp4Dtemperature,
devicefunc.cu_inl:32-83, run after every odd step as in :866-870): the HIP
path against the oracle on the same seeded inputs.

Decisions use the same counter-based draw on both sides, so the events and
the state agree exactly unless a draw lands within an ulp of the probability
(device vs host expf); the state bar is bit-exact ids and 1e-5 L-inf phi.
"""
import numpy as np
import pytest

from cupf_amd import capi, cases
from tests.helpers import compare_states, gpu_context, oracle_sim

pytestmark = pytest.mark.gpu
PHI_TOL = 1e-5


def _macro_T(c, a=0.5, b=0.4, rate=2.0e7, Nt=4, Dt=1.0e-7):
    """T(x, t) = -(a + b (x - X0)) - rate (t - t0) on a 4^3 x Nt macro grid
    covering the domain: the undercooling grows in time and along x."""
    Nx = Ny = Nz = 4
    X0 = float(c.x[0]) - 0.05
    Dx = (float(c.x[-1]) + 0.1 - X0) / (Nx - 1)
    kx = np.arange(Nx)
    frames = []
    for kt in range(Nt):
        plane = -(a + b * kx * Dx) - rate * kt * Dt
        frames.append(np.broadcast_to(plane, (Nz, Ny, Nx)))
    T3d = np.ascontiguousarray(np.stack(frames).ravel(), np.float32)
    c.params.update(Nx=Nx, Ny=Ny, Nz=Nz, Nt=Nt, X0_mac=np.float32(X0), Y0_mac=np.float32(X0),
                    Z0_mac=np.float32(X0), Dx_mac=np.float32(Dx), t0_mac=np.float32(0.0),
                    Dt_mac=np.float32(Dt))
    return T3d


@pytest.mark.parametrize("K,pts_cell", [(5, 1), (8, 2)])
def test_main_loop_nucleation_from_the_macro_field(K, pts_cell):
    # all liquid at the start; the line law with z0 above the box keeps every
    # point undercooled so the nuclei grow and meet
    c = cases.make_line_case("nucl3d", n=(32, 32, 32), grains=3, K=K, seed=53, z0_cells=-20)
    c.params.update(pts_cell=pts_cell, nuc_Nmax=np.float32(1.5), z0=np.float32(3.0))
    T3d = _macro_T(c)
    ctx = gpu_context(c, T3d=T3d)
    sim = oracle_sim(c, T3d=T3d)
    ctx.nucl_reset(0)
    sim.nucl_reset(0)
    seed = int(c.params["seed"])
    for s in range(1, 49, 8):
        ctx.step(8, s, capi.PF_STEP_NUCLEATE)
        sim.step(8, s, nucleate=True, seed=seed)
        assert ctx.nucl_events() == sim.nucl_events, f"events after step {s + 7}"
    assert sim.nucl_events >= 10
    bad, dphi = compare_states(ctx, sim)
    assert bad == 0 and dphi <= PHI_TOL
    assert ctx.overflow_count() == sim.overflow


def test_main_loop_nucleation_needs_the_macro_field():
    c = cases.make_line_case("nucl3d-missing", n=(12, 12, 12), grains=3, K=5, seed=3, z0_cells=-20)
    ctx = gpu_context(c)
    ctx.nucl_reset(0)
    with pytest.raises(RuntimeError, match="T3d"):
        ctx.step(2, 1, capi.PF_STEP_NUCLEATE)
