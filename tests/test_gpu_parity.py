"""HIP path vs CPU oracle (float32 restatement) on identical seeded inputs.

Bar: grain ids (arg) bit-exact everywhere; phi bit-exact as well, since both
sides evaluate the same float operations in the same order (-ffp-contract=off,
correctly rounded division/sqrt); the stated tolerance is 1e-5 L-inf.
"""
import numpy as np
import pytest

from cupf_amd import capi, cases
from tests.helpers import compare_states, first_mismatch, gpu_context, oracle_sim

pytestmark = pytest.mark.gpu
PHI_TOL = 1e-5


def _run_pair(case, nsteps, flags=0, chunk=None, psi_mac=None, T3d=None, nucleate=False):
    ctx = gpu_context(case, psi_mac=psi_mac, T3d=T3d)
    sim = oracle_sim(case, psi_mac=psi_mac, T3d=T3d)
    chunk = chunk or nsteps
    s = 1
    while s <= nsteps:
        n = min(chunk, nsteps - s + 1)
        ctx.step(n, s, flags)
        sim.step(n, s, use_initial_uc=bool(flags & capi.PF_STEP_INITIAL_UC), nucleate=nucleate,
                 seed=int(case.params["seed"]))
        s += n
    return ctx, sim


@pytest.mark.parametrize("K", [5, 8])
def test_step_small_bitexact(K):
    c = cases.make_line_case("small", n=(16, 16, 16), grains=5, K=K, seed=3, z0_cells=8)
    ctx, sim = _run_pair(c, 20)
    bad, dphi = compare_states(ctx, sim)
    assert bad == 0, first_mismatch(ctx, sim, K, c.params["fnx"], c.params["fny"], c.params["fnz"])
    assert dphi <= PHI_TOL
    assert dphi == 0.0
    # the other buffer (previous step) must agree on the interior as well
    ph_g, arg_g = ctx.download_state(1)
    ph_o, arg_o = sim.state(1)
    fn = (c.params["fnz"], c.params["fny"], c.params["fnx"])
    gi = arg_g.reshape(K, *fn)[:, 1:-1, 1:-1, 1:-1]
    oi = arg_o.reshape(K, *fn)[:, 1:-1, 1:-1, 1:-1]
    assert np.array_equal(gi, oi)


@pytest.mark.parametrize("shape,K,g27,bc", [
    ((127, 7, 10), 8, 0, (1, 1, 0)),    # 2x2 full 64x4 tiles, no edge columns
    ((130, 9, 12), 5, 0, (0, 1, 0)),    # tiles plus leftover columns in x and rows in y
    ((63, 12, 40), 8, 0, (1, 0, 1)),    # one tile column, leftover row, two z chunks
    ((130, 9, 12), 8, 1, (1, 1, 0)),    # 27-point gather through the LDS tiles
    ((95, 10, 20), 8, 0, (1, 1, 0)),    # one full tile and a 32-column partial tile in x
    ((32, 32, 16), 5, 1, (0, 1, 0)),    # a single 33-column partial tile, 27-point gather
])
def test_tiled_path_shapes(shape, K, g27, bc):
    c = cases.make_line_case("tiles", n=shape, grains=12, K=K, seed=43, z0_cells=shape[2] // 2,
                             bc=bc, gather27=g27,
                             thermal_type=capi.THERMAL_LINE_3DGRAIN if g27 else capi.THERMAL_LINE)
    ctx, sim = _run_pair(c, 12)
    bad, dphi = compare_states(ctx, sim)
    assert bad == 0, first_mismatch(ctx, sim, K, c.params["fnx"], c.params["fny"], c.params["fnz"])
    assert dphi == 0.0


@pytest.mark.parametrize("bc", [(1, 1, 0), (0, 0, 0), (1, 0, 1)])
def test_step_boundary_conditions(bc):
    c = cases.make_line_case("bc", n=(12, 14, 10), grains=4, K=5, seed=7, z0_cells=5, bc=bc)
    ctx, sim = _run_pair(c, 10)
    bad, dphi = compare_states(ctx, sim)
    assert bad == 0 and dphi == 0.0


def test_c1_parity_through_overflow_transient():
    """64^3, K=8, ~100 grains: the first 200 steps carry thousands of
    neighbourhoods with more than K ids, so the overflow policy is exercised."""
    c = cases.c1(K=8)
    ctx, sim = _run_pair(c, 200, chunk=50)
    bad, dphi = compare_states(ctx, sim)
    assert sim.overflow > 0
    assert ctx.overflow_count() == sim.overflow
    assert bad == 0, first_mismatch(ctx, sim, 8, 67, 67, 67)
    assert dphi <= PHI_TOL


@pytest.mark.parametrize("K", [5, 8])
def test_gather27_variant(K):
    c = cases.make_line_case("g27", n=(14, 14, 14), grains=5, K=K, seed=5, z0_cells=7, gather27=1,
                             thermal_type=capi.THERMAL_LINE_3DGRAIN)
    ctx, sim = _run_pair(c, 12)
    bad, dphi = compare_states(ctx, sim)
    assert bad == 0, first_mismatch(ctx, sim, K, c.params["fnx"], c.params["fny"], c.params["fnz"])
    assert dphi == 0.0


def _with(case, **kw):
    case.params.update(kw)
    return case


def test_thermal_cylinder():
    c = cases.make_line_case("cyl", n=(12, 16, 12), grains=4, K=5, seed=11, z0_cells=6,
                             thermal_type=capi.THERMAL_CYLINDER)
    _with(c, r0=np.float32(0.6), G=np.float32(2.0))
    ctx, sim = _run_pair(c, 10)
    assert compare_states(ctx, sim) == (0, 0.0)


def test_thermal_cone_skips_slots():
    """Type 3 `continue`s past slots with x beyond the moving cone tip, leaving
    the destination buffer as it was two steps before."""
    c = cases.make_line_case("cone", n=(20, 12, 12), grains=4, K=5, seed=13, z0_cells=6,
                             thermal_type=capi.THERMAL_CONE)
    ang = np.float32(0.5)
    _with(c, r0=np.float32(1.2), V=np.float32(2.0), tan_angle=np.float32(np.tan(ang)),
          cos_angle=np.float32(np.cos(ang)), G=np.float32(3.0))
    ctx, sim = _run_pair(c, 10)
    assert compare_states(ctx, sim) == (0, 0.0)


def test_thermal_paraboloid_macro_field():
    c = cases.make_line_case("parab", n=(12, 12, 12), grains=4, K=5, seed=17, z0_cells=6,
                             thermal_type=capi.THERMAL_PARABOLOID)
    Nx, Ny, Nz = 7, 6, 8
    rng = np.random.default_rng(1)
    psi_mac = rng.standard_normal(Nx * Ny * Nz).astype(np.float32) * 0.3
    _with(c, Nx=Nx, Ny=Ny, Nz=Nz, X0_mac=np.float32(-0.1), Y0_mac=np.float32(-0.1),
          Z0_mac=np.float32(-0.1), Dx_mac=np.float32(0.2), mp_len=np.float32(0.6),
          x0=np.float32(0.1), V=np.float32(1.5), G=np.float32(2.0), order=np.float32(2.0))
    ctx, sim = _run_pair(c, 8, psi_mac=psi_mac)
    assert compare_states(ctx, sim) == (0, 0.0)


def test_initial_undercooling_flag():
    c = cases.make_line_case("iuc", n=(12, 12, 12), grains=4, K=5, seed=19, z0_cells=6)
    ctx, sim = _run_pair(c, 10, flags=capi.PF_STEP_INITIAL_UC)
    assert compare_states(ctx, sim) == (0, 0.0)


def test_reference_angle_orientations():
    c = cases.make_line_case("angles", n=(16, 16, 12), grains=6, K=8, seed=23, z0_cells=6,
                             orient="angles")
    ctx, sim = _run_pair(c, 15)
    assert compare_states(ctx, sim) == (0, 0.0)


@pytest.mark.parametrize("clear", [0, 1])
def test_collect(clear):
    c = cases.make_line_case("collect", n=(12, 12, 12), grains=4, K=5, seed=29, z0_cells=6)
    ctx, sim = _run_pair(c, 6)
    phi_g, al_g = ctx.collect(0, capi.PF_COLLECT_CLEAR_ALPHA if clear else 0)
    phi_o, al_o = sim.collect(0, clear_alpha=bool(clear))
    assert np.array_equal(al_g, al_o)
    assert np.array_equal(phi_g, phi_o.astype(np.float32))


def test_moving_domain_shift():
    c = cases.make_line_case("md", n=(12, 12, 24), grains=4, K=5, seed=31, z0_cells=12,
                             line_config=True, nz_full=48)
    ctx, sim = _run_pair(c, 6)
    ctx.md_enable(c.d["fnz_f"])
    sim.md_enable(c.d["fnz_f"])
    tip_g = ctx.md_tip(1)
    tip_o = sim.md_tip(1)
    assert tip_g == tip_o
    assert 1 < tip_o < c.params["fnz"] - 1
    for mc in range(3):
        ctx.md_shift(mc)
        sim.md_shift(mc)
    af_g, loss_g, z_g = ctx.md_download()
    assert np.array_equal(af_g, sim.alpha_full)
    assert np.array_equal(loss_g, sim.loss)
    assert np.array_equal(z_g, sim.z)
    ph_g, arg_g = ctx.download_state(0)
    ph_o, arg_o = sim.state(0)
    assert np.array_equal(arg_g, arg_o) and np.array_equal(ph_g, ph_o)
    # stepping continues identically on the shifted window (z moved on the device too)
    ctx.step(4, 7)
    sim.step(4, 7)
    assert compare_states(ctx, sim) == (0, 0.0)


@pytest.mark.parametrize("z0,start", [(30, 1), (30, 3), (30, 29), (8, 40), (8, 47), (8, 48)])
def test_tip_scan_windows(z0, start):
    """The device tip scan sums planes 16 at a time as it reaches them: a tip
    more than one window above the start, one inside the first window, a
    start above the interface and at the top plane give the oracle's tip."""
    c = cases.make_line_case("tipw", n=(12, 12, 46), grains=4, K=5, seed=37, z0_cells=z0,
                             line_config=True, nz_full=96)
    ctx, sim = _run_pair(c, 4)
    tip_g = ctx.md_tip(start)
    tip_o = sim.md_tip(start)
    print(f"[tip] z0 {z0} start {start}: tip {tip_o}", flush=True)
    assert tip_g == tip_o
    if start == 1 and z0 == 30:
        assert tip_o > 17  # crossed into the second window


def test_nucleation_initial_undercooling():
    """Liquid start with uniform cooling: add_nucl after odd steps with the
    counter-based draw.  Decisions match unless a draw lands within an ulp of
    the probability (device vs host expf); the state must agree."""
    c = cases.make_line_case("nucl", n=(32, 32, 32), grains=3, K=5, seed=37, z0_cells=-20)
    _with(c, uc0=np.float32(8.0), nuc_Nmax0=np.float32(5.0))
    ctx = gpu_context(c)
    sim = oracle_sim(c)
    ctx.nucl_reset(0)
    sim.nucl_reset(0)
    for s in range(1, 41, 10):
        ctx.step(10, s, capi.PF_STEP_INITIAL_UC | capi.PF_STEP_NUCLEATE)
        sim.step(10, s, use_initial_uc=True, nucleate=True, seed=int(c.params["seed"]))
    assert sim.nucl_events > 0
    assert ctx.nucl_events() == sim.nucl_events
    bad, dphi = compare_states(ctx, sim)
    assert bad == 0 and dphi <= PHI_TOL


def test_upload_download_roundtrip():
    c = cases.make_line_case("rt", n=(10, 10, 10), grains=3, K=8, seed=41, z0_cells=5)
    ctx = gpu_context(c)
    rng = np.random.default_rng(0)
    n = 8 * c.L
    ph = rng.standard_normal(n).astype(np.float32)
    arg = rng.integers(-1, 5, n).astype(np.int32)
    ctx.upload_state(ph, arg, 1)
    ph2, arg2 = ctx.download_state(1)
    assert np.array_equal(ph, ph2) and np.array_equal(arg, arg2)
