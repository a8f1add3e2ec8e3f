"""Pins of the CPU oracle (no GPU).

The reference ships no tests, fixtures or golden vectors for this path, so the
oracle is pinned by closed forms, brute force on tiny grids, symmetries and an
independent dense multi-field restatement (numpy), each chosen so that a
plausible slip (dropped term, wrong sign, wrong neighbour, wrong slot order,
missing double promotion) fails at least one of them.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle
from cupf_amd import cases

f32 = np.float32
LS = -0.995


def _params(fn=(7, 7, 7), K=5, **kw):
    d = dict(fnx=fn[0], fny=fn[1], fnz=fn[2], K=K, num_theta=8, bcx=1, bcy=1, bcz=1,
             thermal_type=1, gather27=0, dt=f32(0.03), hi=f32(1.25), lamd=f32(58.3),
             L_cp=f32(229.0), kin_delta=f32(0.11), eps=f32(1e-8), G=f32(0.0), R=f32(0.0),
             uc=f32(0.0), uc0=f32(0.0), z0=f32(0.0), r0=f32(0.0), lyd=f32(1.0), lzd=f32(1.0),
             V=f32(0.0), mp_len=f32(0.0), x0=f32(0.0), order=f32(1.0), tan_angle=f32(1.0),
             cos_angle=f32(1.0), Nx=2, Ny=2, Nz=2, Nt=2, X0_mac=f32(0), Y0_mac=f32(0), Z0_mac=f32(0),
             t0_mac=f32(0), Dx_mac=f32(1), Dt_mac=f32(1), pts_cell=1, nuc_rad=f32(0.4),
             undcool_mean=f32(2.0), undcool_std=f32(0.5), nuc_Nmax=f32(0.01), nuc_Nmax0=f32(0.01),
             dx=f32(0.8), sqrt2=f32(math.sqrt(2.0)), tau0=f32(1.328e-7))
    d.update(kw)
    return d


def _ident(n):
    R = np.zeros((n + 1, 9), np.float32)
    R[:, [0, 4, 8]] = 1
    return R


def _coords(p):
    return (np.arange(p["fnx"], dtype=np.float32) * f32(0.08), np.arange(p["fny"], dtype=np.float32) * f32(0.08),
            np.arange(p["fnz"], dtype=np.float32) * f32(0.08))


def _sim(p, precision="f64", R9=None):
    x, y, z = _coords(p)
    return oracle.OracleSim(p, x, y, z, _ident(p["num_theta"]) if R9 is None else R9, precision)


def _set(sim, ids, phs):
    """ids/phs: (K, fnz, fny, fnx) arrays."""
    sim.set_state(np.asarray(phs, sim.dtype).ravel(), np.asarray(ids, np.int32).ravel())


def _kine_ref(u, R, kd):
    """kine_ani in float64 from its definition (cubic anisotropy in the rotated frame)."""
    a_s = 1 + 3 * kd
    eps = -4 * kd / a_s
    v = R.reshape(3, 3).astype(np.float64) @ np.asarray(u, np.float64)
    m2 = float(v @ v)
    if m2 <= 1e-8:
        return 1.0
    return a_s * (1 + eps * float(np.sum(v ** 4)) / m2 ** 2)


# --------------------------------------------------------------- kine_ani ---
def test_kine_ani_principal_axes():
    p = oracle.OrcParams.from_dict(_params())
    lib = oracle.load("f64")
    I = np.eye(3, dtype=np.float32).ravel()
    kd = float(np.float32(0.11))
    a_s = 1 + 3 * kd
    eps = -4 * kd / a_s
    Rp = I.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    assert lib.orc_kine_ani(p, 1.0, 0.0, 0.0, Rp) == pytest.approx(a_s * (1 + eps), rel=1e-7)
    s = 1 / math.sqrt(3)
    assert lib.orc_kine_ani(p, s, s, s, Rp) == pytest.approx(a_s * (1 + eps / 3), rel=1e-7)
    assert lib.orc_kine_ani(p, 1e-5, 0.0, 0.0, Rp) == 1.0  # |grad|^2 <= eps


def test_kine_ani_rotated_frame_matches_definition():
    lib = oracle.load("f64")
    p = oracle.OrcParams.from_dict(_params())
    rng = np.random.default_rng(5)
    R9 = cases.quaternion_rotations(20, 3)[1:]
    for r in R9:
        u = rng.standard_normal(3)
        got = lib.orc_kine_ani(p, *u, r.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        assert got == pytest.approx(_kine_ref(u, r, float(np.float32(0.11))), rel=1e-6)


def test_kine_ani_matrix_form_bitexact_with_angle_form():
    """The reference evaluates cosa*cosb*ux + sina*cosb*uy - sinb*uz etc.
    (3DgrainMPI/APTPhaseField.cu:41-46); the R9 rows must reproduce it bit for bit in float32."""
    lib = oracle.load("f32")
    p = oracle.OrcParams.from_dict(_params())
    rng = np.random.default_rng(7)
    for _ in range(300):
        a, b = rng.uniform(0, math.pi, 2).astype(np.float32)
        sa, ca, sb, cb = np.sin(a), np.cos(a), np.sin(b), np.cos(b)
        R = np.array([ca * cb, sa * cb, -sb, -sa, ca, 0, ca * sb, sa * sb, cb], np.float32)
        ux, uy, uz = rng.standard_normal(3).astype(np.float32)
        kd = f32(0.11)
        a_s = f32(f32(1) + f32(3) * kd)
        epsv = f32(f32(f32(-4) * kd) / a_s)
        x2 = f32(f32(f32(ca * cb) * ux + f32(sa * cb) * uy) - f32(sb * uz))
        y2 = f32(f32(f32(-sa) * ux + f32(ca * uy)) + f32(0))
        z2 = f32(f32(f32(ca * sb) * ux + f32(sa * sb) * uy) + f32(cb * uz))
        x2, y2, z2 = x2 * x2, y2 * y2, z2 * z2
        mag = f32(f32(x2 + y2) + z2)
        want = f32(a_s * f32(f32(1) + f32(f32(epsv * f32(f32(x2 * x2 + y2 * y2) + z2 * z2)) / f32(mag * mag))))
        got = lib.orc_kine_ani(p, float(ux), float(uy), float(uz), R.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        assert np.float32(got) == want


# ------------------------------------------------------ single-point answers ---
def _uniform_state(p, slots):
    """Every point holds the same (id, phi) list in slots 0..len-1."""
    K = p["K"]
    shape = (K, p["fnz"], p["fny"], p["fnx"])
    ids = np.full(shape, -1, np.int32)
    phs = np.full(shape, -1.0)
    for q, (g, v) in enumerate(slots):
        ids[q] = g
        phs[q] = v
    return ids, phs


def _T_line(p, z, t):
    return float(p["G"]) * (float(z) - float(p["z0"])) - float(p["uc"]) * 1e6 * t


@pytest.mark.parametrize("phi0", [1.0, 0.3, -0.4])
def test_uniform_single_grain_closed_form(phi0):
    """diff = 0, one slot: phi' = phi + dt*((1-phi^2)phi - lamd*Up*(1-phi^2)^2)/A2 with A2 = 1
    (zero gradient)."""
    p = _params(G=f32(2.0), uc=f32(0.05), z0=f32(0.2))
    sim = _sim(p)
    _set(sim, *_uniform_state(p, [(3, phi0)]))
    t = 0.7e-6
    sim.lib.orc_rhs(sim.P, oracle._f(sim.x), oracle._f(sim.y), oracle._f(sim.z), oracle._f(sim.R9), None,
                    oracle._v(sim.ph[0]), oracle._i(sim.arg[0]), oracle._v(sim.ph[1]), oracle._i(sim.arg[1]), t, 0)
    out = sim.ph[1].reshape(p["K"], p["fnz"], p["fny"], p["fnx"])
    for k in range(1, p["fnz"] - 1):
        Up = _T_line(p, sim.z[k], t) / float(p["L_cp"])
        om = 1 - phi0 ** 2
        want = phi0 + float(p["dt"]) * (om * phi0 - float(p["lamd"]) * Up * om * om)
        assert out[0, k, 3, 3] == pytest.approx(want, abs=1e-12)
    assert (sim.arg[1].reshape(out.shape)[0, 1:-1, 1:-1, 1:-1] == 3).all()
    assert (sim.arg[1].reshape(out.shape)[1:, 1:-1, 1:-1, 1:-1] == -1).all()


def test_uniform_two_grains_repulsion_closed_form():
    a, b = 0.6, -0.2
    p = _params(G=f32(0.0), uc=f32(0.1))
    sim = _sim(p)
    _set(sim, *_uniform_state(p, [(2, a), (5, b)]))
    t = 2e-6
    sim.lib.orc_rhs(sim.P, oracle._f(sim.x), oracle._f(sim.y), oracle._f(sim.z), oracle._f(sim.R9), None,
                    oracle._v(sim.ph[0]), oracle._i(sim.arg[0]), oracle._v(sim.ph[1]), oracle._i(sim.arg[1]), t, 0)
    out = sim.ph[1].reshape(p["K"], p["fnz"], p["fny"], p["fnx"])
    Up = -float(p["uc"]) * 1e6 * t / float(p["L_cp"])
    lam, dt = float(p["lamd"]), float(p["dt"])

    def upd(u, o):
        om = 1 - u * u
        return u + dt * (om * u - lam * Up * (om * om - 6 * (u + 1) * 0.25 * (o + 1) ** 2))

    assert out[0, 3, 3, 3] == pytest.approx(upd(a, b), abs=1e-12)
    assert out[1, 3, 3, 3] == pytest.approx(upd(b, a), abs=1e-12)


def test_bulk_solid_and_liquid_are_fixed_points():
    """phi = 1 everywhere (one grain): diff = 0 and 1 - phi^2 = 0, so every term vanishes
    exactly whatever the temperature; an empty neighbourhood stays empty."""
    p = _params(G=f32(3.0), uc=f32(0.2), K=8, z0=f32(0.3))
    sim = _sim(p, "f32")
    _set(sim, *_uniform_state(p, [(4, 1.0)]))
    sim.step(3, 1)
    ph, arg = sim.state()
    ph = ph.reshape(8, 7, 7, 7)[:, 1:-1, 1:-1, 1:-1]
    arg = arg.reshape(8, 7, 7, 7)[:, 1:-1, 1:-1, 1:-1]
    assert (ph[0] == 1.0).all() and (arg[0] == 4).all()
    assert (arg[1:] == -1).all() and (ph[1:] == -1.0).all()
    liq = _sim(p, "f32")
    _set(liq, *_uniform_state(p, []))
    liq.step(3, 1)
    ph, arg = liq.state()
    assert (arg == -1).all() and (ph == -1.0).all()


def test_anisotropic_gradient_closed_form():
    """phi varies along x only: Laplacian and A2 = a_s(1+eps) with the identity rotation."""
    p = _params(fn=(9, 5, 5), K=5)
    sim = _sim(p)
    K = 5
    ids = np.full((K, 5, 5, 9), -1, np.int32)
    phs = np.full((K, 5, 5, 9), -1.0)
    prof = np.array([0.1, 0.3, 0.45, 0.5, 0.4, 0.2, 0.0, -0.1, -0.3])
    ids[0] = 1
    phs[0] = prof[None, None, :]
    _set(sim, ids, phs)
    sim.lib.orc_rhs(sim.P, oracle._f(sim.x), oracle._f(sim.y), oracle._f(sim.z), oracle._f(sim.R9), None,
                    oracle._v(sim.ph[0]), oracle._i(sim.arg[0]), oracle._v(sim.ph[1]), oracle._i(sim.arg[1]), 0.0, 0)
    out = sim.ph[1].reshape(K, 5, 5, 9)
    kd = float(np.float32(0.11))
    a_s = 1 + 3 * kd
    A2 = a_s * (1 - 4 * kd / a_s)
    for i in range(1, 8):
        lap = prof[i + 1] + prof[i - 1] - 2 * prof[i]
        c = prof[i]
        want = c + float(p["dt"]) * (lap * 1.25 * 1.25 + (1 - c * c) * c) / A2
        assert out[0, 2, 2, i] == pytest.approx(want, abs=1e-12)


# ------------------------------------------------------------- slot order ---
def _one_point_case(K, entries):
    """A 3x3x3 interior, periodic; `entries` maps stencil name -> list of (slot, id, phi)
    placed at the neighbours of the centre point (2,2,2)."""
    p = _params(fn=(5, 5, 5), K=K, bcx=0, bcy=0, bcz=0)
    ids = np.full((K, 5, 5, 5), -1, np.int32)
    phs = np.full((K, 5, 5, 5), -1.0)
    off = {"D": (-1, 0, 0), "B": (0, -1, 0), "L": (0, 0, -1), "C": (0, 0, 0), "R": (0, 0, 1),
           "T": (0, 1, 0), "U": (1, 0, 0)}
    for name, lst in entries.items():
        dk, dj, di = off[name]
        for slot, g, v in lst:
            ids[slot, 2 + dk, 2 + dj, 2 + di] = g
            phs[slot, 2 + dk, 2 + dj, 2 + di] = v
    sim = _sim(p)
    _set(sim, ids, phs)
    sim.lib.orc_rhs(sim.P, oracle._f(sim.x), oracle._f(sim.y), oracle._f(sim.z), oracle._f(sim.R9), None,
                    oracle._v(sim.ph[0]), oracle._i(sim.arg[0]), oracle._v(sim.ph[1]), oracle._i(sim.arg[1]), 0.0, 0)
    arg = sim.arg[1].reshape(K, 5, 5, 5)[:, 2, 2, 2]
    return list(arg)


def test_slot_order_first_appearance():
    # scan order D,B,L,C,R,T,U then slot index; holes (-1) are skipped
    got = _one_point_case(5, {
        "U": [(0, 9, 0.2)],
        "B": [(2, 4, 0.1), (0, 6, 0.3)],
        "D": [(1, 6, 0.2)],
        "C": [(0, 7, 0.5), (3, 4, 0.1)],
        "R": [(0, 8, 0.2)],
    })
    assert got == [6, 4, 7, 8, 9]


def test_slot_order_duplicates_and_holes():
    got = _one_point_case(5, {"C": [(1, 3, 0.4), (3, 3, 0.2)], "L": [(4, 5, 0.1)], "T": [(0, 3, 0.3)]})
    assert got == [5, 3, -1, -1, -1]


def test_overflow_keeps_top_k_by_phi_in_first_appearance_order():
    # 7 distinct ids for K=5: ids with the largest phi anywhere are kept
    got = _one_point_case(5, {
        "D": [(0, 11, -0.9), (1, 12, 0.8)],
        "B": [(0, 13, 0.7)],
        "L": [(0, 14, -0.95)],
        "C": [(0, 15, 0.9), (1, 16, 0.1)],
        "U": [(0, 17, 0.2), (1, 11, 0.95)],
    })
    # max phi: 11:0.95 15:0.9 12:0.8 13:0.7 17:0.2 16:0.1 14:-0.95 -> keep 11,15,12,13,17
    assert got == [11, 12, 13, 15, 17]


def test_overflow_tie_prefers_earlier_entry():
    got = _one_point_case(5, {
        "D": [(0, 21, 0.5), (1, 22, 0.5), (2, 23, 0.5)],
        "B": [(0, 24, 0.5), (1, 25, 0.5), (2, 26, 0.5)],
    })
    assert got == [21, 22, 23, 24, 25]


# -------------------------------------------------------------------- BC ---
@pytest.mark.parametrize("bc", [(0, 0, 0), (1, 1, 1), (1, 0, 1), (0, 1, 0)])
def test_bc_fill_matches_numpy_pad(bc):
    p = _params(fn=(6, 7, 8), K=2, bcx=bc[0], bcy=bc[1], bcz=bc[2])
    sim = _sim(p)
    rng = np.random.default_rng(1)
    ph = rng.standard_normal((2, 8, 7, 6))
    arg = rng.integers(0, 50, (2, 8, 7, 6)).astype(np.int32)
    _set(sim, arg, ph)
    sim.apply_bc()
    got_ph, got_arg = sim.state()
    exp = ph[:, 1:-1, 1:-1, 1:-1]
    expa = arg[:, 1:-1, 1:-1, 1:-1]
    for ax, b in zip((3, 2, 1), bc):
        mode = "wrap" if b else "reflect"
        pad = [(0, 0)] * 4
        pad[ax] = (1, 1)
        exp = np.pad(exp, pad, mode=mode)
        expa = np.pad(expa, pad, mode=mode)
    assert np.array_equal(got_ph.reshape(exp.shape), exp)
    assert np.array_equal(got_arg.reshape(expa.shape), expa)


# --------------------------------------------------- dense restatement -------
def _dense_step(phis, p, t, Rs):
    """Multi-phase-field step over every grain (no slots, no gather), float64,
    periodic in all axes: the APT step must agree wherever no slot is killed."""
    G, z0, uc, Lcp, lam, dt, hi = (float(p[k]) for k in ("G", "z0", "uc", "L_cp", "lamd", "dt", "hi"))
    kd = float(p["kin_delta"])
    a_s = 1 + 3 * kd
    eps = -4 * kd / a_s
    z = _coords(p)[2][1:-1].astype(np.float64)  # interior planes, as stored in float32
    out = {}
    rolls = lambda f, s, ax: np.roll(f, s, axis=ax)  # noqa: E731
    for g, f in phis.items():
        fR, fL = rolls(f, -1, 2), rolls(f, 1, 2)
        fT, fB = rolls(f, -1, 1), rolls(f, 1, 1)
        fU, fD = rolls(f, -1, 0), rolls(f, 1, 0)
        u = np.stack([(fR - fL) * 0.5, (fT - fB) * 0.5, (fU - fD) * 0.5])
        v = np.einsum("ab,bkji->akji", Rs[g].reshape(3, 3).astype(np.float64), u)
        vI = u
        sel = f > LS
        vv = np.where(sel[None], v, vI)
        m2 = (vv ** 2).sum(0)
        A2 = np.where(m2 > 1e-8, a_s * (1 + eps * (vv ** 4).sum(0) / np.where(m2 > 0, m2, 1) ** 2), 1.0)
        lap = fR + fL + fT + fB + fU + fD - 6 * f
        T = G * (z[:, None, None] - z0) - uc * 1e6 * t
        Up = T / Lcp
        rep = sum(0.25 * (h + 1) ** 2 for gg, h in phis.items() if gg != g)
        om = 1 - f * f
        out[g] = f + dt * (lap * hi * hi + om * f - lam * Up * (om * om - 6 * (f + 1) * rep)) / A2
    return out


def test_apt_matches_dense_multifield():
    n = 10
    p = _params(fn=(n + 2, n + 2, n + 2), K=8, G=f32(0.5), uc=f32(0.02), z0=f32(0.4), num_theta=3)
    rng = np.random.default_rng(3)
    kk, jj, ii = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    # three smooth periodic grain fields; values stay above the kill threshold
    fields = {
        1: 0.8 * np.sin(2 * np.pi * ii / n) * np.cos(2 * np.pi * jj / n) + 0.1,
        2: 0.6 * np.cos(2 * np.pi * (ii + kk) / n) - 0.2,
        3: 0.5 * np.sin(2 * np.pi * (jj - kk) / n),
    }
    Rs = cases.quaternion_rotations(3, 11)
    sim = _sim(p, "f64", R9=Rs)
    K = 8
    ids = np.full((K, n + 2, n + 2, n + 2), -1, np.int32)
    phs = np.full((K, n + 2, n + 2, n + 2), -1.0)
    for q, (g, f) in enumerate(fields.items()):
        ids[q, 1:-1, 1:-1, 1:-1] = g
        phs[q, 1:-1, 1:-1, 1:-1] = f
    _set(sim, ids, phs)
    t = 1e-6
    sim.apply_bc()
    sim.lib.orc_rhs(sim.P, oracle._f(sim.x), oracle._f(sim.y), oracle._f(sim.z), oracle._f(sim.R9), None,
                    oracle._v(sim.ph[0]), oracle._i(sim.arg[0]), oracle._v(sim.ph[1]), oracle._i(sim.arg[1]), t, 0)
    dense = _dense_step(fields, p, t, Rs)
    got_ph = sim.ph[1].reshape(K, n + 2, n + 2, n + 2)[:, 1:-1, 1:-1, 1:-1]
    got_id = sim.arg[1].reshape(K, n + 2, n + 2, n + 2)[:, 1:-1, 1:-1, 1:-1]
    for q, g in enumerate(fields):
        assert (got_id[q] == g).all()
        assert np.abs(got_ph[q] - dense[g]).max() < 1e-12


# ---------------------------------------------------------------- symmetry ---
def test_translation_invariance_periodic():
    c = cases.make_line_case("tr", n=(11, 11, 11), grains=5, K=5, seed=2, z0_cells=5, bc=(1, 1, 1))
    c.params.update(G=np.float32(0.0))
    a = oracle.OracleSim(c.params, c.x, c.y, c.z, c.R9, "f32")
    a.init_from_phi_alpha(c.phi0, c.alpha0)
    a.step(3, 1)
    ph, arg = a.state()
    fn = (5, 14, 14, 14)
    roll = lambda v: np.roll(v.reshape(fn)[:, 1:-1, 1:-1, 1:-1], 3, axis=3)  # noqa: E731
    b = oracle.OracleSim(c.params, c.x, c.y, c.z, c.R9, "f32")
    shp = np.full(fn, -1.0, np.float32)
    sid = np.full(fn, -1, np.int32)
    shp[:, 1:-1, 1:-1, 1:-1] = roll(ph)
    sid[:, 1:-1, 1:-1, 1:-1] = roll(arg)
    b.set_state(shp.ravel(), sid.ravel())
    b.step(2, 4)
    a.step(2, 4)
    pa, ia = a.state()
    pb, ib = b.state()
    assert np.array_equal(roll(pa), pb.reshape(fn)[:, 1:-1, 1:-1, 1:-1])
    assert np.array_equal(roll(ia), ib.reshape(fn)[:, 1:-1, 1:-1, 1:-1])


# ---------------------------------------------------------- thermal field ---
def test_thermal_line_keeps_double_promotion():
    """T = G*(z-z0) - UC*1e6*t with UC*1e6*t in double (3DgrainMPI/APTPhaseField.cu:286).
    Checked at step times where a float-only evaluation would round differently."""
    p = _params(G=f32(1.0), uc=f32(0.1), z0=f32(2.0))
    lib = oracle.load("f32")
    P = oracle.OrcParams.from_dict(p)
    z = np.array([0, 0, 1.7371], np.float32)
    x = np.zeros(3, np.float32)
    out = ctypes.c_double()
    g = np.float32(np.float32(1.0) * np.float32(np.float32(1.7371) - np.float32(2.0)))
    dt, tau0 = f32(0.0363668), f32(1.3280389e-7)
    hits = 0
    for s_ in range(1, 3000):
        t = f32(f32(f32(s_) * dt) * tau0)
        promoted = np.float32(float(g) - float(np.float32(0.1)) * 1e6 * float(t))
        pure32 = np.float32(g - np.float32(np.float32(np.float32(0.1) * np.float32(1e6)) * t))
        if promoted == pure32:
            continue
        lib.orc_thermal(P, oracle._f(x), oracle._f(x), oracle._f(z), None, 0, 0, 2, float(t), 0,
                        ctypes.byref(out))
        assert np.float32(out.value) == promoted
        hits += 1
        if hits == 20:
            break
    assert hits == 20


def test_thermal_cone_skip_beyond_tip():
    p = _params(thermal_type=3, r0=f32(1.2), z0=f32(0.2), V=f32(2.0), tan_angle=f32(0.5), cos_angle=f32(0.9))
    lib = oracle.load("f32")
    P = oracle.OrcParams.from_dict(p)
    x = np.array([0.5, 1.9, 2.1], np.float32)
    out = ctypes.c_double()
    skips = [lib.orc_thermal(P, oracle._f(x), oracle._f(x), oracle._f(x), None, i, 0, 0, 0.0, 0, ctypes.byref(out))
             for i in range(3)]
    lm = (1.2 - 0.2) / 0.5
    assert skips == [int(v > lm) for v in x]


# -------------------------------------------------------- collect / ini ---
def test_collect_tie_takes_lowest_slot_and_ls_threshold():
    p = _params(fn=(4, 4, 4), K=3)
    sim = _sim(p, "f32")
    ids = np.full((3, 4, 4, 4), -1, np.int32)
    phs = np.full((3, 4, 4, 4), -1.0, np.float32)
    ids[0], ids[1], ids[2] = 5, 6, 7
    phs[0], phs[1], phs[2] = 0.3, 0.3, 0.1          # tie -> slot 0
    phs[:, 0, 0, 0] = np.float32(-0.995)            # float(-0.995) < -0.995 in double: not collected
    phs[:, 0, 0, 1] = np.nextafter(np.float32(-0.995), np.float32(0))
    _set(sim, ids, phs)
    sim.alpha[:] = 99
    phi, alpha = sim.collect(bc=False)
    alpha = alpha.reshape(4, 4, 4)
    assert alpha[1, 1, 1] == 5
    assert alpha[0, 0, 0] == 99   # untouched
    assert alpha[0, 0, 1] == 5


def test_ini_threshold_is_double():
    p = _params(fn=(4, 4, 4), K=3)
    sim = _sim(p, "f32")
    phi0 = np.full(64, -1.0, np.float32)
    phi0[0] = np.float32(-0.995)
    phi0[1] = np.nextafter(np.float32(-0.995), np.float32(0))
    alpha0 = np.arange(64, dtype=np.int32)
    sim.init_from_phi_alpha(phi0, alpha0)
    ph, arg = sim.state()
    assert arg[0] == -1 and arg[1] == 1 and (arg[64:] == -1).all()


# -------------------------------------------------------- moving domain ---
def test_move_frame_shift_identity():
    p = _params(fn=(5, 5, 9), K=2, num_theta=6)
    sim = _sim(p, "f32")
    rng = np.random.default_rng(4)
    ph = rng.standard_normal((2, 9, 5, 5)).astype(np.float32)
    arg = rng.integers(-1, 6, (2, 9, 5, 5)).astype(np.int32)
    _set(sim, arg, ph)
    sim.md_enable(12)
    alpha = rng.integers(0, 7, 225).astype(np.int32)
    sim.alpha[:] = alpha
    z_before = sim.z.copy()
    lib = sim.lib
    lib.orc_move_frame(sim.P, oracle._v(sim.ph[0]), oracle._i(sim.arg[0]), oracle._v(sim.ph[1]),
                       oracle._i(sim.arg[1]), oracle._f(sim.z), oracle._i(sim.alpha), oracle._i(sim.alpha_full),
                       oracle._i(sim.loss), 2)
    g = sim.ph[0].reshape(2, 9, 5, 5)
    ga = sim.arg[0].reshape(2, 9, 5, 5)
    assert np.array_equal(g[:, 1:6, 1:-1, 1:-1], ph[:, 2:7, 1:-1, 1:-1])
    assert (g[:, 7, 1:-1, 1:-1] == -1).all() and (ga[:, 7, 1:-1, 1:-1] == -1).all()
    assert np.array_equal(g[:, 8], ph[:, 8]) and np.array_equal(g[:, 0], ph[:, 0])   # halo untouched
    assert np.array_equal(sim.z[1:7], z_before[2:8])
    assert sim.z[7] == np.float32(2 * z_before[7] - z_before[6])
    plane1 = alpha[25:50]
    assert np.array_equal(sim.alpha_full[75:100], plane1)
    want = np.bincount(plane1[(plane1 >= 1) & (plane1 <= 6)] - 1, minlength=6)
    assert np.array_equal(sim.loss, want)


def test_plane_sums_exact_and_tip_scan():
    p = _params(fn=(4, 4, 6), K=2)
    lib = oracle.load("f32")
    P = oracle.OrcParams.from_dict(p)
    ph = np.full((2, 6, 4, 4), -1.0, np.float32)
    ph[:, :3] = 0.5
    ph[0, 3] = np.float32(-0.99)      # mean (-0.99-1)/2 = -0.995 -> not > LS
    sums = np.zeros(6, np.int64)
    lib.orc_plane_sums(P, oracle._v(ph.ravel()), sums.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    assert sums[0] == int(0.5 * 32 * 2 ** 40)
    assert sums[3] == int(round((float(np.float32(-0.99)) - 1.0) * 16 * 2 ** 40))
    tip = lib.orc_tip_scan(P, sums.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 1)
    assert tip == 3


# ------------------------------------------------------------- nucleation ---
def test_nucleation_probability_closed_form():
    lib = oracle.load("f64")
    P = oracle.OrcParams.from_dict(_params())
    for dT, d in [(2.0, 0.1), (1.3, 0.05), (2.7, 0.2)]:
        want = 8 * 0.4 ** 3 * 0.01 / (math.sqrt(2 * math.pi) * 0.5) * math.exp(-0.5 * ((dT - 2) / 0.5) ** 2) * d
        assert lib.orc_nucl_prob(P, dT, d, 0.01) == pytest.approx(want, rel=1e-6)


def test_nucleation_ball_closed_form():
    p = _params(pts_cell=3, dx=f32(0.8))
    lib = oracle.load("f32")
    P = oracle.OrcParams.from_dict(p)
    ball = np.zeros(343, np.float32)
    lib.orc_nucl_ball(P, oracle._f(ball))
    b = ball.reshape(7, 7, 7)
    r = np.sqrt(np.add.outer(np.add.outer(np.arange(-3, 4) ** 2, np.arange(-3, 4) ** 2), np.arange(-3, 4) ** 2))
    want = np.tanh(0.8 * (2.0 - r) / math.sqrt(2))
    assert np.abs(b - want).max() < 2e-6
    assert b[3, 3, 3] == b.max()


def test_draws_are_uniform_and_deterministic():
    lib = oracle.load("f32")
    u = ctypes.c_float()
    gid = ctypes.c_int()
    vals = []
    for c in range(4000):
        lib.orc_draw(7, c, 3, ctypes.byref(u), ctypes.byref(gid))
        vals.append(u.value)
        assert 0 < u.value <= 1 and 1 <= gid.value < 2 ** 31 - 1
    v = np.array(vals)
    assert abs(v.mean() - 0.5) < 0.02 and abs(v.var() - 1 / 12) < 0.01
    lib.orc_draw(7, 5, 3, ctypes.byref(u), ctypes.byref(gid))
    assert u.value == vals[5]


# ---------------------------------------------------------- precision ---
def test_f32_build_tracks_f64_build():
    c = cases.make_line_case("prec", n=(16, 16, 16), grains=5, K=8, seed=9, z0_cells=8)
    a = oracle.OracleSim(c.params, c.x, c.y, c.z, c.R9, "f32")
    b = oracle.OracleSim(c.params, c.x, c.y, c.z, c.R9, "f64")
    a.init_from_phi_alpha(c.phi0, c.alpha0)
    b.init_from_phi_alpha(c.phi0, c.alpha0)
    a.step(20, 1)
    b.step(20, 1)
    pa, ia = a.state()
    pb, ib = b.state()
    assert np.abs(pa.astype(np.float64) - pb).max() < 1e-3
    assert (ia != ib).mean() < 1e-3


def test_float_thresholds_equal_double_compares():
    """The step kernels compare in float against -0.995f and -0.9999f where the
    reference compares (double)x against -0.995 and -0.9999: the same answer
    for every float in [-1, -0.99] (all of them), and on both sides of it."""
    lo, hi = np.float32(-1.0).view(np.uint32), np.float32(-0.99).view(np.uint32)
    bits = np.arange(min(lo, hi), max(lo, hi) + 1, dtype=np.uint32)
    x = bits.view(np.float32)
    x = np.concatenate([x, np.float32([-2.0, -1.5, 0.0, 0.5, 1.0, -0.5])])
    xd = x.astype(np.float64)
    assert np.array_equal(xd > -0.995, x > np.float32(-0.995))
    assert np.array_equal(xd < -0.9999, x < np.float32(-0.9999))


def test_line_qois_restatement_against_loops_and_a_hand_case():
    """oracle/qoi.py::line_qois against a literal transcription of the loops of
    3DgrainMPI/QOI.cpp:47-90 on random fields, and a hand-checked field."""
    from oracle import qoi
    rng = np.random.default_rng(0)
    fnx, fny, fnz = 7, 6, 9
    for trial in range(6):
        A = rng.integers(0, 6, size=(fnz, fny, fnx)).astype(np.int32)
        A[1:1 + trial % 5, 1:-1, 1:-1] = rng.integers(1, 6, size=(trial % 5, fny - 2, fnx - 2))
        for cur in (0, 1, 2, 5):
            tip = cur
            while True:
                tip += 1
                if tip >= fnz - 1:
                    break
                if any(A[tip, j, i] == 0 for j in range(1, fny - 1) for i in range(1, fnx - 1)):
                    break
            tip -= 1
            tot, ext = [0] * 4, [0] * 4
            for k in range(1, fnz - 1):
                for j in range(1, fny - 1):
                    for i in range(1, fnx - 1):
                        a = A[k, j, i]
                        if 0 < a <= 4:
                            tot[a - 1] += 1
                            if k > tip:
                                ext[a - 1] += 1
            t2, tot2, ext2 = qoi.line_qois(A.ravel(), fnx, fny, fnz, cur, 4)
            assert t2 == tip and list(tot2) == tot and list(ext2) == ext
    # hand case: 3x3 interior, planes 1-2 solid grain 1, plane 3 grain 2 with
    # one liquid point, plane 4 liquid: the tip stops at plane 3 - 1 = 2;
    # grain 1 has 18 points (none above the tip), grain 2 has 8 (all above)
    H = np.zeros((6, 5, 5), dtype=np.int32)
    H[1:3, 1:4, 1:4] = 1
    H[3, 1:4, 1:4] = 2
    H[3, 2, 2] = 0
    tip, total, extra = qoi.line_qois(H.ravel(), 5, 5, 6, 0, 2)
    assert tip == 2 and list(total) == [18, 8] and list(extra) == [0, 8]


def _volume_loops(A):
    """Literal transcription of QOI::calculateQoIs (3DgrainMPI/QOI.cpp:198-231)."""
    fnz, fny, fnx = A.shape
    index, ids, vol = {}, [], []
    for k in range(1, fnz - 1):
        for j in range(1, fny - 1):
            for i in range(1, fnx - 1):
                a = int(A[k, j, i])
                if a > 0:
                    if a not in index:
                        index[a] = len(ids)
                        ids.append(a)
                        vol.append(0)
                    vol[index[a]] += 1
    return ids, vol


def _junction_loops(A):
    """Literal transcription of QOI3D::searchJunctionsOnImage
    (3DgrainMPI/QOI.cpp:233-311); ids as a sorted tuple."""
    fnz, fny, fnx = A.shape
    out = []
    for k in range(1, fnz - 1):
        for j in range(1, fny - 1):
            for i in range(1, fnx - 1):
                if A[k, j, i] == 0:
                    continue
                occur, zero = {}, 0
                for dk in (-1, 0, 1):
                    for dj in (-1, 0, 1):
                        for di in (-1, 0, 1):
                            if di * di + dj * dj + dk * dk > 2:
                                continue
                            a = int(A[k + dk, j + dj, i + di])
                            zero += a == 0
                            occur[a] = occur.get(a, 0) + 1
                if zero > 0:
                    continue
                if len(occur) == 4 and max(occur.values()) <= 10:
                    out.append((i, j, k, max(occur.values()), tuple(sorted(occur))))
    return out


def test_volume_and_junction_restatements_against_loops_and_a_hand_case():
    """oracle/qoi.py::volume and ::junctions3d against the loops above on
    random block fields (many 4-id neighbourhoods, some liquid, counts on
    both sides of the <= 10 bound), and a hand-worked corner of 4 grains."""
    from oracle import qoi
    rng = np.random.default_rng(1)
    seen_big = seen_junction = 0
    for trial in range(8):
        fnx, fny, fnz = 9 + trial % 3, 8, 7 + trial % 2
        k, j, i = np.meshgrid(np.arange(fnz), np.arange(fny), np.arange(fnx), indexing="ij")
        b = 2 + trial % 2
        table = rng.integers(1, 9, size=(fnz // b + 1, fny // b + 1, fnx // b + 1))
        A = table[k // b, j // b, i // b].astype(np.int32)
        A[rng.random(A.shape) < 0.02] = 0
        ids, vol = _volume_loops(A)
        assert qoi.volume(A.ravel(), fnx, fny, fnz) == (ids, vol)
        rows = _junction_loops(A)
        assert qoi.junctions3d(A.ravel(), fnx, fny, fnz) == rows
        seen_junction += len(rows)
        seen_big += sum(r[3] == 10 for r in rows)
    assert seen_junction > 20 and seen_big > 0
    # hand case: interior 4^3, four quadrants in (i, j) split between 2 and 3,
    # the same in every plane.  The 19-point neighbourhood of (2, 2, k) holds
    # id 1 ten times, ids 2 and 3 four times, id 4 once (the dk = 0 diagonal):
    # the four points (2|3, 2|3, k) of each plane are junctions.  A liquid
    # point at (2, 2, 1) removes every point whose neighbourhood holds it:
    # the four of plane 1 and (2, 2, 2), (3, 2, 2), (2, 3, 2); (3, 3, 2) is
    # at distance^2 3 and stays.
    k, j, i = np.meshgrid(np.arange(6), np.arange(6), np.arange(6), indexing="ij")
    H = (1 + (i >= 3) + 2 * (j >= 3)).astype(np.int32)
    full = qoi.junctions3d(H.ravel(), 6, 6, 6)
    assert full == [(ii, jj, kk, 10, (1, 2, 3, 4)) for kk in range(1, 5) for jj in (2, 3)
                    for ii in (2, 3)]
    H[1, 2, 2] = 0
    want = [(3, 3, 2, 10, (1, 2, 3, 4))] + [(ii, jj, kk, 10, (1, 2, 3, 4)) for kk in (3, 4)
                                            for jj in (2, 3) for ii in (2, 3)]
    assert qoi.junctions3d(H.ravel(), 6, 6, 6) == want
    assert qoi.volume(H.ravel(), 6, 6, 6) == ([1, 2, 3, 4], [15, 16, 16, 16])


def _line_junction_loops(A, tip_y, num_nodes, nfeat=8, repeated=5):
    """QOILine::searchJunctionsOnImage (3DgrainMPI/QOI.cpp:103-160) as loops
    with a dict count, the decisions of this build marked: ids ascending, at
    most nfeat - 3 of them, rows past a frame's share skipped."""
    _, fny, fnx = A.shape
    per = repeated * num_nodes * nfeat
    nodes = [-1] * (len(tip_y) * per)
    cross = []
    for kt, tip in enumerate(tip_y):
        cross.extend(int(v) for v in A[tip].ravel())
        node_cnt = 0
        for j in range(1, fny - 1):
            for i in range(1, fnx - 1):
                occur = {}
                for dj in (-1, 0, 1):
                    for di in (-1, 0, 1):
                        a = int(A[tip, j + dj, i + di])
                        occur[a] = occur.get(a, 0) + 1
                if len(occur) >= 3 and max(occur.values()) <= 5:
                    off = per * kt + node_cnt * nfeat
                    if off + nfeat > per * (kt + 1):      # decision: no write past the frame
                        continue
                    nodes[off], nodes[off + 1], nodes[off + 2] = i, j, max(occur.values())
                    for q, a in enumerate(sorted(occur)[:nfeat - 3]):   # decision: ascending, <= 5
                        nodes[off + 3 + q] = a
                    node_cnt += 1
    return cross, nodes


def test_line_junction_restatement_against_loops_and_a_hand_case():
    """oracle/qoi.py::line_junctions against the loops above on random tip
    planes (3-9 values per window, liquid included, frames that overflow
    their share), and a hand-worked plane of four quadrant grains."""
    from oracle import qoi
    rng = np.random.default_rng(4)
    seen = over = wide = 0
    for trial in range(6):
        fnx, fny, nz = 9 + trial, 7, 12
        b = 1 + trial % 3
        jj, ii = np.meshgrid(np.arange(fny), np.arange(fnx), indexing="ij")
        A = np.stack([rng.integers(0, 7, size=(fny // b + 1, fnx // b + 1))[jj // b, ii // b]
                      for _ in range(nz)]).astype(np.int32)
        tips = [int(t) for t in rng.integers(0, nz, size=3)]
        nn = 1 + trial % 2
        cross, nodes = qoi.line_junctions(A.ravel(), fnx, fny, tips, nn)
        c2, n2 = _line_junction_loops(A, tips, nn)
        assert cross.tolist() == c2 and nodes.tolist() == n2
        rows = np.asarray(n2).reshape(-1, 8)
        seen += int((rows[:, 0] != -1).sum())
        over += int(all((rows[q * 5 * nn:(q + 1) * 5 * nn, 0] != -1).all() for q in range(3)))
        wide += int((rows[:, 7] != -1).sum())
    assert seen > 30 and over > 0 and wide > 0
    # hand case: a 6x6 plane (4x4 interior) of four quadrant grains split at
    # i >= 3 and j >= 3.  The 3x3 windows at (2|3, 2|3) hold all four ids,
    # one of them 4 times; every other interior window holds at most 2 ids.
    jj, ii = np.meshgrid(np.arange(6), np.arange(6), indexing="ij")
    H = (1 + (ii >= 3) + 2 * (jj >= 3)).astype(np.int32)[None]
    cross, nodes = qoi.line_junctions(H.ravel(), 6, 6, [0], 1)
    rows = nodes.reshape(-1, 8)
    assert rows[:4].tolist() == [[2, 2, 4, 1, 2, 3, 4, -1], [3, 2, 4, 1, 2, 3, 4, -1],
                                 [2, 3, 4, 1, 2, 3, 4, -1], [3, 3, 4, 1, 2, 3, 4, -1]]
    assert (rows[4:] == -1).all() and cross.tolist() == H.ravel().tolist()
