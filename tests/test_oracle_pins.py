"""More pins of the CPU oracle (no GPU): the 27-point gather order, the thermal
laws of types 2-5, the macro-field interpolations and the main-loop
(macro temperature field) branch of nucleation.

Each check states what the reference's law fixes at chosen points (a circle
of known radius, a node of the macro grid, a multilinear field that
interpolation must reproduce exactly, a probability in closed form), so a
plausible slip in the restatement -- a wrong axis offset, sign, clamp, scan
order or time shift -- fails it.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle
from tests.test_oracle import _params, _sim, _set

f32 = np.float32


def _thermal(p, x, y, z, t, i=0, j=0, k=0, psi=None, prec="f32"):
    lib = oracle.load(prec)
    P = oracle.OrcParams.from_dict(p)
    out = ctypes.c_double()
    xs = np.asarray([x], np.float32) if np.ndim(x) == 0 else np.asarray(x, np.float32)
    ys = np.asarray([y], np.float32) if np.ndim(y) == 0 else np.asarray(y, np.float32)
    zs = np.asarray([z], np.float32) if np.ndim(z) == 0 else np.asarray(z, np.float32)
    skip = lib.orc_thermal(P, oracle._f(xs), oracle._f(ys), oracle._f(zs), oracle._f(psi), i, j, k,
                           float(t), 0, ctypes.byref(out))
    return skip, out.value


def _step_times(n=3000):
    dt, tau0 = f32(0.0363668), f32(1.3280389e-7)
    return [float(f32(f32(f32(s) * dt) * tau0)) for s in range(1, n)]


# ------------------------------------------------- 27-point gather order ---
def _gather27_point(K, entries, G=0.0):
    """5^3 grid, 27-point gather (3Dgrain/APTPhaseField.cu:241-268).  entries:
    list of (q, slot, id, phi) with q = (xi+1) + 3(yi+1) + 9(zi+1) the z-major
    neighbour position of the centre point (2,2,2).  Returns the centre's new
    (ids, phis) over its K slots."""
    p = _params(fn=(5, 5, 5), K=K, bcx=0, bcy=0, bcz=0, gather27=1, G=f32(G))
    ids = np.full((K, 5, 5, 5), -1, np.int32)
    phs = np.full((K, 5, 5, 5), -1.0)
    for q, slot, g, v in entries:
        xi, yi, zi = q % 3 - 1, (q // 3) % 3 - 1, q // 9 - 1
        ids[slot, 2 + zi, 2 + yi, 2 + xi] = g
        phs[slot, 2 + zi, 2 + yi, 2 + xi] = v
    sim = _sim(p)
    _set(sim, ids, phs)
    sim.lib.orc_rhs(sim.P, oracle._f(sim.x), oracle._f(sim.y), oracle._f(sim.z), oracle._f(sim.R9), None,
                    oracle._v(sim.ph[0]), oracle._i(sim.arg[0]), oracle._v(sim.ph[1]), oracle._i(sim.arg[1]),
                    0.0, 0)
    a = sim.arg[1].reshape(K, 5, 5, 5)[:, 2, 2, 2]
    f = sim.ph[1].reshape(K, 5, 5, 5)[:, 2, 2, 2]
    return list(a), list(f), p


def _lone_neighbour_value(p):
    """New phi of a slot whose id sits at exactly one of the six face
    neighbours with phi = 1 and nowhere else (centre phi -1, so the identity
    rotation): diff = 2, |grad| = 1 along one axis, A2 = a_s (1 + eps), no
    thermal term (G = 0): -1 + dt * 2 hi^2 / A2."""
    kd = float(p["kin_delta"])
    a_s = 1 + 3 * kd
    eps = -4 * kd / a_s
    return -1.0 + float(p["dt"]) * 2 * float(p["hi"]) ** 2 / (a_s * (1 + eps))


def test_gather27_scans_z_major_and_keeps_phi_only_on_the_stencil():
    # one id per position, at corners (0, 8, 26), in-plane and out-of-plane
    # edges (11, 20) and the stencil rows D=4, C=13, U=22.  Slots fill in
    # ascending q (z outer, then y, then x); ids seen only off the stencil get
    # a slot but no phi rows, so they evolve from all -1 and are killed
    # (arg -1, phi -1) while keeping their place in the slot order.
    entries = [(26, 0, 7, 0.9), (13, 0, 5, 1.0), (0, 0, 3, 0.9), (22, 0, 6, 1.0),
               (8, 0, 9, 0.9), (4, 0, 4, 1.0), (20, 0, 8, 0.9), (11, 0, 2, 0.9)]
    ids, phs, p = _gather27_point(8, entries)
    # ascending q: 0:3 4:4 8:9 11:2 13:5 20:8 22:6 26:7
    assert ids == [-1, 4, -1, -1, 5, -1, 6, -1]
    assert [phs[s] for s in (0, 2, 3, 5, 7)] == [-1.0] * 5
    lone = _lone_neighbour_value(p)
    assert phs[1] == pytest.approx(lone, abs=1e-12)          # id 4 only at D
    assert phs[6] == pytest.approx(lone, abs=1e-12)          # id 6 only at U
    # id 5 only at C with phi 1: diff = -12, 1 - C^2 = 0, |grad| = 0 -> A2 = 1
    assert phs[4] == pytest.approx(1.0 - 12 * float(p["dt"]) * float(p["hi"]) ** 2, abs=1e-12)


@pytest.mark.parametrize("q", [4, 10, 12, 14, 16, 22])
def test_gather27_each_face_neighbour_feeds_phi(q):
    # the same id at a corner first (slot order) and at one face neighbour
    # with phi 1: the slot takes the face value, never the corner's
    ids, phs, p = _gather27_point(5, [(0, 0, 3, 0.99), (q, 0, 3, 1.0)])
    assert ids[0] == 3 and ids[1:] == [-1] * 4
    assert phs[0] == pytest.approx(_lone_neighbour_value(p), abs=1e-12)


def test_gather27_slot_order_within_a_position_and_holes():
    # at one position the neighbour's slots are read in slot order, holes
    # skipped; a repeated id keeps its first slot
    entries = [(1, 2, 11, 0.5), (1, 0, 12, 0.5), (9, 1, 11, 0.5), (9, 3, 13, 0.5), (17, 0, 12, 0.5)]
    ids, _, _ = _gather27_point(5, entries)
    # q=1: slot0 12, slot2 11; q=9: 11 (seen), 13 -> order 12, 11, 13; all off-stencil -> killed
    assert ids == [-1, -1, -1, -1, -1]
    entries += [(13, 0, 13, 1.0), (13, 1, 11, 1.0), (13, 2, 12, 1.0)]
    ids, phs, _ = _gather27_point(5, entries)
    assert ids == [12, 11, 13, -1, -1]


# ----------------------------------------------------- thermal, types 2-5 ---
def test_thermal_cylinder_on_and_off_the_surface():
    """Type 2 (3DgrainMPI/APTPhaseField.cu:288-292): T = -G (rho - r0) - UC 1e6 t,
    rho the distance from the axis y = lyd/2, z = z0 + lzd."""
    p = _params(thermal_type=2, G=f32(2.0), uc=f32(0.5), z0=f32(1.0), lzd=f32(3.0), lyd=f32(4.0),
                r0=f32(5.0))
    t = 2.0 ** -20
    yc, zc = 2.0, 4.0
    for dy, dz, rho in [(3, 4, 5), (-4, 3, 5), (0, -5, 5), (6, 8, 10), (0, 1.5, 1.5)]:
        skip, T = _thermal(p, 0.0, yc + dy, zc + dz, t)
        assert skip == 0
        assert T == float(f32(-2.0 * (rho - 5.0) - 0.5 * 1e6 * t))
    # the UC*1e6*t term is evaluated in double, as in the line law
    p = dict(p, uc=f32(0.1))
    hits = 0
    for t in _step_times():
        g = f32(f32(-2.0) * f32(f32(10.0) - f32(5.0)))
        promoted = f32(float(g) - float(f32(0.1)) * 1e6 * t)
        pure = f32(g - f32(f32(f32(0.1) * f32(1e6)) * f32(t)))
        if promoted == pure:
            continue
        assert _thermal(p, 0.0, yc + 6, zc + 8, t)[1] == float(promoted)
        hits += 1
        if hits == 10:
            break
    assert hits == 10


def test_thermal_cone_radius_grows_along_x_and_skips_past_the_tip():
    """Type 3 (3DgrainMPI/APTPhaseField.cu:293-306): the radius r0_x grows
    linearly from z0 at x = 0 to r0 at x = lm = (r0 - z0) / tan(angle); the
    distance to it is scaled by cos(angle); points past lm + V 1e6 t are
    skipped."""
    p = _params(thermal_type=3, G=f32(2.0), uc=f32(0.25), z0=f32(1.0), r0=f32(3.0), lzd=f32(6.0),
                lyd=f32(2.0), tan_angle=f32(0.5), cos_angle=f32(0.5), V=f32(1.0))
    lm = 4.0
    y, z = 1.0 + 3.0, 1.0 + 6.0 - 4.0     # axis at y = lyd/2, z = z0 + lzd; rho = 5
    t = 2.0 ** -22                        # V 1e6 t = 0.238...
    for x, r in [(0.0, 1.0), (2.0, 2.0), (4.0, 3.0)]:
        skip, T = _thermal(p, x, y, z, t)
        assert skip == 0
        assert T == float(f32(-2.0 * 0.5 * (5.0 - r) - 0.25 * 1e6 * t))
    assert _thermal(p, 4.2, y, z, t)[0] == 0
    assert _thermal(p, 4.25, y, z, t)[0] == 1
    assert _thermal(p, 4.2, y, z, 0.0)[0] == 1


def _trilinear(c, X, Y, Z):
    return (c[0] + c[1] * X + c[2] * Y + c[3] * Z + c[4] * X * Y + c[5] * Y * Z + c[6] * X * Z
            + c[7] * X * Y * Z)


def test_thermal_paraboloid_reproduces_a_trilinear_macro_field():
    """Type 4 (3DgrainMPI/APTPhaseField.cu:307-325): T = -G psi(x - V 1e6 t,
    y, z) (order 100: -psi) with psi interpolated trilinearly on the macro grid;
    outside [0, mp_len + x0] in the shifted x the point is skipped.  A
    trilinear psi (dyadic values) must come back exactly."""
    Nx, Ny, Nz, Dx = 6, 5, 4, 0.5
    c = [0.25, 1.0, -0.5, 0.75, 0.5, -0.25, 0.125, 0.5]
    X0, Y0, Z0 = -0.5, 0.25, -0.25
    gx, gy, gz = np.meshgrid(np.arange(Nx) * Dx, np.arange(Ny) * Dx, np.arange(Nz) * Dx, indexing="ij")
    psi = np.ascontiguousarray(_trilinear(c, gx, gy, gz).transpose(2, 1, 0).ravel(), np.float32)
    base = dict(thermal_type=4, Nx=Nx, Ny=Ny, Nz=Nz, X0_mac=f32(X0), Y0_mac=f32(Y0), Z0_mac=f32(Z0),
                Dx_mac=f32(Dx), mp_len=f32(1.5), x0=f32(0.5), V=f32(0.0), G=f32(3.0))
    for order, scale in [(f32(2.0), -3.0), (f32(100.0), -1.0)]:
        p = _params(order=order, **base)
        for x, y, z in [(0.125, 0.375, 0.0), (1.0, 1.25, 0.5), (1.875, 2.0, 0.625), (0.0, 0.25, -0.25)]:
            skip, T = _thermal(p, x, y, z, 0.0, psi=psi)
            assert skip == 0
            assert T == scale * _trilinear(c, x - X0, y - Y0, z - Z0)
    p = _params(order=f32(2.0), **base)
    assert _thermal(p, 2.0, 1.0, 0.5, 0.0, psi=psi)[0] == 0
    assert _thermal(p, 2.125, 1.0, 0.5, 0.0, psi=psi)[0] == 1   # past mp_len + x0
    assert _thermal(p, -0.125, 1.0, 0.5, 0.0, psi=psi)[0] == 1  # below 0
    # the frame moves with V 1e6 t: x = 2.5 at V 1e6 t = 1 samples x - 1 = 1.5
    pv = _params(order=f32(2.0), **{**base, "V": f32(1.0)})
    t = 1e-6
    skip, T = _thermal(pv, 2.5, 1.0, 0.5, t, psi=psi)
    assert skip == 0
    xq = float(f32(2.5 - 1.0 * t * 1e6))
    assert T == pytest.approx(-3.0 * _trilinear(c, xq - X0, 1.0 - Y0, 0.5 - Z0), abs=2e-6)


def test_thermal_3dgrain_line_law():
    """Type 5, the 3Dgrain law (3Dgrain/APTPhaseField.cu:303): T = G (z - R 1e6 t - 2),
    evaluated in double and rounded once."""
    p = _params(thermal_type=5, G=f32(1.5), R=f32(0.25))
    t = 2.0 ** -21
    shift = 0.25 * 1e6 * t
    for z in [2.0 + shift, 3.0, 0.5, 2.0]:
        assert _thermal(p, 0.0, 0.0, z, t)[1] == float(f32(1.5 * (float(f32(z)) - shift - 2.0)))
    assert _thermal(p, 0.0, 0.0, 2.0 + shift, t)[1] == 0.0
    hits = 0
    for t in _step_times():
        z = f32(2.7)
        promoted = f32(1.5 * (float(z) - float(f32(0.25)) * 1e6 * t - 2.0))
        pure = f32(f32(1.5) * f32(f32(z - f32(f32(f32(0.25) * f32(1e6)) * f32(t))) - f32(2.0)))
        if promoted == pure:
            continue
        assert _thermal(p, 0.0, 0.0, z, t)[1] == float(promoted)
        hits += 1
        if hits == 10:
            break
    assert hits == 10


# ----------------------------------------------------- macro interpolation ---
def test_interp3_exact_on_trilinear_fields_at_and_between_nodes():
    lib = oracle.load("f32")
    Nx, Ny, Nz, Dx = 5, 4, 6, 0.25
    c = [1.0, -2.0, 0.5, 1.5, 0.25, -0.5, 1.0, 0.75]
    gx, gy, gz = np.meshgrid(np.arange(Nx) * Dx, np.arange(Ny) * Dx, np.arange(Nz) * Dx, indexing="ij")
    u = np.ascontiguousarray(_trilinear(c, gx, gy, gz).transpose(2, 1, 0).ravel(), np.float32)
    pts = [(0, 0, 0), (0.25, 0.5, 0.75), (0.125, 0.375, 0.0625), (0.9375, 0.6875, 1.1875),
           (1.0, 0.75, 1.25),            # the last node on every axis: kx = Nx-1 -> Nx-2, dx = 1
           (1.0, 0.125, 0.5), (0.5, 0.75, 0.3125)]
    for X, Y, Z in pts:
        got = lib.orc_interp3(oracle._f(u), X, Y, Z, Nx, Ny, Nz, Dx)
        assert got == _trilinear(c, X, Y, Z), (X, Y, Z)
    # a slip that swaps the y and z offsets or the Nx stride fails on a
    # field whose value depends on the node index only through that axis
    v = np.zeros((Nz, Ny, Nx), np.float32)
    v[:, 2, :] = 1.0
    v = np.ascontiguousarray(v.ravel())
    assert lib.orc_interp3(oracle._f(v), 0.5, 0.5, 0.5, Nx, Ny, Nz, Dx) == 1.0
    assert lib.orc_interp3(oracle._f(v), 0.5, 0.375, 0.5, Nx, Ny, Nz, Dx) == 0.5
    assert lib.orc_interp3(oracle._f(v), 0.5, 0.25, 0.5, Nx, Ny, Nz, Dx) == 0.0


def test_interp4_exact_on_quadrilinear_fields_and_clamps_the_last_frame():
    lib = oracle.load("f32")
    Nx, Ny, Nz, Nt, Dx, Dt = 4, 3, 5, 3, 0.5, 0.25
    c = [0.5, 1.0, -1.5, 0.25, 0.5, 0.0, -0.25, 0.125]
    gx, gy, gz, gt = np.meshgrid(np.arange(Nx) * Dx, np.arange(Ny) * Dx, np.arange(Nz) * Dx,
                                 np.arange(Nt) * Dt, indexing="ij")

    def f(X, Y, Z, T):
        return _trilinear(c, X, Y, Z) * (1 + 2 * T) - 3 * T

    u = np.ascontiguousarray(f(gx, gy, gz, gt).transpose(3, 2, 1, 0).ravel(), np.float32)
    for X, Y, Z, T in [(0, 0, 0, 0), (0.25, 0.75, 1.25, 0.125), (1.5, 1.0, 2.0, 0.5),
                       (0.75, 0.25, 0.5, 0.375), (1.25, 0.5, 1.75, 0.0625)]:
        got = lib.orc_interp4(oracle._f(u), X, Y, Z, T, Nx, Ny, Nz, Nt, Dx, Dt)
        assert got == f(X, Y, Z, T), (X, Y, Z, T)
    # frames are Nx*Ny*Nz apart: a field that changes only with the frame
    w = np.repeat(np.array([0.0, 2.0, 6.0], np.float32), Nx * Ny * Nz)
    assert lib.orc_interp4(oracle._f(w), 0.3, 0.2, 0.7, 0.125, Nx, Ny, Nz, Nt, Dx, Dt) == 1.0
    assert lib.orc_interp4(oracle._f(w), 0.3, 0.2, 0.7, 0.375, Nx, Ny, Nz, Nt, Dx, Dt) == 4.0
    assert lib.orc_interp4(oracle._f(w), 0.3, 0.2, 0.7, 0.5, Nx, Ny, Nz, Nt, Dx, Dt) == 6.0


# ------------------------------------------- nucleation, macro-field branch ---
def _nucl_case(cells=6, p_cell=1, Nmax=4.0, seed=5):
    """An all-liquid box of cells^3 coarse blocks and a macro temperature that
    is linear in x and cools at B K per unit t from the first frame at t0:
    T(x, t) = -(m + s x) - B (t - t0), so dT = m + s x + B (t + dt - t0) and
    d(dT) = B dt at every block."""
    w = 2 * p_cell + 1
    n = cells * w + 3
    p = _params(fn=(n, n, n), K=3, pts_cell=p_cell, nuc_rad=f32(0.5), undcool_mean=f32(2.0),
                undcool_std=f32(0.5), nuc_Nmax=f32(Nmax), Nx=4, Ny=2, Nz=2, Nt=3, Dx_mac=f32(4.0),
                Dt_mac=f32(1.0), X0_mac=f32(-0.5), Y0_mac=f32(-0.5), Z0_mac=f32(-0.5), t0_mac=f32(0.125))
    m, s, B = 1.0, 0.25, 0.5
    X = np.arange(4) * 4.0
    frames = []
    for kt in range(3):
        plane = -(m + s * X) - B * kt
        frames.append(np.broadcast_to(plane, (2, 2, 4)))
    T3d = np.ascontiguousarray(np.stack(frames).ravel(), np.float32)
    x = (np.arange(n) * 0.25).astype(np.float32)
    return p, x, T3d, (m, s, B)


def _nucl_expect(p, x, mSB, t, dt, p_cell, cells):
    """float64 closed form of the nucleation probability per coarse block
    (3DgrainMPI/APTPhaseField.cu:446-455,486-496)."""
    m, s, B = mSB
    w = 2 * p_cell + 1
    out = {}
    for k in range(cells):
        for j in range(cells):
            for i in range(cells):
                gi = w * i + p_cell + 1
                X = float(x[gi]) - float(p["X0_mac"])
                dT = m + s * X + B * (t + dt - float(p["t0_mac"]))
                ddT = B * dt
                mu, sd, r = float(p["undcool_mean"]), float(p["undcool_std"]), float(p["nuc_rad"])
                P = 8 * r ** 3 * float(p["nuc_Nmax"]) / (math.sqrt(2 * math.pi) * sd) \
                    * math.exp(-0.5 * ((dT - mu) / sd) ** 2) * ddT
                out[i + j * cells + k * cells * cells] = (P, dT)
    return out


def test_add_nucl_macro_field_decisions_and_ball():
    cells, pc = 6, 1
    p, x, T3d, mSB = _nucl_case(cells, pc)
    n = p["fnx"]
    sim = oracle.OracleSim(p, x, x, x, np.tile(np.eye(3, dtype=np.float32).ravel(), (9, 1)), "f32", T3d=T3d)
    L = n ** 3
    phi0 = np.full(L, -1.0, np.float32)
    sim.init_from_phi_alpha(phi0, np.zeros(L, np.int32))
    # one block starts solid at its centre: never nucleates, status set
    w = 2 * pc + 1
    solid_cell = 1 + 2 * cells + 3 * cells * cells
    ci, cj, ck = 1, 2, 3
    g = (w * ck + pc + 1) * n * n + (w * cj + pc + 1) * n + (w * ci + pc + 1)
    sim.ph[sim.cur][g] = 0.5
    sim.arg[sim.cur][g] = 77
    sim.nucl_reset(0)
    t, dt, seed, call = 0.5, 0.25, 11, 0
    sim.ncall = call
    sim.lib.orc_add_nucl(sim.P, oracle._f(sim.x), oracle._f(sim.y), oracle._f(sim.z), oracle._f(sim.T3d),
                         oracle._f(sim.ball), oracle._v(sim.ph[sim.cur]), oracle._i(sim.arg[sim.cur]),
                         oracle._i(sim.nstatus), sim.cnx, sim.cny, sim.cnz, dt, t, 0, seed, call)
    want = _nucl_expect(p, x, mSB, t, dt, pc, cells)
    lib = sim.lib
    u = ctypes.c_float()
    gid = ctypes.c_int()
    ph = sim.ph[sim.cur].reshape(3, n, n, n)
    ar = sim.arg[sim.cur].reshape(3, n, n, n)
    ball = sim.ball.reshape(w, w, w)
    n_yes = n_no = 0
    for C, (P, dT) in want.items():
        i, j, k = C % cells, (C // cells) % cells, C // (cells * cells)
        sl = (slice(w * k + 1, w * k + 1 + w), slice(w * j + 1, w * j + 1 + w), slice(w * i + 1, w * i + 1 + w))
        if C == solid_cell:
            assert sim.nstatus[C] == 1
            assert (ar[0][sl] == -1).sum() == w ** 3 - 1
            continue
        lib.orc_draw(seed, C, call, ctypes.byref(u), ctypes.byref(gid))
        if abs(u.value - P) < 1e-5 * max(P, 1e-3):
            continue   # within rounding of the float32 evaluation
        if u.value < P:
            n_yes += 1
            assert sim.nstatus[C] == 1
            assert np.array_equal(ph[0][sl], ball)
            assert np.array_equal(ar[0][sl], np.where(ball > np.float32(-0.9999), gid.value, -1))
        else:
            n_no += 1
            assert sim.nstatus[C] == 0
            assert (ar[0][sl] == -1).all() and (ph[0][sl] == -1).all()
    assert n_yes > 10 and n_no > 10


def test_nucleation_counts_follow_the_expected_sum_of_probabilities():
    """Many independent blocks and calls: the number of events per undercooling
    bin is binomial with mean sum(P); the counts stay within 4 standard
    deviations of it in every bin (the law of 3DgrainMPI/APTPhaseField.cu:446-455
    driven by the counter-based draw)."""
    cells, pc = 8, 1
    p, x, T3d, mSB = _nucl_case(cells, pc, Nmax=8.0)
    n = p["fnx"]
    sim = oracle.OracleSim(p, x, x, x, np.tile(np.eye(3, dtype=np.float32).ravel(), (9, 1)), "f32", T3d=T3d)
    L = n ** 3
    bins = np.linspace(1.3, 3.1, 5)
    expect = np.zeros(len(bins) - 1)
    var = np.zeros(len(bins) - 1)
    got = np.zeros(len(bins) - 1)
    for call in range(40):
        sim.init_from_phi_alpha(np.full(L, -1.0, np.float32), np.zeros(L, np.int32))
        sim.nucl_reset(0)
        t, dt = 0.02 * call, 0.2
        sim.lib.orc_add_nucl(sim.P, oracle._f(sim.x), oracle._f(sim.y), oracle._f(sim.z),
                             oracle._f(sim.T3d), oracle._f(sim.ball), oracle._v(sim.ph[sim.cur]),
                             oracle._i(sim.arg[sim.cur]), oracle._i(sim.nstatus), sim.cnx, sim.cny, sim.cnz,
                             dt, t, 0, 3, call)
        for C, (P, dT) in _nucl_expect(p, x, mSB, t, dt, pc, cells).items():
            b = np.searchsorted(bins, dT) - 1
            if 0 <= b < len(expect):
                expect[b] += min(P, 1.0)
                var[b] += min(P, 1.0) * (1 - min(P, 1.0))
                got[b] += sim.nstatus[C]
    assert expect.min() > 20
    assert np.all(np.abs(got - expect) <= 4 * np.sqrt(var) + 1), (got, expect)
