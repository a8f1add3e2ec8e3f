"""Synthetic inputs for the APT step: parameters, grid, substrate, orientations.

Everything the reference takes from an input ``.py`` file plus the
``forcing/case<seed>/`` folder is produced here from a small dict of physical
inputs, with the derived constants computed in float32 exactly as
``PhaseField::parseInputParams`` / ``cpuSetup`` / ``initField`` do
(3DgrainMPI/PhaseFieldCpu.cpp:141-199, :285-355, :377-449).

The substrate generator is this project's own: a seeded periodic (or
bounded) 2-D Voronoi partition of the (nx+1) x (ny+1) base plane, extruded in
z below the initial interface -- the rule of 3Dgrain/PhaseFieldCpu.cpp:389-436.
Orientations are seeded random unit quaternions turned into rotation matrices,
or the reference's two-angle law (3DgrainMPI/graph_datastruct.py:256-269).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

f32 = np.float32
LS = -0.995

# 3DgrainMPI/line.py:1-61 (physical and numerical constants of the line case)
LINE_PY = dict(
    Tmelt=933.3, Dh=8.43e7, L_cp=229.0, GT=0.347, beta0=1e-7, mu_k=0.217e6, delta=0.01,
    kin_delta=0.11, undcool_mean=2.0, undcool_std=0.5, nuc_Nmax=0.01, nuc_rad=0.4,
    nx=13, ny=13, nz=23, nt=11, haloWidth=1, xmin=0.0, ymin=0.0, zmin=0.0, eta=0.0,
    noi_period=200, eps=1e-8, ictype=0, dx=0.8, W0=0.1, cfl=1.5, asp_ratio_yx=1.0,
    asp_ratio_zx=2.0, moving_ratio=0.2, nts=120, Lx=40.0, underCoolingRate0=20.0,
    nuc_Nmax0=0.01, preMt=0,
)

# 3DgrainMPI/cone.py:1-63
CONE_PY = dict(LINE_PY, W0=0.15625, cfl=0.6, asp_ratio_yx=0.32, asp_ratio_zx=0.16, nts=1,
               Lx=250.0, nx=253, ny=83, nz=43, nt=5, preMt=4000)


def _cround(v) -> int:
    """C++ std::round of a float: halves away from zero."""
    v = float(v)
    return int(math.copysign(math.floor(abs(v) + 0.5), v))


def derive(inp: dict, *, line_config: bool) -> dict:
    """Derived constants of parseInputParams (3DgrainMPI/PhaseFieldCpu.cpp:141-199), float32."""
    p = {k: (f32(v) if isinstance(v, float) else v) for k, v in inp.items()}
    R = f32(p["R"])
    cfl = f32(p["cfl"])
    uc0 = f32(p["underCoolingRate0"])
    if R < 0.3:
        cfl = f32(cfl * f32(2))
        uc0 = f32(uc0 / f32(2))
    if R > 0.3 and R < 1:
        cfl = f32(np.float64(cfl) * 1.5)
        uc0 = f32(np.float64(uc0) / 1.5)
    dx, W0 = f32(p["dx"]), f32(p["W0"])
    dxd = f32(dx * W0)
    L_cp, GT, mu_k = f32(p["L_cp"]), f32(p["GT"]), f32(p["mu_k"])
    d0 = f32(GT / L_cp)
    beta0 = f32(1.0 / np.float64(f32(mu_k * L_cp)))
    lamd = f32(0.8839 * np.float64(W0) / np.float64(d0))
    tau0 = f32(np.float64(f32(f32(beta0 * lamd) * W0)) / 0.8839)
    beta0_tilde = f32(f32(beta0 * W0) / tau0)
    dt = f32(f32(cfl * dx) * beta0_tilde)
    hi = f32(1.0 / np.float64(dx))
    sqrt2 = f32(math.sqrt(2.0))
    lxd = f32(p["Lx"])
    nx = _cround(f32(f32(lxd / dx) / W0))
    ny = _cround(f32(f32(p["asp_ratio_yx"]) * f32(nx)))
    if line_config:
        nz = _cround(f32(f32(p["moving_ratio"]) * f32(nx)))
    else:
        nz = _cround(f32(f32(p["asp_ratio_zx"]) * f32(nx)))
    nz_full = _cround(f32(f32(p["asp_ratio_zx"]) * f32(nx)))
    out = dict(p)
    out.update(
        cfl=cfl, uc0=uc0, dxd=dxd, d0=d0, beta0=beta0, lamd=lamd, tau0=tau0, dt=dt, hi=hi,
        sqrt2=sqrt2, nx=nx, ny=ny, nz=nz, nz_full=nz_full,
        lxd=f32(nx * dxd), lyd=f32(ny * dxd), lzd=f32(nz * dxd),
        pts_cell=int(f32(p["nuc_rad"]) / dxd),
        fnx=nx + 3, fny=ny + 3, fnz=nz + 3, fnz_f=nz_full + 3,
    )
    return out


def coords(d: dict):
    """x/y/z coordinate vectors of cpuSetup (3DgrainMPI/PhaseFieldCpu.cpp:341-355)."""
    dxd = d["dxd"]
    x = (np.arange(d["fnx"], dtype=np.int64) - 1).astype(f32) * dxd + f32(d["xmin"])
    y = (np.arange(d["fny"], dtype=np.int64) - 1).astype(f32) * dxd + f32(d["ymin"])
    z = (np.arange(d["fnz"], dtype=np.int64) - 1).astype(f32) * dxd + f32(d["zmin"])
    return x.astype(f32), y.astype(f32), z.astype(f32)


def voronoi_substrate(nxp: int, nyp: int, n_grains: int, seed: int, periodic: bool = True) -> np.ndarray:
    """Seeded 2-D Voronoi labels 1..n_grains on an nyp x nxp grid (row j, column i)."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    pts = rng.random((n_grains, 2)) * np.array([nxp, nyp], dtype=np.float64)
    ii, jj = np.meshgrid(np.arange(nxp), np.arange(nyp), indexing="xy")
    q = np.stack([ii.ravel() + 0.0, jj.ravel() + 0.0], axis=1)
    tree = cKDTree(pts, boxsize=[nxp, nyp]) if periodic else cKDTree(pts)
    _, lab = tree.query(q)
    img = (lab + 1).astype(np.int32).reshape(nyp, nxp)
    # relabel so every label 1..N appears (tiny grains may vanish from the raster)
    uniq = np.unique(img)
    remap = np.zeros(n_grains + 1, dtype=np.int32)
    remap[uniq] = np.arange(1, uniq.size + 1, dtype=np.int32)
    return remap[img]


def quaternion_rotations(n: int, seed: int) -> np.ndarray:
    """(n+1, 9) row-major rotations from seeded uniform unit quaternions; row 0 = identity."""
    rng = np.random.default_rng(seed + 0x5EED)
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 9))
    R[:, 0] = 1 - 2 * (y * y + z * z); R[:, 1] = 2 * (x * y - z * w); R[:, 2] = 2 * (x * z + y * w)
    R[:, 3] = 2 * (x * y + z * w); R[:, 4] = 1 - 2 * (x * x + z * z); R[:, 5] = 2 * (y * z - x * w)
    R[:, 6] = 2 * (x * z - y * w); R[:, 7] = 2 * (y * z + x * w); R[:, 8] = 1 - 2 * (x * x + y * y)
    out = np.zeros((n + 1, 9), dtype=f32)
    out[0] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    out[1:] = R.astype(f32)
    return out


def reference_angles(n: int, seed: int) -> np.ndarray:
    """theta = [0, thx(1..n), thz(1..n)] by the reference's law
    (3DgrainMPI/graph_datastruct.py:256-269) from seeded normal vectors."""
    rng = np.random.default_rng(seed + 0xA9)
    ux, uy, uz = rng.standard_normal((3, n))
    thx = np.arctan2(uy, ux) % math.pi
    thz = np.arctan2(np.sqrt(ux ** 2 + uy ** 2), uz) % math.pi
    return np.concatenate([[0.0], thx, thz]).astype(f32)


@dataclass
class Case:
    name: str
    d: dict                      # derived reference parameters
    params: dict                 # pf_params / orc_params field values
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    R9: np.ndarray
    phi0: np.ndarray
    alpha0: np.ndarray
    n_grains: int
    extra: dict = field(default_factory=dict)

    @property
    def L(self) -> int:
        return self.params["fnx"] * self.params["fny"] * self.params["fnz"]

    def t_of_step(self, s: int, precision: str = "f32"):
        """t_cur_step = s*dt*tau0 (3DgrainMPI/APTPhaseField.cu:863), float32 or float64."""
        if precision == "f32":
            return f32(f32(f32(s) * self.params["dt"]) * self.params["tau0"])
        return float(s) * float(self.params["dt"]) * float(self.params["tau0"])


def kernel_params(d: dict, *, K: int, num_theta: int, bc=(1, 1, 0), thermal_type: int = 1,
                  gather27: int = 0, seed: int = 0) -> dict:
    """Flat parameter block shared by the C ABI and the oracle."""
    angle = f32(d.get("angle", 0.0))
    return dict(
        fnx=d["fnx"], fny=d["fny"], fnz=d["fnz"], K=K, num_theta=num_theta,
        bcx=bc[0], bcy=bc[1], bcz=bc[2], thermal_type=thermal_type, gather27=gather27,
        dt=d["dt"], hi=d["hi"], lamd=d["lamd"], L_cp=f32(d["L_cp"]), kin_delta=f32(d["kin_delta"]),
        eps=f32(d["eps"]), G=f32(d["G"]), R=f32(d["R"]), uc=f32(d["UC"]), uc0=d["uc0"],
        z0=f32(d["z0"]), r0=f32(d.get("r0", 0.0)), lyd=d["lyd"], lzd=d["lzd"], V=f32(d.get("V", 0.0)),
        mp_len=f32(d.get("mp_len", 0.0)), x0=f32(d.get("x0", 0.0)), order=f32(d.get("order", 1.0)),
        tan_angle=f32(math.tan(float(angle))) if angle != 0 else f32(0.0),
        cos_angle=f32(math.cos(float(angle))),
        Nx=int(d.get("Nx", 2)), Ny=int(d.get("Ny", 2)), Nz=int(d.get("Nz", 2)), Nt=int(d.get("Nt", 2)),
        X0_mac=f32(d.get("X0_mac", 0.0)), Y0_mac=f32(d.get("Y0_mac", 0.0)),
        Z0_mac=f32(d.get("Z0_mac", 0.0)), t0_mac=f32(d.get("t0_mac", 0.0)),
        Dx_mac=f32(d.get("Dx_mac", 1.0)), Dt_mac=f32(d.get("Dt_mac", 1.0)),
        pts_cell=int(d["pts_cell"]), nuc_rad=f32(d["nuc_rad"]), undcool_mean=f32(d["undcool_mean"]),
        undcool_std=f32(d["undcool_std"]), nuc_Nmax=f32(d["nuc_Nmax"]), nuc_Nmax0=f32(d["nuc_Nmax0"]),
        dx=f32(d["dx"]), sqrt2=d["sqrt2"], tau0=d["tau0"], seed=seed,
    )


def line_initial_fields(d: dict, x, y, z, alpha_img: np.ndarray | None):
    """phi0/alpha0 of initField for psi = z0 - z (3DgrainMPI/PhaseFieldCpu.cpp:397-449,
    TemperatureProfile3DAnalytic.py:73-75); alpha extruded where phi > LS."""
    fnx, fny, fnz = d["fnx"], d["fny"], d["fnz"]
    psi = (f32(d["z0"]) - z).astype(f32) / f32(d["W0"])
    phi_z = np.tanh((psi / d["sqrt2"]).astype(f32)).astype(f32)
    phi = np.zeros((fnz, fny, fnx), dtype=f32)
    alpha = np.zeros((fnz, fny, fnx), dtype=np.int32)
    phi[1:-1, 1:-1, 1:-1] = phi_z[1:-1, None, None]
    if alpha_img is not None:
        solid = phi[1:-1, 1:-1, 1:-1].astype(np.float64) > LS
        alpha[1:-1, 1:-1, 1:-1] = np.where(solid, alpha_img[None, :, :], 0)
    return phi.ravel(), alpha.ravel()


def make_line_case(name: str, *, n: tuple, grains: int, K: int = 8, seed: int = 0,
                   G: float = 1.0, R: float = 0.1, z0_cells: float | None = None, z0: float = 2.0,
                   bc=(1, 1, 0), line_config: bool = False, nz_full: int | None = None,
                   orient: str = "quat", gather27: int = 0, thermal_type: int = 1,
                   kin_delta: float | None = None) -> Case:
    """Plain/line APT case on line.py physics with G in K/um and R in m/s
    (UC = G*R*1e6/1e6, 3DgrainMPI/grains3D.py:187)."""
    nx, ny, nz = n
    inp = dict(LINE_PY)
    inp.update(Lx=float(nx) * 0.8 * 0.1, asp_ratio_yx=ny / nx, asp_ratio_zx=(nz_full or nz) / nx,
               moving_ratio=nz / nx, G=float(G), R=float(R), UC=float(G) * float(R), z0=float(z0))
    if kin_delta is not None:
        inp["kin_delta"] = float(kin_delta)
    d = derive(inp, line_config=line_config)
    if d["nx"] != nx or d["ny"] != ny or d["nz"] != nz:
        raise ValueError(f"grid derivation gave {d['nx']}x{d['ny']}x{d['nz']}, wanted {n}")
    if z0_cells is not None:
        d["z0"] = f32(z0_cells * float(d["dxd"]))
    x, y, z = coords(d)
    img = voronoi_substrate(nx + 1, ny + 1, grains, seed, periodic=bool(bc[0] and bc[1]))
    ng = int(img.max())
    if orient == "quat":
        R9 = quaternion_rotations(ng, seed)
    else:
        from .capi import orientations_from_angles
        R9 = orientations_from_angles(reference_angles(ng, seed), ng)
    phi0, alpha0 = line_initial_fields(d, x, y, z, img)
    params = kernel_params(d, K=K, num_theta=ng, bc=bc, thermal_type=thermal_type, gather27=gather27,
                           seed=seed)
    return Case(name, d, params, x, y, z, R9, phi0, alpha0, ng)


def make_cone_case(name: str, *, n: tuple, K: int = 8, seed: int = 0, G: float = 5.0,
                   R: float = 0.1, angle: float = 0.5, r0_frac: float = 0.35, z0: float = 2.0,
                   V: float = 0.5, orientations: int = 1000) -> Case:
    """All-liquid start under the cone melt-pool field (thermal type 3) on cone.py
    physics (3DgrainMPI/cone.py:1-63: W0 = 0.15625 um, cfl 0.6), for nucleation.
    The cone's radius r0 is r0_frac of the domain width; its tip starts at x = lm
    and moves along x at V (in the units of the type-3 formula,
    3DgrainMPI/APTPhaseField.cu:295-307).  Nucleated ids draw their orientation
    from a pool of `orientations` seeded quaternions (id % NG + 1)."""
    nx, ny, nz = n
    inp = dict(CONE_PY)
    dxd = float(inp["dx"]) * float(inp["W0"])
    inp.update(Lx=float(nx) * dxd, asp_ratio_yx=ny / nx, asp_ratio_zx=nz / nx, moving_ratio=nz / nx,
               G=float(G), R=float(R), UC=float(G) * float(R), z0=float(z0))
    d = derive(inp, line_config=False)
    if d["nx"] != nx or d["ny"] != ny or d["nz"] != nz:
        raise ValueError(f"grid derivation gave {d['nx']}x{d['ny']}x{d['nz']}, wanted {n}")
    d["r0"] = f32(r0_frac * float(d["lyd"]))
    d["V"] = f32(V)
    d["angle"] = f32(angle)
    x, y, z = coords(d)
    L = d["fnx"] * d["fny"] * d["fnz"]
    phi0 = np.full(L, f32(-1.0), dtype=f32)
    alpha0 = np.zeros(L, dtype=np.int32)
    R9 = quaternion_rotations(orientations, seed)
    params = kernel_params(d, K=K, num_theta=orientations, bc=(0, 0, 0), thermal_type=3, seed=seed)
    return Case(name, d, params, x, y, z, R9, phi0, alpha0, 0)


# Named workloads ------------------------------------------------------------------

def c1(K: int = 8, **kw) -> Case:
    """64^3 plain sim, line.py physics, ~100 grains, G=1e6 K/m, R=0.1 m/s, BC 110."""
    return make_line_case("plain-64^3", n=(64, 64, 64), grains=100, K=K, **kw)


def c2(K: int = 8, **kw) -> Case:
    """256x256x512 line-config sim, ~5000 grains, G=1e6 K/m, R=0.1 m/s, BC 110."""
    kw.setdefault("line_config", True)
    kw.setdefault("nz_full", 1024)
    return make_line_case("line-256x256x512", n=(256, 256, 512), grains=5000, K=K, **kw)


def c3(K: int = 8, **kw) -> Case:
    """256^3 all-liquid start, cone melt-pool field moving along x, nucleation."""
    return make_cone_case("cone-256^3", n=(256, 256, 256), K=K, **kw)
