"""Writer of the reference's forcing folder contract (3DgrainMPI/grains3D.py:328-378):
``forcing/case<seed>/`` with x,y,z,t,psi,alpha,theta,G,Rmax,V,NG,NN,UC,Nmax,z0,
r0,x0,top,angle,min_angle,order (one value per line, '%1.4e' / '%d') and
Temp.h5:/Temp, plus the ``<name>.py`` parameter file the drop-in parses.

The substrate is this project's seeded Voronoi generator (cases.voronoi_substrate),
the orientations the reference's two-angle law from seeded normal vectors; the
macro grid and the line-profile psi = z0 - z follow grains3D.py:266-321.
"""
from __future__ import annotations

import os

import numpy as np

from . import cases, h5


def write_input_py(path: str, inp: dict) -> None:
    keys = ["Tmelt", "Dh", "L_cp", "GT", "beta0", "mu_k", "delta", "kin_delta", "undcool_mean",
            "undcool_std", "nuc_Nmax", "nuc_rad", "nx", "ny", "nz", "nt", "haloWidth", "xmin", "ymin",
            "zmin", "eta", "noi_period", "eps", "ictype", "dx", "W0", "cfl", "asp_ratio_yx",
            "asp_ratio_zx", "moving_ratio", "nts", "Lx", "underCoolingRate0", "nuc_Nmax0", "preMt"]
    with open(path, "w") as f:
        for k in keys:
            f.write(f"{k} = {inp[k]}\n")


def _txt(path, arr, fmt="%1.4e"):
    np.savetxt(path, np.atleast_1d(np.asarray(arr)), fmt=fmt, delimiter="\n")


def write_line_case(root: str, seed: int, *, n: tuple, grains: int, G: float, R: float,
                    z0: float, t_end: float, nts: int = 2, line_config: bool = False,
                    nz_full: int | None = None, top: float = 48.0, macro=(13, 13, 23, 11),
                    inp_overrides: dict | None = None, nmax: float | None = None,
                    temp=0.0) -> dict:
    """Write <root>/line.py and <root>/forcing/case<seed>/.  G in K/um, R in m/s,
    z0/top in um, t_end in s.  inp_overrides replace keys of the input file
    (after the grid keys), nmax is Nmax.txt (default: the input's nuc_Nmax) and
    temp the macro temperature Temp.h5:/Temp, uniform in space: one value, or
    one value per time frame of t.txt.  Returns the input dict."""
    nx, ny, nz = n
    inp = dict(cases.LINE_PY)
    Lx = nx * 0.8 * 0.1
    inp.update(Lx=round(Lx, 6), asp_ratio_yx=ny / nx, asp_ratio_zx=(nz_full or nz) / nx,
               moving_ratio=nz / nx, nts=nts, nx=macro[0], ny=macro[1], nz=macro[2], nt=macro[3])
    inp.update(inp_overrides or {})
    os.makedirs(root, exist_ok=True)
    write_input_py(os.path.join(root, "line.py"), inp)
    folder = os.path.join(root, "forcing", f"case{seed}")
    os.makedirs(folder, exist_ok=True)
    Ly = Lx * inp["asp_ratio_yx"]
    Lz = Lx * inp["asp_ratio_zx"]
    Nx, Ny, Nz, Nt = macro
    BC = Lx / (Nx - 3)
    x = np.linspace(0 - BC, Lx + BC, Nx)
    y = np.linspace(0 - BC, Ly + BC, Ny)
    z = np.linspace(0 - BC, Lz + BC, Nz)
    xx, yy, zz = np.meshgrid(x, y, z, indexing="ij")
    psi = (z0 - zz).reshape(Nx * Ny * Nz, order="F")
    t = np.linspace(0, t_end, Nt)
    img = cases.voronoi_substrate(nx + 1, ny + 1, grains, seed, periodic=True)
    ng = int(img.max())
    theta = cases.reference_angles(ng, seed)
    Rmax_um = R * 1e6
    temps = np.broadcast_to(np.asarray(temp, np.float32), (Nt,))
    T = np.repeat(temps, Nx * Ny * Nz).astype(np.float32)  # time is the slowest index
    _txt(f"{folder}/x.txt", x); _txt(f"{folder}/y.txt", y); _txt(f"{folder}/z.txt", z)
    _txt(f"{folder}/t.txt", t); _txt(f"{folder}/psi.txt", psi)
    _txt(f"{folder}/alpha.txt", img.ravel(), "%d")
    _txt(f"{folder}/theta.txt", theta)
    _txt(f"{folder}/G.txt", [G]); _txt(f"{folder}/Rmax.txt", [Rmax_um * 1e-6]); _txt(f"{folder}/V.txt", [0.0])
    _txt(f"{folder}/Nmax.txt", [inp["nuc_Nmax"] if nmax is None else nmax])
    _txt(f"{folder}/UC.txt", [G * Rmax_um / 1e6])
    _txt(f"{folder}/NG.txt", [ng], "%d"); _txt(f"{folder}/NN.txt", [2 * ng], "%d")
    _txt(f"{folder}/z0.txt", [z0]); _txt(f"{folder}/r0.txt", [0.9 * Lz]); _txt(f"{folder}/x0.txt", [0.0])
    _txt(f"{folder}/top.txt", [top]); _txt(f"{folder}/angle.txt", [0.0]); _txt(f"{folder}/min_angle.txt", [0.0])
    _txt(f"{folder}/order.txt", [1.0])
    h5.write(f"{folder}/Temp.h5", {"Temp": T})
    return inp
