"""CPU oracle of the cuPF APT step -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
may import this package, and only as the checker.  The product (cupf_amd)
never imports, links or runs it.

``liboracle_f32.so`` / ``liboracle_f64.so`` are the same C restatement
(apt_oracle.c) built in float32 (reference op order, -ffp-contract=off, the
reference's implicit double promotions kept) and in float64.  ``OracleSim``
drives them in the reference's loop order (3DgrainMPI/APTPhaseField.cu:852-922):
halo fill of the source, APTrhs_psi into the other buffer, nucleation after
odd steps.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_HERE, "lib")

# Same leading layout as struct orc_params (apt_oracle.h).
ORC_FIELDS = [
    ("fnx", ctypes.c_int), ("fny", ctypes.c_int), ("fnz", ctypes.c_int),
    ("K", ctypes.c_int), ("num_theta", ctypes.c_int),
    ("bcx", ctypes.c_int), ("bcy", ctypes.c_int), ("bcz", ctypes.c_int),
    ("thermal_type", ctypes.c_int), ("gather27", ctypes.c_int),
    ("dt", ctypes.c_float), ("hi", ctypes.c_float), ("lamd", ctypes.c_float),
    ("L_cp", ctypes.c_float), ("kin_delta", ctypes.c_float), ("eps", ctypes.c_float),
    ("G", ctypes.c_float), ("R", ctypes.c_float), ("uc", ctypes.c_float), ("uc0", ctypes.c_float),
    ("z0", ctypes.c_float), ("r0", ctypes.c_float), ("lyd", ctypes.c_float), ("lzd", ctypes.c_float),
    ("V", ctypes.c_float), ("mp_len", ctypes.c_float), ("x0", ctypes.c_float), ("order", ctypes.c_float),
    ("tan_angle", ctypes.c_float), ("cos_angle", ctypes.c_float),
    ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int), ("Nt", ctypes.c_int),
    ("X0_mac", ctypes.c_float), ("Y0_mac", ctypes.c_float), ("Z0_mac", ctypes.c_float),
    ("t0_mac", ctypes.c_float), ("Dx_mac", ctypes.c_float), ("Dt_mac", ctypes.c_float),
    ("pts_cell", ctypes.c_int),
    ("nuc_rad", ctypes.c_float), ("undcool_mean", ctypes.c_float), ("undcool_std", ctypes.c_float),
    ("nuc_Nmax", ctypes.c_float), ("nuc_Nmax0", ctypes.c_float), ("dx", ctypes.c_float),
    ("sqrt2", ctypes.c_float),
]


class OrcParams(ctypes.Structure):
    _fields_ = ORC_FIELDS

    @classmethod
    def from_dict(cls, d: dict) -> "OrcParams":
        p = cls()
        for name, _ in ORC_FIELDS:
            if name in d:
                setattr(p, name, d[name])
        return p


_libs: dict = {}


def build() -> None:
    """Compile both precisions with oracle/Makefile."""
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def load(precision: str = "f32") -> ctypes.CDLL:
    if precision in _libs:
        return _libs[precision]
    path = os.path.join(LIBDIR, f"liboracle_{precision}.so")
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    vp, ip, fp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)
    P = ctypes.POINTER(OrcParams)
    i64p = ctypes.POINTER(ctypes.c_int64)
    c_int, c_double, u64 = ctypes.c_int, ctypes.c_double, ctypes.c_uint64
    sig = {
        "orc_set_bc": ([P, vp, ip], None),
        "orc_rhs": ([P, fp, fp, fp, fp, fp, vp, ip, vp, ip, c_double, c_int], ctypes.c_long),
        "orc_collect": ([P, vp, ip, vp, ip], None),
        "orc_ini": ([P, vp, ip, fp, ip], None),
        "orc_kine_ani": ([P, c_double, c_double, c_double, fp], c_double),
        "orc_thermal": ([P, fp, fp, fp, fp, c_int, c_int, c_int, c_double, c_int,
                         ctypes.POINTER(c_double)], c_int),
        "orc_interp3": ([fp, c_double, c_double, c_double, c_int, c_int, c_int, c_double], c_double),
        "orc_interp4": ([fp, c_double, c_double, c_double, c_double, c_int, c_int, c_int, c_int,
                         c_double, c_double], c_double),
        "orc_plane_sums": ([P, vp, i64p], None),
        "orc_tip_scan": ([P, i64p, c_int], c_int),
        "orc_move_frame": ([P, vp, ip, vp, ip, fp, ip, ip, ip, c_int], None),
        "orc_init_nucl_status": ([P, vp, ip, c_int, c_int, c_int], None),
        "orc_add_nucl": ([P, fp, fp, fp, fp, fp, vp, ip, ip, c_int, c_int, c_int, c_double, c_double,
                          c_int, u64, u64], c_int),
        "orc_nucl_ball": ([P, fp], None),
        "orc_nucl_prob": ([P, c_double, c_double, c_double], c_double),
        "orc_draw": ([u64, u64, u64, fp, ip], None),
    }
    for name, (a, r) in sig.items():
        f = getattr(lib, name)
        f.argtypes = a
        f.restype = r
    _libs[precision] = lib
    return lib


def _f(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _i(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _v(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class OracleSim:
    """Two-buffer CPU simulation in the reference's order."""

    def __init__(self, params: dict, x, y, z, R9, precision: str = "f32", psi_mac=None, T3d=None):
        self.precision = precision
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.lib = load(precision)
        self.params = dict(params)
        self.P = OrcParams.from_dict(params)
        self.K = int(params["K"])
        self.L = int(params["fnx"]) * int(params["fny"]) * int(params["fnz"])
        self.x = np.ascontiguousarray(x, np.float32)
        self.y = np.ascontiguousarray(y, np.float32)
        self.z = np.ascontiguousarray(z, np.float32)
        self.R9 = np.ascontiguousarray(R9, np.float32).reshape(-1, 9)
        self.psi_mac = None if psi_mac is None else np.ascontiguousarray(psi_mac, np.float32)
        self.T3d = None if T3d is None else np.ascontiguousarray(T3d, np.float32)
        n = self.K * self.L
        self.ph = [np.full(n, -1, self.dtype), np.full(n, -1, self.dtype)]
        self.arg = [np.full(n, -1, np.int32), np.full(n, -1, np.int32)]
        self.cur = 0
        self.overflow = 0
        self.phi = np.zeros(self.L, self.dtype)
        self.alpha = np.zeros(self.L, np.int32)
        self.phi0 = None
        self.nstatus = None
        self.ncall = 0
        self.nucl_events = 0

    # --- state ------------------------------------------------------------------
    def init_from_phi_alpha(self, phi0, alpha0):
        self.phi0 = np.ascontiguousarray(phi0, np.float32).ravel()
        self.alpha = np.ascontiguousarray(alpha0, np.int32).ravel().copy()
        self.phi = self.phi0.astype(self.dtype)
        for b in (0, 1):
            self.lib.orc_ini(self.P, _v(self.ph[b]), _i(self.arg[b]), _f(self.phi0), _i(self.alpha))
        self.cur = 0

    def reinit_from_collect(self):
        for b in (0, 1):
            self.lib.orc_ini(self.P, _v(self.ph[b]), _i(self.arg[b]), _f(self.phi0), _i(self.alpha))
        self.cur = 0

    def set_state(self, ph, arg, which=0):
        b = self.cur if which == 0 else 1 - self.cur
        self.ph[b] = np.ascontiguousarray(ph, self.dtype).ravel().copy()
        self.arg[b] = np.ascontiguousarray(arg, np.int32).ravel().copy()

    def state(self, which=0):
        b = self.cur if which == 0 else 1 - self.cur
        return self.ph[b], self.arg[b]

    def apply_bc(self, which=0):
        b = self.cur if which == 0 else 1 - self.cur
        self.lib.orc_set_bc(self.P, _v(self.ph[b]), _i(self.arg[b]))

    # --- stepping ------------------------------------------------------------------
    def t_of_step(self, s: int):
        if self.precision == "f32":
            f32 = np.float32
            return float(f32(f32(f32(s) * f32(self.params["dt"])) * f32(self.params["tau0"])))
        return float(s) * float(self.params["dt"]) * float(self.params["tau0"])

    def step(self, nsteps: int, first_step: int = 1, use_initial_uc: bool = False,
             nucleate: bool = False, seed: int = 0):
        for s in range(first_step, first_step + nsteps):
            src, dst = self.cur, 1 - self.cur
            self.lib.orc_set_bc(self.P, _v(self.ph[src]), _i(self.arg[src]))
            t = self.t_of_step(s)
            self.overflow += self.lib.orc_rhs(
                self.P, _f(self.x), _f(self.y), _f(self.z), _f(self.R9), _f(self.psi_mac),
                _v(self.ph[src]), _i(self.arg[src]), _v(self.ph[dst]), _i(self.arg[dst]), t,
                int(use_initial_uc))
            self.cur = dst
            if nucleate and (s & 1):
                self.add_nucl(t, use_initial_uc, seed)

    # --- single-field outputs ---------------------------------------------------------
    def collect(self, which=0, clear_alpha=False, bc=True):
        b = self.cur if which == 0 else 1 - self.cur
        if bc:
            self.lib.orc_set_bc(self.P, _v(self.ph[b]), _i(self.arg[b]))
        if clear_alpha:
            self.alpha[:] = 0
        self.lib.orc_collect(self.P, _v(self.ph[b]), _i(self.arg[b]), _v(self.phi), _i(self.alpha))
        return self.phi.copy(), self.alpha.copy()

    # --- moving domain ------------------------------------------------------------------
    def md_enable(self, fnz_full: int):
        fnx, fny = int(self.params["fnx"]), int(self.params["fny"])
        self.alpha_full = np.zeros(fnx * fny * fnz_full, np.int32)
        self.loss = np.zeros(int(self.params["num_theta"]), np.int32)

    def md_tip(self, tip: int) -> int:
        b = self.cur
        self.lib.orc_set_bc(self.P, _v(self.ph[b]), _i(self.arg[b]))
        sums = np.zeros(int(self.params["fnz"]), np.int64)
        self.lib.orc_plane_sums(self.P, _v(self.ph[b]),
                                sums.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        return self.lib.orc_tip_scan(self.P, sums.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), tip)

    def md_shift(self, move_count: int):
        b = self.cur
        self.lib.orc_collect(self.P, _v(self.ph[b]), _i(self.arg[b]), _v(self.phi), _i(self.alpha))
        self.lib.orc_move_frame(self.P, _v(self.ph[b]), _i(self.arg[b]), _v(self.ph[1 - b]),
                                _i(self.arg[1 - b]), _f(self.z), _i(self.alpha), _i(self.alpha_full),
                                _i(self.loss), move_count)

    # --- nucleation --------------------------------------------------------------------
    def nucl_reset(self, mode: int):
        p = int(self.params["pts_cell"])
        w = 2 * p + 1
        self.cnx = (int(self.params["fnx"]) - 3) // w
        self.cny = (int(self.params["fny"]) - 3) // w
        self.cnz = (int(self.params["fnz"]) - 3) // w
        self.nstatus = np.zeros(max(self.cnx * self.cny * self.cnz, 1), np.int32)
        self.ball = np.zeros(w ** 3, np.float32)
        self.lib.orc_nucl_ball(self.P, _f(self.ball))
        if mode == 1:
            self.lib.orc_init_nucl_status(self.P, _v(self.phi0.astype(self.dtype)), _i(self.nstatus),
                                          self.cnx, self.cny, self.cnz)

    def add_nucl(self, t, use_initial_uc, seed):
        dt_nuc = np.float32(np.float32(2.0) * np.float32(self.params["dt"])) * np.float32(self.params["tau0"])
        if self.precision == "f64":
            dt_nuc = 2.0 * float(self.params["dt"]) * float(self.params["tau0"])
        b = self.cur
        self.nucl_events += self.lib.orc_add_nucl(
            self.P, _f(self.x), _f(self.y), _f(self.z), _f(self.T3d), _f(self.ball), _v(self.ph[b]),
            _i(self.arg[b]), _i(self.nstatus), self.cnx, self.cny, self.cnz, float(dt_nuc), float(t),
            int(use_initial_uc), seed, self.ncall)
        self.ncall += 1
