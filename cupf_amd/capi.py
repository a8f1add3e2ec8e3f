"""ctypes binding of libcupf_amd.so (the C ABI declared in include/pf_capi.h).

The shared library is built in-tree (cupf_amd/lib/libcupf_amd.so) by
``__graft_entry__.build()``.  There is no fallback: if the library is missing
or no GPU is present, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcupf_amd.so")

PF_OK = 0
PF_STEP_INITIAL_UC = 1
PF_STEP_NUCLEATE = 2
PF_BATCH_GRAPH = 4
PF_BATCH_AUTOTUNE = 8
PF_COLLECT_CLEAR_ALPHA = 1
PF_COLLECT_NO_BC = 2

THERMAL_NONE, THERMAL_LINE, THERMAL_CYLINDER, THERMAL_CONE, THERMAL_PARABOLOID, THERMAL_LINE_3DGRAIN = range(6)

# Field names and order of `struct pf_params` (include/pf_capi.h).
PARAM_FIELDS = [
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
    ("tau0", ctypes.c_float),
    ("seed", ctypes.c_uint64),
]


class PFParams(ctypes.Structure):
    _fields_ = PARAM_FIELDS

    @classmethod
    def from_dict(cls, d: dict) -> "PFParams":
        p = cls()
        for name, _ in PARAM_FIELDS:
            if name in d:
                setattr(p, name, d[name])
        return p


class PFError(RuntimeError):
    pass


_lib = None


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if a is not None else None


def load() -> ctypes.CDLL:
    """Load libcupf_amd.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PFError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    lib = ctypes.CDLL(LIB_PATH)
    c_int, c_float, vp = ctypes.c_int, ctypes.c_float, ctypes.c_void_p
    fp, ip = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)
    llp = ctypes.POINTER(ctypes.c_longlong)
    sig = {
        "pf_last_error": ([], ctypes.c_char_p),
        "pf_build_info": ([], ctypes.c_char_p),
        "pf_create": ([ctypes.POINTER(PFParams), c_int, ctypes.POINTER(vp)], c_int),
        "pf_destroy": ([vp], c_int),
        "pf_set_coords": ([vp, fp, fp, fp], c_int),
        "pf_set_orientations": ([vp, fp, c_int], c_int),
        "pf_orientations_from_angles": ([fp, c_int, fp], c_int),
        "pf_set_thermal_macro": ([vp, fp, fp], c_int),
        "pf_init_from_phi_alpha": ([vp, fp, ip], c_int),
        "pf_reinit_from_collect": ([vp], c_int),
        "pf_upload_state": ([vp, c_int, fp, ip], c_int),
        "pf_download_state": ([vp, c_int, fp, ip], c_int),
        "pf_apply_bc": ([vp, c_int], c_int),
        "pf_step": ([vp, c_int, c_int, c_int], c_int),
        "pf_step_timed": ([vp, c_int, c_int, c_int, fp], c_int),
        "pf_step_profiled": ([vp, c_int, c_int, c_int, fp, fp, fp], c_int),
        "pf_collect": ([vp, c_int, c_int, fp, ip], c_int),
        "pf_qoi_volume": ([vp, c_int, ip, ip, llp], c_int),
        "pf_qoi_volume_noalloc": ([vp, c_int, ip, ip, llp], c_int),
        "pf_qoi_junctions3d": ([vp, ip, c_int, ip], c_int),
        "pf_qoi_line": ([vp, c_int, ip, ip, ip], c_int),
        "pf_qoi_tip": ([vp, c_int, ip], c_int),
        "pf_md_enable": ([vp, c_int], c_int),
        "pf_md_tip": ([vp, ip], c_int),
        "pf_md_shift": ([vp, c_int], c_int),
        "pf_md_download": ([vp, ip, ip, fp], c_int),
        "pf_nucl_reset": ([vp, c_int], c_int),
        "pf_nucl_events": ([vp, llp], c_int),
        "pf_overflow_count": ([vp, llp], c_int),
        "pf_sync": ([vp], c_int),
        "pf_activity": ([vp, c_int, ctypes.c_void_p, llp, llp], c_int),
        "pf_batch_create": ([ctypes.POINTER(vp), c_int, ctypes.POINTER(vp)], c_int),
        "pf_batch_destroy": ([vp], c_int),
        "pf_batch_step": ([vp, c_int, ip, c_int], c_int),
        "pf_batch_set_kchunk": ([vp, c_int], c_int),
        "pf_batch_get_kchunk": ([vp, ip], c_int),
        "pf_copy_bandwidth": ([c_int, ctypes.c_ulonglong, c_int, fp], c_int),
    }
    for name, (argt, rest) in sig.items():
        f = getattr(lib, name)
        f.argtypes = argt
        f.restype = rest
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    """Names the header declares as entry points (parsed from include/pf_capi.h)."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "pf_capi.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pf_[a-z0-9_]+)\s*\(", txt, re.M)))


def _check(rc: int) -> None:
    if rc != PF_OK:
        msg = load().pf_last_error().decode(errors="replace")
        raise PFError(f"cupf_amd error {rc}: {msg}")


def orientations_from_angles(theta: np.ndarray, num_theta: int) -> np.ndarray:
    """R9 rows (NG+1, 9) from the reference's [0, thx(1..NG), thz(1..NG)] array."""
    th = np.ascontiguousarray(theta, dtype=np.float32)
    out = np.zeros((num_theta + 1, 9), dtype=np.float32)
    _check(load().pf_orientations_from_angles(_fptr(th), num_theta, _fptr(out)))
    return out


def copy_bandwidth(nbytes: int = 1 << 30, iters: int = 10, device: int = 0) -> float:
    """GB/s (read + write) of a plain copy kernel on the device."""
    g = ctypes.c_float()
    _check(load().pf_copy_bandwidth(device, nbytes, iters, ctypes.byref(g)))
    return g.value


class APTContext:
    """One device-resident simulation behind the C ABI."""

    def __init__(self, params: dict, device: int = 0):
        lib = load()
        self.params = dict(params)
        self.K = int(params["K"])
        self.fnx, self.fny, self.fnz = int(params["fnx"]), int(params["fny"]), int(params["fnz"])
        self.L = self.fnx * self.fny * self.fnz
        self._p = PFParams.from_dict(params)
        h = ctypes.c_void_p()
        _check(lib.pf_create(ctypes.byref(self._p), device, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            load().pf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- set-up -------------------------------------------------------------
    def set_coords(self, x, y, z):
        x, y, z = (np.ascontiguousarray(v, dtype=np.float32) for v in (x, y, z))
        assert x.size == self.fnx and y.size == self.fny and z.size == self.fnz
        _check(load().pf_set_coords(self._h, _fptr(x), _fptr(y), _fptr(z)))

    def set_orientations(self, R9):
        R9 = np.ascontiguousarray(R9, dtype=np.float32).reshape(-1, 9)
        _check(load().pf_set_orientations(self._h, _fptr(R9), R9.shape[0]))

    def set_thermal_macro(self, psi_mac=None, T3d=None):
        psi = np.ascontiguousarray(psi_mac, dtype=np.float32) if psi_mac is not None else None
        T = np.ascontiguousarray(T3d, dtype=np.float32) if T3d is not None else None
        _check(load().pf_set_thermal_macro(self._h, _fptr(psi), _fptr(T)))

    def init_from_phi_alpha(self, phi0, alpha0):
        phi0 = np.ascontiguousarray(phi0, dtype=np.float32).ravel()
        alpha0 = np.ascontiguousarray(alpha0, dtype=np.int32).ravel()
        assert phi0.size == self.L and alpha0.size == self.L
        _check(load().pf_init_from_phi_alpha(self._h, _fptr(phi0), _iptr(alpha0)))

    def reinit_from_collect(self):
        _check(load().pf_reinit_from_collect(self._h))

    def upload_state(self, ph, arg, which=0):
        ph = np.ascontiguousarray(ph, dtype=np.float32).ravel()
        arg = np.ascontiguousarray(arg, dtype=np.int32).ravel()
        assert ph.size == self.K * self.L and arg.size == self.K * self.L
        _check(load().pf_upload_state(self._h, which, _fptr(ph), _iptr(arg)))

    def download_state(self, which=0):
        ph = np.empty(self.K * self.L, dtype=np.float32)
        arg = np.empty(self.K * self.L, dtype=np.int32)
        _check(load().pf_download_state(self._h, which, _fptr(ph), _iptr(arg)))
        return ph, arg

    def apply_bc(self, which=0):
        _check(load().pf_apply_bc(self._h, which))

    # --- stepping -------------------------------------------------------------
    def step(self, nsteps: int, first_step: int = 1, flags: int = 0):
        _check(load().pf_step(self._h, nsteps, first_step, flags))

    def step_timed(self, nsteps: int, first_step: int = 1, flags: int = 0) -> float:
        ms = ctypes.c_float()
        _check(load().pf_step_timed(self._h, nsteps, first_step, flags, ctypes.byref(ms)))
        return ms.value

    def step_profiled(self, nsteps: int, first_step: int = 1, flags: int = 0):
        """(ms in the main step kernel, ms in the fix-up kernels, ms in the
        work-list plan kernel), each summed over nsteps."""
        m, f, pl = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        _check(load().pf_step_profiled(self._h, nsteps, first_step, flags, ctypes.byref(m),
                                       ctypes.byref(f), ctypes.byref(pl)))
        return m.value, f.value, pl.value

    def sync(self):
        _check(load().pf_sync(self._h))

    def collect(self, which=0, flags=0):
        phi = np.empty(self.L, dtype=np.float32)
        alpha = np.empty(self.L, dtype=np.int32)
        _check(load().pf_collect(self._h, which, flags, _fptr(phi), _iptr(alpha)))
        return phi, alpha

    def collect_on_device(self, which=0, flags=0):
        """pf_collect keeping phi/alpha on the device (for the device QoIs)."""
        _check(load().pf_collect(self._h, which, flags, None, None))

    def qoi_volume(self, max_id: int):
        """Device QOI::calculateQoIs on the collected alpha: (ids in
        first-seen scan order, their point counts, points with an id above
        max_id)."""
        counts = np.zeros(max_id + 1, dtype=np.int32)
        first = np.zeros(max_id + 1, dtype=np.int32)
        above = ctypes.c_longlong()
        _check(load().pf_qoi_volume(self._h, max_id, _iptr(counts), _iptr(first), ctypes.byref(above)))
        ids = np.nonzero(counts)[0]
        ids = ids[np.argsort(first[ids], kind="stable")]
        return ids.tolist(), counts[ids].tolist(), above.value

    def qoi_volume_noalloc(self, max_id: int):
        """qoi_volume through pf_qoi_volume_noalloc (no device allocation;
        max_id <= PF_NOALLOC_MAX_ID)."""
        counts = np.zeros(max_id + 1, dtype=np.int32)
        first = np.zeros(max_id + 1, dtype=np.int32)
        above = ctypes.c_longlong()
        _check(load().pf_qoi_volume_noalloc(self._h, max_id, _iptr(counts), _iptr(first), ctypes.byref(above)))
        ids = np.nonzero(counts)[0]
        ids = ids[np.argsort(first[ids], kind="stable")]
        return ids.tolist(), counts[ids].tolist(), above.value

    def qoi_tip(self, cur_tip: int) -> int:
        """Device QOI::sampleHeights on the collected alpha: the tip plane after cur_tip."""
        tip = ctypes.c_int()
        _check(load().pf_qoi_tip(self._h, cur_tip, ctypes.byref(tip)))
        return tip.value

    def qoi_line(self, cur_tip: int):
        """Device line QoIs on the collected alpha: (new tip plane, total and
        extra point counts per grain id 1..num_theta)."""
        ng = int(self.params["num_theta"])
        tip = ctypes.c_int()
        total = np.zeros(ng, dtype=np.int32)
        extra = np.zeros(ng, dtype=np.int32)
        _check(load().pf_qoi_line(self._h, cur_tip, ctypes.byref(tip), _iptr(total), _iptr(extra)))
        return tip.value, total, extra

    def qoi_junction_rows(self) -> np.ndarray:
        """Device QOI3D::searchJunctionsOnImage on the collected alpha: an
        int32 array of rows (i, j, k, largest count, 4 ascending ids) in scan
        order.  One call when the previous call's count still fits the row
        buffer, else a second with the exact count."""
        n = ctypes.c_int()
        cap = getattr(self, "_junction_cap", 1024)
        while True:
            rows = np.zeros(8 * cap, dtype=np.int32)
            _check(load().pf_qoi_junctions3d(self._h, _iptr(rows), cap, ctypes.byref(n)))
            if n.value <= cap:
                break
            cap = n.value + n.value // 4
        self._junction_cap = max(cap, 1024)
        return rows[:8 * n.value].reshape(-1, 8)

    def qoi_junctions3d(self):
        """qoi_junction_rows as tuples (i, j, k, largest count, (4 ids))."""
        return [(int(x[0]), int(x[1]), int(x[2]), int(x[3]), tuple(int(v) for v in x[4:]))
                for x in self.qoi_junction_rows()]

    # --- moving domain / nucleation / diagnostics --------------------------------
    def md_enable(self, fnz_full: int):
        self.fnz_full = fnz_full
        _check(load().pf_md_enable(self._h, fnz_full))

    def md_tip(self, tip: int) -> int:
        t = ctypes.c_int(tip)
        _check(load().pf_md_tip(self._h, ctypes.byref(t)))
        return t.value

    def md_shift(self, move_count: int):
        _check(load().pf_md_shift(self._h, move_count))

    def md_download(self):
        af = np.empty(self.fnx * self.fny * self.fnz_full, dtype=np.int32)
        loss = np.empty(int(self.params["num_theta"]), dtype=np.int32)
        z = np.empty(self.fnz, dtype=np.float32)
        _check(load().pf_md_download(self._h, _iptr(af), _iptr(loss), _fptr(z)))
        return af, loss, z

    def nucl_reset(self, mode: int):
        _check(load().pf_nucl_reset(self._h, mode))

    def nucl_events(self) -> int:
        v = ctypes.c_longlong()
        _check(load().pf_nucl_events(self._h, ctypes.byref(v)))
        return v.value

    def overflow_count(self) -> int:
        v = ctypes.c_longlong()
        _check(load().pf_overflow_count(self._h, ctypes.byref(v)))
        return v.value

    def activity(self, which: int = 0, with_map: bool = False):
        """(active tile planes, all interior tile planes[, map]) of a buffer's
        activity map; map is uint8 [fnz, nty, ntx] (64x4 column tiles)."""
        a, t = ctypes.c_longlong(), ctypes.c_longlong()
        p = self.params
        ntx, nty = (int(p["fnx"]) - 2 + 63) // 64, (int(p["fny"]) - 2 + 3) // 4
        m = np.zeros((int(p["fnz"]), nty, ntx), np.uint8) if with_map else None
        _check(load().pf_activity(self._h, which, None if m is None else m.ctypes.data,
                                  ctypes.byref(a), ctypes.byref(t)))
        return (a.value, t.value, m) if with_map else (a.value, t.value)


class Batch:
    """Several APTContexts of one device stepped together (pf_batch_*).

    Each member keeps its own step counter here; step(n) advances every member
    by n steps from its counter.  Members remain usable on their own between
    calls (download, collect, single steps).
    """

    def __init__(self, members, first_step=None):
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m._h for m in self.members])
        h = ctypes.c_void_p()
        _check(load().pf_batch_create(arr, len(self.members), ctypes.byref(h)))
        self._h = h
        self.next_step = list(first_step) if first_step is not None else [1] * len(self.members)

    def step(self, n: int, flags: int = 0) -> None:
        fs = (ctypes.c_int * len(self.members))(*self.next_step)
        _check(load().pf_batch_step(self._h, int(n), fs, int(flags)))
        self.next_step = [s + int(n) for s in self.next_step]

    @property
    def kchunk(self) -> int:
        v = ctypes.c_int()
        _check(load().pf_batch_get_kchunk(self._h, ctypes.byref(v)))
        return v.value

    @kchunk.setter
    def kchunk(self, v: int) -> None:
        _check(load().pf_batch_set_kchunk(self._h, int(v)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            load().pf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
