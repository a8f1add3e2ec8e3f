"""Shared drivers for the parity tests: the same case on the HIP path (C ABI)
and on the CPU oracle, stepped in the same order."""
from __future__ import annotations

import numpy as np

from cupf_amd import capi, cases


def interior(a: np.ndarray, K: int, fnx: int, fny: int, fnz: int) -> np.ndarray:
    v = a.reshape(K, fnz, fny, fnx)
    return v[:, 1:-1, 1:-1, 1:-1]


def gpu_context(case: cases.Case, psi_mac=None, T3d=None) -> capi.APTContext:
    ctx = capi.APTContext(case.params)
    ctx.set_coords(case.x, case.y, case.z)
    ctx.set_orientations(case.R9)
    if psi_mac is not None or T3d is not None:
        ctx.set_thermal_macro(psi_mac, T3d)
    ctx.init_from_phi_alpha(case.phi0, case.alpha0)
    return ctx


def oracle_sim(case: cases.Case, precision="f32", psi_mac=None, T3d=None):
    import oracle
    sim = oracle.OracleSim(case.params, case.x, case.y, case.z, case.R9, precision,
                           psi_mac=psi_mac, T3d=T3d)
    sim.init_from_phi_alpha(case.phi0, case.alpha0)
    return sim


def compare_states(ctx: capi.APTContext, sim, label=""):
    """Fill the halo of both current buffers, then compare everything.
    Returns (n_arg_mismatch, max_abs_phi_diff)."""
    ctx.apply_bc(0)
    sim.apply_bc(0)
    ph_g, arg_g = ctx.download_state(0)
    ph_o, arg_o = sim.state(0)
    n_bad = int((arg_g != arg_o).sum())
    dphi = float(np.abs(ph_g.astype(np.float64) - ph_o.astype(np.float64)).max())
    return n_bad, dphi


def first_mismatch(ctx, sim, K, fnx, fny, fnz):
    ph_g, arg_g = ctx.download_state(0)
    ph_o, arg_o = sim.state(0)
    bad = np.nonzero((arg_g != arg_o) | (ph_g != ph_o))[0]
    if bad.size == 0:
        return None
    q = int(bad[0])
    L = fnx * fny * fnz
    s, C = divmod(q, L)
    k, r = divmod(C, fnx * fny)
    j, i = divmod(r, fnx)
    return dict(slot=s, i=i, j=j, k=k, gpu=(float(ph_g[q]), int(arg_g[q])),
                oracle=(float(ph_o[q]), int(arg_o[q])),
                gpu_point=arg_g[C::L].tolist(), oracle_point=arg_o[C::L].tolist())
