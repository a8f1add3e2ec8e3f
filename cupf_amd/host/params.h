// params.h -- the reference's parameter block (3DgrainMPI/params.h:10-56),
// kept field for field so that input parsing and the derived constants read
// like the reference and hand over to the C ABI (pf_params) in one place.
#pragma once

struct GlobalConstants {
    // processing parameters
    int thermalType = 0;
    float G = 0, R = 0, delta = 0, kin_delta = 0, underCoolingRate = 0, V = 0;
    float underCoolingRate0 = 0, nuc_Nmax0 = 0;
    // physical parameters
    float k = 0, c_infty = 0, m_slope = 0, c_infm = 0, Dl = 0, GT = 0;
    float Dh = 0, d0 = 0, W0 = 0, L_cp = 0;
    float lamd = 0, tau0 = 0, beta0 = 0, mu_k = 0, lT = 0;
    float R_tilde = 0, Dl_tilde = 0, lT_tilde = 0, beta0_tilde = 0;
    float alpha0 = 0, U0 = 0, eps = 0, eta = 0;
    // geometry parameters
    float lx = 0, lxd = 0, lyd = 0, lzd = 0, tmax = 0;
    float xmin = 0, ymin = 0, zmin = 0;
    float asp_ratio_yx = 0, asp_ratio_zx = 0, moving_ratio = 0;
    float z0 = 0, r0 = 0, x0 = 0, top = 0, angle = 0, min_angle = 0;
    float mp_len = 0, order = 0;
    // grid parameters
    int nx = 0, ny = 0, nz = 0, nz_full = 0, fnx = 0, fny = 0, fnz = 0, fnz_f = 0, length = 0,
        full_length = 0, haloWidth = 1;
    int Mt = 0, preMt = 0, nts = 1, NUM_PF = 8;
    float dx = 0, dt = 0, hi = 0, cfl = 0, dt_sqrt = 0;
    int Nx = 0, Ny = 0, Nz = 0, Nt = 0;
    // initial/boundary condition parameters
    int ictype = 0;
    int bcX = 0, bcY = 0, bcZ = 0;
    int num_theta = 0, num_nodes = 0;
    // anisotropy parameters
    bool includeInterfaceAnisotropy = false;
    float cosa = 0, sina = 0, sqrt2 = 0, a_s = 0, epsilon = 0, a_12 = 0;
    // noise parameters
    int noi_period = 0, seed_val = 0;
    // nucleation parameters
    float Tmelt = 0, Ti = 0, Tliq = 0, Tsol = 0;
    float undcool_mean = 0, undcool_std = 0, nuc_Nmax = 0, nuc_rad = 0, pts_cell = 0;
    int walltime = 0;
};
