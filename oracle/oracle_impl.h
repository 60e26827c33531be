/*
 * Precision-generic body of the CPU oracle. Included twice by
 * homelbm_oracle.c with R = float / double and SFX = f32 / f64.
 * TEST INFRASTRUCTURE ONLY (see homelbm_oracle.h).
 */
#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

static void FN(wall_velocity_2d)(const hlo_env* e, int sid, R hx, R hy, R up[2], R r[2]);
static void FN(wall_velocity_3d)(const hlo_env* e, int sid, R hx, R hy, R hz, R up[3], R r[3]);
static void FN(d3q27_feq_S)(R rho, R ux, R uy, R uz, R S[6]);
static int FN(capsule_cut)(const hlo_env* e, const float* caps, int ncap, R x, R y, R z, int cx, int cy, int cz, R* t_hit);
static void FN(link_force_3d)(const hlo_env* e, int sid, int i, R hx, R hy, R hz, const R m[10], double o[6]);

/* ------------------------------------------------------------------ D2Q9 */

/* mlCalDistribution, envs/lbm/lbm_func.py:685-788 (same operation order). */
static R FN(d2q9_recon)(R rho, R ux, R uy, R Sxx, R Syy, R Sxy, int i)
{
    const R one = 1, two = 2, three = 3, half = (R)0.5, c45 = (R)4.5, nine = 9;
    R Axxy = Sxx * uy + two * Sxy * ux - two * ux * ux * uy;
    R Axyy = Syy * ux + two * Sxy * uy - two * ux * uy * uy;
    const R w0 = (R)4 / (R)9, w1 = (R)1 / (R)9, w5 = (R)1 / (R)36;
    switch (i) {
    case 0: return w0 * rho * (one - half * three * Sxx - half * three * Syy);
    case 1: return w1 * rho * (one + three * ux + three * Sxx - half * three * Syy - c45 * Axyy);
    case 2: return w1 * rho * (one - three * ux + three * Sxx - half * three * Syy + c45 * Axyy);
    case 3: return w1 * rho * (one + three * uy - half * three * Sxx + three * Syy - c45 * Axxy);
    case 4: return w1 * rho * (one - three * uy - half * three * Sxx + three * Syy + c45 * Axxy);
    case 5: return w5 * rho * (one + three * (ux + uy) + three * (Sxx + Syy + three * Sxy) + nine * (Axyy + Axxy));
    case 6: return w5 * rho * (one - three * ux + three * uy + three * (Sxx + Syy - three * Sxy) + nine * (Axxy - Axyy));
    case 7: return w5 * rho * (one - three * ux - three * uy + three * (Sxx + Syy + three * Sxy) - nine * (Axxy + Axyy));
    default: return w5 * rho * (one + three * ux - three * uy + three * (Sxx + Syy - three * Sxy) - nine * (Axxy - Axyy));
    }
}

/* Equilibrium second moment (S part) of a D2Q9 cell: f_eq_home,
 * envs/lbm/lbm_func.py:660-682 (populations clamped at 0). */
static void FN(d2q9_feq_S)(R rho, R ux, R uy, R S[3])
{
    R pop[9];
    R U_sqr = ux * ux + uy * uy;
    for (int i = 0; i < 9; ++i) {
        R cu = (R)D2_CX[i] * ux + (R)D2_CY[i] * uy;
        pop[i] = rho * FN(d2_w)[i] * ((R)1 + (R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * U_sqr);
        if (pop[i] < (R)0) pop[i] = 0;
    }
    R invRho = (R)1 / rho;
    R pixx = pop[1] + pop[2] + pop[5] + pop[6] + pop[7] + pop[8];
    R piyy = pop[3] + pop[4] + pop[5] + pop[6] + pop[7] + pop[8];
    R pixy = pop[5] + pop[7] - pop[6] - pop[8];
    S[0] = pixx * invRho - FN(cs2_2d);
    S[1] = piyy * invRho - FN(cs2_2d);
    S[2] = pixy * invRho;
}

/* Post-collision moments of interior cell (x, y): stream_and_collide,
 * envs/lbm/lbm_func.py:322-467, with the cut-link test replaced by the mask
 * rule occ(x) != occ(x - c_i) and the wall point at the link midpoint. */
static void FN(d2q9_cell)(const hlo_env* e, const R* src, int x, int y, R out[6])
{
    const int64_t st = e->stride;
    const int nx = e->nx;
    const int64_t c = (int64_t)x + (int64_t)nx * y;
    R pop[9];
    R rho_c = src[c], ux_c = src[st + c], uy_c = src[2 * st + c];
    R Sxx_c = src[3 * st + c], Syy_c = src[4 * st + c], Sxy_c = src[5 * st + c];
    int occ_c = e->mask ? (e->mask[c] != 0) : 0;
    for (int i = 0; i < 9; ++i) {
        int x1 = x - D2_CX[i], y1 = y - D2_CY[i];
        int64_t s = (int64_t)x1 + (int64_t)nx * y1;
        int cut = 0, sid = -1;
        if (e->mask) {
            int occ_s = e->mask[s] != 0;
            if (occ_s != occ_c) {
                cut = 1;
                sid = occ_c ? e->mask[c] - 1 : e->mask[s] - 1;
            }
        }
        if (cut) {
            R up[2];
            FN(wall_velocity_2d)(e, sid, (R)x - (R)0.5 * (R)D2_CX[i], (R)y - (R)0.5 * (R)D2_CY[i], up, NULL);
            R Sxx_s = up[0] * up[0] + Sxx_c - ux_c * ux_c;
            R Syy_s = up[1] * up[1] + Syy_c - uy_c * uy_c;
            R Sxy_s = up[0] * up[1] + Sxy_c - ux_c * uy_c;
            pop[i] = FN(d2q9_recon)(rho_c, up[0], up[1], Sxx_s, Syy_s, Sxy_s, i);
            if (pop[i] < (R)0) pop[i] = 0;
        } else {
            pop[i] = FN(d2q9_recon)(src[s], src[st + s], src[2 * st + s], src[3 * st + s],
                                    src[4 * st + s], src[5 * st + s], i);
            if (pop[i] < (R)0) pop[i] = 0;
        }
    }
    R Fx = 0, Fy = 0;
    if (e->force) { Fx = (R)e->force[c]; Fy = (R)e->force[st + c]; }
    R rhoVar = pop[0] + pop[1] + pop[2] + pop[3] + pop[4] + pop[5] + pop[6] + pop[7] + pop[8];
    R invRho = (R)1 / rhoVar;
    R ux = (pop[1] - pop[2] + pop[5] - pop[6] - pop[7] + pop[8] + (R)0.5 * Fx) * invRho;
    R uy = (pop[3] - pop[4] + pop[5] + pop[6] - pop[7] - pop[8] + (R)0.5 * Fy) * invRho;
    R pixx = pop[1] + pop[2] + pop[5] + pop[6] + pop[7] + pop[8];
    R piyy = pop[3] + pop[4] + pop[5] + pop[6] + pop[7] + pop[8];
    R pixy = pop[5] + pop[7] - pop[6] - pop[8];
    R Omega = (R)1 / (((R)e->nu) * (R)3 + (R)0.5);
    R Ru2 = rhoVar * ux * ux;
    R Rv2 = rhoVar * uy * uy;
    R pixx_part = (pixx - piyy) / (R)2;
    R piyy_part = (piyy - pixx) / (R)2;
    R one = 1;
    pixy = (one - Omega) * pixy + ux * uy * rhoVar * Omega + (one - (R)0.5 * Omega) * (Fy * ux + Fx * uy);
    R pixx_n = rhoVar / (R)3 + pixx_part * (one - Omega) + (one + Omega) / (R)2 * Ru2
             + (one - Omega) / (R)2 * Rv2 + Fx * ux + (one - Omega) / (R)2 * (Fx * ux - Fy * uy);
    R piyy_n = rhoVar / (R)3 + piyy_part * (one - Omega) + (one + Omega) / (R)2 * Rv2
             + (one - Omega) / (R)2 * Ru2 + Fy * uy + (one - Omega) / (R)2 * (Fy * uy - Fx * ux);
    out[0] = rhoVar;
    out[1] = ux + Fx * invRho / (R)2;
    out[2] = uy + Fy * invRho / (R)2;
    out[3] = pixx_n * invRho - FN(cs2_2d);
    out[4] = piyy_n * invRho - FN(cs2_2d);
    out[5] = pixy * invRho;
}

/* One boundary cell b from its inner neighbour n: apply_bc_core,
 * envs/lbm/lbm_func.py:634-657 (2D) and envs/lbm3d/lbm_func_3d.py:627-648 (3D).
 * Type 0: velocity, type 1: zero gradient (reference). Type 2 (free slip,
 * no reference counterpart): u_b = u_n with the face-normal component zeroed. */
static void FN(bc_apply)(const hlo_env* e, R* dst, int64_t b, int64_t n, int face, int axis)
{
    const int64_t st = e->stride;
    const int d3 = e->lattice == 27;
    const int nu_ = d3 ? 3 : 2;
    int type = e->bc_type[face];
    R u_b[3], u_n[3];
    for (int a = 0; a < nu_; ++a) u_n[a] = dst[(1 + a) * st + n];
    for (int a = 0; a < nu_; ++a) u_b[a] = dst[(1 + a) * st + b];
    if (type == 0) {
        for (int a = 0; a < nu_; ++a) u_b[a] = (R)e->bc_value[face][a];
    } else if (type == 1) {
        for (int a = 0; a < nu_; ++a) u_b[a] = u_n[a];
    } else if (type == 2) {
        for (int a = 0; a < nu_; ++a) u_b[a] = (a == axis) ? (R)0 : u_n[a];
    }
    for (int a = 0; a < nu_; ++a) dst[(1 + a) * st + b] = u_b[a];
    R rho_n = dst[n];
    dst[b] = rho_n;
    if (!d3) {
        R Sb[3], Sn[3];
        FN(d2q9_feq_S)(rho_n, u_b[0], u_b[1], Sb);
        FN(d2q9_feq_S)(rho_n, u_n[0], u_n[1], Sn);
        for (int k = 0; k < 3; ++k)
            dst[(3 + k) * st + b] = Sb[k] - Sn[k] + dst[(3 + k) * st + n];
    } else {
        /* f_eq_home_3d returns (xx, xy, xz, yy, yz, zz); storage is (xx, yy, zz, xy, xz, yz). */
        R Sb[6], Sn[6];
        FN(d3q27_feq_S)(rho_n, u_b[0], u_b[1], u_b[2], Sb);
        FN(d3q27_feq_S)(rho_n, u_n[0], u_n[1], u_n[2], Sn);
        static const int slot[6] = {4, 7, 8, 5, 9, 6};
        for (int k = 0; k < 6; ++k)
            dst[slot[k] * st + b] = Sb[k] - Sn[k] + dst[slot[k] * st + n];
    }
}

/* Boundary passes in a fixed order: x faces, then y faces, then z faces.
 * The reference's apply_bc / apply_bc_3d run all faces in one launch and
 * race at edges and corners (lbm_func.py:616-631, lbm_func_3d.py:613-624);
 * this sequential order is one of the outcomes that launch can produce. */
static void FN(bc_passes)(const hlo_env* e, R* dst)
{
    const int nx = e->nx, ny = e->ny, nz = e->nz;
    const int d3 = e->lattice == 27;
    /* face indices: 2D left 0, top 1, right 2, bottom 3;
     *               3D left 0, right 1, top 2, bottom 3, front 4, back 5. */
    const int f_left = 0, f_right = d3 ? 1 : 2, f_top = d3 ? 2 : 1, f_bottom = 3;
#define IDX(X, Y, Z) ((int64_t)(X) + (int64_t)nx * ((int64_t)(Y) + (int64_t)ny * (Z)))
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
            FN(bc_apply)(e, dst, IDX(0, y, z), IDX(1, y, z), f_left, 0);
            FN(bc_apply)(e, dst, IDX(nx - 1, y, z), IDX(nx - 2, y, z), f_right, 0);
        }
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x) {
            FN(bc_apply)(e, dst, IDX(x, ny - 1, z), IDX(x, ny - 2, z), f_top, 1);
            FN(bc_apply)(e, dst, IDX(x, 0, z), IDX(x, 1, z), f_bottom, 1);
        }
    if (d3) {
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                FN(bc_apply)(e, dst, IDX(x, y, nz - 1), IDX(x, y, nz - 2), 4, 2);
                FN(bc_apply)(e, dst, IDX(x, y, 0), IDX(x, y, 1), 5, 2);
            }
    }
#undef IDX
}

/* ----------------------------------------------------------------- D3Q27 */

/* mlCalEquilibrium3D_s, envs/lbm3d/lbm_func_3d.py:224-251 (argument order
 * rho, ux, uy, uz, Sxx, Sxy, Sxz, Syy, Syz, Szz as in the reference). */
static R FN(d3q27_recon_s)(R rho, R ux, R uy, R uz, R Sxx, R Sxy, R Sxz, R Syy, R Syz, R Szz,
                           R dx, R dy, R dz, R wi)
{
    R cu = dx * ux + dy * uy + dz * uz;
    R U_sqr = ux * ux + uy * uy + uz * uz;
    R third = FN(third);
    R Qxx = dx * dx - third, Qyy = dy * dy - third, Qzz = dz * dz - third;
    R Qxy = dx * dy, Qxz = dx * dz, Qyz = dy * dz;
    R f_neq = (R)1.5 * (Qxx * Sxx + Qyy * Syy + Qzz * Szz + (R)2 * (Qxy * Sxy + Qxz * Sxz + Qyz * Syz));
    return wi * rho * ((R)1 + (R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * U_sqr + f_neq);
}

/* Equilibrium S of a D3Q27 cell: f_eq_home_3d, envs/lbm3d/lbm_func_3d.py:651-686
 * (no clamp). S = (xx, xy, xz, yy, yz, zz). */
static void FN(d3q27_feq_S)(R rho, R ux, R uy, R uz, R S[6])
{
    R p[27];
    R U_sqr = ux * ux + uy * uy + uz * uz;
    for (int i = 0; i < 27; ++i) {
        R cu = (R)D3_CX[i] * ux + (R)D3_CY[i] * uy + (R)D3_CZ[i] * uz;
        p[i] = FN(d3_w)[i] * rho * ((R)1 + (R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * U_sqr);
    }
    R invRho = (R)1 / rho;
    R pixx = p[1] + p[2] + p[7] + p[8] + p[9] + p[10] + p[13] + p[14] + p[15] + p[16] + p[19] + p[20]
           + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
    R pixy = (p[7] + p[8] + p[19] + p[20] + p[21] + p[22]) - (p[13] + p[14] + p[23] + p[24] + p[25] + p[26]);
    R pixz = (p[9] + p[10] + p[19] + p[20] + p[23] + p[24]) - (p[15] + p[16] + p[21] + p[22] + p[25] + p[26]);
    R piyy = p[3] + p[4] + p[7] + p[8] + p[11] + p[12] + p[13] + p[14] + p[17] + p[18] + p[19] + p[20]
           + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
    R piyz = (p[11] + p[12] + p[19] + p[20] + p[25] + p[26]) - (p[17] + p[18] + p[21] + p[22] + p[23] + p[24]);
    R pizz = p[5] + p[6] + p[9] + p[10] + p[11] + p[12] + p[15] + p[16] + p[17] + p[18] + p[19] + p[20]
           + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
    S[0] = pixx * invRho - FN(cs2_3d);
    S[1] = pixy * invRho;
    S[2] = pixz * invRho;
    S[3] = piyy * invRho - FN(cs2_3d);
    S[4] = piyz * invRho;
    S[5] = pizz * invRho - FN(cs2_3d);
}

/* Post-collision moments of interior cell (x, y, z): stream_and_collide_3d,
 * envs/lbm3d/lbm_func_3d.py:364-604. Cut links use the mask rule and the link
 * midpoint; the cut branch is not clamped (lbm_func_3d.py:487-489), the plain
 * branch is (525-526); moments are accumulated by sign group in ascending i
 * (528-562). */
static void FN(d3q27_cell_rule)(const hlo_env* e, const R* src, int x, int y, int z, const float* caps, int ncap,
                                double* forces, R out[10])
{
    const int64_t st = e->stride;
    const int nx = e->nx, ny = e->ny;
    const int64_t c = (int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z);
    R rho_c = src[c], ux_c = src[st + c], uy_c = src[2 * st + c], uz_c = src[3 * st + c];
    R Sxx_c = src[4 * st + c], Syy_c = src[5 * st + c], Szz_c = src[6 * st + c];
    R Sxy_c = src[7 * st + c], Sxz_c = src[8 * st + c], Syz_c = src[9 * st + c];
    int occ_c = e->mask ? (e->mask[c] != 0) : 0;
    R rho_acc = 0, uxp = 0, uxm = 0, uyp = 0, uym = 0, uzp = 0, uzm = 0;
    R pixx = 0, piyy = 0, pizz = 0, pixy_p = 0, pixy_m = 0, pixz_p = 0, pixz_m = 0, piyz_p = 0, piyz_m = 0;
    for (int i = 0; i < 27; ++i) {
        R dx = (R)D3_CX[i], dy = (R)D3_CY[i], dz = (R)D3_CZ[i], wi = FN(d3_w)[i];
        int x1 = x - D3_CX[i], y1 = y - D3_CY[i], z1 = z - D3_CZ[i];
        int64_t s = (int64_t)x1 + (int64_t)nx * ((int64_t)y1 + (int64_t)ny * z1);
        int cut = 0, sid = -1;
        /* the wall point: the link midpoint under the mask rule, the hit
         * point x - c_i t (lbm_func_3d.py:446) with capsules */
        R hx = (R)x - (R)0.5 * dx, hy = (R)y - (R)0.5 * dy, hz = (R)z - (R)0.5 * dz;
        if (caps) {
            R t = 0;
            if (i > 0) sid = FN(capsule_cut)(e, caps, ncap, (R)x, (R)y, (R)z, D3_CX[i], D3_CY[i], D3_CZ[i], &t);
            if (sid >= 0) {
                cut = 1;
                hx = (R)x - dx * t;
                hy = (R)y - dy * t;
                hz = (R)z - dz * t;
            }
        } else if (e->mask) {
            int occ_s = e->mask[s] != 0;
            if (occ_s != occ_c) {
                cut = 1;
                sid = occ_c ? e->mask[c] - 1 : e->mask[s] - 1;
            }
        }
        R pop_i;
        if (cut) {
            R up[3];
            FN(wall_velocity_3d)(e, sid, hx, hy, hz, up, NULL);
            R Sxx_s = up[0] * up[0] + Sxx_c - ux_c * ux_c;
            R Sxy_s = up[0] * up[1] + Sxy_c - ux_c * uy_c;
            R Sxz_s = up[0] * up[2] + Sxz_c - ux_c * uz_c;
            R Syy_s = up[1] * up[1] + Syy_c - uy_c * uy_c;
            R Syz_s = up[1] * up[2] + Syz_c - uy_c * uz_c;
            R Szz_s = up[2] * up[2] + Szz_c - uz_c * uz_c;
            pop_i = FN(d3q27_recon_s)(rho_c, up[0], up[1], up[2], Sxx_s, Sxy_s, Sxz_s, Syy_s, Syz_s, Szz_s,
                                      dx, dy, dz, wi);
            if (forces && sid < e->nsolid) {
                const R mc[10] = {rho_c, ux_c, uy_c, uz_c, Sxx_c, Syy_c, Szz_c, Sxy_c, Sxz_c, Syz_c};
                FN(link_force_3d)(e, sid, i, hx, hy, hz, mc, forces + 6 * sid);
            }
        } else {
            pop_i = FN(d3q27_recon_s)(src[s], src[st + s], src[2 * st + s], src[3 * st + s],
                                      src[4 * st + s], src[7 * st + s], src[8 * st + s],
                                      src[5 * st + s], src[9 * st + s], src[6 * st + s], dx, dy, dz, wi);
            if (pop_i < (R)0) pop_i = 0;
        }
        rho_acc = rho_acc + pop_i;
        if (dx > 0) { uxp = uxp + pop_i; pixx = pixx + pop_i; }
        else if (dx < 0) { uxm = uxm + pop_i; pixx = pixx + pop_i; }
        if (dy > 0) { uyp = uyp + pop_i; piyy = piyy + pop_i; }
        else if (dy < 0) { uym = uym + pop_i; piyy = piyy + pop_i; }
        if (dz > 0) { uzp = uzp + pop_i; pizz = pizz + pop_i; }
        else if (dz < 0) { uzm = uzm + pop_i; pizz = pizz + pop_i; }
        R dxy = dx * dy, dxz = dx * dz, dyz = dy * dz;
        if (dxy > 0) pixy_p = pixy_p + pop_i; else if (dxy < 0) pixy_m = pixy_m + pop_i;
        if (dxz > 0) pixz_p = pixz_p + pop_i; else if (dxz < 0) pixz_m = pixz_m + pop_i;
        if (dyz > 0) piyz_p = piyz_p + pop_i; else if (dyz < 0) piyz_m = piyz_m + pop_i;
    }
    R Fx = 0, Fy = 0, Fz = 0;
    if (e->force) { Fx = (R)e->force[c]; Fy = (R)e->force[st + c]; Fz = (R)e->force[2 * st + c]; }
    R rhoVar = rho_acc;
    R invRho = (R)1 / rhoVar;
    R ux = (uxp - uxm + (R)0.5 * Fx) * invRho;
    R uy = (uyp - uym + (R)0.5 * Fy) * invRho;
    R uz = (uzp - uzm + (R)0.5 * Fz) * invRho;
    R pixy = pixy_p - pixy_m, pixz = pixz_p - pixz_m, piyz = piyz_p - piyz_m;
    R Omega = (R)1 / (((R)e->nu) * (R)3 + (R)0.5);
    R one = 1, two = 2, three = 3;
    R pixx_part = (two * pixx - piyy - pizz) / three;
    R piyy_part = (two * piyy - pixx - pizz) / three;
    R pizz_part = (two * pizz - pixx - piyy) / three;
    R RU2 = rhoVar * ux * ux, RV2 = rhoVar * uy * uy, RW2 = rhoVar * uz * uz;
    R RUVW2 = (RU2 + RV2 + RW2) / three;
    R pixx_n = rhoVar / three + pixx_part * (one - Omega) + RUVW2 + (two * RU2 * Omega) / three
             - (RV2 * Omega) / three - (RW2 * Omega) / three + Fx * ux;
    R piyy_n = rhoVar / three + piyy_part * (one - Omega) + RUVW2 - (RU2 * Omega) / three
             + (two * RV2 * Omega) / three - (RW2 * Omega) / three + Fy * uy;
    R pizz_n = rhoVar / three + pizz_part * (one - Omega) + RUVW2 - (RU2 * Omega) / three
             - (RV2 * Omega) / three + (two * RW2 * Omega) / three + Fz * uz;
    R pixy_n = pixy - pixy * Omega + ux * uy * rhoVar * Omega + (Fy * ux + Fx * uy) / two;
    R pixz_n = pixz - pixz * Omega + ux * uz * rhoVar * Omega + (Fz * ux + Fx * uz) / two;
    R piyz_n = piyz - piyz * Omega + uy * uz * rhoVar * Omega + (Fz * uy + Fy * uz) / two;
    out[0] = rhoVar;
    out[1] = ux + Fx * invRho / two;
    out[2] = uy + Fy * invRho / two;
    out[3] = uz + Fz * invRho / two;
    out[4] = pixx_n * invRho - FN(cs2_3d);
    out[5] = piyy_n * invRho - FN(cs2_3d);
    out[6] = pizz_n * invRho - FN(cs2_3d);
    out[7] = pixy_n * invRho;
    out[8] = pixz_n * invRho;
    out[9] = piyz_n * invRho;
}

static void FN(d3q27_cell)(const hlo_env* e, const R* src, int x, int y, int z, R out[10])
{
    FN(d3q27_cell_rule)(e, src, x, y, z, NULL, 0, NULL, out);
}

/* --------------------------------------------------------- wall velocity */

/* Velocity of the body material point at h (time step 1): the point's
 * current position minus the position the same body-frame point had one
 * step earlier. 2D follows the material-point difference of
 * envs/lbm/lbm_func.py:354-370; 3D follows lbm_func_3d.py:456-463.
 * Pose table 2D: px, py, cos, sin | last: px, py, cos, sin.
 * Pose table 3D: p[3], R[9] row-major | last: p[3], R[9]. */
static void FN(wall_velocity_2d)(const hlo_env* e, int sid, R hx, R hy, R up[2], R r[2])
{
    if (sid < 0 || sid >= e->nsolid) { up[0] = up[1] = 0; if (r) { r[0] = r[1] = 0; } return; }
    const float* P = e->solids + 8 * sid;
    R px = P[0], py = P[1], cn = P[2], sn = P[3];
    R pxl = P[4], pyl = P[5], cl = P[6], sl = P[7];
    R dx = hx - px, dy = hy - py;
    R qx = cn * dx + sn * dy;
    R qy = cn * dy - sn * dx;
    R cx = cn * qx - sn * qy + px;
    R cy = sn * qx + cn * qy + py;
    R lx = cl * qx - sl * qy + pxl;
    R ly = sl * qx + cl * qy + pyl;
    up[0] = cx - lx;
    up[1] = cy - ly;
    if (r) { r[0] = hx - px; r[1] = hy - py; }
}

static void FN(wall_velocity_3d)(const hlo_env* e, int sid, R hx, R hy, R hz, R up[3], R r[3])
{
    if (sid < 0 || sid >= e->nsolid) {
        up[0] = up[1] = up[2] = 0;
        if (r) { r[0] = r[1] = r[2] = 0; }
        return;
    }
    const float* P = e->solids + 24 * sid;
    const float* M = P + 3;
    const float* L = P + 12;
    const float* ML = P + 15;
    R d0 = hx - (R)P[0], d1 = hy - (R)P[1], d2 = hz - (R)P[2];
    /* q = R^T (h - p) */
    R q0 = (R)M[0] * d0 + (R)M[3] * d1 + (R)M[6] * d2;
    R q1 = (R)M[1] * d0 + (R)M[4] * d1 + (R)M[7] * d2;
    R q2 = (R)M[2] * d0 + (R)M[5] * d1 + (R)M[8] * d2;
    R c0 = (R)M[0] * q0 + (R)M[1] * q1 + (R)M[2] * q2 + (R)P[0];
    R c1 = (R)M[3] * q0 + (R)M[4] * q1 + (R)M[5] * q2 + (R)P[1];
    R c2 = (R)M[6] * q0 + (R)M[7] * q1 + (R)M[8] * q2 + (R)P[2];
    R l0 = (R)ML[0] * q0 + (R)ML[1] * q1 + (R)ML[2] * q2 + (R)L[0];
    R l1 = (R)ML[3] * q0 + (R)ML[4] * q1 + (R)ML[5] * q2 + (R)L[1];
    R l2 = (R)ML[6] * q0 + (R)ML[7] * q1 + (R)ML[8] * q2 + (R)L[2];
    up[0] = c0 - l0; up[1] = c1 - l1; up[2] = c2 - l2;
    if (r) { r[0] = d0; r[1] = d1; r[2] = d2; }
}

/* Momentum exchange of cut link i of a cell with moments m (rho, u, Sxx,
 * Syy, Szz, Sxy, Sxz, Syz), the wall point h on solid sid, added to o
 * (lbm_func_3d.py:478-511: bndForce and bndTorque, dv = h - solid position). */
static void FN(link_force_3d)(const hlo_env* e, int sid, int i, R hx, R hy, R hz, const R m[10], double o[6])
{
    R rho = m[0], ux = m[1], uy = m[2], uz = m[3];
    R Sxx = m[4], Syy = m[5], Szz = m[6], Sxy = m[7], Sxz = m[8], Syz = m[9];
    R dx = (R)D3_CX[i], dy = (R)D3_CY[i], dz = (R)D3_CZ[i];
    R up[3], r[3];
    FN(wall_velocity_3d)(e, sid, hx, hy, hz, up, r);
    R Sxx_s = up[0] * up[0] + Sxx - ux * ux;
    R Sxy_s = up[0] * up[1] + Sxy - ux * uy;
    R Sxz_s = up[0] * up[2] + Sxz - ux * uz;
    R Syy_s = up[1] * up[1] + Syy - uy * uy;
    R Syz_s = up[1] * up[2] + Syz - uy * uz;
    R Szz_s = up[2] * up[2] + Szz - uz * uz;
    R pop = FN(d3q27_recon_s)(rho, up[0], up[1], up[2], Sxx_s, Sxy_s, Sxz_s, Syy_s, Syz_s, Szz_s,
                              dx, dy, dz, FN(d3_w)[i]);
    int inv = D3_INV[i];
    R ex = (R)D3_CX[inv], ey = (R)D3_CY[inv], ez = (R)D3_CZ[inv];
    R cur_f = FN(d3q27_recon_s)(rho, ux, uy, uz, Sxx, Sxy, Sxz, Syy, Syz, Szz, ex, ey, ez, FN(d3_w)[inv]);
    R Fx = cur_f * (ex - up[0]) - pop * (dx - up[0]);
    R Fy = cur_f * (ey - up[1]) - pop * (dy - up[1]);
    R Fz = cur_f * (ez - up[2]) - pop * (dz - up[2]);
    R Tx = r[1] * Fz - r[2] * Fy;
    R Ty = r[2] * Fx - r[0] * Fz;
    R Tz = r[0] * Fy - r[1] * Fx;
    o[0] += (double)Fx; o[1] += (double)Fy; o[2] += (double)Fz;
    o[3] += (double)Tx; o[4] += (double)Ty; o[5] += (double)Tz;
}

/* ------------------------------------------------- capsule cut links (3D) */

/* The crossing of the link from cell X towards its source, P(t) = X - c t
 * with t in [0, 1], nearest to X, over the surfaces of capsules
 * (ax, ay, az, bx, by, bz, r, solid) with 0 <= solid < nsolid. This is the
 * ray query of get_cutcell_multi_3d (lbm_func_3d.py:254-316: origin X,
 * direction -c, the nearest hit with 0 <= t <= 1 over every solid, the first
 * solid winning ties) with each capsule's surface in closed form instead of
 * a triangle mesh. As with a mesh, every surface counts on its own: a ray
 * from inside a capsule hits where it leaves it, and the surfaces of
 * overlapping capsules all count (the reference concatenates a body's geom
 * meshes, mjcf_parser.py:133).
 *
 * Along the line, a capsule is the interval where P(t) lies in the end
 * sphere at a, the end sphere at b or the cylinder between them (the roots of
 * each quadratic, the cylinder's cut to the segment's parameter range); the
 * capsule being convex, the union is the hull [lo, hi]. The crossing nearest
 * to X is lo when lo >= 0 (X outside), otherwise hi (X inside). A capsule
 * whose bounding sphere (centre (a+b)/2, radius |b-a|/2 + r) lies more than
 * 2 > |c| from X is skipped (no crossing can reach it). Returns the solid id
 * (-1: no crossing) and t in *t_hit. */
static int FN(capsule_cut)(const hlo_env* e, const float* caps, int ncap, R x, R y, R z, int cx, int cy, int cz, R* t_hit)
{
    const R dx = (R)-cx, dy = (R)-cy, dz = (R)-cz;
    const R dd = (R)(cx * cx + cy * cy + cz * cz);
    const R big = (R)1e30;
    R best = -1;
    int sid = -1;
    for (int k = 0; k < ncap; ++k) {
        const float* C = caps + 8 * k;
        const int s = (int)C[7];
        if (s < 0 || s >= e->nsolid) continue;
        const R ax = C[0], ay = C[1], az = C[2], bx = C[3], by = C[4], bz = C[5], rad = C[6];
        const R bax = bx - ax, bay = by - ay, baz = bz - az;
        const R baba = bax * bax + bay * bay + baz * baz;
        const R mx = x - (ax + bx) * (R)0.5, my = y - (ay + by) * (R)0.5, mz = z - (az + bz) * (R)0.5;
        const R reach = RSQRT(baba) * (R)0.5 + rad + (R)2;
        if (mx * mx + my * my + mz * mz > reach * reach) continue;
        const R rr = rad * rad;
        R lo = big, hi = -big;
        /* end sphere at a */
        const R ox = x - ax, oy = y - ay, oz = z - az;
        const R rdoa = dx * ox + dy * oy + dz * oz;
        const R oaoa = ox * ox + oy * oy + oz * oz;
        {
            const R disc = rdoa * rdoa - dd * (oaoa - rr);
            if (disc >= (R)0) {
                const R q = RSQRT(disc);
                const R t0 = (-rdoa - q) / dd, t1 = (-rdoa + q) / dd;
                if (t0 < lo) lo = t0;
                if (t1 > hi) hi = t1;
            }
        }
        if (baba > (R)0) {
            /* end sphere at b */
            const R px = x - bx, py = y - by, pz = z - bz;
            const R b = dx * px + dy * py + dz * pz;
            const R disc = b * b - dd * (px * px + py * py + pz * pz - rr);
            if (disc >= (R)0) {
                const R q = RSQRT(disc);
                const R t0 = (-b - q) / dd, t1 = (-b + q) / dd;
                if (t0 < lo) lo = t0;
                if (t1 > hi) hi = t1;
            }
            /* cylinder: |P - a|^2 baba - ((P - a).ba)^2 <= r^2 baba with
             * 0 <= (P - a).ba <= baba; a ray parallel to the axis meets the
             * cylinder only between the two end spheres */
            const R bard = bax * dx + bay * dy + baz * dz;
            const R baoa = bax * ox + bay * oy + baz * oz;
            const R qa = baba * dd - bard * bard;
            if (qa > (R)0) {
                const R qb = baba * rdoa - baoa * bard;
                const R qc = baba * (oaoa - rr) - baoa * baoa;
                const R disc = qb * qb - qa * qc;
                if (disc >= (R)0) {
                    const R q = RSQRT(disc);
                    R c0 = (-qb - q) / qa, c1 = (-qb + q) / qa;
                    int ok = 1;
                    if (bard != (R)0) {
                        R s0 = -baoa / bard, s1 = (baba - baoa) / bard;
                        if (bard < (R)0) { const R sw = s0; s0 = s1; s1 = sw; }
                        if (s0 > c0) c0 = s0;
                        if (s1 < c1) c1 = s1;
                        ok = c0 <= c1;
                    } else {
                        ok = baoa >= (R)0 && baoa <= baba;
                    }
                    if (ok) {
                        if (c0 < lo) lo = c0;
                        if (c1 > hi) hi = c1;
                    }
                }
            }
        }
        if (lo > hi) continue;
        const R t = lo >= (R)0 ? lo : hi;
        if (t >= (R)0 && t <= (R)1 && (best < (R)0 || t < best)) {
            best = t;
            sid = s;
        }
    }
    *t_hit = best;
    return sid;
}

int FN(hlo_capsule_cut)(const hlo_env* e, const float* caps, int32_t ncap, R x, R y, R z, int32_t i, R* t_hit)
{
    *t_hit = -1;
    if (e->lattice != 27 || i <= 0 || i >= 27) return -1;
    return FN(capsule_cut)(e, caps, ncap, x, y, z, D3_CX[i], D3_CY[i], D3_CZ[i], t_hit);
}

/* One D3Q27 step with the capsules' cut links: a link is cut when
 * capsule_cut finds a crossing, and its wall point is the crossing
 * x - c_i t (lbm_func_3d.py:424-511); every interior cell, inside a body or
 * not, takes the update (the reference flags no cell solid in 3D). Forces
 * (nullable, [nsolid][6] FP64) of the state read, summed in cell order
 * (z, y, x) and ascending direction. env.mask is not read. */
void FN(hlo_step_capsules)(const hlo_env* e, const float* caps, int32_t ncap, const R* src, R* dst, double* forces)
{
    const int64_t st = e->stride;
    if (forces)
        for (int k = 0; k < 6 * e->nsolid; ++k) forces[k] = 0.0;
    if (e->lattice != 27) return;
    for (int z = 1; z < e->nz - 1; ++z)
        for (int y = 1; y < e->ny - 1; ++y)
            for (int x = 1; x < e->nx - 1; ++x) {
                R out[10];
                FN(d3q27_cell_rule)(e, src, x, y, z, caps, ncap, forces, out);
                int64_t c = (int64_t)x + (int64_t)e->nx * ((int64_t)y + (int64_t)e->ny * z);
                for (int f = 0; f < 10; ++f) dst[f * st + c] = out[f];
            }
    FN(bc_passes)(e, dst);
}

/* ------------------------------------------------------------ public API */

void FN(hlo_step)(const hlo_env* e, const R* src, R* dst)
{
    const int64_t st = e->stride;
    if (e->lattice == 9) {
        for (int y = 1; y < e->ny - 1; ++y)
            for (int x = 1; x < e->nx - 1; ++x) {
                R out[6];
                FN(d2q9_cell)(e, src, x, y, out);
                int64_t c = (int64_t)x + (int64_t)e->nx * y;
                for (int f = 0; f < 6; ++f) dst[f * st + c] = out[f];
            }
    } else {
        for (int z = 1; z < e->nz - 1; ++z)
            for (int y = 1; y < e->ny - 1; ++y)
                for (int x = 1; x < e->nx - 1; ++x) {
                    R out[10];
                    FN(d3q27_cell)(e, src, x, y, z, out);
                    int64_t c = (int64_t)x + (int64_t)e->nx * ((int64_t)y + (int64_t)e->ny * z);
                    for (int f = 0; f < 10; ++f) dst[f * st + c] = out[f];
                }
    }
    FN(bc_passes)(e, dst);
}

/* Momentum exchange per cut link, accumulated in FP64 in cell order
 * (z, y, x) and ascending direction: lbm_func.py:383-401 (2D, torque scaled
 * by grid_length = 1/nx) and lbm_func_3d.py:491-511 (3D). */
void FN(hlo_forces)(const hlo_env* e, const R* src, double* out)
{
    const int64_t st = e->stride;
    const int nx = e->nx, ny = e->ny, nz = e->nz;
    const int d3 = e->lattice == 27;
    for (int k = 0; k < 6 * e->nsolid; ++k) out[k] = 0.0;
    if (!e->mask) return;
    const int q = d3 ? 27 : 9;
    const R grid_length = (R)1 / (R)nx;
    for (int z = d3 ? 1 : 0; z < (d3 ? nz - 1 : 1); ++z)
        for (int y = 1; y < ny - 1; ++y)
            for (int x = 1; x < nx - 1; ++x) {
                int64_t c = (int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z);
                int occ_c = e->mask[c] != 0;
                for (int i = 0; i < q; ++i) {
                    int cx = d3 ? D3_CX[i] : D2_CX[i];
                    int cy = d3 ? D3_CY[i] : D2_CY[i];
                    int cz = d3 ? D3_CZ[i] : 0;
                    int64_t s = (int64_t)(x - cx) + (int64_t)nx * ((int64_t)(y - cy) + (int64_t)ny * (z - cz));
                    int occ_s = e->mask[s] != 0;
                    if (occ_s == occ_c) continue;
                    int sid = occ_c ? e->mask[c] - 1 : e->mask[s] - 1;
                    if (sid < 0 || sid >= e->nsolid) continue;
                    double* o = out + 6 * sid;
                    if (!d3) {
                        R rho = src[c], ux = src[st + c], uy = src[2 * st + c];
                        R Sxx = src[3 * st + c], Syy = src[4 * st + c], Sxy = src[5 * st + c];
                        R up[2], r[2];
                        FN(wall_velocity_2d)(e, sid, (R)x - (R)0.5 * (R)cx, (R)y - (R)0.5 * (R)cy, up, r);
                        R Sxx_s = up[0] * up[0] + Sxx - ux * ux;
                        R Syy_s = up[1] * up[1] + Syy - uy * uy;
                        R Sxy_s = up[0] * up[1] + Sxy - ux * uy;
                        R pop = FN(d2q9_recon)(rho, up[0], up[1], Sxx_s, Syy_s, Sxy_s, i);
                        if (pop < (R)0) pop = 0;
                        int inv = D2_INV[i];
                        R cur_f = FN(d2q9_recon)(rho, ux, uy, Sxx, Syy, Sxy, inv);
                        R Fx = cur_f * ((R)D2_CX[inv] - up[0]) - pop * ((R)cx - up[0]);
                        R Fy = cur_f * ((R)D2_CY[inv] - up[1]) - pop * ((R)cy - up[1]);
                        R Tz = (r[0] * Fy - r[1] * Fx) * grid_length;
                        o[0] += (double)Fx; o[1] += (double)Fy; o[5] += (double)Tz;
                    } else {
                        const R m[10] = {src[c], src[st + c], src[2 * st + c], src[3 * st + c], src[4 * st + c],
                                         src[5 * st + c], src[6 * st + c], src[7 * st + c], src[8 * st + c],
                                         src[9 * st + c]};
                        FN(link_force_3d)(e, sid, i, (R)x - (R)0.5 * (R)cx, (R)y - (R)0.5 * (R)cy,
                                          (R)z - (R)0.5 * (R)cz, m, o);
                    }
                }
            }
}

/* Uniform-flow initialisation: set_uniform_flow_kernel,
 * envs/lbm/karman/karman_lbm_env.py:15-53, and set_uniform_flow_3d_kernel,
 * envs/lbm3d/karman/karman_lbm_env_3d.py:15-64. With u = 0 this is also
 * InitFlow / InitFlow3D (lbm_func.py:26-63, lbm_func_3d.py:22-83). The
 * diagonal S subtracts the raw pixx (cs2_local = pixx), so Sxx = 0 and
 * Syy = piyy - pixx at rho = 1. */
void FN(hlo_init_uniform)(const hlo_env* e, R ux, R uy, R uz, R* dst)
{
    const int64_t st = e->stride;
    const int64_t n = (int64_t)e->nx * e->ny * e->nz;
    R m[10];
    R rho = 1;
    if (e->lattice == 9) {
        R pop[9];
        R U2 = ux * ux + uy * uy;
        for (int i = 0; i < 9; ++i) {
            R cu = (R)D2_CX[i] * ux + (R)D2_CY[i] * uy;
            pop[i] = rho * FN(d2_w)[i] * ((R)1 + (R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * U2);
        }
        R inv = (R)1 / rho;
        R pixx = pop[1] + pop[2] + pop[5] + pop[6] + pop[7] + pop[8];
        R piyy = pop[3] + pop[4] + pop[5] + pop[6] + pop[7] + pop[8];
        R pixy = pop[5] + pop[7] - pop[6] - pop[8];
        R cs2_local = pixx;
        m[0] = rho; m[1] = ux; m[2] = uy;
        m[3] = pixx * inv - cs2_local; m[4] = piyy * inv - cs2_local; m[5] = pixy * inv;
        for (int f = 0; f < 6; ++f)
            for (int64_t c = 0; c < n; ++c) dst[f * st + c] = m[f];
    } else {
        R p[27];
        R U2 = ux * ux + uy * uy + uz * uz;
        for (int i = 0; i < 27; ++i) {
            R cu = (R)D3_CX[i] * ux + (R)D3_CY[i] * uy + (R)D3_CZ[i] * uz;
            p[i] = FN(d3_w)[i] * rho * ((R)1 + (R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * U2);
        }
        R inv = (R)1 / rho;
        R pixx = p[1] + p[2] + p[7] + p[8] + p[9] + p[10] + p[13] + p[14] + p[15] + p[16] + p[19] + p[20]
               + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
        R pixy = (p[7] + p[8] + p[19] + p[20] + p[21] + p[22]) - (p[13] + p[14] + p[23] + p[24] + p[25] + p[26]);
        R pixz = (p[9] + p[10] + p[19] + p[20] + p[23] + p[24]) - (p[15] + p[16] + p[21] + p[22] + p[25] + p[26]);
        R piyy = p[3] + p[4] + p[7] + p[8] + p[11] + p[12] + p[13] + p[14] + p[17] + p[18] + p[19] + p[20]
               + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
        R piyz = (p[11] + p[12] + p[19] + p[20] + p[25] + p[26]) - (p[17] + p[18] + p[21] + p[22] + p[23] + p[24]);
        R pizz = p[5] + p[6] + p[9] + p[10] + p[11] + p[12] + p[15] + p[16] + p[17] + p[18] + p[19] + p[20]
               + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
        R cs2_local = pixx;
        m[0] = rho; m[1] = ux; m[2] = uy; m[3] = uz;
        m[4] = pixx * inv - cs2_local; m[5] = piyy * inv - cs2_local; m[6] = pizz * inv - cs2_local;
        m[7] = pixy * inv; m[8] = pixz * inv; m[9] = piyz * inv;
        for (int f = 0; f < 10; ++f)
            for (int64_t c = 0; c < n; ++c) dst[f * st + c] = m[f];
    }
}

/* ------------------------------------------------- polygon cut cells (2D) */

static int FN(poly_count)(const hlo_polys* p, int s)
{
    int n = p->counts[s];
    return n < 0 ? 0 : (n > p->maxv ? p->maxv : n);
}

/* get_cutcell, envs/lbm/lbm_func.py:545-607: the crossing nearest to a of
 * the link a = (x - dx, y - dy) -> b = (x, y) with a segment c -> d of the
 * solids whose bounding circle holds b (a solid is skipped when
 * |b - position|^2 > radius2). Returns the solid (-1: no crossing), the
 * position along c -> d and the index of c. */
static int FN(poly_cut)(const hlo_env* e, const hlo_polys* p, R x, R y, R dx, R dy, R* ratio, int* seg)
{
    R cutcell = -1, sr = -1;
    int sid = -1, sg = -1;
    const R ax = x - dx, ay = y - dy, bx = x, by = y;
    for (int s = 0; s < e->nsolid; ++s) {
        const R ox = bx - (R)e->solids[8 * s], oy = by - (R)e->solids[8 * s + 1];
        const R d2 = ox * ox + oy * oy;
        if (d2 > (R)p->radius2[s]) continue;
        const int num = FN(poly_count)(p, s);
        const float* v = p->verts + (int64_t)s * p->maxv * 2;
        for (int n = 0; n < num; ++n) {
            const int m = (n + 1) % num;
            const R cx = v[2 * n], cy = v[2 * n + 1], ddx = v[2 * m], ddy = v[2 * m + 1];
            const R abc = (ax - cx) * (by - cy) - (ay - cy) * (bx - cx);
            const R abd = (ax - ddx) * (by - ddy) - (ay - ddy) * (bx - ddx);
            if (abc * abd >= (R)0) continue;
            const R cda = (cx - ax) * (ddy - ay) - (cy - ay) * (ddx - ax);
            const R cdb = cda + abc - abd;
            if (cda * cdb > (R)0) continue;
            const R t = (R)1 - cda / (abd - abc);
            if (cutcell < (R)0 || t < cutcell) {
                cutcell = t;
                sid = s;
                sr = abc / (abc - abd);
                sg = n;
            }
        }
    }
    if (cutcell >= (R)0 && sid >= 0) {
        *ratio = sr;
        *seg = sg;
        return sid;
    }
    return -1;
}

/* The cut branch of stream_and_collide (lbm_func.py:346-401): the crossing
 * as a material point of segment seg, its velocity from this step's and the
 * previous step's vertices (time step 1), the lever arm from the solid's
 * position. */
static void FN(poly_point)(const hlo_env* e, const hlo_polys* p, int sid, int seg, R ratio, R up[2], R r[2])
{
    const int num = FN(poly_count)(p, sid), m = (seg + 1) % num;
    const float* v = p->verts + (int64_t)sid * p->maxv * 2;
    const float* l = p->verts_last + (int64_t)sid * p->maxv * 2;
    const R pcx = (R)v[2 * seg] + ((R)v[2 * m] - (R)v[2 * seg]) * ratio;
    const R pcy = (R)v[2 * seg + 1] + ((R)v[2 * m + 1] - (R)v[2 * seg + 1]) * ratio;
    const R plx = (R)l[2 * seg] + ((R)l[2 * m] - (R)l[2 * seg]) * ratio;
    const R ply = (R)l[2 * seg + 1] + ((R)l[2 * m + 1] - (R)l[2 * seg + 1]) * ratio;
    up[0] = pcx - plx;
    up[1] = pcy - ply;
    r[0] = pcx - (R)e->solids[8 * sid];
    r[1] = pcy - (R)e->solids[8 * sid + 1];
}

void FN(hlo_step_polygons)(const hlo_env* e, const hlo_polys* p, const R* src, R* dst, double* forces)
{
    const int64_t st = e->stride;
    const int nx = e->nx, ny = e->ny;
    const R grid_length = (R)1 / (R)nx;
    if (forces)
        for (int k = 0; k < 6 * e->nsolid; ++k) forces[k] = 0.0;
    for (int y = 1; y < ny - 1; ++y)
        for (int x = 1; x < nx - 1; ++x) {
            const int64_t c = (int64_t)x + (int64_t)nx * y;
            R pop[9];
            const R rho_c = src[c], ux_c = src[st + c], uy_c = src[2 * st + c];
            const R Sxx_c = src[3 * st + c], Syy_c = src[4 * st + c], Sxy_c = src[5 * st + c];
            for (int i = 0; i < 9; ++i) {
                R ratio = 0;
                int seg = -1;
                const int sid = FN(poly_cut)(e, p, (R)x, (R)y, (R)D2_CX[i], (R)D2_CY[i], &ratio, &seg);
                if (sid >= 0) {
                    R up[2], r[2];
                    FN(poly_point)(e, p, sid, seg, ratio, up, r);
                    R Sxx_s = up[0] * up[0] + Sxx_c - ux_c * ux_c;
                    R Syy_s = up[1] * up[1] + Syy_c - uy_c * uy_c;
                    R Sxy_s = up[0] * up[1] + Sxy_c - ux_c * uy_c;
                    pop[i] = FN(d2q9_recon)(rho_c, up[0], up[1], Sxx_s, Syy_s, Sxy_s, i);
                    if (pop[i] < (R)0) pop[i] = 0;
                    const int inv = D2_INV[i];
                    R cur_f = FN(d2q9_recon)(rho_c, ux_c, uy_c, Sxx_c, Syy_c, Sxy_c, inv);
                    R Fx = cur_f * ((R)D2_CX[inv] - up[0]) - pop[i] * ((R)D2_CX[i] - up[0]);
                    R Fy = cur_f * ((R)D2_CY[inv] - up[1]) - pop[i] * ((R)D2_CY[i] - up[1]);
                    R Tz = (r[0] * Fy - r[1] * Fx) * grid_length;
                    if (forces) {
                        forces[6 * sid] += (double)Fx;
                        forces[6 * sid + 1] += (double)Fy;
                        forces[6 * sid + 5] += (double)Tz;
                    }
                } else {
                    const int64_t s = (int64_t)(x - D2_CX[i]) + (int64_t)nx * (y - D2_CY[i]);
                    pop[i] = FN(d2q9_recon)(src[s], src[st + s], src[2 * st + s], src[3 * st + s],
                                            src[4 * st + s], src[5 * st + s], i);
                    if (pop[i] < (R)0) pop[i] = 0;
                }
            }
            R Fx = 0, Fy = 0;
            if (e->force) { Fx = (R)e->force[c]; Fy = (R)e->force[st + c]; }
            R rhoVar = pop[0] + pop[1] + pop[2] + pop[3] + pop[4] + pop[5] + pop[6] + pop[7] + pop[8];
            R invRho = (R)1 / rhoVar;
            R ux = (pop[1] - pop[2] + pop[5] - pop[6] - pop[7] + pop[8] + (R)0.5 * Fx) * invRho;
            R uy = (pop[3] - pop[4] + pop[5] + pop[6] - pop[7] - pop[8] + (R)0.5 * Fy) * invRho;
            R pixx = pop[1] + pop[2] + pop[5] + pop[6] + pop[7] + pop[8];
            R piyy = pop[3] + pop[4] + pop[5] + pop[6] + pop[7] + pop[8];
            R pixy = pop[5] + pop[7] - pop[6] - pop[8];
            R Omega = (R)1 / (((R)e->nu) * (R)3 + (R)0.5);
            R Ru2 = rhoVar * ux * ux;
            R Rv2 = rhoVar * uy * uy;
            R pixx_part = (pixx - piyy) / (R)2;
            R piyy_part = (piyy - pixx) / (R)2;
            R one = 1;
            pixy = (one - Omega) * pixy + ux * uy * rhoVar * Omega + (one - (R)0.5 * Omega) * (Fy * ux + Fx * uy);
            R pixx_n = rhoVar / (R)3 + pixx_part * (one - Omega) + (one + Omega) / (R)2 * Ru2
                     + (one - Omega) / (R)2 * Rv2 + Fx * ux + (one - Omega) / (R)2 * (Fx * ux - Fy * uy);
            R piyy_n = rhoVar / (R)3 + piyy_part * (one - Omega) + (one + Omega) / (R)2 * Rv2
                     + (one - Omega) / (R)2 * Ru2 + Fy * uy + (one - Omega) / (R)2 * (Fy * uy - Fx * ux);
            dst[c] = rhoVar;
            dst[st + c] = ux + Fx * invRho / (R)2;
            dst[2 * st + c] = uy + Fy * invRho / (R)2;
            dst[3 * st + c] = pixx_n * invRho - FN(cs2_2d);
            dst[4 * st + c] = piyy_n * invRho - FN(cs2_2d);
            dst[5 * st + c] = pixy * invRho;
        }
    FN(bc_passes)(e, dst);
}

R FN(hlo_recon)(int lattice, const R* m, int i)
{
    if (lattice == 9) return FN(d2q9_recon)(m[0], m[1], m[2], m[3], m[4], m[5], i);
    return FN(d3q27_recon_s)(m[0], m[1], m[2], m[3], m[4], m[7], m[8], m[5], m[9], m[6],
                             (R)D3_CX[i], (R)D3_CY[i], (R)D3_CZ[i], FN(d3_w)[i]);
}

#undef FN
#undef CAT
#undef CAT2
