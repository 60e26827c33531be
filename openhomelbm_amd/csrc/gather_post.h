// Post-collision moments of one interior cell in the reference's operation
// order (stream_and_collide, envs/lbm/lbm_func.py:322-467, and
// stream_and_collide_3d, envs/lbm3d/lbm_func_3d.py:364-604), with the mask
// wall rule on cut links. Shared by the gather step kernel and the near-body
// fix-up kernel. Contraction is off so the FP32 build matches a literal FP32
// evaluation.
#pragma once
#pragma clang fp contract(off)

#include "lbm_device.h"

namespace hl {

// Population i pulled into interior cell (x, y) (lbm_func.py:335-412) once
// the link is classified: on a cut link (solid sid) the wall-rule population
// from the cell's own moments mc, else the reconstruction from the source
// cell's moments ms; clamped at 0 on both branches as in the reference.
template <typename T>
__device__ __forceinline__ T d2q9_link_pop(const EnvArgs& a, int x, int y, int i, const T* mc, int cut, int sid,
                                           const T* ms, const float* poses = nullptr)
{
    T pop;
    if (cut) {
        T up[2], r[2];
        wall_velocity_2d<T>(poses ? poses : a.solids, a.nsolid, sid, T(x) - T(0.5) * T(D2CX[i]),
                            T(y) - T(0.5) * T(D2CY[i]), up, r);
        T Sxx_s = up[0] * up[0] + mc[3] - mc[1] * mc[1];
        T Syy_s = up[1] * up[1] + mc[4] - mc[2] * mc[2];
        T Sxy_s = up[0] * up[1] + mc[5] - mc[1] * mc[2];
        pop = d2q9_recon<T>(mc[0], up[0], up[1], Sxx_s, Syy_s, Sxy_s, i);
    } else {
        pop = d2q9_recon<T>(ms[0], ms[1], ms[2], ms[3], ms[4], ms[5], i);
    }
    if (pop < T(0)) pop = T(0);
    return pop;
}

// The same with the source moments read from global memory.
template <typename T, int DT>
__device__ __forceinline__ T d2q9_pull(const EnvArgs& a, int x, int y, int i, const T* mc, bool masked, int mval_c)
{
    using S = Storage<DT>;
    using ST = typename S::type;
    const ST* src = static_cast<const ST*>(a.src);
    const int64_t st = a.stride;
    const int64_t s = (int64_t)(x - D2CX[i]) + (int64_t)a.nx * (y - D2CY[i]);
    int cut = 0, sid = -1;
    if (masked) {
        const int mval_s = a.mask[s];
        if ((mval_s != 0) != (mval_c != 0)) { cut = 1; sid = mval_c ? mval_c - 1 : mval_s - 1; }
    }
    T m[6];
    if (!cut) {
#pragma unroll
        for (int f = 0; f < 6; ++f) m[f] = S::template load<T>(src, f * st + s, f, a.inv_scale[f], a.offset[f]);
    }
    return d2q9_link_pop<T>(a, x, y, i, mc, cut, sid, m);
}

// Moments and regularized collision of the 9 pulled populations
// (lbm_func.py:414-467), same operation order.
template <typename T>
__device__ __forceinline__ void d2q9_collide(const EnvArgs& a, const T* pop, T Fx, T Fy, T out[6])
{
    T rhoVar = pop[0] + pop[1] + pop[2] + pop[3] + pop[4] + pop[5] + pop[6] + pop[7] + pop[8];
    T invRho = T(1) / rhoVar;
    T ux = (pop[1] - pop[2] + pop[5] - pop[6] - pop[7] + pop[8] + T(0.5) * Fx) * invRho;
    T uy = (pop[3] - pop[4] + pop[5] + pop[6] - pop[7] - pop[8] + T(0.5) * Fy) * invRho;
    T pixx = pop[1] + pop[2] + pop[5] + pop[6] + pop[7] + pop[8];
    T piyy = pop[3] + pop[4] + pop[5] + pop[6] + pop[7] + pop[8];
    T pixy = pop[5] + pop[7] - pop[6] - pop[8];
    T Omega = T(1) / (T(a.nu) * T(3) + T(0.5));
    T Ru2 = rhoVar * ux * ux;
    T Rv2 = rhoVar * uy * uy;
    T pixx_part = (pixx - piyy) / T(2);
    T piyy_part = (piyy - pixx) / T(2);
    const T one = 1;
    pixy = (one - Omega) * pixy + ux * uy * rhoVar * Omega + (one - T(0.5) * Omega) * (Fy * ux + Fx * uy);
    T pixx_n = rhoVar / T(3) + pixx_part * (one - Omega) + (one + Omega) / T(2) * Ru2
             + (one - Omega) / T(2) * Rv2 + Fx * ux + (one - Omega) / T(2) * (Fx * ux - Fy * uy);
    T piyy_n = rhoVar / T(3) + piyy_part * (one - Omega) + (one + Omega) / T(2) * Rv2
             + (one - Omega) / T(2) * Ru2 + Fy * uy + (one - Omega) / T(2) * (Fy * uy - Fx * ux);
    out[0] = rhoVar;
    out[1] = ux + Fx * invRho / T(2);
    out[2] = uy + Fy * invRho / T(2);
    out[3] = pixx_n * invRho - cs2_2d<T>();
    out[4] = piyy_n * invRho - cs2_2d<T>();
    out[5] = pixy * invRho;
}

// The nf moments of the cell at field offset so (cell_at) of the source state.
template <typename T, int DT>
__device__ __forceinline__ void load_cell(const EnvArgs& a, int64_t so, int nf, T* mc)
{
    using S = Storage<DT>;
    const typename S::type* src = static_cast<const typename S::type*>(a.src);
    for (int f = 0; f < nf; ++f) mc[f] = S::template load<T>(src, f * a.stride + so, f, a.inv_scale[f], a.offset[f]);
}

template <typename T, int DT>
__device__ __forceinline__ void d2q9_post(const EnvArgs& a, int x, int y, T out[6])
{
    const int64_t c = (int64_t)x + (int64_t)a.nx * y;
    T mc[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) mc[f] = Storage<DT>::template load<T>(static_cast<const typename Storage<DT>::type*>(a.src), f * a.stride + c, f, a.inv_scale[f], a.offset[f]);
    const bool masked = a.mask != nullptr && in_box_halo(a, x, y, 0);
    const int mval_c = masked ? a.mask[c] : 0;
    T pop[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) pop[i] = d2q9_pull<T, DT>(a, x, y, i, mc, masked, mval_c);
    T Fx = 0, Fy = 0;
    if (a.force) { Fx = T(a.force[c]); Fy = T(a.force[a.fs + c]); }
    d2q9_collide<T>(a, pop, Fx, Fy, out);
}

// The cut state of one link: the solid (-1: not cut) and the wall point.
template <typename T>
struct WallPoint {
    int sid;
    T h[3];
};

// Population i pulled into interior cell (x, y, z) (lbm_func_3d.py:414-526):
// the reconstruction from the source cell's moments, clamped at 0, or on a
// cut link the unclamped wall-rule population from the cell's moments mc.
// A link is cut by the capsules' crossings when a.caps is set (the mask is
// then not read), else by the mask rule when masked; wall (optional)
// receives the link's cut state.
template <typename T, int DT>
__device__ __forceinline__ T d3q27_pull(const EnvArgs& a, int x, int y, int z, int i, const T* mc, bool masked, int mval_c,
                                        const float* poses = nullptr, WallPoint<T>* wall = nullptr)
{
    using S = Storage<DT>;
    using ST = typename S::type;
    const ST* src = static_cast<const ST*>(a.src);
    const int64_t st = a.stride;
    const T dx = T(D3CX[i]), dy = T(D3CY[i]), dz = T(D3CZ[i]), wi = d3w<T>(i);
    const int64_t s = (int64_t)(x - D3CX[i]) + (int64_t)a.nx * ((int64_t)(y - D3CY[i]) + (int64_t)a.ny * (z - D3CZ[i]));
    const int64_t so = cell_at(a, x - D3CX[i], y - D3CY[i], z - D3CZ[i]);
    // the source moments are loaded in the same round trip as the mask test
    // (the source cell is inside the lattice either way)
    T m[10];
#pragma unroll
    for (int f = 0; f < 10; ++f) m[f] = S::template load<T>(src, f * st + so, f, a.inv_scale[f], a.offset[f]);
    int cut = 0, sid = -1;
    // the wall point: the link midpoint under the mask rule, the crossing
    // x - c_i t (lbm_func_3d.py:446) with capsule bodies
    T hx = T(x) - T(0.5) * dx, hy = T(y) - T(0.5) * dy, hz = T(z) - T(0.5) * dz;
    if (a.caps) {
        T t = T(0);
        if (i > 0) sid = capsule_cut<T>(a.caps, a.ncap, a.nsolid, T(x), T(y), T(z), D3CX[i], D3CY[i], D3CZ[i], t);
        if (sid >= 0) {
            cut = 1;
            hx = T(x) - dx * t;
            hy = T(y) - dy * t;
            hz = T(z) - dz * t;
        }
    } else if (masked) {
        const int mval_s = a.mask[s];
        if ((mval_s != 0) != (mval_c != 0)) { cut = 1; sid = mval_c ? mval_c - 1 : mval_s - 1; }
    }
    if (wall) {
        wall->sid = cut ? sid : -1;
        wall->h[0] = hx; wall->h[1] = hy; wall->h[2] = hz;
    }
    T pop_i;
    if (cut) {
        T up[3], r[3];
        wall_velocity_3d<T>(poses ? poses : a.solids, a.nsolid, sid, hx, hy, hz, up, r);
        T Sxx_s = up[0] * up[0] + mc[4] - mc[1] * mc[1];
        T Sxy_s = up[0] * up[1] + mc[7] - mc[1] * mc[2];
        T Sxz_s = up[0] * up[2] + mc[8] - mc[1] * mc[3];
        T Syy_s = up[1] * up[1] + mc[5] - mc[2] * mc[2];
        T Syz_s = up[1] * up[2] + mc[9] - mc[2] * mc[3];
        T Szz_s = up[2] * up[2] + mc[6] - mc[3] * mc[3];
        pop_i = d3q27_recon<T>(mc[0], up[0], up[1], up[2], Sxx_s, Sxy_s, Sxz_s, Syy_s, Syz_s, Szz_s, dx, dy, dz, wi);
    } else {
        pop_i = d3q27_recon<T>(m[0], m[1], m[2], m[3], m[4], m[7], m[8], m[5], m[9], m[6], dx, dy, dz, wi);
        if (pop_i < T(0)) pop_i = T(0);
    }
    return pop_i;
}

// Sign-group moment sums of lbm_func_3d.py:528-562: sixteen accumulators,
// each a sum in ascending direction order over the directions of its group.
// k: 0 rho; 1/2 +x/-x; 3/4 +y/-y; 5/6 +z/-z; 7/8/9 pixx/piyy/pizz (c_a != 0);
// 10/11 c_x c_y > 0 / < 0; 12/13 c_x c_z; 14/15 c_y c_z. Each group is an
// independent sequential sum, so lanes may compute different k in parallel.
__device__ __forceinline__ bool d3q27_in_group(int k, int i)
{
    const int cx = D3CX[i], cy = D3CY[i], cz = D3CZ[i];
    switch (k) {
    case 0: return true;
    case 1: return cx > 0;
    case 2: return cx < 0;
    case 3: return cy > 0;
    case 4: return cy < 0;
    case 5: return cz > 0;
    case 6: return cz < 0;
    case 7: return cx != 0;
    case 8: return cy != 0;
    case 9: return cz != 0;
    case 10: return cx * cy > 0;
    case 11: return cx * cy < 0;
    case 12: return cx * cz > 0;
    case 13: return cx * cz < 0;
    case 14: return cy * cz > 0;
    default: return cy * cz < 0;
    }
}

// Group membership as 27-bit masks (bit i set when direction i belongs).
__device__ __forceinline__ uint32_t d3q27_group_bits(int k)
{
    uint32_t bits = 0;
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
        uint32_t b = 0;
#pragma unroll
        for (int i = 0; i < 27; ++i)
            if (d3q27_in_group(kk, i)) b |= 1u << i;
        if (kk == k) bits = b;
    }
    return bits;
}

// Group k's sum over pop (27 entries, e.g. in LDS): all 27 read at once, the
// members added in ascending order and +0 added for the others. From +0 a
// sum of IEEE additions never reaches -0, so adding +0 leaves it unchanged:
// the result is bitwise that of adding the members alone.
template <typename T> __device__ __forceinline__ T d3q27_group_sum(const T* pop, int k)
{
    const uint32_t bits = d3q27_group_bits(k);
    T p[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) p[i] = pop[i];
    T acc = 0;
#pragma unroll
    for (int i = 0; i < 27; ++i) acc = acc + (((bits >> i) & 1u) ? p[i] : T(0));
    return acc;
}

// The 3D regularized collision (lbm_func_3d.py:568-604) from the sixteen
// group sums, same operation order.
template <typename T>
__device__ __forceinline__ void d3q27_collide_sums(const EnvArgs& a, const T* g, T Fx, T Fy, T Fz, T out[10])
{
    const T rho_acc = g[0], uxp = g[1], uxm = g[2], uyp = g[3], uym = g[4], uzp = g[5], uzm = g[6];
    const T pixx = g[7], piyy = g[8], pizz = g[9];
    T rhoVar = rho_acc;
    T invRho = T(1) / rhoVar;
    T ux = (uxp - uxm + T(0.5) * Fx) * invRho;
    T uy = (uyp - uym + T(0.5) * Fy) * invRho;
    T uz = (uzp - uzm + T(0.5) * Fz) * invRho;
    T pixy = g[10] - g[11], pixz = g[12] - g[13], piyz = g[14] - g[15];
    T Omega = T(1) / (T(a.nu) * T(3) + T(0.5));
    const T one = 1, two = 2, three = 3;
    T pixx_part = (two * pixx - piyy - pizz) / three;
    T piyy_part = (two * piyy - pixx - pizz) / three;
    T pizz_part = (two * pizz - pixx - piyy) / three;
    T RU2 = rhoVar * ux * ux, RV2 = rhoVar * uy * uy, RW2 = rhoVar * uz * uz;
    T RUVW2 = (RU2 + RV2 + RW2) / three;
    T pixx_n = rhoVar / three + pixx_part * (one - Omega) + RUVW2 + (two * RU2 * Omega) / three
             - (RV2 * Omega) / three - (RW2 * Omega) / three + Fx * ux;
    T piyy_n = rhoVar / three + piyy_part * (one - Omega) + RUVW2 - (RU2 * Omega) / three
             + (two * RV2 * Omega) / three - (RW2 * Omega) / three + Fy * uy;
    T pizz_n = rhoVar / three + pizz_part * (one - Omega) + RUVW2 - (RU2 * Omega) / three
             - (RV2 * Omega) / three + (two * RW2 * Omega) / three + Fz * uz;
    T pixy_n = pixy - pixy * Omega + ux * uy * rhoVar * Omega + (Fy * ux + Fx * uy) / two;
    T pixz_n = pixz - pixz * Omega + ux * uz * rhoVar * Omega + (Fz * ux + Fx * uz) / two;
    T piyz_n = piyz - piyz * Omega + uy * uz * rhoVar * Omega + (Fz * uy + Fy * uz) / two;
    out[0] = rhoVar;
    out[1] = ux + Fx * invRho / two;
    out[2] = uy + Fy * invRho / two;
    out[3] = uz + Fz * invRho / two;
    out[4] = pixx_n * invRho - cs2_3d<T>();
    out[5] = piyy_n * invRho - cs2_3d<T>();
    out[6] = pizz_n * invRho - cs2_3d<T>();
    out[7] = pixy_n * invRho;
    out[8] = pixz_n * invRho;
    out[9] = piyz_n * invRho;
}

template <typename T>
__device__ __forceinline__ void d3q27_collide(const EnvArgs& a, const T* pop, T Fx, T Fy, T Fz, T out[10])
{
    T g[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        T acc = 0;
#pragma unroll
        for (int i = 0; i < 27; ++i)
            if (d3q27_in_group(k, i)) acc = acc + pop[i];
        g[k] = acc;
    }
    d3q27_collide_sums<T>(a, g, Fx, Fy, Fz, out);
}

template <typename T, int DT>
__device__ __forceinline__ void d3q27_post(const EnvArgs& a, int x, int y, int z, T out[10])
{
    const int64_t c = (int64_t)x + (int64_t)a.nx * ((int64_t)y + (int64_t)a.ny * z);
    T mc[10];
    load_cell<T, DT>(a, cell_at(a, x, y, z), 10, mc);
    const bool masked = a.mask != nullptr && in_box_halo(a, x, y, z);
    const int mval_c = masked ? a.mask[c] : 0;
    T pop[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) pop[i] = d3q27_pull<T, DT>(a, x, y, z, i, mc, masked, mval_c);
    T Fx = 0, Fy = 0, Fz = 0;
    if (a.force) { Fx = T(a.force[c]); Fy = T(a.force[a.fs + c]); Fz = T(a.force[2 * a.fs + c]); }
    d3q27_collide<T>(a, pop, Fx, Fy, Fz, out);
}

// Momentum exchange of cut link i of a D3Q27 cell with moments mc, the wall
// point h on solid sid (lbm_func_3d.py:478-511: bndForce, and bndTorque with
// dv = h - solid position).
template <typename T>
__device__ __forceinline__ void link_force_at(const EnvArgs& a, const T* mc, int i, int sid, T hx, T hy, T hz, T F[6],
                                              const float* poses)
{
    const T dx = T(D3CX[i]), dy = T(D3CY[i]), dz = T(D3CZ[i]);
    T up[3], r[3];
    wall_velocity_3d<T>(poses ? poses : a.solids, a.nsolid, sid, hx, hy, hz, up, r);
    T Sxx_s = up[0] * up[0] + mc[4] - mc[1] * mc[1];
    T Sxy_s = up[0] * up[1] + mc[7] - mc[1] * mc[2];
    T Sxz_s = up[0] * up[2] + mc[8] - mc[1] * mc[3];
    T Syy_s = up[1] * up[1] + mc[5] - mc[2] * mc[2];
    T Syz_s = up[1] * up[2] + mc[9] - mc[2] * mc[3];
    T Szz_s = up[2] * up[2] + mc[6] - mc[3] * mc[3];
    T pop = d3q27_recon<T>(mc[0], up[0], up[1], up[2], Sxx_s, Sxy_s, Sxz_s, Syy_s, Syz_s, Szz_s, dx, dy, dz, d3w<T>(i));
    const int inv = D3INV[i];
    const T ex = T(D3CX[inv]), ey = T(D3CY[inv]), ez = T(D3CZ[inv]);
    T cur_f = d3q27_recon<T>(mc[0], mc[1], mc[2], mc[3], mc[4], mc[7], mc[8], mc[5], mc[9], mc[6], ex, ey, ez,
                             d3w<T>(inv));
    T Fx = cur_f * (ex - up[0]) - pop * (dx - up[0]);
    T Fy = cur_f * (ey - up[1]) - pop * (dy - up[1]);
    T Fz = cur_f * (ez - up[2]) - pop * (dz - up[2]);
    F[0] = Fx; F[1] = Fy; F[2] = Fz;
    F[3] = r[1] * Fz - r[2] * Fy;
    F[4] = r[2] * Fx - r[0] * Fz;
    F[5] = r[0] * Fy - r[1] * Fx;
}

// Momentum exchange of one cut link (lbm_func.py:383-401 in 2D with the
// torque scaled by grid_length = 1/nx; lbm_func_3d.py:491-511 in 3D).
template <typename T, int DT, int Q>
__device__ __forceinline__ void link_force(const EnvArgs& a, const T* mc, int x, int y, int z, int i, int sid, T F[6],
                                           const float* poses = nullptr)
{
    if (Q == 9) {
        T up[2], r[2];
        wall_velocity_2d<T>(poses ? poses : a.solids, a.nsolid, sid, T(x) - T(0.5) * T(D2CX[i]),
                            T(y) - T(0.5) * T(D2CY[i]), up, r);
        T Sxx_s = up[0] * up[0] + mc[3] - mc[1] * mc[1];
        T Syy_s = up[1] * up[1] + mc[4] - mc[2] * mc[2];
        T Sxy_s = up[0] * up[1] + mc[5] - mc[1] * mc[2];
        T pop = d2q9_recon<T>(mc[0], up[0], up[1], Sxx_s, Syy_s, Sxy_s, i);
        if (pop < T(0)) pop = T(0);
        const int inv = D2INV[i];
        T cur_f = d2q9_recon<T>(mc[0], mc[1], mc[2], mc[3], mc[4], mc[5], inv);
        T Fx = cur_f * (T(D2CX[inv]) - up[0]) - pop * (T(D2CX[i]) - up[0]);
        T Fy = cur_f * (T(D2CY[inv]) - up[1]) - pop * (T(D2CY[i]) - up[1]);
        T grid_length = T(1) / T(a.nx);
        F[0] = Fx; F[1] = Fy; F[2] = 0; F[3] = 0; F[4] = 0;
        F[5] = (r[0] * Fy - r[1] * Fx) * grid_length;
    } else {
        const T dx = T(D3CX[i]), dy = T(D3CY[i]), dz = T(D3CZ[i]);
        link_force_at<T>(a, mc, i, sid, T(x) - T(0.5) * dx, T(y) - T(0.5) * dy, T(z) - T(0.5) * dz, F, poses);
    }
}

// Cut link i of D2Q9 interior cell (x, y) with moments mc on solid sid: the
// wall-rule population (d2q9_link_pop's cut branch, clamped) and, FORCE, its
// momentum exchange Fx, Fy, Tz (link_force), the two sharing one wall
// velocity. The expressions are theirs, so are the results.
template <bool FORCE>
__device__ __forceinline__ float d2q9_cut_link(const EnvArgs& a, const float* mc, int x, int y, int i, int sid,
                                               const float* poses, float F[3])
{
    float up[2], r[2];
    wall_velocity_2d<float>(poses, a.nsolid, sid, float(x) - 0.5f * float(D2CX[i]), float(y) - 0.5f * float(D2CY[i]),
                            up, r);
    const float Sxx_s = up[0] * up[0] + mc[3] - mc[1] * mc[1];
    const float Syy_s = up[1] * up[1] + mc[4] - mc[2] * mc[2];
    const float Sxy_s = up[0] * up[1] + mc[5] - mc[1] * mc[2];
    float pop = d2q9_recon<float>(mc[0], up[0], up[1], Sxx_s, Syy_s, Sxy_s, i);
    if (pop < 0.0f) pop = 0.0f;
    if (FORCE) {
        const int inv = D2INV[i];
        const float cur_f = d2q9_recon<float>(mc[0], mc[1], mc[2], mc[3], mc[4], mc[5], inv);
        const float Fx = cur_f * (float(D2CX[inv]) - up[0]) - pop * (float(D2CX[i]) - up[0]);
        const float Fy = cur_f * (float(D2CY[inv]) - up[1]) - pop * (float(D2CY[i]) - up[1]);
        const float grid_length = 1.0f / float(a.nx);
        F[0] = Fx;
        F[1] = Fy;
        F[2] = (r[0] * Fy - r[1] * Fx) * grid_length;
    }
    return pop;
}

// Cell (x, y) of a D2Q9 gather step (x, y inside the lattice): a face cell
// takes the update of its interior projection, then the face chain.
template <typename T, int DT>
__device__ __forceinline__ void gather_d2q9_at(const EnvArgs& a, int x, int y)
{
    const int px = min(max(x, 1), a.nx - 2), py = min(max(y, 1), a.ny - 2);
    T m[6];
    d2q9_post<T, DT>(a, px, py, m);
    if (px != x || py != y) bc_chain<T, 2>(m, x, y, 0, a.nx, a.ny, 1, a.bc_type, a.bc_value);
    using ST = typename Storage<DT>::type;
    ST* dst = static_cast<ST*>(a.dst);
    const int64_t c = (int64_t)x + (int64_t)a.nx * y;
#pragma unroll
    for (int f = 0; f < 6; ++f)
        Storage<DT>::template store<T>(dst, f * a.stride + c, m[f], f, a.scale[f], a.offset[f], HL_ROUNDING_KEY(a, f, x, y, 0));
}

// Cell (region x0 + bx, y0 + by) of a D2Q9 gather step.
template <typename T, int DT>
__device__ __forceinline__ void gather_d2q9_cell(const EnvArgs& a, int bx, int by)
{
    const int x = a.region[0] + bx, y = a.region[2] + by;
    if (x >= a.region[1] || y >= a.region[3]) return;
    gather_d2q9_at<T, DT>(a, x, y);
}

}  // namespace hl
