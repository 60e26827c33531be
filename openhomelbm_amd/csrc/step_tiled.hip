// LDS-staged source-population step kernel for D3Q27 (HL_KERNEL_TILED).
//
// The reference pulls, for every cell, the full moment set of each of its 27
// neighbours and rebuilds one population from each (27 x 10 loads and 27
// reconstructions per cell, stream_and_collide_3d, lbm_func_3d.py:414-526).
// A population f_i(x) depends only on the moments at the source x - c_i, so
// here every cell rebuilds its own 27 outgoing populations once, with the
// shared subexpressions of the 13 direction pairs, and hands them to its
// neighbours through LDS. Arithmetically this is the same pull; rounding
// differs from the reference's order at the 1e-7 level.
//
// Work decomposition (one block = a TX x TY column tile and a chunk of z):
//   * 27 population slots per column in LDS (one plane's worth, twice);
//   * the block marches z upward. At plane z it rebuilds the sources of
//     plane z+1; their cz = -1 ("down") populations go to LDS at once, their
//     cz = 0 ("mid") ones are held in registers one step and their cz = +1
//     ("up") ones two steps, so the slots always hold down(z+1), mid(z),
//     up(z-1): exactly what the cells of plane z pull;
//   * source columns cover the tile plus a one-cell halo, one per thread;
//     the first TX*TY threads also own the destination cells;
//   * moments of the next source plane are prefetched into registers;
//   * face cells gather at their projection onto the interior and apply the
//     face conditions in the order x, y, z (plane 0 and nz-1 are written by
//     the chunks that own planes 1 and nz-2);
//   * cells with cut links (next to a body mask) are not written by the
//     tiled pass: the first workgroups of the same launch update them with
//     the mask wall rule in the reference's operation order and reduce the
//     body forces before their own tiled pass (fixup_block), so the tiled
//     pass carries no wall-rule code and stays small in registers;
//   * global memory goes through raw buffer loads and stores whose
//     out-of-range offset replaces every per-lane branch (halo columns
//     outside the lattice, lanes without a destination cell, cells owned by
//     the near-body pass), so every wave issues the same memory operations
//     on every path and the vector-memory counter (in order, loads and
//     stores alike) waits for the prefetched plane only, never for the
//     previous plane's stores;
//   * LDS is double-buffered and the loop is software-pipelined: while the
//     cells of plane z pull from one buffer, the sources of plane z+2 fill
//     the other, so each plane costs one block barrier and the two halves
//     of the work can overlap within a wavefront.
// Contraction is off: every fused multiply-add below is written out, so the
// arithmetic of a cell is the same in every tile-shape instantiation.
#pragma clang fp contract(off)

#include <cstring>

#include "gather_post.h"
#include "kernels.h"
#include "lbm_device.h"

namespace hl {

namespace {

constexpr int kUp[9] = {5, 9, 11, 16, 18, 19, 22, 23, 25};     // cz = +1
constexpr int kMid[9] = {0, 1, 2, 3, 4, 7, 8, 13, 14};        // cz = 0
constexpr int kDown[9] = {6, 10, 12, 15, 17, 20, 21, 24, 26}; // cz = -1

// All 27 populations of one cell (lbm_func_3d.py:224-251 for every i),
// grouped by direction pairs: f_{+c} = w rho (E + O), f_{-c} = w rho (E - O)
// with even part E = A + 4.5 (c.u)^2 + 1.5 sum_a c_a^2 S_aa
// + 3 sum_{a<b} c_a c_b S_ab, A = 1 - 1.5 u.u - 0.5 tr S, and odd part
// O = 3 c.u, evaluated as f_{+/-c} = wrho E +/- (3 wrho) c.u with fused
// multiply-adds. Populations are clamped at 0 as on the reference's plain
// streaming branch (lbm_func_3d.py:525-526). The clamp cannot change a
// population when A - 1/2 exceeds the largest possible stress term
// (4.5 p^2 -/+ 3 p >= -1/2 for every projection p = c.u, and |c_a| <= 1), so
// a wavefront whose cells all satisfy that bound skips it (same results).
__device__ __forceinline__ void recon27(const float* m, float* f)
{
    const float rho = m[0], ux = m[1], uy = m[2], uz = m[3];
    const float sxx = m[4], syy = m[5], szz = m[6], sxy = m[7], sxz = m[8], syz = m[9];
    const float trS = sxx + syy + szz;
    const float usq = fmaf(ux, ux, fmaf(uy, uy, uz * uz));
    const float A = fmaf(-1.5f, usq, fmaf(-0.5f, trS, 1.0f));
    const float r0 = (float)(8.0 / 27.0) * rho, r1 = (float)(2.0 / 27.0) * rho;
    const float r2 = (float)(1.0 / 54.0) * rho, r3 = (float)(1.0 / 216.0) * rho;
    const float o1 = 3.0f * r1, o2 = 3.0f * r2, o3 = 3.0f * r3;
    // pair (p, q) = (+c, -c) with projection s = c.u and even part E
    auto pair = [&](int p, int q, float r, float o, float E, float s) {
        const float e = r * E;
        f[p] = fmaf(o, s, e);
        f[q] = fmaf(-o, s, e);
    };
    f[0] = r0 * A;
    pair(1, 2, r1, o1, fmaf(4.5f * ux, ux, fmaf(1.5f, sxx, A)), ux);
    pair(3, 4, r1, o1, fmaf(4.5f * uy, uy, fmaf(1.5f, syy, A)), uy);
    pair(5, 6, r1, o1, fmaf(4.5f * uz, uz, fmaf(1.5f, szz, A)), uz);
    const float Pxy = fmaf(1.5f, sxx + syy, A), Pxz = fmaf(1.5f, sxx + szz, A), Pyz = fmaf(1.5f, syy + szz, A);
    float t = ux + uy;
    pair(7, 8, r2, o2, fmaf(4.5f * t, t, fmaf(3.0f, sxy, Pxy)), t);
    t = ux - uy;
    pair(13, 14, r2, o2, fmaf(4.5f * t, t, fmaf(-3.0f, sxy, Pxy)), t);
    t = ux + uz;
    pair(9, 10, r2, o2, fmaf(4.5f * t, t, fmaf(3.0f, sxz, Pxz)), t);
    t = ux - uz;
    pair(15, 16, r2, o2, fmaf(4.5f * t, t, fmaf(-3.0f, sxz, Pxz)), t);
    t = uy + uz;
    pair(11, 12, r2, o2, fmaf(4.5f * t, t, fmaf(3.0f, syz, Pyz)), t);
    t = uy - uz;
    pair(17, 18, r2, o2, fmaf(4.5f * t, t, fmaf(-3.0f, syz, Pyz)), t);
    const float R = fmaf(1.5f, trS, A);
    t = ux + uy + uz;
    pair(19, 20, r3, o3, fmaf(4.5f * t, t, fmaf(3.0f, sxy + sxz + syz, R)), t);
    t = ux + uy - uz;
    pair(21, 22, r3, o3, fmaf(4.5f * t, t, fmaf(3.0f, sxy - sxz - syz, R)), t);
    t = ux - uy + uz;
    pair(23, 24, r3, o3, fmaf(4.5f * t, t, fmaf(3.0f, sxz - sxy - syz, R)), t);
    t = uy + uz - ux;
    pair(25, 26, r3, o3, fmaf(4.5f * t, t, fmaf(3.0f, syz - sxy - sxz, R)), t);
    const float sd = (fabsf(sxx) + fabsf(syy)) + fabsf(szz), so = (fabsf(sxy) + fabsf(sxz)) + fabsf(syz);
    const float bound = fmaf(-3.0f, so, fmaf(-1.5f, sd, A - 0.5f));
    const bool positive = bound > 1e-3f && rho > 0.0f;
    if (__builtin_amdgcn_ballot_w64(!positive)) {
#pragma unroll
        for (int i = 0; i < 27; ++i) f[i] = fmaxf(f[i], 0.0f);
    }
}

// Direction index of velocity (cx, cy, cz) in the reference's order.
__host__ __device__ constexpr int dir_index(int cx, int cy, int cz)
{
    for (int i = 0; i < 27; ++i)
        if (D3CX[i] == cx && D3CY[i] == cy && D3CZ[i] == cz) return i;
    return -1;
}

// Zeroth, first and second moments of the 27 pulled populations g: rho,
// j = sum c g, Pi = sum c c g (the reference's sign-group sums,
// lbm_func_3d.py:528-562), reduced hierarchically: first along z for each
// (cx, cy) column (sum, z-difference, z-pair sum), then along y for each
// cx, then along x. 66 additions instead of one accumulator update per
// population and component. out: rho, jx, jy, jz, Pxx, Pyy, Pzz, Pxy, Pxz, Pyz.
__device__ __forceinline__ void moments27(const float* g, float* out)
{
    float T[3][3], Dz[3][3], Qz[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float m = g[dir_index(a - 1, b - 1, -1)], o = g[dir_index(a - 1, b - 1, 0)];
            const float p = g[dir_index(a - 1, b - 1, 1)];
            Qz[a][b] = m + p;
            T[a][b] = Qz[a][b] + o;
            Dz[a][b] = p - m;
        }
    float Ty[3], QyT[3], DyT[3], DzS[3], DzDy[3], QzS[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        QyT[a] = T[a][0] + T[a][2];
        Ty[a] = QyT[a] + T[a][1];
        DyT[a] = T[a][2] - T[a][0];
        DzS[a] = (Dz[a][0] + Dz[a][2]) + Dz[a][1];
        DzDy[a] = Dz[a][2] - Dz[a][0];
        QzS[a] = (Qz[a][0] + Qz[a][2]) + Qz[a][1];
    }
    const float pxx = Ty[0] + Ty[2];
    out[0] = pxx + Ty[1];
    out[1] = Ty[2] - Ty[0];
    out[2] = (DyT[0] + DyT[2]) + DyT[1];
    out[3] = (DzS[0] + DzS[2]) + DzS[1];
    out[4] = pxx;
    out[5] = (QyT[0] + QyT[2]) + QyT[1];
    out[6] = (QzS[0] + QzS[2]) + QzS[1];
    out[7] = DyT[2] - DyT[0];
    out[8] = DzS[2] - DzS[0];
    out[9] = (DzDy[0] + DzDy[2]) + DzDy[1];
}

// Regularized collision of lbm_func_3d.py:564-604 on the moments (rho, j, Pi)
// of the pulled populations, written as trace / deviator relaxation:
//   S'_aa = (1 - Om) (dev_aa / rho + u.u / 3) + Om u_a^2 + F_a u_a / rho
//   S'_ab = (1 - Om) Pi_ab / rho + Om u_a u_b + (F_a u_b + F_b u_a) / (2 rho)
// which equals the reference's Pi'/rho - 1/3 without the cancellation
// against 1/3. out: rho, u + F/(2 rho), S (storage order). Without a
// body-force field the forcing terms are left out at compile time.
template <bool FORCE>
__device__ __forceinline__ void collide27(const float* mo, const float* F, float omega, float* out)
{
    const float rho = mo[0], pxx = mo[4], pyy = mo[5], pzz = mo[6];
    const float inv = 1.0f / rho;
    const float t3 = (pxx + pyy + pzz) * (1.0f / 3.0f);
    const float om1 = 1.0f - omega;
    out[0] = rho;
    if (FORCE) {
        const float hx = 0.5f * F[0] * inv, hy = 0.5f * F[1] * inv, hz = 0.5f * F[2] * inv;  // F / (2 rho)
        const float ux = fmaf(mo[1], inv, hx), uy = fmaf(mo[2], inv, hy), uz = fmaf(mo[3], inv, hz);
        const float usq3 = fmaf(ux, ux, fmaf(uy, uy, uz * uz)) * (1.0f / 3.0f);
        out[1] = ux + hx;
        out[2] = uy + hy;
        out[3] = uz + hz;
        // forcing terms F_a u_b / rho written as 2 h_a u_b
        out[4] = fmaf(om1, fmaf(pxx - t3, inv, usq3), fmaf(omega * ux, ux, 2.0f * hx * ux));
        out[5] = fmaf(om1, fmaf(pyy - t3, inv, usq3), fmaf(omega * uy, uy, 2.0f * hy * uy));
        out[6] = fmaf(om1, fmaf(pzz - t3, inv, usq3), fmaf(omega * uz, uz, 2.0f * hz * uz));
        out[7] = fmaf(om1 * mo[7], inv, fmaf(omega * ux, uy, fmaf(hx, uy, hy * ux)));
        out[8] = fmaf(om1 * mo[8], inv, fmaf(omega * ux, uz, fmaf(hx, uz, hz * ux)));
        out[9] = fmaf(om1 * mo[9], inv, fmaf(omega * uy, uz, fmaf(hy, uz, hz * uy)));
    } else {
        const float ux = mo[1] * inv, uy = mo[2] * inv, uz = mo[3] * inv;
        const float usq3 = fmaf(ux, ux, fmaf(uy, uy, uz * uz)) * (1.0f / 3.0f);
        out[1] = ux;
        out[2] = uy;
        out[3] = uz;
        out[4] = fmaf(om1, fmaf(pxx - t3, inv, usq3), omega * ux * ux);
        out[5] = fmaf(om1, fmaf(pyy - t3, inv, usq3), omega * uy * uy);
        out[6] = fmaf(om1, fmaf(pzz - t3, inv, usq3), omega * uz * uz);
        out[7] = fmaf(om1 * mo[7], inv, omega * ux * uy);
        out[8] = fmaf(om1 * mo[8], inv, omega * ux * uz);
        out[9] = fmaf(om1 * mo[9], inv, omega * uy * uz);
    }
}

// One face condition with the equilibrium S in closed form: f_eq_home_3d
// (lbm_func_3d.py:651-686) builds unclamped populations whose second moment
// is rho (I/3 + u u), so S_eq = u u and
// S_b = u_b u_b - u_n u_n + S_n (apply_bc_core_3d, lbm_func_3d.py:643-648).
__device__ __forceinline__ void bc_face_closed(float* m, int type, const float* value, int axis)
{
    const float un[3] = {m[1], m[2], m[3]};
    float ub[3] = {m[1], m[2], m[3]};
    if (type == 0) { ub[0] = value[0]; ub[1] = value[1]; ub[2] = value[2]; }
    else if (type == 2) ub[axis] = 0.0f;
    m[4] += ub[0] * ub[0] - un[0] * un[0];
    m[5] += ub[1] * ub[1] - un[1] * un[1];
    m[6] += ub[2] * ub[2] - un[2] * un[2];
    m[7] += ub[0] * ub[1] - un[0] * un[1];
    m[8] += ub[0] * ub[2] - un[0] * un[2];
    m[9] += ub[1] * ub[2] - un[1] * un[2];
    m[1] = ub[0]; m[2] = ub[1]; m[3] = ub[2];
}

// Raw buffer access to a state array: one resource spans all its fields, the
// per-lane offset selects the column (or kOOB: loads read 0, stores are
// dropped by the hardware range check), the scalar offset the field and
// plane. Every lane issues the same loads and stores whatever its role, so
// the step's memory operations are the same on every control-flow path and
// the in-order vector-memory counter waits for exactly the loads in use.
constexpr uint32_t kOOB = 0x80000000u;
#ifndef WAVES_ATTR
#define WAVES_ATTR
#endif
constexpr int kNT = 2;  // non-temporal: the step streams the whole lattice once

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}

template <int DT> struct BufIO;
template <> struct BufIO<HL_F32> {
    static constexpr uint32_t size = 4;
    __device__ static float load(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s, int, float, float)
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)v, (int)s, 0));
    }
    __device__ static void store(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s, float x, int, float, float, uint32_t)
    {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, x), r, (int)v, (int)s, kNT);
    }
};
// 16-bit storage: the encodings of Storage<HL_F16> / Storage<HL_BF16>
template <> struct BufIO<HL_F16> {
    static constexpr uint32_t size = 2;
    __device__ static float load(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s, int, float inv_scale, float offset)
    {
        const uint16_t h = __builtin_amdgcn_raw_buffer_load_b16(r, (int)v, (int)s, 0);
        return (float)__builtin_bit_cast(_Float16, h) * inv_scale + offset;
    }
    __device__ static void store(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s, float x, int, float scale,
                                 float offset, uint32_t)
    {
        const _Float16 h = (_Float16)((x - offset) * scale);
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, h), r, (int)v, (int)s, kNT);
    }
};
template <> struct BufIO<HL_BF16> {
    static constexpr uint32_t size = 2;
    __device__ static float load(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s, int, float inv_scale, float offset)
    {
        return bf16_to_f32(__builtin_amdgcn_raw_buffer_load_b16(r, (int)v, (int)s, 0)) * inv_scale + offset;
    }
    __device__ static void store(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s, float x, int, float scale,
                                 float offset, uint32_t key)
    {
        __builtin_amdgcn_raw_buffer_store_b16(f32_to_bf16_sr((x - offset) * scale, key), r, (int)v, (int)s, kNT);
    }
};

// Bits [o, o + 64) of a column's cut bitstring (nw words; bits outside it are 0).
__device__ __forceinline__ uint64_t column_window(const uint64_t* col, int nw, int o)
{
    const int j = o >> 6, sh = o & 63;  // floor division, also for o < 0
    const uint64_t lo = (j >= 0 && j < nw) ? col[j] : 0ull;
    const uint64_t hi = (j + 1 >= 0 && j + 1 < nw) ? col[j + 1] : 0ull;
    return sh == 0 ? lo : (lo >> sh) | (hi << (64 - sh));
}

}  // namespace

// Orders a wavefront's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Near-body share of workgroup blk of the tiled launch (the cells of fx.list
// dealt two per wavefront and pass over the first fx.nfix workgroups, one
// lane per direction): the wall rule and the collision in the reference's
// operation order, as the stand-alone fix-up kernel does (aux_kernels.hip),
// with per-wave FP64 force sums reduced in wave order to one partial row per
// workgroup (summed by the next launch). lds: the launch's LDS, used before
// the workgroup's tiled pass.
template <int DT, int NT>
__device__ __forceinline__ void fixup_block(const EnvArgs& a, const FixArgs& fx, float* lds, int blk)
{
    constexpr int W = NT / 64;
    float(*pops)[2][32] = reinterpret_cast<float(*)[2][32]>(lds);
    float(*sums)[2][16] = reinterpret_cast<float(*)[2][16]>(lds + W * 64);
    double* accs = reinterpret_cast<double*>(lds + W * 96);
    // the pose table, copied to LDS with the first loads: the wall rule of a
    // cut link reads it without another global-memory round trip
    float* poses = lds + W * 96 + 2 * W * HL_MAX_SOLIDS * 6;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, h = lane >> 5, i = lane & 31;
    const int nw = fx.nfix * W, gw = blk * W + wave;
    const int width = fx.partial ? fx.width : 0;
    double* acc = accs + wave * width;
    for (int k = lane; k < width; k += 64) acc[k] = 0.0;
    for (int k = tid; k < a.nsolid * 24; k += NT) poses[k] = a.solids[k];
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const Region rg = fx.rg;
    // the count and this lane's first list entry in one round trip: the entry
    // is read inside the list's capacity (the region's cell count) and used
    // only when it is below the count
    const int first = gw * 2 + h;
    const int ncut = *fx.count;
    const int tt_first = fx.list[min((int64_t)first, rg.n() - 1)];
    __syncthreads();  // the pose table is in LDS
    for (int base = gw * 2; base < ncut; base += nw * 2) {
        const int e = base + h;
        const bool cell = e < ncut;
        int x = 0, y = 0, z = 0, px = 0, py = 0, pz = 0, mval_c = 0, sid = -1;
        int64_t c = 0;
        float F[6] = {0, 0, 0, 0, 0, 0};
        if (cell) {
            const int tt = e == first ? tt_first : fx.list[e];
            x = rg.x0 + tt % rg.w;
            y = rg.y0 + (tt / rg.w) % rg.h;
            z = rg.z0 + tt / (rg.w * rg.h);
            px = min(max(x, 1), nx - 2);
            py = min(max(y, 1), ny - 2);
            pz = min(max(z, 1), nz - 2);
            c = (int64_t)px + (int64_t)nx * ((int64_t)py + (int64_t)ny * pz);
            mval_c = a.caps ? 0 : a.mask[c];
        }
        if (cell && i < 27) {
            float mc[10];
            load_cell<float, DT>(a, cell_at(a, px, py, pz), 10, mc);
            WallPoint<float> w{-1, {0.f, 0.f, 0.f}};
            pops[wave][h][i] = d3q27_pull<float, DT>(a, px, py, pz, i, mc, true, mval_c, poses, &w);
            // force of this lane's link (interior cells only: a face cell repeats its projection)
            if (width && px == x && py == y && pz == z && w.sid >= 0 && w.sid < a.nsolid) {
                sid = w.sid;
                link_force_at<float>(a, mc, i, sid, w.h[0], w.h[1], w.h[2], F, poses);
            }
        }
        wave_sync();
        if (cell && i < 16) sums[wave][h][i] = d3q27_group_sum<float>(pops[wave][h], i);
        wave_sync();
        if (cell && i == 0) {
            float Fx = 0, Fy = 0, Fz = 0, m[10];
            if (a.force) { Fx = a.force[c]; Fy = a.force[a.fs + c]; Fz = a.force[2 * a.fs + c]; }
            d3q27_collide_sums<float>(a, sums[wave][h], Fx, Fy, Fz, m);
            if (px != x || py != y || pz != z) bc_chain<float, 3>(m, x, y, z, nx, ny, nz, a.bc_type, a.bc_value);
            using ST = typename Storage<DT>::type;
            ST* dst = static_cast<ST*>(a.dst);
            const int64_t d = cell_at(a, x, y, z);
#pragma unroll
            for (int f = 0; f < 10; ++f)
                Storage<DT>::template store<float>(dst, f * a.stride + d, m[f], f, a.scale[f], a.offset[f],
                                                   HL_ROUNDING_KEY(a, f, x, y, z));
        }
        if (width) wave_solid_sums(sid, F, kRow6, lane, acc);
        wave_sync();  // pops and sums are rewritten by the next pass
    }
    if (!width) return;
    // the workgroup's waves in order, one partial row per workgroup
    __syncthreads();
    for (int k = tid; k < width; k += NT) {
        double v = 0.0;
        for (int w = 0; w < W; ++w) v += accs[w * width + k];
        fx.partial[(int64_t)blk * width + k] = v;
    }
}

// The previous step's force rows: output k is summed by wave k % W of
// workgroup (k / W) % nblk (the world's workgroups), its lanes taking rows
// lane, lane + 64, ... in order and a butterfly combining them (the order of
// sum_partials, so the sums agree bitwise).
template <int NT>
__device__ __forceinline__ void sum_previous(const FixArgs& fx, int blk, int nblk)
{
    constexpr int W = NT / 64;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int k = blk * W + wave; k < fx.width; k += nblk * W) {
        const double v = wave_sum(lane_rows_sum(fx.prev, fx.prev_rows, fx.width, k, lane));
        if (lane == 0) fx.prev_out[k] = v;
    }
}

// The tiled pass of workgroup L of one lattice's ntile workgroups.
template <int DT, int TX, int TY, bool FORCE>
__device__ __forceinline__ void tiled_pass(const EnvArgs& a, int zc, int gx, int gy, int ntile, const FixArgs& fx,
                                           int L)
{
    constexpr int NX = TX + 2, NY = TY + 2, NCOL = NX * NY;
    constexpr int NT = (NCOL + 63) / 64 * 64;
    using IO = BufIO<DT>;
    constexpr uint32_t E = IO::size;
    __shared__ float sp[2][27][NCOL];  // double-buffered: one barrier per plane
    static_assert(NT / 64 * (96 * 4 + HL_MAX_SOLIDS * 6 * 8) + HL_MAX_SOLIDS * 24 * 4 <= sizeof(sp),
                  "near-body work fits the LDS");
    const int tid = threadIdx.x;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    // XCD-aware tile order: workgroups are dealt to the 8 XCDs round-robin in
    // dispatch order, so dispatch slot L runs on XCD L % 8. Give XCD k the
    // contiguous run of tiles [k * per, (k + 1) * per) (x fastest, then y,
    // then z) so that tiles sharing halo cache lines share one L2.
    const int per = ntile / 8;
    const int T = L < per * 8 ? (L % 8) * per + L / 8 : L;
    const int x0 = (T % gx) * TX, y0 = ((T / gx) % gy) * TY;
    const int zb = 1 + (T / (gx * gy)) * zc;
    const int ze = min(zb + zc, nz - 1);  // destination planes [zb, ze), at most 64

    // byte offsets: per field and per plane (scalar), per column (lane)
    const int rp = (int)a.rp;
    const uint32_t fbytes = (uint32_t)a.stride * E, pbytes = (uint32_t)(rp * ny) * E;
    const uint32_t sbytes = (uint32_t)state_elems(a) * E;
    const __amdgpu_buffer_rsrc_t rsrc = buffer_rsrc(a.src, sbytes);
    const __amdgpu_buffer_rsrc_t rdst = buffer_rsrc(a.dst, sbytes);
    const __amdgpu_buffer_rsrc_t rfor = buffer_rsrc(a.force, FORCE ? 3 * (uint32_t)a.fs * 4 : 0);

    // source column of this thread (tile + halo) and its LDS slot (the
    // column's index in the window); lanes past the halo keep a slot of
    // their own out of the way. With 64-wide tiles each of the first NY
    // waves loads one whole window row from x0 (two 128-byte lines per FP32
    // field when rows start on a line, as for nx a multiple of 32) and the
    // next 2 * NY threads take the side halo columns.
    int wsl = tid;
    if (TX == 64) {
        if (tid < NY * TX) wsl = (tid / TX) * NX + 1 + tid % TX;
        else if (tid < NCOL) {
            const int j = tid - NY * TX;
            wsl = (j >> 1) * NX + ((j & 1) ? NX - 1 : 0);
        }
    }
    const int sgx = x0 - 1 + wsl % NX, sgy = y0 - 1 + wsl / NX;
    const bool src_lds = tid < NCOL;
    const bool src_ok = src_lds && sgx >= 0 && sgx < nx && sgy >= 0 && sgy < ny;
    const uint32_t vsrc = src_ok ? (uint32_t)(sgy * rp + sgx) * E : kOOB;

    // destination cell of this thread
    const int dx = x0 + tid % TX, dy = y0 + tid / TX;
    const bool dst_ok = tid < TX * TY && dx < nx && dy < ny;
    const int px = min(max(dx, 1), nx - 2), py = min(max(dy, 1), ny - 2);
    const int cp = dst_ok ? (py - y0 + 1) * NX + (px - x0 + 1) : 0;
    const uint32_t vdst = dst_ok ? (uint32_t)(dy * rp + dx) * E : kOOB;
    const uint32_t vfor = dst_ok ? (uint32_t)(py * nx + px) * 4 : kOOB;

    const float omega = 1.0f / (a.nu * 3.0f + 0.5f);
    // cells of this column the near-body pass owns, bit k for plane zb + k
    // (from the per-column cut bitmap; only tiles that meet its region)
    uint64_t own = 0;
    if (a.cut_cols != nullptr && x0 + TX > a.cut_origin[0] && x0 < a.cut_origin[0] + a.cut_extent[0] &&
        y0 + TY > a.cut_origin[1] && y0 < a.cut_origin[1] + a.cut_extent[1]) {
        const int rx = dx - a.cut_origin[0], ry = dy - a.cut_origin[1];
        if (dst_ok && rx >= 0 && rx < a.cut_extent[0] && ry >= 0 && ry < a.cut_extent[1])
            own = column_window(a.cut_cols + (int64_t)(ry * a.cut_extent[0] + rx) * a.cut_colwords, a.cut_colwords,
                                zb - a.cut_origin[2]);
    }

    auto load_plane = [&](int z, float* m) {
        const uint32_t sz = (uint32_t)z * pbytes;
#pragma unroll
        for (int f = 0; f < 10; ++f) m[f] = IO::load(rsrc, vsrc, sz + f * fbytes, f, a.inv_scale[f], a.offset[f]);
    };
    auto load_force = [&](int z, float* F) {
        const uint32_t sz = (uint32_t)z * (uint32_t)(nx * ny) * 4;
#pragma unroll
        for (int q = 0; q < 3; ++q) F[q] = BufIO<HL_F32>::load(rfor, vfor, sz + q * (uint32_t)a.fs * 4, 0, 1.0f, 0.0f);
    };
    // BF16 rounding key of this thread's destination cell in plane 0
    const uint32_t kcell = (uint32_t)dx + (uint32_t)nx * (uint32_t)dy, kplane = (uint32_t)nx * (uint32_t)ny;
    auto store_plane = [&](int z, uint32_t v, const float* m) {
        const uint32_t sz = (uint32_t)z * pbytes, k0 = kcell + (uint32_t)z * kplane;
#pragma unroll
        for (int f = 0; f < 10; ++f)
            IO::store(rdst, v, sz + f * fbytes, m[f], f, a.scale[f], a.offset[f], k0 + (uint32_t)f * kplane * (uint32_t)nz);
    };

    // up pops of the last two source planes, alternating by plane parity so
    // that no hold has to be copied into another
    float hold_mid[9], up_a[9], up_b[9];
    float pfa[10], pfb[10];  // prefetched source moments, alternating by plane
    float fa[3] = {0.0f, 0.0f, 0.0f}, fb[3] = {0.0f, 0.0f, 0.0f};  // body force of the next dest plane

    // Source work for plane k: rebuild its 27 populations, write to buf the
    // down pops of plane k, the held mid pops of plane k-1 and the held up
    // pops of plane k-2, then shift the holds.
    // U holds the up pops of plane k-2 on entry and those of plane k on exit.
    auto source_plane = [&](float(*buf)[NCOL], const float* m, float* U) {
        float f[27];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            buf[kMid[k]][wsl] = hold_mid[k];
            buf[kUp[k]][wsl] = U[k];
        }
        recon27(m, f);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            buf[kDown[k]][wsl] = f[kDown[k]];
            U[k] = f[kUp[k]];
            hold_mid[k] = f[kMid[k]];
        }
    };

    // Pull, collide and apply the x and y faces for the cell of plane z
    // (lanes without a destination cell leave out untouched).
    auto dest_plane = [&](float(*buf)[NCOL], const float* F, float* out) {
        if (dst_ok) {
            float g[27], mo[10];
#pragma unroll
            for (int i = 0; i < 27; ++i) g[i] = buf[i][cp - D3CX[i] - D3CY[i] * NX];
            moments27(g, mo);
            collide27<FORCE>(mo, F, omega, out);
            if (dx == 0) bc_face_closed(out, a.bc_type[0], a.bc_value[0], 0);
            else if (dx == nx - 1) bc_face_closed(out, a.bc_type[1], a.bc_value[1], 0);
            if (dy == ny - 1) bc_face_closed(out, a.bc_type[2], a.bc_value[2], 1);
            else if (dy == 0) bc_face_closed(out, a.bc_type[3], a.bc_value[3], 1);
        }
    };

    // One plane step: source plane z+2 from `use` into the other buffer,
    // prefetch plane z+3 into `into` and the next plane's body force into
    // fin, pull/collide/store plane z; one barrier. EDGE: the first or last
    // steps of the chunk (source/prefetch limits, the z faces); the steady
    // steps issue the same memory operations on every path.
    auto step = [&](int z, const float* use, float* into, float* U, const float* fuse, float* fin, bool edge) {
        const int cur = (z - zb) & 1;
        if (!edge || z + 1 < ze) {
            // the set of plane z+3 was consumed a step ago: refill it before
            // the source work, so the loads have a whole step to land
            if (!edge || z + 3 <= ze) load_plane(z + 3, into);
            if (FORCE && (!edge || z + 1 < ze)) load_force(z + 1, fin);
            if (src_lds) source_plane(sp[cur ^ 1], use, U);
        }
        float out[10] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        dest_plane(sp[cur], fuse, out);
        // cells the near-body pass owns are left to it
        store_plane(z, (own >> (z - zb)) & 1ull ? kOOB : vdst, out);
        if (edge && z == 1) {  // face z = 0 from the cell above it
            float face[10];
#pragma unroll
            for (int k = 0; k < 10; ++k) face[k] = out[k];
            bc_face_closed(face, a.bc_type[5], a.bc_value[5], 2);
            store_plane(0, dst_ok && !fixup_owns(a, dx, dy, 0) ? vdst : kOOB, face);
        }
        if (edge && z == nz - 2) {  // face z = nz-1 from the cell below it
            bc_face_closed(out, a.bc_type[4], a.bc_value[4], 2);
            store_plane(nz - 1, dst_ok && !fixup_owns(a, dx, dy, nz - 1) ? vdst : kOOB, out);
        }
        __syncthreads();
    };

    // the near-body cells of this step, spread over the first fx.nfix
    // workgroups ahead of their tiled pass (tiled cells skip them). They run
    // before the tiled prologue's loads are issued: vector-memory loads
    // complete in order, so behind those loads each of their dependent round
    // trips would wait for whole planes
    if (fx.list) {
        if (L < fx.nfix) fixup_block<DT, NT>(a, fx, &sp[0][0][0], L);
        __syncthreads();  // the LDS is free again
    }
    // prologue: planes zb-1 (up pops), zb (mid/up pops), zb+1 (into buffer 0)
    {
        float m0[10], m1[10], m2[10];
        load_plane(zb - 1, m0);
        load_plane(zb, m1);
        load_plane(zb + 1, m2);
        if (zb + 2 <= ze) load_plane(zb + 2, pfa);
        if (FORCE) load_force(zb, fa);
        if (src_lds) {
            float f[27];
            recon27(m0, f);
#pragma unroll
            for (int k = 0; k < 9; ++k) up_a[k] = f[kUp[k]];
            recon27(m1, f);
#pragma unroll
            for (int k = 0; k < 9; ++k) { hold_mid[k] = f[kMid[k]]; up_b[k] = f[kUp[k]]; }
            source_plane(sp[0], m2, up_a);  // plane zb+1: up(zb-1) out, up(zb+1) in
        }
    }
    __syncthreads();

    // Software pipeline: step z pulls plane z from buffer (z-zb)&1 while
    // rebuilding the sources of plane z+2 (prefetched one step earlier) into
    // the other buffer, then prefetches plane z+3; one barrier per plane. The
    // steady loop is unrolled by two so each prefetch lands in the registers
    // it is consumed from (no loop-carried copies that would wait on the
    // loads); step z rebuilds source plane z+2, whose up-pop hold is up_b
    // when z - zb is even and up_a when odd. The first step (face z = 0) and
    // the last three (fewer sources, face z = nz-1) are peeled.
    step(zb, pfa, pfb, up_b, fa, fb, true);
    int z = zb + 1;
    for (; z + 4 <= ze; z += 2) {
        step(z, pfb, pfa, up_a, fb, fa, false);
        step(z + 1, pfa, pfb, up_b, fa, fb, false);
    }
    if (z < ze) {
        step(z, pfb, pfa, up_a, fb, fa, true);
        if (z + 1 < ze) {
            step(z + 1, pfa, pfb, up_b, fa, fb, true);
            if (z + 2 < ze) step(z + 2, pfb, pfa, up_a, fb, fa, true);
        }
    }
    // the previous step's force rows (a few waves of the grid), after the march
    if (fx.list && fx.prev) sum_previous<NT>(fx, L, ntile);
}

template <int DT, int TX, int TY, bool FORCE>
__global__ void __launch_bounds__(((TX + 2) * (TY + 2) + 63) / 64 * 64) WAVES_ATTR
k_tiled_d3q27(EnvArgs a, int zc, int gx, int gy, int ntile, FixArgs fx)
{
    tiled_pass<DT, TX, TY, FORCE>(a, zc, gx, gy, ntile, fx, (int)blockIdx.x);
}

// Several lattices in one launch: world k owns workgroups [base[k],
// base[k] + ntile[k]) of the grid; base[k] is a multiple of 8, so a world's
// local workgroup index has the same XCD as its global one and each world's
// tiles keep the XCD-aware order (the workgroups past ntile[k] up to the next
// world's base return at once).
template <int DT, int TX, int TY, bool FORCE>
__global__ void __launch_bounds__(((TX + 2) * (TY + 2) + 63) / 64 * 64) WAVES_ATTR
k_tiled_worlds(TiledWorlds w)
{
    int k = 0;
    while (k + 1 < w.n && (int)blockIdx.x >= w.base[k + 1]) ++k;
    const int L = (int)blockIdx.x - w.base[k];
    if (L >= w.ntile[k]) return;
    tiled_pass<DT, TX, TY, FORCE>(w.a[k], w.zc, w.gx[k], w.gy[k], w.ntile[k], w.fx[k], L);
}

// Tile shape of the D3Q27 kernel (set by hl_set_tiling / the autotuner).
static TileCfg g_tile3d{64, 8, 16};

bool tiling_supported(int lattice, int tx, int ty, int zc)
{
    if (lattice != HL_D3Q27 || zc < 1 || zc > 64) return false;
    return (tx == 32 && ty == 16) || (tx == 32 && ty == 8) || (tx == 64 && ty == 8) || (tx == 64 && ty == 4) ||
           (tx == 16 && ty == 16);
}

void set_tiling(int lattice, TileCfg c)
{
    if (lattice == HL_D3Q27) g_tile3d = c;
}

TileCfg get_tiling(int lattice) { return g_tile3d; }

namespace {

TileCfg tile_config(const EnvArgs&) { return g_tile3d; }

template <int DT, int TX, int TY>
hipError_t launch_tiled(const EnvArgs& a, int zc, FixArgs* fxp, hipStream_t s)
{
    // every tile must hold at least 3 columns per axis for the face projection
    if ((a.nx % TX) == 1 || (a.nx % TX) == 2 || (a.ny % TY) == 1 || (a.ny % TY) == 2) return hipErrorNotSupported;
    if (zc < 1 || zc > 64) return hipErrorInvalidValue;
    // buffer offsets are 32-bit with the out-of-range marker at 2^31
    if ((uint64_t)state_elems(a) * BufIO<DT>::size >= (1ull << 31) || (a.force && (uint64_t)a.fs * 3 * 4 >= (1ull << 31)))
        return hipErrorNotSupported;
    constexpr int threads = ((TX + 2) * (TY + 2) + 63) / 64 * 64;
    const int gx = (a.nx + TX - 1) / TX, gy = (a.ny + TY - 1) / TY, gz = (a.nz - 2 + zc - 1) / zc;
    const int ntile = gx * gy * gz;
    FixArgs fx;
    std::memset(&fx, 0, sizeof(fx));
    if (fxp && fxp->list) {
        fx = *fxp;
        // every workgroup (up to the partial buffer's rows) takes a share of the
        // near-body cells before its tiled pass
        fx.nfix = ntile < fx.max_waves ? ntile : fx.max_waves;
        if (fx.nfix < 1) return hipErrorInvalidValue;
        fxp->nfix = fx.nfix;
    }
    const dim3 grid(ntile);
    // the body-force field is a template switch: a run-time branch around its
    // loads would make the compiler wait on the in-flight prefetch at the merge
    if (a.force) hipLaunchKernelGGL((k_tiled_d3q27<DT, TX, TY, true>), grid, dim3(threads), 0, s, a, zc, gx, gy, ntile, fx);
    else hipLaunchKernelGGL((k_tiled_d3q27<DT, TX, TY, false>), grid, dim3(threads), 0, s, a, zc, gx, gy, ntile, fx);
    return hipGetLastError();
}

// Launch of several worlds: each world's tiles (padded to a multiple of 8
// workgroups) in one grid, FixArgs set up as launch_tiled does per world.
template <int DT, int TX, int TY>
hipError_t launch_tiled_worlds(const EnvArgs* a, FixArgs* const* fxp, int n, int zc, hipStream_t s)
{
    if (n < 1 || n > kTiledWorlds || zc < 1 || zc > 64) return hipErrorInvalidValue;
    TiledWorlds w;
    std::memset(&w, 0, sizeof(w));
    w.zc = zc;
    w.n = n;
    const bool force = a[0].force != nullptr;
    int base = 0;
    for (int k = 0; k < n; ++k) {
        const EnvArgs& e = a[k];
        if ((e.force != nullptr) != force) return hipErrorNotSupported;
        if ((e.nx % TX) == 1 || (e.nx % TX) == 2 || (e.ny % TY) == 1 || (e.ny % TY) == 2) return hipErrorNotSupported;
        if ((uint64_t)state_elems(e) * BufIO<DT>::size >= (1ull << 31) || (e.force && (uint64_t)e.fs * 3 * 4 >= (1ull << 31)))
            return hipErrorNotSupported;
        w.a[k] = e;
        w.gx[k] = (e.nx + TX - 1) / TX;
        w.gy[k] = (e.ny + TY - 1) / TY;
        w.ntile[k] = w.gx[k] * w.gy[k] * ((e.nz - 2 + zc - 1) / zc);
        w.base[k] = base;
        if (fxp[k] && fxp[k]->list) {
            w.fx[k] = *fxp[k];
            w.fx[k].nfix = w.ntile[k] < w.fx[k].max_waves ? w.ntile[k] : w.fx[k].max_waves;
            if (w.fx[k].nfix < 1) return hipErrorInvalidValue;
            fxp[k]->nfix = w.fx[k].nfix;
        }
        base += (w.ntile[k] + 7) / 8 * 8;
    }
    w.base[n] = base;
    constexpr int threads = ((TX + 2) * (TY + 2) + 63) / 64 * 64;
    if (force) hipLaunchKernelGGL((k_tiled_worlds<DT, TX, TY, true>), dim3(base), dim3(threads), 0, s, w);
    else hipLaunchKernelGGL((k_tiled_worlds<DT, TX, TY, false>), dim3(base), dim3(threads), 0, s, w);
    return hipGetLastError();
}

template <int DT>
hipError_t dispatch_tile_worlds(const EnvArgs* a, FixArgs* const* fx, int n, hipStream_t s)
{
    const TileCfg c = tile_config(a[0]);
    if (c.tx == 32 && c.ty == 16) return launch_tiled_worlds<DT, 32, 16>(a, fx, n, c.zc, s);
    if (c.tx == 32 && c.ty == 8) return launch_tiled_worlds<DT, 32, 8>(a, fx, n, c.zc, s);
    if (c.tx == 64 && c.ty == 8) return launch_tiled_worlds<DT, 64, 8>(a, fx, n, c.zc, s);
    if (c.tx == 64 && c.ty == 4) return launch_tiled_worlds<DT, 64, 4>(a, fx, n, c.zc, s);
    if (c.tx == 16 && c.ty == 16) return launch_tiled_worlds<DT, 16, 16>(a, fx, n, c.zc, s);
    return hipErrorNotSupported;
}

template <int DT>
hipError_t dispatch_tile(const EnvArgs& a, FixArgs* fx, hipStream_t s)
{
    const TileCfg c = tile_config(a);
    if (c.tx == 32 && c.ty == 16) return launch_tiled<DT, 32, 16>(a, c.zc, fx, s);
    if (c.tx == 32 && c.ty == 8) return launch_tiled<DT, 32, 8>(a, c.zc, fx, s);
    if (c.tx == 64 && c.ty == 8) return launch_tiled<DT, 64, 8>(a, c.zc, fx, s);
    if (c.tx == 64 && c.ty == 4) return launch_tiled<DT, 64, 4>(a, c.zc, fx, s);
    if (c.tx == 16 && c.ty == 16) return launch_tiled<DT, 16, 16>(a, c.zc, fx, s);
    return hipErrorNotSupported;
}

}  // namespace

bool tiled_fits(const EnvArgs& a, int dtype)
{
    if (dtype != HL_F32 && dtype != HL_F16 && dtype != HL_BF16) return false;
    const TileCfg c = tile_config(a);
    if (!tiling_supported(HL_D3Q27, c.tx, c.ty, c.zc)) return false;
    if ((a.nx % c.tx) == 1 || (a.nx % c.tx) == 2 || (a.ny % c.ty) == 1 || (a.ny % c.ty) == 2) return false;
    const uint64_t esize = dtype == HL_F32 ? 4 : 2;
    return (uint64_t)state_elems(a) * esize < (1ull << 31) && !(a.force && (uint64_t)a.fs * 3 * 4 >= (1ull << 31));
}

hipError_t step_tiled_worlds(int dtype, const EnvArgs* a, FixArgs* const* fx, int n, hipStream_t s)
{
    switch (dtype) {
    case HL_F32: return dispatch_tile_worlds<HL_F32>(a, fx, n, s);
    case HL_F16: return dispatch_tile_worlds<HL_F16>(a, fx, n, s);
    case HL_BF16: return dispatch_tile_worlds<HL_BF16>(a, fx, n, s);
    default: return hipErrorNotSupported;
    }
}

hipError_t step_tiled(int lattice, int dtype, const EnvArgs& a, FixArgs* fx, hipStream_t s)
{
    if (lattice != HL_D3Q27) return hipErrorNotSupported;
    switch (dtype) {
    case HL_F32: return dispatch_tile<HL_F32>(a, fx, s);
    case HL_F16: return dispatch_tile<HL_F16>(a, fx, s);
    case HL_BF16: return dispatch_tile<HL_BF16>(a, fx, s);
    default: return hipErrorNotSupported;  // FP64 runs the gather kernel (the oracle path)
    }
}

}  // namespace hl
