// Device-side building blocks of the HOME-LBM step for gfx950: lattice
// tables, moment storage codecs, population reconstruction, collision, face
// boundary conditions and the wall velocity of a rigid body. Every function is
// templated on the arithmetic type T (float or double) so the same source
// gives the FP32 production path and the FP64 GPU oracle.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "homelbm.h"

namespace hl {

// D2Q9 / D3Q27 velocity sets in the reference's index order
// (envs/lbm/lbm_core.py:127-130, envs/lbm3d/lbm_core_3d.py:12-35). The moment
// sums and the inverse-direction map depend on this order.
constexpr int D2CX[9] = {0, 1, -1, 0, 0, 1, -1, -1, 1};
constexpr int D2CY[9] = {0, 0, 0, 1, -1, 1, 1, -1, -1};
constexpr int D2INV[9] = {0, 2, 1, 4, 3, 7, 8, 5, 6};
constexpr int D3CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
constexpr int D3CY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
constexpr int D3CZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
constexpr int D3INV[27] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25};

// Weights: the exact fractions rounded once to T.
template <typename T> __host__ __device__ constexpr T d2w(int i)
{
    return i == 0 ? T(4.0 / 9.0) : (i <= 4 ? T(1.0 / 9.0) : T(1.0 / 36.0));
}
template <typename T> __host__ __device__ constexpr T d3w(int i)
{
    return i == 0 ? T(8.0 / 27.0) : (i <= 6 ? T(2.0 / 27.0) : (i <= 18 ? T(1.0 / 54.0) : T(1.0 / 216.0)));
}
// c_s^2 as the reference rounds it: the float32 sum of D2Q9 weights
// (lbm_func.py:4) and float32(1/3) (lbm_core_3d.py:37).
template <typename T> __host__ __device__ constexpr T cs2_2d()
{
    return sizeof(T) == 4 ? T(1.0 / 9 + 1.0 / 9 + 1.0 / 36 + 1.0 / 36 + 1.0 / 36 + 1.0 / 36) : T(1.0 / 3.0);
}
template <typename T> __host__ __device__ constexpr T cs2_3d() { return T(1.0 / 3.0); }

// ------------------------------------------------------------ storage

// Moment codecs. 32/64-bit fields are stored as is. 16-bit fields hold
// (value - offset) * scale, offset and scale per field (the density's offset
// includes its 1); decode multiplies by 1/scale and adds the offset (both
// precomputed by the host). FP16 rounds to nearest even. BF16 rounds
// stochastically: up with the probability of the discarded 16-bit fraction,
// the random bits a hash of the element index and the value, so the rounding
// is deterministic and unbiased and increments below half a BF16 ulp survive
// on average instead of being lost every step.
template <int DT> struct Storage;
template <> struct Storage<HL_F32> {
    using type = float;
    template <typename T> __device__ static T load(const float* p, int64_t i, int, float, float) { return T(p[i]); }
    template <typename T> __device__ static void store(float* p, int64_t i, T v, int, float, float, uint32_t)
    {
        p[i] = float(v);
    }
};
template <> struct Storage<HL_F64> {
    using type = double;
    template <typename T> __device__ static T load(const double* p, int64_t i, int, float, float) { return T(p[i]); }
    template <typename T> __device__ static void store(double* p, int64_t i, T v, int, float, float, uint32_t)
    {
        p[i] = double(v);
    }
};
template <> struct Storage<HL_F16> {
    using type = __half;
    template <typename T> __device__ static T load(const __half* p, int64_t i, int, float inv_scale, float offset)
    {
        return T(__half2float(p[i]) * inv_scale + offset);
    }
    template <typename T>
    __device__ static void store(__half* p, int64_t i, T v, int, float scale, float offset, uint32_t)
    {
        p[i] = __float2half_rn((float(v) - offset) * scale);
    }
};
// bfloat16 as raw 16-bit words. key: the logical element index
// f * ncell + x + nx (y + ny z) (rounding_key), the same in every kernel,
// layout and tiling. The hash is a 32-bit integer mix of key ^ value bits
// (an xor-shift, then an odd multiplier); the product's high 16 bits are the
// random fraction added before truncation.
__device__ __forceinline__ uint16_t f32_to_bf16_sr(float v, uint32_t key)
{
    const uint32_t u = __float_as_uint(v);
    uint32_t h = key ^ u;
    h ^= h >> 16;
    h *= 0x7feb352du;  // one odd multiplier; the product's high half is the well-mixed part
    // gfx950's stochastic-rounding conversion: the bits (u + (h >> 16)) >> 16
    // for every non-NaN input (measured bit for bit over a million values,
    // denormals and infinities included: the high half carries into the kept
    // bits with probability fraction / 2^16); a NaN stays a NaN
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bf16x2 z = {(__bf16)0.0f, (__bf16)0.0f};
    return __builtin_bit_cast(uint16_t, __builtin_amdgcn_cvt_sr_bf16_f32(z, v, h, false)[0]);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

template <> struct Storage<HL_BF16> {
    using type = uint16_t;
    template <typename T> __device__ static T load(const uint16_t* p, int64_t i, int, float inv_scale, float offset)
    {
        return T(bf16_to_f32(p[i]) * inv_scale + offset);
    }
    template <typename T>
    __device__ static void store(uint16_t* p, int64_t i, T v, int, float scale, float offset, uint32_t key)
    {
        p[i] = f32_to_bf16_sr((float(v) - offset) * scale, key);
    }
};

// BF16 rounding key of field f at cell (x, y, z): its logical element index.
#define HL_ROUNDING_KEY(a, f, x, y, z)                                                                         \
    ((uint32_t)(f) * ((uint32_t)(a).nx * (uint32_t)(a).ny * (uint32_t)(a).nz) + (uint32_t)(x) +                  \
     (uint32_t)(a).nx * ((uint32_t)(y) + (uint32_t)(a).ny * (uint32_t)(z)))

// ------------------------------------------------------- reconstruction

// D2Q9 population i from (rho, u, S) and the third-order terms A_xxy, A_xyy:
// mlCalDistribution, envs/lbm/lbm_func.py:685-788, same operation order.
template <typename T>
__device__ __forceinline__ T d2q9_recon_a(T rho, T ux, T uy, T Sxx, T Syy, T Sxy, T Axxy, T Axyy, int i)
{
    const T one = 1, three = 3, half = T(0.5), c45 = T(4.5), nine = 9;
    const T w0 = T(4) / T(9), w1 = T(1) / T(9), w5 = T(1) / T(36);
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

template <typename T>
__device__ __forceinline__ void d2q9_third(T ux, T uy, T Sxx, T Syy, T Sxy, T& Axxy, T& Axyy)
{
    const T two = 2;
    Axxy = Sxx * uy + two * Sxy * ux - two * ux * ux * uy;
    Axyy = Syy * ux + two * Sxy * uy - two * ux * uy * uy;
}

template <typename T>
__device__ __forceinline__ T d2q9_recon(T rho, T ux, T uy, T Sxx, T Syy, T Sxy, int i)
{
    T Axxy, Axyy;
    d2q9_third<T>(ux, uy, Sxx, Syy, Sxy, Axxy, Axyy);
    return d2q9_recon_a<T>(rho, ux, uy, Sxx, Syy, Sxy, Axxy, Axyy, i);
}

// All 9 populations of one cell's moments m (rho, ux, uy, Sxx, Syy, Sxy),
// each exactly as d2q9_recon computes it, clamped at 0 as on the
// reference's streaming branch (lbm_func.py:409-412).
template <typename T>
__device__ __forceinline__ void d2q9_pops(const T* m, T* f)
{
    T Axxy, Axyy;
    d2q9_third<T>(m[1], m[2], m[3], m[4], m[5], Axxy, Axyy);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        T p = d2q9_recon_a<T>(m[0], m[1], m[2], m[3], m[4], m[5], Axxy, Axyy, i);
        if (p < T(0)) p = T(0);
        f[i] = p;
    }
}

// D3Q27 population with velocity (dx, dy, dz) and weight wi:
// mlCalEquilibrium3D_s, envs/lbm3d/lbm_func_3d.py:224-251 (argument order
// rho, u, Sxx, Sxy, Sxz, Syy, Syz, Szz as in the reference).
template <typename T>
__device__ __forceinline__ T d3q27_recon(T rho, T ux, T uy, T uz, T Sxx, T Sxy, T Sxz, T Syy, T Syz, T Szz,
                                         T dx, T dy, T dz, T wi)
{
    T cu = dx * ux + dy * uy + dz * uz;
    T U_sqr = ux * ux + uy * uy + uz * uz;
    const T third = T(1) / T(3);
    T Qxx = dx * dx - third, Qyy = dy * dy - third, Qzz = dz * dz - third;
    T Qxy = dx * dy, Qxz = dx * dz, Qyz = dy * dz;
    T f_neq = T(1.5) * (Qxx * Sxx + Qyy * Syy + Qzz * Szz + T(2) * (Qxy * Sxy + Qxz * Sxz + Qyz * Syz));
    return wi * rho * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * U_sqr + f_neq);
}

// --------------------------------------------------------- equilibrium S

// S of the second-order equilibrium from its populations: f_eq_home
// (lbm_func.py:660-682, populations clamped at 0) and f_eq_home_3d
// (lbm_func_3d.py:651-686, no clamp). 2D out: xx, yy, xy; 3D out in
// storage order xx, yy, zz, xy, xz, yz.
template <typename T> __device__ __forceinline__ void d2q9_feq_S(T rho, T ux, T uy, T S[3])
{
    T pop[9];
    T U_sqr = ux * ux + uy * uy;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        T cu = T(D2CX[i]) * ux + T(D2CY[i]) * uy;
        pop[i] = rho * d2w<T>(i) * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * U_sqr);
        if (pop[i] < T(0)) pop[i] = T(0);
    }
    T invRho = T(1) / rho;
    T pixx = pop[1] + pop[2] + pop[5] + pop[6] + pop[7] + pop[8];
    T piyy = pop[3] + pop[4] + pop[5] + pop[6] + pop[7] + pop[8];
    T pixy = pop[5] + pop[7] - pop[6] - pop[8];
    S[0] = pixx * invRho - cs2_2d<T>();
    S[1] = piyy * invRho - cs2_2d<T>();
    S[2] = pixy * invRho;
}

template <typename T> __device__ __forceinline__ void d3q27_feq_S(T rho, T ux, T uy, T uz, T S[6])
{
    T p[27];
    T U_sqr = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int i = 0; i < 27; ++i) {
        T cu = T(D3CX[i]) * ux + T(D3CY[i]) * uy + T(D3CZ[i]) * uz;
        p[i] = d3w<T>(i) * rho * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * U_sqr);
    }
    T invRho = T(1) / rho;
    T pixx = p[1] + p[2] + p[7] + p[8] + p[9] + p[10] + p[13] + p[14] + p[15] + p[16] + p[19] + p[20]
           + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
    T pixy = (p[7] + p[8] + p[19] + p[20] + p[21] + p[22]) - (p[13] + p[14] + p[23] + p[24] + p[25] + p[26]);
    T pixz = (p[9] + p[10] + p[19] + p[20] + p[23] + p[24]) - (p[15] + p[16] + p[21] + p[22] + p[25] + p[26]);
    T piyy = p[3] + p[4] + p[7] + p[8] + p[11] + p[12] + p[13] + p[14] + p[17] + p[18] + p[19] + p[20]
           + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
    T piyz = (p[11] + p[12] + p[19] + p[20] + p[25] + p[26]) - (p[17] + p[18] + p[21] + p[22] + p[23] + p[24]);
    T pizz = p[5] + p[6] + p[9] + p[10] + p[11] + p[12] + p[15] + p[16] + p[17] + p[18] + p[19] + p[20]
           + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
    S[0] = pixx * invRho - cs2_3d<T>();
    S[1] = piyy * invRho - cs2_3d<T>();
    S[2] = pizz * invRho - cs2_3d<T>();
    S[3] = pixy * invRho;
    S[4] = pixz * invRho;
    S[5] = piyz * invRho;
}

// ----------------------------------------------------------- face BCs

// One face boundary condition: the boundary cell takes its value from the
// inner neighbour m (apply_bc_core, lbm_func.py:634-657 and
// lbm_func_3d.py:627-648). m is replaced by the boundary cell's value.
// Type 0: velocity; 1: zero gradient; 2: free slip (normal velocity zero,
// no reference counterpart). D = 2 or 3; m = rho, u[D], S[...].
template <typename T, int D>
__device__ __forceinline__ void bc_face(T* m, int type, const float* value, int axis)
{
    T u_n[D], u_b[D];
#pragma unroll
    for (int a = 0; a < D; ++a) { u_n[a] = m[1 + a]; u_b[a] = m[1 + a]; }
    if (type == 0) {
#pragma unroll
        for (int a = 0; a < D; ++a) u_b[a] = T(value[a]);
    } else if (type == 2) {
#pragma unroll
        for (int a = 0; a < D; ++a) u_b[a] = (a == axis) ? T(0) : u_n[a];
    }
    const T rho_n = m[0];
    if (D == 2) {
        T Sb[3], Sn[3];
        d2q9_feq_S<T>(rho_n, u_b[0], u_b[1], Sb);
        d2q9_feq_S<T>(rho_n, u_n[0], u_n[1], Sn);
#pragma unroll
        for (int k = 0; k < 3; ++k) m[3 + k] = Sb[k] - Sn[k] + m[3 + k];
    } else {
        T Sb[6], Sn[6];
        d3q27_feq_S<T>(rho_n, u_b[0], u_b[1], u_b[D - 1], Sb);
        d3q27_feq_S<T>(rho_n, u_n[0], u_n[1], u_n[D - 1], Sn);
#pragma unroll
        for (int k = 0; k < 6; ++k) m[4 + k] = Sb[k] - Sn[k] + m[4 + k];
    }
#pragma unroll
    for (int a = 0; a < D; ++a) m[1 + a] = u_b[a];
}

// Face value of a cell from the post-collision value m of its projection
// onto the interior: x faces, then y faces, then z faces. This is the
// sequential outcome of the reference's one-launch apply_bc{,_3d}
// (lbm_func.py:610-631, lbm_func_3d.py:607-624), whose edge and corner
// threads race; the fixed order removes the race.
template <typename T, int D>
__device__ __forceinline__ void bc_chain(T* m, int x, int y, int z, int nx, int ny, int nz,
                                         const int32_t* bc_type, const float (*bc_value)[3])
{
    // face ids: 2D left 0, top 1, right 2, bottom 3; 3D left 0, right 1, top 2, bottom 3, front 4, back 5
    const int f_right = D == 3 ? 1 : 2, f_top = D == 3 ? 2 : 1;
    if (x == 0) bc_face<T, D>(m, bc_type[0], bc_value[0], 0);
    else if (x == nx - 1) bc_face<T, D>(m, bc_type[f_right], bc_value[f_right], 0);
    if (y == ny - 1) bc_face<T, D>(m, bc_type[f_top], bc_value[f_top], 1);
    else if (y == 0) bc_face<T, D>(m, bc_type[3], bc_value[3], 1);
    if (D == 3) {
        if (z == nz - 1) bc_face<T, D>(m, bc_type[4], bc_value[4], 2);
        else if (z == 0) bc_face<T, D>(m, bc_type[5], bc_value[5], 2);
    }
}

// -------------------------------------------------------- wall velocity

// Velocity of the body material point at h over one step (time step 1):
// its current position minus the position of the same body-frame point one
// step earlier (lbm_func.py:354-370, lbm_func_3d.py:456-463). Pose tables:
// 2D px, py, cos, sin | last px, py, cos, sin; 3D p[3], R[9] | last p[3], R[9].
// r = h - p (torque arm).
template <typename T>
__device__ __forceinline__ void wall_velocity_2d(const float* solids, int nsolid, int sid, T hx, T hy, T up[2], T r[2])
{
    if (sid < 0 || sid >= nsolid) { up[0] = up[1] = r[0] = r[1] = T(0); return; }
    const float* P = solids + 8 * sid;
    T px = P[0], py = P[1], cn = P[2], sn = P[3];
    T pxl = P[4], pyl = P[5], cl = P[6], sl = P[7];
    T dx = hx - px, dy = hy - py;
    T qx = cn * dx + sn * dy;
    T qy = cn * dy - sn * dx;
    T cx = cn * qx - sn * qy + px;
    T cy = sn * qx + cn * qy + py;
    T lx = cl * qx - sl * qy + pxl;
    T ly = sl * qx + cl * qy + pyl;
    up[0] = cx - lx;
    up[1] = cy - ly;
    r[0] = dx;
    r[1] = dy;
}

template <typename T>
__device__ __forceinline__ void wall_velocity_3d(const float* solids, int nsolid, int sid, T hx, T hy, T hz, T up[3], T r[3])
{
    if (sid < 0 || sid >= nsolid) { up[0] = up[1] = up[2] = r[0] = r[1] = r[2] = T(0); return; }
    const float* P = solids + 24 * sid;
    const float* M = P + 3;
    const float* L = P + 12;
    const float* ML = P + 15;
    T d0 = hx - T(P[0]), d1 = hy - T(P[1]), d2 = hz - T(P[2]);
    T q0 = T(M[0]) * d0 + T(M[3]) * d1 + T(M[6]) * d2;
    T q1 = T(M[1]) * d0 + T(M[4]) * d1 + T(M[7]) * d2;
    T q2 = T(M[2]) * d0 + T(M[5]) * d1 + T(M[8]) * d2;
    T c0 = T(M[0]) * q0 + T(M[1]) * q1 + T(M[2]) * q2 + T(P[0]);
    T c1 = T(M[3]) * q0 + T(M[4]) * q1 + T(M[5]) * q2 + T(P[1]);
    T c2 = T(M[6]) * q0 + T(M[7]) * q1 + T(M[8]) * q2 + T(P[2]);
    T l0 = T(ML[0]) * q0 + T(ML[1]) * q1 + T(ML[2]) * q2 + T(L[0]);
    T l1 = T(ML[3]) * q0 + T(ML[4]) * q1 + T(ML[5]) * q2 + T(L[1]);
    T l2 = T(ML[6]) * q0 + T(ML[7]) * q1 + T(ML[8]) * q2 + T(L[2]);
    up[0] = c0 - l0; up[1] = c1 - l1; up[2] = c2 - l2;
    r[0] = d0; r[1] = d1; r[2] = d2;
}

// Correctly rounded square roots (HIP's default for FP32 and FP64).
__device__ __forceinline__ float exact_sqrt(float v) { return __builtin_sqrtf(v); }
__device__ __forceinline__ double exact_sqrt(double v) { return __builtin_sqrt(v); }

// The crossing of the link from cell (x, y, z) towards its source,
// P(t) = X - c t with t in [0, 1], nearest to the cell, over the surfaces of
// the capsules caps[k] = (ax, ay, az, bx, by, bz, r, solid) with
// 0 <= solid < nsolid: the ray query of get_cutcell_multi_3d
// (lbm_func_3d.py:254-316: origin X, direction -c, nearest hit with
// 0 <= t <= 1, the first solid winning ties) with each capsule surface in
// closed form. Every capsule surface counts on its own, as every triangle of
// the reference's concatenated body meshes does: a ray from inside a capsule
// hits where it leaves it.
//
// Along the line a capsule is the hull [lo, hi] of the intervals inside its
// end spheres and inside the cylinder between them (cut to the segment's
// parameter range); the crossing nearest to X is lo when lo >= 0, else hi.
// Capsules whose bounding sphere lies more than 2 > |c| from X are skipped.
// Operation order and rounding follow the C restatement of the same query,
// with contraction off. Returns the solid (-1: no crossing), t in t_hit.
template <typename T>
__device__ __forceinline__ int capsule_cut(const float* caps, int ncap, int nsolid, T x, T y, T z, int cx, int cy,
                                           int cz, T& t_hit)
{
    const T dx = T(-cx), dy = T(-cy), dz = T(-cz);
    const T dd = T(cx * cx + cy * cy + cz * cz);
    const T big = T(1e30);
    T best = T(-1);
    int sid = -1;
    for (int k = 0; k < ncap; ++k) {
        const float* C = caps + 8 * k;
        const int s = (int)C[7];
        if (s < 0 || s >= nsolid) continue;
        const T ax = C[0], ay = C[1], az = C[2], bx = C[3], by = C[4], bz = C[5], rad = C[6];
        const T bax = bx - ax, bay = by - ay, baz = bz - az;
        const T baba = bax * bax + bay * bay + baz * baz;
        const T mx = x - (ax + bx) * T(0.5), my = y - (ay + by) * T(0.5), mz = z - (az + bz) * T(0.5);
        const T reach = exact_sqrt(baba) * T(0.5) + rad + T(2);
        if (mx * mx + my * my + mz * mz > reach * reach) continue;
        const T rr = rad * rad;
        T lo = big, hi = -big;
        const T ox = x - ax, oy = y - ay, oz = z - az;
        const T rdoa = dx * ox + dy * oy + dz * oz;
        const T oaoa = ox * ox + oy * oy + oz * oz;
        {  // end sphere at a
            const T disc = rdoa * rdoa - dd * (oaoa - rr);
            if (disc >= T(0)) {
                const T q = exact_sqrt(disc);
                const T t0 = (-rdoa - q) / dd, t1 = (-rdoa + q) / dd;
                if (t0 < lo) lo = t0;
                if (t1 > hi) hi = t1;
            }
        }
        if (baba > T(0)) {
            {  // end sphere at b
                const T px = x - bx, py = y - by, pz = z - bz;
                const T b = dx * px + dy * py + dz * pz;
                const T disc = b * b - dd * (px * px + py * py + pz * pz - rr);
                if (disc >= T(0)) {
                    const T q = exact_sqrt(disc);
                    const T t0 = (-b - q) / dd, t1 = (-b + q) / dd;
                    if (t0 < lo) lo = t0;
                    if (t1 > hi) hi = t1;
                }
            }
            // cylinder: |P - a|^2 baba - ((P - a).ba)^2 <= r^2 baba with
            // 0 <= (P - a).ba <= baba; a ray parallel to the axis meets it
            // only between the end spheres
            const T bard = bax * dx + bay * dy + baz * dz;
            const T baoa = bax * ox + bay * oy + baz * oz;
            const T qa = baba * dd - bard * bard;
            if (qa > T(0)) {
                const T qb = baba * rdoa - baoa * bard;
                const T qc = baba * (oaoa - rr) - baoa * baoa;
                const T disc = qb * qb - qa * qc;
                if (disc >= T(0)) {
                    const T q = exact_sqrt(disc);
                    T c0 = (-qb - q) / qa, c1 = (-qb + q) / qa;
                    bool ok;
                    if (bard != T(0)) {
                        T s0 = -baoa / bard, s1 = (baba - baoa) / bard;
                        if (bard < T(0)) { const T sw = s0; s0 = s1; s1 = sw; }
                        if (s0 > c0) c0 = s0;
                        if (s1 < c1) c1 = s1;
                        ok = c0 <= c1;
                    } else {
                        ok = baoa >= T(0) && baoa <= baba;
                    }
                    if (ok) {
                        if (c0 < lo) lo = c0;
                        if (c1 > hi) hi = c1;
                    }
                }
            }
        }
        if (lo > hi) continue;
        const T t = lo >= T(0) ? lo : hi;
        if (t >= T(0) && t <= T(1) && (best < T(0) || t < best)) {
            best = t;
            sid = s;
        }
    }
    t_hit = best;
    return sid;
}

// Wavefront-wide FP64 sum and 64-bit OR (butterfly over the 64 lanes).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// A lane's share of the sum of column k over nrows rows of `width`: rows
// lane, lane + 64, ... added in that order, eight loads issued before their
// adds (the same sum as the plain strided loop, without a load latency
// per row).
__device__ __forceinline__ double lane_rows_sum(const double* rows, int nrows, int width, int k, int lane)
{
    double v = 0.0;
    int r = lane;
    for (; r + 64 * 7 < nrows; r += 64 * 8) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = rows[(int64_t)(r + 64 * u) * width + k];
#pragma unroll
        for (int u = 0; u < 8; ++u) v += t[u];
    }
    for (; r < nrows; r += 64) v += rows[(int64_t)r * width + k];
    return v;
}

// N wavefront-wide FP64 sums at once: wave_sum's butterfly with the N
// exchanges of each stage issued together, so their latencies overlap (each
// sum is bitwise that of wave_sum).
template <int N> __device__ __forceinline__ void wave_sums(double* v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double t[N];
#pragma unroll
        for (int q = 0; q < N; ++q) t[q] = __shfl_xor(v[q], off, 64);
#pragma unroll
        for (int q = 0; q < N; ++q) v[q] += t[q];
    }
}

// Per-solid FP64 sums of the lanes' link forces into the solid's force row:
// F[q] goes to acc[6 sid + Q[q]] (lane 0 adds). One pass per solid present
// in the wave (the lowest lane's solid first), each a wave_sums over the
// lanes of that solid; every solid gets one addition per call, so the sums
// are those of one wave_sum per solid and component. sid < 0: no link.
template <int N, typename T>
__device__ __forceinline__ void wave_solid_sums(int sid, const T* F, const int (&Q)[N], int lane, double* acc)
{
    int s = sid;
    for (;;) {
        const uint64_t act = __ballot(s >= 0);
        if (!act) break;
        const int sv = __builtin_amdgcn_readlane(s, (int)__builtin_ctzll(act));
        double v[N];
#pragma unroll
        for (int q = 0; q < N; ++q) v[q] = s == sv ? (double)F[q] : 0.0;
        wave_sums<N>(v);
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < N; ++q) acc[sv * 6 + Q[q]] += v[q];
        }
        if (s == sv) s = -1;
    }
}
constexpr int kRow6[6] = {0, 1, 2, 3, 4, 5};   // F, T of a 3D link (2D: Fx, Fy, 0, 0, 0, Tz)
constexpr int kRow2D[3] = {0, 1, 5};           // Fx, Fy, Tz of a 2D link

__device__ __forceinline__ uint64_t wave_or(uint64_t v)
{
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        lo |= (uint32_t)__shfl_xor((int)lo, off, 64);
        hi |= (uint32_t)__shfl_xor((int)hi, off, 64);
    }
    return ((uint64_t)hi << 32) | lo;
}

// Global-memory view of a generic pointer. Pointers read from memory (not
// kernel arguments) are generic to the compiler, so their accesses become
// flat instructions, which count against the LDS wait counter as well and
// make LDS reads wait for outstanding stores. Casting them lets the compiler
// emit global loads and stores.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* as_global(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}

// Kernel-side view of an hl_env for one precision (by value, so it can be a
// kernel argument and a graph node parameter).
struct EnvArgs {
    int32_t nx, ny, nz, nsolid;
    int64_t stride;  // elements between moment fields of one cell
    int64_t rp;      // elements between consecutive x-rows of a field (nx, or 10 nx row-interleaved)
    int64_t fs;      // elements between body-force components (linear cell index)
    const void* src;
    void* dst;
    const uint8_t* mask;
    int32_t box[6];
    const float* force;
    const float* solids;
    float nu;
    int32_t bc_type[6];
    float bc_value[6][3];
    float scale[10];
    float inv_scale[10];
    float offset[10];  // 16-bit storage offset (the density's includes its 1)
    int32_t region[6];  // cells the gather kernel updates: x0, x1, y0, y1, z0, z1 (half-open)
    // cells whose update the near-body fix-up owns: bit t of cut_bits marks
    // region cell t = (x - cx0) + cw * ((y - cy0) + ch * (z - cz0)); null = none
    const uint64_t* cut_bits;
    int32_t cut_origin[3];
    int32_t cut_extent[3];
    // the same cells per column: word j of cut_cols[(ry * cw + rx) * cut_colwords ...]
    // holds bit k for z = cz0 + 64 j + k
    const uint64_t* cut_cols;
    int32_t cut_colwords;
    // capsule bodies whose crossings cut links (null: the mask rule): [ncap][8]
    const float* caps;
    int32_t ncap;
};

// Offset of cell (x, y, z) within a moment field: x + rp (y + ny z). The
// mask and the body force use the linear index x + nx (y + ny z) instead.
__device__ __forceinline__ int64_t cell_at(const EnvArgs& a, int x, int y, int z)
{
    return (int64_t)x + a.rp * ((int64_t)y + (int64_t)a.ny * z);
}

// Elements one state spans: the last field's offset plus one field's extent.
__host__ __device__ __forceinline__ int64_t state_elems(const EnvArgs& a, int nf = 10)
{
    return (nf - 1) * a.stride + a.nx + a.rp * ((int64_t)a.ny * a.nz - 1);
}

// cell_at of the cell with linear index c.
__device__ __forceinline__ int64_t cell_of(const EnvArgs& a, int64_t c)
{
    return a.rp == a.nx ? c : c % a.nx + a.rp * (c / a.nx);
}

__device__ __forceinline__ bool fixup_owns(const EnvArgs& a, int x, int y, int z)
{
    if (!a.cut_bits) return false;
    const int rx = x - a.cut_origin[0], ry = y - a.cut_origin[1], rz = z - a.cut_origin[2];
    if (rx < 0 || ry < 0 || rz < 0 || rx >= a.cut_extent[0] || ry >= a.cut_extent[1] || rz >= a.cut_extent[2])
        return false;
    const int t = rx + a.cut_extent[0] * (ry + a.cut_extent[1] * rz);
    return (a.cut_bits[t >> 6] >> (t & 63)) & 1ull;
}

__device__ __forceinline__ bool in_box_halo(const EnvArgs& a, int x, int y, int z)
{
    return x >= a.box[0] - 1 && x <= a.box[1] && y >= a.box[2] - 1 && y <= a.box[3] &&
           z >= a.box[4] - 1 && z <= a.box[5];
}

}  // namespace hl
