// Coupling of the lattice with a rigid-body engine, all on the device:
//   * body poses of nworld worlds -> each world's pose table
//     (convert_and_update_solid_batch_2d, envs/lbm/lbm_fluid_env_func.py:72-105;
//      convert_and_update_solid_batch_3d, envs/lbm3d/lbm_fluid_env_3d_func.py:42-99);
//   * per-solid FP64 force rows -> engine force / torque buffers and the
//     spatial-force buffer (extract_forces_torques_batch, lbm_fluid_env_func.py:51-69;
//      extract_forces_torques_batch_3d / _physical_3d, lbm_fluid_env_3d_func.py:102-185;
//      fill_xfrc_kernel / fill_xfrc_3d_kernel, lbm_fluid_env_func.py:28-48,
//      lbm_fluid_env_3d_func.py:188-214);
//   * a free-rigid-body integrator standing in for the engine (the reference
//     steps MuJoCo-Warp, which does not run here), and the posed capsules of
//     each world for the voxelizer.
// One thread per (world, body); these are launch-latency sized kernels.
#pragma clang fp contract(off)

#include "kernels.h"
#include "lbm_device.h"

namespace hl {

namespace {

// Rotation matrix (row-major) of a (w, x, y, z) quaternion, normalised first.
__device__ void quat_matrix(float w, float x, float y, float z, float* R)
{
    const float n = sqrtf(w * w + x * x + y * y + z * z);
    w /= n; x /= n; y /= n; z /= n;
    R[0] = 1.0f - 2.0f * (y * y + z * z); R[1] = 2.0f * (x * y - w * z);        R[2] = 2.0f * (x * z + w * y);
    R[3] = 2.0f * (x * y + w * z);        R[4] = 1.0f - 2.0f * (x * x + z * z); R[5] = 2.0f * (y * z - w * x);
    R[6] = 2.0f * (x * z - w * y);        R[7] = 2.0f * (y * z + w * x);        R[8] = 1.0f - 2.0f * (x * x + y * y);
}

struct PoseArgs {
    int lattice, nworld, nbody, nsolid;
    const float* pos;       // engine frame, body t at pos + pstride * t
    const float* quat;      // (w, x, y, z), body t at quat + qstride * t
    int pstride, qstride;
    const int32_t* solid;   // [nbody] solid index in the tables
    const float* eorigin;   // [nbody][3] engine origin of the mapping
    const float* lorigin;   // [nbody][3] lattice origin of the mapping
    const float* scale;     // [nbody] lattice cells per engine unit
    float* const* tables;   // [nworld] device pose tables
    int32_t* init;          // [nworld][nsolid]: 0 until a pose has been written
};

__global__ void __launch_bounds__(64) k_couple_poses(PoseArgs p)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= p.nworld * p.nbody) return;
    const int w = t / p.nbody, b = t % p.nbody;
    const int sid = p.solid[b];
    if (sid < 0 || sid >= p.nsolid) return;  // not a row of the tables
    const float s = p.scale[b];
    const float* P = p.pos + (int64_t)p.pstride * t;
    const float* Q = p.quat + (int64_t)p.qstride * t;
    const float lx = (P[0] - p.eorigin[3 * b + 0]) * s + p.lorigin[3 * b + 0];
    const float ly = (P[1] - p.eorigin[3 * b + 1]) * s + p.lorigin[3 * b + 1];
    int32_t& flag = p.init[(int64_t)w * p.nsolid + sid];
    if (p.lattice == HL_D2Q9) {
        // planar angle of the quaternion (lbm_fluid_env_func.py:101-105)
        const float qw = Q[0], qx = Q[1], qy = Q[2], qz = Q[3];
        const float ang = atan2f(2.0f * (qw * qz + qx * qy), 1.0f - 2.0f * (qy * qy + qz * qz));
        float* T = p.tables[w] + 8 * sid;
        const float now[4] = {lx, ly, cosf(ang), sinf(ang)};
        for (int k = 0; k < 4; ++k) T[4 + k] = flag ? T[k] : now[k];
        for (int k = 0; k < 4; ++k) T[k] = now[k];
    } else {
        const float lz = (P[2] - p.eorigin[3 * b + 2]) * s + p.lorigin[3 * b + 2];
        float* T = p.tables[w] + 24 * sid;
        float now[12] = {lx, ly, lz};
        quat_matrix(Q[0], Q[1], Q[2], Q[3], now + 3);
        // the current transform becomes the last one (lbm_fluid_env_3d_func.py:87-99)
        for (int k = 0; k < 12; ++k) T[12 + k] = flag ? T[k] : now[k];
        for (int k = 0; k < 12; ++k) T[k] = now[k];
    }
    flag = 1;
}

struct ForceArgs {
    int lattice, mode, nworld, nbody, nsolid;
    const double* rows;     // world w, solid k: rows + w * wstride + 6 k: F(3), tau(3)
    int64_t wstride;
    const int32_t* solid;
    const float* scale;
    float fconv, tconv;
    float* fout;            // [nworld][nbody][3]
    float* tout;            // [nworld][nbody][3]
    float* xfrc;            // [nworld][xbodies][6] or null
    int xbodies;
    const int32_t* body;    // [nbody] engine body index in xfrc
};

__global__ void __launch_bounds__(64) k_couple_forces(ForceArgs f)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= f.nworld * f.nbody) return;
    const int w = t / f.nbody, b = t % f.nbody;
    const int sid = f.solid[b];
    if (sid < 0 || sid >= f.nsolid) return;  // not a row of the force table
    const double* r = f.rows + (int64_t)w * f.wstride + 6 * sid;
    float F[3] = {(float)r[0], (float)r[1], (float)r[2]};
    float T[3] = {(float)r[3], (float)r[4], (float)r[5]};
    if (f.lattice == HL_D2Q9) { F[2] = 0.0f; T[0] = 0.0f; T[1] = 0.0f; }  // planar loads
    if (f.mode == HL_COUPLE_SCALED) {
        const float s = f.scale[b];
        const float ts = f.lattice == HL_D2Q9 ? s : s * s;  // the 2D torque is divided by scale once
        for (int k = 0; k < 3; ++k) { F[k] = F[k] / s; T[k] = T[k] / ts; }
        if (f.lattice == HL_D2Q9) { F[2] = 0.0f; T[0] = 0.0f; T[1] = 0.0f; }
    } else {
        for (int k = 0; k < 3; ++k) { F[k] = F[k] * f.fconv; T[k] = T[k] * f.tconv; }
    }
    for (int k = 0; k < 3; ++k) {
        f.fout[3 * (int64_t)t + k] = F[k];
        f.tout[3 * (int64_t)t + k] = T[k];
    }
    if (f.xfrc && f.body[b] >= 0 && f.body[b] < f.xbodies) {
        float* X = f.xfrc + 6 * ((int64_t)w * f.xbodies + f.body[b]);
        for (int k = 0; k < 3; ++k) { X[k] = F[k]; X[3 + k] = T[k]; }
    }
}

// Free rigid bodies, semi-implicit Euler: state [p(3), q(4) (w,x,y,z),
// v(3), omega(3) world frame]; params [mass, Ixx, Iyy, Izz] (principal, body
// frame). v += dt (F/m + g); p += dt v; in the body frame
// I domega/dt = tau - omega x I omega; q += dt/2 (0, omega) q, renormalised.
__global__ void __launch_bounds__(64) k_free_bodies(int nworld, int nbody, float* state, const float* params,
                                                     const float* force, const float* torque, float gx, float gy,
                                                     float gz, float dt)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= nworld * nbody) return;
    const int b = t % nbody;
    float* S = state + 13 * (int64_t)t;
    const float* M = params + 4 * b;
    const float* F = force + 3 * (int64_t)t;
    const float* Tq = torque + 3 * (int64_t)t;
    const float g[3] = {gx, gy, gz};
    for (int k = 0; k < 3; ++k) {
        S[7 + k] = S[7 + k] + dt * (F[k] / M[0] + g[k]);
        S[k] = S[k] + dt * S[7 + k];
    }
    float R[9];
    quat_matrix(S[3], S[4], S[5], S[6], R);
    // body-frame angular velocity and torque
    float wb[3], tb[3];
    for (int i = 0; i < 3; ++i) {
        wb[i] = R[i] * S[10] + R[3 + i] * S[11] + R[6 + i] * S[12];
        tb[i] = R[i] * Tq[0] + R[3 + i] * Tq[1] + R[6 + i] * Tq[2];
    }
    const float Iw[3] = {M[1] * wb[0], M[2] * wb[1], M[3] * wb[2]};
    const float gyro[3] = {wb[1] * Iw[2] - wb[2] * Iw[1], wb[2] * Iw[0] - wb[0] * Iw[2], wb[0] * Iw[1] - wb[1] * Iw[0]};
    for (int i = 0; i < 3; ++i) wb[i] = wb[i] + dt * (tb[i] - gyro[i]) / M[1 + i];
    for (int i = 0; i < 3; ++i) S[10 + i] = R[3 * i] * wb[0] + R[3 * i + 1] * wb[1] + R[3 * i + 2] * wb[2];
    const float qw = S[3], qx = S[4], qy = S[5], qz = S[6];
    const float ox = S[10], oy = S[11], oz = S[12], h = 0.5f * dt;
    float nw = qw + h * (-ox * qx - oy * qy - oz * qz);
    float nx = qx + h * (ox * qw + oy * qz - oz * qy);
    float ny = qy + h * (oy * qw + oz * qx - ox * qz);
    float nz = qz + h * (oz * qw + ox * qy - oy * qx);
    const float n = sqrtf(nw * nw + nx * nx + ny * ny + nz * nz);
    S[3] = nw / n; S[4] = nx / n; S[5] = ny / n; S[6] = nz / n;
}

// Posed capsules of every world for the voxelizer: body-frame segments
// shapes [nbody][7] (a(3), b(3), radius, engine units) -> caps
// [nworld][nbody][8] in lattice coordinates, solid id in slot 7.
__global__ void __launch_bounds__(64) k_body_capsules(int nworld, int nbody, const float* pos, int pstride,
                                                      const float* quat, int qstride, const float* shapes,
                                                      const int32_t* solid, const float* eorigin,
                                                      const float* lorigin, const float* scale, float* caps)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= nworld * nbody) return;
    const int b = t % nbody;
    const float* P = pos + (int64_t)pstride * t;
    const float* Q = quat + (int64_t)qstride * t;
    const float* Sh = shapes + 7 * b;
    const float s = scale[b];
    float R[9];
    quat_matrix(Q[0], Q[1], Q[2], Q[3], R);
    float* C = caps + 8 * (int64_t)t;
    for (int e = 0; e < 2; ++e)
        for (int i = 0; i < 3; ++i) {
            const float wld = P[i] + (R[3 * i] * Sh[3 * e] + R[3 * i + 1] * Sh[3 * e + 1] + R[3 * i + 2] * Sh[3 * e + 2]);
            C[3 * e + i] = (wld - eorigin[3 * b + i]) * s + lorigin[3 * b + i];
        }
    C[6] = Sh[6] * s;
    C[7] = (float)solid[b];
}

}  // namespace

hipError_t couple_poses(int lattice, int nworld, int nbody, int nsolid, const float* pos, int pstride,
                        const float* quat, int qstride, const int32_t* solid, const float* eorigin,
                        const float* lorigin, const float* scale, float* const* tables, int32_t* init, hipStream_t s)
{
    const PoseArgs p{lattice, nworld, nbody, nsolid, pos, quat, pstride, qstride, solid, eorigin, lorigin, scale,
                     tables, init};
    hipLaunchKernelGGL(k_couple_poses, (nworld * nbody + 63) / 64, 64, 0, s, p);
    return hipGetLastError();
}

hipError_t couple_forces(int lattice, int mode, int nworld, int nbody, int nsolid, const double* rows,
                         int64_t wstride, const int32_t* solid, const float* scale, float fconv, float tconv,
                         float* fout, float* tout, float* xfrc, int xbodies, const int32_t* body, hipStream_t s)
{
    const ForceArgs f{lattice, mode, nworld, nbody, nsolid, rows, wstride, solid, scale, fconv, tconv,
                      fout, tout, xfrc, xbodies, body};
    hipLaunchKernelGGL(k_couple_forces, (nworld * nbody + 63) / 64, 64, 0, s, f);
    return hipGetLastError();
}

hipError_t free_bodies_step(int nworld, int nbody, float* state, const float* params, const float* force,
                            const float* torque, const float g[3], float dt, hipStream_t s)
{
    hipLaunchKernelGGL(k_free_bodies, (nworld * nbody + 63) / 64, 64, 0, s, nworld, nbody, state, params, force,
                       torque, g[0], g[1], g[2], dt);
    return hipGetLastError();
}

hipError_t body_capsules(int nworld, int nbody, const float* pos, int pstride, const float* quat, int qstride,
                         const float* shapes, const int32_t* solid, const float* eorigin, const float* lorigin,
                         const float* scale, float* caps, hipStream_t s)
{
    hipLaunchKernelGGL(k_body_capsules, (nworld * nbody + 63) / 64, 64, 0, s, nworld, nbody, pos, pstride, quat,
                       qstride, shapes, solid, eorigin, lorigin, scale, caps);
    return hipGetLastError();
}

}  // namespace hl
