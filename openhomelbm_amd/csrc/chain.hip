// Planar articulated chains stepped on the device (hl_chain_*): one 64-lane
// workgroup per world, FP64 throughout.
//
// Each stage of the integrator evaluates the joint accelerations of the
// world's state in five phases separated by workgroup barriers:
//   1. lanes k < ndof: cos / sin of each hinge's angle;
//   2. lane 0: forward kinematics of the tree in order (body offset, then the
//      body's joints: slides along an axis of the frame before them, hinges
//      about an anchor), each frame origin's velocity, each dof's world
//      anchor / axis and the angular velocity of the frame before it, and
//      the centres of mass with their velocities;
//   3. lanes over (body, dof): the centre-of-mass Jacobian columns; lanes
//      b < nbody: the centre-of-mass acceleration at zero joint acceleration;
//   4. lanes i < ndof: row i of the joint-space inertia (plus armature, plus
//      dt * damping for implicit Euler) and the force: passive + motors +
//      J^T xfrc + qfrc_applied - bias (gravity included);
//   5. lane 0: Cholesky solve of M qacc = f.
// The state update of the integrator (RK4 or semi-implicit Euler) runs on
// lanes i < ndof. Everything a stage reads lives in LDS (about 9 KB).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <new>
#include <string>

#include "homelbm.h"
#include "kernels.h"

namespace hl {
namespace {

constexpr int NB = HL_CHAIN_MAX_BODIES, NV = HL_CHAIN_MAX_DOFS;

struct ChainShared {
    double q[NV], v[NV];          // the state this stage evaluates
    double cq[NV], sq[NV];        // cos / sin of hinge angles
    double P[NB][2], CS[NB][2], V[NB][2], W[NB];
    double C[NB][2], Cd[NB][2], a[NB][2];
    double A[NV][2], Ad[NV][2], U[NV][2], wb[NV];
    double J[NB][NV][2];
    double M[NV][NV];
    double f[NV], acc[NV];
    double applied[NV];           // qfrc_applied of this step
    double xf[NB][3];             // planar loads fx, fy, tz
    double u[NV];                 // motor force per dof
    uint32_t anc[NB];             // dofs that move each body
};

__device__ __forceinline__ void perp(const double* v, double* o) { o[0] = -v[1]; o[1] = v[0]; }

// Joint accelerations of the state in S.q / S.v into S.acc. first: also set
// S.applied = factor * J^T xfrc at this configuration.
__device__ void chain_accel(const hl_chain_model& m, ChainShared& S, bool first, double implicit_dt)
{
    const int t = threadIdx.x;
    const int nb = m.nbody, nv = m.ndof;
    if (t < nv && m.dof_type[t] == HL_CHAIN_HINGE) {
        const double ang = m.dof_sign[t] * S.q[t];
        S.cq[t] = cos(ang);
        S.sq[t] = sin(ang);
    }
    __syncthreads();
    if (t == 0) {
        for (int b = 0; b < nb; ++b) {
            const int pa = m.parent[b];
            double p[2] = {0.0, 0.0}, v[2] = {0.0, 0.0}, c = 1.0, s = 0.0, w = 0.0;
            if (pa >= 0) {
                p[0] = S.P[pa][0]; p[1] = S.P[pa][1];
                v[0] = S.V[pa][0]; v[1] = S.V[pa][1];
                c = S.CS[pa][0]; s = S.CS[pa][1];
                w = S.W[pa];
            }
            if (!m.free_body[b]) {
                const double off[2] = {c * m.pos[b][0] - s * m.pos[b][1], s * m.pos[b][0] + c * m.pos[b][1]};
                p[0] = p[0] + off[0]; p[1] = p[1] + off[1];
                v[0] = v[0] + w * -off[1]; v[1] = v[1] + w * off[0];
                const double ca = cos(m.angle[b]), sa = sin(m.angle[b]);
                const double c2 = c * ca - s * sa, s2 = s * ca + c * sa;
                c = c2; s = s2;
            }
            for (int k = 0; k < nv; ++k) {
                if (m.dof_body[k] != b) continue;
                S.wb[k] = w;
                if (m.dof_type[k] == HL_CHAIN_HINGE) {
                    const double* an = m.dof_anchor[k];
                    const double a0[2] = {c * an[0] - s * an[1], s * an[0] + c * an[1]};
                    S.A[k][0] = p[0] + a0[0]; S.A[k][1] = p[1] + a0[1];
                    S.Ad[k][0] = v[0] + w * -a0[1]; S.Ad[k][1] = v[1] + w * a0[0];
                    const double c2 = c * S.cq[k] - s * S.sq[k], s2 = s * S.cq[k] + c * S.sq[k];
                    c = c2; s = s2;
                    w = w + m.dof_sign[k] * S.v[k];
                    const double a1[2] = {c * an[0] - s * an[1], s * an[0] + c * an[1]};
                    p[0] = S.A[k][0] - a1[0]; p[1] = S.A[k][1] - a1[1];
                    v[0] = S.Ad[k][0] - w * -a1[1]; v[1] = S.Ad[k][1] - w * a1[0];
                } else {
                    const double* ax = m.dof_axis[k];
                    S.U[k][0] = c * ax[0] - s * ax[1]; S.U[k][1] = s * ax[0] + c * ax[1];
                    const double d[2] = {S.U[k][0] * S.q[k], S.U[k][1] * S.q[k]};
                    p[0] = p[0] + d[0]; p[1] = p[1] + d[1];
                    v[0] = v[0] + w * -d[1] + S.U[k][0] * S.v[k];
                    v[1] = v[1] + w * d[0] + S.U[k][1] * S.v[k];
                }
            }
            S.P[b][0] = p[0]; S.P[b][1] = p[1];
            S.CS[b][0] = c; S.CS[b][1] = s;
            S.V[b][0] = v[0]; S.V[b][1] = v[1];
            S.W[b] = w;
            const double l[2] = {c * m.com[b][0] - s * m.com[b][1], s * m.com[b][0] + c * m.com[b][1]};
            S.C[b][0] = p[0] + l[0]; S.C[b][1] = p[1] + l[1];
            S.Cd[b][0] = v[0] + w * -l[1]; S.Cd[b][1] = v[1] + w * l[0];
        }
    }
    __syncthreads();
    for (int e = t; e < nb * nv; e += blockDim.x) {
        const int b = e / nv, k = e % nv;
        double j0 = 0.0, j1 = 0.0;
        if ((S.anc[b] >> k) & 1u) {
            if (m.dof_type[k] == HL_CHAIN_HINGE) {
                const double r[2] = {S.C[b][0] - S.A[k][0], S.C[b][1] - S.A[k][1]};
                j0 = m.dof_sign[k] * -r[1];
                j1 = m.dof_sign[k] * r[0];
            } else {
                j0 = S.U[k][0];
                j1 = S.U[k][1];
            }
        }
        S.J[b][k][0] = j0;
        S.J[b][k][1] = j1;
    }
    if (t < nb) {
        double a0 = 0.0, a1 = 0.0;
        for (int k = 0; k < nv; ++k) {
            if (!((S.anc[t] >> k) & 1u)) continue;
            double g[2];
            if (m.dof_type[k] == HL_CHAIN_HINGE) {
                const double r[2] = {S.Cd[t][0] - S.Ad[k][0], S.Cd[t][1] - S.Ad[k][1]};
                perp(r, g);
                g[0] = m.dof_sign[k] * g[0];
                g[1] = m.dof_sign[k] * g[1];
            } else {
                perp(S.U[k], g);
                g[0] = S.wb[k] * g[0];
                g[1] = S.wb[k] * g[1];
            }
            a0 = a0 + S.v[k] * g[0];
            a1 = a1 + S.v[k] * g[1];
        }
        S.a[t][0] = a0;
        S.a[t][1] = a1;
    }
    __syncthreads();
    if (t < nv) {
        const int i = t;
        for (int j = 0; j < nv; ++j) S.M[i][j] = (i == j) ? m.armature[i] : 0.0;
        double bias = 0.0, xg = 0.0;
        for (int b = 0; b < nb; ++b) {
            if (!((S.anc[b] >> i) & 1u)) continue;
            const double jp_i = m.dof_type[i] == HL_CHAIN_HINGE ? m.dof_sign[i] : 0.0;
            for (int j = 0; j < nv; ++j) {
                const double jp_j = ((S.anc[b] >> j) & 1u) && m.dof_type[j] == HL_CHAIN_HINGE ? m.dof_sign[j] : 0.0;
                S.M[i][j] = S.M[i][j] + m.mass[b] * (S.J[b][i][0] * S.J[b][j][0] + S.J[b][i][1] * S.J[b][j][1]) +
                            m.inertia[b] * (jp_i * jp_j);
            }
            bias = bias + m.mass[b] * (S.J[b][i][0] * (S.a[b][0] - m.gravity[0]) +
                                       S.J[b][i][1] * (S.a[b][1] - m.gravity[1]));
            xg = xg + (S.J[b][i][0] * S.xf[b][0] + S.J[b][i][1] * S.xf[b][1]) + jp_i * S.xf[b][2];
        }
        if (first) S.applied[i] = m.applied_factor * xg;
        const double passive = -m.damping[i] * S.v[i] - m.stiffness[i] * (S.q[i] - m.springref[i]);
        S.f[i] = passive + S.u[i] + xg + S.applied[i] - bias;
        if (implicit_dt != 0.0) S.M[i][i] = S.M[i][i] + implicit_dt * m.damping[i];
    }
    __syncthreads();
    if (t == 0) {
        // Cholesky M = L L^T in place (lower triangle), then two triangular solves
        for (int k = 0; k < nv; ++k) {
            double d = S.M[k][k];
            for (int j = 0; j < k; ++j) d = d - S.M[k][j] * S.M[k][j];
            d = sqrt(d);
            S.M[k][k] = d;
            for (int i = k + 1; i < nv; ++i) {
                double x = S.M[i][k];
                for (int j = 0; j < k; ++j) x = x - S.M[i][j] * S.M[k][j];
                S.M[i][k] = x / d;
            }
        }
        for (int i = 0; i < nv; ++i) {
            double x = S.f[i];
            for (int j = 0; j < i; ++j) x = x - S.M[i][j] * S.acc[j];
            S.acc[i] = x / S.M[i][i];
        }
        for (int i = nv - 1; i >= 0; --i) {
            double x = S.acc[i];
            for (int j = i + 1; j < nv; ++j) x = x - S.M[j][i] * S.acc[j];
            S.acc[i] = x / S.M[i][i];
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(64) k_chain_step(const hl_chain_model* __restrict__ mp, int nworld, double* qpos,
                                                   double* qvel, const float* ctrl, const float* xfrc, int nsteps,
                                                   double* applied_out, float* xpos, float* xipos, float* xquat)
{
    __shared__ hl_chain_model m;
    __shared__ ChainShared S;
    const int w = blockIdx.x, t = threadIdx.x;
    static_assert(sizeof(hl_chain_model) % 4 == 0, "model copy");
    for (int i = t; i < (int)(sizeof(hl_chain_model) / 4); i += blockDim.x)
        reinterpret_cast<uint32_t*>(&m)[i] = reinterpret_cast<const uint32_t*>(mp)[i];
    __syncthreads();
    const int nb = m.nbody, nv = m.ndof;
    if (t < nb) {
        uint32_t mask = 0;
        for (int b = t; b >= 0; b = m.parent[b])
            for (int k = 0; k < nv; ++k)
                if (m.dof_body[k] == b) mask |= 1u << k;
        S.anc[t] = mask;
        const float* X = xfrc ? xfrc + ((int64_t)w * nb + t) * 6 : nullptr;
        S.xf[t][0] = X ? (double)X[0] : 0.0;
        S.xf[t][1] = X ? (double)X[1] : 0.0;
        S.xf[t][2] = X ? (double)X[5] : 0.0;
    }
    if (t < nv) {
        double u = 0.0;
        for (int a = 0; a < m.nact; ++a) {
            if (m.act_dof[a] != t) continue;
            double c = ctrl ? (double)ctrl[(int64_t)w * m.nact + a] : 0.0;
            if (m.ctrl_limited[a]) c = fmin(fmax(c, m.ctrl_lo[a]), m.ctrl_hi[a]);
            u = u + m.gear[a] * c;
        }
        S.u[t] = u;
    }
    double q0 = 0.0, v0 = 0.0;
    if (t < nv) {
        S.applied[t] = 0.0;
        q0 = qpos[(int64_t)w * nv + t];
        v0 = qvel[(int64_t)w * nv + t];
    }
    const double h = m.dt;
    for (int n = 0; n < nsteps; ++n) {
        if (t < nv) {
            S.q[t] = q0;
            S.v[t] = v0;
        }
        __syncthreads();
        if (m.integrator == HL_CHAIN_RK4) {
            // k1 at the start; X_i = X0 + c_i h k_{i-1}; weights 1/6, 1/3, 1/3, 1/6
            double kq[4], kv[4];
            chain_accel(m, S, true, 0.0);
            kq[0] = v0;
            kv[0] = t < nv ? S.acc[t] : 0.0;
            const double cst[3] = {0.5, 0.5, 1.0};
            for (int st = 1; st < 4; ++st) {
                if (t < nv) {
                    S.q[t] = q0 + cst[st - 1] * h * kq[st - 1];
                    S.v[t] = v0 + cst[st - 1] * h * kv[st - 1];
                }
                kq[st] = t < nv ? v0 + cst[st - 1] * h * kv[st - 1] : 0.0;
                __syncthreads();
                chain_accel(m, S, false, 0.0);
                kv[st] = t < nv ? S.acc[t] : 0.0;
            }
            q0 = q0 + h * (kq[0] / 6.0 + kq[1] / 3.0 + kq[2] / 3.0 + kq[3] / 6.0);
            v0 = v0 + h * (kv[0] / 6.0 + kv[1] / 3.0 + kv[2] / 3.0 + kv[3] / 6.0);
        } else {
            chain_accel(m, S, true, h);
            v0 = v0 + h * (t < nv ? S.acc[t] : 0.0);
            q0 = q0 + h * v0;
        }
    }
    if (t < nv) {
        qpos[(int64_t)w * nv + t] = q0;
        qvel[(int64_t)w * nv + t] = v0;
        if (applied_out) applied_out[(int64_t)w * nv + t] = S.applied[t];
        S.q[t] = q0;
        S.v[t] = 0.0;
    }
    __syncthreads();
    if (!(xpos || xipos || xquat)) return;
    chain_accel(m, S, false, 0.0);  // kinematics of the final configuration
    if (t < nb) {
        const int64_t o = (int64_t)w * nb + t;
        if (xpos) {
            xpos[o * 3 + 0] = (float)S.P[t][0];
            xpos[o * 3 + 1] = (float)S.P[t][1];
            xpos[o * 3 + 2] = (float)m.posz[t];
        }
        if (xipos) {
            xipos[o * 3 + 0] = (float)S.C[t][0];
            xipos[o * 3 + 1] = (float)S.C[t][1];
            xipos[o * 3 + 2] = (float)m.comz[t];
        }
        if (xquat) {
            const double phi = atan2(S.CS[t][1], S.CS[t][0]);
            xquat[o * 4 + 0] = (float)cos(0.5 * phi);
            xquat[o * 4 + 1] = 0.0f;
            xquat[o * 4 + 2] = 0.0f;
            xquat[o * 4 + 3] = (float)sin(0.5 * phi);
        }
    }
}

}  // namespace
}  // namespace hl

struct hl_chain {
    hl_chain_model host;
    hl_chain_model* dev = nullptr;
};

extern "C" {

int32_t hl_chain_create(const hl_chain_model* m, hl_chain** out)
{
    using hl::report;
    if (!m || !out) return report(HL_E_INVALID, "hl_chain_create: null argument");
    *out = nullptr;
    if (m->nbody < 1 || m->nbody > HL_CHAIN_MAX_BODIES || m->ndof < 1 || m->ndof > HL_CHAIN_MAX_DOFS ||
        m->nact < 0 || m->nact > HL_CHAIN_MAX_ACTS)
        return report(HL_E_INVALID, "hl_chain_create: 1..16 bodies, 1..16 dofs and 0..16 actuators");
    if (m->integrator != HL_CHAIN_RK4 && m->integrator != HL_CHAIN_EULER)
        return report(HL_E_INVALID, "hl_chain_create: integrator is HL_CHAIN_RK4 or HL_CHAIN_EULER");
    if (!(m->dt > 0.0)) return report(HL_E_INVALID, "hl_chain_create: dt must be positive");
    for (int b = 0; b < m->nbody; ++b) {
        if (m->parent[b] < -1 || m->parent[b] >= b)
            return report(HL_E_INVALID, ("hl_chain_create: body " + std::to_string(b) +
                                         " must come after its parent (tree order)").c_str());
        if (!(m->mass[b] >= 0.0) || !(m->inertia[b] >= 0.0))
            return report(HL_E_INVALID, "hl_chain_create: negative mass or inertia");
    }
    int prev = -1;
    for (int k = 0; k < m->ndof; ++k) {
        const int b = m->dof_body[k];
        if (b < 0 || b >= m->nbody || b < prev)
            return report(HL_E_INVALID, "hl_chain_create: dofs grouped by body, bodies in order");
        prev = b;
        if (m->dof_type[k] != HL_CHAIN_SLIDE && m->dof_type[k] != HL_CHAIN_HINGE)
            return report(HL_E_INVALID, "hl_chain_create: dof type is HL_CHAIN_SLIDE or HL_CHAIN_HINGE");
        if (m->dof_type[k] == HL_CHAIN_HINGE && m->dof_sign[k] != 1.0 && m->dof_sign[k] != -1.0)
            return report(HL_E_INVALID, "hl_chain_create: hinge sign is +1 or -1");
    }
    for (int a = 0; a < m->nact; ++a)
        if (m->act_dof[a] < 0 || m->act_dof[a] >= m->ndof)
            return report(HL_E_INVALID, "hl_chain_create: actuator dof out of range");
    hl_chain* c = new (std::nothrow) hl_chain;
    if (!c) return report(HL_E_NOMEM, "hl_chain_create: out of host memory");
    c->host = *m;
    hipError_t e = hipMalloc(&c->dev, sizeof(hl_chain_model));
    if (e == hipSuccess) e = hipMemcpy(c->dev, m, sizeof(hl_chain_model), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (c->dev) (void)hipFree(c->dev);
        delete c;
        return report(HL_E_HIP, (std::string("hl_chain_create: ") + hipGetErrorString(e)).c_str());
    }
    *out = c;
    return HL_OK;
}

int32_t hl_chain_destroy(hl_chain* c)
{
    if (!c) return HL_OK;
    hipError_t e = hipFree(c->dev);
    delete c;
    if (e != hipSuccess) return hl::report(HL_E_HIP, (std::string("hl_chain_destroy: ") + hipGetErrorString(e)).c_str());
    return HL_OK;
}

int32_t hl_chain_step(hl_chain* c, int32_t nworld, double* qpos, double* qvel, const float* ctrl, const float* xfrc,
                      int32_t nsteps, double* qfrc_applied, float* xpos, float* xipos, float* xquat, void* stream)
{
    using hl::report;
    if (!c || !qpos || !qvel) return report(HL_E_INVALID, "hl_chain_step: null chain or state");
    if (nworld < 1 || nsteps < 0) return report(HL_E_INVALID, "hl_chain_step: nworld >= 1 and nsteps >= 0");
    hipLaunchKernelGGL(hl::k_chain_step, dim3(nworld), dim3(64), 0, static_cast<hipStream_t>(stream), c->dev, nworld,
                       qpos, qvel, ctrl, xfrc, nsteps, qfrc_applied, xpos, xipos, xquat);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return report(HL_E_HIP, (std::string("hl_chain_step: ") + hipGetErrorString(e)).c_str());
    return HL_OK;
}

}  // extern "C"
