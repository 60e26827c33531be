// Triangle-mesh voxelizer for D3Q27 bodies: a cell centre is inside a closed
// mesh when a ray from it along +x crosses the mesh an odd number of times.
// One thread per (y, z) column of the box: it intersects its ray with every
// triangle once (projected edge functions; a triangle is crossed when the ray
// lies strictly inside its yz projection, the crossing x from the barycentric
// weights), then sweeps the column's cells, toggling a parity bit per solid
// for each crossing ahead of the cell; mask = lowest solid with odd parity
// + 1. Adjacent triangles share edges computed from the same two products,
// so a ray near a shared edge is inside exactly one of them; the ray is
// offset by (1.2e-4, 2.3e-4) off the lattice planes so that it does not pass
// exactly through vertices or edges of meshes authored on grid coordinates.
//
// The reference rasterises nothing: it ray-casts each link against a Warp BVH
// of the posed mesh (get_cutcell_multi_3d, envs/lbm3d/lbm_func_3d.py:254-316).
// Here the posed mesh becomes a mask and the wall rule runs on mask links, as
// for every other body.
#pragma clang fp contract(off)

#include "kernels.h"
#include "lbm_device.h"

namespace hl {

namespace {

constexpr int kMaxCrossings = 64;
constexpr float kRayDy = 1.2345679e-4f, kRayDz = 2.3456789e-4f;

__global__ void __launch_bounds__(64) k_voxelize_mesh(uint8_t* mask, int nx, int ny, const float* tris,
                                                       const int32_t* tsolid, int ntri, int bx0, int by0, int bz0,
                                                       int bw, int bh, int bd, int merge, int32_t* overflow)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= bh * bd) return;
    const int y = by0 + t % bh, z = bz0 + t / bh;
    const float py = (float)y + kRayDy, pz = (float)z + kRayDz;
    float xs[kMaxCrossings];
    int ss[kMaxCrossings];
    int nc = 0;
    bool full = false;
    for (int k = 0; k < ntri; ++k) {
        const float* T = tris + 9 * (int64_t)k;
        const float ay = T[1] - py, az = T[2] - pz, by = T[4] - py, bz = T[5] - pz, cy = T[7] - py, cz = T[8] - pz;
        const float w0 = by * cz - bz * cy;  // edge b -> c, weight of a
        const float w1 = cy * az - cz * ay;  // edge c -> a, weight of b
        const float w2 = ay * bz - az * by;  // edge a -> b, weight of c
        const bool in = (w0 > 0.0f && w1 > 0.0f && w2 > 0.0f) || (w0 < 0.0f && w1 < 0.0f && w2 < 0.0f);
        if (!in) continue;
        const float xc = (w0 * T[0] + w1 * T[3] + w2 * T[6]) / (w0 + w1 + w2);
        if (nc == kMaxCrossings) {
            full = true;
            continue;
        }
        xs[nc] = xc;
        ss[nc] = tsolid[k];
        ++nc;
    }
    if (full && overflow) *overflow = 1;
    for (int i = 0; i < bw; ++i) {
        const int x = bx0 + i;
        const float px = (float)x;
        uint64_t bits = 0;
        for (int j = 0; j < nc; ++j)
            if (xs[j] > px && ss[j] >= 0 && ss[j] < 64) bits ^= 1ull << ss[j];
        const uint8_t v = bits ? (uint8_t)(__builtin_ctzll(bits) + 1) : (uint8_t)0;
        uint8_t* m = mask + ((int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z));
        if (!merge) *m = v;
        else if (v && *m == 0) *m = v;
    }
}

}  // namespace

hipError_t voxelize_mesh(const EnvArgs& a, uint8_t* mask, const float* tris, const int32_t* tsolid, int ntri,
                         const int32_t box[6], int merge, int32_t* overflow, hipStream_t s)
{
    const int bw = box[1] - box[0], bh = box[3] - box[2], bd = box[5] - box[4];
    if (bw <= 0 || bh <= 0 || bd <= 0) return hipSuccess;
    const int cols = bh * bd;
    hipLaunchKernelGGL(k_voxelize_mesh, (cols + 63) / 64, 64, 0, s, mask, a.nx, a.ny, tris, tsolid, ntri, box[0],
                       box[2], box[4], bw, bh, bd, merge, overflow);
    return hipGetLastError();
}

}  // namespace hl
