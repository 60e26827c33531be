// Render and diagnostic images of the present state: velocity magnitude, RGB
// velocity, vorticity (central differences inside, one-sided at the faces, over
// the grid length 1/nx), with the bodies overlaid, on the 2D lattice, on 3D
// slices and as 3D projections along an axis.
//
// Reference kernels (same formulas and operation order; the squares u**2.0
// are products):
//   get_u_img, get_velocity_rgb, get_vorticity_img,
//   get_vorticity_with_solid_img, get_solid_boundary_img   lbm_func.py:163-319, 804-864
//   get_u_img_xy/yz_3d, get_vorticity_xy/yz_3d(_with_solid) lbm_func_3d.py:753-988
//   get_u_projection_{topdown,max_topdown,side,front}_3d,
//   get_vorticity_projection_{topdown,side,front}_3d        lbm_func_3d.py:990-1313
// Bodies are the voxel mask here: a cell is solid where the mask is set, or
// (with thickness t > 0) where a masked cell lies within distance t of it,
// in place of the reference's polygon / mesh distance queries; a projected
// ray is solid where it meets any masked cell, in place of its mesh ray hit.
#pragma clang fp contract(off)

#include "kernels.h"
#include "lbm_device.h"

namespace hl {

namespace {

constexpr float kSolidMark = 1000.0f;  // the reference's solid sentinel in vorticity images
constexpr int kMaxThickness = 4;

struct RenderArgs {
    float gl;         // grid length 1/nx, rounded to FP32 as the reference's struct field
    float thickness;
    float vmax;       // RGB scale
};

// Velocity component comp of the cell with linear index c.
template <int DT>
__device__ __forceinline__ float vel(const EnvArgs& a, int64_t c, int comp)
{
    using S = Storage<DT>;
    using ST = typename S::type;
    const ST* src = static_cast<const ST*>(a.src);
    return S::template load<float>(src, (int64_t)(1 + comp) * a.stride + cell_of(a, c), 1 + comp, a.inv_scale[1 + comp], a.offset[1 + comp]);
}

__device__ __forceinline__ int64_t cell(const EnvArgs& a, int x, int y, int z)
{
    return (int64_t)x + (int64_t)a.nx * ((int64_t)y + (int64_t)a.ny * z);
}

// d u_comp / d axis at (x, y, z) (lbm_func.py:208-230 for one axis)
template <int DT>
__device__ float deriv(const EnvArgs& a, int x, int y, int z, int comp, int axis, float gl)
{
    const int n = axis == 0 ? a.nx : axis == 1 ? a.ny : a.nz;
    const int p = axis == 0 ? x : axis == 1 ? y : z;
    const int64_t st = axis == 0 ? 1 : axis == 1 ? (int64_t)a.nx : (int64_t)a.nx * a.ny;
    const int64_t c = cell(a, x, y, z);
    if (p == 0) return (vel<DT>(a, c + st, comp) - vel<DT>(a, c, comp)) / gl;
    if (p == n - 1) return (vel<DT>(a, c, comp) - vel<DT>(a, c - st, comp)) / gl;
    return (vel<DT>(a, c + st, comp) - vel<DT>(a, c - st, comp)) / (2.0f * gl);
}

// Vorticity component about `normal` (2: dv/dx - du/dy, 0: dw/dy - dv/dz,
// 1: du/dz - dw/dx), as the reference's xy / yz / front kernels form it.
template <int DT>
__device__ float curl(const EnvArgs& a, int x, int y, int z, int normal, float gl)
{
    if (normal == 2) return deriv<DT>(a, x, y, z, 1, 0, gl) - deriv<DT>(a, x, y, z, 0, 1, gl);
    if (normal == 0) return deriv<DT>(a, x, y, z, 2, 1, gl) - deriv<DT>(a, x, y, z, 1, 2, gl);
    return deriv<DT>(a, x, y, z, 0, 2, gl) - deriv<DT>(a, x, y, z, 2, 0, gl);
}

template <int DT>
__device__ float speed(const EnvArgs& a, int64_t c, int dim)
{
    const float ux = vel<DT>(a, c, 0), uy = vel<DT>(a, c, 1);
    float s = ux * ux + uy * uy;
    if (dim == 3) {
        const float uz = vel<DT>(a, c, 2);
        s = s + uz * uz;
    }
    return sqrtf(s);
}

// Mask value of a cell (0 outside the mask's box, which is all it covers).
__device__ __forceinline__ bool masked(const EnvArgs& a, int x, int y, int z)
{
    if (x < a.box[0] || x >= a.box[1] || y < a.box[2] || y >= a.box[3]) return false;
    if (a.nz > 1 && (z < a.box[4] || z >= a.box[5])) return false;
    return a.mask[cell(a, x, y, z)] != 0;
}

// Masked cell, or a masked cell within distance t (cell centres; in the
// plane for 2D lattices, in 3D otherwise).
__device__ bool solid_near(const EnvArgs& a, int x, int y, int z, float t)
{
    if (a.mask == nullptr) return false;
    if (masked(a, x, y, z)) return true;
    if (!(t > 0.0f)) return false;
    const int r = min((int)t, kMaxThickness);
    const int rz = a.nz > 1 ? r : 0;
    const float t2 = t * t;
    for (int dz = -rz; dz <= rz; ++dz)
        for (int dy = -r; dy <= r; ++dy)
            for (int dx = -r; dx <= r; ++dx) {
                const int qx = x + dx, qy = y + dy, qz = z + dz;
                if (qx < 0 || qy < 0 || qz < 0 || qx >= a.nx || qy >= a.ny || qz >= a.nz) continue;
                if ((float)(dx * dx + dy * dy + dz * dz) <= t2 && masked(a, qx, qy, qz)) return true;
            }
    return false;
}

__device__ __forceinline__ bool face_cell(const EnvArgs& a, int x, int y, int z)
{
    return x == 0 || y == 0 || x == a.nx - 1 || y == a.ny - 1 || (a.nz > 1 && (z == 0 || z == a.nz - 1));
}

// One image of a plane: normal 2 = the xy plane at z = index (the 2D
// lattice itself: index 0), out[y * nx + x]; normal 0 = the yz plane at
// x = index, out[z * ny + y]. RGB writes out[3 * pixel + channel].
template <int DT>
__global__ void __launch_bounds__(256) k_render_plane(EnvArgs a, int kind, int normal, int index, RenderArgs r,
                                                      float* out)
{
    const int w = normal == 2 ? a.nx : a.ny, h = normal == 2 ? a.ny : a.nz;
    const int64_t pix = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (pix >= (int64_t)w * h) return;
    const int i = (int)(pix % w), j = (int)(pix / w);
    const int x = normal == 2 ? i : index, y = normal == 2 ? j : i, z = normal == 2 ? index : j;
    const int dim = a.nz > 1 ? 3 : 2;
    const int64_t c = cell(a, x, y, z);
    switch (kind) {
    case HL_RENDER_SPEED: out[pix] = speed<DT>(a, c, dim); break;
    case HL_RENDER_RGB: {
        const float ux = vel<DT>(a, c, 0), uy = vel<DT>(a, c, 1);
        const float mag = sqrtf(ux * ux + uy * uy);
        out[3 * pix + 0] = fminf(fmaxf(mag / r.vmax, 0.0f), 1.0f);
        out[3 * pix + 1] = fminf(fmaxf((ux + r.vmax) / (2.0f * r.vmax), 0.0f), 1.0f);
        out[3 * pix + 2] = fminf(fmaxf((uy + r.vmax) / (2.0f * r.vmax), 0.0f), 1.0f);
        break;
    }
    case HL_RENDER_VORTICITY: out[pix] = curl<DT>(a, x, y, z, normal, r.gl); break;
    case HL_RENDER_VORTICITY_SOLID:
        out[pix] = solid_near(a, x, y, z, r.thickness) ? kSolidMark : curl<DT>(a, x, y, z, normal, r.gl);
        break;
    default: out[pix] = solid_near(a, x, y, z, r.thickness) ? 1.0f : 0.0f; break;  // HL_RENDER_SOLID
    }
}

// One ray per pixel along `axis` (0: x, image out[z * ny + y]; 1: y, image
// out[z * nx + x]; 2: z, image out[y * nx + x]). Face cells are skipped as
// the reference skips ML_WALL cells; a ray through the mask is solid.
template <int DT>
__global__ void __launch_bounds__(256) k_render_projection(EnvArgs a, int kind, int axis, RenderArgs r, float* out)
{
    const int w = axis == 0 ? a.ny : a.nx, h = axis == 2 ? a.ny : a.nz;
    const int64_t pix = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (pix >= (int64_t)w * h) return;
    const int i = (int)(pix % w), j = (int)(pix / w);
    const int n = axis == 0 ? a.nx : axis == 1 ? a.ny : a.nz;
    auto at = [&](int k, int& x, int& y, int& z) {
        x = axis == 0 ? k : i;
        y = axis == 0 ? i : axis == 1 ? k : j;
        z = axis == 2 ? k : j;
    };
    bool hit = false;
    if (a.mask != nullptr)
        for (int k = 0; k < n && !hit; ++k) {
            int x, y, z;
            at(k, x, y, z);
            hit = masked(a, x, y, z);
        }
    if (hit) {
        out[pix] = kSolidMark;
        return;
    }
    float v = 0.0f;
    if (kind == HL_PROJECT_FIRST_SPEED) {  // from the viewer's side (high index) inward
        for (int k = n - 1; k >= 0; --k) {
            int x, y, z;
            at(k, x, y, z);
            if (face_cell(a, x, y, z)) continue;
            const float s = speed<DT>(a, cell(a, x, y, z), 3);
            if (s > 1e-8f) {
                v = s;
                break;
            }
        }
    } else if (kind == HL_PROJECT_MAX_SPEED) {
        for (int k = n - 1; k >= 0; --k) {
            int x, y, z;
            at(k, x, y, z);
            if (face_cell(a, x, y, z)) continue;
            const float s = speed<DT>(a, cell(a, x, y, z), 3);
            if (s > v) v = s;
        }
    } else {  // HL_PROJECT_MAX_VORTICITY: signed value of the largest |curl| about the axis
        for (int k = 1; k < n - 1; ++k) {
            int x, y, z;
            at(k, x, y, z);
            if (face_cell(a, x, y, z)) continue;
            const float q = curl<DT>(a, x, y, z, axis, r.gl);
            if (fabsf(q) > fabsf(v)) v = q;
        }
    }
    out[pix] = v;
}

}  // namespace

hipError_t render_plane(int dtype, const EnvArgs& a, int kind, int normal, int index, float gl, float thickness,
                        float vmax, float* out, hipStream_t s)
{
    const int64_t n = normal == 2 ? (int64_t)a.nx * a.ny : (int64_t)a.ny * a.nz;
    const RenderArgs r{gl, thickness, vmax};
    const unsigned grid = (unsigned)((n + 255) / 256);
    switch (dtype) {
    case HL_F32: hipLaunchKernelGGL(k_render_plane<HL_F32>, grid, 256, 0, s, a, kind, normal, index, r, out); break;
    case HL_F64: hipLaunchKernelGGL(k_render_plane<HL_F64>, grid, 256, 0, s, a, kind, normal, index, r, out); break;
    case HL_F16: hipLaunchKernelGGL(k_render_plane<HL_F16>, grid, 256, 0, s, a, kind, normal, index, r, out); break;
    case HL_BF16: hipLaunchKernelGGL(k_render_plane<HL_BF16>, grid, 256, 0, s, a, kind, normal, index, r, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t render_projection(int dtype, const EnvArgs& a, int kind, int axis, float gl, float* out, hipStream_t s)
{
    const int64_t n = axis == 0 ? (int64_t)a.ny * a.nz : axis == 1 ? (int64_t)a.nx * a.nz : (int64_t)a.nx * a.ny;
    const RenderArgs r{gl, 0.0f, 1.0f};
    const unsigned grid = (unsigned)((n + 255) / 256);
    switch (dtype) {
    case HL_F32: hipLaunchKernelGGL(k_render_projection<HL_F32>, grid, 256, 0, s, a, kind, axis, r, out); break;
    case HL_F64: hipLaunchKernelGGL(k_render_projection<HL_F64>, grid, 256, 0, s, a, kind, axis, r, out); break;
    case HL_F16: hipLaunchKernelGGL(k_render_projection<HL_F16>, grid, 256, 0, s, a, kind, axis, r, out); break;
    case HL_BF16: hipLaunchKernelGGL(k_render_projection<HL_BF16>, grid, 256, 0, s, a, kind, axis, r, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int max_render_thickness() { return kMaxThickness; }

}  // namespace hl
