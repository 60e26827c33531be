// Per-cell gather step kernels (HL_KERNEL_GATHER).
//
// One thread per lattice cell. Interior cells pull the 9/27 neighbour moment
// sets, rebuild one population from each, apply the mask wall rule on cut
// links, take moments and collide. Face cells compute the same update for
// their projection onto the interior and then apply the face conditions in
// the order x, y, z (no separate boundary launch, no edge race).
//
// The arithmetic follows the reference's operation order with contraction
// off, so the FP32 build reproduces a literal FP32 evaluation of
// stream_and_collide{,_3d} (envs/lbm/lbm_func.py:322-467,
// envs/lbm3d/lbm_func_3d.py:364-604), and the FP64 build is the GPU FP64
// oracle used to state the error of the fast kernels.
#pragma clang fp contract(off)

#include "gather_post.h"
#include "kernels.h"
#include "lbm_device.h"

namespace hl {

template <typename T, int DT>
__global__ void __launch_bounds__(256) k_gather_d2q9(EnvArgs a)
{
    const int x = a.region[0] + blockIdx.x * blockDim.x + threadIdx.x;
    const int y = a.region[2] + blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= a.region[1] || y >= a.region[3]) return;
    const int px = min(max(x, 1), a.nx - 2), py = min(max(y, 1), a.ny - 2);
    T m[6];
    d2q9_post<T, DT>(a, px, py, m);
    if (px != x || py != y) bc_chain<T, 2>(m, x, y, 0, a.nx, a.ny, 1, a.bc_type, a.bc_value);
    using ST = typename Storage<DT>::type;
    ST* dst = static_cast<ST*>(a.dst);
    const int64_t c = (int64_t)x + (int64_t)a.nx * y;
#pragma unroll
    for (int f = 0; f < 6; ++f) Storage<DT>::template store<T>(dst, f * a.stride + c, m[f], f, a.scale[f]);
}

template <typename T, int DT>
__global__ void __launch_bounds__(256) k_gather_d3q27(EnvArgs a)
{
    const int x = a.region[0] + blockIdx.x * blockDim.x + threadIdx.x;
    const int y = a.region[2] + blockIdx.y * blockDim.y + threadIdx.y;
    const int z = a.region[4] + blockIdx.z;
    if (x >= a.region[1] || y >= a.region[3]) return;
    const int px = min(max(x, 1), a.nx - 2), py = min(max(y, 1), a.ny - 2), pz = min(max(z, 1), a.nz - 2);
    T m[10];
    d3q27_post<T, DT>(a, px, py, pz, m);
    if (px != x || py != y || pz != z) bc_chain<T, 3>(m, x, y, z, a.nx, a.ny, a.nz, a.bc_type, a.bc_value);
    using ST = typename Storage<DT>::type;
    ST* dst = static_cast<ST*>(a.dst);
    const int64_t c = (int64_t)x + (int64_t)a.nx * ((int64_t)y + (int64_t)a.ny * z);
#pragma unroll
    for (int f = 0; f < 10; ++f) Storage<DT>::template store<T>(dst, f * a.stride + c, m[f], f, a.scale[f]);
}

template <typename T, int DT>
static hipError_t launch_gather(int lattice, const EnvArgs& a, hipStream_t s)
{
    const int w = a.region[1] - a.region[0], h = a.region[3] - a.region[2], d = a.region[5] - a.region[4];
    if (w <= 0 || h <= 0 || d <= 0) return hipSuccess;
    dim3 block(64, 4, 1);
    dim3 grid((w + 63) / 64, (h + 3) / 4, lattice == HL_D2Q9 ? 1 : d);
    if (lattice == HL_D2Q9) hipLaunchKernelGGL((k_gather_d2q9<T, DT>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_gather_d3q27<T, DT>), grid, block, 0, s, a);
    return hipGetLastError();
}

hipError_t step_gather(int lattice, int dtype, const EnvArgs& a, hipStream_t s)
{
    switch (dtype) {
    case HL_F32: return launch_gather<float, HL_F32>(lattice, a, s);
    case HL_F64: return launch_gather<double, HL_F64>(lattice, a, s);
    case HL_F16: return launch_gather<float, HL_F16>(lattice, a, s);
    case HL_BF16: return launch_gather<float, HL_BF16>(lattice, a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace hl
