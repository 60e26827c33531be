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
    gather_d2q9_cell<T, DT>(a, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y);
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
    const int64_t c = cell_at(a, x, y, z);
#pragma unroll
    for (int f = 0; f < 10; ++f)
        Storage<DT>::template store<T>(dst, f * a.stride + c, m[f], f, a.scale[f], a.offset[f], HL_ROUNDING_KEY(a, f, x, y, z));
}

// D2Q9 gather step that also produces the body forces of the state it reads
// (lbm_func.py:383-401): the first fx.nfix wavefronts of the launch walk the
// near-body list two cells per wave, one lane per direction, with the link
// force and per-solid wave sums of the stand-alone fix-up kernel, one FP64
// row per wave; every wave also takes a share of summing the previous
// step's rows (one wave per output, the order of k_sum_partials). The
// near-body cells themselves are exact in the gather pass, so nothing is
// rewritten. One launch per step instead of three.
template <typename T, int DT>
__global__ void __launch_bounds__(256) k_gather_d2q9_fx(EnvArgs a, FixArgs fx)
{
    __shared__ double accs[4][HL_MAX_SOLIDS * 6];
    const int wave = threadIdx.y, lane = threadIdx.x;
    const int gw = (blockIdx.y * gridDim.x + blockIdx.x) * 4 + wave;
    const int nwaves = gridDim.x * gridDim.y * 4;
    const int width = fx.width;
    if (gw < fx.nfix) {
        double* acc = accs[wave];
        for (int k = lane; k < width; k += 64) acc[k] = 0.0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        const int h = lane >> 5, i = lane & 31;
        const int nx = a.nx, ny = a.ny;
        const Region rg = fx.rg;
        const int ncut = *fx.count;
        for (int base = gw * 2; base < ncut; base += fx.nfix * 2) {
            const int e = base + h;
            int sid = -1;
            T F[6] = {0, 0, 0, 0, 0, 0};
            if (e < ncut && i > 0 && i < 9) {
                const int tt = fx.list[e];
                const int x = rg.x0 + tt % rg.w, y = rg.y0 + (tt / rg.w) % rg.h;
                const int px = min(max(x, 1), nx - 2), py = min(max(y, 1), ny - 2);
                if (px == x && py == y) {  // a face cell repeats its projection: no force
                    const int64_t c = (int64_t)x + (int64_t)nx * y;
                    const int mval_c = a.mask[c];
                    const int mv = a.mask[c - D2CX[i] - (int64_t)nx * D2CY[i]];
                    if ((mv != 0) != (mval_c != 0)) {
                        const int s2 = mval_c ? mval_c - 1 : mv - 1;
                        if (s2 < a.nsolid) {
                            sid = s2;
                            T mc[6];
                            load_cell<T, DT>(a, c, 6, mc);
                            link_force<T, DT, 9>(a, mc, x, y, 0, i, sid, F);
                        }
                    }
                }
            }
            wave_solid_sums(sid, F, kRow6, lane, acc);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        for (int k = lane; k < width; k += 64) fx.partial[(int64_t)gw * width + k] = acc[k];
    }
    if (fx.prev) {
        for (int k = gw; k < fx.width; k += nwaves) {
            const double v = wave_sum(lane_rows_sum(fx.prev, fx.prev_rows, fx.width, k, lane));
            if (lane == 0) fx.prev_out[k] = v;
        }
    }
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
    for (int f = 0; f < 6; ++f)
        Storage<DT>::template store<T>(dst, f * a.stride + c, m[f], f, a.scale[f], a.offset[f], HL_ROUNDING_KEY(a, f, x, y, 0));
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

hipError_t step_gather_forces(int dtype, const EnvArgs& a, FixArgs& fx, hipStream_t s)
{
    const int w = a.region[1] - a.region[0], h = a.region[3] - a.region[2];
    if (w <= 0 || h <= 0 || fx.list == nullptr || fx.partial == nullptr) return hipErrorInvalidValue;
    const dim3 block(64, 4, 1), grid((w + 63) / 64, (h + 3) / 4, 1);
    const int64_t waves = (int64_t)grid.x * grid.y * 4, pairs = (fx.rg.n() + 1) / 2;
    int64_t n = pairs < fx.max_waves ? pairs : fx.max_waves;
    if (n > waves) n = waves;
    fx.nfix = (int)(n > 0 ? n : 1);
    switch (dtype) {
    case HL_F32: hipLaunchKernelGGL((k_gather_d2q9_fx<float, HL_F32>), grid, block, 0, s, a, fx); break;
    case HL_F64: hipLaunchKernelGGL((k_gather_d2q9_fx<double, HL_F64>), grid, block, 0, s, a, fx); break;
    case HL_F16: hipLaunchKernelGGL((k_gather_d2q9_fx<float, HL_F16>), grid, block, 0, s, a, fx); break;
    case HL_BF16: hipLaunchKernelGGL((k_gather_d2q9_fx<float, HL_BF16>), grid, block, 0, s, a, fx); break;
    default: return hipErrorInvalidValue;
    }
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
