// Batched D2Q9 engine: many resident environments of uneven grid sizes
// advanced together, one launch per kernel kind per step for all of them.
//
// Work-lists map each block to an (environment, tile). Per-environment state
// that changes during an advance (ping-pong parity, kinematics time index)
// lives in device memory and is advanced by the last kernel of each step, so
// a multi-step advance needs no host involvement and can be captured in a
// graph and replayed.
//
// Per step, for all active environments:
//   k_bvox      rebuild the body mask in the body box from the capsule table
//               at the environment's time t
//   k_bstep     the D2Q9 update of one 64x8 tile staged with its halo in LDS
//               (reference operation order, mask wall rule); tiles that
//               meet the body also reduce the per-solid momentum-exchange
//               force of the state they read into an FP64 partial
//   k_badvance  per environment: fixed-order sum of its tiles' partials into
//               the force history, then flip the parity and advance time.
// The same kernels can be launched one environment at a time (the
// per-environment baseline); every cell's arithmetic and every sum's order
// are the same, so the two schedules give bitwise equal results.
//
// The persistent engine (k_brun, hl_batch_run) advances every environment K
// steps in one cooperative launch: the workgroups deal out the step tiles of
// every environment, each tile doing k_bstep's update with k_bvox fused into
// its staging (each staged cell's mask value is computed from the capsule
// table at the environment's time and the tile's own cells are written to
// the mask), then a grid barrier ends the step; the first pass of step k+1
// sums step k's force rows into the history. The cell arithmetic, the tile
// slots and the summation order are those of the three-kernel schedule, so
// states and forces agree with it bit for bit.
#pragma clang fp contract(off)

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "gather_post.h"
#include "homelbm.h"
#include "kernels.h"
#include "lbm_device.h"

namespace hl {

constexpr int kBatchMaxSolids = 16;
// step tile: kTX x kTY cells, 512 threads; lane = column, each wave kRows rows
constexpr int kTX = 64, kTY = 8;
constexpr int kThreads = 512, kWaves = kThreads / 64, kRows = kTY / kWaves;
constexpr int kLX = kTX + 3, kLY = kTY + 3;  // staged region: 2 cells below, 1 above
constexpr int kStage = (kLX * kLY + kThreads - 1) / kThreads;
constexpr int kVoxW = 16;         // voxelizer item: 16 x 16 cells
constexpr int kPartialWidth = kBatchMaxSolids * 3;  // Fx, Fy, Tz per solid
constexpr int kRunCaps = 32;  // capsules per environment the persistent engine stages in LDS

struct BatchEnv {
    EnvArgs base;         // arguments for parity 0 (moments[0] -> moments[1])
    const float* caps;    // [T][ncap][8] capsule table (null: static mask)
    const float* poses;   // [T][nsolid][8] pose table (null: static poses in base.solids)
    int32_t ncap, T;
    int32_t vox_begin, vox_count;    // voxelizer items
    int32_t step_begin, step_count;  // step tiles
    int32_t slot_begin, slot_count;  // force partial slots (step tiles meeting the body)
};

struct BatchItem {  // env, origin and extent of a rectangle of cells; partial slot or -1
    int32_t env, x0, y0, w, h, slot;
};

struct BatchArgs {
    const BatchEnv* envs;
    const BatchItem* step_items;
    const BatchItem* vox_items;
    int32_t* parity;        // per env
    int32_t* time;          // per env
    const uint8_t* active;  // per env (null = all)
    double* partial;        // [slots][kPartialWidth]
};

__device__ __forceinline__ bool env_active(const BatchArgs& b, int e) { return b.active == nullptr || b.active[e]; }

// A block-uniform value read from LDS, moved to a scalar register (so
// addresses built from it are a scalar base plus a 32-bit lane offset).
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <typename T>
__device__ __forceinline__ T* uni(T* p)
{
    const uint64_t a = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a), hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

// Per-block copy of the environment's arguments in LDS, with the source,
// destination and poses of its current step.
__device__ __forceinline__ void load_env(const BatchArgs& b, int e, EnvArgs& ea)
{
    static_assert(sizeof(EnvArgs) % 4 == 0 && sizeof(EnvArgs) / 4 <= 256, "EnvArgs copy");
    const BatchEnv& be = b.envs[e];
    if (threadIdx.x < sizeof(EnvArgs) / 4)
        reinterpret_cast<uint32_t*>(&ea)[threadIdx.x] = reinterpret_cast<const uint32_t*>(&be.base)[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        if (b.parity[e] & 1) {
            const void* t = ea.src;
            ea.src = ea.dst;
            ea.dst = const_cast<void*>(t);
        }
        if (be.poses && be.T > 0) ea.solids = be.poses + (int64_t)(b.time[e] % be.T) * ea.nsolid * 8;
    }
    __syncthreads();
}

// Mask value of cell (x, y) (the same evaluation as the single-lattice
// voxelizer): solid id + 1 of the lowest-index capsule whose segment lies
// within its radius of the cell centre, else 0.
template <typename P>
__device__ __forceinline__ uint8_t capsule_voxel(P caps, int ncap, int x, int y)
{
    const float px = (float)x, py = (float)y, pz = 0.0f;
    for (int j = 0; j < ncap; ++j) {
        const auto C = caps + 8 * j;
        float abx = C[3] - C[0], aby = C[4] - C[1], abz = C[5] - C[2];
        float apx = px - C[0], apy = py - C[1], apz = pz - C[2];
        float len2 = abx * abx + aby * aby + abz * abz;
        float tt = 0.0f;
        if (len2 > 0.0f) {
            tt = (apx * abx + apy * aby + apz * abz) / len2;
            tt = tt < 0.0f ? 0.0f : (tt > 1.0f ? 1.0f : tt);
        }
        float dx = apx - tt * abx, dy = apy - tt * aby, dz = apz - tt * abz;
        float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 <= C[6] * C[6]) return (uint8_t)((int)C[7] + 1);
    }
    return 0;
}

// Capsule voxelizer over one item.
__global__ void __launch_bounds__(256) k_bvox(BatchArgs b, int item0)
{
    const BatchItem it = b.vox_items[item0 + blockIdx.x];
    if (!env_active(b, it.env)) return;
    const BatchEnv& be = b.envs[it.env];
    if (!be.caps || be.T <= 0) return;
    const float* caps = be.caps + (int64_t)(b.time[it.env] % be.T) * be.ncap * 8;
    const int k = threadIdx.x;
    if (k >= it.w * it.h) return;
    const int x = it.x0 + k % it.w, y = it.y0 + k / it.w;
    const_cast<uint8_t*>(be.base.mask)[(int64_t)x + (int64_t)be.base.nx * y] =
        capsule_voxel(as_global(caps), be.ncap, x, y);
}

__device__ __forceinline__ double bwave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One step of tile `it` of environment be at buffer parity par and time t,
// its force row (if any) into part. VOX: the staged cells are voxelized from
// the capsule table at t on the spot (and the tile's own cells of the mask
// written) instead of read from the mask. The block's threads all call it
// for the same tile.
template <bool VOX>
__device__ __forceinline__ void bstep_tile(const BatchEnv& be, const BatchItem it, int par, int t, double* part)
{
    __shared__ EnvArgs ea;
    __shared__ float tile[6][kLY][kLX];   // moments (near the body only)
    __shared__ float ptile[9][kLY][kLX];  // outgoing populations of every staged cell
    __shared__ uint8_t mtile[kLY][kLX + 1];
    __shared__ float pose[kBatchMaxSolids * 8];
    __shared__ double acc[kWaves][kPartialWidth];
    __shared__ float vcaps[kRunCaps * 8];  // VOX: the capsule table at t
    const int tid = threadIdx.x;
    const bool near = it.slot >= 0;
    const int nx = uni(be.base.nx), ny = uni(be.base.ny), nsolid = uni(be.base.nsolid);
    // an environment's fields hold fewer than 2^31 values: 32-bit offsets
    const int st = uni((int)be.base.stride);
    const float* src_p = uni(static_cast<const float*>(par ? be.base.dst : be.base.src));
    const auto src = as_global(src_p);
    const auto mask = as_global(uni(const_cast<uint8_t*>(be.base.mask)));
    // the staged region [x0-2, x0+kTX] x [y0-2, y0+kTY] reaches the body box
    // (the mask is 0 outside it): stage the mask and evaluate the wall rule
    const bool masked = mask != nullptr && it.x0 - 2 <= be.base.box[1] && it.x0 + kTX >= be.base.box[0] - 1 &&
                        it.y0 - 2 <= be.base.box[3] && it.y0 + kTY >= be.base.box[2] - 1;
    const int T = uni(be.T), ncap = uni(be.ncap);
    const float* bposes = uni(be.poses);
    const float* caps = uni(be.caps);
    const float* poses = bposes && T > 0 ? bposes + (t % T) * nsolid * 8 : uni(be.base.solids);
    const bool vox = VOX && masked && caps && T > 0;
    const int bx0 = uni(be.base.box[0]), bx1 = uni(be.base.box[1]), by0 = uni(be.base.box[2]), by1 = uni(be.base.box[3]);
    // issue every load of the tile before the first wait: the capsule table
    // and poses at t, then the tile with a halo of two cells below and one
    // above in x and y (so every face cell's interior projection finds its
    // sources here) and its mask
    float cv = 0.0f, pv = 0.0f;
    if (vox && tid < ncap * 8) cv = as_global(caps + (t % T) * ncap * 8)[tid];
    if (masked && poses && tid < nsolid * 8) pv = as_global(poses)[tid];
    float v[kStage][6];
    int mv[kStage];
    bool in[kStage];
#pragma unroll
    for (int r = 0; r < kStage; ++r) {
        const int k = tid + r * kThreads;
        const int lx = k % kLX, ly = k / kLX;
        const int gx = it.x0 - 2 + lx, gy = it.y0 - 2 + ly;
        in[r] = k < kLX * kLY && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
        const int c = (in[r] ? gx : 0) + nx * (in[r] ? gy : 0);
#pragma unroll
        for (int f = 0; f < 6; ++f) v[r][f] = in[r] ? src[f * st + c] : 0.0f;
        mv[r] = (!vox && masked && in[r]) ? mask[c] : 0;
    }
    if (vox) {
        if (tid < ncap * 8) vcaps[tid] = cv;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kStage; ++r) {
            const int k = tid + r * kThreads;
            const int lx = k % kLX, ly = k / kLX;
            const int gx = it.x0 - 2 + lx, gy = it.y0 - 2 + ly;
            if (in[r] && gx >= bx0 && gx < bx1 && gy >= by0 && gy < by1) {
                mv[r] = capsule_voxel(vcaps, ncap, gx, gy);
                // the tile's own cells of the mask
                if (lx >= 2 && lx < kTX + 2 && ly >= 2 && ly < kTY + 2) mask[gx + nx * gy] = (uint8_t)mv[r];
            }
        }
    }
    static_assert(sizeof(EnvArgs) % 4 == 0 && sizeof(EnvArgs) / 4 <= kThreads, "EnvArgs copy");
    if (tid < sizeof(EnvArgs) / 4)
        reinterpret_cast<uint32_t*>(&ea)[tid] = reinterpret_cast<const uint32_t*>(&be.base)[tid];
    if (near)
        for (int k = tid; k < kWaves * kPartialWidth; k += kThreads) (&acc[0][0])[k] = 0.0;
    // each staged cell rebuilds its 9 outgoing populations once (the pull of
    // a neighbour reads them instead of rebuilding from the moments); the
    // values are those the reference's per-link reconstruction gives
#pragma unroll
    for (int r = 0; r < kStage; ++r) {
        const int k = tid + r * kThreads;
        if (k < kLX * kLY) {
            const int lx = k % kLX, ly = k / kLX;
            if (masked) {
#pragma unroll
                for (int f = 0; f < 6; ++f) tile[f][ly][lx] = v[r][f];
                mtile[ly][lx] = (uint8_t)mv[r];
            }
            if (in[r]) {
                float f9[9];
                d2q9_pops<float>(v[r], f9);
#pragma unroll
                for (int i = 0; i < 9; ++i) ptile[i][ly][lx] = f9[i];
            }
        }
    }
    if (tid < kBatchMaxSolids * 8) pose[tid] = pv;
    __syncthreads();
    if (tid == 0) {
        ea.src = src_p;
        ea.dst = const_cast<void*>(par ? be.base.src : be.base.dst);
        ea.solids = pose;  // wall velocities read the staged pose table
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    const int x = it.x0 + lane;
    const auto dst = as_global(uni(static_cast<float*>(const_cast<void*>(par ? be.base.src : be.base.dst))));
    uint32_t cutbits[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
        const int y = it.y0 + wave * kRows + r;
        cutbits[r] = 0;
        if (x >= nx || y >= ny) continue;
        // a face cell takes its interior projection's update, then the face chain
        const int px = min(max(x, 1), nx - 2), py = min(max(y, 1), ny - 2);
        const int tx = px - it.x0 + 2, ty = py - it.y0 + 2;
        // stream_and_collide (lbm_func.py:322-467): the staged populations,
        // or the wall rule from the cell's own moments on a cut link
        float pop[9];
        uint32_t cb = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) pop[i] = ptile[i][ty - D2CY[i]][tx - D2CX[i]];
        if (masked) {
            const int mval_c = mtile[ty][tx];
            float mc[6];
#pragma unroll
            for (int f = 0; f < 6; ++f) mc[f] = tile[f][ty][tx];
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                const int mval_s = mtile[ty - D2CY[i]][tx - D2CX[i]];
                if ((mval_s != 0) != (mval_c != 0)) {
                    cb |= 1u << i;
                    pop[i] = d2q9_link_pop<float>(ea, px, py, i, mc, 1, mval_c ? mval_c - 1 : mval_s - 1, mc);
                }
            }
        }
        float m[6];
        d2q9_collide<float>(ea, pop, 0.0f, 0.0f, m);
        if (px != x || py != y) bc_chain<float, 2>(m, x, y, 0, nx, ny, 1, ea.bc_type, ea.bc_value);
        else cutbits[r] = cb;  // forces come from interior cells only
        const int c = x + nx * y;
#pragma unroll
        for (int f = 0; f < 6; ++f) dst[f * st + c] = m[f];
    }
    if (!near) return;
    // momentum exchange of the tile's cut links per solid: each lane's cells
    // and links in order, a wavefront butterfly, then the block's waves in order
    uint32_t sids = 0;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
        const int ty = wave * kRows + r + 2, tx = lane + 2;
        for (int i = 1; i < 9; ++i)
            if ((cutbits[r] >> i) & 1u) {
                const int mvc = mtile[ty][tx], mvs = mtile[ty - D2CY[i]][tx - D2CX[i]];
                const int s = mvc ? mvc - 1 : mvs - 1;
                if (s < nsolid) sids |= 1u << s;
            }
    }
    uint32_t uni = sids;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) uni |= (uint32_t)__shfl_xor((int)uni, off, 64);
    while (uni) {
        const int s = __builtin_ctz(uni);
        uni &= uni - 1;
        double sum[3] = {0.0, 0.0, 0.0};
        if ((sids >> s) & 1u) {
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                if (!cutbits[r]) continue;
                const int ty = wave * kRows + r + 2, tx = lane + 2;
                const int y = it.y0 + ty - 2;
                const int mvc = mtile[ty][tx];
                float mc[6];
#pragma unroll
                for (int f = 0; f < 6; ++f) mc[f] = tile[f][ty][tx];
                for (int i = 1; i < 9; ++i) {
                    if (!((cutbits[r] >> i) & 1u)) continue;
                    const int mvs = mtile[ty - D2CY[i]][tx - D2CX[i]];
                    if ((mvc ? mvc - 1 : mvs - 1) != s) continue;
                    float F[6];
                    link_force<float, HL_F32, 9>(ea, mc, x, y, 0, i, s, F);
                    sum[0] += (double)F[0];
                    sum[1] += (double)F[1];
                    sum[2] += (double)F[5];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) sum[q] = bwave_sum(sum[q]);
        if (lane == 0)
            for (int q = 0; q < 3; ++q) acc[wave][s * 3 + q] = sum[q];
    }
    __syncthreads();
    if (tid < kPartialWidth) {
        double sum = 0.0;
        for (int w = 0; w < kWaves; ++w) sum += acc[w][tid];
        as_global(part)[(int64_t)it.slot * kPartialWidth + tid] = sum;
    }
}

__global__ void __launch_bounds__(kThreads) k_bstep(BatchArgs b, int item0)
{
    const BatchItem it = b.step_items[item0 + blockIdx.x];
    if (!env_active(b, it.env)) return;
    bstep_tile<false>(b.envs[it.env], it, b.parity[it.env] & 1, b.time[it.env], b.partial);
}

// Per environment (one block): the fixed-order sum of its partial slots into
// the force history [k][env][16][6] (Fx, Fy, 0, 0, 0, Tz), then flip the
// parity and advance the time.
__global__ void __launch_bounds__(64) k_badvance(BatchArgs b, int env0, double* history, int nenv_total, int k)
{
    __shared__ double vals[kPartialWidth];
    const int e = env0 + blockIdx.x;
    if (!env_active(b, e)) return;
    const BatchEnv& be = b.envs[e];
    const int q = threadIdx.x;
    if (history) {
        if (q < kPartialWidth) {
            const double* p = b.partial + (int64_t)be.slot_begin * kPartialWidth + q;
            const int n = be.slot_count;
            double v = 0.0;
            int j = 0;
            for (; j + 8 <= n; j += 8) {  // loads issued together, added in slot order
                double t[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) t[u] = p[(int64_t)(j + u) * kPartialWidth];
#pragma unroll
                for (int u = 0; u < 8; ++u) v += t[u];
            }
            for (; j < n; ++j) v += p[(int64_t)j * kPartialWidth];
            vals[q] = v;
        }
        __syncthreads();
        double* h = history + ((int64_t)k * nenv_total + e) * kBatchMaxSolids * 6;
        for (int o = q; o < kBatchMaxSolids * 6; o += 64) {
            const int s = o / 6, comp = o % 6;
            h[o] = comp == 0 ? vals[s * 3] : comp == 1 ? vals[s * 3 + 1] : comp == 5 ? vals[s * 3 + 2] : 0.0;
        }
    }
    if (q == 0) {
        b.parity[e] ^= 1;
        b.time[e] += 1;
    }
}

// ------------------------------------------------- persistent engine

// grid barrier: iterations of a 256-cycle sleep before a wait gives up and
// flags the launch (about half a second)
constexpr unsigned kSpinLimit = 1u << 21;

// Every workgroup arrives once per barrier; target = arrivals expected so
// far. Stores before the barrier are released to the device (every XCD's
// L2) and loads after it acquire them. A wait that outlasts kSpinLimit
// sleeps sets *flag and lets the launch finish (wrong results, no hang).
__device__ __forceinline__ void grid_barrier(unsigned* bar, unsigned target, int* flag)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned n = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(4);
            if (++n > kSpinLimit) {
                __hip_atomic_fetch_or(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Step k's force rows (bank part) of environments [e0, e0 + ne) summed in
// slot order into history[k][env][16][6] (Fx, Fy, 0, 0, 0, Tz), one thread
// per output: the order of k_badvance, so the sums agree bitwise.
__device__ __forceinline__ void sum_output(const BatchArgs& b, const double* part, int k, double* history,
                                           int nenv_total, int e, int r);
__device__ __forceinline__ void run_sums(const BatchArgs& b, const double* part, int k, double* history,
                                         int nenv_total, int e0, int ne)
{
    const int gt = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    for (int o = gt; o < ne * kBatchMaxSolids * 6; o += nt)
        sum_output(b, part, k, history, nenv_total, e0 + o / (kBatchMaxSolids * 6), o % (kBatchMaxSolids * 6));
}

// Output r of environment e's step-k force history row (see run_sums).
__device__ __forceinline__ void sum_output(const BatchArgs& b, const double* part, int k, double* history,
                                           int nenv_total, int e, int r)
{
    if (!env_active(b, e)) return;
    const int so = r / 6, comp = r % 6;
    double v = 0.0;
    if (comp == 0 || comp == 1 || comp == 5) {
        const BatchEnv& be = b.envs[e];
        const double* p = part + (int64_t)be.slot_begin * kPartialWidth + so * 3 + (comp == 5 ? 2 : comp);
        const int n = be.slot_count;
        int j = 0;
        for (; j + 8 <= n; j += 8) {  // loads issued together, added in slot order
            double tv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) tv[u] = p[(int64_t)(j + u) * kPartialWidth];
#pragma unroll
            for (int u = 0; u < 8; ++u) v += tv[u];
        }
        for (; j < n; ++j) v += p[(int64_t)j * kPartialWidth];
    }
    history[((int64_t)k * nenv_total + e) * kBatchMaxSolids * 6 + r] = v;
}

// K steps of the step tiles [i0, i0 + ni) (environments [e0, e0 + ne)) in
// one launch: the tiles dealt to the workgroups in order, each tile the
// three-kernel schedule's step with the voxelizer fused into its staging; a
// grid barrier after every step; step k's force rows go to bank k & 1 and
// are summed by step k + 1 (the last step's after the loop); then parity and
// time advance by K.
__global__ void __launch_bounds__(kThreads) k_brun(BatchArgs b, int i0, int ni, int K, double* history,
                                                   int nenv_total, int e0, int ne, int nslots, unsigned* bar,
                                                   int* flag)
{
    static_assert(sizeof(BatchEnv) % 4 == 0 && sizeof(BatchEnv) / 4 < kThreads, "BatchEnv copy");
    __shared__ BatchEnv hdr;
    __shared__ int hst[3];  // parity, time, active
    for (int k = 0; k < K; ++k) {
        if (k > 0 && history)
            run_sums(b, b.partial + (size_t)((k - 1) & 1) * nslots * kPartialWidth, k - 1, history, nenv_total, e0, ne);
        for (int i = i0 + blockIdx.x; i < i0 + ni; i += gridDim.x) {
            // the tile's environment copied to LDS in one round of loads (its
            // fields then read from LDS, not one dependent load at a time)
            const BatchItem it = b.step_items[i];
            const int e = it.env;
            __syncthreads();  // the previous tile's LDS reads are done
            if (threadIdx.x < sizeof(BatchEnv) / 4) {
                reinterpret_cast<uint32_t*>(&hdr)[threadIdx.x] =
                    as_global(reinterpret_cast<const uint32_t*>(b.envs + e))[threadIdx.x];
            } else if (threadIdx.x >= kThreads - 64) {  // the last wave: loads issued together
                const int p = as_global(b.parity)[e], tm = as_global(b.time)[e];
                const int a = b.active ? (int)as_global(b.active)[e] : 1;
                if (threadIdx.x == kThreads - 1) {
                    hst[0] = p;
                    hst[1] = tm;
                    hst[2] = a;
                }
            }
            __syncthreads();
            if (!hst[2]) continue;
            bstep_tile<true>(hdr, it, (hst[0] + k) & 1, hst[1] + k,
                             b.partial + (size_t)(k & 1) * nslots * kPartialWidth);
        }
        grid_barrier(bar, (unsigned)(k + 1) * gridDim.x, flag);
    }
    if (K > 0 && history)
        run_sums(b, b.partial + (size_t)((K - 1) & 1) * nslots * kPartialWidth, K - 1, history, nenv_total, e0, ne);
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    for (int e = e0 + gt; e < e0 + ne; e += gridDim.x * blockDim.x)
        if (env_active(b, e)) {
            b.parity[e] = (b.parity[e] + K) & 1;
            b.time[e] += K;
        }
}

__global__ void k_brun_reset(unsigned* bar, int* flag)
{
    *bar = 0u;
    *flag = 0;
}

// XCD-affine persistent engine. The workgroups of a launch are dealt to the
// 8 XCDs round-robin (workgroup w runs on XCD w % 8; the host checks this
// once per batch and grid), and each XCD has its own L2. Every environment
// is assigned to one XCD group (the host balances their tiles), and only
// that group's workgroups touch its state, masks and force rows, so a step
// needs a barrier among one XCD's workgroups only: the stores completed
// (no L2 writeback: the writers share the readers' L2), an arrival count
// kept in that L2, and a device-scope cache invalidate afterwards. An
// invalidate of the L1 alone was measured to leave stale lines (the step
// after went wrong), the device-scope one does not; the agent-scope release
// and acquire of a grid-wide barrier (L2 writeback each step) measured 14.1
// GLUPS on the 64-eel batch, this barrier 18.0.
struct XcdPlan {
    int goff[9];   // step items of group g: xitems[goff[g] .. goff[g + 1])
    int geoff[9];  // environments of group g: genv[geoff[g] .. geoff[g + 1])
};
constexpr int kXcdBarStride = 64;  // one 256-byte line per group's count

// zero: 0, passed in by the host, so that the poll below stays an atomic
// (performed in the L2) instead of becoming a plain load that could hit a
// stale L1 line.
__device__ __forceinline__ void xcd_barrier(unsigned* bar, unsigned target, int* flag, unsigned zero)
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wave's stores are in the L2
    __syncthreads();
    if (threadIdx.x == 0) {
        // workgroup-scope atomics are performed in the XCD's L2, which every
        // workgroup of the group shares
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        unsigned n = 0;
        while (__hip_atomic_fetch_add(bar, zero, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < target) {
            __builtin_amdgcn_s_sleep(2);
            if (++n > kSpinLimit) {
                __hip_atomic_fetch_or(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    __syncthreads();
    asm volatile("buffer_inv sc1" ::: "memory");  // later loads miss every stale line
}

__global__ void __launch_bounds__(kThreads) k_brun_xcd(BatchArgs b, const int* xitems, XcdPlan plan, const int* genv,
                                                       int K, double* history, int nenv_total, int nslots,
                                                       unsigned* bars, int* flag, unsigned zero)
{
    static_assert(sizeof(BatchEnv) % 4 == 0 && sizeof(BatchEnv) / 4 < kThreads, "BatchEnv copy");
    __shared__ BatchEnv hdr;
    __shared__ int hst[3];  // parity, time, active
    const int g = (int)(blockIdx.x & 7), lb = (int)(blockIdx.x >> 3), nb = (int)(gridDim.x >> 3);
    // this workgroup's contiguous share of its group's tiles (consecutive
    // tiles mostly share an environment, whose header then stays in LDS)
    const int ib = plan.goff[g], ie = plan.goff[g + 1];
    const int per = (ie - ib + nb - 1) / nb;
    const int i0 = ib + lb * per, i1 = min(i0 + per, ie);
    const int ge0 = plan.geoff[g], ne = plan.geoff[g + 1] - ge0;
    const int gt = lb * kThreads + (int)threadIdx.x, nt = nb * kThreads;
    unsigned* bar = bars + g * kXcdBarStride;
    int cur = -1;  // environment whose header is in LDS
    for (int k = 0; k < K; ++k) {
        if (k > 0 && history)
            for (int o = gt; o < ne * kBatchMaxSolids * 6; o += nt)
                sum_output(b, b.partial + (size_t)((k - 1) & 1) * nslots * kPartialWidth, k - 1, history, nenv_total,
                           genv[ge0 + o / (kBatchMaxSolids * 6)], o % (kBatchMaxSolids * 6));
        for (int i = i0; i < i1; ++i) {
            const BatchItem it = b.step_items[xitems[i]];
            const int e = it.env;
            __syncthreads();  // the previous tile's LDS reads are done
            if (e != cur) {
                if (threadIdx.x < sizeof(BatchEnv) / 4) {
                    reinterpret_cast<uint32_t*>(&hdr)[threadIdx.x] =
                        as_global(reinterpret_cast<const uint32_t*>(b.envs + e))[threadIdx.x];
                } else if (threadIdx.x >= kThreads - 64) {
                    const int p = as_global(b.parity)[e], tm = as_global(b.time)[e];
                    const int a = b.active ? (int)as_global(b.active)[e] : 1;
                    if (threadIdx.x == kThreads - 1) {
                        hst[0] = p;
                        hst[1] = tm;
                        hst[2] = a;
                    }
                }
                cur = e;
                __syncthreads();
            }
            if (!hst[2]) continue;
            bstep_tile<true>(hdr, it, (hst[0] + k) & 1, hst[1] + k,
                             b.partial + (size_t)(k & 1) * nslots * kPartialWidth);
        }
        xcd_barrier(bar, (unsigned)(k + 1) * nb, flag, zero);
    }
    if (K > 0 && history)
        for (int o = gt; o < ne * kBatchMaxSolids * 6; o += nt)
            sum_output(b, b.partial + (size_t)((K - 1) & 1) * nslots * kPartialWidth, K - 1, history, nenv_total,
                       genv[ge0 + o / (kBatchMaxSolids * 6)], o % (kBatchMaxSolids * 6));
    // parity and time of this group's environments (no other group reads them)
    for (int j = gt; j < ne; j += nt) {
        const int e = genv[ge0 + j];
        if (env_active(b, e)) {
            b.parity[e] = (b.parity[e] + K) & 1;
            b.time[e] += K;
        }
    }
}

__global__ void k_brun_xcd_reset(unsigned* bars, int* flag)
{
    if (threadIdx.x < 8) bars[threadIdx.x * kXcdBarStride] = 0u;
    if (threadIdx.x == 0) *flag = 0;
}

// XCD of each workgroup (its hardware XCC id), for the host's check of the
// round-robin dealing the XCD-affine engine relies on.
__global__ void __launch_bounds__(kThreads) k_xcc_ids(int* out)
{
    if (threadIdx.x == 0) {
        unsigned x;
        asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
        out[blockIdx.x] = (int)(x & 0xf);
    }
}

}  // namespace hl

// ----------------------------------------------------------------- C ABI

struct hl_batch {
    int n = 0;
    std::vector<hl::BatchEnv> envs;  // host mirror of the device table
    std::vector<hl::BatchItem> step_items, vox_items;
    int nslots = 0;
    unsigned* d_bar = nullptr;  // grid barrier counter of the persistent engine
    int* d_flag = nullptr;      // set when a barrier wait gave up
    hl::BatchEnv* d_envs = nullptr;
    hl::BatchItem *d_step = nullptr, *d_vox = nullptr;
    size_t cap_envs = 0, cap_step = 0, cap_vox = 0, cap_partial = 0;
    int32_t *d_parity = nullptr, *d_time = nullptr;
    size_t cap_parity = 0, cap_time = 0;
    double* d_partial = nullptr;
    bool dirty = true;  // table or item lists changed since the last upload
    // XCD-affine persistent engine: environments dealt to the 8 XCD groups,
    // their step items grouped (d_xitems), the groups' environments (d_genv)
    bool plan_stale = true;
    hl::XcdPlan plan;
    int *d_xitems = nullptr, *d_genv = nullptr;
    size_t cap_xitems = 0, cap_genv = 0;
    unsigned* d_bars = nullptr;  // one arrival count per XCD group
    int* d_xcc = nullptr;        // workgroup -> XCC ids of the dealing check
    int xcd_grid = 0;            // grid the check ran for (0: none yet)
    bool xcd_ok = false;         // workgroup w ran on XCD w % 8 for that grid
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

namespace {

using hl::report;

int32_t hip_report(hipError_t e, const char* where)
{
    std::string m = std::string(where) + ": " + hipGetErrorString(e);
    return report(HL_E_HIP, m.c_str());
}

template <typename T> hipError_t upload(T** dptr, size_t* cap, const T* data, size_t count)
{
    const size_t need = count ? count : 1;
    if (!*dptr || *cap < need) {
        if (*dptr) {
            hipError_t e = hipFree(*dptr);
            if (e != hipSuccess) return e;
        }
        *dptr = nullptr;
        *cap = 0;
        hipError_t e = hipMalloc(dptr, need * sizeof(T));
        if (e != hipSuccess) return e;
        *cap = need;
    }
    if (!count) return hipSuccess;
    return hipMemcpy(*dptr, data, count * sizeof(T), hipMemcpyHostToDevice);
}

// Step tiles cover every environment; tiles that meet the force region
// (interior cells within one cell of the mask box, where every cut link has
// its fluid end) get a partial slot. Voxelizer items tile the mask box.
void rebuild_items(hl_batch* b)
{
    b->step_items.clear();
    b->vox_items.clear();
    int slot = 0;
    for (int e = 0; e < b->n; ++e) {
        hl::BatchEnv& be = b->envs[e];
        const hl::EnvArgs& a = be.base;
        const bool body = a.mask && a.box[0] < a.box[1] && a.box[2] < a.box[3];
        const int fx0 = std::max(a.box[0] - 1, 1), fx1 = std::min(a.box[1] + 1, a.nx - 1);
        const int fy0 = std::max(a.box[2] - 1, 1), fy1 = std::min(a.box[3] + 1, a.ny - 1);
        const bool forces = body && a.nsolid > 0 && fx0 < fx1 && fy0 < fy1;
        be.step_begin = (int)b->step_items.size();
        be.slot_begin = slot;
        for (int y = 0; y < a.ny; y += hl::kTY)
            for (int x = 0; x < a.nx; x += hl::kTX) {
                const bool meets = forces && x < fx1 && x + hl::kTX > fx0 && y < fy1 && y + hl::kTY > fy0;
                b->step_items.push_back({e, x, y, hl::kTX, hl::kTY, meets ? slot++ : -1});
            }
        be.step_count = (int)b->step_items.size() - be.step_begin;
        be.slot_count = slot - be.slot_begin;
        be.vox_begin = (int)b->vox_items.size();
        if (body && be.caps)
            for (int y = a.box[2]; y < a.box[3]; y += hl::kVoxW)
                for (int x = a.box[0]; x < a.box[1]; x += hl::kVoxW)
                    b->vox_items.push_back(
                        {e, x, y, std::min(hl::kVoxW, a.box[1] - x), std::min(hl::kVoxW, a.box[3] - y), -1});
        be.vox_count = (int)b->vox_items.size() - be.vox_begin;
    }
    b->nslots = slot;
    b->dirty = true;
}

int32_t sync_device(hl_batch* b)
{
    if (!b->dirty) return HL_OK;
    hipError_t e;
    if ((e = upload(&b->d_envs, &b->cap_envs, b->envs.data(), b->envs.size())) != hipSuccess ||
        (e = upload(&b->d_step, &b->cap_step, b->step_items.data(), b->step_items.size())) != hipSuccess ||
        (e = upload(&b->d_vox, &b->cap_vox, b->vox_items.data(), b->vox_items.size())) != hipSuccess)
        return hip_report(e, "batch work-list upload");
    // force rows of the tiles: two banks (the persistent engine sums one
    // step's while it writes the next)
    const size_t np = (size_t)std::max(2 * b->nslots, 1) * hl::kPartialWidth;
    if (!b->d_partial || b->cap_partial < np) {
        if (b->d_partial && (e = hipFree(b->d_partial)) != hipSuccess) return hip_report(e, "hipFree");
        b->d_partial = nullptr;
        b->cap_partial = 0;
        if ((e = hipMalloc(&b->d_partial, np * sizeof(double))) != hipSuccess)
            return hip_report(e, "batch partial-sum allocation");
        b->cap_partial = np;
    }
    b->dirty = false;
    b->plan_stale = true;
    return HL_OK;
}

// Environments dealt to the 8 XCD groups, most tiles first, each to the group
// with the fewest tiles so far; each group's step items in environment order.
int32_t build_plan(hl_batch* b)
{
    std::vector<int> order((size_t)b->n);
    for (int e = 0; e < b->n; ++e) order[(size_t)e] = e;
    std::stable_sort(order.begin(), order.end(),
                     [&](int p, int q) { return b->envs[(size_t)p].step_count > b->envs[(size_t)q].step_count; });
    std::vector<std::vector<int>> groups(8);
    int64_t load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int e : order) {
        int g = 0;
        for (int h = 1; h < 8; ++h)
            if (load[h] < load[g]) g = h;
        groups[(size_t)g].push_back(e);
        load[g] += b->envs[(size_t)e].step_count;
    }
    std::vector<int> xitems, genv;
    for (int g = 0; g < 8; ++g) {
        std::sort(groups[(size_t)g].begin(), groups[(size_t)g].end());
        b->plan.goff[g] = (int)xitems.size();
        b->plan.geoff[g] = (int)genv.size();
        for (int e : groups[(size_t)g]) {
            genv.push_back(e);
            const hl::BatchEnv& be = b->envs[(size_t)e];
            for (int t = 0; t < be.step_count; ++t) xitems.push_back(be.step_begin + t);
        }
    }
    b->plan.goff[8] = (int)xitems.size();
    b->plan.geoff[8] = (int)genv.size();
    hipError_t e;
    if ((e = upload(&b->d_xitems, &b->cap_xitems, xitems.data(), xitems.size())) != hipSuccess ||
        (e = upload(&b->d_genv, &b->cap_genv, genv.data(), genv.size())) != hipSuccess)
        return hip_report(e, "persistent engine plan upload");
    b->plan_stale = false;
    return HL_OK;
}

hl::BatchArgs batch_args(const hl_batch* b, const uint8_t* active)
{
    hl::BatchArgs a;
    a.envs = b->d_envs;
    a.step_items = b->d_step;
    a.vox_items = b->d_vox;
    a.parity = b->d_parity;
    a.time = b->d_time;
    a.active = active;
    a.partial = b->d_partial;
    return a;
}

int32_t enqueue(hl_batch* b, const uint8_t* active, int32_t nsteps, double* history, int32_t per_env, hipStream_t s)
{
    const hl::BatchArgs a = batch_args(b, active);
    const int nvox = (int)b->vox_items.size(), nstep = (int)b->step_items.size();
    for (int k = 0; k < nsteps; ++k) {
        if (!per_env) {
            if (nvox) hipLaunchKernelGGL(hl::k_bvox, dim3(nvox), dim3(256), 0, s, a, 0);
            hipLaunchKernelGGL(hl::k_bstep, dim3(nstep), dim3(hl::kThreads), 0, s, a, 0);
            hipLaunchKernelGGL(hl::k_badvance, dim3(b->n), dim3(64), 0, s, a, 0, history, b->n, k);
        } else {
            // the same kernels, one environment per launch
            for (int e = 0; e < b->n; ++e) {
                const hl::BatchEnv& be = b->envs[e];
                if (be.vox_count) hipLaunchKernelGGL(hl::k_bvox, dim3(be.vox_count), dim3(256), 0, s, a, be.vox_begin);
                hipLaunchKernelGGL(hl::k_bstep, dim3(be.step_count), dim3(hl::kThreads), 0, s, a, be.step_begin);
                hipLaunchKernelGGL(hl::k_badvance, dim3(1), dim3(64), 0, s, a, e, history, b->n, k);
            }
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_report(e, "batch kernel launch");
    }
    return HL_OK;
}

void drop_graph(hl_batch* b)
{
    if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
}

}  // namespace

extern "C" {

int32_t hl_batch_create(const hl_env* envs, int32_t n, hl_batch** out)
{
    if (!out) return report(HL_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!envs || n <= 0) return report(HL_E_INVALID, "a batch needs at least one environment");
    for (int e = 0; e < n; ++e) {
        int32_t rc = hl::check_env(&envs[e]);
        if (rc) return rc;
        const hl_env& x = envs[e];
        if (x.lattice != HL_D2Q9 || x.dtype != HL_F32 || x.force || x.nsolid > hl::kBatchMaxSolids) {
            std::string m = "batch environment " + std::to_string(e) +
                            ": batches hold D2Q9 FP32 lattices with no body-force field and at most 16 solids";
            return report(HL_E_UNSUPPORTED, m.c_str());
        }
    }
    hl_batch* b = new (std::nothrow) hl_batch();
    if (!b) return report(HL_E_NOMEM, "out of host memory");
    b->n = n;
    b->envs.resize(n);
    for (int e = 0; e < n; ++e) {
        hl::BatchEnv& be = b->envs[e];
        std::memset(&be, 0, sizeof(be));
        be.base = hl::env_args(&envs[e], 0);
        be.base.cut_bits = nullptr;
    }
    rebuild_items(b);
    std::vector<int32_t> zeros(n, 0);
    hipError_t he;
    int32_t rc = sync_device(b);
    if (!rc && ((he = upload(&b->d_parity, &b->cap_parity, zeros.data(), zeros.size())) != hipSuccess ||
                (he = upload(&b->d_time, &b->cap_time, zeros.data(), zeros.size())) != hipSuccess))
        rc = hip_report(he, "batch state allocation");
    if (rc) {
        hl_batch_destroy(b);
        return rc;
    }
    *out = b;
    return HL_OK;
}

void hl_batch_destroy(hl_batch* b)
{
    if (!b) return;
    drop_graph(b);
    if (b->cap_stream) hipStreamDestroy(b->cap_stream);
    hipFree(b->d_envs);
    hipFree(b->d_step);
    hipFree(b->d_vox);
    hipFree(b->d_parity);
    hipFree(b->d_time);
    hipFree(b->d_partial);
    hipFree(b->d_bar);
    hipFree(b->d_flag);
    hipFree(b->d_xitems);
    hipFree(b->d_genv);
    hipFree(b->d_bars);
    hipFree(b->d_xcc);
    delete b;
}

int32_t hl_batch_set_bodies(hl_batch* b, int32_t env, const uint8_t* mask, const int32_t box[6], const float* caps,
                            int32_t ncap, const float* poses, int32_t nposes)
{
    if (!b) return report(HL_E_INVALID, "batch is NULL");
    if (env < 0 || env >= b->n) return report(HL_E_INVALID, "environment index out of range");
    hl::BatchEnv& be = b->envs[env];
    hl::EnvArgs& a = be.base;
    if (!mask || !box) return report(HL_E_INVALID, "mask and box are required");
    if (box[0] < 0 || box[1] > a.nx || box[2] < 0 || box[3] > a.ny || box[0] > box[1] || box[2] > box[3])
        return report(HL_E_INVALID, "body box outside the grid");
    if (ncap < 0 || nposes < 0) return report(HL_E_INVALID, "negative table length");
    if (ncap > 0 && (!caps || nposes <= 0)) return report(HL_E_INVALID, "a capsule table needs caps and nposes > 0");
    if (poses && nposes <= 0) return report(HL_E_INVALID, "a pose table needs nposes > 0");
    if (a.nsolid > 0 && !poses && !a.solids) return report(HL_E_INVALID, "no pose table for the solids");
    a.mask = mask;
    a.box[0] = box[0]; a.box[1] = box[1]; a.box[2] = box[2]; a.box[3] = box[3];
    a.box[4] = 0;
    a.box[5] = 1;
    be.caps = ncap > 0 ? caps : nullptr;
    be.ncap = ncap;
    be.poses = poses;
    be.T = nposes;
    drop_graph(b);
    rebuild_items(b);
    return sync_device(b);
}

int32_t hl_batch_set_state(hl_batch* b, const int32_t* parity, const int32_t* time)
{
    if (!b || !parity || !time) return report(HL_E_INVALID, "batch, parity and time are required");
    for (int e = 0; e < b->n; ++e)
        if ((parity[e] != 0 && parity[e] != 1) || time[e] < 0) return report(HL_E_INVALID, "parity must be 0/1, time >= 0");
    hipError_t e;
    if ((e = hipMemcpy(b->d_parity, parity, b->n * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_time, time, b->n * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_report(e, "batch state upload");
    return HL_OK;
}

int32_t hl_batch_get_state(hl_batch* b, int32_t* parity, int32_t* time)
{
    if (!b || !parity || !time) return report(HL_E_INVALID, "batch, parity and time are required");
    hipError_t e;
    if ((e = hipMemcpy(parity, b->d_parity, b->n * sizeof(int32_t), hipMemcpyDeviceToHost)) != hipSuccess ||
        (e = hipMemcpy(time, b->d_time, b->n * sizeof(int32_t), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_report(e, "batch state download");
    return HL_OK;
}

int32_t hl_batch_step(hl_batch* b, const uint8_t* active, int32_t nsteps, double* force_history, int32_t per_env,
                      void* stream)
{
    if (!b) return report(HL_E_INVALID, "batch is NULL");
    if (nsteps < 0) return report(HL_E_INVALID, "nsteps must be >= 0");
    int32_t rc = sync_device(b);
    if (rc) return rc;
    return enqueue(b, active, nsteps, force_history, per_env, static_cast<hipStream_t>(stream));
}

int32_t hl_batch_capture(hl_batch* b, const uint8_t* active, int32_t nsteps, double* force_history)
{
    if (!b) return report(HL_E_INVALID, "batch is NULL");
    if (nsteps <= 0) return report(HL_E_INVALID, "nsteps must be > 0");
    int32_t rc = sync_device(b);
    if (rc) return rc;
    drop_graph(b);
    hipError_t e;
    if (!b->cap_stream && (e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking)) != hipSuccess)
        return hip_report(e, "hipStreamCreate");
    if ((e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return hip_report(e, "hipStreamBeginCapture");
    rc = enqueue(b, active, nsteps, force_history, 0, b->cap_stream);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->cap_stream, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_report(e, "hipStreamEndCapture");
    b->graph = g;
    if ((e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0)) != hipSuccess) {
        drop_graph(b);
        return hip_report(e, "hipGraphInstantiate");
    }
    return HL_OK;
}

int32_t hl_batch_replay(hl_batch* b, void* stream)
{
    if (!b) return report(HL_E_INVALID, "batch is NULL");
    if (!b->graph_exec) return report(HL_E_INVALID, "no captured graph (hl_batch_capture first)");
    if (b->dirty) return report(HL_E_INVALID, "bodies changed since the capture; capture again");
    hipError_t e = hipGraphLaunch(b->graph_exec, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_report(e, "hipGraphLaunch");
    return HL_OK;
}

int32_t hl_batch_run(hl_batch* b, const uint8_t* active, int32_t nsteps, double* force_history, int32_t per_env,
                     void* stream)
{
    if (!b) return report(HL_E_INVALID, "batch is NULL");
    if (nsteps < 0) return report(HL_E_INVALID, "nsteps must be >= 0");
    if (nsteps == 0) return HL_OK;
    for (int env = 0; env < b->n; ++env)
        if (b->envs[env].caps && b->envs[env].ncap > hl::kRunCaps) {
            std::string m = "environment " + std::to_string(env) + ": the persistent engine takes at most " +
                            std::to_string(hl::kRunCaps) + " capsules (hl_batch_step takes any number)";
            return report(HL_E_UNSUPPORTED, m.c_str());
        }
    int32_t rc = sync_device(b);
    if (rc) return rc;
    hipError_t e;
    if (!b->d_bar && (e = hipMalloc(&b->d_bar, sizeof(unsigned))) != hipSuccess) return hip_report(e, "hipMalloc");
    if (!b->d_flag && (e = hipMalloc(&b->d_flag, sizeof(int))) != hipSuccess) return hip_report(e, "hipMalloc");
    int dev = 0, cus = 0, occ = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess ||
        (e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, hl::k_brun, hl::kThreads, 0)) != hipSuccess)
        return hip_report(e, "persistent engine occupancy");
    if (occ < 1) return report(HL_E_UNSUPPORTED, "the persistent batch kernel does not fit on a CU");
    // at most two workgroups per CU: the occupancy calculator's four per CU
    // (8 waves of 8 a SIMD) were not all resident at once on MI355X, and the
    // barrier waits gave up; two per CU were
    occ = std::min(occ, 2);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const hl::BatchArgs a = batch_args(b, active);
    auto launch = [&](int i0, int ni, int e0, int ne) -> int32_t {
        if (ni <= 0) return HL_OK;
        // every workgroup resident (a cooperative launch); no more than the tiles
        const int grid = std::max(1, std::min(occ * cus, ni));
        hipLaunchKernelGGL(hl::k_brun_reset, dim3(1), dim3(1), 0, s, b->d_bar, b->d_flag);
        int i0v = i0, niv = ni, K = nsteps, ntot = b->n, e0v = e0, nev = ne, nsl = b->nslots;
        double* hist = force_history;
        unsigned* bar = b->d_bar;
        int* flag = b->d_flag;
        hl::BatchArgs av = a;
        void* args[] = {&av, &i0v, &niv, &K, &hist, &ntot, &e0v, &nev, &nsl, &bar, &flag};
        hipError_t le = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(hl::k_brun), dim3(grid),
                                                   dim3(hl::kThreads), args, 0, s);
        if (le != hipSuccess) return hip_report(le, "persistent batch launch");
        return HL_OK;
    };
    if (!per_env) {
        // every environment at once: the XCD-affine engine when the dealing
        // of workgroups to XCDs is the round-robin it assumes
        const int ni = (int)b->step_items.size();
        if (ni <= 0) return HL_OK;
        const int g8 = std::max(8, std::min(occ * cus, ni) / 8 * 8);
        if (g8 <= occ * cus) {
            if (b->plan_stale && (rc = build_plan(b)) != HL_OK) return rc;
            if (!b->d_bars && (e = hipMalloc(&b->d_bars, 8 * hl::kXcdBarStride * sizeof(unsigned))) != hipSuccess)
                return hip_report(e, "hipMalloc");
            if (b->xcd_grid != g8) {
                if (b->d_xcc && (e = hipFree(b->d_xcc)) != hipSuccess) return hip_report(e, "hipFree");
                b->d_xcc = nullptr;
                if ((e = hipMalloc(&b->d_xcc, (size_t)g8 * sizeof(int))) != hipSuccess) return hip_report(e, "hipMalloc");
                int* xcc = b->d_xcc;
                void* pargs[] = {&xcc};
                if ((e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(hl::k_xcc_ids), dim3(g8),
                                                    dim3(hl::kThreads), pargs, 0, s)) != hipSuccess)
                    return hip_report(e, "XCD dealing check launch");
                std::vector<int> ids((size_t)g8);
                if ((e = hipMemcpyAsync(ids.data(), b->d_xcc, (size_t)g8 * sizeof(int), hipMemcpyDeviceToHost, s)) !=
                        hipSuccess ||
                    (e = hipStreamSynchronize(s)) != hipSuccess)
                    return hip_report(e, "XCD dealing check");
                bool ok = true;
                for (int w = 0; w < g8; ++w) ok = ok && ids[(size_t)w] == w % 8;
                b->xcd_grid = g8;
                b->xcd_ok = ok;
            }
            if (b->xcd_ok) {
                hipLaunchKernelGGL(hl::k_brun_xcd_reset, dim3(1), dim3(64), 0, s, b->d_bars, b->d_flag);
                hl::BatchArgs av = a;
                const int* xi = b->d_xitems;
                const int* ge = b->d_genv;
                hl::XcdPlan plan = b->plan;
                int K = nsteps, ntot = b->n, nsl = b->nslots;
                double* hist = force_history;
                unsigned* bars = b->d_bars;
                int* flag = b->d_flag;
                unsigned zero = 0u;
                void* args[] = {&av, &xi, &plan, &ge, &K, &hist, &ntot, &nsl, &bars, &flag, &zero};
                hipError_t le = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(hl::k_brun_xcd), dim3(g8),
                                                           dim3(hl::kThreads), args, 0, s);
                if (le != hipSuccess) return hip_report(le, "persistent batch launch");
                return HL_OK;
            }
        }
        return launch(0, ni, 0, b->n);
    }
    for (int env = 0; env < b->n; ++env) {
        rc = launch(b->envs[env].step_begin, b->envs[env].step_count, env, 1);
        if (rc) return rc;
    }
    return HL_OK;
}

int32_t hl_batch_run_status(hl_batch* b, int32_t* timed_out)
{
    if (!b || !timed_out) return report(HL_E_INVALID, "batch and timed_out are required");
    *timed_out = 0;
    if (!b->d_flag) return HL_OK;
    int v = 0;
    hipError_t e = hipMemcpy(&v, b->d_flag, sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_report(e, "batch run status");
    *timed_out = v;
    return HL_OK;
}

int32_t hl_batch_size(const hl_batch* b, int32_t* nenv, int32_t* step_tiles, int32_t* force_tiles)
{
    if (!b) return report(HL_E_INVALID, "batch is NULL");
    if (nenv) *nenv = b->n;
    if (step_tiles) *step_tiles = (int32_t)b->step_items.size();
    if (force_tiles) *force_tiles = b->nslots;
    return HL_OK;
}

}  // extern "C"
