// Auxiliary kernels: uniform initialisation, capsule voxelizer, per-solid
// momentum-exchange force/torque and the FP64 mass sum.
//
// Forces and mass are reduced in FP64 in a fixed order (wavefront butterfly,
// then waves of a block in index order, then blocks in index order), so the
// result does not depend on scheduling: repeated runs are bitwise equal,
// unlike the reference's FP32 atomics (lbm_func.py:399-401,
// lbm_func_3d.py:510-511).
#pragma clang fp contract(off)

#include <cstring>

#include "gather_post.h"

#include "kernels.h"
#include "lbm_device.h"

namespace hl {

constexpr int kReduceBlock = 256;
constexpr int kWaves = kReduceBlock / 64;
constexpr int kMaxReduceBlocks = 1024;


// ---------------------------------------------------------------- init

template <int DT>
__global__ void k_fill(void* buf, int64_t stride, int64_t n, int nf, EnvArgs a, double m0, double m1, double m2,
                       double m3, double m4, double m5, double m6, double m7, double m8, double m9)
{
    using ST = typename Storage<DT>::type;
    ST* p = static_cast<ST*>(buf);
    const double m[10] = {m0, m1, m2, m3, m4, m5, m6, m7, m8, m9};
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t so = cell_of(a, c);
        for (int f = 0; f < nf; ++f) {
            const uint32_t key = (uint32_t)f * (uint32_t)n + (uint32_t)c;
            if (DT == HL_F64) Storage<DT>::template store<double>(p, f * stride + so, m[f], f, a.scale[f], a.offset[f], key);
            else Storage<DT>::template store<float>(p, f * stride + so, (float)m[f], f, a.scale[f], a.offset[f], key);
        }
    }
}

hipError_t init_uniform(int lattice, int dtype, const EnvArgs& a, void* buf, const double m[10], hipStream_t s)
{
    const int64_t n = (int64_t)a.nx * a.ny * a.nz;
    const int nf = lattice == HL_D2Q9 ? 6 : 10;
    const int blocks = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
    switch (dtype) {
    case HL_F32: hipLaunchKernelGGL(k_fill<HL_F32>, blocks, 256, 0, s, buf, a.stride, n, nf, a, m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9]); break;
    case HL_F64: hipLaunchKernelGGL(k_fill<HL_F64>, blocks, 256, 0, s, buf, a.stride, n, nf, a, m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9]); break;
    case HL_F16: hipLaunchKernelGGL(k_fill<HL_F16>, blocks, 256, 0, s, buf, a.stride, n, nf, a, m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9]); break;
    case HL_BF16: hipLaunchKernelGGL(k_fill<HL_BF16>, blocks, 256, 0, s, buf, a.stride, n, nf, a, m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9]); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------- state copy

// Streaming copy of 16-byte words: four non-temporal loads in flight per
// lane before their stores, over a large grid (the fastest of the shapes
// measured for a whole-state copy on MI355X: 6.4 TB/s, against 5.3-5.9 TB/s
// with one load in flight per lane).
typedef unsigned int copy_word __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy16(const copy_word* __restrict__ src, copy_word* __restrict__ dst,
                                                int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * 256;
    int64_t i = blockIdx.x * 256ll + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        copy_word v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
        for (int u = 0; u < 4; ++u) __builtin_nontemporal_store(v[u], dst + i + u * stride);
    }
    for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

__global__ void __launch_bounds__(256) k_copy1(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n)
{
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

hipError_t copy_state(int lattice, int dtype, const EnvArgs& a, hipStream_t s)
{
    const int64_t esize = dtype == HL_F64 ? 8 : (dtype == HL_F32 ? 4 : 2);
    const int64_t bytes = state_elems(a, lattice == HL_D2Q9 ? 6 : 10) * esize;
    const uint8_t* src = static_cast<const uint8_t*>(a.src);
    uint8_t* dst = static_cast<uint8_t*>(a.dst);
    int64_t body = 0;
    if ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) % 16 == 0) {
        body = bytes / 16 * 16;
        if (body > 0) {
            const int64_t n = body / 16, want = (n + 255) / 256;
            hipLaunchKernelGGL(k_copy16, (int)(want < 16384 ? want : 16384), 256, 0, s,
                               reinterpret_cast<const copy_word*>(src), reinterpret_cast<copy_word*>(dst), n);
        }
    }
    if (bytes > body) {
        const int64_t n = bytes - body, want = (n + 255) / 256;
        hipLaunchKernelGGL(k_copy1, (int)(want < 2048 ? want : 2048), 256, 0, s, src + body, dst + body, n);
    }
    return hipGetLastError();
}

// ----------------------------------------------------------- voxelizer

// Cell centre p is inside capsule (a, b, r) iff its squared distance to the
// segment ab is <= r*r; the lowest-index capsule containing p sets
// mask = solid_id + 1. FP32 without contraction, same order as the CPU
// voxelizer, so masks are bitwise identical.
__global__ void k_voxelize(uint8_t* mask, int nx, int ny, const float* caps, int ncap, int bx0, int by0, int bz0,
                           int bw, int bh, int bd)
{
    const int64_t n = (int64_t)bw * bh * bd;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int x = bx0 + (int)(t % bw);
        const int y = by0 + (int)((t / bw) % bh);
        const int z = bz0 + (int)(t / ((int64_t)bw * bh));
        const float px = (float)x, py = (float)y, pz = (float)z;
        uint8_t v = 0;
        for (int k = 0; k < ncap; ++k) {
            const float* C = caps + 8 * k;
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
            if (d2 <= C[6] * C[6]) { v = (uint8_t)((int)C[7] + 1); break; }
        }
        mask[(int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z)] = v;
    }
}

hipError_t voxelize(int lattice, const EnvArgs& a, uint8_t* mask, const float* caps, int ncap, const int32_t box[6],
                    hipStream_t s)
{
    const int bw = box[1] - box[0], bh = box[3] - box[2], bd = box[5] - box[4];
    if (bw <= 0 || bh <= 0 || bd <= 0) return hipSuccess;
    const int64_t n = (int64_t)bw * bh * bd;
    const int blocks = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
    hipLaunchKernelGGL(k_voxelize, blocks, 256, 0, s, mask, a.nx, a.ny, caps, ncap, box[0], box[2], box[4], bw, bh, bd);
    return hipGetLastError();
}

// Even-odd +x ray test of get_solid_boundary_img (envs/lbm/lbm_func.py:804-864,
// border thickness 0) for closed 2D polygons in lattice coordinates; the
// lowest-index polygon containing the cell centre sets mask = index + 1.
// FP32 without contraction, same order as the CPU voxelizer.
__global__ void k_voxelize_polygons(uint8_t* mask, int nx, const float* verts, const int32_t* counts, int npoly,
                                    int maxv, int bx0, int by0, int bw, int bh)
{
    const int64_t n = (int64_t)bw * bh;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int x = bx0 + (int)(t % bw), y = by0 + (int)(t / bw);
        const float px = (float)x, py = (float)y;
        uint8_t v = 0;
        for (int k = 0; k < npoly && !v; ++k) {
            const float* P = verts + 2 * (int64_t)maxv * k;
            const int nv = counts[k];
            int crossings = 0;
            for (int j = 0; j < nv; ++j) {
                const float x1 = P[2 * j], y1 = P[2 * j + 1];
                const int j2 = (j + 1) % nv;
                const float x2 = P[2 * j2], y2 = P[2 * j2 + 1];
                if ((y1 > py) != (y2 > py)) {
                    const float ix = x1 + (py - y1) * (x2 - x1) / (y2 - y1);
                    if (px < ix) ++crossings;
                }
            }
            if (crossings & 1) v = (uint8_t)(k + 1);
        }
        mask[(int64_t)x + (int64_t)nx * y] = v;
    }
}

hipError_t voxelize_polygons(const EnvArgs& a, uint8_t* mask, const float* verts, const int32_t* counts, int npoly,
                             int maxv, const int32_t box[6], hipStream_t s)
{
    const int bw = box[1] - box[0], bh = box[3] - box[2];
    if (bw <= 0 || bh <= 0) return hipSuccess;
    const int64_t n = (int64_t)bw * bh;
    const int blocks = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
    hipLaunchKernelGGL(k_voxelize_polygons, blocks, 256, 0, s, mask, a.nx, verts, counts, npoly, maxv, box[0], box[2],
                       bw, bh);
    return hipGetLastError();
}

// -------------------------------------------------- near-body fix-up

// Cells whose update can involve a cut link: interior cells within one cell
// of the mask box, plus the face cells that project onto them.
static void axis_range(int b0, int b1, int n, int& lo, int& hi)
{
    lo = b0 - 1 > 1 ? b0 - 1 : 1;
    hi = b1 + 1 < n - 1 ? b1 + 1 : n - 1;
    if (lo >= hi) { lo = hi = 0; return; }
    if (lo == 1) lo = 0;
    if (hi == n - 1) hi = n;
}

// Bodies that cut links: a mask, or capsules (which take precedence).
static bool has_bodies(const EnvArgs& a) { return a.mask != nullptr || a.caps != nullptr; }

static Region fixup_region(const EnvArgs& a, int lattice)
{
    Region r{0, 0, 0, 0, 0, 0};
    int x0, x1, y0, y1, z0 = 0, z1 = 1;
    axis_range(a.box[0], a.box[1], a.nx, x0, x1);
    axis_range(a.box[2], a.box[3], a.ny, y0, y1);
    if (lattice == HL_D3Q27) axis_range(a.box[4], a.box[5], a.nz, z0, z1);
    r.x0 = x0; r.y0 = y0; r.z0 = z0;
    r.w = x1 - x0; r.h = y1 - y0; r.d = z1 - z0;
    return r;
}

static int reduce_blocks(int64_t n, int block = kReduceBlock, int maxb = kMaxReduceBlocks)
{
    int64_t b = (n + block - 1) / block;
    if (b < 1) b = 1;
    return (int)(b < maxb ? b : maxb);
}




// ------------------------------------------------------------------------
// Near-body fix-up.
//
// A cell needs the wall rule when its projection onto the interior has a cut
// link (a link whose two ends differ in mask occupancy). Those cells are
// listed once per mask (hl_step builds the list at the start of each call):
//   k_cut_count  one wave per 64 region cells: ballot of "has a cut link";
//   k_cut_scan   one block: exclusive scan of the per-wave counts;
//   k_cut_write  each wave writes its cells in lane order at its offset,
// so the list is in region order whatever the scheduling. Per step,
// k_fixup runs a fixed grid of wavefronts over the list, two cells per wave
// and one lane per lattice direction: each lane pulls its population (the
// source cell's loads run in parallel across lanes); the group's first lane
// sums them in ascending direction order and collides exactly as the
// reference does, then writes the cell (overwriting the tiled kernel's
// value). Each lane also owns the momentum-exchange force of its own link;
// per solid the wave sums with a butterfly into a per-wave FP64 partial.

__device__ __forceinline__ bool region_cell_cut(const EnvArgs& a, const Region& rg, int t, int Q)
{
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int x = rg.x0 + t % rg.w;
    const int y = rg.y0 + (t / rg.w) % rg.h;
    const int z = rg.z0 + t / (rg.w * rg.h);
    const int px = min(max(x, 1), nx - 2), py = min(max(y, 1), ny - 2);
    const int pz = Q == 9 ? 0 : min(max(z, 1), nz - 2);
    if (Q == 27 && a.caps) {  // a crossing of any capsule surface
        // in FP32 or FP64, so the list holds the cells of either precision
        bool has = false;
        for (int i = 1; i < 27 && !has; ++i) {
            float t32;
            double t64;
            has = capsule_cut<float>(a.caps, a.ncap, a.nsolid, float(px), float(py), float(pz), D3CX[i], D3CY[i],
                                     D3CZ[i], t32) >= 0 ||
                  capsule_cut<double>(a.caps, a.ncap, a.nsolid, double(px), double(py), double(pz), D3CX[i],
                                      D3CY[i], D3CZ[i], t64) >= 0;
        }
        return has;
    }
    if (!in_box_halo(a, px, py, pz)) return false;
    const int64_t c = (int64_t)px + (int64_t)nx * ((int64_t)py + (int64_t)ny * pz);
    const int mv_c = a.mask[c] != 0;
    bool has = false;
    for (int i = 1; i < Q; ++i) {
        const int cx = Q == 9 ? D2CX[i] : D3CX[i], cy = Q == 9 ? D2CY[i] : D3CY[i], cz = Q == 9 ? 0 : D3CZ[i];
        has |= (a.mask[c - cx - (int64_t)nx * (cy + (int64_t)ny * cz)] != 0) != mv_c;
    }
    return has;
}

__global__ void __launch_bounds__(64) k_cut_count(EnvArgs a, Region rg, int Q, uint64_t* ballots, int* counts)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    const bool has = t < (int)rg.n() && region_cell_cut(a, rg, t, Q);
    const uint64_t b = __ballot(has);
    if (threadIdx.x == 0) {
        ballots[blockIdx.x] = b;
        counts[blockIdx.x] = __popcll(b);
    }
}

// Exclusive scan of counts[0..n) into offsets, total into *total (one block).
__global__ void __launch_bounds__(1024) k_cut_scan(const int* counts, int n, int* offsets, int* total)
{
    __shared__ int part[1024];
    const int tid = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b0 = tid * per, b1 = min(b0 + per, n);
    int s = 0;
    for (int b = b0; b < b1; ++b) s += counts[b];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
        const int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = tid > 0 ? part[tid - 1] : 0;
    for (int b = b0; b < b1; ++b) {
        offsets[b] = run;
        run += counts[b];
    }
    if (tid == 1023) *total = part[1023];
}

__global__ void __launch_bounds__(64) k_cut_write(const uint64_t* ballots, const int* offsets, int nblk, int* list)
{
    const int blk = blockIdx.x, lane = threadIdx.x;
    if (blk >= nblk) return;
    const uint64_t b = ballots[blk];
    if ((b >> lane) & 1ull) list[offsets[blk] + __popcll(b & ((1ull << lane) - 1ull))] = blk * 64 + lane;
}

// Per-column form of the same bitmap, for kernels that march along z: for
// region column (rx, ry), word j of cols[(ry * w + rx) * nw ...] holds bit k
// for region plane 64 j + k.
__global__ void __launch_bounds__(64) k_cut_columns(const uint64_t* ballots, Region rg, int nw, uint64_t* cols)
{
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= rg.w * rg.h) return;
    const int64_t layer = (int64_t)rg.w * rg.h;
    for (int j = 0; j < nw; ++j) {
        uint64_t word = 0;
        for (int k = 0; k < 64 && 64 * j + k < rg.d; ++k) {
            const int64_t t = c + layer * (64 * j + k);
            word |= ((ballots[t >> 6] >> (t & 63)) & 1ull) << k;
        }
        cols[(int64_t)c * nw + j] = word;
    }
}

template <typename T, int DT, int Q>
__global__ void __launch_bounds__(64) k_fixup(EnvArgs a, Region rg, const int* list, const int* count, int store,
                                               double* partial)
{
    __shared__ T pops[2][32];
    __shared__ T sums[2][16];
    __shared__ double acc[HL_MAX_SOLIDS * 6];
    const int lane = threadIdx.x;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int ncut = *count;
    const int h = lane >> 5, i = lane & 31;
    if (partial)
        for (int k = lane; k < a.nsolid * 6; k += 64) acc[k] = 0.0;
    for (int base = blockIdx.x * 2; base < ncut; base += gridDim.x * 2) {
        const int e = base + h;
        const bool cell = e < ncut;
        const bool act = cell && i < Q;
        int x = 0, y = 0, z = 0, px = 0, py = 0, pz = 0, mval_c = 0, sid = -1;
        int64_t c = 0;
        T F[6] = {0, 0, 0, 0, 0, 0};
        if (cell) {
            const int tt = list[e];
            x = rg.x0 + tt % rg.w;
            y = rg.y0 + (tt / rg.w) % rg.h;
            z = rg.z0 + tt / (rg.w * rg.h);
            px = min(max(x, 1), nx - 2);
            py = min(max(y, 1), ny - 2);
            pz = Q == 9 ? 0 : min(max(z, 1), nz - 2);
            c = (int64_t)px + (int64_t)nx * ((int64_t)py + (int64_t)ny * pz);
            mval_c = a.caps ? 0 : a.mask[c];
        }
        if (act) {
            T mc[10];
            load_cell<T, DT>(a, cell_at(a, px, py, pz), Q == 9 ? 6 : 10, mc);
            WallPoint<T> w{-1, {T(0), T(0), T(0)}};
            const T pop = Q == 9 ? d2q9_pull<T, DT>(a, px, py, i, mc, true, mval_c)
                                 : d3q27_pull<T, DT>(a, px, py, pz, i, mc, true, mval_c, nullptr, &w);
            pops[h][i] = pop;
            // force of this lane's link (interior cells only: a face cell repeats its projection)
            if (Q == 27 && a.caps) {
                if (partial && px == x && py == y && pz == z && w.sid >= 0) {
                    sid = w.sid;
                    link_force_at<T>(a, mc, i, sid, w.h[0], w.h[1], w.h[2], F, nullptr);
                }
            } else if (partial && i > 0 && px == x && py == y && pz == z) {
                const int cx = Q == 9 ? D2CX[i] : D3CX[i], cy = Q == 9 ? D2CY[i] : D3CY[i], cz = Q == 9 ? 0 : D3CZ[i];
                const int mv = a.mask[c - cx - (int64_t)nx * (cy + (int64_t)ny * cz)];
                if ((mv != 0) != (mval_c != 0)) {
                    const int s2 = mval_c ? mval_c - 1 : mv - 1;
                    if (s2 < a.nsolid) {
                        sid = s2;
                        link_force<T, DT, Q>(a, mc, x, y, z, i, sid, F);
                    }
                }
            }
        }
        __syncthreads();
        if (Q == 27 && store && cell && i < 16) sums[h][i] = d3q27_group_sum<T>(pops[h], i);
        __syncthreads();
        if (store && cell && i == 0) {
            using ST = typename Storage<DT>::type;
            ST* dst = static_cast<ST*>(a.dst);
            const int64_t d = cell_at(a, x, y, z);
            if (Q == 9) {
                T Fx = 0, Fy = 0, m[6];
                if (a.force) { Fx = T(a.force[c]); Fy = T(a.force[a.fs + c]); }
                d2q9_collide<T>(a, pops[h], Fx, Fy, m);
                if (px != x || py != y) bc_chain<T, 2>(m, x, y, 0, nx, ny, 1, a.bc_type, a.bc_value);
#pragma unroll
                for (int f = 0; f < 6; ++f)
                    Storage<DT>::template store<T>(dst, f * a.stride + d, m[f], f, a.scale[f], a.offset[f],
                                                   HL_ROUNDING_KEY(a, f, x, y, z));
            } else {
                T Fx = 0, Fy = 0, Fz = 0, m[10];
                if (a.force) { Fx = T(a.force[c]); Fy = T(a.force[a.fs + c]); Fz = T(a.force[2 * a.fs + c]); }
                d3q27_collide_sums<T>(a, sums[h], Fx, Fy, Fz, m);
                if (px != x || py != y || pz != z) bc_chain<T, 3>(m, x, y, z, nx, ny, nz, a.bc_type, a.bc_value);
#pragma unroll
                for (int f = 0; f < 10; ++f)
                    Storage<DT>::template store<T>(dst, f * a.stride + d, m[f], f, a.scale[f], a.offset[f],
                                                   HL_ROUNDING_KEY(a, f, x, y, z));
            }
        }
        if (partial) wave_solid_sums(sid, F, kRow6, lane, acc);
        __syncthreads();  // pops are rewritten by the next pair
    }
    if (!partial) return;
    __syncthreads();
    for (int k = lane; k < a.nsolid * 6; k += 64) partial[(int64_t)blockIdx.x * a.nsolid * 6 + k] = acc[k];
}

// out[k] = sum over blocks b of partial[b][k]: each lane sums a fixed strided
// subset in order (lane_rows_sum), then a wavefront butterfly; one wave per
// output.
__global__ void __launch_bounds__(64) k_sum_partials(const double* partial, int nblocks, int width, double* out)
{
    const int k = blockIdx.x;
    const double v = wave_sum(lane_rows_sum(partial, nblocks, width, k, threadIdx.x));
    if (threadIdx.x == 0) out[k] = v;
}

__global__ void k_zero(double* out, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = 0.0;
}

constexpr int kFixupWaves = 1024;

struct FixupLayout {
    size_t partial, ballots, counts, offsets, total, list, cols, bytes;
    int nblk, colwords;
};

static FixupLayout fixup_layout(const Region& rg)
{
    FixupLayout L;
    const int64_t n = rg.n();
    L.nblk = (int)((n + 63) / 64);
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off = (off + bytes + 255) / 256 * 256; return o; };
    L.partial = take((size_t)kFixupWaves * HL_MAX_SOLIDS * 6 * sizeof(double));
    L.ballots = take((size_t)(L.nblk > 0 ? L.nblk : 1) * sizeof(uint64_t));
    L.counts = take((size_t)(L.nblk > 0 ? L.nblk : 1) * sizeof(int));
    L.offsets = take((size_t)(L.nblk > 0 ? L.nblk : 1) * sizeof(int));
    L.total = take(sizeof(int));
    L.list = take((size_t)(n > 0 ? n : 1) * sizeof(int));
    L.colwords = (rg.d + 63) / 64;
    L.cols = take((size_t)(n > 0 ? (int64_t)rg.w * rg.h * L.colwords : 1) * sizeof(uint64_t));
    L.bytes = off;
    return L;
}

size_t forces_workspace_bytes(const EnvArgs& a, int lattice)
{
    return fixup_layout(fixup_region(a, lattice)).bytes;
}

void cut_bits_view(int lattice, EnvArgs& a, void* ws)
{
    const Region rg = fixup_region(a, lattice);
    if (!has_bodies(a) || rg.n() == 0 || ws == nullptr) {
        a.cut_bits = nullptr;
        a.cut_cols = nullptr;
        return;
    }
    const FixupLayout L = fixup_layout(rg);
    a.cut_bits = reinterpret_cast<const uint64_t*>(static_cast<char*>(ws) + L.ballots);
    a.cut_cols = reinterpret_cast<const uint64_t*>(static_cast<char*>(ws) + L.cols);
    a.cut_colwords = L.colwords;
    a.cut_origin[0] = rg.x0; a.cut_origin[1] = rg.y0; a.cut_origin[2] = rg.z0;
    a.cut_extent[0] = rg.w; a.cut_extent[1] = rg.h; a.cut_extent[2] = rg.d;
}

hipError_t cut_list(int lattice, const EnvArgs& a, void* ws, hipStream_t s)
{
    const Region rg = fixup_region(a, lattice);
    if (!has_bodies(a) || rg.n() == 0) return hipSuccess;
    if (rg.n() > (int64_t)1 << 30) return hipErrorInvalidValue;
    const FixupLayout L = fixup_layout(rg);
    char* base = static_cast<char*>(ws);
    uint64_t* ballots = reinterpret_cast<uint64_t*>(base + L.ballots);
    int* counts = reinterpret_cast<int*>(base + L.counts);
    int* offsets = reinterpret_cast<int*>(base + L.offsets);
    int* total = reinterpret_cast<int*>(base + L.total);
    int* list = reinterpret_cast<int*>(base + L.list);
    const int Q = lattice == HL_D2Q9 ? 9 : 27;
    hipLaunchKernelGGL(k_cut_count, L.nblk, 64, 0, s, a, rg, Q, ballots, counts);
    hipLaunchKernelGGL(k_cut_scan, 1, 1024, 0, s, counts, L.nblk, offsets, total);
    hipLaunchKernelGGL(k_cut_write, L.nblk, 64, 0, s, ballots, offsets, L.nblk, list);
    uint64_t* cols = reinterpret_cast<uint64_t*>(base + L.cols);
    hipLaunchKernelGGL(k_cut_columns, (rg.w * rg.h + 63) / 64, 64, 0, s, ballots, rg, L.colwords, cols);
    return hipGetLastError();
}

template <typename T, int DT>
static hipError_t launch_fixup(int lattice, const EnvArgs& a, const Region& rg, const int* list, const int* count,
                               int store, double* part, int nb, hipStream_t s)
{
    if (lattice == HL_D2Q9) hipLaunchKernelGGL((k_fixup<T, DT, 9>), nb, 64, 0, s, a, rg, list, count, store, part);
    else hipLaunchKernelGGL((k_fixup<T, DT, 27>), nb, 64, 0, s, a, rg, list, count, store, part);
    return hipGetLastError();
}

hipError_t fixup(int lattice, int dtype, const EnvArgs& a, bool store, double* out, void* ws, hipStream_t s)
{
    const int width = a.nsolid * 6;
    const bool want_forces = out != nullptr && width > 0;
    const Region rg = fixup_region(a, lattice);
    if (!has_bodies(a) || rg.n() == 0) {
        if (want_forces) {
            hipLaunchKernelGGL(k_zero, (width + 63) / 64, 64, 0, s, out, width);
            return hipGetLastError();
        }
        return hipSuccess;
    }
    if (!store && !want_forces) return hipSuccess;
    const FixupLayout L = fixup_layout(rg);
    char* base = static_cast<char*>(ws);
    const int* total = reinterpret_cast<const int*>(base + L.total);
    const int* list = reinterpret_cast<const int*>(base + L.list);
    double* part = want_forces ? reinterpret_cast<double*>(base + L.partial) : nullptr;
    int64_t pairs = (rg.n() + 1) / 2;
    const int nb = (int)(pairs < kFixupWaves ? (pairs > 0 ? pairs : 1) : kFixupWaves);
    hipError_t e;
    switch (dtype) {
    case HL_F32: e = launch_fixup<float, HL_F32>(lattice, a, rg, list, total, store, part, nb, s); break;
    case HL_F64: e = launch_fixup<double, HL_F64>(lattice, a, rg, list, total, store, part, nb, s); break;
    case HL_F16: e = launch_fixup<float, HL_F16>(lattice, a, rg, list, total, store, part, nb, s); break;
    case HL_BF16: e = launch_fixup<float, HL_BF16>(lattice, a, rg, list, total, store, part, nb, s); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess || !want_forces) return e;
    hipLaunchKernelGGL(k_sum_partials, width, 64, 0, s, part, nb, width, out);
    return hipGetLastError();
}

FixArgs fixup_args(int lattice, const EnvArgs& a, void* ws, bool forces, int parity)
{
    FixArgs f;
    std::memset(&f, 0, sizeof(f));
    const Region rg = fixup_region(a, lattice);
    if (!has_bodies(a) || rg.n() == 0 || ws == nullptr) return f;
    const FixupLayout L = fixup_layout(rg);
    char* base = static_cast<char*>(ws);
    f.rg = rg;
    f.list = reinterpret_cast<const int*>(base + L.list);
    f.count = reinterpret_cast<const int*>(base + L.total);
    f.width = a.nsolid * 6;
    // two banks of rows, by step parity, so that a launch can sum the
    // previous step's rows while it writes its own
    f.max_waves = kFixupWaves / 2;
    if (forces && f.width > 0)
        f.partial = reinterpret_cast<double*>(base + L.partial) + (size_t)(parity & 1) * f.max_waves * f.width;
    return f;
}

hipError_t sum_partials(const double* partial, int rows, int width, double* out, hipStream_t s)
{
    if (width <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sum_partials, width, 64, 0, s, partial, rows, width, out);
    return hipGetLastError();
}

hipError_t forces(int lattice, int dtype, const EnvArgs& a, double* out, void* ws, hipStream_t s)
{
    hipError_t e = cut_list(lattice, a, ws, s);
    if (e != hipSuccess) return e;
    return fixup(lattice, dtype, a, false, out, ws, s);
}

// ---------------------------------------------------------------- mass

size_t mass_workspace_bytes(const EnvArgs& a, int lattice) { return (size_t)kMaxReduceBlocks * sizeof(double); }

template <int DT>
__global__ void __launch_bounds__(kReduceBlock) k_mass(EnvArgs a, int64_t n, double* partial)
{
    __shared__ double wsum[kWaves];
    using ST = typename Storage<DT>::type;
    const ST* src = static_cast<const ST*>(a.src);
    double v = 0.0;
    for (int64_t c = blockIdx.x * (int64_t)kReduceBlock + threadIdx.x; c < n; c += (int64_t)gridDim.x * kReduceBlock)
        v += Storage<DT>::template load<double>(src, cell_of(a, c), 0, a.inv_scale[0], a.offset[0]);
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kWaves; ++w) t += wsum[w];
        partial[blockIdx.x] = t;
    }
}

hipError_t mass(int lattice, int dtype, const EnvArgs& a, double* out, void* ws, hipStream_t s)
{
    const int64_t n = (int64_t)a.nx * a.ny * a.nz;
    const int nb = reduce_blocks(n);
    double* part = static_cast<double*>(ws);
    switch (dtype) {
    case HL_F32: hipLaunchKernelGGL(k_mass<HL_F32>, nb, kReduceBlock, 0, s, a, n, part); break;
    case HL_F64: hipLaunchKernelGGL(k_mass<HL_F64>, nb, kReduceBlock, 0, s, a, n, part); break;
    case HL_F16: hipLaunchKernelGGL(k_mass<HL_F16>, nb, kReduceBlock, 0, s, a, n, part); break;
    case HL_BF16: hipLaunchKernelGGL(k_mass<HL_BF16>, nb, kReduceBlock, 0, s, a, n, part); break;
    default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sum_partials, 1, 64, 0, s, part, nb, 1, out);
    return hipGetLastError();
}

}  // namespace hl
