// Batched D2Q9 engine: many resident environments of uneven grid sizes
// advanced together, one launch per kernel kind per step for all of them.
//
// Work-lists map each wavefront to an (environment, strip). Per-environment state
// that changes during an advance (ping-pong parity, kinematics time index)
// lives in device memory and is advanced by the last kernel of each step, so
// a multi-step advance needs no host involvement and can be captured in a
// graph and replayed.
//
// Per step, for all active environments:
//   k_bvox      the body masks rebuilt in the body boxes from the capsule
//               tables at each environment's time t
//   k_bstep     the D2Q9 update in strips of up to 62 columns, one wavefront
//               each marching up its rows (march): every source row's
//               populations rebuilt once and passed between lanes, the
//               reference's operation order, the mask wall rule. The strips
//               that reach a body box run the masked march (k_bstep<true>,
//               on the batch's side stream), and those near the body reduce
//               the per-solid momentum-exchange force of the state they read
//               into an FP64 partial; the others the plain march
//               (k_bstep<false>), which reads no mask
//   k_badvance  (one environment per launch) per environment: fixed-order
//               sum of its strips' partials into the force history, then
//               flip the parity and advance time. The batched schedule runs
//               step k of a call k steps past each environment's parity and
//               time instead, sums step k's rows in step k + 1's plain-strip
//               launch and closes the call with k_bclose (the last step's
//               sums, parity and time advanced once).
// The same kernels can be launched one environment at a time (the
// per-environment baseline); every cell's arithmetic and every sum's order
// are the same, so the two schedules give bitwise equal results.
//
// The persistent engine (k_brun, hl_batch_run) advances every environment K
// steps in one cooperative launch: the wavefronts deal out the strips of
// every environment (the march with the voxelizer, as in k_bstep), then a
// barrier ends the step; the first pass of step k+1 sums step k's force rows
// into the history. The cell arithmetic, the strips and the summation order
// are those of the three-kernel schedule, so states and forces agree with it
// bit for bit.
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
// step item: a strip of up to kSW columns and kSH rows, one wavefront; lane l
// holds column x0 - 1 + l, so lanes 1..w are the strip's cells and lanes 0
// and w + 1 its halo columns
constexpr int kSW = 62, kSH = 16, kSHB = 4;  // default strip heights: kSH rows, kSHB at a body
constexpr int kStepWaves = 4;                  // k_bstep: one step item per wavefront
constexpr int kThreads = 512, kWaves = kThreads / 64;  // persistent engine workgroup
constexpr int kVoxW = 16;         // voxelizer item: 16 x 16 cells
constexpr int kRunCaps = 32;  // capsules per environment the persistent engine stages in LDS
constexpr int kPartialWidth = kBatchMaxSolids * 3;  // Fx, Fy, Tz per solid

struct BatchEnv {
    EnvArgs base;         // arguments for parity 0 (moments[0] -> moments[1])
    const float* caps;    // [T][ncap][8] capsule table (null: static mask)
    const float* poses;   // [T][nsolid][8] pose table (null: static poses in base.solids)
    int32_t ncap, T;
    int32_t vox_begin, vox_count;    // voxelizer items
    int32_t step_begin, step_count;    // plain strips
    int32_t mstep_begin, mstep_count;  // strips that reach the body box (masked march)
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
__global__ void __launch_bounds__(256) k_bvox(BatchArgs b, int item0, int k0)
{
    const BatchItem it = b.vox_items[item0 + blockIdx.x];
    if (!env_active(b, it.env)) return;
    const BatchEnv& be = b.envs[it.env];
    if (!be.caps || be.T <= 0) return;
    const float* caps = be.caps + (int64_t)((b.time[it.env] + k0) % be.T) * be.ncap * 8;
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

// Raw buffer access to one environment's fields: the lane offset selects the
// column (kOOB: loads read 0, the range check drops stores), the scalar
// offset the field and row.
constexpr uint32_t kOOB = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t batch_rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}

// Wavefront lane shifts (DPP): lane l receives lane l - 1's value (from_lower;
// lane 0 receives 0) or lane l + 1's (from_upper; lane 63 receives 0).
__device__ __forceinline__ int from_lower(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int from_upper(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ float from_lower(float v) { return __builtin_bit_cast(float, from_lower(__builtin_bit_cast(int, v))); }
__device__ __forceinline__ float from_upper(float v) { return __builtin_bit_cast(float, from_upper(__builtin_bit_cast(int, v))); }

// Slot n of a list held across the lanes of a wavefront (six moments and a
// row) takes lane src's moments m and row y.
__device__ __forceinline__ void list_put(float* L, int& Ly, int n, const float* m, int src, int y)
{
    const bool mine = (int)(threadIdx.x & 63) == n;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const float v = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, m[f]), src));
        L[f] = mine ? v : L[f];
    }
    Ly = mine ? y : Ly;
}

// Whether strip `it` reads mask values: its source rows and columns reach
// the body box (the mask is 0 outside it).
__device__ __forceinline__ bool strip_masked(const BatchEnv& be, const BatchItem& it)
{
    const int x0 = uni(it.x0), w = uni(it.w), ny = uni(be.base.ny);
    const int ya = max(uni(it.y0), 1), yb = min(uni(it.y0) + uni(it.h), ny - 1);
    return uni(be.base.mask) != nullptr && x0 - 2 <= uni(be.base.box[1]) && x0 + w + 1 >= uni(be.base.box[0]) - 1 &&
           ya - 2 <= uni(be.base.box[3]) && yb + 1 >= uni(be.base.box[2]) - 1;
}

// One step of step item `it` (a strip of environment be) at buffer parity par
// and time t, by one wavefront, its force row (if any) into part. Lane l holds
// column x = x0 - 1 + l; the wavefront marches up the strip's rows, and every
// source row rebuilds its 9 outgoing populations once (d2q9_pops, the
// reference's per-link reconstruction and clamp), shifting those that move in
// x by one lane; a row's populations are pulled by the rows below, at and
// above it over three consecutive steps. Each cell then takes the reference's
// stream_and_collide (lbm_func.py:322-467) on those populations or, MASKED,
// the wall rule from its own moments on a cut link. Face cells take their
// interior projection's update and then the face chain: the face columns
// (one cell per row) are collected across the lanes and chained together at
// the end, with the face rows. The loads of a row are issued three rows
// ahead (a ring of four buffers, the march unrolled by four so that each
// load lands in the buffer it is consumed from). VOX: the source rows' mask
// values come from the capsule table at t (the strip's own cells written to
// the mask) instead of the mask.
template <bool VOX, bool MASKED, int NW>
__device__ __forceinline__ void march(const BatchEnv& be, const BatchItem& it, int par, int t, double* part)
{
    const int lane = (int)(threadIdx.x & 63);
    // every value below is the same across the wavefront: moved to scalar
    // registers (buffer descriptors built from vector registers would make
    // the compiler loop over the lanes around every load and store)
    const int nx = uni(be.base.nx), ny = uni(be.base.ny), nsolid = uni(be.base.nsolid);
    const uint32_t fbytes = (uint32_t)uni((int)be.base.stride) * 4u;  // below 2^31 / 6 (checked on the host)
    const void* src = uni(par ? be.base.dst : be.base.src);
    void* dst = uni(const_cast<void*>(par ? be.base.src : be.base.dst));
    const auto rs = batch_rsrc(src, 6u * fbytes);
    const auto rd = batch_rsrc(dst, 6u * fbytes);
    uint8_t* mask = uni(const_cast<uint8_t*>(be.base.mask));
    const int x0 = uni(it.x0), w = uni(it.w), y0 = uni(it.y0), h = uni(it.h), slot = uni(it.slot);
    const int bx0 = uni(be.base.box[0]), bx1 = uni(be.base.box[1]), by0 = uni(be.base.box[2]), by1 = uni(be.base.box[3]);
    const int T = uni(be.T), ncap = uni(be.ncap);
    const float* bcaps = uni(be.caps);
    const float* bposes = uni(be.poses);
    // (more capsules than the LDS table holds: the mask, written by k_bvox)
    const bool vox = VOX && MASKED && bcaps != nullptr && T > 0 && ncap <= kRunCaps;
    const float* gposes = bposes && T > 0 ? bposes + (int64_t)(t % T) * nsolid * 8 : uni(be.base.solids);
    // MASKED: this wavefront's copies of the pose table at t and (VOX) the
    // capsule table at t in LDS, read by the wall rule, the forces and the
    // voxelizer (one round of loads instead of one round trip per link)
    __shared__ float spose[NW][kBatchMaxSolids * 8];
    __shared__ float scaps[NW][kRunCaps * 8];
    // MASKED: a row's cut links listed across the wavefront (lane | link << 6
    // | (solid + 1) << 10), the cells' moments and the links' populations
    __shared__ uint32_t clist[NW][8 * 64];
    __shared__ float cmom[NW][6][64];
    __shared__ float cpop[NW][9][64];
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const float* poses = gposes;
    const float* caps = nullptr;
    if (MASKED) {
        if (gposes && lane < nsolid * 8) spose[wave][lane] = as_global(gposes)[lane];
        if (gposes && lane + 64 < nsolid * 8) spose[wave][lane + 64] = as_global(gposes)[lane + 64];
        poses = gposes ? &spose[wave][0] : nullptr;
        if (vox) {
            const float* gcaps = bcaps + (int64_t)(t % T) * ncap * 8;
#pragma unroll
            for (int k = 0; k < kRunCaps * 8; k += 64)
                if (k + lane < ncap * 8) scaps[wave][k + lane] = as_global(gcaps)[k + lane];
            caps = &scaps[wave][0];
        }
    }
    // the fields the cell functions read (the viscosity, the pose table, the
    // solid count, the width and the face conditions)
    EnvArgs ea = {};
    ea.nx = nx;
    ea.ny = ny;
    ea.nsolid = nsolid;
    ea.solids = poses;
    ea.nu = __builtin_bit_cast(float, uni(__builtin_bit_cast(int, be.base.nu)));
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        ea.bc_type[k] = uni(be.base.bc_type[k]);
#pragma unroll
        for (int c = 0; c < 3; ++c) ea.bc_value[k][c] = __builtin_bit_cast(float, uni(__builtin_bit_cast(int, be.base.bc_value[k][c])));
    }
    const int x = x0 - 1 + lane;
    const bool xin = x >= 0 && x < nx;
    const uint32_t vcol = xin ? (uint32_t)x * 4u : kOOB;
    const bool cell = lane >= 1 && lane <= w;          // a cell of this strip
    const bool inner = cell && x >= 1 && x <= nx - 2;  // an interior column
    const uint32_t vst = inner ? (uint32_t)x * 4u : kOOB;
    const int ya = max(y0, 1), yb = min(y0 + h, ny - 1);  // interior rows [ya, yb)
    const bool near = slot >= 0;
    const auto rm = batch_rsrc(mask, MASKED ? (uint32_t)nx * (uint32_t)ny : 0u);
    const uint32_t vmc = (MASKED && !vox && xin) ? (uint32_t)x : kOOB;  // mask bytes read from memory
    const bool xbox = xin && x >= bx0 && x < bx1;
    double acc = 0.0;  // lane 3 s + q: component q of solid s's force row
    if (ya >= yb) {  // no interior row (the host never builds such a strip)
        if (near && lane < kPartialWidth) as_global(part)[(int64_t)slot * kPartialWidth + lane] = acc;
        return;
    }

    auto load_row = [&](int r, float* M, int& K) {
        const uint32_t so = (uint32_t)(r * nx) * 4u;
#pragma unroll
        for (int f = 0; f < 6; ++f)
            M[f] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)vcol, (int)(so + f * fbytes), 0));
        if (MASKED) K = (int)__builtin_amdgcn_raw_buffer_load_b8(rm, (int)vmc, r * nx, 0);
    };
    auto store_row = [&](int r, uint32_t v, const float* m) {
        const uint32_t so = (uint32_t)(r * nx) * 4u;
#pragma unroll
        for (int f = 0; f < 6; ++f)
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, m[f]), rd, (int)v, (int)(so + f * fbytes), 0);
    };
    // mask values of source row r (VOX: from the capsules, the strip's own cells written)
    auto row_mask = [&](int r, int K) -> int {
        if (!vox) return K;
        int k = 0;
        if (r >= by0 && r < by1 && xbox) {
            k = capsule_voxel(caps, ncap, x, r);
            if (cell && r >= y0 && r < y0 + h) as_global(mask)[x + nx * r] = (uint8_t)k;
        }
        return k;
    };
    // a source row's populations as the lanes pull them: kept (cy = 0), sent
    // down (cy = -1, pulled by the row below) and up (cy = 1, by the row above)
    auto send = [&](const float* M, float* dn, float* md, float* up) {
        float f[9];
        d2q9_pops<float>(M, f);
        md[0] = f[0];
        md[1] = from_lower(f[1]);
        md[2] = from_upper(f[2]);
        up[0] = f[3];
        up[1] = from_lower(f[5]);
        up[2] = from_upper(f[6]);
        dn[0] = f[4];
        dn[1] = from_upper(f[7]);
        dn[2] = from_lower(f[8]);
    };

    // the march's carried state: moments of row y (My), the populations row
    // y keeps (md) and sends up (uy), those row y - 1 sends up (um); the mask
    // values of rows y - 1 and y and of their lower and upper neighbours; the
    // updates of the first and last rows (the face rows' projections)
    float My[6], md[3], um[3], uy[3], mfirst[6], mlast[6];
    int Km = 0, KmL = 0, KmU = 0, Ky = 0, KyL = 0, KyU = 0;
    // face columns: x = 0 from the projection x = 1 (lane 2), x = nx - 1 from
    // x = nx - 2 (lane nx - x0 - 1); slot n of a list in lane n
    const bool left = x0 == 0, right = x0 + w == nx;
    const int lsrc = 2, rsrc = nx - x0 - 1;
    float Lf[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, Rf[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    int Ly = 0, Ry = 0, nl = 0, nr = 0;

    // near the body: each lane's FP64 force sums in up to kSlots solid slots
    // (solids in order of first appearance); a slot's sums go through a
    // wavefront butterfly into acc when the slots run out and at the end
    constexpr int kSlots = 4;
    double fa0[kSlots], fa1[kSlots], fa2[kSlots];
    uint32_t ssl = 0xffffffffu;  // the slots' solids, one byte each (0xff: free)
    int nsl = 0;
#pragma unroll
    for (int j = 0; j < kSlots; ++j) fa0[j] = fa1[j] = fa2[j] = 0.0;
    auto flush = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < kSlots; ++j)
            if (j < nsl) {
                const int sv = (int)((ssl >> (8 * j)) & 0xffu);
                const double s0 = bwave_sum(fa0[j]), s1 = bwave_sum(fa1[j]), s2 = bwave_sum(fa2[j]);
                const int q = lane - 3 * sv;
                acc = q == 0 ? acc + s0 : q == 1 ? acc + s1 : q == 2 ? acc + s2 : acc;
                fa0[j] = fa1[j] = fa2[j] = 0.0;
            }
        nsl = 0;
        ssl = 0xffffffffu;
    };
    auto slot_of = [&](int sv) __attribute__((always_inline)) -> int {
        int j = -1;
#pragma unroll
        for (int jj = kSlots - 1; jj >= 0; --jj)
            if ((int)((ssl >> (8 * jj)) & 0xffu) == sv) j = jj;
        if (j < 0) {
            if (nsl == kSlots) flush();
            j = nsl++;
            ssl = (ssl & ~(0xffu << (8 * j))) | ((uint32_t)sv << (8 * j));
        }
        return j;
    };

    auto dest_row = [&](int y, const float* Mn, int Kraw, float* mout) {
        float dn[3], mdn[3], uyn[3];
        send(Mn, dn, mdn, uyn);
        int Kn = 0, KnL = 0, KnU = 0;
        if (MASKED) {
            Kn = row_mask(y + 1, Kraw);
            KnL = from_lower(Kn);
            KnU = from_upper(Kn);
        }
        float pop[9] = {md[0], md[1], md[2], um[0], dn[0], um[1], um[2], dn[1], dn[2]};
        if (MASKED) {
            // mask values of the sources (x - c_ix, y - c_iy) in direction order
            const int ks[9] = {Ky, KyL, KyU, Km, Kn, KmL, KmU, KnU, KnL};
            uint32_t cutm = 0;  // cut links of an interior cell (the others' pulls are not kept)
#pragma unroll
            for (int i = 1; i < 9; ++i)
                if (inner && (ks[i] != 0) != (Ky != 0)) cutm |= 1u << i;
            if (__builtin_amdgcn_ballot_w64(cutm != 0)) {
                // the row's cut links listed in lane and link order: the
                // exclusive prefix count of the lanes' links (four bit planes)
                const uint32_t cnt = (uint32_t)__builtin_popcount(cutm);
                uint32_t off = 0, K = 0;
#pragma unroll
                for (int bp = 0; bp < 4; ++bp) {
                    const uint64_t bal = __builtin_amdgcn_ballot_w64((cnt >> bp) & 1u);
                    off += __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u)) << bp;
                    K += (uint32_t)__builtin_popcountll(bal) << bp;
                }
#pragma unroll
                for (int i = 1; i < 9; ++i)
                    if ((cutm >> i) & 1u) {
                        const int sd = Ky ? Ky - 1 : ks[i] - 1;
                        clist[wave][off++] = (uint32_t)lane | ((uint32_t)i << 6) | ((uint32_t)(sd + 1) << 10);
                    }
#pragma unroll
                for (int f = 0; f < 6; ++f) cmom[wave][f][lane] = My[f];
                // each listed link by one lane: the wall-rule population and,
                // near the body, the force into its solid's slot
                for (uint32_t k0 = 0; k0 < K; k0 += 64) {
                    const uint32_t k = k0 + (uint32_t)lane;
                    float F[3] = {0.0f, 0.0f, 0.0f};
                    int sd = -1;
                    if (k < K) {
                        const uint32_t en = clist[wave][k];
                        const int l = (int)(en & 63u), i = (int)((en >> 6) & 15u);
                        sd = (int)(en >> 10) - 1;
                        float mc[6];
#pragma unroll
                        for (int f = 0; f < 6; ++f) mc[f] = cmom[wave][f][l];
                        const int xl = x0 - 1 + l;
                        float pl;
                        if (near) pl = d2q9_cut_link<true>(ea, mc, xl, y, i, sd, poses, F);
                        else pl = d2q9_cut_link<false>(ea, mc, xl, y, i, sd, poses, F);
                        cpop[wave][i][l] = pl;
                        if (sd >= nsolid) sd = -1;
                    }
                    if (near) {
                        // this pass's solids in the order of their first lanes
                        // (ballots, no shuffle chain), each lane's link into its
                        // solid's slot
                        uint64_t pend = __builtin_amdgcn_ballot_w64(sd >= 0);
                        while (pend) {
                            const int sv = __builtin_amdgcn_readlane(sd, (int)__builtin_ctzll(pend));
                            const bool mine = sd == sv;
                            pend &= ~__builtin_amdgcn_ballot_w64(mine);
                            const int j = slot_of(sv);
#pragma unroll
                            for (int jj = 0; jj < kSlots; ++jj)
                                if (jj == j) {
                                    fa0[jj] += mine ? (double)F[0] : 0.0;
                                    fa1[jj] += mine ? (double)F[1] : 0.0;
                                    fa2[jj] += mine ? (double)F[2] : 0.0;
                                }
                        }
                    }
                }
#pragma unroll
                for (int i = 1; i < 9; ++i)
                    if ((cutm >> i) & 1u) pop[i] = cpop[wave][i][lane];
            }
        }
        d2q9_collide<float>(ea, pop, 0.0f, 0.0f, mout);
        store_row(y, vst, mout);
        // face columns to the lists
        if (left) list_put(Lf, Ly, nl++, mout, lsrc, y);
        if (right) list_put(Rf, Ry, nr++, mout, rsrc, y);
        // march on
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            um[k] = uy[k];
            uy[k] = uyn[k];
            md[k] = mdn[k];
        }
#pragma unroll
        for (int f = 0; f < 6; ++f) My[f] = Mn[f];
        if (MASKED) {
            Km = Ky; KmL = KyL; KmU = KyU;
            Ky = Kn; KyL = KnL; KyU = KnU;
        }
    };

    // prologue: rows ya - 1 and ya rebuilt, rows ya + 1 .. ya + 3 requested
    float R0[6], R1[6], R2[6], R3[6];
    int K0 = 0, K1 = 0, K2 = 0, K3 = 0;
    {
        float Ma[6], Mb[6];
        int Ka = 0, Kb = 0;
        load_row(ya - 1, Ma, Ka);
        load_row(ya, Mb, Kb);
        load_row(ya + 1, R0, K0);
        load_row(min(ya + 2, yb), R1, K1);
        load_row(min(ya + 3, yb), R2, K2);
        float dn[3], mk[3];
        send(Ma, dn, mk, um);
        send(Mb, dn, md, uy);
#pragma unroll
        for (int f = 0; f < 6; ++f) My[f] = Mb[f];
        if (MASKED) {
            Km = row_mask(ya - 1, Ka);
            KmL = from_lower(Km);
            KmU = from_upper(Km);
            Ky = row_mask(ya, Kb);
            KyL = from_lower(Ky);
            KyU = from_upper(Ky);
        }
    }
    // step y: row y + 1 in ring buffer (y - ya) % 4, row y + 4 (past the last
    // row, the last row again, so every step issues the same loads) requested
    // into the buffer row y left
    auto step = [&](int y, const float* Mn, int Kn, float* Mp, int& Kp, float* mout) {
        load_row(min(y + 4, yb), Mp, Kp);
        dest_row(y, Mn, Kn, mout);
    };
    step(ya, R0, K0, R3, K3, mfirst);
#pragma unroll
    for (int f = 0; f < 6; ++f) mlast[f] = mfirst[f];
    int y = ya + 1;
    for (; y + 3 < yb; y += 4) {
        step(y, R1, K1, R0, K0, mlast);
        step(y + 1, R2, K2, R1, K1, mlast);
        step(y + 2, R3, K3, R2, K2, mlast);
        step(y + 3, R0, K0, R3, K3, mlast);
    }
    if (y < yb) {
        step(y, R1, K1, R0, K0, mlast);
        if (y + 1 < yb) {
            step(y + 1, R2, K2, R1, K1, mlast);
            if (y + 2 < yb) step(y + 2, R3, K3, R2, K2, mlast);
        }
    }

    // face rows from their projection rows over the interior columns, the
    // corners to the face-column lists
    if (y0 == 0) {  // row 0 from row 1 (the first row)
        float mf[6] = {mfirst[0], mfirst[1], mfirst[2], mfirst[3], mfirst[4], mfirst[5]};
        bc_chain<float, 2>(mf, inner ? x : 1, 0, 0, nx, ny, 1, ea.bc_type, ea.bc_value);
        store_row(0, vst, mf);
        if (left) list_put(Lf, Ly, nl++, mfirst, lsrc, 0);
        if (right) list_put(Rf, Ry, nr++, mfirst, rsrc, 0);
    }
    if (y0 + h == ny) {  // row ny - 1 from row ny - 2 (the last row)
        float mf[6] = {mlast[0], mlast[1], mlast[2], mlast[3], mlast[4], mlast[5]};
        bc_chain<float, 2>(mf, inner ? x : 1, ny - 1, 0, nx, ny, 1, ea.bc_type, ea.bc_value);
        store_row(ny - 1, vst, mf);
        if (left) list_put(Lf, Ly, nl++, mlast, lsrc, ny - 1);
        if (right) list_put(Rf, Ry, nr++, mlast, rsrc, ny - 1);
    }
    // face columns: the face chain over the collected cells, one lane each
    if (nl) {
        float mf[6] = {Lf[0], Lf[1], Lf[2], Lf[3], Lf[4], Lf[5]};
        if (lane < nl) bc_chain<float, 2>(mf, 0, Ly, 0, nx, ny, 1, ea.bc_type, ea.bc_value);
        const uint32_t v = lane < nl ? (uint32_t)(Ly * nx) * 4u : kOOB;
#pragma unroll
        for (int f = 0; f < 6; ++f)
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, mf[f]), rd, (int)v, (int)(f * fbytes), 0);
    }
    if (nr) {
        float mf[6] = {Rf[0], Rf[1], Rf[2], Rf[3], Rf[4], Rf[5]};
        if (lane < nr) bc_chain<float, 2>(mf, nx - 1, Ry, 0, nx, ny, 1, ea.bc_type, ea.bc_value);
        const uint32_t v = lane < nr ? (uint32_t)(Ry * nx + nx - 1) * 4u : kOOB;
#pragma unroll
        for (int f = 0; f < 6; ++f)
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, mf[f]), rd, (int)v, (int)(f * fbytes), 0);
    }
    if (MASKED && near) flush();
    if (near && lane < kPartialWidth) as_global(part)[(int64_t)slot * kPartialWidth + lane] = acc;
}

template <bool VOX>
__device__ __forceinline__ void march_strip(const BatchEnv& be, const BatchItem& it, int par, int t, double* part)
{
    if (strip_masked(be, it)) march<VOX, true, kWaves>(be, it, par, t, part);
    else march<VOX, false, kWaves>(be, it, par, t, part);
}

// The step of the strips [item0, item0 + count), one per wavefront: the
// strips that reach a body box (MASKED) and the others are listed apart, so
// the plain strips' kernel keeps the registers of the plain march only.
__device__ __forceinline__ void sum_output(const BatchArgs& b, const double* part, int k, double* history,
                                           int nenv_total, int e, int r);

// Step k of a call (parity and time k steps past the environment's), its
// force rows into bank k & 1 of nslots rows. The blocks past the strips'
// (sums > 0) sum step k - 1's rows into the history, one output per thread.
template <bool MASKED>
__global__ void __launch_bounds__(64 * kStepWaves) k_bstep(BatchArgs b, int item0, int count, int k, int nslots,
                                                            double* history, int nenv_total)
{
    const int nb = (count + kStepWaves - 1) / kStepWaves;
    if ((int)blockIdx.x >= nb) {
        const int o = ((int)blockIdx.x - nb) * 64 * kStepWaves + (int)threadIdx.x;
        if (o < nenv_total * kBatchMaxSolids * 6)
            sum_output(b, b.partial + (size_t)((k - 1) & 1) * nslots * kPartialWidth, k - 1, history, nenv_total,
                       o / (kBatchMaxSolids * 6), o % (kBatchMaxSolids * 6));
        return;
    }
    const int i = (int)blockIdx.x * kStepWaves + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (i >= count) return;
    const BatchItem it = b.step_items[item0 + i];
    const int e = uni(it.env);
    if (!env_active(b, e)) return;
    march<false, MASKED, kStepWaves>(b.envs[e], it, (uni(b.parity[e]) + k) & 1, uni(b.time[e]) + k,
                                     b.partial + (size_t)(k & 1) * nslots * kPartialWidth);
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

// The end of a call of K steps (environments [0, nenv_total), one thread per
// output): step K - 1's force rows summed into the history as k_badvance
// sums them, then parity and time advanced by K.
__global__ void __launch_bounds__(64 * kStepWaves) k_bclose(BatchArgs b, int K, int nslots, double* history,
                                                            int nenv_total)
{
    const int o = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (history && o < nenv_total * kBatchMaxSolids * 6)
        sum_output(b, b.partial + (size_t)((K - 1) & 1) * nslots * kPartialWidth, K - 1, history, nenv_total,
                   o / (kBatchMaxSolids * 6), o % (kBatchMaxSolids * 6));
    if (o < nenv_total && env_active(b, o)) {
        b.parity[o] = (b.parity[o] + K) & 1;
        b.time[o] += K;
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

// K steps of the step items [i0, i0 + ni) and [j0, j0 + nj) (environments
// [e0, e0 + ne)) in
// one launch: the items dealt to the wavefronts in order, each the
// three-kernel schedule's step with the voxelizer fused into its row loads; a
// grid barrier after every step; step k's force rows go to bank k & 1 and
// are summed by step k + 1 (the last step's after the loop); then parity and
// time advance by K.
__global__ void __launch_bounds__(kThreads) k_brun(BatchArgs b, int i0, int ni, int j0, int nj, int K,
                                                   double* history, int nenv_total, int e0, int ne, int nslots,
                                                   unsigned* bar, int* flag)
{
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (int k = 0; k < K; ++k) {
        if (k > 0 && history)
            run_sums(b, b.partial + (size_t)((k - 1) & 1) * nslots * kPartialWidth, k - 1, history, nenv_total, e0, ne);
        // step items dealt to the wavefronts in order
        for (int u = (int)blockIdx.x * kWaves + wave; u < ni + nj; u += (int)gridDim.x * kWaves) {
            const BatchItem it = b.step_items[u < ni ? i0 + u : j0 + (u - ni)];
            const int e = uni(it.env);
            if (!env_active(b, e)) continue;
            march_strip<true>(b.envs[e], it, (uni(b.parity[e]) + k) & 1, uni(b.time[e]) + k,
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
    const int g = (int)(blockIdx.x & 7), lb = (int)(blockIdx.x >> 3), nb = (int)(gridDim.x >> 3);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // this workgroup's contiguous share of its group's step items, dealt to
    // its wavefronts in order
    const int ib = plan.goff[g], ie = plan.goff[g + 1];
    const int per = (ie - ib + nb - 1) / nb;
    const int i0 = ib + lb * per, i1 = min(i0 + per, ie);
    const int ge0 = plan.geoff[g], ne = plan.geoff[g + 1] - ge0;
    const int gt = lb * kThreads + (int)threadIdx.x, nt = nb * kThreads;
    unsigned* bar = bars + g * kXcdBarStride;
    for (int k = 0; k < K; ++k) {
        if (k > 0 && history)
            for (int o = gt; o < ne * kBatchMaxSolids * 6; o += nt)
                sum_output(b, b.partial + (size_t)((k - 1) & 1) * nslots * kPartialWidth, k - 1, history, nenv_total,
                           genv[ge0 + o / (kBatchMaxSolids * 6)], o % (kBatchMaxSolids * 6));
        for (int i = i0 + wave; i < i1; i += kWaves) {
            const BatchItem it = b.step_items[xitems[i]];
            const int e = uni(it.env);
            if (!env_active(b, e)) continue;
            march_strip<true>(b.envs[e], it, (uni(b.parity[e]) + k) & 1, uni(b.time[e]) + k,
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
    std::vector<hl::BatchItem> step_items, vox_items;  // step items: every plain strip, then the masked ones
    int nplain = 0;
    int nslots = 0;
    int rows = hl::kSH, body_rows = hl::kSHB;  // strip heights (hl_batch_set_strips)
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
    // the body strips' side of a step (voxelizer, masked march), forked from
    // and joined back to the caller's stream
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_plain = nullptr;
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

// Whether a strip reads mask values, as the device decides for the masked
// march: its source rows and columns reach the body box.
bool strip_reaches_box(const hl::EnvArgs& a, const hl::BatchItem& it)
{
    const int ya = std::max(it.y0, 1), yb = std::min(it.y0 + it.h, a.ny - 1);
    return a.mask != nullptr && it.x0 - 2 <= a.box[1] && it.x0 + it.w + 1 >= a.box[0] - 1 && ya - 2 <= a.box[3] &&
           yb + 1 >= a.box[2] - 1;
}

// Step items (strips) cover every environment: columns in strips of
// near-equal widths (at most kSW, so at least 31 columns, or the whole width
// of at least 3), rows in strips of at most `rows` rows, and of at most
// `body_rows` where a column strip crosses the body box (the wall rule and the forces of
// those rows are then shared by more wavefronts); the face rows' strips hold
// at least 2 rows, so every face cell's projection lies in its strip. Each
// environment lists its plain strips, then the strips that reach the box
// (the masked march), which alone can meet the force region (interior cells
// within one cell of the box, where every cut link has its fluid end) and
// then get a partial slot. Voxelizer items tile the mask box.
void rebuild_items(hl_batch* b)
{
    b->step_items.clear();
    b->vox_items.clear();
    std::vector<std::vector<hl::BatchItem>> masked((size_t)b->n);
    for (int e = 0; e < b->n; ++e) {
        hl::BatchEnv& be = b->envs[e];
        const hl::EnvArgs& a = be.base;
        const bool body = a.mask && a.box[0] < a.box[1] && a.box[2] < a.box[3];
        be.step_begin = (int)b->step_items.size();
        const int nsx = (a.nx + hl::kSW - 1) / hl::kSW;
        for (int i = 0; i < nsx; ++i) {
            const int x = (int)((int64_t)a.nx * i / nsx), xw = (int)((int64_t)a.nx * (i + 1) / nsx) - x;
            // row segments: the rows within 3 of the box apart when the strip
            // column crosses it (cuts kept within [2, ny - 2])
            const bool xb = body && x - 2 <= a.box[1] && x + xw + 1 >= a.box[0] - 1;
            std::vector<int> cuts = {0, a.ny};
            const int c0 = a.box[2] - 3, c1 = a.box[3] + 3;
            if (xb && c0 >= 2 && c0 <= a.ny - 2) cuts.push_back(c0);
            if (xb && c1 >= 2 && c1 <= a.ny - 2) cuts.push_back(c1);
            std::sort(cuts.begin(), cuts.end());
            cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
            for (size_t k = 0; k + 1 < cuts.size(); ++k) {
                const int s0 = cuts[k], s1 = cuts[k + 1];
                const bool inner = xb && s0 < c1 && s1 > c0;
                const int hmax = inner ? b->body_rows : b->rows, n = (s1 - s0 + hmax - 1) / hmax;
                for (int j = 0; j < n; ++j) {
                    const int y = s0 + (int)((int64_t)(s1 - s0) * j / n);
                    const int yh = s0 + (int)((int64_t)(s1 - s0) * (j + 1) / n) - y;
                    const hl::BatchItem it{e, x, y, xw, yh, -1};
                    if (strip_reaches_box(a, it)) masked[(size_t)e].push_back(it);
                    else b->step_items.push_back(it);
                }
            }
        }
        be.step_count = (int)b->step_items.size() - be.step_begin;
    }
    b->nplain = (int)b->step_items.size();
    int slot = 0;
    for (int e = 0; e < b->n; ++e) {
        hl::BatchEnv& be = b->envs[e];
        const hl::EnvArgs& a = be.base;
        const bool body = a.mask && a.box[0] < a.box[1] && a.box[2] < a.box[3];
        const int fx0 = std::max(a.box[0] - 1, 1), fx1 = std::min(a.box[1] + 1, a.nx - 1);
        const int fy0 = std::max(a.box[2] - 1, 1), fy1 = std::min(a.box[3] + 1, a.ny - 1);
        const bool forces = body && a.nsolid > 0 && fx0 < fx1 && fy0 < fy1;
        be.mstep_begin = (int)b->step_items.size();
        be.slot_begin = slot;
        for (hl::BatchItem it : masked[(size_t)e]) {
            if (forces && it.x0 < fx1 && it.x0 + it.w > fx0 && it.y0 < fy1 && it.y0 + it.h > fy0) it.slot = slot++;
            b->step_items.push_back(it);
        }
        be.mstep_count = (int)b->step_items.size() - be.mstep_begin;
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
                     [&](int p, int q) {
                         const hl::BatchEnv &u = b->envs[(size_t)p], &v = b->envs[(size_t)q];
                         return u.step_count + u.mstep_count > v.step_count + v.mstep_count;
                     });
    std::vector<std::vector<int>> groups(8);
    int64_t load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int e : order) {
        int g = 0;
        for (int h = 1; h < 8; ++h)
            if (load[h] < load[g]) g = h;
        groups[(size_t)g].push_back(e);
        load[g] += b->envs[(size_t)e].step_count + b->envs[(size_t)e].mstep_count;
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
            for (int t = 0; t < be.mstep_count; ++t) xitems.push_back(be.mstep_begin + t);
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
    const int W = hl::kStepWaves, T = 64 * hl::kStepWaves;
    const int nout = b->n * hl::kBatchMaxSolids * 6, sum_blocks = (nout + T - 1) / T;
    if (!per_env) {
        // Step k of the call runs k steps past each environment's parity and
        // time (advanced once, by the closing kernel); its force rows go to
        // bank k & 1 and are summed by step k + 1's plain-strip launch (the
        // last step's by the closing kernel), so a step is three launches.
        // The side stream carries the voxelizer and the body strips (masked
        // march), s the plain strips: they touch different cells and the
        // plain march reads no mask. Step k + 1's voxelizer follows step k's
        // body strips (the last mask readers) while the plain strips still
        // run; each kind of strip of step k + 1 waits for the other kind of
        // step k (the pulls cross strip borders).
        const int nm = nstep - b->nplain;
        hipError_t fe = hipSuccess;
        const bool side = nm > 0 || nvox > 0;
        if (side && ((fe = hipEventRecord(b->ev_fork, s)) != hipSuccess ||
                     (fe = hipStreamWaitEvent(b->side, b->ev_fork, 0)) != hipSuccess))
            return hip_report(fe, "batch stream fork");
        for (int k = 0; k < nsteps; ++k) {
            if (k == 0 && nvox) hipLaunchKernelGGL(hl::k_bvox, dim3(nvox), dim3(256), 0, b->side, a, 0, 0);
            if (nm > 0)
                hipLaunchKernelGGL(hl::k_bstep<true>, dim3((nm + W - 1) / W), dim3(T), 0, b->side, a, b->nplain, nm, k,
                                   b->nslots, history, b->n);
            if (side && (fe = hipEventRecord(b->ev_join, b->side)) != hipSuccess)  // step k's body strips
                return hip_report(fe, "batch stream event");
            if (k + 1 < nsteps && nvox) hipLaunchKernelGGL(hl::k_bvox, dim3(nvox), dim3(256), 0, b->side, a, 0, k + 1);
            const int extra = (k > 0 && history) ? sum_blocks : 0;
            if (b->nplain + extra > 0)
                hipLaunchKernelGGL(hl::k_bstep<false>, dim3((b->nplain + W - 1) / W + extra), dim3(T), 0, s, a, 0,
                                   b->nplain, k, b->nslots, history, b->n);
            if (side && k + 1 < nsteps &&
                ((fe = hipEventRecord(b->ev_plain, s)) != hipSuccess ||
                 (fe = hipStreamWaitEvent(b->side, b->ev_plain, 0)) != hipSuccess))
                return hip_report(fe, "batch stream event");
            if (side && (fe = hipStreamWaitEvent(s, b->ev_join, 0)) != hipSuccess)
                return hip_report(fe, "batch stream join");
        }
        if (nsteps > 0)
            hipLaunchKernelGGL(hl::k_bclose, dim3(std::max((std::max(nout, b->n) + T - 1) / T, 1)), dim3(T), 0, s, a,
                               (int)nsteps, b->nslots, history, b->n);
    } else {
        // the same kernels, one environment per launch, each step closed by
        // k_badvance (parity and time advanced per step)
        for (int k = 0; k < nsteps; ++k)
            for (int e = 0; e < b->n; ++e) {
                const hl::BatchEnv& be = b->envs[e];
                if (be.vox_count) hipLaunchKernelGGL(hl::k_bvox, dim3(be.vox_count), dim3(256), 0, s, a, be.vox_begin, 0);
                if (be.mstep_count)
                    hipLaunchKernelGGL(hl::k_bstep<true>, dim3((be.mstep_count + W - 1) / W), dim3(T), 0, s, a,
                                       be.mstep_begin, be.mstep_count, 0, b->nslots, history, b->n);
                if (be.step_count)
                    hipLaunchKernelGGL(hl::k_bstep<false>, dim3((be.step_count + W - 1) / W), dim3(T), 0, s, a,
                                       be.step_begin, be.step_count, 0, b->nslots, history, b->n);
                hipLaunchKernelGGL(hl::k_badvance, dim3(1), dim3(64), 0, s, a, e, history, b->n, k);
            }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_report(e, "batch kernel launch");
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
        // the step's buffer offsets are 32-bit with the out-of-range marker at 2^31
        const int64_t st = x.stride > 0 ? x.stride : (int64_t)x.nx * x.ny;
        if (st * 6 * 4 >= (int64_t)1 << 31) {
            std::string m = "batch environment " + std::to_string(e) + ": fields of more than 2^31 bytes";
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
    if ((he = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking)) != hipSuccess ||
        (he = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming)) != hipSuccess ||
        (he = hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming)) != hipSuccess ||
        (he = hipEventCreateWithFlags(&b->ev_plain, hipEventDisableTiming)) != hipSuccess) {
        hl_batch_destroy(b);
        return hip_report(he, "batch stream");
    }
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
    if (b->side) hipStreamDestroy(b->side);
    if (b->ev_fork) hipEventDestroy(b->ev_fork);
    if (b->ev_join) hipEventDestroy(b->ev_join);
    if (b->ev_plain) hipEventDestroy(b->ev_plain);
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

int32_t hl_batch_set_strips(hl_batch* b, int32_t rows, int32_t body_rows)
{
    if (!b) return report(HL_E_INVALID, "batch is NULL");
    // at least 3 rows, so every face-row strip keeps its projection row; at
    // most 64, the lanes that collect a strip's face-column cells
    if (rows < 3 || rows > 64 || body_rows < 3 || body_rows > 64)
        return report(HL_E_INVALID, "strip heights must be in [3, 64]");
    b->rows = rows;
    b->body_rows = body_rows;
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
    auto launch = [&](int i0, int ni, int j0, int nj, int e0, int ne) -> int32_t {
        if (ni + nj <= 0) return HL_OK;
        // every workgroup resident (a cooperative launch); no more wavefronts than items
        const int grid = std::max(1, std::min(occ * cus, (ni + nj + hl::kWaves - 1) / hl::kWaves));
        hipLaunchKernelGGL(hl::k_brun_reset, dim3(1), dim3(1), 0, s, b->d_bar, b->d_flag);
        int i0v = i0, niv = ni, j0v = j0, njv = nj, K = nsteps, ntot = b->n, e0v = e0, nev = ne, nsl = b->nslots;
        double* hist = force_history;
        unsigned* bar = b->d_bar;
        int* flag = b->d_flag;
        hl::BatchArgs av = a;
        void* args[] = {&av, &i0v, &niv, &j0v, &njv, &K, &hist, &ntot, &e0v, &nev, &nsl, &bar, &flag};
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
        const int g8 = std::max(8, std::min(occ * cus, (ni + hl::kWaves - 1) / hl::kWaves) / 8 * 8);
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
        return launch(0, ni, 0, 0, 0, b->n);
    }
    for (int env = 0; env < b->n; ++env) {
        const hl::BatchEnv& be = b->envs[env];
        rc = launch(be.step_begin, be.step_count, be.mstep_begin, be.mstep_count, env, 1);
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
