// Exact polygon cut cells for D2Q9 (the reference's boundary, no mask).
//
// The reference tests every link of every fluid cell against the segments of
// the polygon bodies whose bounding circle holds the cell (get_cutcell,
// envs/lbm/lbm_func.py:545-607) and, on a crossing, applies the wall rule at
// the crossing's material point (stream_and_collide, lbm_func.py:346-401).
// Far from bodies that is the plain pull step, so a step here is one
// launch of k_poly_step, whose workgroups take one of two roles:
//   1. the band (the first kPolyWaves workgroups, one wave each) over the
//      bounding box of the bounding circles (each workgroup finds it first,
//      poly_bbox): cells whose interior projection lies in a circle are
//      computed with the segment predicate, face cells through the same
//      face chain as the gather kernel; interior cells also emit the
//      momentum exchange of their cut links. G = 8, 4 or 2 lanes per link
//      (8 / G cells per wavefront pass) or, for a large box, one lane per
//      cell; the vertex tables in LDS;
//   2. the plain pull (gather with no mask) of every other cell, 64 cells
//      of one row per workgroup.
// The two roles write disjoint cells and read only the source state. The
// band's fixed number of per-wave FP64 rows is summed by the next step's
// band workgroups (k_sum_partials after the last step). The band strides
// over the box with a fixed number of waves, so the FP64 force order does
// not depend on scheduling. Arithmetic follows the reference's operation
// order with contraction off.
#pragma clang fp contract(off)

#include <math.h>

#include <vector>

#include "gather_post.h"
#include "kernels.h"
#include "lbm_device.h"

namespace hl {

namespace {

constexpr int kPolyWaves = 1024;  // band grid: one row of force partials per wave

// A lane's share of the sum of column k over the kPolyWaves rows: rows
// lane, lane + 64, ... added in that order (the order of k_sum_partials),
// every load issued before the first add.
__device__ __forceinline__ double rows_lane_sum(const double* rows, int width, int k, int lane)
{
    static_assert(kPolyWaves % 64 == 0, "whole rows per lane");
    constexpr int kPer = kPolyWaves / 64;
    double t[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) t[j] = rows[(int64_t)(lane + 64 * j) * width + k];
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) v += t[j];
    return v;
}

struct PolyArgs {
    const float* verts;       // [nsolid][maxv][2], this step
    const float* verts_last;  // same layout, previous step
    const int32_t* counts;    // [nsolid]
    int32_t maxv;
    const float* radius2;     // [nsolid]
};

__device__ __forceinline__ int poly_count(const PolyArgs& p, int s)
{
    const int n = p.counts[s];
    return n < 0 ? 0 : (n > p.maxv ? p.maxv : n);
}

// Solids whose bounding circle holds (x, y): the early exit of get_cutcell
// (a NaN distance is not skipped, as there).
template <typename T>
__device__ __forceinline__ uint64_t poly_near(const EnvArgs& a, const PolyArgs& p, int x, int y)
{
    uint64_t near = 0;
    for (int s = 0; s < a.nsolid; ++s) {
        if (poly_count(p, s) == 0) continue;
        const T ox = T(x) - T(a.solids[8 * s]), oy = T(y) - T(a.solids[8 * s + 1]);
        const T d2 = ox * ox + oy * oy;
        if (!(d2 > T(p.radius2[s]))) near |= 1ull << s;
    }
    return near;
}

// Nearest crossing (to the cell) of link i of cell (x, y) with a segment of
// the near solids; returns the solid or -1, with the crossing's position
// along its segment and the segment's first vertex.
template <typename T>
__device__ __forceinline__ int poly_cut(const PolyArgs& p, const float* verts, uint64_t near, int x, int y, int i,
                                       T* ratio, int* seg)
{
    T cutcell = T(-1), sr = T(-1);
    int sid = -1, sg = -1;
    const T bx = T(x), by = T(y);
    const T ax = bx - T(D2CX[i]), ay = by - T(D2CY[i]);
    while (near) {
        const int s = __builtin_ctzll(near);
        near &= near - 1;
        const int num = poly_count(p, s);
        const float* v = verts + (int64_t)s * p.maxv * 2;
        T cx = T(v[0]), cy = T(v[1]);  // segment n runs from vertex n to vertex n+1 (mod num)
        for (int n = 0; n < num; ++n) {
            const int m = n + 1 == num ? 0 : n + 1;
            const T dx = T(v[2 * m]), dy = T(v[2 * m + 1]);
            const T abc = (ax - cx) * (by - cy) - (ay - cy) * (bx - cx);
            const T abd = (ax - dx) * (by - dy) - (ay - dy) * (bx - dx);
            if (!(abc * abd >= T(0))) {  // the reference skips on >= 0 (a NaN product goes on)
                const T cda = (cx - ax) * (dy - ay) - (cy - ay) * (dx - ax);
                const T cdb = cda + abc - abd;
                if (!(cda * cdb > T(0))) {
                    const T t = T(1) - cda / (abd - abc);
                    if (cutcell < T(0) || t < cutcell) {
                        cutcell = t;
                        sid = s;
                        sr = abc / (abc - abd);
                        sg = n;
                    }
                }
            }
            cx = dx;
            cy = dy;
        }
    }
    if (cutcell >= T(0) && sid >= 0) {
        *ratio = sr;
        *seg = sg;
        return sid;
    }
    return -1;
}

// The same nearest crossing with the segment loop shared by G lanes (G a
// power of two up to 8): lane g takes the segments n = g, g + G, ... of the
// near solids, then the group
// keeps the smallest t, ties to the earliest (solid, segment) in the loop
// order. With every crossing at t >= 0 that is exactly the sequential rule
// (the first crossing is taken, later ones only when strictly nearer); a
// crossing with t < 0 or NaN makes the group run the sequential loop.
constexpr int kMaxG = 8;

template <typename T>
__device__ __forceinline__ int poly_cut_split(const PolyArgs& p, const float* verts, uint64_t near, int x, int y,
                                              int i, int g, int G, T* ratio, int* seg)
{
    T bt = T(-1), bsr = T(-1);
    int bsid = -1, bsg = -1, bkey = 0x7fffffff, odd = 0;
    const T bx = T(x), by = T(y);
    const T ax = bx - T(D2CX[i]), ay = by - T(D2CY[i]);
    uint64_t nr = near;
    while (nr) {
        const int s = __builtin_ctzll(nr);
        nr &= nr - 1;
        const int num = poly_count(p, s);
        const float* v = verts + (int64_t)s * p.maxv * 2;
        for (int n = g; n < num; n += G) {
            const int m = n + 1 == num ? 0 : n + 1;
            const T cx = T(v[2 * n]), cy = T(v[2 * n + 1]), dx = T(v[2 * m]), dy = T(v[2 * m + 1]);
            const T abc = (ax - cx) * (by - cy) - (ay - cy) * (bx - cx);
            const T abd = (ax - dx) * (by - dy) - (ay - dy) * (bx - dx);
            if (!(abc * abd >= T(0))) {
                const T cda = (cx - ax) * (dy - ay) - (cy - ay) * (dx - ax);
                const T cdb = cda + abc - abd;
                if (!(cda * cdb > T(0))) {
                    const T t = T(1) - cda / (abd - abc);
                    if (!(t >= T(0))) {
                        odd = 1;
                    } else if (bt < T(0) || t < bt) {
                        bt = t;
                        bsid = s;
                        bsr = abc / (abc - abd);
                        bsg = n;
                        bkey = (s << 20) | n;
                    }
                }
            }
        }
    }
    for (int off = 1; off < G; off <<= 1) {
        const T ot = __shfl_xor(bt, off), osr = __shfl_xor(bsr, off);
        const int okey = __shfl_xor(bkey, off), osid = __shfl_xor(bsid, off), osg = __shfl_xor(bsg, off);
        odd |= __shfl_xor(odd, off);
        if (ot >= T(0) && (bt < T(0) || ot < bt || (ot == bt && okey < bkey))) {
            bt = ot;
            bsr = osr;
            bkey = okey;
            bsid = osid;
            bsg = osg;
        }
    }
    if (odd) return poly_cut<T>(p, verts, near, x, y, i, ratio, seg);
    if (bt >= T(0) && bsid >= 0) {
        *ratio = bsr;
        *seg = bsg;
        return bsid;
    }
    return -1;
}

// Velocity (time step 1) and lever arm of the crossing's material point.
template <typename T>
__device__ __forceinline__ void poly_point(const EnvArgs& a, const PolyArgs& p, int sid, int seg, T ratio, T up[2],
                                           T r[2])
{
    const int num = poly_count(p, sid), m = seg + 1 == num ? 0 : seg + 1;
    const float* v = p.verts + (int64_t)sid * p.maxv * 2;
    const float* l = p.verts_last + (int64_t)sid * p.maxv * 2;
    const T pcx = T(v[2 * seg]) + (T(v[2 * m]) - T(v[2 * seg])) * ratio;
    const T pcy = T(v[2 * seg + 1]) + (T(v[2 * m + 1]) - T(v[2 * seg + 1])) * ratio;
    const T plx = T(l[2 * seg]) + (T(l[2 * m]) - T(l[2 * seg])) * ratio;
    const T ply = T(l[2 * seg + 1]) + (T(l[2 * m + 1]) - T(l[2 * seg + 1])) * ratio;
    up[0] = pcx - plx;
    up[1] = pcy - ply;
    r[0] = pcx - T(a.solids[8 * sid]);
    r[1] = pcy - T(a.solids[8 * sid + 1]);
}

// Bounding box (half-open [x0, x1) x [y0, y1)) of the cells whose interior
// projection can lie in a bounding circle; a circle with a non-finite centre
// or radius takes the whole lattice.
__device__ void poly_bbox(const EnvArgs& a, const PolyArgs& p, int* bbox)
{
    const float hx = (float)(a.nx - 2), hy = (float)(a.ny - 2);
    float x0 = 1e30f, x1 = -1e30f, y0 = 1e30f, y1 = -1e30f;
    for (int s = 0; s < a.nsolid; ++s) {
        if (poly_count(p, s) == 0) continue;
        const float cx = a.solids[8 * s], cy = a.solids[8 * s + 1], r2 = p.radius2[s];
        if (!isfinite(cx) || !isfinite(cy) || !isfinite(r2)) {
            x0 = 1.0f; x1 = hx; y0 = 1.0f; y1 = hy;
            continue;
        }
        if (r2 < 0.0f) continue;
        const float r = sqrtf(r2);
        x0 = fminf(x0, floorf(cx - r) - 2.0f);
        x1 = fmaxf(x1, ceilf(cx + r) + 2.0f);
        y0 = fminf(y0, floorf(cy - r) - 2.0f);
        y1 = fmaxf(y1, ceilf(cy + r) + 2.0f);
    }
    // clip to the interior, then one more cell for the faces
    x0 = fmaxf(x0, 1.0f); x1 = fminf(x1, hx);
    y0 = fmaxf(y0, 1.0f); y1 = fminf(y1, hy);
    if (x0 > x1 || y0 > y1) {
        bbox[0] = bbox[1] = bbox[2] = bbox[3] = 0;
        bbox[4] = 1;
        return;
    }
    bbox[0] = (int)x0 - 1;
    bbox[1] = (int)x1 + 2;
    bbox[2] = (int)y0 - 1;
    bbox[3] = (int)y1 + 2;
    // lanes per link: a small band gets eight lanes sharing each link's
    // segment loop, a large one one lane per link (the grid stays full)
    const int64_t cells = (int64_t)(bbox[1] - bbox[0]) * (bbox[3] - bbox[2]);
    bbox[4] = cells <= 1024 ? 8 : cells <= 4096 ? 4 : cells <= 16384 ? 2 : 1;
}

// LDS for the vertex tables: larger tables are read from global memory.
constexpr int kPolyLdsFloats = 8192;  // 32 KiB

// One wavefront per workgroup, 8 / G cells per pass (G lanes per link, set
// by poly_bbox from the band's size): the G lanes 8 G c + G (i - 1) + g
// share the segment test of link i of cell c; lane
// g = 0 then applies the wall rule or the pull and keeps that link's
// momentum exchange; the lane of link 1 also pulls the rest population,
// gathers the nine through LDS and collides. Force rows:
// per solid a wavefront sum per pass, accumulated in pass order (FP64).
// The band with one lane per cell, for bands too large to share links
// among lanes: lane c of a pass pulls all eight links of its cell, then
// tests them in turn and replaces the cut ones by the wall rule, each value
// with the arithmetic of the lane-per-link path (so the same state); the
// per-solid wave sums of each link's momentum exchange go into acc (lane 0)
// link by link.
template <typename T, int DT>
__device__ __forceinline__ void poly_band_cells(const EnvArgs& a, const PolyArgs& p, const float* verts,
                                                double* partial, double* acc, int64_t n, int bw, int bx0, int by0,
                                                int wg, int nwg)
{
    using S = Storage<DT>;
    using ST = typename S::type;
    const ST* src = static_cast<const ST*>(a.src);
    ST* dst = static_cast<ST*>(a.dst);
    const int lane = threadIdx.x;
    const int nx = a.nx, ny = a.ny;
    const T grid_length = T(1) / T(nx);
    for (int64_t base = (int64_t)wg * 64; base < n; base += (int64_t)nwg * 64) {
        const int64_t t = base + lane;
        int x = 0, y = 0, px = 0, py = 0;
        uint64_t near = 0;
        if (t < n) {
            x = bx0 + (int)(t % bw);
            y = by0 + (int)(t / bw);
            px = min(max(x, 1), nx - 2);
            py = min(max(y, 1), ny - 2);
            near = poly_near<T>(a, p, px, py);
        }
        const bool interior = px == x && py == y;
        const int64_t c = (int64_t)px + (int64_t)nx * py;
        T mc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
        T pp[9];
        if (near) {
            // every link's pull first, so that their loads issue together;
            // a cut link's value is replaced below
            load_cell<T, DT>(a, c, 6, mc);
#pragma unroll
            for (int i = 1; i < 9; ++i) {
                const int64_t s = (int64_t)(px - D2CX[i]) + (int64_t)nx * (py - D2CY[i]);
                T m[6];
#pragma unroll
                for (int f = 0; f < 6; ++f) m[f] = S::template load<T>(src, f * a.stride + s, f, a.inv_scale[f], a.offset[f]);
                T q = d2q9_recon<T>(m[0], m[1], m[2], m[3], m[4], m[5], i);
                if (q < T(0)) q = T(0);
                pp[i] = q;
            }
        }
#pragma unroll
        for (int i = 1; i < 9; ++i) {
            int sid = -1;
            T F[3] = {T(0), T(0), T(0)};
            if (near) {
                T ratio = 0;
                int seg = -1;
                sid = poly_cut_split<T>(p, verts, near, px, py, i, 0, 1, &ratio, &seg);
                if (sid >= 0) {
                    T q;
                    T up[2], r[2];
                    poly_point<T>(a, p, sid, seg, ratio, up, r);
                    const T Sxx_s = up[0] * up[0] + mc[3] - mc[1] * mc[1];
                    const T Syy_s = up[1] * up[1] + mc[4] - mc[2] * mc[2];
                    const T Sxy_s = up[0] * up[1] + mc[5] - mc[1] * mc[2];
                    q = d2q9_recon<T>(mc[0], up[0], up[1], Sxx_s, Syy_s, Sxy_s, i);
                    if (q < T(0)) q = T(0);
                    pp[i] = q;
                    if (partial && interior) {
                        const int inv = D2INV[i];
                        const T cur_f = d2q9_recon<T>(mc[0], mc[1], mc[2], mc[3], mc[4], mc[5], inv);
                        const T Fx = cur_f * (T(D2CX[inv]) - up[0]) - q * (T(D2CX[i]) - up[0]);
                        const T Fy = cur_f * (T(D2CY[inv]) - up[1]) - q * (T(D2CY[i]) - up[1]);
                        F[0] = Fx;
                        F[1] = Fy;
                        F[2] = (r[0] * Fy - r[1] * Fx) * grid_length;
                    } else {
                        sid = -1;
                    }
                }
            }
            if (partial) wave_solid_sums(sid, F, kRow2D, lane, acc);
        }
        if (near) {
            T q0 = d2q9_recon<T>(mc[0], mc[1], mc[2], mc[3], mc[4], mc[5], 0);
            if (q0 < T(0)) q0 = T(0);
            pp[0] = q0;
            T Fx = 0, Fy = 0, out[6];
            if (a.force) { Fx = T(a.force[c]); Fy = T(a.force[a.fs + c]); }
            d2q9_collide<T>(a, pp, Fx, Fy, out);
            if (!interior) bc_chain<T, 2>(out, x, y, 0, nx, ny, 1, a.bc_type, a.bc_value);
            const int64_t d = (int64_t)x + (int64_t)nx * y;
#pragma unroll
            for (int f = 0; f < 6; ++f)
                S::template store<T>(dst, f * a.stride + d, out[f], f, a.scale[f], a.offset[f], HL_ROUNDING_KEY(a, f, x, y, 0));
        }
    }
}

template <typename T, int DT, bool LDSV>
__device__ __forceinline__ void poly_band(const EnvArgs& a, const PolyArgs& p, double* partial, const double* prev,
                                          double* prev_out, int wg, int nwg)
{
    extern __shared__ float lds_verts[];
    __shared__ int bbox[5];
    __shared__ double acc[HL_MAX_SOLIDS * 6];
    __shared__ T pops[8][9];
    using S = Storage<DT>;
    using ST = typename S::type;
    const ST* src = static_cast<const ST*>(a.src);
    ST* dst = static_cast<ST*>(a.dst);
    const int lane = threadIdx.x;
    const int nx = a.nx, ny = a.ny;
    const int width = a.nsolid * 6;
    // the vertex tables in LDS when they fit (a compile-time choice, so the
    // segment loop reads LDS directly rather than through generic pointers);
    // the first 64 values are loaded before the box is found, so that the
    // two sets of loads overlap
    const int nv = a.nsolid * p.maxv * 2;
    const float v0 = LDSV && lane < nv ? p.verts[lane] : 0.0f;
    if (threadIdx.x == 0) poly_bbox(a, p, bbox);  // every workgroup finds the same box
    if (LDSV) {
        if (lane < nv) lds_verts[lane] = v0;
        for (int k = lane + 64; k < nv; k += 64) lds_verts[k] = p.verts[k];
    }
    const float* verts = LDSV ? static_cast<const float*>(lds_verts) : p.verts;
    if (partial)
        for (int k = lane; k < width; k += 64) acc[k] = 0.0;
    __syncthreads();  // bbox, vertex tables and acc are ready
    const int G = min(max(bbox[4], 1), kMaxG), cells = 8 / G;
    const int g = lane % G, slot = lane / (8 * G), i = (lane / G) % 8 + 1;
    const int bx0 = bbox[0], by0 = bbox[2];
    const int bw = bbox[1] - bbox[0], bh = bbox[3] - bbox[2];
    const int64_t n = bw > 0 && bh > 0 ? (int64_t)bw * bh : 0;
    const T grid_length = T(1) / T(nx);
    if (G == 1)  // a large band: one lane per cell
        poly_band_cells<T, DT>(a, p, verts, partial, acc, n, bw, bx0, by0, wg, nwg);
    else
    for (int64_t base = (int64_t)wg * cells; base < n; base += (int64_t)nwg * cells) {
        const int64_t t = base + slot;
        int sid = -1;
        T F[3] = {T(0), T(0), T(0)};
        bool work = false, interior = false;
        int x = 0, y = 0;
        int64_t c = 0;
        if (t < n) {
            x = bx0 + (int)(t % bw);
            y = by0 + (int)(t / bw);
            const int px = min(max(x, 1), nx - 2), py = min(max(y, 1), ny - 2);
            const uint64_t near = poly_near<T>(a, p, px, py);
            if (near) {
                work = g == 0;
                interior = px == x && py == y;
                c = (int64_t)px + (int64_t)nx * py;
                T mc[6];
                load_cell<T, DT>(a, c, 6, mc);
                T ratio = 0;
                int seg = -1;
                sid = poly_cut_split<T>(p, verts, near, px, py, i, g, G, &ratio, &seg);
                if (g != 0) sid = -1;  // lane g = 0 carries the link from here
                T q = T(0);
                if (sid >= 0) {
                    T up[2], r[2];
                    poly_point<T>(a, p, sid, seg, ratio, up, r);
                    const T Sxx_s = up[0] * up[0] + mc[3] - mc[1] * mc[1];
                    const T Syy_s = up[1] * up[1] + mc[4] - mc[2] * mc[2];
                    const T Sxy_s = up[0] * up[1] + mc[5] - mc[1] * mc[2];
                    q = d2q9_recon<T>(mc[0], up[0], up[1], Sxx_s, Syy_s, Sxy_s, i);
                    if (q < T(0)) q = T(0);
                    if (partial && interior) {
                        const int inv = D2INV[i];
                        const T cur_f = d2q9_recon<T>(mc[0], mc[1], mc[2], mc[3], mc[4], mc[5], inv);
                        const T Fx = cur_f * (T(D2CX[inv]) - up[0]) - q * (T(D2CX[i]) - up[0]);
                        const T Fy = cur_f * (T(D2CY[inv]) - up[1]) - q * (T(D2CY[i]) - up[1]);
                        F[0] = Fx;
                        F[1] = Fy;
                        F[2] = (r[0] * Fy - r[1] * Fx) * grid_length;
                    } else {
                        sid = -1;
                    }
                } else if (g == 0) {
                    const int64_t s = (int64_t)(px - D2CX[i]) + (int64_t)nx * (py - D2CY[i]);
                    T m[6];
#pragma unroll
                    for (int f = 0; f < 6; ++f) m[f] = S::template load<T>(src, f * a.stride + s, f, a.inv_scale[f], a.offset[f]);
                    q = d2q9_recon<T>(m[0], m[1], m[2], m[3], m[4], m[5], i);
                    if (q < T(0)) q = T(0);
                }
                if (g == 0) pops[slot][i] = q;
                if (g == 0 && i == 1) {  // the rest population: never cut (its link has no length)
                    T q0 = d2q9_recon<T>(mc[0], mc[1], mc[2], mc[3], mc[4], mc[5], 0);
                    if (q0 < T(0)) q0 = T(0);
                    pops[slot][0] = q0;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        if (work && i == 1) {
            T pp[9], Fx = 0, Fy = 0, out[6];
#pragma unroll
            for (int k = 0; k < 9; ++k) pp[k] = pops[slot][k];
            if (a.force) { Fx = T(a.force[c]); Fy = T(a.force[a.fs + c]); }
            d2q9_collide<T>(a, pp, Fx, Fy, out);
            if (!interior) bc_chain<T, 2>(out, x, y, 0, nx, ny, 1, a.bc_type, a.bc_value);
            const int64_t d = (int64_t)x + (int64_t)nx * y;
#pragma unroll
            for (int f = 0; f < 6; ++f)
                S::template store<T>(dst, f * a.stride + d, out[f], f, a.scale[f], a.offset[f], HL_ROUNDING_KEY(a, f, x, y, 0));
        }
        if (partial) wave_solid_sums(sid, F, kRow2D, lane, acc);
        // the lead lanes' reads of pops end before the next pass writes them
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
    // the previous step's rows: one wave per output, the order of k_sum_partials
    if (prev)
        for (int k = wg; k < width; k += nwg) {
            const double v = wave_sum(rows_lane_sum(prev, width, k, lane));
            if (lane == 0) prev_out[k] = v;
        }
    if (!partial) return;
    __syncthreads();
    for (int k = lane; k < width; k += 64) partial[(int64_t)wg * width + k] = acc[k];
}

// Per-world band arguments of a chunk of worlds (with WorldArgs: 3.3 KB of
// kernel arguments for eight worlds).
struct PolyWorlds {
    PolyArgs p[kWorldChunk];
    double* rows[kWorldChunk];         // this step's force rows, or null
    const double* prev[kWorldChunk];   // the previous step's rows, or null
    double* prev_out[kWorldChunk];     // where their sums go
    // the previous chunk's rows of the call's last step and where their sums
    // go (null: none), summed by this launch's band workgroups of world 0
    const double* xrows[kWorldChunk];
    double* xout[kWorldChunk];
    int32_t xwidth[kWorldChunk];
};
static_assert(sizeof(WorldArgs) + sizeof(PolyWorlds) <= 4000, "chunk arguments exceed the kernel-argument block");

// The closing sums of a chunk (k_sum_partials of each world's last rows):
// output blockIdx.x of world blockIdx.y.
__global__ void __launch_bounds__(64) k_sum_rows_worlds(WorldArgs w, PolyWorlds p)
{
    const int j = blockIdx.y, k = blockIdx.x, width = w.a[j].nsolid * 6;
    const double* rows = p.prev[j];
    if (!rows || k >= width) return;
    const double v = wave_sum(rows_lane_sum(rows, width, k, threadIdx.x));
    if (threadIdx.x == 0) p.prev_out[j][k] = v;
}

// The plain pull of the cells the band leaves alone: a cell whose interior
// projection lies in a bounding circle is the band's (the band's box holds
// every such cell), so it is skipped here.
template <typename T, int DT>
__device__ __forceinline__ void poly_gather_cell(const EnvArgs& a, const PolyArgs& p, int bx, int by)
{
    const int x = a.region[0] + bx, y = a.region[2] + by;
    if (x >= a.region[1] || y >= a.region[3]) return;
    const int px = min(max(x, 1), a.nx - 2), py = min(max(y, 1), a.ny - 2);
    if (poly_near<T>(a, p, px, py)) return;
    gather_d2q9_at<T, DT>(a, x, y);
}

// One step in one launch: the band and the plain pull write disjoint cells
// and both only read the source state, so workgroups 0 .. kPolyWaves - 1
// run the band (workgroup b in the role of band wave b of kPolyWaves) and
// the rest pull 64 cells of one row each (gw workgroups per row).
template <typename T, int DT, bool LDSV>
__device__ __forceinline__ void poly_step(const EnvArgs& a, const PolyArgs& p, double* partial, const double* prev,
                                          double* prev_out, int b, int gw)
{
    if (b < kPolyWaves) {
        poly_band<T, DT, LDSV>(a, p, partial, prev, prev_out, b, kPolyWaves);
        return;
    }
    const int g = b - kPolyWaves;
    poly_gather_cell<T, DT>(a, p, (g % gw) * 64 + (int)threadIdx.x, g / gw);
}

template <typename T, int DT, bool LDSV>
__global__ void __launch_bounds__(64) k_poly_step(EnvArgs a, PolyArgs p, double* partial, const double* prev,
                                                   double* prev_out, int gw)
{
    poly_step<T, DT, LDSV>(a, p, partial, prev, prev_out, blockIdx.x, gw);
}

// World blockIdx.y of the chunk; the grid covers its largest region.
template <typename T, int DT, bool LDSV>
__global__ void __launch_bounds__(64) k_poly_step_worlds(WorldArgs w, PolyWorlds p, int gw)
{
    const int k = blockIdx.y;
    poly_step<T, DT, LDSV>(w.a[k], p.p[k], p.rows[k], p.prev[k], p.prev_out[k], blockIdx.x, gw);
    if (k != 0 || blockIdx.x >= kPolyWaves) return;
    // output o of the previous chunk's columns (worlds in order) by band
    // workgroup o mod kPolyWaves, in the order of k_sum_partials
    int base = 0;
    for (int j = 0; j < kWorldChunk; ++j) {
        const double* rows = p.xrows[j];
        const int wd = rows ? p.xwidth[j] : 0;
        for (int o = (int)blockIdx.x; o < base + wd; o += kPolyWaves) {
            if (o < base) continue;
            const double v = wave_sum(rows_lane_sum(rows, wd, o - base, threadIdx.x));
            if (threadIdx.x == 0) p.xout[j][o - base] = v;
        }
        base += wd;
    }
}

template <typename T, int DT>
hipError_t launch_step_worlds(const WorldArgs& w, const PolyWorlds& p, int n, hipStream_t s)
{
    int nv = 0, wmax = 0, hmax = 0;
    for (int k = 0; k < n; ++k) {
        nv = max(nv, w.a[k].nsolid * p.p[k].maxv * 2);
        const int* r = w.a[k].region;
        wmax = max(wmax, r[1] - r[0]);
        hmax = max(hmax, r[3] - r[2]);
    }
    const int gw = (max(wmax, 0) + 63) / 64;
    const dim3 grid(kPolyWaves + gw * max(hmax, 0), n);
    if (nv <= kPolyLdsFloats)
        hipLaunchKernelGGL((k_poly_step_worlds<T, DT, true>), grid, 64, (size_t)nv * sizeof(float), s, w, p,
                           max(gw, 1));
    else
        hipLaunchKernelGGL((k_poly_step_worlds<T, DT, false>), grid, 64, 0, s, w, p, max(gw, 1));
    return hipGetLastError();
}

template <typename T, int DT>
hipError_t launch_step(const EnvArgs& a, const PolyArgs& p, double* partial, const double* prev, double* prev_out,
                       hipStream_t s)
{
    const int nv = a.nsolid * p.maxv * 2;
    const int w = max(a.region[1] - a.region[0], 0), h = max(a.region[3] - a.region[2], 0);
    const int gw = (w + 63) / 64;
    const dim3 grid(kPolyWaves + gw * h);
    if (nv <= kPolyLdsFloats)
        hipLaunchKernelGGL((k_poly_step<T, DT, true>), grid, 64, (size_t)nv * sizeof(float), s, a, p, partial, prev,
                           prev_out, max(gw, 1));
    else
        hipLaunchKernelGGL((k_poly_step<T, DT, false>), grid, 64, 0, s, a, p, partial, prev, prev_out, max(gw, 1));
    return hipGetLastError();
}

}  // namespace

// bbox, then two banks of force rows (a step writes one while summing the other)
size_t polygon_workspace_bytes() { return 256 + 2 * (size_t)kPolyWaves * HL_MAX_SOLIDS * 6 * sizeof(double); }

hipError_t step_polygons(int dtype, const EnvArgs& a0, int nsteps, const float* verts, const float* verts_last,
                         const int32_t* counts, int maxv, const float* radius2, double* force_history, void* ws,
                         hipStream_t s)
{
    double* partial = reinterpret_cast<double*>(static_cast<char*>(ws) + 256);
    PolyArgs p{verts, verts_last, counts, maxv, radius2};
    hipError_t e = hipSuccess;
    const int width = a0.nsolid * 6;
    const double* prev = nullptr;  // the previous step's rows, summed by this step's band launch
    double* prev_out = nullptr;
    for (int k = 0; k < nsteps && e == hipSuccess; ++k) {
        EnvArgs a = a0;
        if (k & 1) {
            a.src = a0.dst;
            a.dst = const_cast<void*>(a0.src);
        }
        a.mask = nullptr;
        a.cut_bits = nullptr;
        a.cut_cols = nullptr;
        a.box[0] = a.box[2] = a.box[4] = 1 << 29;
        a.box[1] = a.box[3] = a.box[5] = -(1 << 29);
        // steps after the first see the same vertices twice (bodies at rest)
        p.verts_last = k == 0 ? verts_last : verts;
        double* fout = force_history && width > 0 ? force_history + (int64_t)k * width : nullptr;
        double* rows = fout ? partial + (size_t)(k & 1) * kPolyWaves * HL_MAX_SOLIDS * 6 : nullptr;
        switch (dtype) {
        case HL_F32: e = launch_step<float, HL_F32>(a, p, rows, prev, prev_out, s); break;
        case HL_F64: e = launch_step<double, HL_F64>(a, p, rows, prev, prev_out, s); break;
        case HL_F16: e = launch_step<float, HL_F16>(a, p, rows, prev, prev_out, s); break;
        case HL_BF16: e = launch_step<float, HL_BF16>(a, p, rows, prev, prev_out, s); break;
        default: e = hipErrorInvalidValue;
        }
        prev = rows;
        prev_out = fout;
    }
    if (e == hipSuccess && prev) e = sum_partials(prev, kPolyWaves, width, prev_out, s);
    return e;
}

size_t polygon_worlds_workspace_bytes(int64_t total_width)
{
    return 256 + (size_t)(total_width > 0 ? total_width : 0) * 2 * kPolyWaves * sizeof(double);
}

hipError_t step_polygons_worlds(int dtype, const EnvArgs* a0, int nworld, int nsteps, const float* const* verts,
                                const float* const* verts_last, const int32_t* const* counts, const int32_t* maxv,
                                const float* const* radius2, double* const* force_history, void* ws, hipStream_t s)
{
    if (nworld < 1 || nsteps < 0) return hipErrorInvalidValue;
    // a world with forces keeps two banks of kPolyWaves rows of its width,
    // as step_polygons keeps them in its own workspace; worlds in order
    double* rows0 = reinterpret_cast<double*>(static_cast<char*>(ws) + 256);
    std::vector<size_t> at((size_t)nworld);
    size_t off = 0;
    for (int wi = 0; wi < nworld; ++wi) {
        at[(size_t)wi] = off;
        if (force_history && force_history[wi]) off += (size_t)2 * kPolyWaves * a0[wi].nsolid * 6;
    }
    hipError_t e = hipSuccess;
    PolyWorlds before{};  // the previous chunk's arguments in this step
    for (int k = 0; k <= nsteps && e == hipSuccess; ++k) {
        for (int c0 = 0; c0 < nworld && e == hipSuccess; c0 += kWorldChunk) {
            const int n = nworld - c0 < kWorldChunk ? nworld - c0 : kWorldChunk;
            WorldArgs w{};
            PolyWorlds p{};
            if (k == nsteps - 1 && c0 > 0)  // the call's last step: sum the previous chunk's rows here
                for (int j = 0; j < kWorldChunk; ++j) {
                    p.xrows[j] = before.rows[j];
                    p.xout[j] = before.xout[j];
                    p.xwidth[j] = before.xwidth[j];
                }
            int maxwidth = 0;
            for (int j = 0; j < n; ++j) {
                const int wi = c0 + j;
                EnvArgs a = a0[wi];
                if (k & 1) {
                    a.src = a0[wi].dst;
                    a.dst = const_cast<void*>(a0[wi].src);
                }
                a.mask = nullptr;
                a.cut_bits = nullptr;
                a.cut_cols = nullptr;
                a.box[0] = a.box[2] = a.box[4] = 1 << 29;
                a.box[1] = a.box[3] = a.box[5] = -(1 << 29);
                w.a[j] = a;
                const int width = a.nsolid * 6;
                maxwidth = width > maxwidth ? width : maxwidth;
                p.p[j] = PolyArgs{verts[wi], k == 0 ? verts_last[wi] : verts[wi], counts[wi], maxv[wi], radius2[wi]};
                double* fh = force_history && width > 0 ? force_history[wi] : nullptr;
                double* wr = rows0 + at[(size_t)wi];
                const size_t bank = (size_t)kPolyWaves * width;
                p.rows[j] = fh && k < nsteps ? wr + (size_t)(k & 1) * bank : nullptr;
                p.prev[j] = fh && k > 0 ? wr + (size_t)((k - 1) & 1) * bank : nullptr;
                p.prev_out[j] = fh && k > 0 ? fh + (int64_t)(k - 1) * width : nullptr;
            }
            // for the next chunk: where this chunk's rows of this step go
            before = p;
            for (int j = 0; j < kWorldChunk; ++j) {
                const int wi = c0 + j;
                const int width = j < n ? w.a[j].nsolid * 6 : 0;
                before.xout[j] = p.rows[j] ? force_history[wi] + (int64_t)k * width : nullptr;
                before.xwidth[j] = width;
            }
            if (k == nsteps) {  // the last chunk's last rows: summed by a launch of their own
                if (maxwidth > 0 && c0 + kWorldChunk >= nworld) {
                    hipLaunchKernelGGL(k_sum_rows_worlds, dim3(maxwidth, n), 64, 0, s, w, p);
                    e = hipGetLastError();
                }
                continue;
            }
            switch (dtype) {
            case HL_F32: e = launch_step_worlds<float, HL_F32>(w, p, n, s); break;
            case HL_F64: e = launch_step_worlds<double, HL_F64>(w, p, n, s); break;
            case HL_F16: e = launch_step_worlds<float, HL_F16>(w, p, n, s); break;
            case HL_BF16: e = launch_step_worlds<float, HL_BF16>(w, p, n, s); break;
            default: e = hipErrorInvalidValue;
            }
        }
    }
    return e;
}

}  // namespace hl
