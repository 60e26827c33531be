/*
 * CPU oracle for the HOME-LBM time step. TEST INFRASTRUCTURE ONLY: used by
 * tests/, by the smoke check and by the benchmark's CPU-baseline leg; never
 * by the product. Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include "homelbm_oracle.h"

#include <math.h>
#include <stddef.h>

/* Lattice tables: D2Q9 from envs/lbm/lbm_core.py:127-130, D3Q27 from
 * envs/lbm3d/lbm_core_3d.py:12-35 (same direction order; the order is part
 * of the moment-sum and wall-rule index arithmetic). */
static const int D2_CX[9] = {0, 1, -1, 0, 0, 1, -1, -1, 1};
static const int D2_CY[9] = {0, 0, 0, 1, -1, 1, 1, -1, -1};
static const int D2_INV[9] = {0, 2, 1, 4, 3, 7, 8, 5, 6};
static const int D3_CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
static const int D3_CY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
static const int D3_CZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
static const int D3_INV[27] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25};

/* Weights are the FP64 fractions rounded once to the working precision, as
 * the reference builds them (wp.types.vector(..., dtype=float32)(4/9, ...)). */
#define D2W(T) {(T)(4.0 / 9.0), (T)(1.0 / 9.0), (T)(1.0 / 9.0), (T)(1.0 / 9.0), (T)(1.0 / 9.0), \
                (T)(1.0 / 36.0), (T)(1.0 / 36.0), (T)(1.0 / 36.0), (T)(1.0 / 36.0)}
#define D3W(T) {(T)(8.0 / 27.0), (T)(2.0 / 27.0), (T)(2.0 / 27.0), (T)(2.0 / 27.0), (T)(2.0 / 27.0), \
                (T)(2.0 / 27.0), (T)(2.0 / 27.0), (T)(1.0 / 54.0), (T)(1.0 / 54.0), (T)(1.0 / 54.0), \
                (T)(1.0 / 54.0), (T)(1.0 / 54.0), (T)(1.0 / 54.0), (T)(1.0 / 54.0), (T)(1.0 / 54.0), \
                (T)(1.0 / 54.0), (T)(1.0 / 54.0), (T)(1.0 / 54.0), (T)(1.0 / 54.0), (T)(1.0 / 216.0), \
                (T)(1.0 / 216.0), (T)(1.0 / 216.0), (T)(1.0 / 216.0), (T)(1.0 / 216.0), (T)(1.0 / 216.0), \
                (T)(1.0 / 216.0), (T)(1.0 / 216.0)}

static const float d2_w_f32[9] = D2W(float);
static const double d2_w_f64[9] = D2W(double);
static const float d3_w_f32[27] = D3W(float);
static const double d3_w_f64[27] = D3W(double);
/* cs2: lbm_func.py:4 (sum of weights, rounded to float32) and lbm_core_3d.py:37. */
static const float cs2_2d_f32 = (float)(1.0 / 9 + 1.0 / 9 + 1.0 / 36 + 1.0 / 36 + 1.0 / 36 + 1.0 / 36);
static const double cs2_2d_f64 = 1.0 / 3.0;
static const float cs2_3d_f32 = (float)(1.0 / 3.0);
static const double cs2_3d_f64 = 1.0 / 3.0;
/* the literal 1.0/3.0 inside mlCalEquilibrium3D_s, evaluated in float32 */
static const float third_f32 = 1.0f / 3.0f;
static const double third_f64 = 1.0 / 3.0;

#define R float
#define SFX f32
#define RSQRT sqrtf
#include "oracle_impl.h"
#undef R
#undef SFX
#undef RSQRT

#define R double
#define SFX f64
#define RSQRT sqrt
#include "oracle_impl.h"
#undef R
#undef SFX
#undef RSQRT

void hlo_step_omp_f32(const hlo_env* e, const float* src, float* dst)
{
    const int64_t st = e->stride;
    if (e->lattice == 9) {
#pragma omp parallel for schedule(static)
        for (int y = 1; y < e->ny - 1; ++y)
            for (int x = 1; x < e->nx - 1; ++x) {
                float out[6];
                d2q9_cell_f32(e, src, x, y, out);
                int64_t c = (int64_t)x + (int64_t)e->nx * y;
                for (int f = 0; f < 6; ++f) dst[f * st + c] = out[f];
            }
    } else {
        const int ny = e->ny, nz = e->nz;
#pragma omp parallel for schedule(static) collapse(2)
        for (int z = 1; z < nz - 1; ++z)
            for (int y = 1; y < ny - 1; ++y)
                for (int x = 1; x < e->nx - 1; ++x) {
                    float out[10];
                    d3q27_cell_f32(e, src, x, y, z, out);
                    int64_t c = (int64_t)x + (int64_t)e->nx * ((int64_t)y + (int64_t)e->ny * z);
                    for (int f = 0; f < 10; ++f) dst[f * st + c] = out[f];
                }
    }
    bc_passes_f32(e, dst);
}

/* hlo_step_capsules_f32 without forces, parallel over cells with OpenMP
 * (identical results). */
void hlo_step_capsules_omp_f32(const hlo_env* e, const float* caps, int32_t ncap, const float* src, float* dst)
{
    const int64_t st = e->stride;
    if (e->lattice != 27) return;
    const int ny = e->ny, nz = e->nz;
#pragma omp parallel for schedule(static) collapse(2)
    for (int z = 1; z < nz - 1; ++z)
        for (int y = 1; y < ny - 1; ++y)
            for (int x = 1; x < e->nx - 1; ++x) {
                float out[10];
                d3q27_cell_rule_f32(e, src, x, y, z, caps, ncap, NULL, out);
                int64_t c = (int64_t)x + (int64_t)e->nx * ((int64_t)y + (int64_t)e->ny * z);
                for (int f = 0; f < 10; ++f) dst[f * st + c] = out[f];
            }
    bc_passes_f32(e, dst);
}

/* Capsule voxelizer (no reference counterpart: the reference tests link
 * crossings against polygon edges / a BVH instead). A cell centre p is inside
 * capsule (a, b, r) when its squared distance to segment ab is <= r*r; the
 * mask takes the lowest-index capsule containing the cell. FP32 with no
 * contraction, evaluated in the same order as the GPU voxelizer. */
void hlo_voxelize(const hlo_env* e, const float* caps, int32_t ncap, const int32_t box[6], uint8_t* mask)
{
    const int nx = e->nx, ny = e->ny;
    for (int z = box[4]; z < box[5]; ++z)
        for (int y = box[2]; y < box[3]; ++y)
            for (int x = box[0]; x < box[1]; ++x) {
                float px = (float)x, py = (float)y, pz = (float)z;
                uint8_t v = 0;
                for (int k = 0; k < ncap; ++k) {
                    const float* C = caps + 8 * k;
                    float abx = C[3] - C[0], aby = C[4] - C[1], abz = C[5] - C[2];
                    float apx = px - C[0], apy = py - C[1], apz = pz - C[2];
                    float len2 = abx * abx + aby * aby + abz * abz;
                    float t = 0.0f;
                    if (len2 > 0.0f) {
                        t = (apx * abx + apy * aby + apz * abz) / len2;
                        t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
                    }
                    float dx = apx - t * abx, dy = apy - t * aby, dz = apz - t * abz;
                    float d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 <= C[6] * C[6]) { v = (uint8_t)((int)C[7] + 1); break; }
                }
                mask[(int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z)] = v;
            }
}

/* Even-odd +x ray test of get_solid_boundary_img (envs/lbm/lbm_func.py:804-864,
 * with border thickness 0): an edge (p1, p2) counts when it straddles the
 * row (y1 > py) != (y2 > py) and its crossing x1 + (py - y1)(x2 - x1)/(y2 - y1)
 * lies right of px; an odd count means inside. FP32, no contraction. */
void hlo_voxelize_polygons(const hlo_env* e, const float* verts, const int32_t* counts, int32_t npoly, int32_t maxv,
                           const int32_t box[6], uint8_t* mask)
{
    const int nx = e->nx;
    for (int y = box[2]; y < box[3]; ++y)
        for (int x = box[0]; x < box[1]; ++x) {
            const float px = (float)x, py = (float)y;
            uint8_t v = 0;
            for (int k = 0; k < npoly && !v; ++k) {
                const float* P = verts + 2 * (int64_t)maxv * k;
                const int n = counts[k];
                int crossings = 0;
                for (int j = 0; j < n; ++j) {
                    const float x1 = P[2 * j], y1 = P[2 * j + 1];
                    const int j2 = (j + 1) % n;
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

double hlo_mass_f32(const hlo_env* e, const float* src)
{
    const int64_t n = (int64_t)e->nx * e->ny * e->nz;
    double s = 0.0;
    for (int64_t c = 0; c < n; ++c) s += (double)src[c];
    return s;
}
