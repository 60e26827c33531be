/*
 * CPU oracle for the HOME-LBM time step (D2Q9 and D3Q27).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP path is
 * compared against, and the scalar/OpenMP CPU baseline timed beside it. The
 * product (openhomelbm_amd) never links, loads or calls it.
 *
 * It restates the reference's Warp kernels in plain C. Each function keeps the
 * reference's operation order, so with -ffp-contract=off the FP32 build gives
 * the rounding of a literal evaluation of the Warp source. Citations name the
 * reference file:line each function follows.
 *
 * Storage: structure-of-arrays. Cell (x, y, z) is at x + nx*(y + ny*z), and
 * moment field f starts at f*stride. Field order:
 *   D2Q9  : rho, ux, uy, Sxx, Syy, Sxy                      (6)
 *   D3Q27 : rho, ux, uy, uz, Sxx, Syy, Szz, Sxy, Sxz, Syz    (10)
 *
 * Parity status per function (how each is pinned; tests/ hold the checks):
 *   lattice tables, reconstruction, collision, BC, init : closed forms and
 *     moment identities, plus agreement with the independent NumPy
 *     restatement (oracle/np_oracle.py) on seeded tiny grids;
 *   mask voxelizer : integer closed forms (disc/sphere) and brute force;
 *   wall rule / force : closed forms on a static body (u_p = 0) and the
 *     mirror-symmetry zero-lift invariant, plus the NumPy restatement.
 *   The reference ships no golden vectors and cannot run here (Warp is not
 *   installed), so no function is pinned against reference outputs.
 */
#ifndef HOMELBM_ORACLE_H
#define HOMELBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hlo_env {
    int32_t lattice;          /* 9 or 27 */
    int32_t nx, ny, nz;       /* nz = 1 for D2Q9 */
    int64_t stride;           /* elements between moment fields (>= cells) */
    const uint8_t* mask;      /* nullable: 0 = fluid, k>0 = solid k-1 */
    const float* force;       /* nullable: per-cell body force [dim][stride] */
    const float* solids;      /* pose table: 8 floats/solid (2D), 24 (3D) */
    int32_t nsolid;
    float nu;                 /* lattice kinematic viscosity */
    int32_t bc_type[6];       /* 2D: left, top, right, bottom; 3D: left, right, top, bottom, front, back */
    float bc_value[6][3];
} hlo_env;

/* One time step: src -> dst (both [nfields][stride]). */
void hlo_step_f32(const hlo_env* e, const float* src, float* dst);
void hlo_step_f64(const hlo_env* e, const double* src, double* dst);
/* Same step, parallel over cells with OpenMP (identical results). */
void hlo_step_omp_f32(const hlo_env* e, const float* src, float* dst);

/* Momentum-exchange force and torque per solid from the state `src` (the
 * state the step reads). out: [nsolid][6] = Fx, Fy, Fz, Tx, Ty, Tz (FP64). */
void hlo_forces_f32(const hlo_env* e, const float* src, double* out);
void hlo_forces_f64(const hlo_env* e, const double* src, double* out);

/* Uniform-flow initialisation of one moment buffer. */
void hlo_init_uniform_f32(const hlo_env* e, float ux, float uy, float uz, float* dst);
void hlo_init_uniform_f64(const hlo_env* e, double ux, double uy, double uz, double* dst);

/* Per-direction reconstruction (exposed for moment-identity tests). */
float hlo_recon_f32(int lattice, const float* m, int i);
double hlo_recon_f64(int lattice, const double* m, int i);

/* Capsule voxelizer. caps: [ncap][8] = ax, ay, az, bx, by, bz, radius, solid_id.
 * Writes mask over the half-open box [x0,x1)x[y0,y1)x[z0,z1); other cells untouched. */
void hlo_voxelize(const hlo_env* e, const float* caps, int32_t ncap,
                  const int32_t box[6], uint8_t* mask);

/* Even-odd polygon voxelizer (2D). verts: [npoly][maxv][2] lattice
 * coordinates of closed polygons, counts[npoly] vertices each. Writes mask
 * over the half-open box; lowest polygon index wins. */
void hlo_voxelize_polygons(const hlo_env* e, const float* verts, const int32_t* counts, int32_t npoly,
                           int32_t maxv, const int32_t box[6], uint8_t* mask);

/* Polygon bodies for the exact cut-cell step (2D). verts / verts_last:
 * [nsolid][maxv][2] transformed vertices of this step and of the previous
 * one (precompute_transformed_segments, envs/lbm/lbm_func.py:509-542),
 * counts[nsolid] vertices each; radius2[nsolid] = the squared bounding
 * radius in lattice units, (solid_max_radius * scale) * scale in FP32. */
typedef struct hlo_polys {
    const float* verts;
    const float* verts_last;
    const int32_t* counts;
    int32_t maxv;
    const float* radius2;
} hlo_polys;

/* One D2Q9 step with the reference's segment-link predicate (get_cutcell)
 * and material-point wall rule instead of the mask rule; forces (nullable,
 * [nsolid][6] FP64, Fx, Fy at 0, 1 and the torque at 5) of the state read. */
void hlo_step_polygons_f32(const hlo_env* e, const hlo_polys* p, const float* src, float* dst, double* forces);
void hlo_step_polygons_f64(const hlo_env* e, const hlo_polys* p, const double* src, double* dst, double* forces);

/* Capsule cut links (3D). caps: [ncap][8] = ax, ay, az, bx, by, bz, radius,
 * solid_id in lattice coordinates at the current pose (the voxelizer's
 * table). hlo_capsule_cut: the crossing of link i from (x, y, z) towards
 * x - c_i nearest to the cell, over every capsule surface (the ray query of
 * get_cutcell_multi_3d, lbm_func_3d.py:254-316): returns the solid (-1: none)
 * and t in [0, 1]. hlo_step_capsules: one D3Q27 step cutting those links,
 * the wall point at the crossing; forces (nullable, [nsolid][6] FP64) of the
 * state read. env.mask is not read. */
int32_t hlo_capsule_cut_f32(const hlo_env* e, const float* caps, int32_t ncap, float x, float y, float z, int32_t i,
                            float* t_hit);
int32_t hlo_capsule_cut_f64(const hlo_env* e, const float* caps, int32_t ncap, double x, double y, double z, int32_t i,
                            double* t_hit);
void hlo_step_capsules_f32(const hlo_env* e, const float* caps, int32_t ncap, const float* src, float* dst,
                           double* forces);
void hlo_step_capsules_f64(const hlo_env* e, const float* caps, int32_t ncap, const double* src, double* dst,
                           double* forces);
void hlo_step_capsules_omp_f32(const hlo_env* e, const float* caps, int32_t ncap, const float* src, float* dst);

/* FP64 sum of rho over all cells. */
double hlo_mass_f32(const hlo_env* e, const float* src);

#ifdef __cplusplus
}
#endif
#endif
