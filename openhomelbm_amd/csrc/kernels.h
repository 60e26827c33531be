// Internal launcher interface between the C ABI (capi.hip) and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "homelbm.h"
#include "lbm_device.h"

namespace hl {

struct TileCfg {
    int tx, ty, zc;
};

// Half-open block of cells [x0, x0+w) x [y0, y0+h) x [z0, z0+d).
struct Region {
    int x0, y0, z0, w, h, d;
    __host__ __device__ int64_t n() const { return (int64_t)w * h * d; }
};

// Near-body work that the tiled D3Q27 kernel's workgroups share ahead of
// their tiled pass: the cells of `list` (count in *count, region order),
// updated with the wall rule in the reference's operation order; with forces,
// each of the nfix participating workgroups writes one FP64 partial row
// (sum_partials adds them).
struct FixArgs {
    Region rg;
    const int* list;
    const int* count;
    double* partial;   // [nfix][width], null without forces
    int width;         // nsolid * 6
    int max_waves;     // rows available in partial
    int nfix;          // workgroups sharing the near-body work (set by the launcher)
    // the previous step's force rows, summed by this launch (one wave per
    // output, the order of sum_partials): out[k] = sum_r prev[r * width + k]
    const double* prev;
    int prev_rows;
    double* prev_out;
};
bool tiling_supported(int lattice, int tx, int ty, int zc);
void set_tiling(int lattice, TileCfg c);
TileCfg get_tiling(int lattice);

hipError_t step_gather(int lattice, int dtype, const EnvArgs& a, hipStream_t s);
// Worlds stepped by one launch: their arguments travel in the kernel-argument
// block (2.8 KB for eight), the world index is a grid dimension.
constexpr int kWorldChunk = 8;
struct WorldArgs {
    EnvArgs a[kWorldChunk];
};
// D2Q9 gather step that also writes fx.nfix (set here) force rows of the
// state it reads into fx.partial and sums fx.prev into fx.prev_out.
hipError_t step_gather_forces(int dtype, const EnvArgs& a, FixArgs& fx, hipStream_t s);
// Several D3Q27 lattices stepped by one tiled launch: their arguments
// travel in the kernel-argument block (about 3.1 KB for six).
constexpr int kTiledWorlds = 6;
struct TiledWorlds {
    EnvArgs a[kTiledWorlds];
    FixArgs fx[kTiledWorlds];
    int base[kTiledWorlds + 1];  // first workgroup of each world (multiples of 8)
    int gx[kTiledWorlds], gy[kTiledWorlds], ntile[kTiledWorlds];
    int zc, n;
};
// step_tiled of n <= kTiledWorlds worlds (one dtype) in one launch; fx[k] is
// null or world k's near-body arguments (its nfix is set here). The results
// are those of step_tiled world by world. hipErrorNotSupported when a world
// has no tiled kernel or the worlds disagree on the body-force field.
hipError_t step_tiled_worlds(int dtype, const EnvArgs* a, FixArgs* const* fx, int n, hipStream_t s);
// Whether the tiled kernel can step this lattice with the current tile shape.
bool tiled_fits(const EnvArgs& a, int dtype);
// Returns hipErrorNotSupported when no tiled kernel exists for this case.
// With fx (near-body list built by cut_list, a.cut_bits set by
// cut_bits_view), the same launch also updates the listed cells and, with
// fx->partial, writes fx->nfix force rows (set here) for sum_partials.
hipError_t step_tiled(int lattice, int dtype, const EnvArgs& a, FixArgs* fx, hipStream_t s);
// Near-body arguments for one step over ws (null list when there is no mask;
// partial rows only when forces are wanted).
FixArgs fixup_args(int lattice, const EnvArgs& a, void* ws, bool forces, int parity = 0);
// out[k] = sum over rows r of partial[r * width + k], in a fixed order.
hipError_t sum_partials(const double* partial, int rows, int width, double* out, hipStream_t s);


hipError_t init_uniform(int lattice, int dtype, const EnvArgs& a, void* buf, const double m[10], hipStream_t s);
// a.src -> a.dst over the whole state span (plain streaming copy).
hipError_t copy_state(int lattice, int dtype, const EnvArgs& a, hipStream_t s);
hipError_t voxelize(int lattice, const EnvArgs& a, uint8_t* mask, const float* caps, int ncap, const int32_t box[6],
                    hipStream_t s);
hipError_t voxelize_polygons(const EnvArgs& a, uint8_t* mask, const float* verts, const int32_t* counts, int npoly,
                             int maxv, const int32_t box[6], hipStream_t s);
// ws: workspace of at least forces_workspace_bytes() bytes.
size_t forces_workspace_bytes(const EnvArgs& a, int lattice);
hipError_t forces(int lattice, int dtype, const EnvArgs& a, double* out, void* ws, hipStream_t s);
// Lists, in region order, the cells whose interior projection has a cut link
// (into ws, sized by forces_workspace_bytes); valid until the mask changes.
hipError_t cut_list(int lattice, const EnvArgs& a, void* ws, hipStream_t s);
// Points a.cut_bits at the list's cell bitmap in ws (null when no mask).
void cut_bits_view(int lattice, EnvArgs& a, void* ws);
// Near-body pass over that list: re-runs each listed cell in the reference's
// operation order with the wall rule (when store), and/or writes the
// per-solid force/torque of the state read (when out != NULL).
hipError_t fixup(int lattice, int dtype, const EnvArgs& a, bool store, double* out, void* ws, hipStream_t s);
size_t mass_workspace_bytes(const EnvArgs& a, int lattice);

// Shared with the batch entry points (defined in capi.hip): set the calling
// thread's hl_last_error message and return code; validate a descriptor;
// kernel arguments of a descriptor reading state cur.
int32_t report(int32_t code, const char* msg);
int32_t check_env(const hl_env* e);
EnvArgs env_args(const hl_env* e, int cur);
hipError_t mass(int lattice, int dtype, const EnvArgs& a, double* out, void* ws, hipStream_t s);

// Render images (render.hip): one plane (normal 2: xy at z = index, normal 0:
// yz at x = index) or one projection along an axis; gl is the grid length.
hipError_t render_plane(int dtype, const EnvArgs& a, int kind, int normal, int index, float gl, float thickness,
                        float vmax, float* out, hipStream_t s);
hipError_t render_projection(int dtype, const EnvArgs& a, int kind, int axis, float gl, float* out, hipStream_t s);
int max_render_thickness();

// Triangle-mesh voxelizer (mesh.hip): tris [ntri][9] posed vertices, solid
// ids [ntri]; rewrites the box (merge: only sets cells still 0).
hipError_t voxelize_mesh(const EnvArgs& a, uint8_t* mask, const float* tris, const int32_t* tsolid, int ntri,
                         const int32_t box[6], int merge, int32_t* overflow, hipStream_t s);

// Exact polygon cut cells for D2Q9 (polygon.hip): nsteps steps from a0 (its
// src/dst swap every step) with the segment-link predicate; ws holds at
// least polygon_workspace_bytes(); force_history [nsteps][nsolid * 6] or null.
size_t polygon_workspace_bytes();
hipError_t step_polygons(int dtype, const EnvArgs& a0, int nsteps, const float* verts, const float* verts_last,
                         const int32_t* counts, int maxv, const float* radius2, double* force_history, void* ws,
                         hipStream_t s);
// step_polygons of nworld worlds that share nothing but the stream: per step,
// one gather and one band launch per kWorldChunk worlds; the results are
// those of step_polygons world by world. Per-world arrays of device
// pointers (force_history entries may be null); ws at least
// polygon_worlds_workspace_bytes(sum of nsolid * 6 over the worlds with forces).
size_t polygon_worlds_workspace_bytes(int64_t total_width);
hipError_t step_polygons_worlds(int dtype, const EnvArgs* a0, int nworld, int nsteps, const float* const* verts,
                                const float* const* verts_last, const int32_t* const* counts, const int32_t* maxv,
                                const float* const* radius2, double* const* force_history, void* ws, hipStream_t s);

// Coupling with a rigid-body engine (coupling.hip); see include/homelbm.h.
hipError_t couple_poses(int lattice, int nworld, int nbody, int nsolid, const float* pos, int pstride,
                        const float* quat, int qstride, const int32_t* solid, const float* eorigin,
                        const float* lorigin, const float* scale, float* const* tables, int32_t* init, hipStream_t s);
hipError_t couple_forces(int lattice, int mode, int nworld, int nbody, int nsolid, const double* rows,
                         int64_t wstride, const int32_t* solid, const float* scale, float fconv, float tconv,
                         float* fout, float* tout, float* xfrc, int xbodies, const int32_t* body, hipStream_t s);
hipError_t free_bodies_step(int nworld, int nbody, float* state, const float* params, const float* force,
                            const float* torque, const float g[3], float dt, hipStream_t s);
hipError_t body_capsules(int nworld, int nbody, const float* pos, int pstride, const float* quat, int qstride,
                         const float* shapes, const int32_t* solid, const float* eorigin, const float* lorigin,
                         const float* scale, float* caps, hipStream_t s);

}  // namespace hl
