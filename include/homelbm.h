/*
 * homelbm: MI355X (gfx950) HOME-LBM time step, C ABI.
 *
 * Plain C types only: device pointers, sizes and a descriptor struct. Every
 * entry point returns 0 on success or a negative HL_E_* code; the message of
 * the last failure on the calling thread is returned by hl_last_error().
 * Nothing throws across the ABI. Work is enqueued on the caller's HIP stream
 * (a hipStream_t passed as void*, NULL = default stream) and is asynchronous.
 *
 * The reference exposes a Python class API over NVIDIA Warp and has no C
 * layer; each entry point below names the reference call it replaces. The
 * Python package openhomelbm_amd binds these with ctypes behind the
 * reference's LBM_Solver / LBM_Solver3D / HomeFlow / HomeFlow3D API.
 *
 * Moment storage (owned by the caller, e.g. PyTorch tensors): two ping-pong
 * buffers; field f of cell (x, y, z) at f*stride + x + rpitch*(y + ny*z).
 * Two layouts:
 *   structure of arrays (rpitch = nx, or 0): stride >= nx*ny*nz;
 *   row-interleaved (D3Q27 only; rpitch = 10*nx, stride = nx): the ten fields
 *     of one x-row are consecutive, so a row of cells with all its fields is
 *     one contiguous span (fewer concurrent HBM streams per step).
 * Field order:
 *   D2Q9  : rho, ux, uy, Sxx, Syy, Sxy
 *   D3Q27 : rho, ux, uy, uz, Sxx, Syy, Szz, Sxy, Sxz, Syz
 * 16-bit storage holds (rho - 1 - offset[0])*scale[0] and
 * (value - offset[f])*scale[f] for f > 0 (offsets of zero: rho - 1 and the raw
 * values). FP16 rounds to nearest even; BF16 rounds stochastically (up with
 * the probability of the discarded fraction, the random bits a hash of the
 * element index and the value: deterministic, unbiased, so updates smaller
 * than half a BF16 ulp are not lost).
 */
#ifndef HOMELBM_H
#define HOMELBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HL_D2Q9 9
#define HL_D3Q27 27

#define HL_F32 0
#define HL_F64 1
#define HL_F16 2
#define HL_BF16 3

/* Step kernels. */
#define HL_KERNEL_AUTO 0   /* fastest kernel for the lattice / dtype */
#define HL_KERNEL_GATHER 1 /* per-cell gather in the reference's operation order */
#define HL_KERNEL_TILED 2  /* LDS-staged source-population kernel */

#define HL_OK 0
#define HL_E_INVALID (-1)
#define HL_E_UNSUPPORTED (-2)
#define HL_E_HIP (-3)
#define HL_E_NOMEM (-4)

#define HL_MAX_SOLIDS 64
#define HL_MAX_CAPSULES 4096

typedef struct hl_env {
    int32_t lattice;      /* HL_D2Q9 or HL_D3Q27 */
    int32_t dtype;        /* HL_F32, HL_F64, HL_F16, HL_BF16 */
    int32_t nx, ny, nz;   /* nz = 1 for D2Q9; every extent >= 3 */
    int32_t nsolid;       /* solids in the pose table (<= HL_MAX_SOLIDS) */
    int64_t stride;       /* elements between moment fields (see the layouts above) */
    void* moments[2];     /* ping-pong moment buffers */
    const uint8_t* mask;  /* nullable; 0 = fluid, k > 0 = solid k-1 */
    int32_t mask_box[6];  /* x0, x1, y0, y1, z0, z1 (half-open): mask is 0 outside */
    const float* force;   /* nullable per-cell body force, FP32, component d of cell
                             x + nx*(y + ny*z) at d*force_stride + cell */
    const float* solids;  /* device pose table: 8 floats/solid (2D), 24 (3D) */
    float nu;             /* kinematic viscosity (lattice units) */
    int32_t bc_type[6];   /* 0 velocity, 1 zero gradient, 2 free slip */
    float bc_value[6][3]; /* face velocity for type 0 */
    float scale[10];      /* 16-bit storage scale per field */
    float offset[10];     /* 16-bit storage offset per field (see above) */
    uint64_t mask_epoch;  /* 0, or a caller-kept version of mask + mask_box: hl_step calls on one
                             stream with the same nonzero epoch, mask and box reuse the near-body
                             list the first of them built (a new epoch whenever the mask changes) */
    int64_t rpitch;       /* elements between consecutive x-rows of a field: 0 or nx (structure of
                             arrays) or 10*nx with stride = nx (row-interleaved, D3Q27) */
    int64_t force_stride; /* elements between force components (>= nx*ny*nz); 0 = stride */
} hl_env;

/* Library version (major*10000 + minor*100 + patch). */
int32_t hl_version(void);
/* sizeof(hl_env) as compiled, so bindings can check their struct layout. */
int32_t hl_env_bytes(void);
/* Message of the last failed call on this thread ("" if none). */
const char* hl_last_error(void);

/* Advance nsteps time steps. Buffer `cur` (0/1) holds the current state;
 * afterwards the current state is in buffer cur ^ (nsteps & 1).
 * Replaces LBM_Solver.step (envs/lbm/lbm_solver.py:97-141) and
 * LBM_Solver3D.step (envs/lbm3d/lbm_solver_3d.py:88-127): the fused kernel
 * does stream_and_collide{,_3d} + apply_bc{,_3d}; Swap_Mom{,_3D} becomes the
 * buffer parity. `kernel` selects HL_KERNEL_*. If force_history is not NULL
 * (device, [nsteps][nsolid][6]) step k also stores the per-solid force and
 * torque of the state it reads, as the reference accumulates them inside the
 * step (lbm_func.py:392-401, lbm_func_3d.py:497-511). */
int32_t hl_step(const hl_env* env, int32_t cur, int32_t nsteps, int32_t kernel, double* force_history,
                void* stream);

/* hl_step without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. Same arguments, steps
 * and results as hl_step; its near-body list and force rows live in a
 * file-scope device array of 8 MiB per GPU (held until the process ends)
 * instead of a per-stream workspace. A lattice whose near-body work needs
 * more (a mask box of more than about 1.2 million cells) is rejected with
 * HL_E_NOMEM. Every call on a GPU shares that GPU's array, so calls whose
 * device work overlaps on one GPU (several streams, or host threads) race;
 * the array keeps the last call's contents (a call builds what it reads
 * first, so nothing needs clearing). */
int32_t hl_step_noalloc(const hl_env* env, int32_t cur, int32_t nsteps, int32_t kernel, double* force_history,
                        void* stream);

/* hl_step of nworld lattices (one lattice type and dtype, sizes free), all
 * at buffer `cur`: the world dimension of LBM_Solver3D.step, which launches
 * stream_and_collide_3d over (nworld, nx, ny, nz) in one grid
 * (envs/lbm3d/lbm_solver_3d.py:104-108). D3Q27 worlds that the tiled kernel
 * fits share launches (up to six worlds per launch, each world's near-body
 * cells and force rows as in hl_step); other worlds step one by one. Each
 * world ends bitwise where its own hl_step call would. force_history: NULL
 * or an array of nworld device pointers (each NULL or [nsteps][nsolid][6]).
 * Errors name the world at fault. */
int32_t hl_step_worlds(const hl_env* envs, int32_t nworld, int32_t cur, int32_t nsteps, int32_t kernel,
                       double* const* force_history, void* stream);

/* Per-solid momentum-exchange force and torque of state `cur`, FP64,
 * deterministic. out: device [nsolid][6] = Fx, Fy, Fz, Tx, Ty, Tz.
 * Replaces the atomics of lbm_func.py:392-401 / lbm_func_3d.py:497-511 and the
 * readback LBM_Solver3D.get_forces_and_torques (lbm_solver_3d.py:308-346). */
int32_t hl_forces(const hl_env* env, int32_t cur, double* out, void* stream);

/* Uniform flow into buffer `buf`: set_uniform_flow_kernel
 * (envs/lbm/karman/karman_lbm_env.py:15-53) / set_uniform_flow_3d_kernel
 * (envs/lbm3d/karman/karman_lbm_env_3d.py:15-64); u = 0 is InitFlow{,3D}
 * (lbm_func.py:26-63, lbm_func_3d.py:22-83). */
int32_t hl_init_uniform(const hl_env* env, int32_t buf, float ux, float uy, float uz, void* stream);

/* Capsule voxelizer: writes mask over the half-open box (device caps:
 * [ncap][8] = ax, ay, az, bx, by, bz, radius, solid_id). No reference
 * counterpart (the reference tests links against polygon edges,
 * get_cutcell lbm_func.py:545-607, or a BVH, lbm_func_3d.py:254-316). */
int32_t hl_voxelize(const hl_env* env, uint8_t* mask, const float* caps, int32_t ncap,
                    const int32_t box[6], void* stream);

/* Tile shape (TX x TY columns, ZC planes per block) of the tiled kernel for a
 * lattice; used by the autotuner. Results do not depend on the tile shape.
 * No reference counterpart (Warp chooses its own launch shape). */
int32_t hl_set_tiling(int32_t lattice, int32_t tx, int32_t ty, int32_t zc);
int32_t hl_get_tiling(int32_t lattice, int32_t* tx, int32_t* ty, int32_t* zc);

/* Step-kernel timer: with max_launches > 0, later hl_step calls record a
 * HIP event pair on their stream around one step-kernel launch in every
 * `every` (up to max_launches recorded launches; not while the stream is
 * being captured); max_launches = 0 stops recording. A recorded launch does
 * not overlap its neighbours, so sampling (every > 1) keeps the timed
 * sequence close to an unrecorded one. hl_kernel_timer_read waits for the
 * recorded events, returns the summed launch durations (ms) and their count,
 * and starts a new count. Measurement only: no reference counterpart. */
int32_t hl_kernel_timer(int32_t max_launches, int32_t every);
int32_t hl_kernel_timer_read(double* total_ms, int32_t* launches);

/* Copy state `from` into the other buffer: a plain streaming kernel over the
 * whole state span (16-byte loads and stores). Duplicates a state into both
 * ping-pong buffers, as the reference's init writes both (InitFlow3D,
 * lbm_func_3d.py:22-83); also the HBM streaming ceiling a step is measured
 * against (it moves the step's algorithmic bytes). */
int32_t hl_copy_state(const hl_env* env, int32_t from, void* stream);

/* Even-odd polygon voxelizer (D2Q9): device verts [npoly][maxv][2] closed
 * polygons in lattice coordinates, counts [npoly]; mask = index + 1 of the
 * lowest polygon containing the cell centre, 0 elsewhere in the box.
 * Restates the inside test of get_solid_boundary_img (lbm_func.py:804-864);
 * the mask then drives the wall rule that get_cutcell (lbm_func.py:545-607)
 * drives in the reference. */
int32_t hl_voxelize_polygons(const hl_env* env, uint8_t* mask, const float* verts, const int32_t* counts,
                             int32_t npoly, int32_t maxv, const int32_t box[6], void* stream);

/* Triangle-mesh voxelizer (D3Q27): device tris [ntri][9] (the posed vertices
 * a, b, c of each triangle in lattice coordinates) of closed meshes, solid
 * ids [ntri] (< 64); a cell centre is inside when a +x ray from it crosses
 * its solid's triangles an odd number of times; mask = lowest such solid + 1,
 * 0 elsewhere in the box (merge != 0: only cells still 0 are set). A column
 * meeting more than 64 crossings sets device *overflow (if not NULL).
 * Replaces the BVH link queries of get_cutcell_multi_3d
 * (envs/lbm3d/lbm_func_3d.py:254-316) behind create_solid_from_mesh
 * (envs/lbm3d/lbm_solver_3d.py:129-242): the posed mesh becomes a mask. */
int32_t hl_voxelize_mesh(const hl_env* env, uint8_t* mask, const float* tris, const int32_t* tri_solid,
                         int32_t ntri, const int32_t box[6], int32_t merge, int32_t* overflow, void* stream);

/* D2Q9 steps with the reference's exact polygon cut cells instead of the
 * mask: each link of an interior cell is tested against the segments of the
 * polygons whose bounding circle holds the cell (get_cutcell,
 * envs/lbm/lbm_func.py:545-607), and a cut link takes the wall rule at the
 * crossing's material point, its velocity from this step's and the previous
 * step's vertices, its lever arm from the pose table's position
 * (stream_and_collide, lbm_func.py:346-401); the per-step graph of
 * LBM_Solver.step (lbm_solver.py:97-141) with precompute_transformed_segments
 * done by the caller. Device inputs: verts / verts_last [nsolid][maxv][2]
 * transformed vertices (lbm_func.py:509-542), counts [nsolid], radius2
 * [nsolid] = (solid_max_radius * solid_scale * nx) * (solid_scale * nx) in
 * FP32. verts_last applies to the first step; later steps of the call see
 * verts twice (bodies at rest). env->mask is ignored; force_history as in
 * hl_step (Fx, Fy at 0, 1, the torque scaled by 1/nx at 5). */
int32_t hl_step_polygons(const hl_env* env, int32_t cur, int32_t nsteps, const float* verts,
                         const float* verts_last, const int32_t* counts, int32_t maxv, const float* radius2,
                         double* force_history, void* stream);

/* D3Q27 steps whose cut links are the crossings of capsule surfaces instead
 * of the mask: link i of a cell is cut when the ray from the cell along -c_i
 * meets a capsule surface at t in [0, 1] (the nearest such crossing over
 * every capsule, the first capsule winning ties), and a cut link takes the
 * wall rule at the hit point x - c_i t, its velocity and lever arm from the
 * pose table at that point. Replaces the ray query of get_cutcell_multi_3d
 * (envs/lbm3d/lbm_func_3d.py:254-316) and its use in stream_and_collide_3d
 * (lbm_func_3d.py:424-511) for bodies that are capsules and spheres (a
 * capsule with a == b), whose surfaces cross links in closed form. Every
 * interior cell takes the update, inside a body or not (the reference flags
 * no 3D cell solid). Device input caps: [ncap][8] = ax, ay, az, bx, by, bz,
 * radius, solid_id in lattice coordinates at this pose (hl_voxelize's table;
 * capsules with solid_id outside [0, nsolid) are ignored), at most
 * HL_MAX_CAPSULES. Host input box: a half-open cell box holding every capsule
 * (hl_voxelize's box); links farther than one cell from it are never cut.
 * env->mask is ignored; kernel and force_history as in hl_step (forces of
 * the state each step reads; the torque about the pose table's position). The
 * near-body cells are listed from the capsules at the start of every call.
 * FP32 / FP16 / BF16 take the tiled kernel, FP64 the gather kernel and a
 * near-body pass. */
int32_t hl_step_capsules(const hl_env* env, const float* caps, int32_t ncap, const int32_t box[6], int32_t cur,
                         int32_t nsteps, int32_t kernel, double* force_history, void* stream);

/* hl_step_polygons for nworld independent worlds of one dtype, all at buffer
 * `cur`: the world loop of LBM_Solver.step (envs/lbm/lbm_solver.py:97-141)
 * in one call. Host arrays of nworld entries (maxv, and device pointers for
 * the others); verts_last, force_history or any of their entries may be
 * NULL. Per step, one gather and one band launch serve up to eight worlds;
 * each world's state and force_history are bitwise those of its own
 * hl_step_polygons call. */
int32_t hl_step_polygons_worlds(const hl_env* envs, int32_t nworld, int32_t cur, int32_t nsteps,
                                const float* const* verts, const float* const* verts_last,
                                const int32_t* const* counts, const int32_t* maxv, const float* const* radius2,
                                double* const* force_history, void* stream);

/* FP64 sum of rho over all cells of state `cur` into device out[0]. */
int32_t hl_mass(const hl_env* env, int32_t cur, double* out, void* stream);

/* ------------------------------------------------------------------------
 * Render and diagnostic images of state `cur` into device FP32 out (the
 * reference's u_img / u_img_rgb / u_img_xy / u_img_xz / u_img_xz_front
 * buffers, here with x (or the first in-plane axis) fastest). Vorticity uses
 * central differences inside and one-sided ones at the faces over the grid
 * length 1/nx; solid cells read 1000 in vorticity images. Bodies are the
 * voxel mask: solid where masked or (thickness t, 0 <= t <= 4) within
 * distance t of a masked cell; a projected ray is solid if it meets the mask.
 */
#define HL_RENDER_SPEED 0            /* get_u_img (lbm_func.py:163-166), get_u_img_xy/yz_3d (lbm_func_3d.py:753-776) */
#define HL_RENDER_VORTICITY 1        /* get_vorticity_img (lbm_func.py:196-233), get_vorticity_xy/yz_3d (lbm_func_3d.py:779-840) */
#define HL_RENDER_VORTICITY_SOLID 2  /* get_vorticity_with_solid_img (lbm_func.py:236-319), ..._with_solid_3d (lbm_func_3d.py:843-988) */
#define HL_RENDER_RGB 3              /* get_velocity_rgb (lbm_func.py:169-193): out[3 * pixel + channel] */
#define HL_RENDER_SOLID 4            /* get_solid_boundary_img (lbm_func.py:804-864): 1 solid, 0 fluid */
#define HL_PROJECT_FIRST_SPEED 0     /* get_u_projection_topdown_3d (lbm_func_3d.py:990-1030) */
#define HL_PROJECT_MAX_SPEED 1       /* get_u_projection_{max_topdown,side,front}_3d (lbm_func_3d.py:1033-1201) */
#define HL_PROJECT_MAX_VORTICITY 2   /* get_vorticity_projection_{topdown,side,front}_3d (lbm_func_3d.py:1072-1313) */

/* One plane: normal 2 = the xy plane at z = index (a D2Q9 lattice: index 0),
 * out[y * nx + x]; normal 0 (D3Q27) = the yz plane at x = index,
 * out[z * ny + y]. max_velocity scales HL_RENDER_RGB (> 0). */
int32_t hl_render_plane(const hl_env* env, int32_t cur, int32_t kind, int32_t normal, int32_t index,
                        float thickness, float max_velocity, float* out, void* stream);
/* One D3Q27 projection along axis (0: x, out[z * ny + y]; 1: y,
 * out[z * nx + x]; 2: z, out[y * nx + x]), face cells skipped; the first
 * speed is taken from the high-index end, vorticity is the component about
 * the axis (signed value of the largest magnitude over the interior). */
int32_t hl_render_projection(const hl_env* env, int32_t cur, int32_t kind, int32_t axis, float* out, void* stream);

/* ------------------------------------------------------------------------
 * Coupling with a rigid-body engine, device buffers throughout (no host
 * round trip between the engine's step and the lattice's). Entries whose
 * solid or body index is out of range are skipped.
 */
#define HL_COUPLE_SCALED 0    /* F / scale, torque / scale^2 (2D torque: / scale) */
#define HL_COUPLE_PHYSICAL 1  /* F * force_conv, torque * torque_conv */

/* Engine body poses of nworld worlds -> each world's pose table (tables: a
 * device array of nworld device pose-table pointers, nsolid rows each). Body
 * b of world w is at pos + pos_stride * (w * nbody + b) and quat (w, x, y,
 * z) likewise; its table row solid_ids[b] gets (pos - engine_origin[b]) *
 * scale[b] + lattice_origin[b] and the rotation (2D: the planar angle's cos
 * and sin), after the row's current pose moved to its "last" half (a row not
 * yet initialized gets last = current; initialized [nworld][nsolid] is set).
 * Replaces convert_and_update_solid_batch_2d (envs/lbm/lbm_fluid_env_func.py:72-105)
 * and convert_and_update_solid_batch_3d (envs/lbm3d/lbm_fluid_env_3d_func.py:42-99). */
int32_t hl_couple_poses(int32_t lattice, int32_t nworld, int32_t nbody, int32_t nsolid, const float* pos,
                        int32_t pos_stride, const float* quat, int32_t quat_stride, const int32_t* solid_ids,
                        const float* engine_origin, const float* lattice_origin, const float* scale,
                        float* const* tables, int32_t* initialized, void* stream);

/* FP64 force rows (world w, solid k at rows + w * world_stride + 6 k: F(3),
 * torque(3)) -> engine loads forces / torques [nworld][nbody][3] in `mode`
 * (2D: planar, Fz = 0 and torque about z only), and when xfrc is not NULL
 * also the spatial forces xfrc [nworld][xfrc_bodies][6] at body_ids[b].
 * Replaces extract_forces_torques_batch (envs/lbm/lbm_fluid_env_func.py:51-69),
 * extract_forces_torques_batch_3d / _physical_3d
 * (envs/lbm3d/lbm_fluid_env_3d_func.py:102-185) and fill_xfrc_kernel /
 * fill_xfrc_3d_kernel (lbm_fluid_env_func.py:28-48; lbm_fluid_env_3d_func.py:188-214). */
int32_t hl_couple_forces(int32_t lattice, int32_t mode, int32_t nworld, int32_t nbody, int32_t nsolid,
                         const double* rows, int64_t world_stride, const int32_t* solid_ids, const float* scale,
                         float force_conv, float torque_conv, float* forces, float* torques, float* xfrc,
                         int32_t xfrc_bodies, const int32_t* body_ids, void* stream);

/* One semi-implicit Euler step of free rigid bodies (state [nworld][nbody][13]:
 * position, quaternion (w, x, y, z), velocity, angular velocity in the world
 * frame; params [nbody][4]: mass, principal inertia) under forces / torques
 * [nworld][nbody][3] and gravity. No reference counterpart: it stands in for
 * the MuJoCo-Warp step the reference's coupled environments take, which does
 * not run here. */
int32_t hl_free_bodies_step(int32_t nworld, int32_t nbody, float* state, const float* params, const float* forces,
                            const float* torques, float gx, float gy, float gz, float dt, void* stream);

/* Posed capsules for hl_voxelize: body-frame segments shapes [nbody][7]
 * (a, b, radius in engine units) of every world's bodies -> caps
 * [nworld][nbody][8] in lattice coordinates (the mapping of hl_couple_poses),
 * solid id solid_ids[b]. No reference counterpart (the reference ray-casts
 * the posed mesh, lbm_func_3d.py:254-316). */
int32_t hl_body_capsules(int32_t nworld, int32_t nbody, const float* pos, int32_t pos_stride, const float* quat,
                         int32_t quat_stride, const float* shapes, const int32_t* solid_ids,
                         const float* engine_origin, const float* lattice_origin, const float* scale, float* caps,
                         void* stream);

/* ------------------------------------------------------------------------
 * Articulated chains: planar kinematic trees (a free or fixed root, hinges
 * about z, slides in the plane) stepped on the device in FP64, one workgroup
 * per world. Stands in for the MuJoCo-Warp step of the reference's coupled
 * loop (envs/lbm/lbm_fluid_env.py:441-533: poses out, fluid loads in as
 * xfrc_applied and, accumulated, qfrc_applied; mjw.step) on chains such as
 * envs/lbm/eel/eel_2d.xml (free root, five z hinges, motors with gears, joint
 * damping and armature, RK4 at 0.01). Per world: joint-space inertia from the
 * centre-of-mass Jacobians plus armature, bias forces at zero acceleration
 * with gravity, passive damping and stiffness, motor forces gear *
 * clamp(ctrl), the loads xfrc [nbody][6] (force at the centre of mass,
 * torque) mapped through J^T at every stage plus applied_factor times that
 * mapping at the step's start (the reference's qfrc_applied); MuJoCo's RK4
 * (loads and controls held over the step) or semi-implicit Euler with
 * implicit damping. A free joint is three dofs of its body: slides along
 * world x and y, then a hinge about the body origin; its qpos0 is the body's
 * position and angle. Bodies in tree order, each body's dofs consecutive. */
#define HL_CHAIN_MAX_BODIES 16
#define HL_CHAIN_MAX_DOFS 16
#define HL_CHAIN_MAX_ACTS 16
#define HL_CHAIN_SLIDE 0
#define HL_CHAIN_HINGE 1
#define HL_CHAIN_RK4 0
#define HL_CHAIN_EULER 1

typedef struct hl_chain_model {
    int32_t nbody, ndof, nact, integrator;
    int32_t parent[HL_CHAIN_MAX_BODIES];     /* -1: the world */
    int32_t free_body[HL_CHAIN_MAX_BODIES];  /* 1: the body offset is absorbed in its dofs */
    int32_t dof_body[HL_CHAIN_MAX_DOFS];
    int32_t dof_type[HL_CHAIN_MAX_DOFS];     /* HL_CHAIN_SLIDE / HL_CHAIN_HINGE */
    int32_t act_dof[HL_CHAIN_MAX_ACTS];
    int32_t ctrl_limited[HL_CHAIN_MAX_ACTS];
    double pos[HL_CHAIN_MAX_BODIES][2];      /* body offset in the parent frame */
    double angle[HL_CHAIN_MAX_BODIES];       /* body rotation about z relative to the parent */
    double com[HL_CHAIN_MAX_BODIES][2];      /* centre of mass, body frame */
    double mass[HL_CHAIN_MAX_BODIES];
    double inertia[HL_CHAIN_MAX_BODIES];     /* zz moment about the centre of mass */
    double posz[HL_CHAIN_MAX_BODIES];        /* world z of the body origin */
    double comz[HL_CHAIN_MAX_BODIES];        /* world z of the centre of mass */
    double dof_axis[HL_CHAIN_MAX_DOFS][2];   /* slide axis in the frame before the joint */
    double dof_anchor[HL_CHAIN_MAX_DOFS][2]; /* hinge anchor, body frame */
    double dof_sign[HL_CHAIN_MAX_DOFS];      /* hinge axis +z (1) or -z (-1) */
    double damping[HL_CHAIN_MAX_DOFS], stiffness[HL_CHAIN_MAX_DOFS], springref[HL_CHAIN_MAX_DOFS];
    double armature[HL_CHAIN_MAX_DOFS];
    double gear[HL_CHAIN_MAX_ACTS], ctrl_lo[HL_CHAIN_MAX_ACTS], ctrl_hi[HL_CHAIN_MAX_ACTS];
    double gravity[2];
    double dt;
    double applied_factor; /* 1: the reference's loop (xfrc_applied and its accumulation both act) */
} hl_chain_model;

typedef struct hl_chain hl_chain;

/* Validate a host model and keep a device copy. */
int32_t hl_chain_create(const hl_chain_model* model, hl_chain** out);
int32_t hl_chain_destroy(hl_chain* c);

/* nsteps engine steps of nworld worlds. qpos, qvel: device FP64
 * [nworld][ndof], advanced in place. ctrl: device FP32 [nworld][nact] or
 * NULL (zero); xfrc: device FP32 [nworld][nbody][6] (fx, fy, fz, tx, ty, tz;
 * the planar parts act) or NULL, both held over the call. Outputs after the
 * last step (each may be NULL): qfrc_applied FP64 [nworld][ndof] (the last
 * step's), and the body poses as MuJoCo's data keeps them, FP32: xpos and
 * xipos [nworld][nbody][3] (body origin, centre of mass), xquat
 * [nworld][nbody][4] (w, x, y, z). */
int32_t hl_chain_step(hl_chain* c, int32_t nworld, double* qpos, double* qvel, const float* ctrl, const float* xfrc,
                      int32_t nsteps, double* qfrc_applied, float* xpos, float* xipos, float* xquat, void* stream);

/* ------------------------------------------------------------------------
 * Batches: many resident D2Q9 FP32 lattices, of uneven sizes, advanced
 * together with one launch per kernel kind per step (work-lists of
 * (environment, tile) items). Replaces the nworld dimension of LBM_Solver
 * (envs/lbm/lbm_solver.py:41-141: precompute_transformed_segments,
 * stream_and_collide, apply_bc and Swap_Mom launched over (nworld, nx, ny),
 * every world the same size) and its per-world init_force clears. Each
 * environment's buffer parity and time index live in device memory and are
 * advanced by the step itself, so a captured multi-step graph can be
 * replayed. Results are bitwise equal to stepping each environment alone.
 * ---------------------------------------------------------------------- */
typedef struct hl_batch hl_batch;

/* Build a batch over n descriptors (copied; the buffers they point to stay
 * owned by the caller). Every environment starts with parity 0 (state in
 * moments[0]) and time 0. */
int32_t hl_batch_create(const hl_env* envs, int32_t n, hl_batch** out);
void hl_batch_destroy(hl_batch* b);

/* Bodies of environment env: device mask (0 outside the half-open box);
 * optional device capsule table [nposes][ncap][8] re-voxelized into the box
 * at the start of every step from entry time % nposes; optional device pose
 * table [nposes][nsolid][8] (NULL = the descriptor's static poses). The
 * prescribed-kinematics tables replace advance_prescribed_solid_motion_kernel
 * (envs/lbm/karman/karman_lbm_env.py:65-77) and the pose-to-solid kernels
 * (envs/lbm/lbm_fluid_env_func.py:72-105). Invalidates a captured graph. */
int32_t hl_batch_set_bodies(hl_batch* b, int32_t env, const uint8_t* mask, const int32_t box[6], const float* caps,
                            int32_t ncap, const float* poses, int32_t nposes);

/* Strip heights of the batched step: rows (default 16) away from the body
 * boxes, body_rows (default 4) where a column strip crosses one; both in
 * [3, 64]. No counterpart in the reference (its worlds share one grid): a
 * tuning knob for uneven batches. States do not depend on it; the force
 * history's per-strip partial sums are grouped by it, so it agrees to
 * summation-order rounding. Invalidates a captured graph. */
int32_t hl_batch_set_strips(hl_batch* b, int32_t rows, int32_t body_rows);

/* Host copies of every environment's buffer parity (0/1) and time index. */
int32_t hl_batch_set_state(hl_batch* b, const int32_t* parity, const int32_t* time);
int32_t hl_batch_get_state(hl_batch* b, int32_t* parity, int32_t* time);

/* Advance nsteps. active: device uint8 [n] (NULL = all); inactive
 * environments are untouched. force_history: NULL or device FP64
 * [nsteps][n][16][6] receiving each step's force/torque of the state it reads
 * (as hl_forces). per_env != 0 launches the same kernels one environment at
 * a time (the baseline schedule; same results). */
int32_t hl_batch_step(hl_batch* b, const uint8_t* active, int32_t nsteps, double* force_history, int32_t per_env,
                      void* stream);

/* Capture nsteps batched steps (arguments as hl_batch_step) in a graph owned
 * by the batch, and replay it on a stream. Replaces the Warp graph capture
 * of LBM_Solver.step (lbm_solver.py:107-141). */
int32_t hl_batch_capture(hl_batch* b, const uint8_t* active, int32_t nsteps, double* force_history);
int32_t hl_batch_replay(hl_batch* b, void* stream);

/* The persistent engine: nsteps steps of every active environment in one
 * cooperative launch (the reference's per-world loop of LBM_Solver.step,
 * lbm_solver.py:97-141, with all worlds in one grid and all steps in one
 * launch). The wavefronts deal out every environment's step strips, the
 * strips at a body voxelizing the capsule table at their step's time as they
 * load their rows (environments with at most 32 capsules); a barrier
 * separates the steps. Arguments, results and the force history as
 * hl_batch_step, bit for bit (the same strips, cell arithmetic and summation
 * order). per_env != 0:
 * one launch (of nsteps steps) per environment. The barrier waits are
 * bounded; hl_batch_run_status reports (after a stream synchronize) whether
 * the last launch gave one up. Not capturable in a graph. */
int32_t hl_batch_run(hl_batch* b, const uint8_t* active, int32_t nsteps, double* force_history, int32_t per_env,
                     void* stream);
int32_t hl_batch_run_status(hl_batch* b, int32_t* timed_out);

/* Environments, step items (strips) and step items that reduce forces (any may be NULL). */
int32_t hl_batch_size(const hl_batch* b, int32_t* nenv, int32_t* step_tiles, int32_t* force_tiles);

#ifdef __cplusplus
}
#endif
#endif
