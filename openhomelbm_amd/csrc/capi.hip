// C ABI of libhomelbm (declared in include/homelbm.h): argument checks,
// launch of the step / auxiliary kernels, and the library-owned reduction
// workspace. Errors are reported by return code and hl_last_error().
#pragma clang fp contract(off)

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "homelbm.h"
#include "kernels.h"

namespace {

thread_local std::string g_error;

int32_t fail(int32_t code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int32_t hip_fail(hipError_t e, const char* where)
{
    return fail(HL_E_HIP, "%s: %s", where, hipGetErrorString(e));
}

int nfields(int lattice) { return lattice == HL_D2Q9 ? 6 : 10; }

int32_t validate(const hl_env* e, bool need_moments = true)
{
    if (!e) return fail(HL_E_INVALID, "env is NULL");
    if (e->lattice != HL_D2Q9 && e->lattice != HL_D3Q27) return fail(HL_E_INVALID, "lattice must be 9 or 27, got %d", e->lattice);
    if (e->dtype < HL_F32 || e->dtype > HL_BF16) return fail(HL_E_INVALID, "unknown dtype %d", e->dtype);
    if (e->nx < 3 || e->ny < 3) return fail(HL_E_INVALID, "grid %dx%d: every extent must be >= 3", e->nx, e->ny);
    if (e->lattice == HL_D2Q9 && e->nz != 1) return fail(HL_E_INVALID, "D2Q9 needs nz == 1, got %d", e->nz);
    if (e->lattice == HL_D3Q27 && e->nz < 3) return fail(HL_E_INVALID, "D3Q27 needs nz >= 3, got %d", e->nz);
    const int64_t n = (int64_t)e->nx * e->ny * e->nz;
    if (n > ((int64_t)1 << 40)) return fail(HL_E_INVALID, "grid too large");
    const int64_t nf = nfields(e->lattice);
    if (e->rpitch == 0 || e->rpitch == e->nx) {
        if (e->stride < n) return fail(HL_E_INVALID, "stride %lld < cell count %lld", (long long)e->stride, (long long)n);
    } else if (e->rpitch == nf * e->nx) {
        if (e->lattice != HL_D3Q27) return fail(HL_E_UNSUPPORTED, "the row-interleaved layout is for D3Q27 lattices");
        if (e->stride != e->nx)
            return fail(HL_E_INVALID, "row-interleaved layout needs stride == nx, got %lld", (long long)e->stride);
    } else {
        return fail(HL_E_INVALID, "rpitch %lld is neither nx (%d) nor %lld * nx", (long long)e->rpitch, e->nx,
                    (long long)nf);
    }
    if (e->force) {
        const int64_t fs = e->force_stride ? e->force_stride : e->stride;
        if (fs < n) return fail(HL_E_INVALID, "force stride %lld < cell count %lld", (long long)fs, (long long)n);
    }
    if (need_moments && (!e->moments[0] || !e->moments[1])) return fail(HL_E_INVALID, "moment buffers are NULL");
    if (e->nsolid < 0 || e->nsolid > HL_MAX_SOLIDS) return fail(HL_E_INVALID, "nsolid %d outside [0, %d]", e->nsolid, HL_MAX_SOLIDS);
    if (e->nsolid > 0 && !e->solids) return fail(HL_E_INVALID, "nsolid > 0 but the pose table is NULL");
    if (!(e->nu > 0.0f) || !std::isfinite(e->nu)) return fail(HL_E_INVALID, "viscosity must be > 0, got %g", (double)e->nu);
    const int nfaces = e->lattice == HL_D2Q9 ? 4 : 6;
    for (int f = 0; f < nfaces; ++f)
        if (e->bc_type[f] < 0 || e->bc_type[f] > 2) return fail(HL_E_INVALID, "bc_type[%d] = %d not in {0, 1, 2}", f, e->bc_type[f]);
    if (e->mask) {
        const int32_t* b = e->mask_box;
        if (b[0] < 0 || b[1] > e->nx || b[2] < 0 || b[3] > e->ny || b[4] < 0 || b[5] > e->nz || b[0] > b[1] ||
            b[2] > b[3] || b[4] > b[5])
            return fail(HL_E_INVALID, "mask_box (%d,%d,%d,%d,%d,%d) outside the grid", b[0], b[1], b[2], b[3], b[4], b[5]);
    }
    if (e->dtype == HL_F16 || e->dtype == HL_BF16)
        for (int f = 0; f < nfields(e->lattice); ++f) {
            if (!(e->scale[f] > 0.0f) || !std::isfinite(e->scale[f]))
                return fail(HL_E_INVALID, "16-bit storage needs scale[%d] > 0", f);
            if (!std::isfinite(e->offset[f])) return fail(HL_E_INVALID, "16-bit storage needs a finite offset[%d]", f);
        }
    return HL_OK;
}

hl::EnvArgs make_args(const hl_env* e, int cur)
{
    hl::EnvArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nx = e->nx; a.ny = e->ny; a.nz = e->nz; a.nsolid = e->nsolid;
    a.stride = e->stride;
    a.rp = e->rpitch ? e->rpitch : e->nx;
    a.fs = e->force_stride ? e->force_stride : e->stride;
    a.src = e->moments[cur & 1];
    a.dst = e->moments[(cur & 1) ^ 1];
    a.mask = e->mask;
    for (int k = 0; k < 6; ++k) a.box[k] = e->mask_box[k];
    if (!e->mask) { a.box[0] = a.box[2] = a.box[4] = 1 << 29; a.box[1] = a.box[3] = a.box[5] = -(1 << 29); }
    a.force = e->force;
    a.solids = e->solids;
    a.nu = e->nu;
    for (int k = 0; k < 6; ++k) {
        a.bc_type[k] = e->bc_type[k];
        for (int c = 0; c < 3; ++c) a.bc_value[k][c] = e->bc_value[k][c];
    }
    for (int f = 0; f < 10; ++f) {
        const bool sixteen = e->dtype == HL_F16 || e->dtype == HL_BF16;
        a.scale[f] = sixteen ? e->scale[f] : 1.0f;
        a.inv_scale[f] = sixteen ? 1.0f / e->scale[f] : 1.0f;
        a.offset[f] = sixteen ? (f == 0 ? 1.0f : 0.0f) + e->offset[f] : 0.0f;
    }
    a.region[0] = 0; a.region[1] = e->nx; a.region[2] = 0; a.region[3] = e->ny; a.region[4] = 0; a.region[5] = e->nz;
    return a;
}

// Library-owned scratch per (device, caller stream): a grow-only reduction
// workspace (cut-cell list, force partials).
// Calls on different caller streams never share scratch, so their work may
// overlap; calls on one stream are ordered by that stream (concurrent calls
// from several host threads on one stream are not supported). Created on first use, kept until the process ends. Call a step or
// reduction once outside any stream capture so the allocation does not happen
// while capturing.
// The near-body list a workspace holds and what it was built for (hl_step
// with a nonzero mask_epoch); any other use of the workspace clears it.
struct ListCache {
    bool valid = false;
    uint64_t epoch = 0;
    const void* mask = nullptr;
    int32_t key[10] = {0};
};
// Near-body lists of the worlds of hl_step_worlds in the same workspace, one
// per world at its offset; valid while no other call used the workspace.
struct WorldsCache {
    bool valid = false;
    void* ws = nullptr;
    std::vector<size_t> offsets;
    std::vector<ListCache> lists;
};
struct StreamCtx {
    std::mutex mutex;  // held for the whole call that uses this context
    void* ws = nullptr;
    size_t bytes = 0;
    ListCache list;
    WorldsCache worlds;
};
std::mutex g_ctx_mutex;
std::map<std::pair<int, hipStream_t>, StreamCtx> g_ctx;

int32_t stream_ctx(hipStream_t s, StreamCtx** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    *out = &g_ctx[{dev, s}];  // std::map nodes are stable
    return HL_OK;
}

// Grows c's workspace to at least `bytes`. Work already enqueued on the
// caller's stream may still read the old buffer, so it is drained before the
// buffer is freed.
int32_t workspace(StreamCtx* c, hipStream_t s, size_t bytes, void** out)
{
    hipError_t e;
    if (c->bytes < bytes) {
        if (c->ws) {
            if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
            if ((e = hipFree(c->ws)) != hipSuccess) return hip_fail(e, "hipFree(workspace)");
            c->ws = nullptr;
            c->bytes = 0;
        }
        c->list.valid = false;
        c->worlds.valid = false;
        e = hipMalloc(&c->ws, bytes);
        if (e != hipSuccess) return fail(HL_E_NOMEM, "hipMalloc(%zu) for the reduction workspace: %s", bytes, hipGetErrorString(e));
        c->bytes = bytes;
    }
    *out = c->ws;
    c->worlds.valid = false;  // the caller may overwrite the worlds' lists
    return HL_OK;
}

// Step-kernel timer (hl_kernel_timer): event pairs recorded around the step
// kernel launches of hl_step, read back by hl_kernel_timer_read.
struct KernelTimer {
    std::mutex mutex;
    std::vector<hipEvent_t> start, end;
    size_t used = 0;
    int32_t cap = 0;
    int32_t every = 1;   // record one launch in every `every`
    int64_t seen = 0;    // launches since the timer was armed
};
KernelTimer g_timer;

// Records one launch's start event on s; returns the slot or -1.
int timer_begin(hipStream_t s)
{
    std::lock_guard<std::mutex> lock(g_timer.mutex);
    if (g_timer.used >= (size_t)g_timer.cap) return -1;
    if (g_timer.seen++ % g_timer.every != 0) return -1;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return -1;
    const int k = (int)g_timer.used;
    if (hipEventRecord(g_timer.start[k], s) != hipSuccess) return -1;
    return k;
}

void timer_end(int k, hipStream_t s)
{
    if (k < 0) return;
    std::lock_guard<std::mutex> lock(g_timer.mutex);
    if (hipEventRecord(g_timer.end[k], s) == hipSuccess) g_timer.used = (size_t)k + 1;
}

// Uniform-flow moments in working precision T, in the operation order of
// set_uniform_flow_kernel (karman_lbm_env.py:15-53) and
// set_uniform_flow_3d_kernel (karman_lbm_env_3d.py:15-64).
template <typename T> void uniform_moments(int lattice, T ux, T uy, T uz, double m[10])
{
    T rho = 1;
    for (int f = 0; f < 10; ++f) m[f] = 0.0;
    if (lattice == HL_D2Q9) {
        T pop[9];
        T U2 = ux * ux + uy * uy;
        for (int i = 0; i < 9; ++i) {
            T cu = T(hl::D2CX[i]) * ux + T(hl::D2CY[i]) * uy;
            pop[i] = rho * hl::d2w<T>(i) * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * U2);
        }
        T inv = T(1) / rho;
        T pixx = pop[1] + pop[2] + pop[5] + pop[6] + pop[7] + pop[8];
        T piyy = pop[3] + pop[4] + pop[5] + pop[6] + pop[7] + pop[8];
        T pixy = pop[5] + pop[7] - pop[6] - pop[8];
        T cs2_local = pixx;
        m[0] = rho; m[1] = ux; m[2] = uy;
        m[3] = pixx * inv - cs2_local; m[4] = piyy * inv - cs2_local; m[5] = pixy * inv;
    } else {
        T p[27];
        T U2 = ux * ux + uy * uy + uz * uz;
        for (int i = 0; i < 27; ++i) {
            T cu = T(hl::D3CX[i]) * ux + T(hl::D3CY[i]) * uy + T(hl::D3CZ[i]) * uz;
            p[i] = hl::d3w<T>(i) * rho * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * U2);
        }
        T inv = T(1) / rho;
        T pixx = p[1] + p[2] + p[7] + p[8] + p[9] + p[10] + p[13] + p[14] + p[15] + p[16] + p[19] + p[20]
               + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
        T pixy = (p[7] + p[8] + p[19] + p[20] + p[21] + p[22]) - (p[13] + p[14] + p[23] + p[24] + p[25] + p[26]);
        T pixz = (p[9] + p[10] + p[19] + p[20] + p[23] + p[24]) - (p[15] + p[16] + p[21] + p[22] + p[25] + p[26]);
        T piyy = p[3] + p[4] + p[7] + p[8] + p[11] + p[12] + p[13] + p[14] + p[17] + p[18] + p[19] + p[20]
               + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
        T piyz = (p[11] + p[12] + p[19] + p[20] + p[25] + p[26]) - (p[17] + p[18] + p[21] + p[22] + p[23] + p[24]);
        T pizz = p[5] + p[6] + p[9] + p[10] + p[11] + p[12] + p[15] + p[16] + p[17] + p[18] + p[19] + p[20]
               + p[21] + p[22] + p[23] + p[24] + p[25] + p[26];
        T cs2_local = pixx;
        m[0] = rho; m[1] = ux; m[2] = uy; m[3] = uz;
        m[4] = pixx * inv - cs2_local; m[5] = piyy * inv - cs2_local; m[6] = pizz * inv - cs2_local;
        m[7] = pixy * inv; m[8] = pixz * inv; m[9] = piyz * inv;
    }
}

}  // namespace

namespace {

// Capsule bodies whose surface crossings cut links (hl_step_capsules).
struct CapsuleBodies {
    const float* caps;
    int32_t ncap;
    int32_t box[6];
};

// The step arguments of env, with capsule bodies in place of the mask.
hl::EnvArgs body_args(const hl_env* env, int cur, const CapsuleBodies* cb)
{
    hl::EnvArgs a = make_args(env, cur);
    if (cb) {
        a.mask = nullptr;
        a.caps = cb->caps;
        a.ncap = cb->ncap;
        for (int k = 0; k < 6; ++k) a.box[k] = cb->box[k];
    }
    return a;
}

// hl_step after its argument checks, with the reduction workspace from
// get_ws(bytes, &ptr) and the near-body list cache that goes with it. With
// cb, the capsules' crossings cut links instead of the mask.
template <typename GetWs>
int32_t step_core(const hl_env* env, int32_t cur, int32_t nsteps, int32_t kernel, double* force_history, hipStream_t s,
                  ListCache& cache, GetWs get_ws, const CapsuleBodies* cb = nullptr)
{
    int32_t rc = HL_OK;
    // the near-body pass needs the cut-cell list when a mask is present and
    // either the tiled kernel runs or forces are requested; capsule bodies
    // always need it (no other kernel applies their wall rule)
    const bool tiled = kernel != HL_KERNEL_GATHER && hl::tiling_supported(env->lattice, 32, 8, 1) &&
                       env->dtype != HL_F64;
    const bool want_forces = force_history && env->nsolid > 0;
    void* ws = nullptr;
    if (cb && nsteps > 0) {
        hl::EnvArgs a0 = body_args(env, cur, cb);
        rc = get_ws(hl::forces_workspace_bytes(a0, env->lattice), &ws);
        if (rc) return rc;
        cache.valid = false;  // the capsules live in device memory: their list is rebuilt per call
        hipError_t e = hl::cut_list(env->lattice, a0, ws, s);
        if (e != hipSuccess) return hip_fail(e, "cut-list kernel launch");
    } else if (env->mask && nsteps > 0 && (tiled || want_forces)) {
        hl::EnvArgs a0 = make_args(env, cur);
        rc = get_ws(hl::forces_workspace_bytes(a0, env->lattice), &ws);
        if (rc) return rc;
        // the list depends on the mask, its box and the lattice; a caller that
        // versions its mask (mask_epoch) gets it rebuilt only when that changes.
        // Under stream capture it is always recorded, so graphs stay self-contained.
        const int32_t key[10] = {env->lattice, env->nx, env->ny, env->nz, env->mask_box[0], env->mask_box[1],
                                 env->mask_box[2], env->mask_box[3], env->mask_box[4], env->mask_box[5]};
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cap) != hipSuccess) cap = hipStreamCaptureStatusActive;
        const bool reuse = env->mask_epoch != 0 && cap == hipStreamCaptureStatusNone && cache.valid &&
                           cache.epoch == env->mask_epoch && cache.mask == env->mask &&
                           std::memcmp(cache.key, key, sizeof(key)) == 0;
        if (!reuse) {
            hipError_t e = hl::cut_list(env->lattice, a0, ws, s);
            if (e != hipSuccess) return hip_fail(e, "cut-list kernel launch");
            cache.valid = env->mask_epoch != 0 && cap == hipStreamCaptureStatusNone;
            cache.epoch = env->mask_epoch;
            cache.mask = env->mask;
            std::memcpy(cache.key, key, sizeof(key));
        }
    }
    // force rows a tiled launch left for the next launch to sum
    struct Pending {
        const double* rows = nullptr;
        int nrows = 0, width = 0;
        double* out = nullptr;
    } pend;
    auto flush = [&]() -> hipError_t {
        hipError_t e = hipSuccess;
        if (pend.rows) e = hl::sum_partials(pend.rows, pend.nrows, pend.width, pend.out, s);
        pend = Pending();
        return e;
    };
    for (int k = 0; k < nsteps; ++k) {
        hl::EnvArgs a = body_args(env, cur, cb);
        double* fout = want_forces ? force_history + (int64_t)k * env->nsolid * 6 : nullptr;
        hipError_t e = hipErrorNotSupported;
        bool fixed = false;
        if (kernel == HL_KERNEL_AUTO || kernel == HL_KERNEL_TILED) {
            // with a body mask, the same launch also updates the near-body
            // cells (wall rule, reference order), writes this step's force
            // rows and sums the previous step's
            hl::FixArgs fx;
            std::memset(&fx, 0, sizeof(fx));
            hl::FixArgs* pfx = nullptr;
            if (ws) {
                hl::cut_bits_view(env->lattice, a, ws);
                fx = hl::fixup_args(env->lattice, a, ws, fout != nullptr, k);
                if (fx.list) {
                    pfx = &fx;
                    fx.prev = pend.rows;
                    fx.prev_rows = pend.nrows;
                    fx.prev_out = pend.out;
                }
            }
            const int tk = timer_begin(s);
            e = hl::step_tiled(env->lattice, env->dtype, a, pfx, s);
            timer_end(e == hipSuccess ? tk : -1, s);
            if (e == hipErrorNotSupported && kernel == HL_KERNEL_TILED)
                return fail(HL_E_UNSUPPORTED, "no tiled kernel for lattice %d dtype %d", env->lattice, env->dtype);
            if (e == hipSuccess) {
                fixed = pfx != nullptr || fout == nullptr;
                if (pfx) pend = Pending();  // summed by this launch
                // the launch's near-body workgroups left one force row each
                if (pfx && fout) {
                    pend.rows = fx.partial;
                    pend.nrows = fx.nfix;
                    pend.width = fx.width;
                    pend.out = fout;
                }
            } else if (e == hipErrorNotSupported) {
                a.cut_bits = nullptr;  // the gather kernel does every cell
                a.cut_cols = nullptr;
            }
        }
        if (e == hipErrorNotSupported && env->lattice == HL_D2Q9 && fout && ws) {
            // one launch: the gather step, this step's force rows and the sum
            // of the previous step's rows
            hl::FixArgs fx = hl::fixup_args(env->lattice, a, ws, true, k);
            if (fx.list && fx.partial) {
                fx.prev = pend.rows;
                fx.prev_rows = pend.nrows;
                fx.prev_out = pend.out;
                const int tk = timer_begin(s);
                e = hl::step_gather_forces(env->dtype, a, fx, s);
                timer_end(e == hipSuccess ? tk : -1, s);
                if (e == hipSuccess) {
                    pend.rows = fx.partial;
                    pend.nrows = fx.nfix;
                    pend.width = fx.width;
                    pend.out = fout;
                    fixed = true;
                }
            }
        }
        if (e == hipErrorNotSupported) {
            e = flush();
            if (e == hipSuccess) {
                // capsule bodies: the gather kernel streams every cell and the
                // near-body pass below rewrites the listed ones
                hl::EnvArgs ag = a;
                ag.caps = nullptr;
                const int tk = timer_begin(s);
                e = hl::step_gather(env->lattice, env->dtype, ag, s);
                timer_end(e == hipSuccess ? tk : -1, s);
            }
        }
        if (e == hipSuccess && !fixed && (fout || cb))
            e = hl::fixup(env->lattice, env->dtype, a, cb != nullptr, fout, ws, s);
        if (e != hipSuccess) return hip_fail(e, "step kernel launch");
        cur ^= 1;
    }
    hipError_t e = flush();
    if (e != hipSuccess) return hip_fail(e, "force sum launch");
    return HL_OK;
}

}  // namespace

namespace {

// Scratch of the allocation-free entry point: one array per GPU (every GPU
// that runs the library holds it until the process ends). In host code the
// name is only a stand-in for the symbol: its device address comes from
// hipGetSymbolAddress, once per device with that device current.
constexpr size_t kNoAllocScratchBytes = (size_t)8 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
struct NoAllocDevice {
    std::mutex mutex;
    void* ptr = nullptr;  // device address of this GPU's scratch
    ListCache list;
};
NoAllocDevice g_noalloc[kMaxDevices];

}  // namespace

namespace hl {
int32_t report(int32_t code, const char* msg) { return fail(code, "%s", msg); }
int32_t check_env(const hl_env* e) { return validate(e); }
EnvArgs env_args(const hl_env* e, int cur) { return make_args(e, cur); }
}  // namespace hl

extern "C" {

int32_t hl_version(void) { return 100; }

int32_t hl_env_bytes(void) { return (int32_t)sizeof(hl_env); }

const char* hl_last_error(void) { return g_error.c_str(); }

int32_t hl_step(const hl_env* env, int32_t cur, int32_t nsteps, int32_t kernel, double* force_history, void* stream)
{
    int32_t rc = validate(env);
    if (rc) return rc;
    if (cur != 0 && cur != 1) return fail(HL_E_INVALID, "cur must be 0 or 1");
    if (nsteps < 0) return fail(HL_E_INVALID, "nsteps must be >= 0");
    if (kernel < HL_KERNEL_AUTO || kernel > HL_KERNEL_TILED) return fail(HL_E_INVALID, "unknown kernel %d", kernel);
    hipStream_t s = static_cast<hipStream_t>(stream);
    StreamCtx* ctx = nullptr;
    rc = stream_ctx(s, &ctx);
    if (rc) return rc;
    std::lock_guard<std::mutex> ctx_lock(ctx->mutex);
    return step_core(env, cur, nsteps, kernel, force_history, s, ctx->list,
                     [&](size_t bytes, void** ws) { return workspace(ctx, s, bytes, ws); });
}

int32_t hl_step_noalloc(const hl_env* env, int32_t cur, int32_t nsteps, int32_t kernel, double* force_history,
                        void* stream)
{
    int32_t rc = validate(env);
    if (rc) return rc;
    if (cur != 0 && cur != 1) return fail(HL_E_INVALID, "cur must be 0 or 1");
    if (nsteps < 0) return fail(HL_E_INVALID, "nsteps must be >= 0");
    if (kernel < HL_KERNEL_AUTO || kernel > HL_KERNEL_TILED) return fail(HL_E_INVALID, "unknown kernel %d", kernel);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(HL_E_UNSUPPORTED, "device %d beyond the %d the library tracks", dev, kMaxDevices);
    NoAllocDevice& d = g_noalloc[dev];
    std::lock_guard<std::mutex> lock(d.mutex);
    if (!d.ptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        d.ptr = p;
    }
    return step_core(env, cur, nsteps, kernel, force_history, static_cast<hipStream_t>(stream), d.list,
                     [&](size_t bytes, void** ws) -> int32_t {
                         if (bytes > kNoAllocScratchBytes)
                             return fail(HL_E_NOMEM,
                                         "hl_step_noalloc: the near-body work of this lattice needs %zu bytes of "
                                         "scratch, more than the %zu bytes the allocation-free entry point holds "
                                         "(a smaller mask box, or hl_step)",
                                         bytes, kNoAllocScratchBytes);
                         *ws = d.ptr;
                         return HL_OK;
                     });
}

int32_t hl_step_worlds(const hl_env* envs, int32_t nworld, int32_t cur, int32_t nsteps, int32_t kernel,
                       double* const* force_history, void* stream)
{
    if (!envs) return fail(HL_E_INVALID, "envs is NULL");
    if (nworld < 1 || nworld > (1 << 16)) return fail(HL_E_INVALID, "nworld must be in [1, 65536], got %d", nworld);
    if (cur != 0 && cur != 1) return fail(HL_E_INVALID, "cur must be 0 or 1");
    if (nsteps < 0) return fail(HL_E_INVALID, "nsteps must be >= 0");
    if (kernel < HL_KERNEL_AUTO || kernel > HL_KERNEL_TILED) return fail(HL_E_INVALID, "unknown kernel %d", kernel);
    for (int32_t w = 0; w < nworld; ++w) {
        const int32_t rc = validate(envs + w);
        if (rc) {
            const std::string m = g_error;
            return fail(rc, "world %d: %s", w, m.c_str());
        }
        if (envs[w].dtype != envs[0].dtype) return fail(HL_E_INVALID, "world %d: every world needs the same dtype", w);
        if (envs[w].lattice != envs[0].lattice)
            return fail(HL_E_INVALID, "world %d: every world needs the same lattice", w);
    }
    if (nsteps == 0) return HL_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    StreamCtx* ctx = nullptr;
    int32_t rc = stream_ctx(s, &ctx);
    if (rc) return rc;
    std::lock_guard<std::mutex> ctx_lock(ctx->mutex);
    const int dtype = envs[0].dtype;
    auto history = [&](int w) { return force_history && envs[w].nsolid > 0 ? force_history[w] : nullptr; };
    // D3Q27 worlds the tiled kernel fits share launches; anything else steps
    // world by world exactly as hl_step does
    bool batched = nworld > 1 && kernel != HL_KERNEL_GATHER && envs[0].lattice == HL_D3Q27;
    for (int32_t w = 0; batched && w < nworld; ++w) batched = hl::tiled_fits(make_args(envs + w, cur), dtype);
    if (!batched) {
        for (int32_t w = 0; w < nworld; ++w) {
            rc = step_core(envs + w, cur, nsteps, kernel, history(w), s, ctx->list,
                           [&](size_t bytes, void** ws) { return workspace(ctx, s, bytes, ws); });
            if (rc) {
                const std::string m = g_error;
                return fail(rc, "world %d: %s", w, m.c_str());
            }
        }
        return HL_OK;
    }
    // one workspace, each world's near-body area at its own 256-byte aligned offset
    std::vector<size_t> off((size_t)nworld);
    size_t total = 0;
    for (int32_t w = 0; w < nworld; ++w) {
        off[(size_t)w] = total;
        if (envs[w].mask) total += (hl::forces_workspace_bytes(make_args(envs + w, cur), HL_D3Q27) + 255) / 256 * 256;
    }
    WorldsCache& wc = ctx->worlds;
    const bool kept = wc.valid;  // no other call used the workspace since the last worlds call
    void* ws = nullptr;
    rc = workspace(ctx, s, total > 256 ? total : 256, &ws);
    if (rc) return rc;
    ctx->list.valid = false;  // the single-lattice list shares this workspace
    if (!kept || wc.ws != ws || wc.offsets != off) {
        wc.ws = ws;
        wc.offsets = off;
        wc.lists.assign((size_t)nworld, ListCache());
    }
    wc.valid = true;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) != hipSuccess) cap = hipStreamCaptureStatusActive;
    auto wsp = [&](int w) { return static_cast<void*>(static_cast<char*>(ws) + off[(size_t)w]); };
    for (int32_t w = 0; w < nworld; ++w) {
        const hl_env* env = envs + w;
        if (!env->mask) continue;
        // the list depends on the mask, its box and the lattice (see step_core)
        const int32_t key[10] = {env->lattice, env->nx, env->ny, env->nz, env->mask_box[0], env->mask_box[1],
                                 env->mask_box[2], env->mask_box[3], env->mask_box[4], env->mask_box[5]};
        ListCache& lc = wc.lists[(size_t)w];
        const bool reuse = env->mask_epoch != 0 && cap == hipStreamCaptureStatusNone && lc.valid &&
                           lc.epoch == env->mask_epoch && lc.mask == env->mask &&
                           std::memcmp(lc.key, key, sizeof(key)) == 0;
        if (reuse) continue;
        hipError_t e = hl::cut_list(HL_D3Q27, make_args(env, cur), wsp(w), s);
        if (e != hipSuccess) return hip_fail(e, "cut-list kernel launch");
        lc.valid = env->mask_epoch != 0 && cap == hipStreamCaptureStatusNone;
        lc.epoch = env->mask_epoch;
        lc.mask = env->mask;
        std::memcpy(lc.key, key, sizeof(key));
    }
    // force rows a launch left for the next launch (or the closing sums)
    struct Pending {
        const double* rows = nullptr;
        int nrows = 0, width = 0;
        double* out = nullptr;
    };
    std::vector<Pending> pend((size_t)nworld);
    for (int k = 0; k < nsteps; ++k) {
        const int c = cur ^ (k & 1);
        // chunks of consecutive worlds that agree on the body-force field
        for (int32_t w0 = 0; w0 < nworld;) {
            int n = 1;
            while (n < hl::kTiledWorlds && w0 + n < nworld && (envs[w0 + n].force != nullptr) == (envs[w0].force != nullptr))
                ++n;
            hl::EnvArgs a[hl::kTiledWorlds];
            hl::FixArgs fx[hl::kTiledWorlds];
            hl::FixArgs* fxp[hl::kTiledWorlds];
            for (int i = 0; i < n; ++i) {
                const int w = w0 + i;
                double* fout = history(w) ? history(w) + (int64_t)k * envs[w].nsolid * 6 : nullptr;
                a[i] = make_args(envs + w, c);
                std::memset(&fx[i], 0, sizeof(fx[i]));
                fxp[i] = nullptr;
                if (envs[w].mask) {
                    hl::cut_bits_view(HL_D3Q27, a[i], wsp(w));
                    fx[i] = hl::fixup_args(HL_D3Q27, a[i], wsp(w), fout != nullptr, k);
                    if (fx[i].list) {
                        fx[i].prev = pend[(size_t)w].rows;
                        fx[i].prev_rows = pend[(size_t)w].nrows;
                        fx[i].prev_out = pend[(size_t)w].out;
                        fxp[i] = &fx[i];
                    }
                }
            }
            hipError_t e = hl::step_tiled_worlds(dtype, a, fxp, n, s);
            if (e != hipSuccess) return hip_fail(e, "worlds step launch");
            for (int i = 0; i < n; ++i) {
                const int w = w0 + i;
                double* fout = history(w) ? history(w) + (int64_t)k * envs[w].nsolid * 6 : nullptr;
                if (fxp[i]) {
                    pend[(size_t)w] = Pending();  // summed by this launch
                    if (fout) pend[(size_t)w] = Pending{fx[i].partial, fx[i].nfix, fx[i].width, fout};
                } else if (fout) {
                    // no near-body cells: the forces are zero
                    e = hl::fixup(HL_D3Q27, dtype, a[i], false, fout, wsp(w), s);
                    if (e != hipSuccess) return hip_fail(e, "force launch");
                }
            }
            w0 += n;
        }
    }
    for (const Pending& p : pend) {
        if (!p.rows) continue;
        hipError_t e = hl::sum_partials(p.rows, p.nrows, p.width, p.out, s);
        if (e != hipSuccess) return hip_fail(e, "force sum launch");
    }
    return HL_OK;
}

int32_t hl_step_capsules(const hl_env* env, const float* caps, int32_t ncap, const int32_t box[6], int32_t cur,
                         int32_t nsteps, int32_t kernel, double* force_history, void* stream)
{
    int32_t rc = validate(env);
    if (rc) return rc;
    if (env->lattice != HL_D3Q27) return fail(HL_E_UNSUPPORTED, "capsule cut links are D3Q27 only");
    if (cur != 0 && cur != 1) return fail(HL_E_INVALID, "cur must be 0 or 1");
    if (nsteps < 0) return fail(HL_E_INVALID, "nsteps must be >= 0");
    if (kernel < HL_KERNEL_AUTO || kernel > HL_KERNEL_TILED) return fail(HL_E_INVALID, "unknown kernel %d", kernel);
    if (ncap < 0 || ncap > HL_MAX_CAPSULES) return fail(HL_E_INVALID, "ncap must be in [0, %d], got %d", HL_MAX_CAPSULES, ncap);
    if (ncap > 0 && !caps) return fail(HL_E_INVALID, "caps is NULL");
    if (!box) return fail(HL_E_INVALID, "box is NULL");
    const int32_t n[3] = {env->nx, env->ny, env->nz};
    for (int k = 0; k < 3; ++k)
        if (box[2 * k] < 0 || box[2 * k] > box[2 * k + 1] || box[2 * k + 1] > n[k])
            return fail(HL_E_INVALID, "box axis %d [%d, %d) is not inside [0, %d]", k, box[2 * k], box[2 * k + 1], n[k]);
    CapsuleBodies cb;
    cb.caps = caps;
    cb.ncap = ncap;
    for (int k = 0; k < 6; ++k) cb.box[k] = box[k];
    hipStream_t s = static_cast<hipStream_t>(stream);
    StreamCtx* ctx = nullptr;
    rc = stream_ctx(s, &ctx);
    if (rc) return rc;
    std::lock_guard<std::mutex> ctx_lock(ctx->mutex);
    return step_core(
        env, cur, nsteps, kernel, force_history, s, ctx->list,
        [&](size_t bytes, void** ws) { return workspace(ctx, s, bytes, ws); }, &cb);
}

int32_t hl_step_polygons(const hl_env* env, int32_t cur, int32_t nsteps, const float* verts,
                         const float* verts_last, const int32_t* counts, int32_t maxv, const float* radius2,
                         double* force_history, void* stream)
{
    int32_t rc = validate(env);
    if (rc) return rc;
    if (env->lattice != HL_D2Q9) return fail(HL_E_UNSUPPORTED, "polygon cut cells are D2Q9 only");
    if (cur != 0 && cur != 1) return fail(HL_E_INVALID, "cur must be 0 or 1");
    if (nsteps < 0) return fail(HL_E_INVALID, "nsteps must be >= 0");
    if (env->nsolid < 1) return fail(HL_E_INVALID, "polygon cut cells need nsolid >= 1");
    if (!verts || !counts || !radius2) return fail(HL_E_INVALID, "verts, counts and radius2 must be non-NULL");
    if (maxv < 1 || maxv > (1 << 20)) return fail(HL_E_INVALID, "maxv must be in [1, 2^20], got %d", maxv);
    if (nsteps == 0) return HL_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    StreamCtx* ctx = nullptr;
    rc = stream_ctx(s, &ctx);
    if (rc) return rc;
    std::lock_guard<std::mutex> ctx_lock(ctx->mutex);
    void* ws = nullptr;
    ctx->list.valid = false;  // ws is reused below
    rc = workspace(ctx, s, hl::polygon_workspace_bytes(), &ws);
    if (rc) return rc;
    hl::EnvArgs a = make_args(env, cur);
    hipError_t e = hl::step_polygons(env->dtype, a, nsteps, verts, verts_last ? verts_last : verts, counts, maxv,
                                     radius2, force_history && env->nsolid > 0 ? force_history : nullptr, ws, s);
    if (e != hipSuccess) return hip_fail(e, "polygon step launch");
    return HL_OK;
}

int32_t hl_step_polygons_worlds(const hl_env* envs, int32_t nworld, int32_t cur, int32_t nsteps,
                                const float* const* verts, const float* const* verts_last,
                                const int32_t* const* counts, const int32_t* maxv, const float* const* radius2,
                                double* const* force_history, void* stream)
{
    if (!envs) return fail(HL_E_INVALID, "envs is NULL");
    if (nworld < 1 || nworld > (1 << 16)) return fail(HL_E_INVALID, "nworld must be in [1, 65536], got %d", nworld);
    if (cur != 0 && cur != 1) return fail(HL_E_INVALID, "cur must be 0 or 1");
    if (nsteps < 0) return fail(HL_E_INVALID, "nsteps must be >= 0");
    if (!verts || !counts || !maxv || !radius2)
        return fail(HL_E_INVALID, "verts, counts, maxv and radius2 must be non-NULL");
    std::vector<hl::EnvArgs> args((size_t)nworld);
    std::vector<const float*> last((size_t)nworld);
    std::vector<double*> fh((size_t)nworld);
    int64_t total_width = 0;
    for (int32_t w = 0; w < nworld; ++w) {
        const hl_env* env = envs + w;
        int32_t rc = validate(env);
        if (rc) return rc;
        if (env->lattice != HL_D2Q9) return fail(HL_E_UNSUPPORTED, "world %d: polygon cut cells are D2Q9 only", w);
        if (env->dtype != envs[0].dtype) return fail(HL_E_INVALID, "world %d: every world needs the same dtype", w);
        if (env->nsolid < 1) return fail(HL_E_INVALID, "world %d: polygon cut cells need nsolid >= 1", w);
        if (!verts[w] || !counts[w] || !radius2[w])
            return fail(HL_E_INVALID, "world %d: verts, counts and radius2 must be non-NULL", w);
        if (maxv[w] < 1 || maxv[w] > (1 << 20))
            return fail(HL_E_INVALID, "world %d: maxv must be in [1, 2^20], got %d", w, maxv[w]);
        args[(size_t)w] = make_args(env, cur);
        last[(size_t)w] = verts_last && verts_last[w] ? verts_last[w] : verts[w];
        fh[(size_t)w] = force_history ? force_history[w] : nullptr;
        if (fh[(size_t)w]) total_width += env->nsolid * 6;
    }
    if (nsteps == 0) return HL_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    StreamCtx* ctx = nullptr;
    int32_t rc = stream_ctx(s, &ctx);
    if (rc) return rc;
    std::lock_guard<std::mutex> ctx_lock(ctx->mutex);
    void* ws = nullptr;
    ctx->list.valid = false;  // ws is reused below
    rc = workspace(ctx, s, hl::polygon_worlds_workspace_bytes(total_width), &ws);
    if (rc) return rc;
    hipError_t e = hl::step_polygons_worlds(envs[0].dtype, args.data(), nworld, nsteps, verts, last.data(), counts,
                                            maxv, radius2, fh.data(), ws, s);
    if (e != hipSuccess) return hip_fail(e, "polygon worlds step launch");
    return HL_OK;
}

int32_t hl_forces(const hl_env* env, int32_t cur, double* out, void* stream)
{
    int32_t rc = validate(env);
    if (rc) return rc;
    if (!out) return fail(HL_E_INVALID, "out is NULL");
    if (cur != 0 && cur != 1) return fail(HL_E_INVALID, "cur must be 0 or 1");
    hl::EnvArgs a = make_args(env, cur);
    hipStream_t s = static_cast<hipStream_t>(stream);
    StreamCtx* ctx = nullptr;
    rc = stream_ctx(s, &ctx);
    if (rc) return rc;
    std::lock_guard<std::mutex> ctx_lock(ctx->mutex);
    void* ws = nullptr;
    ctx->list.valid = false;  // ws is reused below
    rc = workspace(ctx, s, hl::forces_workspace_bytes(a, env->lattice), &ws);
    if (rc) return rc;
    hipError_t e = hl::forces(env->lattice, env->dtype, a, out, ws, s);
    if (e != hipSuccess) return hip_fail(e, "forces kernel launch");
    return HL_OK;
}

int32_t hl_init_uniform(const hl_env* env, int32_t buf, float ux, float uy, float uz, void* stream)
{
    int32_t rc = validate(env);
    if (rc) return rc;
    if (buf != 0 && buf != 1) return fail(HL_E_INVALID, "buf must be 0 or 1");
    double m[10];
    if (env->dtype == HL_F64) uniform_moments<double>(env->lattice, ux, uy, uz, m);
    else uniform_moments<float>(env->lattice, ux, uy, uz, m);
    hl::EnvArgs a = make_args(env, 0);
    hipError_t e = hl::init_uniform(env->lattice, env->dtype, a, env->moments[buf], m, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "init kernel launch");
    return HL_OK;
}

int32_t hl_voxelize(const hl_env* env, uint8_t* mask, const float* caps, int32_t ncap, const int32_t box[6], void* stream)
{
    int32_t rc = validate(env, false);
    if (rc) return rc;
    if (!mask) return fail(HL_E_INVALID, "mask is NULL");
    if (ncap < 0 || (ncap > 0 && !caps)) return fail(HL_E_INVALID, "bad capsule table");
    if (!box) return fail(HL_E_INVALID, "box is NULL");
    if (box[0] < 0 || box[1] > env->nx || box[2] < 0 || box[3] > env->ny || box[4] < 0 || box[5] > env->nz)
        return fail(HL_E_INVALID, "voxelize box outside the grid");
    hl::EnvArgs a = make_args(env, 0);
    hipError_t e = hl::voxelize(env->lattice, a, mask, caps, ncap, box, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "voxelize kernel launch");
    return HL_OK;
}

int32_t hl_set_tiling(int32_t lattice, int32_t tx, int32_t ty, int32_t zc)
{
    if (!hl::tiling_supported(lattice, tx, ty, zc))
        return fail(HL_E_UNSUPPORTED, "no tiled kernel for lattice %d with tile %dx%dx%d", lattice, tx, ty, zc);
    hl::set_tiling(lattice, hl::TileCfg{tx, ty, zc});
    return HL_OK;
}

int32_t hl_kernel_timer(int32_t max_launches, int32_t every)
{
    if (max_launches < 0 || max_launches > (1 << 20)) return fail(HL_E_INVALID, "max_launches must be in [0, 2^20]");
    if (every < 1) return fail(HL_E_INVALID, "every must be >= 1");
    std::lock_guard<std::mutex> lock(g_timer.mutex);
    while ((int32_t)g_timer.start.size() < max_launches) {
        hipEvent_t a = nullptr, b = nullptr;
        hipError_t e = hipEventCreate(&a);
        if (e == hipSuccess) e = hipEventCreate(&b);
        if (e != hipSuccess) {
            if (a) (void)hipEventDestroy(a);
            return hip_fail(e, "hipEventCreate");
        }
        g_timer.start.push_back(a);
        g_timer.end.push_back(b);
    }
    g_timer.cap = max_launches;
    g_timer.every = every;
    g_timer.used = 0;
    g_timer.seen = 0;
    return HL_OK;
}

int32_t hl_kernel_timer_read(double* total_ms, int32_t* launches)
{
    if (!total_ms || !launches) return fail(HL_E_INVALID, "NULL output");
    std::lock_guard<std::mutex> lock(g_timer.mutex);
    double sum = 0.0;
    for (size_t k = 0; k < g_timer.used; ++k) {
        hipError_t e = hipEventSynchronize(g_timer.end[k]);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, g_timer.start[k], g_timer.end[k]);
        if (e != hipSuccess) return hip_fail(e, "kernel timer read");
        sum += ms;
    }
    *total_ms = sum;
    *launches = (int32_t)g_timer.used;
    g_timer.used = 0;
    return HL_OK;
}

int32_t hl_copy_state(const hl_env* env, int32_t from, void* stream)
{
    int32_t rc = validate(env);
    if (rc) return rc;
    if (from != 0 && from != 1) return fail(HL_E_INVALID, "from must be 0 or 1");
    hl::EnvArgs a = make_args(env, from);
    hipError_t e = hl::copy_state(env->lattice, env->dtype, a, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? HL_OK : hip_fail(e, "copy kernel launch");
}

int32_t hl_get_tiling(int32_t lattice, int32_t* tx, int32_t* ty, int32_t* zc)
{
    if (!tx || !ty || !zc) return fail(HL_E_INVALID, "NULL output");
    hl::TileCfg c = hl::get_tiling(lattice);
    *tx = c.tx; *ty = c.ty; *zc = c.zc;
    return HL_OK;
}

int32_t hl_voxelize_polygons(const hl_env* env, uint8_t* mask, const float* verts, const int32_t* counts,
                             int32_t npoly, int32_t maxv, const int32_t box[6], void* stream)
{
    int32_t rc = validate(env, false);
    if (rc) return rc;
    if (env->lattice != HL_D2Q9) return fail(HL_E_UNSUPPORTED, "polygon voxelizer is D2Q9 only");
    if (!mask || !box) return fail(HL_E_INVALID, "mask or box is NULL");
    if (npoly < 0 || npoly > 255 || maxv < 0 || (npoly > 0 && (!verts || !counts)))
        return fail(HL_E_INVALID, "bad polygon table (npoly %d, maxv %d)", npoly, maxv);
    if (box[0] < 0 || box[1] > env->nx || box[2] < 0 || box[3] > env->ny)
        return fail(HL_E_INVALID, "voxelize box outside the grid");
    hl::EnvArgs a = make_args(env, 0);
    hipError_t e = hl::voxelize_polygons(a, mask, verts, counts, npoly, maxv, box, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "polygon voxelize kernel launch");
    return HL_OK;
}

int32_t hl_mass(const hl_env* env, int32_t cur, double* out, void* stream)
{
    int32_t rc = validate(env);
    if (rc) return rc;
    if (!out) return fail(HL_E_INVALID, "out is NULL");
    if (cur != 0 && cur != 1) return fail(HL_E_INVALID, "cur must be 0 or 1");
    hl::EnvArgs a = make_args(env, cur);
    hipStream_t s = static_cast<hipStream_t>(stream);
    StreamCtx* ctx = nullptr;
    rc = stream_ctx(s, &ctx);
    if (rc) return rc;
    std::lock_guard<std::mutex> ctx_lock(ctx->mutex);
    void* ws = nullptr;
    ctx->list.valid = false;  // ws is reused below
    rc = workspace(ctx, s, hl::mass_workspace_bytes(a, env->lattice), &ws);
    if (rc) return rc;
    hipError_t e = hl::mass(env->lattice, env->dtype, a, out, ws, s);
    if (e != hipSuccess) return hip_fail(e, "mass kernel launch");
    return HL_OK;
}

int32_t hl_voxelize_mesh(const hl_env* env, uint8_t* mask, const float* tris, const int32_t* tri_solid,
                         int32_t ntri, const int32_t box[6], int32_t merge, int32_t* overflow, void* stream)
{
    int32_t rc = validate(env);
    if (rc) return rc;
    if (env->lattice != HL_D3Q27) return fail(HL_E_UNSUPPORTED, "mesh bodies are for D3Q27 lattices");
    if (!mask || !box || (ntri > 0 && (!tris || !tri_solid))) return fail(HL_E_INVALID, "a buffer is NULL");
    if (ntri < 0) return fail(HL_E_INVALID, "ntri must be >= 0");
    const int n3[3] = {env->nx, env->ny, env->nz};
    for (int k = 0; k < 3; ++k)
        if (box[2 * k] < 0 || box[2 * k + 1] > n3[k] || box[2 * k] > box[2 * k + 1])
            return fail(HL_E_INVALID, "box [%d, %d) outside axis %d of extent %d", box[2 * k], box[2 * k + 1], k, n3[k]);
    const hl::EnvArgs a = make_args(env, 0);
    hipError_t e = hl::voxelize_mesh(a, mask, tris, tri_solid, ntri, box, merge, overflow,
                                     static_cast<hipStream_t>(stream));
    return e == hipSuccess ? HL_OK : hip_fail(e, "mesh voxelizer launch");
}

int32_t hl_render_plane(const hl_env* env, int32_t cur, int32_t kind, int32_t normal, int32_t index,
                        float thickness, float max_velocity, float* out, void* stream)
{
    int32_t rc = validate(env);
    if (rc) return rc;
    if (!out) return fail(HL_E_INVALID, "out is NULL");
    if (cur != 0 && cur != 1) return fail(HL_E_INVALID, "cur must be 0 or 1");
    if (kind < HL_RENDER_SPEED || kind > HL_RENDER_SOLID) return fail(HL_E_INVALID, "unknown render kind %d", kind);
    const bool three = env->lattice == HL_D3Q27;
    if (normal != 2 && !(three && normal == 0)) return fail(HL_E_INVALID, "plane normal %d not supported", normal);
    const int n = normal == 2 ? env->nz : env->nx;
    if (index < 0 || index >= (three ? n : 1)) return fail(HL_E_INVALID, "plane index %d out of range", index);
    if (!(thickness >= 0.0f) || thickness > (float)hl::max_render_thickness())
        return fail(HL_E_INVALID, "thickness must be in [0, %d]", hl::max_render_thickness());
    if (kind == HL_RENDER_RGB && !(max_velocity > 0.0f)) return fail(HL_E_INVALID, "max_velocity must be > 0");
    const hl::EnvArgs a = make_args(env, cur);
    const float gl = (float)(1.0 / env->nx);
    hipError_t e = hl::render_plane(env->dtype, a, kind, normal, index, gl, thickness, max_velocity, out,
                                    static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "render kernel launch");
    return HL_OK;
}

int32_t hl_render_projection(const hl_env* env, int32_t cur, int32_t kind, int32_t axis, float* out, void* stream)
{
    int32_t rc = validate(env);
    if (rc) return rc;
    if (!out) return fail(HL_E_INVALID, "out is NULL");
    if (cur != 0 && cur != 1) return fail(HL_E_INVALID, "cur must be 0 or 1");
    if (env->lattice != HL_D3Q27) return fail(HL_E_UNSUPPORTED, "projections are for D3Q27 lattices");
    if (kind < HL_PROJECT_FIRST_SPEED || kind > HL_PROJECT_MAX_VORTICITY)
        return fail(HL_E_INVALID, "unknown projection kind %d", kind);
    if (axis < 0 || axis > 2) return fail(HL_E_INVALID, "axis must be 0, 1 or 2");
    const hl::EnvArgs a = make_args(env, cur);
    const float gl = (float)(1.0 / env->nx);
    hipError_t e = hl::render_projection(env->dtype, a, kind, axis, gl, out, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "projection kernel launch");
    return HL_OK;
}

int32_t hl_couple_poses(int32_t lattice, int32_t nworld, int32_t nbody, int32_t nsolid, const float* pos,
                        int32_t pos_stride, const float* quat, int32_t quat_stride, const int32_t* solid_ids,
                        const float* engine_origin, const float* lattice_origin, const float* scale,
                        float* const* tables, int32_t* initialized, void* stream)
{
    if (lattice != HL_D2Q9 && lattice != HL_D3Q27) return fail(HL_E_INVALID, "unknown lattice %d", lattice);
    if (nworld < 1 || nbody < 1 || nbody > HL_MAX_SOLIDS || nsolid < 1 || nsolid > HL_MAX_SOLIDS)
        return fail(HL_E_INVALID, "need nworld >= 1 and 1 <= nbody, nsolid <= %d", HL_MAX_SOLIDS);
    if ((int64_t)nworld * nbody > (1 << 24)) return fail(HL_E_INVALID, "too many bodies");
    if (pos_stride < 3 || quat_stride < 4) return fail(HL_E_INVALID, "pos_stride >= 3 and quat_stride >= 4");
    if (!pos || !quat || !solid_ids || !engine_origin || !lattice_origin || !scale || !tables || !initialized)
        return fail(HL_E_INVALID, "a buffer is NULL");
    hipError_t e = hl::couple_poses(lattice, nworld, nbody, nsolid, pos, pos_stride, quat, quat_stride, solid_ids,
                                    engine_origin, lattice_origin, scale, tables, initialized,
                                    static_cast<hipStream_t>(stream));
    return e == hipSuccess ? HL_OK : hip_fail(e, "pose kernel launch");
}

int32_t hl_couple_forces(int32_t lattice, int32_t mode, int32_t nworld, int32_t nbody, int32_t nsolid,
                         const double* rows, int64_t world_stride, const int32_t* solid_ids, const float* scale,
                         float force_conv, float torque_conv, float* forces, float* torques, float* xfrc,
                         int32_t xfrc_bodies, const int32_t* body_ids, void* stream)
{
    if (lattice != HL_D2Q9 && lattice != HL_D3Q27) return fail(HL_E_INVALID, "unknown lattice %d", lattice);
    if (mode != HL_COUPLE_SCALED && mode != HL_COUPLE_PHYSICAL) return fail(HL_E_INVALID, "unknown mode %d", mode);
    if (nworld < 1 || nbody < 1 || nbody > HL_MAX_SOLIDS || nsolid < 1 || nsolid > HL_MAX_SOLIDS)
        return fail(HL_E_INVALID, "need nworld >= 1 and 1 <= nbody, nsolid <= %d", HL_MAX_SOLIDS);
    if ((int64_t)nworld * nbody > (1 << 24)) return fail(HL_E_INVALID, "too many bodies");
    if (world_stride < 6 * (int64_t)nsolid) return fail(HL_E_INVALID, "world_stride < 6 * nsolid");
    if (!rows || !solid_ids || !forces || !torques || (mode == HL_COUPLE_SCALED && !scale))
        return fail(HL_E_INVALID, "a buffer is NULL");
    if (xfrc && (!body_ids || xfrc_bodies < 1)) return fail(HL_E_INVALID, "xfrc needs body_ids and xfrc_bodies >= 1");
    hipError_t e = hl::couple_forces(lattice, mode, nworld, nbody, nsolid, rows, world_stride, solid_ids, scale,
                                     force_conv, torque_conv, forces, torques, xfrc, xfrc_bodies, body_ids,
                                     static_cast<hipStream_t>(stream));
    return e == hipSuccess ? HL_OK : hip_fail(e, "force kernel launch");
}

int32_t hl_free_bodies_step(int32_t nworld, int32_t nbody, float* state, const float* params, const float* forces,
                            const float* torques, float gx, float gy, float gz, float dt, void* stream)
{
    if (nworld < 1 || nbody < 1 || (int64_t)nworld * nbody > (1 << 24))
        return fail(HL_E_INVALID, "need nworld, nbody >= 1");
    if (!state || !params || !forces || !torques) return fail(HL_E_INVALID, "a buffer is NULL");
    if (!(dt > 0.0f)) return fail(HL_E_INVALID, "dt must be > 0");
    const float g[3] = {gx, gy, gz};
    hipError_t e = hl::free_bodies_step(nworld, nbody, state, params, forces, torques, g, dt,
                                        static_cast<hipStream_t>(stream));
    return e == hipSuccess ? HL_OK : hip_fail(e, "body kernel launch");
}

int32_t hl_body_capsules(int32_t nworld, int32_t nbody, const float* pos, int32_t pos_stride, const float* quat,
                         int32_t quat_stride, const float* shapes, const int32_t* solid_ids,
                         const float* engine_origin, const float* lattice_origin, const float* scale, float* caps,
                         void* stream)
{
    if (nworld < 1 || nbody < 1 || (int64_t)nworld * nbody > (1 << 24))
        return fail(HL_E_INVALID, "need nworld, nbody >= 1");
    if (pos_stride < 3 || quat_stride < 4) return fail(HL_E_INVALID, "pos_stride >= 3 and quat_stride >= 4");
    if (!pos || !quat || !shapes || !solid_ids || !engine_origin || !lattice_origin || !scale || !caps)
        return fail(HL_E_INVALID, "a buffer is NULL");
    hipError_t e = hl::body_capsules(nworld, nbody, pos, pos_stride, quat, quat_stride, shapes, solid_ids,
                                     engine_origin, lattice_origin, scale, caps, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? HL_OK : hip_fail(e, "capsule kernel launch");
}

}  // extern "C"
