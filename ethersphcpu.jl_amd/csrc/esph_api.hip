// esph_api.hip — host runtime behind the C ABI declared in include/ethersph.h.
//
// A scene owns its device buffers (two sorted copies of the fluid state, step
// temporaries, sort scratch, the static wall set) plus a SceneDev table entry
// describing them.  Stepping launches the kernels of esph_kernels.hip on the
// context's stream; nothing is synchronised unless the caller asks for data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ethersph.h"
#include "esph_device.hpp"
#include "esph_launch.hpp"
#include "esph_runtime.hpp"

using namespace esph;

namespace esph {

int fail(esph_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int hip_fail(esph_ctx* ctx, const char* what, hipError_t e) {
    return fail(ctx, ESPH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void* Arena::get(size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) bytes = 16;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    total += bytes;
    return p;
}

void Arena::release() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
    total = 0;
}

// SmoothKernel(R, dim, K) derived quantities, SPHKernel/SmoothKernel/SmoothKernel.jl:37-48.
bool kernel_params(int kind, int dim, double R, double* h, double* sigma) {
    static const double PI = 3.14159265358979323846;
    double ratio, c[3];
    switch (kind) {
        case ESPH_CUBIC_SPLINE: ratio = 2.0; c[0] = 2.0 / 3.0; c[1] = 10.0 / 7.0 / PI; c[2] = 1.0 / PI; break;
        case ESPH_GAUSSIAN: ratio = 3.0; c[0] = 1.0 / std::sqrt(PI); c[1] = 1.0 / PI; c[2] = 1.0 / std::sqrt(PI * PI * PI); break;
        case ESPH_WENDLAND_C2: ratio = 2.0; c[0] = 0.0; c[1] = 7.0 / 4.0 / PI; c[2] = 21.0 / 16.0 / PI; break;
        case ESPH_WENDLAND_C4: ratio = 2.0; c[0] = 5.0 / 8.0; c[1] = 9.0 / 4.0 / PI; c[2] = 495.0 / 256.0 / PI; break;
        default: return false;
    }
    if (dim < 1 || dim > 3) return false;
    *h = R / ratio;
    double hd = *h;
    if (dim == 2) hd = *h * *h;
    if (dim == 3) hd = *h * *h * *h;
    *sigma = c[dim - 1] / hd;
    return true;
}

}  // namespace esph

// ---------------------------------------------------------------------------
// Staging of esph_scene_probe_noalloc: the probe points (dim doubles each)
// and the interpolated values (up to dim doubles each) of at most
// kProbeNoallocPoints points, in a static device array (3 MiB), so that
// entry point needs no allocation.  Each device holds its own copy at its own
// address, resolved once per context at esph_ctx_create with the context's
// device current; host code uses that address, never the array's name.
// ---------------------------------------------------------------------------
constexpr int64_t kProbeNoallocPoints = 65536;
static __device__ __align__(256) double scratch[kProbeNoallocPoints * 6];

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int esph_abi_version(void) { return ESPH_ABI_VERSION; }

extern "C" int esph_ctx_create(int32_t device, uint64_t hbm_budget_bytes, esph_ctx** out) {
    if (!out) return ESPH_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ESPH_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return ESPH_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return ESPH_ERR_HIP;
    esph_ctx* c = new esph_ctx();
    c->device = device;
    c->budget = hbm_budget_bytes;
    {
        void* p = nullptr;
        c->probe_scratch_err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        c->probe_scratch = c->probe_scratch_err == hipSuccess ? static_cast<double*>(p) : nullptr;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return ESPH_ERR_HIP;
    }
    *out = c;
    return ESPH_OK;
}

extern "C" int esph_ctx_destroy(esph_ctx* ctx) {
    if (!ctx) return ESPH_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return ESPH_OK;
}

extern "C" const char* esph_last_error(const esph_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int esph_ctx_set_tuning(esph_ctx* ctx, int32_t mode, int32_t depth) {
    if (!ctx || (mode != 0 && mode != 1) || (mode == 0 && depth != 1 && depth != 2 && depth != 4))
        return fail(ctx, ESPH_ERR_ARG, "set_tuning: mode 0 needs depth 1, 2 or 4; mode 1 tunes");
    ctx->tune = mode;
    ctx->fixed_depth = mode == 0 ? depth : 2;
    ctx->depth.clear();
    ctx->depth_ms.clear();
    ctx->block.clear();
    ctx->block_ms.clear();
    ctx->pass_cfg.clear();
    ctx->pass_ms.clear();
    return ESPH_OK;
}

extern "C" int esph_ctx_set_workgroup(esph_ctx* ctx, int32_t threads) {
    if (!ctx || (threads != 64 && threads != 128 && threads != 256))
        return fail(ctx, ESPH_ERR_ARG, "set_workgroup: 64, 128 or 256 threads");
    ctx->fixed_block = threads;
    return ESPH_OK;
}

extern "C" int esph_ctx_set_brick_walk(esph_ctx* ctx, int32_t enable) {
    if (!ctx) return ESPH_ERR_ARG;
    ctx->bricks = enable != 0;
    return ESPH_OK;
}

extern "C" int esph_ctx_set_wall_lanes(esph_ctx* ctx, int32_t enable) {
    if (!ctx) return ESPH_ERR_ARG;
    ctx->wall_lanes = enable != 0;
    return ESPH_OK;
}

extern "C" int esph_ctx_tuned_workgroup(const esph_ctx* ctx, int32_t dim, int32_t kernel, int32_t precision,
                                        int32_t thermal_mix, int32_t* threads, float* ms3) {
    if (!ctx || !threads) return ESPH_ERR_ARG;
    auto it = ctx->block.find({dim, kernel, precision, thermal_mix});
    if (it == ctx->block.end()) return ESPH_ERR_STATE;
    *threads = it->second;
    if (ms3) {
        const auto& m = ctx->block_ms.at(it->first);
        for (int q = 0; q < 3; ++q) ms3[q] = m[q];
    }
    return ESPH_OK;
}

extern "C" int esph_ctx_tuned_depth(const esph_ctx* ctx, int32_t dim, int32_t kernel, int32_t precision,
                                    int32_t thermal_mix, int32_t* depth, float* ms3) {
    if (!ctx || !depth) return ESPH_ERR_ARG;
    auto it = ctx->depth.find({dim, kernel, precision, thermal_mix});
    if (it == ctx->depth.end()) return ESPH_ERR_STATE;
    *depth = it->second;
    if (ms3) {
        const auto& m = ctx->depth_ms.at(it->first);
        for (int q = 0; q < 3; ++q) ms3[q] = m[q];
    }
    return ESPH_OK;
}

extern "C" int esph_ctx_tuned_pass(const esph_ctx* ctx, int32_t dim, int32_t kernel, int32_t precision,
                                   int32_t thermal_mix, int32_t pass, int32_t* depth, int32_t* threads, float* ms9) {
    if (!ctx || !depth || !threads || pass < 0 || pass > 2) return ESPH_ERR_ARG;
    const std::array<int, 4> key{dim, kernel, precision, thermal_mix};
    if (pass == 0) {
        auto id = ctx->depth.find(key);
        auto ib = ctx->block.find(key);
        if (id == ctx->depth.end() || ib == ctx->block.end()) return ESPH_ERR_STATE;
        *depth = id->second;
        *threads = ib->second;
        if (ms9) {
            // depths 1, 2, 4 at 256 threads, then 64, 128, 256 threads at the chosen depth
            const auto& a = ctx->depth_ms.at(key);
            const auto& b = ctx->block_ms.at(key);
            for (int q = 0; q < 3; ++q) ms9[q] = a[q];
            for (int q = 0; q < 3; ++q) ms9[3 + q] = b[q];
            for (int q = 6; q < 9; ++q) ms9[q] = 0.0f;
        }
        return ESPH_OK;
    }
    auto it = ctx->pass_cfg.find(key);
    if (it == ctx->pass_cfg.end()) return ESPH_ERR_STATE;
    *depth = it->second[pass - 1][0];
    *threads = it->second[pass - 1][1];
    if (ms9) {
        const auto& m = ctx->pass_ms.at(key)[pass - 1];
        for (int q = 0; q < 9; ++q) ms9[q] = m[q];
    }
    return ESPH_OK;
}

extern "C" int esph_ctx_memory_used(const esph_ctx* ctx, uint64_t* bytes) {
    if (!ctx || !bytes) return ESPH_ERR_ARG;
    *bytes = ctx->used;
    return ESPH_OK;
}

// ---------------------------------------------------------------------------
// kernel evaluation on the device
// ---------------------------------------------------------------------------
extern "C" int esph_kernel_eval(esph_ctx* ctx, int32_t kernel, int32_t dim, double influence_radius, int64_t n,
                                const double* r, double* w, double* dw) {
    if (!ctx || n < 0 || (n > 0 && (!r || !w || !dw))) return fail(ctx, ESPH_ERR_ARG, "esph_kernel_eval: bad argument");
    double h, sigma;
    if (!kernel_params(kernel, dim, influence_radius, &h, &sigma) || !(influence_radius > 0))
        return fail(ctx, ESPH_ERR_ARG, "esph_kernel_eval: bad kernel/dim/radius");
    if (n == 0) return ESPH_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    Arena a;
    double* dr = (double*)a.get(sizeof(double) * n);
    double* dw0 = (double*)a.get(sizeof(double) * n);
    double* dw1 = (double*)a.get(sizeof(double) * n);
    if (!dr || !dw0 || !dw1) { a.release(); return fail(ctx, ESPH_ERR_HIP, "esph_kernel_eval: allocation failed"); }
    hipError_t e = hipMemcpyAsync(dr, r, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_kernel_eval(kernel, n, dr, dw0, dw1, h, sigma, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(w, dw0, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dw, dw1, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    a.release();
    if (e != hipSuccess) return hip_fail(ctx, "esph_kernel_eval", e);
    return ESPH_OK;
}

// ---------------------------------------------------------------------------
// scenes
// ---------------------------------------------------------------------------
extern "C" int esph_scene_create(esph_ctx* ctx, const esph_scene_desc* desc, esph_scene** out) {
    if (!ctx || !desc || !out) return fail(ctx, ESPH_ERR_ARG, "esph_scene_create: null argument");
    *out = nullptr;
    const esph_scene_desc& d = *desc;
    if (d.dim != 2 && d.dim != 3) return fail(ctx, ESPH_ERR_ARG, "dim must be 2 or 3");
    double h, sigma;
    if (!(d.influence_radius > 0) || !kernel_params(d.kernel, d.dim, d.influence_radius, &h, &sigma))
        return fail(ctx, ESPH_ERR_ARG, "bad kernel or influence_radius");
    if (d.kernel == ESPH_WENDLAND_C2 && d.dim == 1) return fail(ctx, ESPH_ERR_ARG, "WendlandC2 has no 1D form");
    if (d.solver < ESPH_SOLVER_FLUID_WALL || d.solver > ESPH_SOLVER_FLUID_INLET_OUTLET_WALL)
        return fail(ctx, ESPH_ERR_ARG, "bad solver");
    if (d.check_nplanes < 0 || d.check_nplanes > 8) return fail(ctx, ESPH_ERR_ARG, "check_nplanes must be 0..8");
    if (d.solver == ESPH_SOLVER_FLUID_INLET_OUTLET_WALL) {
        if (d.io_axis < 0 || d.io_axis >= d.dim) return fail(ctx, ESPH_ERR_ARG, "io_axis must be 0..dim-1");
        if (d.profile < ESPH_PROFILE_NONE || d.profile > ESPH_PROFILE_PIPE_3D ||
            (d.profile == ESPH_PROFILE_CHANNEL_2D && d.dim != 2) || (d.profile == ESPH_PROFILE_PIPE_3D && d.dim != 3))
            return fail(ctx, ESPH_ERR_ARG, "bad buffer profile for this dim");
        if (d.profile != ESPH_PROFILE_NONE && (!(d.profile_nu > 0) || !(d.profile_l > 0) || d.profile_terms < 0))
            return fail(ctx, ESPH_ERR_ARG, "profile needs nu > 0, l > 0 and terms >= 0");
    }
    if (d.precision != ESPH_FP64 && d.precision != ESPH_FP32_BULK) return fail(ctx, ESPH_ERR_ARG, "bad precision");
    if (d.gamma < 1) return fail(ctx, ESPH_ERR_ARG, "gamma must be >= 1");
    if (!(d.rho_0 > 0) || !(d.dt > 0)) return fail(ctx, ESPH_ERR_ARG, "rho_0 and dt must be positive");
    esph_scene* s = new esph_scene();
    s->ctx = ctx;
    s->d = d;
    s->kh = h;
    s->ksigma = sigma;
    *out = s;
    return ESPH_OK;
}

static void free_scene_device(esph_scene* s) {
    s->gplan.release();
    if (s->timing_events.size()) {
        for (hipEvent_t e : s->timing_events) (void)hipEventDestroy(e);
        s->timing_events.clear();
    }
    s->ctx->used -= std::min<uint64_t>(s->ctx->used, s->arena.total);
    s->arena.release();
    s->allocated = false;
    s->fluid_ready = false;
}

extern "C" int esph_scene_destroy(esph_scene* s) {
    if (!s) return ESPH_ERR_ARG;
    if (s->batch_refs > 0) return fail(s->ctx, ESPH_ERR_STATE, "scene_destroy: the scene is a member of a live batch");
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    free_scene_device(s);
    delete s;
    return ESPH_OK;
}

extern "C" int esph_scene_set_walls(esph_scene* s, int32_t ngroups, const esph_wall_arrays* g) {
    if (!s || ngroups < 0 || (ngroups > 0 && !g)) return fail(s ? s->ctx : nullptr, ESPH_ERR_ARG, "set_walls: bad argument");
    if (s->allocated) return fail(s->ctx, ESPH_ERR_STATE, "set_walls must precede the first set_fluid");
    const int D = s->d.dim;
    int64_t nw = 0;
    for (int q = 0; q < ngroups; ++q) {
        if (g[q].n < 0 || (g[q].n > 0 && (!g[q].x || !g[q].normal || !g[q].gap)))
            return fail(s->ctx, ESPH_ERR_ARG, "set_walls: group missing x/normal/gap");
        if (g[q].kind < ESPH_WALL_COMPULSIVE || g[q].kind > ESPH_WALL_THERMOSTATIC)
            return fail(s->ctx, ESPH_ERR_ARG, "set_walls: bad wall kind");
        if (g[q].kind == ESPH_WALL_STATIC_VELOCITY && g[q].n > 0 && !g[q].v)
            return fail(s->ctx, ESPH_ERR_ARG, "set_walls: static-velocity walls need v");
        if (g[q].kind == ESPH_WALL_THERMOSTATIC && g[q].n > 0 && (!g[q].t || !g[q].kappa))
            return fail(s->ctx, ESPH_ERR_ARG, "set_walls: thermostatic walls need t and kappa");
        nw += g[q].n;
    }
    if (nw >= ((int64_t)1 << 27)) return fail(s->ctx, ESPH_ERR_ARG, "set_walls: too many wall particles (at most 2^27 - 1)");
    WallHost& w = s->walls;
    w = WallHost();
    w.n = nw;
    w.x.resize(nw * D); w.n_.resize(nw * D); w.v.assign(nw * D, 0.0);
    w.gap.resize(nw); w.T.assign(nw, 0.0); w.kappa.assign(nw, 0.0); w.kind.resize(nw);
    int64_t o = 0;
    for (int q = 0; q < ngroups; ++q) {
        const int64_t n = g[q].n;
        if (n == 0) continue;
        std::memcpy(&w.x[o * D], g[q].x, sizeof(double) * n * D);
        std::memcpy(&w.n_[o * D], g[q].normal, sizeof(double) * n * D);
        std::memcpy(&w.gap[o], g[q].gap, sizeof(double) * n);
        if (g[q].kind == ESPH_WALL_STATIC_VELOCITY) std::memcpy(&w.v[o * D], g[q].v, sizeof(double) * n * D);
        if (g[q].kind == ESPH_WALL_THERMOSTATIC) {
            std::memcpy(&w.T[o], g[q].t, sizeof(double) * n);
            std::memcpy(&w.kappa[o], g[q].kappa, sizeof(double) * n);
        }
        for (int64_t i = 0; i < n; ++i) w.kind[o + i] = g[q].kind;
        o += n;
    }
    s->walls_set = true;
    return ESPH_OK;
}

extern "C" int esph_scene_set_buffers(esph_scene* s, const esph_fluid_arrays* inlet, const esph_fluid_arrays* outlet) {
    if (!s || !inlet || !outlet) return fail(s ? s->ctx : nullptr, ESPH_ERR_ARG, "set_buffers: null argument");
    if (s->d.solver != ESPH_SOLVER_FLUID_INLET_OUTLET_WALL)
        return fail(s->ctx, ESPH_ERR_ARG, "set_buffers: only the inlet/outlet solver has buffers");
    if (s->allocated) return fail(s->ctx, ESPH_ERR_STATE, "set_buffers must precede the first set_fluid");
    const int D = s->d.dim;
    for (const esph_fluid_arrays* b : {inlet, outlet})
        if (b->n < 0 || (b->n > 0 && (!b->x || !b->v || !b->rho || !b->p || !b->c || !b->mass || !b->gap)))
            return fail(s->ctx, ESPH_ERR_ARG, "set_buffers: missing arrays");
    if (inlet->n + outlet->n > INT32_MAX / 4) return fail(s->ctx, ESPH_ERR_ARG, "set_buffers: too many particles");
    s->n_inlet = inlet->n;
    s->n_outlet = outlet->n;
    auto cat = [&](std::vector<double>& dst, const double* a, const double* b, int w) {
        dst.resize((size_t)(inlet->n + outlet->n) * w);
        if (inlet->n) std::memcpy(dst.data(), a, sizeof(double) * inlet->n * w);
        if (outlet->n) std::memcpy(dst.data() + inlet->n * w, b, sizeof(double) * outlet->n * w);
    };
    cat(s->bx, inlet->x, outlet->x, D);
    cat(s->bv, inlet->v, outlet->v, D);
    cat(s->brho, inlet->rho, outlet->rho, 1);
    cat(s->bp, inlet->p, outlet->p, 1);
    cat(s->bc, inlet->c, outlet->c, 1);
    cat(s->bm, inlet->mass, outlet->mass, 1);
    cat(s->bgap, inlet->gap, outlet->gap, 1);
    return ESPH_OK;
}

// Cell grid over the bounding box of fluid, walls and finite check bounds
// (or the caller's domain), one cell of padding.  Cells are 1.0001 Rv / 2
// wide along x (a 5-cell stencil: about 300 candidates per particle in 3D
// instead of about 500 with Rv-wide cells) and `split` times finer along y
// and z, split the smallest with cells under 0.8 of the particle spacing
// there (the first particle's gap; 2 for the examples' h / dx): a row of
// cells along x then holds at most one lattice line of particles, the
// particles of a wavefront (consecutive in cell order) lie along one line,
// and their candidate lists, built row by row, address their neighbours in
// step -- entry k of consecutive lanes lies on a neighbouring line at
// consecutive positions, a few cache lines per gather instead of dozens.
// Cells widen (all axes) while the grid would exceed 2^27 of them.
static void build_grid(esph_scene* s, const esph_fluid_arrays* f, SceneDev& hd) {
    const int D = s->d.dim;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    auto add = [&](const double* x, int64_t n) {
        for (int64_t i = 0; i < n; ++i)
            for (int k = 0; k < D; ++k) {
                const double v = x[i * D + k];
                if (std::isfinite(v)) { lo[k] = std::min(lo[k], v); hi[k] = std::max(hi[k], v); }
            }
    };
    if (s->d.domain_set) {
        for (int k = 0; k < D; ++k) { lo[k] = s->d.domain_min[k]; hi[k] = s->d.domain_max[k]; }
    } else {
        add(f->x, f->n);
        if (s->walls.n) add(s->walls.x.data(), s->walls.n);
        if (!s->bx.empty()) add(s->bx.data(), s->n_inlet + s->n_outlet);
        if (s->d.check_enabled)
            for (int k = 0; k < D; ++k) {
                if (std::isfinite(s->d.check_min[k])) lo[k] = std::min(lo[k], s->d.check_min[k]);
                if (std::isfinite(s->d.check_max[k])) hi[k] = std::max(hi[k], s->d.check_max[k]);
            }
    }
    for (int k = 0; k < D; ++k)
        if (!(lo[k] <= hi[k])) { lo[k] = 0.0; hi[k] = 0.0; }
    double cs0 = hd.Rv * 1.0001 / 2.0;
    const double spacing = (f->n > 0 && f->gap && std::isfinite(f->gap[0]) && f->gap[0] > 0.0) ? f->gap[0] : cs0;
    const int split = (int)std::min(4.0, std::max(1.0, std::ceil(cs0 / (0.8 * spacing))));
    double cs[3];
    for (;;) {
        int64_t ncell = 1;
        for (int k = 0; k < 3; ++k) {
            cs[k] = k == 0 ? cs0 : cs0 / split;
            if (k >= D) { hd.nc[k] = 1; hd.lo[k] = 0.0; continue; }
            hd.lo[k] = lo[k] - cs[k];
            const int64_t n = (int64_t)std::floor((hi[k] - hd.lo[k]) / cs[k]) + 2;
            hd.nc[k] = (int32_t)std::min<int64_t>(n, 1 << 20);
            ncell *= hd.nc[k];
        }
        if (ncell <= (int64_t)1 << 27) { hd.ncell = (int32_t)ncell; break; }
        cs0 *= 1.25;  // wider cells keep the search exact
    }
    double extent = 0.0;
    for (int k = 0; k < 3; ++k) {
        hd.cs[k] = cs[k];
        hd.inv_cs[k] = 1.0 / cs[k];
        hd.stencil[k] = (int32_t)std::ceil(hd.Rv / cs[k]);
        if ((double)hd.stencil[k] * cs[k] <= hd.Rv) hd.stencil[k] += 1;
        if (k < D) extent = std::max(extent, (double)hd.nc[k] * cs[k]);
    }
    // position quantum (FluidSet::xq): the smallest power of two with the
    // grid extent within 2^30 quanta
    const int e = (int)std::ceil(std::log2(std::max(extent, 1e-300) / 1073741824.0));
    hd.qu = std::ldexp(1.0, e);
    hd.inv_qu = std::ldexp(1.0, -e);
    // band for the quantized float distance near R (d <= 1.01 R): each axis
    // of rv is off the binary64 difference by at most one quantum (two
    // half-quantum roundings) plus one float rounding of the integer
    // difference, dx = u + 2^-24 |rv|; then |d2 - d^2| <= 2 sqrt(D) d dx
    // + D dx^2, the float sum adds D + 1 roundings of d^2 and R2q one of R^2.
    // Doubled for margin.
    const double eps = std::ldexp(1.0, -24), dd = 1.01 * hd.R;
    const double dx = hd.qu + eps * dd;
    const double band = 2.0 * std::sqrt((double)D) * dd * dx + D * dx * dx + (D + 1) * eps * dd * dd +
                        eps * hd.R2;
    hd.qbandq = (float)(2.0 * band * hd.inv_qu * hd.inv_qu);
    hd.R2q = (float)(hd.R2 * hd.inv_qu * hd.inv_qu);
}

static int cell_of_host(const SceneDev& hd, int D, const double* x) {
    auto cc = [&](int k) {
        const double f = std::floor((x[k] - hd.lo[k]) * hd.inv_cs[k]);
        if (!(f >= 0.0)) return 0;
        if (f >= (double)(hd.nc[k] - 1)) return hd.nc[k] - 1;
        return (int)f;
    };
    int c = cc(0) + hd.nc[0] * cc(1);
    if (D == 3) c += hd.nc[0] * hd.nc[1] * cc(2);
    return c;
}

// Whether the force pass may walk two records per candidate and take p / rho^2
// from the equation of state (FLAG_EOS_WALK): 3D FP32_BULK, no heat, no
// buffer particles (whose p need not follow the EOS), gamma >= 1 and one gap
// for every fluid particle.
static bool eos_walk_ok(const esph_scene* s, const esph_fluid_arrays* f) {
    const esph_scene_desc& d = s->d;
    const bool thermal = d.solver == ESPH_SOLVER_FLUID_THERMOSTATIC_WALL ||
                         d.solver == ESPH_SOLVER_FLUID_COMPULSIVE_THERMOSTATIC_WALL;
    if (d.dim != 3 || d.precision != ESPH_FP32_BULK || thermal || d.solver == ESPH_SOLVER_FLUID_INLET_OUTLET_WALL ||
        d.gamma < 1 || f->n <= 0)
        return false;
    const double g0 = f->gap[0];
    if (!std::isfinite(g0)) return false;
    for (int64_t i = 1; i < f->n; ++i)
        if (f->gap[i] != g0) return false;
    return true;
}

// Brick shape of a FLAG_BRICK scene (cells per brick along x, y, z): the
// largest brick whose region (widened by the stencil on each side) holds at
// most 92% of kBrickCap particles and whose own particles fit one workgroup,
// counted on a lattice of the given spacing (a window of w cells holds at
// most floor(w cs / spacing) + 1 lattice points along an axis); among equal
// volumes the longest along x (longer contiguous rows).  False if no brick
// fits.
static bool choose_bricks(const SceneDev& hd, double spacing, int bd[3]) {
    if (!(spacing > 0.0) || !(hd.cs[0] > 0.0)) return false;
    const int mx = hd.stencil[0], my = hd.stencil[1], mz = hd.stencil[2];
    auto most = [&](int w, int k) { return std::floor((double)w * hd.cs[k] / spacing) + 1.0; };
    int best = 0;
    for (int bx = 1; bx <= 16; ++bx)
        for (int by = 1; by <= 16; ++by)
            for (int bz = 1; bz <= 16; ++bz) {
                if ((by + 2 * my) * (bz + 2 * mz) > kBrickRows || by * bz > kBrickOwn) continue;
                const double region = most(bx + 2 * mx, 0) * most(by + 2 * my, 1) * most(bz + 2 * mz, 2);
                const double own = most(bx, 0) * most(by, 1) * most(bz, 2);
                if (region > 0.92 * kBrickCap || own > kBrickThreads) continue;
                const int vol = bx * by * bz;
                if (vol > best || (vol == best && bx > bd[0])) {
                    best = vol;
                    bd[0] = bx;
                    bd[1] = by;
                    bd[2] = bz;
                }
            }
    return best > 0;
}

// Allocates every device buffer of the scene (capacity cap) and fills the
// static parts of SceneDev.
static int allocate_scene(esph_scene* s, const esph_fluid_arrays* f) {
    esph_ctx* ctx = s->ctx;
    const esph_scene_desc& d = s->d;
    const int D = d.dim;
    const bool io = d.solver == ESPH_SOLVER_FLUID_INLET_OUTLET_WALL;
    const int64_t n_all = f->n + s->n_inlet + s->n_outlet;  // buffers share the particle arrays
    // the inlet/outlet solver grows by inlet copies: default headroom of 2x
    int64_t cap64 = d.fluid_capacity > 0 ? d.fluid_capacity : (io ? 2 * n_all + 1024 : n_all);
    if (cap64 < n_all) return fail(ctx, ESPH_ERR_ARG, "fluid_capacity smaller than the upload");
    // gathers take 32-bit byte offsets (24 bytes per 3D position, urs)
    if (cap64 >= ((int64_t)1 << 27)) return fail(ctx, ESPH_ERR_ARG, "fluid capacity too large (at most 2^27 - 1)");
    const int cap = (int)std::max<int64_t>(cap64, 1);
    SceneDev hd{};
    hd.dim = D;
    hd.kernel = d.kernel;
    hd.solver = d.solver;
    hd.prec = d.precision;
    const bool thermal = d.solver == ESPH_SOLVER_FLUID_THERMOSTATIC_WALL ||
                         d.solver == ESPH_SOLVER_FLUID_COMPULSIVE_THERMOSTATIC_WALL;
    // wall pressure: every solver but FluidWallSolve and the inlet/outlet one
    // (FluidWallSolve.jl:68-79, FluidInletOutletWallSolve.jl:121-122)
    const bool wall_p = d.solver != ESPH_SOLVER_FLUID_WALL && !io;
    hd.flags = (thermal ? FLAG_THERMAL : 0) | (wall_p ? FLAG_WALL_PRESSURE : 0) | (d.check_enabled && !io ? FLAG_CHECK : 0) |
               (io ? FLAG_IO : 0) | (eos_walk_ok(s, f) ? FLAG_EOS_WALK : 0) |
               (ctx->wall_lanes ? FLAG_WALL_LANES : 0);
    hd.io_axis = d.io_axis;
    hd.inlet_enter = d.inlet_enter;
    hd.inlet_shift = d.inlet_shift;
    hd.outlet_enter = d.outlet_enter;
    hd.outlet_exit = d.outlet_exit;
    hd.profile = d.profile;
    hd.profile_terms = d.profile_terms;
    hd.outlet_profile = d.outlet_profile;
    hd.profile_f = d.profile_f;
    hd.profile_nu = d.profile_nu;
    hd.profile_l = d.profile_l;
    for (int k = 0; k < 3; ++k) hd.profile_c[k] = d.profile_center[k];
    hd.gamma = d.gamma;
    hd.cap = cap;
    hd.kreinit = d.density_reinitialized_interval;
    hd.kcheck = d.check_interval;
    hd.R = d.influence_radius;
    hd.R2 = d.influence_radius * d.influence_radius;
    {
        // candidate lists reused across steps (not for the inlet/outlet
        // solver, whose migrations reorder the particles every step)
        const double skin = d.verlet_skin == 0.0 ? default_skin(D) : d.verlet_skin;
        hd.verlet = (skin > 0.0 && !io) ? 1 : 0;
        const double sk = hd.verlet ? skin : 0.0;
        hd.Rv = d.influence_radius * (1.0 + sk);
        hd.Rv2 = hd.Rv * hd.Rv;
        const double half = 0.5 * sk * d.influence_radius;
        hd.disp2_limit = half * half * (1.0 - 1e-9);
    }
    hd.h = s->kh;
    hd.sigma = s->ksigma;
    hd.W0 = s->ksigma;
    hd.inv_h = 1.0 / s->kh;
    switch (d.kernel) {
        case ESPH_CUBIC_SPLINE: hd.kA = s->ksigma / 4.0; hd.kB = 3.0 * s->ksigma / (4.0 * s->kh); break;
        case ESPH_GAUSSIAN: hd.kA = s->ksigma; hd.kB = 2.0 * s->ksigma / s->kh; break;
        case ESPH_WENDLAND_C2: hd.kA = s->ksigma / 16.0; hd.kB = 5.0 * s->ksigma / (8.0 * s->kh); break;
        default: hd.kA = s->ksigma / 768.0; hd.kB = s->ksigma / s->kh; break;
    }
    hd.rho0 = d.rho_0;
    hd.c0 = d.c_0;
    hd.p0 = d.p_0;
    hd.b = d.c_0 * d.c_0 * d.rho_0 / (double)d.gamma;  // LiquidModel.jl:30
    hd.rho0f = (float)d.rho_0;
    hd.inv_rho0f = (float)(1.0 / d.rho_0);
    hd.bf = (float)hd.b;
    hd.clinf = (float)(hd.b / d.rho_0);
    hd.p0f = (float)d.p_0;
    hd.nu0 = d.mu_0 / d.rho_0;                          // LiquidModel.jl:31
    hd.t0 = d.t_0;
    hd.cp = d.cp;
    hd.beta = d.beta;
    hd.dt = d.dt;
    for (int k = 0; k < 3; ++k) {
        hd.g[k] = k < D ? d.body_force[k] : 0.0;
        hd.check_min[k] = d.check_min[k];
        hd.check_max[k] = d.check_max[k];
    }
    hd.check_nplanes = d.check_enabled ? std::max(0, std::min(d.check_nplanes, 8)) : 0;
    for (int q = 0; q < 8; ++q) {
        for (int k = 0; k < 4; ++k) hd.check_planes[q][k] = d.check_planes[q][k];
    }
    build_grid(s, f, hd);
    const int nw = (int)s->walls.n;
    hd.nw = nw;
    hd.cell_tiles = (hd.ncell + 1 + kCellTile - 1) / kCellTile;
    // neighbour-list widths: 1.5x the lattice neighbour count at the
    // particles' spacing (gap of the first particle) plus 16, multiple of 8
    {
        double spacing = (f->n > 0 && f->gap && f->gap[0] > 0) ? f->gap[0] : d.influence_radius / 1.5;
        const double a = hd.Rv / spacing;
        const double est = D == 2 ? 3.14159265358979 * a * a : 4.18879020478639 * a * a * a;
        int k = (int)std::ceil(1.5 * est + 16.0);
        k = std::min(std::max((k + 15) / 16 * 16, 16), 1024);  // whole 16-entry chunks
        hd.kff = k;
        hd.kfw = nw > 0 ? std::min(std::max((k / 2 + 15) / 16 * 16, 16), 1024) : 0;
    }

    // brick walks (FLAG_BRICK; taken up when FLAG_EOS_WALK holds): 3D
    // FP32_BULK scenes without heat or buffers, bricks sized for the fluid's
    // spacing (the first particle's gap)
    size_t brick_bytes = 0;
    {
        int bd[3] = {0, 0, 0};
        const double spacing = (f->n > 0 && f->gap) ? f->gap[0] : 0.0;
        if (ctx->bricks && D == 3 && d.precision == ESPH_FP32_BULK && !thermal && !io && choose_bricks(hd, spacing, bd)) {
            int64_t nb = 1;
            for (int k = 0; k < 3; ++k) {
                hd.bdim[k] = bd[k];
                hd.nbk[k] = (hd.nc[k] + bd[k] - 1) / bd[k];
                nb *= hd.nbk[k];
            }
            if (nb < ((int64_t)1 << 24)) {
                hd.nbricks = (int32_t)nb;
                brick_bytes = sizeof(int32_t) * (size_t)nb + brick_flag_bytes(hd.nbricks) +
                              sizeof(uint16_t) * (size_t)cap * hd.kff;
            } else {
                hd.nbricks = 0;
            }
        }
    }

    // size check against the context budget
    const bool f32 = d.precision == ESPH_FP32_BULK;
    const size_t F = sizeof(float), G = sizeof(double), I = sizeof(int32_t);
    const size_t P = f32 ? F : G;
    size_t per_set = (size_t)cap * (2 * D * G + 5 * G + I + 8 * P) + (thermal ? (size_t)cap * (3 * G + P) : 0);
    size_t bytes = 2 * per_set + (size_t)cap * (G + 3 * I) +
                   (size_t)(hd.ncell + 1) * I * 2 + I + (size_t)hd.cell_tiles * I +
                   (size_t)nw * (3 * D * G + 3 * G + 2 * I + 9 * P) + (size_t)(hd.ncell + 1) * I +
                   2 * sizeof(SceneDyn) + sizeof(SceneDev) +
                   ((size_t)(cap + kBlock - 1) / kBlock + (size_t)hd.cell_tiles) * sizeof(Tile) +
                   (size_t)cap * I * (hd.kff + hd.kfw + 4) + 2 * (size_t)cap * 4 * F + (f32 ? 2 * (size_t)cap * 4 * I : 0) +
                   (size_t)cap * D * G + brick_bytes;
    if (ctx->budget && ctx->used + bytes > ctx->budget)
        return fail(ctx, ESPH_ERR_BUDGET, "scene needs " + std::to_string(bytes) + " bytes; budget left " +
                                              std::to_string(ctx->budget - ctx->used));
    Arena& a = s->arena;
    bool ok = true;
    auto dal = [&](size_t n) { void* p = a.get(n); ok = ok && p; return p; };
    // set[2 * parity + dpar]: arrays of particle order `parity` (changed by
    // the sort) with per-step copy `dpar` (rewritten every step)
    for (int q = 0; q < 2; ++q) {
        FluidSet S{};
        S.rho = (double*)dal(G * cap);
        S.p = (double*)dal(G * cap);
        S.c = (double*)dal(G * cap);
        S.m = (double*)dal(G * cap);
        S.gap = (double*)dal(G * cap);
        S.id = (int32_t*)dal(I * cap);
        S.kind = io ? (int8_t*)dal(sizeof(int8_t) * cap) : nullptr;
        if (thermal) {
            S.kappa = (double*)dal(G * cap);
            S.cp = (double*)dal(G * cap);
        }
        S.kap = (thermal && D == 3) ? dal(P * cap) : nullptr;
        hd.set[2 * q] = S;
        hd.set[2 * q + 1] = S;
    }
    for (int q = 0; q < 2; ++q) {
        double* x = (double*)dal(G * cap * D);
        double* v = (double*)dal(G * cap * D);
        double* T = thermal ? (double*)dal(G * cap) : nullptr;
        void* A = dal(4 * P * cap);
        void* B = dal(4 * P * cap);
        void* xf = dal(F * (D == 3 ? 4 : 2) * cap);
        void* xq = f32 ? dal(I * (D == 3 ? 4 : 2) * cap) : nullptr;
        for (int par = 0; par < 2; ++par) {
            FluidSet& S = hd.set[2 * par + q];
            S.x = x;
            S.v = v;
            S.T = T;
            S.A = A;
            S.B = B;
            S.xf = xf;
            S.xq = xq;
        }
    }
    hd.rhon = (double*)dal(G * cap);
    hd.key = (int32_t*)dal(I * cap);
    hd.slot = (int32_t*)dal(I * cap);
    hd.tmp = (int32_t*)dal(I * cap);
    // whole scan tiles (the scan reads and writes them in 16-byte words)
    const size_t ncs = (size_t)hd.cell_tiles * kCellTile + 4;
    hd.count = (int32_t*)dal(I * ncs);
    hd.cell_start = (int32_t*)dal(I * ncs);
    hd.cell_local = (int32_t*)dal(I * ncs);
    hd.bsum = (int32_t*)dal(I * hd.cell_tiles);
    hd.vnbr = (int32_t*)dal(I * (size_t)cap * hd.kff);
    hd.nvnb = (int32_t*)dal(I * (size_t)cap);
    hd.xref = (double*)dal(G * (size_t)cap * D);
    hd.wnbr = (int32_t*)dal(I * (size_t)cap * std::max(hd.kfw, 1));
    hd.nnbw = (int32_t*)dal(I * (size_t)cap);
    hd.ovf = (int32_t*)dal(I * (size_t)cap);
    if (io) {
        hd.spawn_src = (int32_t*)dal(I * (size_t)cap);
        hd.spawn_id = (int32_t*)dal(I * (size_t)cap);
        hd.spawn_key = (int32_t*)dal(I * (size_t)cap);
        hd.spawn_slot = (int32_t*)dal(I * (size_t)cap);
        hd.spawn_order = (int32_t*)dal(I * (size_t)cap);
        hd.leave_id = (int32_t*)dal(I * (size_t)cap);
    }
    // walls, binned on the host once (stable by upload index)
    hd.wx = (double*)dal(G * std::max(nw, 1) * D);
    hd.wn = (double*)dal(G * std::max(nw, 1) * D);
    hd.wv = (double*)dal(G * std::max(nw, 1) * D);
    hd.wgap = (double*)dal(G * std::max(nw, 1));
    hd.wT = (double*)dal(G * std::max(nw, 1));
    hd.wkappa = (double*)dal(G * std::max(nw, 1));
    hd.wkind = (int32_t*)dal(I * std::max(nw, 1));
    hd.wid = (int32_t*)dal(I * std::max(nw, 1));
    hd.wcell_start = (int32_t*)dal(I * (hd.ncell + 1));
    hd.wnear = (uint8_t*)dal((size_t)hd.ncell);
    {
        const size_t P = f32 ? F : G;
        hd.wA = dal(4 * P * std::max(nw, 1));
        hd.wB = dal(4 * P * std::max(nw, 1));
        hd.wkap = dal(P * std::max(nw, 1));
        hd.wxf = dal(F * (D == 3 ? 4 : 2) * std::max(nw, 1));
    }
    if (hd.nbricks > 0) {
        hd.blist = (int32_t*)dal(I * (size_t)hd.nbricks);
        hd.bflag = (uint8_t*)dal(brick_flag_bytes(hd.nbricks));
        hd.vslot = (uint16_t*)dal(sizeof(uint16_t) * (size_t)cap * hd.kff);
        if (eos_walk_ok(s, f)) hd.flags |= FLAG_BRICK;
    }
    hd.dyn = (SceneDyn*)dal(2 * sizeof(SceneDyn));
    hd.snap = hd.dyn + 1;
    s->dev = (SceneDev*)dal(sizeof(SceneDev));
    s->nptiles = (cap + kBlock - 1) / kBlock;
    s->nctiles = hd.cell_tiles;
    s->ptiles = (Tile*)dal(sizeof(Tile) * s->nptiles);
    s->ctiles = (Tile*)dal(sizeof(Tile) * s->nctiles);
    s->otiles = (Tile*)dal(sizeof(Tile) * kOvfTiles);
    if (!ok) {
        a.release();
        return fail(ctx, ESPH_ERR_HIP, "scene allocation failed (hipMalloc)");
    }
    ctx->used += a.total;
    hipStream_t st = ctx->stream;
    // the sort expects zero cell counts (its scan clears them after use),
    // and no brick marks
    HIPCHK(ctx, hipMemsetAsync(hd.count, 0, I * ((size_t)hd.cell_tiles * kCellTile + 4), st));
    if (hd.bflag) HIPCHK(ctx, hipMemsetAsync(hd.bflag, 0, brick_flag_bytes(hd.nbricks), st));
    // wall binning and upload
    {
        const WallHost& w = s->walls;
        std::vector<int> key(nw), order(nw);
        for (int i = 0; i < nw; ++i) key[i] = cell_of_host(hd, D, &w.x[(size_t)i * D]);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a1, int b1) { return key[a1] < key[b1]; });
        std::vector<double> x(nw * D), nn(nw * D), v(nw * D), gap(nw), T(nw), kap(nw);
        std::vector<int32_t> kind(nw), wid(nw), wcs(hd.ncell + 1, 0);
        for (int p = 0; p < nw; ++p) {
            const int i = order[p];
            for (int k = 0; k < D; ++k) {
                x[p * D + k] = w.x[i * D + k];
                nn[p * D + k] = w.n_[i * D + k];
                v[p * D + k] = w.v[i * D + k];
            }
            gap[p] = w.gap[i]; T[p] = w.T[i]; kap[p] = w.kappa[i]; kind[p] = w.kind[i]; wid[p] = i;
        }
        for (int i = 0; i < nw; ++i) wcs[key[i] + 1]++;
        for (int c = 0; c < hd.ncell; ++c) wcs[c + 1] += wcs[c];
        s->wall_stage = WallStage{x, nn, v, gap, T, kap, kind, wid, wcs};
        // cells with a wall particle within the search stencil (SceneDev::wnear):
        // the cells holding walls, dilated by stencil[k] along each axis
        {
            std::vector<uint8_t>& nr = s->wall_stage.near;
            nr.assign((size_t)hd.ncell, 0);
            for (int i = 0; i < nw; ++i) nr[key[i]] = 1;
            std::vector<int32_t> line;
            const int64_t stride[3] = {1, hd.nc[0], (int64_t)hd.nc[0] * hd.nc[1]};
            for (int k = 0; k < D; ++k) {
                const int n = hd.nc[k], m = hd.stencil[k];
                line.assign((size_t)n + 1, 0);
                for (int64_t base = 0; base < hd.ncell; ++base) {
                    // one line along axis k per base cell whose coordinate k is 0
                    if ((base / stride[k]) % n != 0) continue;
                    for (int q = 0; q < n; ++q) line[q + 1] = line[q] + nr[base + q * stride[k]];
                    for (int q = 0; q < n; ++q) {
                        const int a = q - m > 0 ? q - m : 0, b = q + m < n - 1 ? q + m : n - 1;
                        nr[base + q * stride[k]] = line[b + 1] - line[a] > 0 ? 1 : 0;
                    }
                }
            }
        }
        WallStage& ws = s->wall_stage;
        hipError_t e = hipSuccess;
        auto up = [&](void* dst, const void* src, size_t n) {
            if (e == hipSuccess && n) e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, st);
        };
        up(hd.wx, ws.x.data(), G * nw * D);
        up(hd.wn, ws.n_.data(), G * nw * D);
        up(hd.wv, ws.v.data(), G * nw * D);
        up(hd.wgap, ws.gap.data(), G * nw);
        up(hd.wT, ws.T.data(), G * nw);
        up(hd.wkappa, ws.kappa.data(), G * nw);
        up(hd.wkind, ws.kind.data(), I * nw);
        up(hd.wid, ws.wid.data(), I * nw);
        up(hd.wcell_start, ws.wcs.data(), I * (hd.ncell + 1));
        up(hd.wnear, ws.near.data(), (size_t)hd.ncell);
        // grid-relative float positions (the neighbour-list prefilter)
        {
            const int W = D == 3 ? 4 : 2;
            ws.fX.assign((size_t)W * std::max(nw, 1), 0.0f);
            for (int p = 0; p < nw; ++p)
                for (int k = 0; k < D; ++k) ws.fX[(size_t)W * p + k] = (float)(ws.x[(size_t)p * D + k] - hd.lo[k]);
            up(hd.wxf, ws.fX.data(), F * ws.fX.size());
        }
        // packed wall records (see SceneDev::wA / wB / wkap)
        {
            std::vector<double> A(4 * (size_t)std::max(nw, 1), 0.0), B(4 * (size_t)std::max(nw, 1), 0.0);
            std::vector<double> K(std::max(nw, 1), 0.0);
            // float records hold T - t0 (see the fluid records)
            const double Toff = f32 ? s->d.t_0 : 0.0;
            for (int p = 0; p < nw; ++p) {
                double* a = &A[4 * (size_t)p];
                double* b = &B[4 * (size_t)p];
                if (D == 3) {
                    a[0] = ws.n_[3 * p]; a[1] = ws.n_[3 * p + 1]; a[2] = ws.n_[3 * p + 2]; a[3] = ws.gap[p];
                    b[0] = ws.v[3 * p]; b[1] = ws.v[3 * p + 1]; b[2] = ws.v[3 * p + 2]; b[3] = ws.T[p] - Toff;
                    K[p] = ws.kappa[p];
                } else {
                    a[0] = ws.n_[2 * p]; a[1] = ws.n_[2 * p + 1]; a[2] = ws.gap[p]; a[3] = ws.T[p] - Toff;
                    b[0] = ws.v[2 * p]; b[1] = ws.v[2 * p + 1]; b[2] = ws.kappa[p]; b[3] = 0.0;
                }
            }
            if (f32) {
                ws.fA.assign(A.begin(), A.end());
                ws.fB.assign(B.begin(), B.end());
                ws.fK.assign(K.begin(), K.end());
                up(hd.wA, ws.fA.data(), F * ws.fA.size());
                up(hd.wB, ws.fB.data(), F * ws.fB.size());
                up(hd.wkap, ws.fK.data(), F * ws.fK.size());
            } else {
                ws.dA = A;
                ws.dB = B;
                ws.dK = K;
                up(hd.wA, ws.dA.data(), G * ws.dA.size());
                up(hd.wB, ws.dB.data(), G * ws.dB.size());
                up(hd.wkap, ws.dK.data(), G * ws.dK.size());
            }
        }
        std::vector<Tile> pt(s->nptiles), ct(s->nctiles);
        for (int t = 0; t < s->nptiles; ++t) pt[t] = Tile{0, t * kBlock};
        for (int t = 0; t < s->nctiles; ++t) ct[t] = Tile{0, t * kCellTile};
        std::vector<Tile> ot(kOvfTiles);
        for (int t = 0; t < kOvfTiles; ++t) ot[t] = Tile{0, t * kBlock};
        s->tile_stage_p = pt;
        s->tile_stage_c = ct;
        s->tile_stage_o = ot;
        up(s->ptiles, s->tile_stage_p.data(), sizeof(Tile) * pt.size());
        up(s->ctiles, s->tile_stage_c.data(), sizeof(Tile) * ct.size());
        up(s->otiles, s->tile_stage_o.data(), sizeof(Tile) * ot.size());
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(ctx, "wall upload", e);
    }
    s->h = hd;
    s->allocated = true;
    return ESPH_OK;
}

extern "C" int esph_scene_set_fluid(esph_scene* s, const esph_fluid_arrays* f) {
    if (!s || !f) return fail(s ? s->ctx : nullptr, ESPH_ERR_ARG, "set_fluid: null argument");
    esph_ctx* ctx = s->ctx;
    const int D = s->d.dim;
    const bool thermal = s->d.solver == ESPH_SOLVER_FLUID_THERMOSTATIC_WALL ||
                         s->d.solver == ESPH_SOLVER_FLUID_COMPULSIVE_THERMOSTATIC_WALL;
    if (f->n < 0 || (f->n > 0 && (!f->x || !f->v || !f->rho || !f->p || !f->c || !f->mass || !f->gap)))
        return fail(ctx, ESPH_ERR_ARG, "set_fluid: missing arrays");
    if (thermal && f->n > 0 && (!f->t || !f->kappa || !f->cp))
        return fail(ctx, ESPH_ERR_ARG, "set_fluid: thermal solvers need t, kappa and cp");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (!s->allocated) {
        int rc = allocate_scene(s, f);
        if (rc) return rc;
        HIPCHK(ctx, hipMemcpy(s->dev, &s->h, sizeof(SceneDev), hipMemcpyHostToDevice));
    } else if (((s->h.flags & FLAG_EOS_WALK) != 0) != eos_walk_ok(s, f)) {
        // a re-upload that changes the gaps' uniformity changes the force walk
        if (s->batch_refs > 0)
            return fail(ctx, ESPH_ERR_STATE, "set_fluid: the new gaps change the scene's force walk while it is in a batch");
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        s->h.flags ^= FLAG_EOS_WALK;
        // brick walks follow (their buffers exist when the scene allowed them)
        if (s->h.nbricks > 0) s->h.flags = (s->h.flags & ~FLAG_BRICK) | ((s->h.flags & FLAG_EOS_WALK) ? FLAG_BRICK : 0);
        HIPCHK(ctx, hipMemcpy(s->dev, &s->h, sizeof(SceneDev), hipMemcpyHostToDevice));
        s->gplan.release();
    }
    SceneDev& hd = s->h;
    const bool io = (hd.flags & FLAG_IO) != 0;
    const int64_t nb = io ? s->n_inlet + s->n_outlet : 0;
    if (f->n + nb > hd.cap) return fail(ctx, ESPH_ERR_CAPACITY, "set_fluid: more particles than the scene capacity");
    const int nf = (int)f->n;
    const int n = (int)(f->n + nb);
    hipStream_t st = ctx->stream;
    HIPCHK(ctx, hipStreamSynchronize(st));
    const size_t G = sizeof(double);
    // the upload starts at parity 0, per-step copy 0: the arrays of order 0
    // take the constant fields; the positions, velocities and temperatures go
    // to copy 1 (the step outputs' place) and the upload's sort gathers them
    FluidSet& S = hd.set[0];
    const FluidSet& N = hd.set[1];
    // fluid, then inlet and outlet buffers; ids are sequence numbers per group
    std::vector<int32_t> ids(n);
    std::vector<int8_t> kinds(n, (int8_t)KIND_FLUID);
    for (int i = 0; i < nf; ++i) ids[i] = i;
    for (int64_t i = 0; i < s->n_inlet; ++i) { ids[nf + i] = (int32_t)i; kinds[nf + i] = KIND_INLET; }
    for (int64_t i = 0; i < s->n_outlet; ++i) {
        ids[nf + s->n_inlet + i] = (int32_t)i;
        kinds[nf + s->n_inlet + i] = KIND_OUTLET;
    }
    hipError_t e = hipSuccess;
    auto up = [&](void* dst, const void* src, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    };
    auto up2 = [&](double* dst, const double* a, int64_t na, const std::vector<double>& b, int w) {
        up(dst, a, G * na * w);
        if (nb) up(dst + na * w, b.data(), G * nb * w);
    };
    up2(N.x, f->x, nf, s->bx, D);
    up2(N.v, f->v, nf, s->bv, D);
    up2(hd.rhon, f->rho, nf, s->brho, 1);
    up2(S.p, f->p, nf, s->bp, 1);
    up2(S.c, f->c, nf, s->bc, 1);
    up2(S.m, f->mass, nf, s->bm, 1);
    up2(S.gap, f->gap, nf, s->bgap, 1);
    up(S.id, ids.data(), sizeof(int32_t) * n);
    if (io) up(S.kind, kinds.data(), sizeof(int8_t) * n);
    if (thermal) {
        up(N.T, f->t, G * n);
        up(S.kappa, f->kappa, G * n);
        up(S.cp, f->cp, G * n);
    }
    SceneDyn dyn{};
    dyn.step = 0;
    dyn.parity = 0;
    dyn.n_live = n;
    dyn.next_fluid_id = nf;
    dyn.next_outlet_id = (int32_t)s->n_outlet;
    up(hd.dyn, &dyn, sizeof(SceneDyn));
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx, "set_fluid upload", e);
    SortPlan plan{s->dev, 1, s->ptiles, s->nptiles, s->ctiles, s->nctiles};
    e = launch_sort(plan, 1, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx, "set_fluid sort", e);
    s->step = 0;
    s->n_upload = n;
    s->fluid_ready = true;
    return ESPH_OK;
}

static Bucket scene_bucket(esph_scene* s) {
    const bool th = (s->h.flags & FLAG_THERMAL) != 0;
    const bool eos = (s->h.flags & FLAG_EOS_WALK) != 0;
    const bool br = (s->h.flags & FLAG_BRICK) != 0;
    Bucket b{s->d.dim, s->d.kernel, s->d.precision, s->dev, s->ptiles, s->nptiles, s->otiles, 1, th, !th && !eos};
    b.has_eos = eos && !br;
    b.has_brick = br;
    b.has_tile = !br;
    b.btiles = s->ptiles;  // its first tile is (scene 0, particle 0)
    b.nbtiles = br ? 1 : 0;
    return b;
}

static hipEvent_t timing_event(esph_scene* s) {
    if (s->ev_next < s->timing_events.size()) return s->timing_events[s->ev_next++];
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    s->timing_events.push_back(e);
    s->ev_next++;
    return e;
}

extern "C" int esph_scene_step(esph_scene* s, int64_t nsteps) {
    if (!s || nsteps < 0) return fail(s ? s->ctx : nullptr, ESPH_ERR_ARG, "step: bad argument");
    if (!s->fluid_ready) return fail(s->ctx, ESPH_ERR_STATE, "step before set_fluid");
    esph_ctx* ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (s->use_graph) {
        if (s->gplan.nmembers == 0) {
            HIPCHK(ctx, hipStreamSynchronize(st));
            s->gplan.use_graph = true;
            int rc = s->gplan.build(ctx, std::vector<esph_scene*>{s});
            if (rc) return rc;
        }
        int rc = s->gplan.step(ctx, nsteps);
        if (rc) return rc;
        s->step += nsteps;
        return ESPH_OK;
    }
    Bucket b = scene_bucket(s);
    apply_depth(ctx, b);
    const SortPlan plan{s->dev, 1, s->ptiles, s->nptiles, s->ctiles, s->nctiles};
    const int64_t K = s->d.density_reinitialized_interval;
    for (int64_t t = 0; t < nsteps; ++t) {
        const bool shep = K > 0 && (s->step + 1) % K == 0;
        hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
        if (s->timing) {
            for (int q = 0; q < 5; ++q) ev[q] = timing_event(s);
            (void)hipEventRecord(ev[0], st);
        }
        hipError_t e = launch_continuity(b, st);
        if (e == hipSuccess && b.depth == 0) e = tune_depth(ctx, b, st);  // first step of an untuned class
        if (s->timing) (void)hipEventRecord(ev[1], st);
        if (e == hipSuccess) e = launch_forces(b, st);
        if (s->timing) (void)hipEventRecord(ev[2], st);
        if (e == hipSuccess && shep) e = launch_shepard(b, st);
        if (s->timing) (void)hipEventRecord(ev[3], st);
        if (e == hipSuccess) e = launch_sort(plan, 0, st);
        if (s->timing) {
            (void)hipEventRecord(ev[4], st);
            s->timed.push_back(TimedStep{{ev[0], ev[1], ev[2], ev[3], ev[4]}, shep});
        }
        if (e != hipSuccess) return hip_fail(ctx, "step launch", e);
        s->step += 1;
    }
    return ESPH_OK;
}

extern "C" int esph_scene_sync(esph_scene* s) {
    if (!s) return ESPH_ERR_ARG;
    esph_ctx* ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (!s->fluid_ready) return ESPH_OK;
    SceneDyn dyn;
    HIPCHK(ctx, hipMemcpy(&dyn, s->h.dyn, sizeof(SceneDyn), hipMemcpyDeviceToHost));
    if (dyn.err & 1) return fail(ctx, ESPH_ERR_NAN, "a non-finite fluid position was met");
    if (dyn.err & 2)
        return fail(ctx, ESPH_ERR_CAPACITY, "inlet copies exceeded the scene capacity (fluid_capacity) and were dropped");
    return ESPH_OK;
}

extern "C" int esph_scene_step_count(const esph_scene* s, int64_t* step) {
    if (!s || !step) return ESPH_ERR_ARG;
    *step = s->step;
    return ESPH_OK;
}

static int read_dyn(esph_scene* s, SceneDyn* dyn) {
    esph_ctx* ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemcpy(dyn, s->h.dyn, sizeof(SceneDyn), hipMemcpyDeviceToHost));
    return ESPH_OK;
}

extern "C" int esph_scene_fluid_count(esph_scene* s, int64_t* n) {
    if (!s || !n) return ESPH_ERR_ARG;
    if (!s->fluid_ready) return fail(s->ctx, ESPH_ERR_STATE, "no fluid uploaded");
    if (s->h.flags & FLAG_IO) return esph_scene_group_count(s, ESPH_GROUP_FLUID, n);
    SceneDyn dyn;
    int rc = read_dyn(s, &dyn);
    if (rc) return rc;
    *n = dyn.n_live;
    return ESPH_OK;
}

// Live particles of one group (fluid, inlet or outlet buffer) in reference
// order: ascending id within the group.  Scenes without buffers hold fluid
// only.
static int group_order(esph_scene* s, int32_t group, SceneDyn& dyn, std::vector<int32_t>& order,
                       std::vector<int32_t>& ids) {
    int rc = read_dyn(s, &dyn);
    if (rc) return rc;
    esph_ctx* ctx = s->ctx;
    const int n = dyn.n_live;
    const FluidSet& S = s->h.set[cur_set(dyn)];
    std::vector<int32_t> all(n);
    std::vector<int8_t> kinds;
    if (n) HIPCHK(ctx, hipMemcpy(all.data(), S.id, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (S.kind && n) {
        kinds.resize(n);
        HIPCHK(ctx, hipMemcpy(kinds.data(), S.kind, sizeof(int8_t) * n, hipMemcpyDeviceToHost));
    }
    order.clear();
    for (int q = 0; q < n; ++q)
        if ((kinds.empty() ? (int)KIND_FLUID : (int)kinds[q]) == group) order.push_back(q);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return all[a] < all[b]; });
    ids.resize(order.size());
    for (size_t q = 0; q < order.size(); ++q) ids[q] = all[order[q]];
    return ESPH_OK;
}

static int get_group(esph_scene* s, int32_t group, int32_t field, double* out, int64_t capacity, int64_t* written) {
    if (!s || !out) return fail(s ? s->ctx : nullptr, ESPH_ERR_ARG, "get_fluid: null argument");
    if (!s->fluid_ready) return fail(s->ctx, ESPH_ERR_STATE, "no fluid uploaded");
    if (group != ESPH_GROUP_FLUID && group != ESPH_GROUP_INLET && group != ESPH_GROUP_OUTLET)
        return fail(s->ctx, ESPH_ERR_ARG, "get_group: bad group");
    esph_ctx* ctx = s->ctx;
    const int D = s->d.dim;
    const bool thermal = s->h.flags & FLAG_THERMAL;
    SceneDyn dyn;
    std::vector<int32_t> order, ids;
    int rc = group_order(s, group, dyn, order, ids);
    if (rc) return rc;
    const int n = (int)order.size();
    const FluidSet& S = s->h.set[cur_set(dyn)];
    const double* src = nullptr;
    int w = 1;
    switch (field) {
        case ESPH_FIELD_X: src = S.x; w = D; break;
        case ESPH_FIELD_V: src = S.v; w = D; break;
        case ESPH_FIELD_RHO: src = S.rho; break;
        case ESPH_FIELD_P: src = S.p; break;
        case ESPH_FIELD_C: src = S.c; break;
        case ESPH_FIELD_MASS: src = S.m; break;
        case ESPH_FIELD_GAP: src = S.gap; break;
        case ESPH_FIELD_T: src = thermal ? S.T : nullptr; break;
        case ESPH_FIELD_KAPPA: src = thermal ? S.kappa : nullptr; break;
        case ESPH_FIELD_CP: src = thermal ? S.cp : nullptr; break;
        case ESPH_FIELD_ID: break;
        default: return fail(ctx, ESPH_ERR_ARG, "get_fluid: bad field");
    }
    if (field != ESPH_FIELD_ID && !src) return fail(ctx, ESPH_ERR_ARG, "get_fluid: field not present for this solver");
    if (written) *written = (int64_t)n * w;
    if (capacity < (int64_t)n * w) return fail(ctx, ESPH_ERR_CAPACITY, "get_fluid: output too small");
    if (field == ESPH_FIELD_ID) {
        for (int q = 0; q < n; ++q) out[q] = (double)ids[q];
        return ESPH_OK;
    }
    const int nl = dyn.n_live;
    std::vector<double> buf((size_t)nl * w);
    if (nl) HIPCHK(ctx, hipMemcpy(buf.data(), src, sizeof(double) * nl * w, hipMemcpyDeviceToHost));
    for (int q = 0; q < n; ++q)
        for (int k = 0; k < w; ++k) out[(size_t)q * w + k] = buf[(size_t)order[q] * w + k];
    return ESPH_OK;
}

extern "C" int esph_scene_get_fluid(esph_scene* s, int32_t field, double* out, int64_t capacity, int64_t* written) {
    return get_group(s, ESPH_GROUP_FLUID, field, out, capacity, written);
}

extern "C" int esph_scene_get_group(esph_scene* s, int32_t group, int32_t field, double* out, int64_t capacity,
                                    int64_t* written) {
    return get_group(s, group, field, out, capacity, written);
}

extern "C" int esph_scene_group_count(esph_scene* s, int32_t group, int64_t* n) {
    if (!s || !n) return ESPH_ERR_ARG;
    if (!s->fluid_ready) return fail(s->ctx, ESPH_ERR_STATE, "no fluid uploaded");
    if (group != ESPH_GROUP_FLUID && group != ESPH_GROUP_INLET && group != ESPH_GROUP_OUTLET)
        return fail(s->ctx, ESPH_ERR_ARG, "group_count: bad group");
    SceneDyn dyn;
    std::vector<int32_t> order, ids;
    int rc = group_order(s, group, dyn, order, ids);
    if (rc) return rc;
    *n = (int64_t)order.size();
    return ESPH_OK;
}

// Pair export through k_pairs (cell scan, binary64 rule) or, with walk, through
// k_walk_pairs (the pair passes' own decisions): count, host scan, write,
// remap, sort.
static int export_pairs(esph_scene* s, int pairset, std::vector<int64_t>& pairs_out, bool walk = false) {
    esph_ctx* ctx = s->ctx;
    SceneDyn dyn;
    int rc = read_dyn(s, &dyn);
    if (rc) return rc;
    const int n = dyn.n_live;
    hipStream_t st = ctx->stream;
    Arena a;
    int32_t* counts = (int32_t*)a.get(sizeof(int32_t) * std::max(n, 1));
    int64_t* offs = (int64_t*)a.get(sizeof(int64_t) * std::max(n, 1));
    if (!counts || !offs) { a.release(); return fail(ctx, ESPH_ERR_HIP, "export: allocation failed"); }
    const Bucket b = scene_bucket(s);
    auto launch = [&](int32_t* c, const int64_t* o, int32_t* out) {
        return walk ? launch_walk_pairs(b, pairset, c, o, out, st) : launch_pairs(b, pairset, c, o, out, st);
    };
    hipError_t e = launch(counts, nullptr, nullptr);
    std::vector<int32_t> hc(n);
    if (e == hipSuccess && n) e = hipMemcpyAsync(hc.data(), counts, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { a.release(); return hip_fail(ctx, "export count", e); }
    std::vector<int64_t> ho(n);
    int64_t m = 0;
    for (int i = 0; i < n; ++i) { ho[i] = m; m += hc[i]; }
    int32_t* dout = (int32_t*)a.get(sizeof(int32_t) * 2 * std::max<int64_t>(m, 1));
    if (!dout) { a.release(); return fail(ctx, ESPH_ERR_HIP, "export: allocation failed"); }
    std::vector<int32_t> hp(2 * m);
    if (n) e = hipMemcpyAsync(offs, ho.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch(nullptr, offs, dout);
    if (e == hipSuccess && m) e = hipMemcpyAsync(hp.data(), dout, sizeof(int32_t) * 2 * m, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    a.release();
    if (e != hipSuccess) return hip_fail(ctx, "export write", e);
    // fluid id -> reference index (rank among the live fluid ids)
    std::vector<int32_t> order, sorted_ids;
    rc = group_order(s, ESPH_GROUP_FLUID, dyn, order, sorted_ids);
    if (rc) return rc;
    auto rank = [&](int32_t id) -> int64_t {
        return std::lower_bound(sorted_ids.begin(), sorted_ids.end(), id) - sorted_ids.begin();
    };
    pairs_out.resize(m);
    for (int64_t q = 0; q < m; ++q) {
        const int64_t i = rank(hp[2 * q]);
        const int64_t j = pairset == ESPH_PAIRS_FLUID_FLUID ? rank(hp[2 * q + 1]) : hp[2 * q + 1];
        pairs_out[q] = (i << 32) | j;
    }
    std::sort(pairs_out.begin(), pairs_out.end());
    return ESPH_OK;
}

extern "C" int esph_scene_neighbours(esph_scene* s, int32_t pairset, int32_t* out_pairs, int64_t capacity, int64_t* m) {
    if (!s || !m || (pairset != ESPH_PAIRS_FLUID_FLUID && pairset != ESPH_PAIRS_FLUID_WALL))
        return fail(s ? s->ctx : nullptr, ESPH_ERR_ARG, "neighbours: bad argument");
    if (!s->fluid_ready) return fail(s->ctx, ESPH_ERR_STATE, "no fluid uploaded");
    std::vector<int64_t> pairs;
    int rc = export_pairs(s, pairset, pairs);
    if (rc) return rc;
    *m = (int64_t)pairs.size();
    if (!out_pairs && capacity == 0) return ESPH_OK;
    if ((int64_t)pairs.size() > capacity) return fail(s->ctx, ESPH_ERR_CAPACITY, "neighbours: output too small");
    for (size_t q = 0; q < pairs.size(); ++q) {
        out_pairs[2 * q] = (int32_t)(pairs[q] >> 32);
        out_pairs[2 * q + 1] = (int32_t)(pairs[q] & 0xffffffff);
    }
    return ESPH_OK;
}

extern "C" int esph_scene_walk_pairs(esph_scene* s, int32_t pairset, int32_t* out_pairs, int64_t capacity,
                                     int64_t* m) {
    if (!s || !m || (pairset != ESPH_PAIRS_FLUID_FLUID && pairset != ESPH_PAIRS_FLUID_WALL))
        return fail(s ? s->ctx : nullptr, ESPH_ERR_ARG, "walk_pairs: bad argument");
    if (!s->fluid_ready) return fail(s->ctx, ESPH_ERR_STATE, "no fluid uploaded");
    if (s->h.flags & FLAG_IO) return fail(s->ctx, ESPH_ERR_ARG, "walk_pairs: not for the inlet/outlet solver");
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    std::vector<int64_t> pairs;
    int rc = export_pairs(s, pairset, pairs, true);
    if (rc) return rc;
    *m = (int64_t)pairs.size();
    if (!out_pairs && capacity == 0) return ESPH_OK;
    if ((int64_t)pairs.size() > capacity) return fail(s->ctx, ESPH_ERR_CAPACITY, "walk_pairs: output too small");
    for (size_t q = 0; q < pairs.size(); ++q) {
        out_pairs[2 * q] = (int32_t)(pairs[q] >> 32);
        out_pairs[2 * q + 1] = (int32_t)(pairs[q] & 0xffffffff);
    }
    return ESPH_OK;
}

extern "C" int esph_scene_probe(esph_scene* s, int32_t field, int64_t npoints, const double* points,
                                double none_value, double* out) {
    if (!s || npoints < 0 || (npoints > 0 && (!points || !out)))
        return fail(s ? s->ctx : nullptr, ESPH_ERR_ARG, "probe: bad argument");
    if (!s->fluid_ready) return fail(s->ctx, ESPH_ERR_STATE, "no fluid uploaded");
    const bool thermal = s->h.flags & FLAG_THERMAL;
    if (field != ESPH_FIELD_V && field != ESPH_FIELD_RHO && field != ESPH_FIELD_P && field != ESPH_FIELD_C &&
        !(field == ESPH_FIELD_T && thermal))
        return fail(s->ctx, ESPH_ERR_ARG, "probe: field must be V, RHO, P, C or (thermal solvers) T");
    if (npoints == 0) return ESPH_OK;
    esph_ctx* ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int D = s->d.dim;
    const int w = field == ESPH_FIELD_V ? D : 1;
    hipStream_t st = ctx->stream;
    Arena a;
    double* dp = (double*)a.get(sizeof(double) * (size_t)npoints * D);
    double* dout = (double*)a.get(sizeof(double) * (size_t)npoints * w);
    if (!dp || !dout) {
        a.release();
        return fail(ctx, ESPH_ERR_HIP, "probe: allocation failed");
    }
    hipError_t e = hipMemcpyAsync(dp, points, sizeof(double) * (size_t)npoints * D, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_probe(scene_bucket(s), npoints, dp, field, none_value, dout, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, sizeof(double) * (size_t)npoints * w, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    a.release();
    if (e != hipSuccess) return hip_fail(ctx, "probe", e);
    return ESPH_OK;
}

extern "C" int esph_scene_probe_noalloc(esph_scene* s, int32_t field, int64_t npoints, const double* points,
                                        double none_value, double* out) {
    if (!s || npoints < 0 || (npoints > 0 && (!points || !out)))
        return fail(s ? s->ctx : nullptr, ESPH_ERR_ARG, "probe_noalloc: bad argument");
    if (!s->fluid_ready) return fail(s->ctx, ESPH_ERR_STATE, "no fluid uploaded");
    const bool thermal = s->h.flags & FLAG_THERMAL;
    if (field != ESPH_FIELD_V && field != ESPH_FIELD_RHO && field != ESPH_FIELD_P && field != ESPH_FIELD_C &&
        !(field == ESPH_FIELD_T && thermal))
        return fail(s->ctx, ESPH_ERR_ARG, "probe_noalloc: field must be V, RHO, P, C or (thermal solvers) T");
    if (npoints > kProbeNoallocPoints)
        return fail(s->ctx, ESPH_ERR_CAPACITY,
                    "probe_noalloc: at most " + std::to_string(kProbeNoallocPoints) + " points per call");
    esph_ctx* ctx = s->ctx;
    if (!ctx->probe_scratch) return hip_fail(ctx, "probe_noalloc: staging address", ctx->probe_scratch_err);
    if (npoints == 0) return ESPH_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int D = s->d.dim;
    const int w = field == ESPH_FIELD_V ? D : 1;
    hipStream_t st = ctx->stream;
    double* dp = ctx->probe_scratch;                              // points: npoints * D
    double* dout = ctx->probe_scratch + kProbeNoallocPoints * 3;  // values: npoints * w
    hipError_t e = hipMemcpyAsync(dp, points, sizeof(double) * (size_t)npoints * D, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_probe(scene_bucket(s), npoints, dp, field, none_value, dout, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, sizeof(double) * (size_t)npoints * w, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx, "probe_noalloc", e);
    return ESPH_OK;
}

extern "C" int esph_neighbor_list(esph_ctx* ctx, int32_t dim, double cutoff, int64_t nx, const double* x, int64_t ny,
                                  const double* y, int32_t* out_pairs, int64_t capacity, int64_t* m) {
    if (!ctx || !m || (dim != 2 && dim != 3) || !(cutoff > 0) || nx < 0 || ny < 0 || (nx > 0 && !x) ||
        (ny > 0 && !y))
        return fail(ctx, ESPH_ERR_ARG, "neighbor_list: bad argument");
    esph_scene_desc d{};
    d.dim = dim;
    d.kernel = ESPH_CUBIC_SPLINE;
    d.influence_radius = cutoff;
    d.solver = ESPH_SOLVER_FLUID_WALL;
    d.precision = ESPH_FP64;
    d.rho_0 = 1.0;
    d.c_0 = 1.0;
    d.gamma = 1;
    d.dt = 1.0;
    esph_scene* s = nullptr;
    int rc = esph_scene_create(ctx, &d, &s);
    if (rc) return rc;
    std::vector<double> ones(std::max<int64_t>(std::max(nx, ny), 1), 1.0);
    std::vector<double> zeros(std::max<int64_t>(std::max(nx, ny), 1) * dim, 0.0);
    if (ny > 0) {
        esph_wall_arrays w{};
        w.kind = ESPH_WALL_COMPULSIVE;
        w.n = ny;
        w.x = y;
        w.normal = zeros.data();
        w.gap = ones.data();
        rc = esph_scene_set_walls(s, 1, &w);
    }
    if (!rc) {
        esph_fluid_arrays f{};
        f.n = nx;
        f.x = x;
        f.v = zeros.data();
        f.rho = ones.data();
        f.p = zeros.data();
        f.c = ones.data();
        f.mass = ones.data();
        f.gap = ones.data();
        rc = esph_scene_set_fluid(s, &f);
    }
    if (!rc) rc = esph_scene_neighbours(s, ny > 0 ? ESPH_PAIRS_FLUID_WALL : ESPH_PAIRS_FLUID_FLUID, out_pairs,
                                        capacity, m);
    esph_scene_destroy(s);
    return rc;
}

extern "C" int esph_scene_enable_timing(esph_scene* s, int32_t enable) {
    if (!s) return ESPH_ERR_ARG;
    s->timing = enable != 0;
    return ESPH_OK;
}

extern "C" int esph_scene_phase_times(esph_scene* s, double* ms5, int64_t* steps) {
    if (!s || !ms5) return ESPH_ERR_ARG;
    esph_ctx* ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    double acc[5] = {0, 0, 0, 0, 0};
    int64_t nshep = 0;
    for (const TimedStep& t : s->timed) {
        float a;
        (void)hipEventElapsedTime(&a, t.ev[0], t.ev[1]); acc[1] += a;
        (void)hipEventElapsedTime(&a, t.ev[1], t.ev[2]); acc[2] += a;
        (void)hipEventElapsedTime(&a, t.ev[2], t.ev[3]); if (t.shepard) { acc[3] += a; nshep++; }
        (void)hipEventElapsedTime(&a, t.ev[3], t.ev[4]); acc[0] += a;
    }
    const int64_t n = (int64_t)s->timed.size();
    for (int q = 0; q < 5; ++q) ms5[q] = 0.0;
    if (n) {
        ms5[0] = acc[0] / n;
        ms5[1] = acc[1] / n;
        ms5[2] = acc[2] / n;
        ms5[3] = nshep ? acc[3] / nshep : 0.0;
    }
    if (steps) *steps = n;
    s->timed.clear();
    s->ev_next = 0;
    return ESPH_OK;
}

extern "C" int esph_scene_stats(esph_scene* s, int64_t* out8) {
    if (!s || !out8) return ESPH_ERR_ARG;
    int64_t* out6 = out8;
    if (!s->fluid_ready) return fail(s->ctx, ESPH_ERR_STATE, "no fluid uploaded");
    SceneDyn dyn;
    int rc = read_dyn(s, &dyn);
    if (rc) return rc;
    out6[0] = dyn.overflow;
    out6[1] = s->h.kff;
    out6[2] = s->h.kfw;
    out6[3] = s->h.ncell;
    out6[4] = dyn.outside;
    out6[5] = dyn.err;
    out8[6] = dyn.nsorts;
    out8[7] = s->h.verlet ? s->h.kff : 0;
    return ESPH_OK;
}

extern "C" int esph_scene_brick_stats(esph_scene* s, int64_t* out6) {
    if (!s || !out6) return ESPH_ERR_ARG;
    if (!s->fluid_ready) return fail(s->ctx, ESPH_ERR_STATE, "no fluid uploaded");
    SceneDyn dyn;
    int rc = read_dyn(s, &dyn);
    if (rc) return rc;
    const bool br = (s->h.flags & FLAG_BRICK) != 0;
    out6[0] = br ? 1 : 0;
    for (int k = 0; k < 3; ++k) out6[1 + k] = br ? s->h.bdim[k] : 0;
    out6[4] = br ? dyn.nbrick : 0;
    out6[5] = br ? dyn.brick_spill : 0;
    return ESPH_OK;
}

extern "C" int esph_scene_enable_graph(esph_scene* s, int32_t enable) {
    if (!s) return ESPH_ERR_ARG;
    s->use_graph = enable != 0;
    if (!s->use_graph) {
        (void)hipSetDevice(s->ctx->device);
        (void)hipStreamSynchronize(s->ctx->stream);
        s->gplan.release();
    }
    return ESPH_OK;
}
