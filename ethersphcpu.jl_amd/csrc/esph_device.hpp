// esph_device.hpp — device-side data structures and per-pair math of the
// gfx950 SPH hot path.  Shared by the kernels (esph_kernels.hip) and the
// host runtime (esph_api.hip), which fills SceneDev and uploads it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ethersph.h"

namespace esph {

constexpr int kBlock = 256;          // particles per tile == threads per block
// cells per thread in the scan kernels: 16,384 cells per scan tile.  Each
// tile's block fences its total before the ticket of k_sort_scan1, and that
// fence writes the L2 back, so fewer, larger tiles make a sort cheaper.
// Measured with cubic cells (a quarter of today's count): 1,024 cells per
// tile 0.586 ms per 3D dam-break step, 4,096 0.580, 8,192 0.579 (but the 2D
// ensemble 1.78 ms per batched step against 1.77 at 4,096)
constexpr int kScanItems = 64;
constexpr int kCellTile = kBlock * kScanItems;
constexpr int kOvfTiles = 16;        // tiles per scene of the overflow (cell-scan) kernels
// waves per SIMD the force pass of FLAG_EOS_WALK scenes is compiled for
constexpr int kEosWalkWaves = 4;
// Brick walk (FLAG_BRICK): workgroups of kBrickThreads threads, each taking
// one brick of cells at a time and staging the quantized positions and A
// records of every particle of the brick's neighbour region (the brick
// widened by the stencil on each side) in LDS, at most kBrickCap particles:
// 2 x 16 bytes each, so two workgroups (sixteen waves) fit the 160 KB of LDS
// of a CU.  kBrickRows / kBrickOwn bound the cell rows (fixed y, z) of the
// region and of the brick.
constexpr int kBrickThreads = 512;
constexpr int kBrickCap = 2432;
constexpr int kBrickRows = 128;
constexpr int kBrickOwn = 64;

// SceneDev::flags
enum : int {
    FLAG_THERMAL = 1,        // thermal fluid + ThermalWeaklyCompressibleLiquidModel
    FLAG_WALL_PRESSURE = 2,  // pressureForce!(fluid, wall) in the wall loop
    FLAG_CHECK = 4,          // AABB deletion predicate enabled
    FLAG_IO = 8,             // inlet/outlet solver: buffer particles and migration
    // 3D FP32_BULK scene without buffer particles, heat or varying gaps: the
    // force pass's fluid walk gathers two records per candidate (quantized
    // position with rho - rho0, velocity with mass) and takes p / rho^2 from
    // the equation of state (every fluid particle's p is EOS(rho) once the
    // continuity pass has run)
    FLAG_EOS_WALK = 16,
    // FLAG_EOS_WALK scene whose continuity and force passes walk bricks of
    // cells with their neighbour regions staged in LDS (SceneDev::bdim);
    // candidates are addressed by their LDS slots (SceneDev::vslot)
    FLAG_BRICK = 32,
    // the force pass walks each particle's wall candidates in its own lane
    // (as brick walks and 2D scenes always do); without it the 3D candidates
    // of a wavefront's particles are shared out over all its lanes
    // (wall_walk_shared)
    FLAG_WALL_LANES = 64,
};

// Particle kinds of the inlet/outlet solver (FluidSet::kind).
enum : int8_t { KIND_FLUID = 0, KIND_INLET = 1, KIND_OUTLET = 2 };

// Candidate-list skin (fraction of R) when the descriptor leaves it at 0.
// Every pair pass walks the candidates, so their excess (1 + skin)^D - 1
// costs on every step and a sort with a list rebuild on every skin R / 2 of
// motion.  Swept over 1,000 steps (FP32_BULK, ms per step, sorts): the 1M-
// particle 3D dam break 0.02: 0.427 (63), 0.03: 0.413 (41), 0.04: 0.406
// (31), 0.05: 0.403 (25), 0.06: 0.3995 (20), 0.07: 0.400 (17), 0.08: 0.402
// (15); the 256-scene 2D ensemble, ms per batched step, 0.03: 1.646, 0.045:
// 1.613, 0.06: 1.665.
__host__ __device__ inline double default_skin(int dim) { return dim == 3 ? 0.06 : 0.045; }

// Per-scene state that changes every step and lives in device memory, so a
// captured step graph can be replayed without the host.
struct SceneDyn {
    int64_t step;    // completed steps (the reference's `step` of the last one)
    int32_t parity;  // particle order (static arrays) of the current state: 0 or 1
    int32_t dpar;    // which copy of the per-step arrays holds the current state: 0 or 1
    int32_t n_live;  // fluid particles in the current set
    int32_t err;     // bit 0: a non-finite position was met
    int32_t overflow;  // particle-steps whose neighbour count exceeded the list width
    int32_t novf;      // this step's overflowed particles (listed in SceneDev::ovf)
    int32_t nspawn;    // inlet/outlet: inlet particles entering the fluid this step
    int32_t nleave;    // inlet/outlet: fluid particles entering the outlet this step
    int32_t next_fluid_id;   // next sequence number of the fluid group
    int32_t next_outlet_id;  // next sequence number of the outlet group
    int32_t outside;   // particles more than one cell outside the grid at the last sort
    int32_t rebuild;   // 1: a sort just ran (new particle order: lists rebuilt, records written whole)
    int32_t sorted;    // this step's decision: sort (and rebuild) or commit in place
    uint32_t disp2_bits;  // float bits of this step's largest squared displacement from xref
    int32_t nsorts;       // sorts (list rebuilds) so far, the upload's included
    int32_t scan_done;    // cell-count tiles scanned in this sort (reset by the last one)
    int32_t outside_acc;  // particles far outside the grid, counted by the running sort
    int32_t shepard_ran;  // 1: this step's Shepard pass wrote rhon (the sort moves it into rho)
    int32_t nbrick;       // FLAG_BRICK: bricks holding particles (SceneDev::blist), set by each sort
    int32_t brick_spill;  // FLAG_BRICK: brick walks whose region exceeded kBrickCap (walked from device memory)
    int32_t pad_dyn;
};

// Bytes of SceneDev::bflag for nb bricks: whole 16-byte words, kBlock per
// 4,096 bricks (the compaction gives each thread of one block a run of words).
__host__ __device__ inline size_t brick_flag_bytes(int nb) {
    return ((size_t)(nb + 16 * kBlock - 1) / (16 * kBlock)) * (16 * kBlock);
}

// Index into SceneDev::set of the current state and of the step's outputs
// (same particle order, the other per-step copy).
__host__ __device__ inline int cur_set(const SceneDyn& d) { return 2 * d.parity + d.dpar; }
__host__ __device__ inline int next_set(const SceneDyn& d) { return 2 * d.parity + (d.dpar ^ 1); }

// A view of the fluid state in one particle order.  Masters are binary64
// SoA arrays (state integration and downloads).  A, B (and kap) are the
// packed records the pair loops gather, in the pair precision T (float in
// FP32_BULK, double in FP64), 4 values per particle so one record is one or
// two vector loads:
//   3D: A = (vx, vy, vz, m)   B = (rho, p/rho^2, gap, T)   kap = kappa
//   2D: A = (vx, vy, m, gap)  B = (rho, p/rho^2, T, kappa)
// The arrays come in two kinds.  Per-step arrays (x, v, T, A, B, xf) are
// rewritten by every step: the force pass writes the new state into the
// other copy (`dpar ^ 1`), and a step without a sort just switches copies.
// The other arrays change only through the sort (which permutes every array
// into the other order) or in place (rho, p, c by the continuity pass).
// SceneDev::set[2 * parity + dpar] pairs the arrays of order `parity` with
// per-step copy `dpar`.  The sort writes A, B.zw and kap; the continuity pass
// writes B.xy; the force pass writes whole A and B records.
struct FluidSet {
    double* x;      // n*dim
    double* v;      // n*dim
    double* rho;
    double* p;
    double* c;
    double* m;
    double* gap;
    double* T;
    double* kappa;
    double* cp;
    int32_t* id;    // sequence number within the particle's group (upload index at first)
    int8_t* kind;   // KIND_* (inlet/outlet solver only; null otherwise)
    void* A;
    void* B;
    void* kap;
    // positions relative to the grid origin in float (float4 in 3D, float2 in
    // 2D), written by the sort for its new order: the neighbour-list build
    // after it tests candidates on these and falls back to the FP64 rule
    // only inside a rounding band around R
    void* xf;
    // FP32_BULK only (null otherwise): positions relative to the grid origin
    // in whole quanta SceneDev::qu (int4 in 3D, int2 in 2D; the position
    // components all kQInvalid when a coordinate is beyond 2^30 quanta or not
    // finite).  Differences of these are exact, so the continuity pass and
    // the 3D force pass's fluid walk take their pair geometry from one
    // 16-byte load per candidate (see qgeom).  In 3D, w holds rho - rho0 as
    // float bits, rewritten in place by the continuity pass (the density the
    // force pass uses; see FLAG_EOS_WALK).
    void* xq;
};

struct SceneDev {
    int32_t dim, kernel, solver, prec;
    int32_t flags;
    int32_t gamma;
    int32_t cap;         // fluid capacity
    int32_t ncell;       // cells of the grid (the tail bucket is index ncell)
    int32_t nc[3];
    int32_t nw;          // wall particles
    int64_t kreinit;     // density_reinitialized_interval (<=0: never)
    int64_t kcheck;      // check_interval (<=0: never)
    // kernel (SmoothKernel.jl:37-48); kA, kB: factors of W(q) and dW/dr(q)
    // in the q-forms below, inv_h = 1 / h
    double R, R2, h, sigma, W0;
    // candidate radius Rv = R (1 + skin): the grid, the candidate lists and
    // their reuse across steps while no particle has moved skin R / 2 since
    // the lists were built (disp2_limit); verlet = 0: rebuilt every step
    double Rv, Rv2, disp2_limit;
    int32_t verlet;
    int32_t pad3;
    double inv_h, kA, kB;
    // model (LiquidModel.jl:10-66)
    double rho0, c0, p0, b, nu0, t0, cp, beta, dt;
    double g[3];
    double check_min[3], check_max[3];
    double check_planes[8][4];  // deletion iff n . x > d for some plane q < check_nplanes
    int32_t check_nplanes;
    int32_t pad4;
    // cell grid: cells cs[k] wide along axis k (cs[k] >= Rv / stencil[k]);
    // neighbours of a cell lie within stencil[k] cells along axis k.  Cells
    // are half of Rv along x and finer along y and z (see build_grid), so a
    // row of cells along x holds about one lattice line of particles
    double lo[3];
    double inv_cs[3];
    double cs[3];
    int32_t stencil[3];
    int32_t pad1;
    // position quantum of FluidSet::xq: a power of two with the grid extent
    // below 2^30 quanta; qbandq: half-width of the band around R2q = R^2 / qu^2
    // inside which the quantized float distance cannot decide the FP64 rule
    // (see qgeom), both in quanta^2
    double qu, inv_qu;
    float qbandq, R2q, pad5, pad6;
    // the equation of state in float for FLAG_EOS_WALK (see eos_pr2)
    float rho0f, inv_rho0f, bf, p0f;
    float clinf;  // b / rho0: p = clinf (rho - rho0) for gamma 1 and p0 0
    float pad8;
    // fluid: set[2 * parity + dpar] (see FluidSet)
    FluidSet set[4];
    double* rhon;    // Shepard output / upload staging
    int32_t* key;
    int32_t* slot;
    int32_t* count;       // ncell + 1
    int32_t* cell_start;  // ncell + 2
    int32_t* cell_local;  // prefix of the counts within their scan tile (the running sort)
    int32_t* tmp;         // cap
    int32_t* bsum;        // scan block sums
    int32_t cell_tiles;   // ceil((ncell+1) / kCellTile)
    int32_t kff;          // neighbour-list width (fluid-fluid)
    int32_t kfw;          // neighbour-list width (fluid-wall)
    int32_t pad0;
    // neighbour lists in 4-entry chunks: entry k of particle i at
    // [((k / 4) * cap + i) * 4 + k % 4]; the pair passes apply the exact rule
    // to the candidates every step
    int32_t* vnbr;        // fluid-fluid candidates within Rv (the neighbours themselves when !verlet)
    int32_t* nvnb;        // their count per particle (overflow when > kff)
    int32_t* wnbr;        // fluid-wall candidates within Rv (filtered by the exact rule when used)
    int32_t* nnbw;        // their count per particle
    double* xref;         // positions when the candidate lists were built
    int32_t* ovf;         // particles whose lists overflowed this step (they scan their cells)
    // walls: static, sorted by cell once
    double* wx;      // nw*dim
    double* wn;      // nw*dim
    double* wgap;
    double* wv;      // nw*dim (zeros for fixed walls)
    double* wT;
    double* wkappa;
    int32_t* wkind;
    int32_t* wid;    // wall index in group-concatenated upload order
    int32_t* wcell_start;  // ncell + 1
    // per cell: 1 if some wall particle lies in a cell within stencil[k] of
    // it along every axis k (the cells a candidate search from it visits),
    // else 0 -- the list rebuild skips the wall search there
    uint8_t* wnear;
    // packed wall records in the pair precision (static):
    //   3D: wA = (nx, ny, nz, gap)  wB = (vx, vy, vz, T)  wkap = kappa
    //   2D: wA = (nx, ny, gap, T)   wB = (vx, vy, kappa, 0)
    void* wA;
    void* wB;
    void* wkap;
    void* wxf;       // wall positions relative to lo in float (float4 / float2)
    // inlet/outlet solver: migration planes along io_axis and buffer profile
    int32_t io_axis, profile, profile_terms, outlet_profile;
    double inlet_enter, inlet_shift, outlet_enter, outlet_exit;
    double profile_f, profile_nu, profile_l, profile_c[3];
    // bricks (FLAG_BRICK): bdim[k] cells per brick along axis k, nbk[k]
    // bricks along it; brick (bx, by, bz) is bx + nbk[0] (by + nbk[1] bz)
    int32_t bdim[3];
    int32_t nbk[3];
    int32_t nbricks;
    int32_t pad7;
    int32_t* blist;   // bricks holding particles, in brick order (SceneDyn::nbrick of them)
    uint8_t* bflag;   // nbricks marks of the running sort (cleared when blist is written)
    // LDS slot of every fluid candidate within its particle's brick region,
    // in the chunk layout of vnbr (entry k of particle i at
    // [((k / 4) * cap + i) * 4 + k % 4]); written with vnbr at each rebuild
    uint16_t* vslot;
    int32_t* spawn_src;   // per spawn (append order): source inlet particle
    int32_t* spawn_id;    //   its inlet sequence number
    int32_t* spawn_key;   //   cell of the copy
    int32_t* spawn_slot;  //   slot in that cell's count
    int32_t* spawn_order; // spawn index by rank (rank = order of the inlet sequence numbers)
    int32_t* leave_id;    // fluid sequence numbers of this step's fluid -> outlet moves
    SceneDyn* dyn;
    // a copy of *dyn taken by the sort pass before its scatter (by
    // k_sort_offsets_place): the scatter reads it while the scene's first
    // tile ends the step in *dyn (sort_finish)
    SceneDyn* snap;
};

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
    using type = float4;
};
template <>
struct Vec4<double> {
    using type = double4;
};

struct Tile {
    int32_t scene;   // index into the SceneDev table
    int32_t begin;   // first particle (or cell) of the tile
};

// ---------------------------------------------------------------------------
// Cell grid.  Cells are a little wider than R (host sets inv_cs), so two
// particles within R are never more than one cell apart along any axis.
// Positions outside the grid are clamped into the border cells, which keeps
// the search exact.  Non-finite coordinates go to cell 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int cell_coord(double x, double lo, double inv, int nc) {
    double f = floor((x - lo) * inv);
    if (!(f >= 0.0)) return 0;
    if (f >= (double)(nc - 1)) return nc - 1;
    return (int)f;
}

template <int D>
__device__ __forceinline__ void cell_of(const SceneDev& s, const double* x, int c[3]) {
    c[0] = cell_coord(x[0], s.lo[0], s.inv_cs[0], s.nc[0]);
    c[1] = cell_coord(x[1], s.lo[1], s.inv_cs[1], s.nc[1]);
    c[2] = D == 3 ? cell_coord(x[2], s.lo[2], s.inv_cs[2], s.nc[2]) : 0;
}

__device__ __forceinline__ int cell_linear(const SceneDev& s, int cx, int cy, int cz) {
    return cx + s.nc[0] * (cy + s.nc[1] * cz);
}

// Neighbour rule of the ABI: d2 in binary64, each operation rounded on its
// own (no contraction), d2 <= R*R.
template <int D>
__device__ __forceinline__ double dist2_exact(const double* a, const double* b, double* rv) {
#pragma clang fp contract(off)
    double d0 = a[0] - b[0];
    double d1 = a[1] - b[1];
    double s = d0 * d0;
    s = s + d1 * d1;
    rv[0] = d0;
    rv[1] = d1;
    if (D == 3) {
        double d2 = a[2] - b[2];
        s = s + d2 * d2;
        rv[2] = d2;
    }
    return s;
}

// Quantized position of x (see FluidSet::xq); also used by the host.
constexpr int32_t kQInvalid = INT32_MIN;
__host__ __device__ inline void quantize(const double* lo, double inv_qu, int D, const double* x, int32_t* q) {
    bool ok = true;
    for (int k = 0; k < D; ++k) {
        const double t = (x[k] - lo[k]) * inv_qu;  // inv_qu is a power of two: exact scaling
        ok = ok && t > -1073741824.0 && t < 1073741824.0;  // |t| < 2^30 (false for NaN)
        q[k] = ok ? (int32_t)rint(t) : 0;
    }
    for (int k = D; k < 3; ++k) q[k] = 0;
    if (!ok)
        for (int k = 0; k < 3; ++k) q[k] = kQInvalid;
}

// ---------------------------------------------------------------------------
// Smoothing kernels, CubicSpline.jl / Gaussian.jl / WendlandC2.jl /
// WendlandC4.jl.  Operation order follows the reference's expressions; the
// binary64 form does not contract, so values equal a plain C evaluation.
// ---------------------------------------------------------------------------
template <typename T, int K>
struct SmoothKernelFn;

template <typename T>
struct SmoothKernelFn<T, ESPH_CUBIC_SPLINE> {
    static constexpr double ratio = 2.0;
    __device__ static __forceinline__ T value(T r, T h, T sigma) {
#pragma clang fp contract(off)
        T q = r / h;
        if (q < T(1)) return sigma / T(4) * (T(3) * (q * q) * (q - T(2)) + T(4));
        if (q < T(2)) { T t = T(2) - q; return sigma / T(4) * (t * t * t); }
        return T(0);
    }
    __device__ static __forceinline__ T gradient(T r, T h, T sigma) {
#pragma clang fp contract(off)
        T q = r / h;
        if (q < T(1)) return sigma / h * T(3) / T(4) * q * (T(3) * q - T(4));
        if (q < T(2)) { T t = T(2) - q; return -sigma / h * T(3) / T(4) * (t * t); }
        return T(0);
    }
};

template <typename T>
struct SmoothKernelFn<T, ESPH_GAUSSIAN> {
    static constexpr double ratio = 3.0;
    __device__ static __forceinline__ T value(T r, T h, T sigma) {
#pragma clang fp contract(off)
        T q = r / h;
        if (q < T(3)) return sigma * exp(-(q * q));
        return T(0);
    }
    __device__ static __forceinline__ T gradient(T r, T h, T sigma) {
#pragma clang fp contract(off)
        T q = r / h;
        if (q < T(3)) return T(-2) * sigma / h * q * exp(-(q * q));
        return T(0);
    }
};

template <typename T>
struct SmoothKernelFn<T, ESPH_WENDLAND_C2> {
    static constexpr double ratio = 2.0;
    __device__ static __forceinline__ T value(T r, T h, T sigma) {
#pragma clang fp contract(off)
        T q = r / h;
        if (q < T(2)) {
            T t = T(2) - q, t2 = t * t;
            return sigma * (t2 * t2) * (T(1) + T(2) * q) / T(16);
        }
        return T(0);
    }
    __device__ static __forceinline__ T gradient(T r, T h, T sigma) {
#pragma clang fp contract(off)
        T q = r / h;
        if (q < T(2)) { T t = T(2) - q; return -sigma / h * T(5) / T(8) * q * (t * t * t); }
        return T(0);
    }
};

template <typename T>
struct SmoothKernelFn<T, ESPH_WENDLAND_C4> {
    static constexpr double ratio = 2.0;
    __device__ static __forceinline__ T value(T r, T h, T sigma) {
#pragma clang fp contract(off)
        T q = r / h;
        if (q < T(2)) {
            T t = T(2) - q, t2 = t * t, t6 = t2 * t2 * t2;
            return sigma * t6 * (T(35) * (q * q) + T(36) * q + T(12)) / T(768);
        }
        return T(0);
    }
    __device__ static __forceinline__ T gradient(T r, T h, T sigma) {
#pragma clang fp contract(off)
        T q = r / h;
        if (q < T(2)) {
            T t = T(2) - q, t2 = t * t, t5 = t2 * t2 * t;
            return -sigma / h * (T(35.0 / 96.0) * (q * q) + T(7.0 / 48.0) * q) * t5;
        }
        return T(0);
    }
};

// ---------------------------------------------------------------------------
// The same kernels as functions of q = r / h with the leading factors folded
// into kA and kB (pair loops: no division per kernel evaluation):
//   CubicSpline  kA = sigma/4,   kB = 3 sigma / (4 h)
//   Gaussian     kA = sigma,     kB = 2 sigma / h
//   WendlandC2   kA = sigma/16,  kB = 5 sigma / (8 h)
//   WendlandC4   kA = sigma/768, kB = sigma / h
// They equal the reference-order forms above up to rounding.
// ---------------------------------------------------------------------------
template <typename T, int K>
struct KernelQ;

template <typename T>
struct KernelQ<T, ESPH_CUBIC_SPLINE> {
    __device__ static __forceinline__ void wdw(T q, T kA, T kB, T& w, T& dw) {
        if (q < T(1)) {
            w = kA * (T(3) * q * q * (q - T(2)) + T(4));
            dw = kB * q * (T(3) * q - T(4));
        } else if (q < T(2)) {
            const T t = T(2) - q;
            w = kA * t * t * t;
            dw = -kB * t * t;
        } else {
            w = T(0);
            dw = T(0);
        }
    }
};

template <typename T>
struct KernelQ<T, ESPH_GAUSSIAN> {
    __device__ static __forceinline__ void wdw(T q, T kA, T kB, T& w, T& dw) {
        if (q < T(3)) {
            const T e = exp(-(q * q));
            w = kA * e;
            dw = -kB * q * e;
        } else {
            w = T(0);
            dw = T(0);
        }
    }
};

// The Wendland polynomials carry a power of t = 2 - q, so t = max(2 - q, 0)
// gives the zero outside the support without a branch (the same values
// inside it; -0 for W' outside).
template <typename T>
struct KernelQ<T, ESPH_WENDLAND_C2> {
    __device__ static __forceinline__ void wdw(T q, T kA, T kB, T& w, T& dw) {
        const T t = fmax(T(2) - q, T(0)), t3 = t * t * t;
        w = kA * t3 * t * (T(1) + T(2) * q);
        dw = -kB * q * t3;
    }
};

template <typename T>
struct KernelQ<T, ESPH_WENDLAND_C4> {
    __device__ static __forceinline__ void wdw(T q, T kA, T kB, T& w, T& dw) {
        const T t = fmax(T(2) - q, T(0)), t2 = t * t, t5 = t2 * t2 * t;
        w = kA * t5 * t * (T(35) * q * q + T(36) * q + T(12));
        dw = -kB * (T(35.0 / 96.0) * q * q + T(7.0 / 48.0) * q) * t5;
    }
};

// Reciprocal, square root and reciprocal square root of the pair arithmetic:
// IEEE in binary64; the hardware approximations (v_rcp_f32, v_sqrt_f32,
// v_rsq_f32, ~1 ulp) in binary32.
__device__ __forceinline__ double prcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float prcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double psqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float psqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double prsqrt(double x) { return 1.0 / sqrt(x); }
__device__ __forceinline__ float prsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
// a / b: IEEE in binary64; through the hardware reciprocal in binary32
__device__ __forceinline__ double pdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float pdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// r = |r_ij| and ri = 1 / r of a pair from its squared distance d2 >= 0.
// FAST (binary32, kernels that vanish continuously at R): one hardware
// reciprocal square root (v_rsq_f32) of d2 + FLT_MIN instead of a square
// root and a reciprocal; ri stays finite and r = 0 at d2 = 0, where every
// kernel's W' vanishes, so the pair gradient W' ri is 0 there without a test
// (normalize(0) -> 0, Neighbour.jl:54-59).  The product d2 ri is a few ulp
// less accurate than the square root, which matters only where W jumps at
// R: the truncated Gaussian (W(3h) = sigma e^-9) keeps the square root, so a
// pair the FP64 rule puts just inside R is not evaluated as q = 3.
// Otherwise (and in binary64: IEEE square root and division) ri = inf at
// d2 = 0 and the caller zeroes the gradient there.
template <bool FAST>
__device__ __forceinline__ void pair_r(float d2, float& r, float& ri) {
    if (FAST) {
        // d2 + FLT_MIN equals d2 for every d2 a pair can have but 0
        ri = __builtin_amdgcn_rsqf(d2 + 1.17549435e-38f);
        r = d2 * ri;
    } else {
        r = __builtin_amdgcn_sqrtf(d2);
        ri = __builtin_amdgcn_rcpf(r);
    }
}
template <bool FAST>
__device__ __forceinline__ void pair_r(double d2, double& r, double& ri) {
    r = sqrt(d2);
    ri = 1.0 / r;
}

// cos(pi a / b) of the wall force (0 <= a <= b): binary32 by the hardware
// cosine (v_cos_f32 takes its argument in revolutions, here in [0, 1/2]) and
// the hardware reciprocal; binary64 by the library, in the reference's order.
__device__ __forceinline__ float cos_pi_ratio(float a, float b) {
    return __builtin_amdgcn_cosf(0.5f * a * __builtin_amdgcn_rcpf(b));
}
__device__ __forceinline__ double cos_pi_ratio(double a, double b) { return cos(3.14159265358979323846 * a / b); }

// Integer power by squaring (UpdatePressure.jl:13-14 uses x^gamma, gamma::Int).
__device__ __forceinline__ double ipow(double x, int n) {
    double y = 1.0;
    while (n > 0) {
        if (n & 1) y *= x;
        x *= x;
        n >>= 1;
    }
    return y;
}

}  // namespace esph
