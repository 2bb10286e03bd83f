// esph_runtime.hpp — host-side objects behind the opaque C ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <string>
#include <vector>

#include "../../include/ethersph.h"
#include "esph_device.hpp"
#include "esph_launch.hpp"

struct esph_scene;

struct esph_ctx {
    int device = 0;
    hipStream_t stream = nullptr;  // default priority
    uint64_t budget = 0;           // 0: unlimited
    uint64_t used = 0;
    std::string err;
    // device address of the static probe staging array on this context's
    // device (esph_scene_probe_noalloc); null if it could not be resolved
    double* probe_scratch = nullptr;
    hipError_t probe_scratch_err = hipSuccess;
    // gather depth and workgroup size of the pair passes: tuned per bucket
    // class (dim, kernel, precision, thermal mix) at its first step, or fixed
    int tune = 1;
    int fixed_depth = 2;
    int fixed_block = 256;
    // brick walks for the 3D FP32_BULK scenes that allow them (FLAG_BRICK),
    // decided when a scene's buffers are allocated; off by default: on the
    // 1M-particle 3D dam break the force pass takes 0.38 ms with them against
    // 0.27 ms gathering from device memory
    int bricks = 0;
    // wall walks of the force pass in each particle's own lane (FLAG_WALL_LANES)
    // instead of shared out over the wavefront, decided at upload; off by default
    int wall_lanes = 0;
    std::map<std::array<int, 4>, int> depth;
    std::map<std::array<int, 4>, std::array<float, 3>> depth_ms;  // forces pass at depth 1, 2, 4 (256 threads)
    std::map<std::array<int, 4>, int> block;
    std::map<std::array<int, 4>, std::array<float, 3>> block_ms;  // forces pass at 64, 128, 256 threads
    // continuity [0] and Shepard [1] passes: (depth, workgroup size) chosen
    // and the times (ms) at depth {1, 2, 4} x {64, 128, 256} threads
    std::map<std::array<int, 4>, std::array<std::array<int, 2>, 2>> pass_cfg;
    std::map<std::array<int, 4>, std::array<std::array<float, 9>, 2>> pass_ms;
};

namespace esph {

// Device allocations owned by one object, released together.
struct Arena {
    std::vector<void*> ptrs;
    uint64_t total = 0;
    void* get(size_t bytes);
    void release();
};

struct WallHost {
    int64_t n = 0;
    std::vector<double> x, n_, v, gap, T, kappa;
    std::vector<int32_t> kind;
};

// Host copies kept alive while their asynchronous uploads are in flight.
struct WallStage {
    std::vector<double> x, n_, v, gap, T, kappa;
    std::vector<int32_t> kind, wid, wcs;
    std::vector<uint8_t> near;
    std::vector<float> fA, fB, fK, fX;
    std::vector<double> dA, dB, dK;
};

struct TimedStep {
    hipEvent_t ev[5];
    bool shepard;
};

// Launch plan of one step for a set of scenes (a batch, or one scene in
// graph mode): device scene table, per-bucket tile tables, sort plan and the
// optionally captured step graph.
// The Shepard pass is launched only on steps where some member re-initialises
// its density (the host knows every member's step count), so a plan holds
// two step graphs, [0] without and [1] with that pass.
struct StepPlan {
    Arena arena;
    std::vector<Bucket> buckets;
    SortPlan plan{};
    int nmembers = 0;
    std::vector<esph_scene*> members;
    bool use_graph = false;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    // runs of steps without a Shepard pass, 2 << k steps in one graph: the
    // launch of a graph costs a gap of several microseconds before its first
    // kernel, the kernels inside one follow each other without it
    static constexpr int kChunkLevels = 4;
    hipGraph_t cgraph[kChunkLevels] = {};
    hipGraphExec_t cexec[kChunkLevels] = {};
    std::vector<int> captured_depth;  // bucket depths and workgroup sizes the graphs were captured with
    std::vector<int> launch_config() const;  // those of the buckets now
    int build(esph_ctx* ctx, const std::vector<esph_scene*>& members);
    hipError_t launch_step(hipStream_t st, bool shepard) const;
    bool shepard_due(int64_t t) const;  // some member's step t from now re-initialises its density
    void drop_graphs();
    int step(esph_ctx* ctx, int64_t nsteps);
    void release();
};

// Pair-pass tuner.  apply_depth sets each pass's gather depth and workgroup
// size in b from the context (fixed values, or the tuned ones of b's class;
// b.depth 0 when that class is not tuned yet).  tune_depth times the force
// pass of b at gather depths 1, 2 and 4 with 256-thread workgroups, then at
// the fastest depth with 64- and 128-thread workgroups, on the current step's
// lists (the pass only writes the step outputs, so repeating it changes
// nothing); then the continuity pass (dry: its result to rhon, the state
// untouched) and the Shepard pass (whether or not due; its output is rhon)
// at every depth and workgroup size; and records the fastest of each.  It
// must follow the step's continuity pass on the stream.  Results depend on
// none of the choices.
std::array<int, 4> depth_key(const Bucket& b);
void apply_depth(const esph_ctx* ctx, Bucket& b);
hipError_t tune_depth(esph_ctx* ctx, Bucket& b, hipStream_t st);

int fail(esph_ctx* ctx, int code, const std::string& msg);
int hip_fail(esph_ctx* ctx, const char* what, hipError_t e);
bool kernel_params(int kind, int dim, double R, double* h, double* sigma);

}  // namespace esph

struct esph_scene {
    esph_ctx* ctx = nullptr;
    esph_scene_desc d{};
    double kh = 0, ksigma = 0;
    esph::SceneDev h{};            // host mirror of the device table entry
    esph::SceneDev* dev = nullptr; // device table entry (one scene)
    esph::Tile* ptiles = nullptr;
    int nptiles = 0;
    esph::Tile* ctiles = nullptr;
    int nctiles = 0;
    esph::Tile* otiles = nullptr;  // overflow-list tiles of the SCAN kernels
    esph::Arena arena;
    esph::WallHost walls;
    esph::WallStage wall_stage;
    std::vector<esph::Tile> tile_stage_p, tile_stage_c, tile_stage_o;
    bool walls_set = false;
    int batch_refs = 0;  // live batches holding this scene
    // inlet/outlet buffers (host copies, inlet then outlet), set before the fluid
    int64_t n_inlet = 0, n_outlet = 0;
    std::vector<double> bx, bv, brho, bp, bc, bm, bgap;
    bool allocated = false;
    bool fluid_ready = false;
    int64_t n_upload = 0;
    int64_t step = 0;  // host mirror of SceneDyn::step
    bool timing = false;
    std::vector<hipEvent_t> timing_events;
    size_t ev_next = 0;
    std::vector<esph::TimedStep> timed;
    bool use_graph = false;     // step through a captured one-step graph
    esph::StepPlan gplan;       // this scene alone (graph mode)
};

#define HIPCHK(ctx, expr)                                          \
    do {                                                           \
        hipError_t e_ = (expr);                                    \
        if (e_ != hipSuccess) return esph::hip_fail((ctx), #expr, e_); \
    } while (0)
