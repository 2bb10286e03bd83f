// esph_batch.hip — batched stepping of many resident scenes.
//
// A batch holds a device table of the member scenes' SceneDev entries and
// tile tables that map thread blocks to (scene, first particle).  One step of
// the batch is one launch per (dim, kernel, precision) bucket and phase plus
// one sort launch for all members, so small scenes share launches instead of
// paying them one by one.  Optionally the step is captured once as a HIP graph
// and replayed.  Members keep their own device state (SceneDyn): a scene
// stepped alone or inside a batch runs the same per-particle arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/ethersph.h"
#include "esph_device.hpp"
#include "esph_launch.hpp"
#include "esph_runtime.hpp"

using namespace esph;

namespace esph {

std::array<int, 4> depth_key(const Bucket& b) {
    return {b.dim, b.kernel, b.prec, (b.has_thermal ? 2 : 0) | (b.has_plain || b.has_eos || b.has_brick ? 1 : 0)};
}

void apply_depth(const esph_ctx* ctx, Bucket& b) {
    if (!ctx->tune) {
        b.depth = b.cdepth = b.sdepth = ctx->fixed_depth;
        b.bs = b.cbs = b.sbs = ctx->fixed_block;
        return;
    }
    auto it = ctx->depth.find(depth_key(b));
    b.depth = it == ctx->depth.end() ? 0 : it->second;
    auto ib = ctx->block.find(depth_key(b));
    b.bs = ib == ctx->block.end() ? kBlock : ib->second;
    auto ip = ctx->pass_cfg.find(depth_key(b));
    if (ip != ctx->pass_cfg.end()) {
        b.cdepth = ip->second[0][0];
        b.cbs = ip->second[0][1];
        b.sdepth = ip->second[1][0];
        b.sbs = ip->second[1][1];
    } else {
        b.cdepth = b.sdepth = b.depth > 0 ? b.depth : 2;
        b.cbs = b.sbs = b.bs;
    }
}

// Shortest of two timed launches (after one untimed) of a pass of b, in ms.
template <typename F>
static hipError_t time_pass(F&& launch, hipStream_t st, hipEvent_t e0, hipEvent_t e1, float& t) {
    hipError_t e = launch();
    t = 1e30f;
    for (int r = 0; r < 2 && e == hipSuccess; ++r) {
        e = hipEventRecord(e0, st);
        if (e == hipSuccess) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        t = std::min(t, ms);
    }
    return e;
}

// Shortest of three timed force passes of b (after one untimed), in ms.
static hipError_t time_forces(Bucket& b, hipStream_t st, hipEvent_t e0, hipEvent_t e1, float& t) {
    hipError_t e = launch_forces(b, st);  // warm-up
    t = 1e30f;
    for (int r = 0; r < 3 && e == hipSuccess; ++r) {
        e = hipEventRecord(e0, st);
        if (e == hipSuccess) e = launch_forces(b, st);
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        t = std::min(t, ms);
    }
    return e;
}

hipError_t tune_depth(esph_ctx* ctx, Bucket& b, hipStream_t st) {
    const int dcand[3] = {1, 2, 4};
    const int bcand[3] = {64, 128, kBlock};
    std::array<float, 3> dbest{}, bbest{};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // gather depth at kBlock-thread workgroups, then the workgroup size at that depth
    b.bs = kBlock;
    for (int c = 0; c < 3 && e == hipSuccess; ++c) {
        b.depth = dcand[c];
        e = time_forces(b, st, e0, e1, dbest[c]);
    }
    int k = 0;
    for (int c = 1; c < 3; ++c)
        if (dbest[c] < dbest[k]) k = c;
    b.depth = dcand[k];
    bbest[2] = dbest[k];
    for (int c = 0; c < 2 && e == hipSuccess; ++c) {
        b.bs = bcand[c];
        e = time_forces(b, st, e0, e1, bbest[c]);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) {
        b.depth = 2;
        b.bs = kBlock;
        return e;
    }
    int kb = 2;
    for (int c = 0; c < 2; ++c)
        if (bbest[c] < bbest[kb]) kb = c;
    b.bs = bcand[kb];
    // continuity (dry) and Shepard (forced) at every depth and workgroup size
    std::array<std::array<int, 2>, 2> cfg{};
    std::array<std::array<float, 9>, 2> pms{};
    e0 = e1 = nullptr;
    e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    b.tuning = true;
    for (int pass = 0; pass < 2 && e == hipSuccess; ++pass) {
        int best = 0;
        for (int c = 0; c < 9 && e == hipSuccess; ++c) {
            int& d = pass == 0 ? b.cdepth : b.sdepth;
            int& w = pass == 0 ? b.cbs : b.sbs;
            d = dcand[c / 3];
            w = bcand[c % 3];
            e = time_pass([&] { return pass == 0 ? launch_continuity(b, st) : launch_shepard(b, st); }, st, e0, e1,
                          pms[pass][c]);
            if (pms[pass][c] < pms[pass][best]) best = c;
        }
        cfg[pass] = {dcand[best / 3], bcand[best % 3]};
    }
    b.tuning = false;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) {
        b.cdepth = b.sdepth = b.depth;
        b.cbs = b.sbs = b.bs;
        return e;
    }
    b.cdepth = cfg[0][0];
    b.cbs = cfg[0][1];
    b.sdepth = cfg[1][0];
    b.sbs = cfg[1][1];
    ctx->depth[depth_key(b)] = b.depth;
    ctx->depth_ms[depth_key(b)] = dbest;
    ctx->block[depth_key(b)] = b.bs;
    ctx->block_ms[depth_key(b)] = bbest;
    ctx->pass_cfg[depth_key(b)] = cfg;
    ctx->pass_ms[depth_key(b)] = pms;
    return hipSuccess;
}

std::vector<int> StepPlan::launch_config() const {
    std::vector<int> v;
    for (const Bucket& b : buckets) v.insert(v.end(), {b.depth, b.bs, b.cdepth, b.cbs, b.sdepth, b.sbs});
    return v;
}

int StepPlan::build(esph_ctx* ctx, const std::vector<esph_scene*>& members) {
    release();
    const int n = (int)members.size();
    std::vector<SceneDev> table(n);
    std::map<std::tuple<int, int, int>, std::vector<int>> groups;
    std::vector<Tile> all_p, all_c;
    for (int k = 0; k < n; ++k) {
        const esph_scene* s = members[k];
        table[k] = s->h;
        groups[{s->d.dim, s->d.kernel, s->d.precision}].push_back(k);
        for (int t = 0; t < s->nptiles; ++t) all_p.push_back(Tile{k, t * kBlock});
        for (int t = 0; t < s->nctiles; ++t) all_c.push_back(Tile{k, t * kCellTile});
    }
    // per-bucket tile tables, laid out after the global ones in one buffer
    std::vector<Tile> tiles(all_p);
    tiles.insert(tiles.end(), all_c.begin(), all_c.end());
    struct Off { size_t p, np, o, no, bk, nbk; bool th, pl, eo, br, tl; int dim, kernel, prec, ns; };
    std::vector<Off> offs;
    for (auto& g : groups) {
        Off o{};
        o.dim = std::get<0>(g.first);
        o.kernel = std::get<1>(g.first);
        o.prec = std::get<2>(g.first);
        o.ns = (int)g.second.size();
        o.p = tiles.size();
        for (int k : g.second) {
            for (int t = 0; t < members[k]->nptiles; ++t) tiles.push_back(Tile{k, t * kBlock});
            const bool th = (members[k]->h.flags & FLAG_THERMAL) != 0;
            const bool eo = (members[k]->h.flags & FLAG_EOS_WALK) != 0;
            const bool br = (members[k]->h.flags & FLAG_BRICK) != 0;
            o.th = o.th || th;
            o.pl = o.pl || (!th && !eo);
            o.eo = o.eo || (eo && !br);
            o.br = o.br || br;
            o.tl = o.tl || !br;
        }
        o.np = tiles.size() - o.p;
        o.o = tiles.size();
        for (int k : g.second)
            for (int t = 0; t < kOvfTiles; ++t) tiles.push_back(Tile{k, t * kBlock});
        o.no = tiles.size() - o.o;
        o.bk = tiles.size();
        for (int k : g.second)
            if (members[k]->h.flags & FLAG_BRICK) tiles.push_back(Tile{k, 0});
        o.nbk = tiles.size() - o.bk;
        offs.push_back(o);
    }
    SceneDev* dtable = (SceneDev*)arena.get(sizeof(SceneDev) * std::max(n, 1));
    Tile* dtiles = (Tile*)arena.get(sizeof(Tile) * std::max<size_t>(tiles.size(), 1));
    if (!dtable || !dtiles) {
        release();
        return fail(ctx, ESPH_ERR_HIP, "batch allocation failed");
    }
    hipError_t e = hipMemcpy(dtable, table.data(), sizeof(SceneDev) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess && !tiles.empty())
        e = hipMemcpy(dtiles, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        release();
        return hip_fail(ctx, "batch upload", e);
    }
    buckets.clear();
    for (const Off& o : offs) {
        buckets.push_back(Bucket{o.dim, o.kernel, o.prec, dtable, dtiles + o.p, (int)o.np, dtiles + o.o, o.ns, o.th,
                                 o.pl});
        buckets.back().has_eos = o.eo;
        buckets.back().has_brick = o.br;
        buckets.back().has_tile = o.tl;
        buckets.back().btiles = dtiles + o.bk;
        buckets.back().nbtiles = (int)o.nbk;
        apply_depth(ctx, buckets.back());
    }
    plan = SortPlan{dtable, n, dtiles, (int)all_p.size(), dtiles + all_p.size(), (int)all_c.size()};
    nmembers = n;
    this->members = members;
    return ESPH_OK;
}

hipError_t StepPlan::launch_step(hipStream_t st, bool shepard) const {
    hipError_t e = hipSuccess;
    for (const Bucket& b : buckets)
        if (e == hipSuccess) e = launch_continuity(b, st);
    for (const Bucket& b : buckets)
        if (e == hipSuccess) e = launch_forces(b, st);
    if (shepard)
        for (const Bucket& b : buckets)  // gated per scene on the device
            if (e == hipSuccess) e = launch_shepard(b, st);
    if (e == hipSuccess) e = launch_sort(plan, 0, st);
    return e;
}

bool StepPlan::shepard_due(int64_t t) const {
    for (const esph_scene* s : members) {
        const int64_t K = s->d.density_reinitialized_interval;
        if (K > 0 && (s->step + t + 1) % K == 0) return true;
    }
    return false;
}

void StepPlan::drop_graphs() {
    for (int q = 0; q < 2; ++q) {
        if (exec[q]) (void)hipGraphExecDestroy(exec[q]);
        if (graph[q]) (void)hipGraphDestroy(graph[q]);
        exec[q] = nullptr;
        graph[q] = nullptr;
    }
    for (int k = 0; k < kChunkLevels; ++k) {
        if (cexec[k]) (void)hipGraphExecDestroy(cexec[k]);
        if (cgraph[k]) (void)hipGraphDestroy(cgraph[k]);
        cexec[k] = nullptr;
        cgraph[k] = nullptr;
    }
}

int StepPlan::step(esph_ctx* ctx, int64_t nsteps) {
    hipStream_t st = ctx->stream;
    // buckets of an untuned class: the first step runs eagerly and tunes
    // their gather depth and workgroup size between its continuity and force passes
    bool untuned = false;
    for (Bucket& b : buckets) {
        apply_depth(ctx, b);
        untuned = untuned || b.depth == 0;
    }
    int64_t t0 = 0;  // steps of this call already launched
    if (untuned && nsteps > 0) {
        hipError_t e = hipSuccess;
        for (const Bucket& b : buckets)
            if (e == hipSuccess) e = launch_continuity(b, st);
        for (Bucket& b : buckets)
            if (e == hipSuccess && b.depth == 0) {
                e = tune_depth(ctx, b, st);
                for (Bucket& o : buckets) apply_depth(ctx, o);  // same class, same depth
            }
        for (const Bucket& b : buckets)
            if (e == hipSuccess) e = launch_forces(b, st);
        if (shepard_due(0))
            for (const Bucket& b : buckets)
                if (e == hipSuccess) e = launch_shepard(b, st);
        if (e == hipSuccess) e = launch_sort(plan, 0, st);
        if (e != hipSuccess) return hip_fail(ctx, "tuning step", e);
        nsteps -= 1;
        t0 = 1;
    }
    // a graph captured with other depths or workgroup sizes (tuning settings
    // changed) is recaptured
    const std::vector<int> depths = launch_config();
    if (depths != captured_depth) drop_graphs();
    // captures `steps` steps (with the Shepard pass when shepard) into g / x
    auto capture = [&](int steps, bool shepard, hipGraph_t& g, hipGraphExec_t& x) -> int {
        captured_depth = depths;
        HIPCHK(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int k = 0; k < steps && e == hipSuccess; ++k) e = launch_step(st, shepard);
        hipError_t e2 = hipStreamEndCapture(st, &g);
        if (e != hipSuccess) return hip_fail(ctx, "graph capture", e);
        if (e2 != hipSuccess) return hip_fail(ctx, "graph capture end", e2);
        HIPCHK(ctx, hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
        return ESPH_OK;
    };
    const int64_t end = t0 + nsteps;
    for (int64_t t = t0; t < end;) {
        const int q = shepard_due(t) ? 1 : 0;
        if (use_graph && q == 0) {
            // the longest chunk (2, 4, 8 or 16 steps) of the run of steps
            // without a Shepard pass that starts here
            int run = 1;
            while (run < (2 << (kChunkLevels - 1)) && t + run < end && !shepard_due(t + run)) ++run;
            int k = kChunkLevels - 1;
            while (k >= 0 && (2 << k) > run) --k;
            if (k >= 0) {
                if (!cexec[k]) {
                    const int rc = capture(2 << k, false, cgraph[k], cexec[k]);
                    if (rc != ESPH_OK) return rc;
                }
                const hipError_t e = hipGraphLaunch(cexec[k], st);
                if (e != hipSuccess) return hip_fail(ctx, "batch step", e);
                t += 2 << k;
                continue;
            }
        }
        if (use_graph && !exec[q]) {
            const int rc = capture(1, q == 1, graph[q], exec[q]);
            if (rc != ESPH_OK) return rc;
        }
        hipError_t e = use_graph ? hipGraphLaunch(exec[q], st) : launch_step(st, q == 1);
        if (e != hipSuccess) return hip_fail(ctx, "batch step", e);
        ++t;
    }
    return ESPH_OK;
}

void StepPlan::release() {
    drop_graphs();
    arena.release();
    buckets.clear();
    members.clear();
    nmembers = 0;
}

}  // namespace esph

struct esph_batch {
    esph_ctx* ctx = nullptr;
    std::vector<esph_scene*> members;
    esph::StepPlan plan;
};

extern "C" int esph_batch_create(esph_ctx* ctx, int32_t nscenes, esph_scene* const* scenes, esph_batch** out) {
    if (!ctx || !out || nscenes <= 0 || !scenes) return fail(ctx, ESPH_ERR_ARG, "batch_create: bad argument");
    *out = nullptr;
    std::vector<esph_scene*> m(scenes, scenes + nscenes);
    for (esph_scene* s : m) {
        if (!s || s->ctx != ctx) return fail(ctx, ESPH_ERR_ARG, "batch_create: scene of another context");
        if (!s->fluid_ready) return fail(ctx, ESPH_ERR_STATE, "batch_create: scene without fluid");
    }
    std::vector<esph_scene*> u(m);
    std::sort(u.begin(), u.end());
    if (std::adjacent_find(u.begin(), u.end()) != u.end()) return fail(ctx, ESPH_ERR_ARG, "batch_create: duplicate scene");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    esph_batch* b = new esph_batch();
    b->ctx = ctx;
    b->members = m;
    int rc = b->plan.build(ctx, m);
    if (rc) {
        delete b;
        return rc;
    }
    for (esph_scene* s : m) s->batch_refs++;
    *out = b;
    return ESPH_OK;
}

extern "C" int esph_batch_destroy(esph_batch* b) {
    if (!b) return ESPH_ERR_ARG;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    b->plan.release();
    for (esph_scene* s : b->members) s->batch_refs--;
    delete b;
    return ESPH_OK;
}

extern "C" int esph_batch_enable_graph(esph_batch* b, int32_t enable) {
    if (!b) return ESPH_ERR_ARG;
    if (!enable) b->plan.drop_graphs();
    b->plan.use_graph = enable != 0;
    return ESPH_OK;
}

extern "C" int esph_batch_step(esph_batch* b, int64_t nsteps) {
    if (!b || nsteps < 0) return fail(b ? b->ctx : nullptr, ESPH_ERR_ARG, "batch_step: bad argument");
    HIPCHK(b->ctx, hipSetDevice(b->ctx->device));
    int rc = b->plan.step(b->ctx, nsteps);
    if (rc) return rc;
    for (esph_scene* s : b->members) s->step += nsteps;
    return ESPH_OK;
}

extern "C" int esph_batch_sync(esph_batch* b) {
    if (!b) return ESPH_ERR_ARG;
    for (esph_scene* s : b->members) {
        int rc = esph_scene_sync(s);
        if (rc) return rc;
    }
    return ESPH_OK;
}
