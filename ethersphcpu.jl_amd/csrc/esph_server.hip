// esph_server.hip — continuous batching of simulation requests.
//
// Requests (a scene descriptor, its particles and a step count) are copied in
// and queued.  esph_server_advance admits queued requests in arrival order
// while the context's HBM budget and the slot limit allow, steps all running
// scenes together as one batch (one launch per kernel class and phase,
// replayed from a captured graph; the batch is rebuilt only when its
// membership changes) until the soonest member finishes or the step limit is
// reached, then reads back and frees the finished scenes.  Requests that
// arrive while others run join the next batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "../../include/ethersph.h"
#include "esph_device.hpp"
#include "esph_runtime.hpp"

using namespace esph;

namespace {

double now_s() {
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

struct WallCopy {
    int32_t kind;
    int64_t n;
    std::vector<double> x, normal, gap, v, t, kappa;
};

struct FluidCopy {
    int64_t n = 0;
    std::vector<double> x, v, rho, p, c, m, gap, t, kappa, cp;

    void take(const esph_fluid_arrays& f, int D) {
        auto cp_ = [](std::vector<double>& dst, const double* src, int64_t k) {
            if (src && k > 0) dst.assign(src, src + k);
        };
        n = f.n;
        cp_(x, f.x, n * D); cp_(v, f.v, n * D); cp_(rho, f.rho, n); cp_(p, f.p, n);
        cp_(c, f.c, n); cp_(m, f.mass, n); cp_(gap, f.gap, n);
        cp_(t, f.t, n); cp_(kappa, f.kappa, n); cp_(cp, f.cp, n);
    }
    esph_fluid_arrays view() const {
        esph_fluid_arrays f{};
        f.n = n;
        f.x = x.data(); f.v = v.data(); f.rho = rho.data(); f.p = p.data(); f.c = c.data();
        f.mass = m.data(); f.gap = gap.data();
        f.t = t.empty() ? nullptr : t.data();
        f.kappa = kappa.empty() ? nullptr : kappa.data();
        f.cp = cp.empty() ? nullptr : cp.data();
        return f;
    }
};

struct Request {
    int64_t id = 0;
    esph_scene_desc desc{};
    FluidCopy fluid, inlet, outlet;
    bool buffers = false;
    std::vector<WallCopy> walls;
    int64_t nsteps = 0;
    int64_t done = 0;
    int32_t state = ESPH_REQ_QUEUED;
    esph_scene* scene = nullptr;
    double t_submit = 0, t_start = 0, t_end = 0;
    int64_t n_out = 0;
    std::map<int32_t, std::vector<double>> result;  // field -> values (reference order)
};

double percentile(std::vector<std::pair<double, double>> s, double q) {
    if (s.empty()) return 0.0;
    std::sort(s.begin(), s.end());
    double total = 0;
    for (auto& e : s) total += e.second;
    double acc = 0;
    for (auto& e : s) {
        acc += e.second;
        if (acc >= q * total) return e.first;
    }
    return s.back().first;
}

}  // namespace

struct esph_server {
    esph_ctx* ctx = nullptr;
    int32_t max_active = 0;
    int64_t next_id = 1;
    std::deque<int64_t> queue;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    std::vector<int64_t> active;
    StepPlan plan;
    bool dirty = true;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<std::pair<double, double>> lat;  // (ms per step, weight = scenes stepped)
    int64_t completed = 0, failed = 0, batch_steps = 0;
    double t_first = -1.0;
};

static int admit(esph_server* sv) {
    esph_ctx* ctx = sv->ctx;
    while (!sv->queue.empty() && (sv->max_active <= 0 || (int)sv->active.size() < sv->max_active)) {
        Request& r = *sv->reqs[sv->queue.front()];
        esph_scene* s = nullptr;
        int rc = esph_scene_create(ctx, &r.desc, &s);
        std::vector<esph_wall_arrays> wa;
        for (const WallCopy& w : r.walls) {
            esph_wall_arrays a{};
            a.kind = w.kind;
            a.n = w.n;
            a.x = w.x.data();
            a.normal = w.normal.data();
            a.gap = w.gap.data();
            a.v = w.v.data();
            a.t = w.t.data();
            a.kappa = w.kappa.data();
            wa.push_back(a);
        }
        if (!rc) rc = esph_scene_set_walls(s, (int32_t)wa.size(), wa.data());
        if (!rc && r.buffers) {
            const esph_fluid_arrays in = r.inlet.view(), out = r.outlet.view();
            rc = esph_scene_set_buffers(s, &in, &out);
        }
        if (!rc) {
            const esph_fluid_arrays f = r.fluid.view();
            rc = esph_scene_set_fluid(s, &f);
        }
        if (rc == ESPH_ERR_BUDGET) {
            if (s) esph_scene_destroy(s);
            if (sv->active.empty()) {  // can never fit
                r.state = ESPH_REQ_FAILED;
                sv->failed++;
                sv->queue.pop_front();
                continue;
            }
            break;  // wait for memory to free up (arrival order is kept)
        }
        sv->queue.pop_front();
        if (rc) {
            if (s) esph_scene_destroy(s);
            r.state = ESPH_REQ_FAILED;
            sv->failed++;
            continue;
        }
        r.scene = s;
        r.state = ESPH_REQ_RUNNING;
        r.t_start = now_s();
        sv->active.push_back(r.id);
        sv->dirty = true;
    }
    return ESPH_OK;
}

static int retire(esph_server* sv, Request& r) {
    esph_scene* s = r.scene;
    int rc = esph_scene_sync(s);
    int64_t n = 0;
    if (!rc) rc = esph_scene_fluid_count(s, &n);
    const int D = r.desc.dim;
    const bool thermal = r.desc.solver == ESPH_SOLVER_FLUID_THERMOSTATIC_WALL ||
                         r.desc.solver == ESPH_SOLVER_FLUID_COMPULSIVE_THERMOSTATIC_WALL;
    std::vector<int32_t> fields{ESPH_FIELD_X, ESPH_FIELD_V, ESPH_FIELD_RHO, ESPH_FIELD_P, ESPH_FIELD_C, ESPH_FIELD_ID};
    if (thermal) fields.push_back(ESPH_FIELD_T);
    for (int32_t f : fields) {
        if (rc) break;
        const int64_t w = (f == ESPH_FIELD_X || f == ESPH_FIELD_V) ? D : 1;
        std::vector<double> out((size_t)std::max<int64_t>(n * w, 1));
        int64_t written = 0;
        rc = esph_scene_get_fluid(s, f, out.data(), (int64_t)out.size(), &written);
        out.resize((size_t)written);
        r.result[f] = std::move(out);
    }
    esph_scene_destroy(s);
    r.scene = nullptr;
    r.n_out = n;
    r.t_end = now_s();
    r.state = rc ? ESPH_REQ_FAILED : ESPH_REQ_DONE;
    if (rc) sv->failed++; else sv->completed++;
    // the particle copies are no longer needed
    r.fluid = FluidCopy();
    r.inlet = FluidCopy();
    r.outlet = FluidCopy();
    r.walls.clear();
    return ESPH_OK;
}

extern "C" int esph_server_create(esph_ctx* ctx, int32_t max_active, esph_server** out) {
    if (!ctx || !out) return fail(ctx, ESPH_ERR_ARG, "server_create: bad argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    esph_server* sv = new esph_server();
    sv->ctx = ctx;
    sv->max_active = max_active;
    if (hipEventCreate(&sv->ev0) != hipSuccess || hipEventCreate(&sv->ev1) != hipSuccess) {
        delete sv;
        return fail(ctx, ESPH_ERR_HIP, "server_create: event creation failed");
    }
    *out = sv;
    return ESPH_OK;
}

extern "C" int esph_server_destroy(esph_server* sv) {
    if (!sv) return ESPH_ERR_ARG;
    (void)hipSetDevice(sv->ctx->device);
    (void)hipStreamSynchronize(sv->ctx->stream);
    sv->plan.release();
    for (auto& kv : sv->reqs)
        if (kv.second->scene) esph_scene_destroy(kv.second->scene);
    (void)hipEventDestroy(sv->ev0);
    (void)hipEventDestroy(sv->ev1);
    delete sv;
    return ESPH_OK;
}

extern "C" int esph_server_submit(esph_server* sv, const esph_scene_desc* desc, int32_t nwall_groups,
                                  const esph_wall_arrays* walls, const esph_fluid_arrays* fluid,
                                  const esph_fluid_arrays* inlet, const esph_fluid_arrays* outlet, int64_t nsteps,
                                  int64_t* request_id) {
    if (!sv || !desc || !fluid || nsteps < 0 || nwall_groups < 0 || (nwall_groups > 0 && !walls) ||
        (!inlet) != (!outlet))
        return fail(sv ? sv->ctx : nullptr, ESPH_ERR_ARG, "server_submit: bad argument");
    const int D = desc->dim;
    if (D != 2 && D != 3) return fail(sv->ctx, ESPH_ERR_ARG, "server_submit: dim must be 2 or 3");
    if ((inlet != nullptr) != (desc->solver == ESPH_SOLVER_FLUID_INLET_OUTLET_WALL))
        return fail(sv->ctx, ESPH_ERR_ARG, "server_submit: buffers go with the inlet/outlet solver and only with it");
    auto r = std::make_unique<Request>();
    r->id = sv->next_id++;
    r->desc = *desc;
    r->nsteps = nsteps;
    r->fluid.take(*fluid, D);
    if (inlet) {
        r->buffers = true;
        r->inlet.take(*inlet, D);
        r->outlet.take(*outlet, D);
    }
    auto cp = [](std::vector<double>& dst, const double* src, int64_t n) {
        if (src && n > 0) dst.assign(src, src + n);
    };
    for (int q = 0; q < nwall_groups; ++q) {
        WallCopy w;
        w.kind = walls[q].kind;
        w.n = walls[q].n;
        cp(w.x, walls[q].x, w.n * D); cp(w.normal, walls[q].normal, w.n * D); cp(w.gap, walls[q].gap, w.n);
        cp(w.v, walls[q].v, w.n * D); cp(w.t, walls[q].t, w.n); cp(w.kappa, walls[q].kappa, w.n);
        r->walls.push_back(std::move(w));
    }
    r->t_submit = now_s();
    if (sv->t_first < 0) sv->t_first = r->t_submit;
    if (request_id) *request_id = r->id;
    sv->queue.push_back(r->id);
    sv->reqs[r->id] = std::move(r);
    return ESPH_OK;
}

extern "C" int esph_server_advance(esph_server* sv, int64_t max_steps, int64_t* stepped) {
    if (!sv || max_steps < 0) return fail(sv ? sv->ctx : nullptr, ESPH_ERR_ARG, "server_advance: bad argument");
    esph_ctx* ctx = sv->ctx;
    if (stepped) *stepped = 0;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    admit(sv);
    // zero-step requests finish at once
    for (size_t q = 0; q < sv->active.size();) {
        Request& r = *sv->reqs[sv->active[q]];
        if (r.done >= r.nsteps) {
            retire(sv, r);
            sv->active.erase(sv->active.begin() + q);
            sv->dirty = true;
        } else {
            ++q;
        }
    }
    if (sv->active.empty() || max_steps == 0) return ESPH_OK;
    int64_t k = max_steps;
    for (int64_t id : sv->active) k = std::min(k, sv->reqs[id]->nsteps - sv->reqs[id]->done);
    if (sv->dirty) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        std::vector<esph_scene*> members;
        for (int64_t id : sv->active) members.push_back(sv->reqs[id]->scene);
        sv->plan.release();
        sv->plan.use_graph = true;
        int rc = sv->plan.build(ctx, members);
        if (rc) return rc;
        sv->dirty = false;
    }
    HIPCHK(ctx, hipEventRecord(sv->ev0, ctx->stream));
    int rc = sv->plan.step(ctx, k);
    if (rc) return rc;
    HIPCHK(ctx, hipEventRecord(sv->ev1, ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(sv->ev1));
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, sv->ev0, sv->ev1));
    sv->lat.emplace_back((double)ms / (double)k, (double)k * (double)sv->active.size());
    sv->batch_steps += k;
    if (stepped) *stepped = k;
    for (size_t q = 0; q < sv->active.size();) {
        Request& r = *sv->reqs[sv->active[q]];
        r.done += k;
        r.scene->step += k;
        // a request whose simulation blew up (non-finite positions, or many
        // particles far outside its grid, which crowd the border cells) fails
        // now instead of slowing every member of the batch down
        SceneDyn dyn{};
        if (hipMemcpy(&dyn, r.scene->h.dyn, sizeof(SceneDyn), hipMemcpyDeviceToHost) == hipSuccess &&
            ((dyn.err & 1) || dyn.outside > std::max(64, dyn.n_live / 64))) {
            esph_scene_destroy(r.scene);
            r.scene = nullptr;
            r.state = ESPH_REQ_FAILED;
            r.t_end = now_s();
            sv->failed++;
            r.fluid = FluidCopy();
            r.inlet = FluidCopy();
            r.outlet = FluidCopy();
            r.walls.clear();
            sv->active.erase(sv->active.begin() + q);
            sv->dirty = true;
            continue;
        }
        if (r.done >= r.nsteps) {
            retire(sv, r);
            sv->active.erase(sv->active.begin() + q);
            sv->dirty = true;
        } else {
            ++q;
        }
    }
    admit(sv);
    return ESPH_OK;
}

extern "C" int esph_server_status(esph_server* sv, int64_t request_id, int32_t* state, int64_t* steps_done) {
    if (!sv) return ESPH_ERR_ARG;
    auto it = sv->reqs.find(request_id);
    if (it == sv->reqs.end()) return fail(sv->ctx, ESPH_ERR_ARG, "server_status: unknown request");
    if (state) *state = it->second->state;
    if (steps_done) *steps_done = it->second->done;
    return ESPH_OK;
}

extern "C" int esph_server_result(esph_server* sv, int64_t request_id, int32_t field, double* out, int64_t capacity,
                                  int64_t* written) {
    if (!sv) return ESPH_ERR_ARG;
    auto it = sv->reqs.find(request_id);
    if (it == sv->reqs.end()) return fail(sv->ctx, ESPH_ERR_ARG, "server_result: unknown request");
    Request& r = *it->second;
    if (r.state != ESPH_REQ_DONE) return fail(sv->ctx, ESPH_ERR_STATE, "server_result: request not done");
    auto f = r.result.find(field);
    if (f == r.result.end()) return fail(sv->ctx, ESPH_ERR_ARG, "server_result: field not kept");
    if (written) *written = (int64_t)f->second.size();
    if (!out) return ESPH_OK;
    if (capacity < (int64_t)f->second.size()) return fail(sv->ctx, ESPH_ERR_CAPACITY, "server_result: output too small");
    std::memcpy(out, f->second.data(), sizeof(double) * f->second.size());
    return ESPH_OK;
}

extern "C" int esph_server_release(esph_server* sv, int64_t request_id) {
    if (!sv) return ESPH_ERR_ARG;
    auto it = sv->reqs.find(request_id);
    if (it == sv->reqs.end()) return fail(sv->ctx, ESPH_ERR_ARG, "server_release: unknown request");
    if (it->second->state == ESPH_REQ_RUNNING) return fail(sv->ctx, ESPH_ERR_STATE, "server_release: request running");
    if (it->second->state == ESPH_REQ_QUEUED) sv->queue.erase(std::find(sv->queue.begin(), sv->queue.end(), request_id));
    sv->reqs.erase(it);
    return ESPH_OK;
}

extern "C" int esph_server_stats(esph_server* sv, double* out8) {
    if (!sv || !out8) return ESPH_ERR_ARG;
    const double t = now_s();
    out8[0] = percentile(sv->lat, 0.50);
    out8[1] = percentile(sv->lat, 0.99);
    out8[2] = (double)sv->completed;
    out8[3] = sv->t_first >= 0 && t > sv->t_first ? sv->completed / (t - sv->t_first) : 0.0;
    out8[4] = (double)sv->active.size();
    out8[5] = (double)sv->queue.size();
    out8[6] = (double)sv->ctx->used;
    out8[7] = (double)sv->batch_steps;
    return ESPH_OK;
}
