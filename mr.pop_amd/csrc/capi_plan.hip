// capi_plan.hip -- planning entry points of the C ABI: one problem (solve),
// many at once (solve_batch) and continuous batching (server).  All three
// run on one Scheduler: a queue of waiting problems, admitted to the engine
// while their device memory fits the budget, and one engine step per tick.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>

#include "engine.h"

using namespace mrpop;

struct mrpop_solution {
    int D = 0;
    SlotResult res;
};

namespace {

struct Request {
    uint64_t ticket;
    int n_arms;
    std::vector<float> bases;  // empty = default
    std::vector<float> spheres, capsules, zcaps, cuboids, zcubs;
    std::vector<float> start, goals;
    int n_goals;
    uint64_t seed;
    mrpop_settings st;
    SolveOptions opt;
    Slot *slot = nullptr;
    std::chrono::steady_clock::time_point t_submit;
    int state = 0;  // 0 waiting, 1 active, 2 done
    SlotResult done_result;
};

int copy_problem(const mrpop_problem &p, Request &r) {
    if (!p.scene || !p.start || !p.goals || p.n_arms < 2 || p.n_arms > MRPOP_MAX_ARMS || p.n_goals < 1)
        return MRPOP_ERR_INVALID;
    const int D = 7 * p.n_arms;
    const mrpop_scene &s = *p.scene;
    if (s.n_spheres < 0 || s.n_capsules < 0 || s.n_z_capsules < 0 || s.n_cuboids < 0 || s.n_z_cuboids < 0)
        return MRPOP_ERR_INVALID;
    r.n_arms = p.n_arms;
    if (p.bases) r.bases.assign(p.bases, p.bases + 12 * p.n_arms);
    auto cp = [](std::vector<float> &v, const float *src, int n, int w) {
        if (n > 0 && src) v.assign(src, src + (size_t)n * w);
    };
    cp(r.spheres, s.spheres, s.n_spheres, 4);
    cp(r.capsules, s.capsules, s.n_capsules, 8);
    cp(r.zcaps, s.z_capsules, s.n_z_capsules, 8);
    cp(r.cuboids, s.cuboids, s.n_cuboids, 15);
    cp(r.zcubs, s.z_cuboids, s.n_z_cuboids, 15);
    if ((s.n_spheres && !s.spheres) || (s.n_capsules && !s.capsules) || (s.n_z_capsules && !s.z_capsules) ||
        (s.n_cuboids && !s.cuboids) || (s.n_z_cuboids && !s.z_cuboids))
        return MRPOP_ERR_INVALID;
    r.start.assign(p.start, p.start + D);
    r.goals.assign(p.goals, p.goals + (size_t)p.n_goals * D);
    r.n_goals = p.n_goals;
    r.seed = p.seed;
    return MRPOP_OK;
}

SolveOptions to_options(const mrpop_solve_options *o) {
    SolveOptions so;
    if (o) {
        so.block_per_robot = o->block_per_robot;
        so.time_budget_s = o->time_budget_s;
        so.roadmap_capacity = o->roadmap_capacity;
        so.tree_capacity = o->tree_capacity;
        so.neighbour_capacity = o->neighbour_capacity;
        so.neighbour_candidates = o->neighbour_candidates;
        so.rounds_per_step = o->rounds_per_step;
        so.probe_samples = o->probe_samples;
        so.graph_rounds = o->graph_rounds;
        so.arm_split = o->arm_split;
        so.biased_iter = o->biased_iter;
        so.roadmap_cache = o->roadmap_cache;
    }
    return so;
}

}  // namespace

struct mrpop_server {
    Context *ctx = nullptr;
    std::unique_ptr<Engine> engine;
    size_t budget = 0;
    int max_active = 0;
    uint64_t next_ticket = 1;
    std::deque<Request *> waiting;
    std::vector<Request *> active;
    std::map<uint64_t, Request *> all;

    ~mrpop_server() {
        engine.reset();
        for (auto &kv : all) delete kv.second;
    }

    int submit(const mrpop_problem &p, const mrpop_settings *st, const mrpop_solve_options *o, uint64_t *ticket) {
        auto *r = new (std::nothrow) Request();
        if (!r) return MRPOP_ERR_BUDGET;
        if (int rc = copy_problem(p, *r)) {
            delete r;
            return ctx->fail(rc, "invalid problem");
        }
        if (st) r->st = *st;
        else mrpop_settings_default(&r->st);
        r->opt = to_options(o);
        r->ticket = next_ticket++;
        r->t_submit = std::chrono::steady_clock::now();
        waiting.push_back(r);
        all[r->ticket] = r;
        if (ticket) *ticket = r->ticket;
        return MRPOP_OK;
    }

    // Admit waiting problems in arrival order while they fit the budget.
    int admit() {
        while (!waiting.empty()) {
            if (max_active > 0 && (int)active.size() >= max_active) break;
            Request *r = waiting.front();
            mrpop_scene sc{(int)(r->spheres.size() / 4), (int)(r->capsules.size() / 8), (int)(r->zcaps.size() / 8),
                           (int)(r->cuboids.size() / 15), (int)(r->zcubs.size() / 15),
                           r->spheres.data(), r->capsules.data(), r->zcaps.data(), r->cuboids.data(), r->zcubs.data()};
            const float *bases = r->bases.empty() ? nullptr : r->bases.data();
            // rejected settings or values end this problem only, before they size anything
            int rc = engine->validate(sc, r->n_arms, bases, r->start.data(), r->n_goals, r->goals.data(), r->st, r->opt);
            if (!rc) {
                const size_t need = engine->slot_bytes(r->n_arms, r->st, r->opt);
                if (need > budget) {
                    // can never fit: refuse this request alone and go on with the queue
                    rc = ctx->fail(MRPOP_ERR_BUDGET, "one problem needs %zu bytes, budget %zu", need, budget);
                    waiting.pop_front();
                    r->state = 2;
                    r->done_result.status = rc;
                    continue;
                }
                if (engine->device_bytes() + need > budget) break;  // wait for memory to free up
            }
            Slot *s = nullptr;
            if (!rc)
                rc = engine->add(sc, r->n_arms, bases, r->start.data(), r->n_goals, r->goals.data(), r->seed, r->st,
                                 r->opt, &s);
            if (rc) {
                if (rc == MRPOP_ERR_BUDGET && !active.empty()) break;  // wait for memory to free up
                waiting.pop_front();
                r->state = 2;
                r->done_result.status = rc;
                continue;
            }
            s->t_submit = r->t_submit;  // time-to-solution counts from submission
            s->user = r;
            r->slot = s;
            r->state = 1;
            waiting.pop_front();
            active.push_back(r);
        }
        return MRPOP_OK;
    }

    // Move finished slots' results out and free their memory.
    void reap() {
        for (size_t i = 0; i < active.size();) {
            Request *r = active[i];
            if (r->slot->phase == Phase::Done) {
                r->done_result = r->slot->res;
                engine->release(r->slot);
                r->slot = nullptr;
                r->state = 2;
                active.erase(active.begin() + i);
            } else {
                ++i;
            }
        }
    }

    int step(int *n_active, int *n_waiting) {
        if (int rc = admit()) return rc;
        if (!active.empty())
            if (int rc = engine->step()) return rc;
        reap();
        if (n_active) *n_active = (int)active.size();
        if (n_waiting) *n_waiting = (int)waiting.size();
        return MRPOP_OK;
    }
};

namespace {

int make_server(mrpop_ctx *c, size_t budget, int max_active, mrpop_server **out) {
    auto *srv = new (std::nothrow) mrpop_server();
    if (!srv) return MRPOP_ERR_BUDGET;
    srv->ctx = (Context *)c;
    srv->engine.reset(new Engine((Context *)c));
    if (budget == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            delete srv;
            return ((Context *)c)->fail(MRPOP_ERR_HIP, "hipMemGetInfo failed");
        }
        budget = free_b / 10 * 9;
    }
    srv->budget = budget;
    srv->max_active = max_active;
    *out = srv;
    return MRPOP_OK;
}

mrpop_solution *solution_of(const Request *r, int D) {
    auto *s = new (std::nothrow) mrpop_solution();
    if (!s) return nullptr;
    s->D = D;
    s->res = r->state == 2 ? r->done_result : (r->slot ? r->slot->res : SlotResult{});
    return s;
}

}  // namespace

extern "C" {

int mrpop_solve_options_default(mrpop_solve_options *o) {
    if (!o) return MRPOP_ERR_INVALID;
    SolveOptions d;
    o->block_per_robot = d.block_per_robot;
    o->time_budget_s = d.time_budget_s;
    o->roadmap_capacity = d.roadmap_capacity;
    o->tree_capacity = d.tree_capacity;
    o->neighbour_capacity = d.neighbour_capacity;
    o->neighbour_candidates = d.neighbour_candidates;
    o->rounds_per_step = d.rounds_per_step;
    o->probe_samples = d.probe_samples;
    o->graph_rounds = d.graph_rounds;
    o->arm_split = d.arm_split;
    o->biased_iter = d.biased_iter;
    o->roadmap_cache = d.roadmap_cache;
    return MRPOP_OK;
}

int mrpop_solve_batch(mrpop_ctx *c, int n, const mrpop_problem *problems, const mrpop_settings *settings,
                      const mrpop_solve_options *options, mrpop_solution **out) {
    if (!c || n < 0 || (n > 0 && (!problems || !out))) return MRPOP_ERR_INVALID;
    if (hipSetDevice(((Context *)c)->device) != hipSuccess) return MRPOP_ERR_HIP;
    mrpop_server *srv = nullptr;
    if (int rc = make_server(c, 0, 0, &srv)) return rc;
    std::vector<uint64_t> tickets(n);
    int rc = MRPOP_OK;
    for (int i = 0; i < n && !rc; ++i)
        rc = srv->submit(problems[i], settings ? &settings[i] : nullptr, options ? &options[i] : nullptr, &tickets[i]);
    while (!rc && (!srv->active.empty() || !srv->waiting.empty())) rc = srv->step(nullptr, nullptr);
    if (!rc)
        for (int i = 0; i < n; ++i) {
            Request *r = srv->all[tickets[i]];
            out[i] = solution_of(r, 7 * problems[i].n_arms);
            if (!out[i]) rc = MRPOP_ERR_BUDGET;
        }
    delete srv;
    return rc;
}

int mrpop_solve(mrpop_ctx *c, const mrpop_problem *problem, const mrpop_settings *settings,
                const mrpop_solve_options *options, mrpop_solution **out) {
    if (!problem || !out) return MRPOP_ERR_INVALID;
    return mrpop_solve_batch(c, 1, problem, settings, options, out);
}

int mrpop_solution_info(const mrpop_solution *s, mrpop_result_info *info) {
    if (!s || !info) return MRPOP_ERR_INVALID;
    const SlotResult &r = s->res;
    info->solved = r.solved;
    info->cost = r.cost;
    info->path_length = s->D ? (int)(r.path.size() / s->D) : 0;
    info->start_tree_size = r.start_tree_size;
    info->goal_tree_size = r.goal_tree_size;
    info->iters = r.iters;
    info->wall_ns = r.wall_ns;
    info->kernel_ns = r.kernel_ns;
    info->copy_ns = r.copy_ns;
    info->ttfs_ns = r.ttfs_ns;
    info->n_seed_paths = (int)r.seed_paths.size();
    info->passes = r.passes;
    info->rounds = r.rounds;
    info->roadmap_rounds = r.roadmap_rounds;
    for (int a = 0; a < MRPOP_MAX_ARMS; ++a) info->roadmap_size[a] = r.roadmap_size[a];
    info->status = r.status;
    info->union_find_bound = r.events[kEvUnionFind];
    info->edge_table_full = r.events[kEvEdgeTable];
    info->connect_skipped = r.events[kEvConnectSkip];
    info->simplify_dropped = r.events[kEvSimplifyDrop];
    for (int a = 0; a < MRPOP_MAX_ARMS; ++a) info->roadmap_cached[a] = r.roadmap_cached[a];
    return MRPOP_OK;
}

int mrpop_solution_path(const mrpop_solution *s, int which, float *out, int cap, int *n_waypoints, float *cost) {
    if (!s || s->D <= 0) return MRPOP_ERR_INVALID;
    const std::vector<float> *p;
    float cst;
    if (which < 0) {
        p = &s->res.path;
        cst = s->res.cost;
    } else {
        if (which >= (int)s->res.seed_paths.size()) return MRPOP_ERR_INVALID;
        p = &s->res.seed_paths[which].path;
        cst = s->res.seed_paths[which].cost;
    }
    const int n = (int)(p->size() / s->D);
    if (n_waypoints) *n_waypoints = n;
    if (cost) *cost = cst;
    if (out && cap > 0) std::copy(p->begin(), p->begin() + (size_t)std::min(n, cap) * s->D, out);
    return MRPOP_OK;
}

int mrpop_solution_improvements(const mrpop_solution *s, uint64_t *out_ns, int cap, int *n) {
    if (!s) return MRPOP_ERR_INVALID;
    const int k = (int)s->res.improve_ns.size();
    if (n) *n = k;
    if (out_ns)
        for (int i = 0; i < std::min(k, cap); ++i) out_ns[i] = s->res.improve_ns[i];
    return MRPOP_OK;
}

int mrpop_solution_free(mrpop_solution *s) {
    delete s;
    return MRPOP_OK;
}

int mrpop_server_create(mrpop_ctx *c, size_t budget, int max_active, mrpop_server **out) {
    if (!c || !out) return MRPOP_ERR_INVALID;
    *out = nullptr;
    if (hipSetDevice(((Context *)c)->device) != hipSuccess) return MRPOP_ERR_HIP;
    return make_server(c, budget, max_active, out);
}

int mrpop_server_destroy(mrpop_server *srv) {
    delete srv;
    return MRPOP_OK;
}

int mrpop_server_submit(mrpop_server *srv, const mrpop_problem *p, const mrpop_settings *st,
                        const mrpop_solve_options *o, uint64_t *ticket) {
    if (!srv || !p) return MRPOP_ERR_INVALID;
    return srv->submit(*p, st, o, ticket);
}

int mrpop_server_step(mrpop_server *srv, int *n_active, int *n_waiting) {
    if (!srv) return MRPOP_ERR_INVALID;
    return srv->step(n_active, n_waiting);
}

int mrpop_server_poll(mrpop_server *srv, uint64_t ticket, int *state, mrpop_solution **out) {
    if (!srv) return MRPOP_ERR_INVALID;
    auto it = srv->all.find(ticket);
    if (it == srv->all.end()) {
        if (state) *state = -1;
        return MRPOP_OK;
    }
    Request *r = it->second;
    if (state) *state = r->state;
    if (out) {
        *out = solution_of(r, 7 * r->n_arms);
        if (!*out) return MRPOP_ERR_BUDGET;
    }
    return MRPOP_OK;
}

int mrpop_server_cancel(mrpop_server *srv, uint64_t ticket) {
    if (!srv) return MRPOP_ERR_INVALID;
    auto it = srv->all.find(ticket);
    if (it == srv->all.end()) return MRPOP_ERR_INVALID;
    Request *r = it->second;
    if (r->state == 0) {
        srv->waiting.erase(std::find(srv->waiting.begin(), srv->waiting.end(), r));
        r->state = 2;
        r->done_result.status = MRPOP_ERR_STATE;
    } else if (r->state == 1) {
        r->slot->phase = Phase::Done;
        r->slot->res.status = MRPOP_ERR_STATE;
        srv->reap();
    }
    return MRPOP_OK;
}

size_t mrpop_server_bytes(const mrpop_server *srv) { return srv ? srv->engine->device_bytes() : 0; }

}  // extern "C"

// Diagnostics (not part of the public ABI): build the roadmaps of a problem
// for `rounds` roadmap rounds and copy arm `arm`'s roadmap out (nodes as
// [n][8] floats, degrees, neighbour lists [n][neighbour_capacity], union-find
// parents [n], the arm's start-goal connectivity flag and the number of
// rounds of the solve's first roadmap stretch the arm sampled in; any output
// may be NULL); *n gets the node count, at most `cap` nodes are copied.
extern "C" int mrpop_diag_roadmap(mrpop_ctx *c, const mrpop_problem *p, const mrpop_settings *settings,
                                  const mrpop_solve_options *options, int rounds, int arm, float *q, int *deg,
                                  int *nbr, int *uf, int *connected, int *n, int cap, int *stretch_rounds) {
    if (!c || !p || !n || rounds < 0) return MRPOP_ERR_INVALID;
    Context *ctx = (Context *)c;
    if (hipSetDevice(ctx->device) != hipSuccess) return MRPOP_ERR_HIP;
    Request r;
    if (int rc = copy_problem(*p, r)) return ctx->fail(rc, "invalid problem");
    mrpop_settings st;
    if (settings) st = *settings;
    else mrpop_settings_default(&st);
    Engine engine(ctx);
    mrpop_scene sc{(int)(r.spheres.size() / 4), (int)(r.capsules.size() / 8), (int)(r.zcaps.size() / 8),
                   (int)(r.cuboids.size() / 15), (int)(r.zcubs.size() / 15),
                   r.spheres.data(), r.capsules.data(), r.zcaps.data(), r.cuboids.data(), r.zcubs.data()};
    Slot *s = nullptr;
    if (int rc = engine.add(sc, r.n_arms, r.bases.empty() ? nullptr : r.bases.data(), r.start.data(), r.n_goals,
                            r.goals.data(), r.seed, st, to_options(options), &s))
        return rc;
    if (int rc = engine.debug_roadmap(s, rounds)) return rc;
    if (int rc = engine.debug_copy_roadmap(s, arm, q, deg, nbr, uf, connected, n, cap)) return rc;
    if (stretch_rounds)
        MRPOP_HIP(ctx, hipMemcpy(stretch_rounds, s->dv.rs_brounds + arm, sizeof(int), hipMemcpyDeviceToHost));
    return MRPOP_OK;
}

// Diagnostics (not part of the public ABI): device bytes one problem of
// n_arms arms holds with these settings and options (what the server's HBM
// budget admits against); 0 if they are rejected.
extern "C" size_t mrpop_diag_slot_bytes(mrpop_ctx *c, int n_arms, const mrpop_settings *settings,
                                        const mrpop_solve_options *options) {
    if (!c || n_arms < 2 || n_arms > MRPOP_MAX_ARMS) return 0;
    mrpop_settings st;
    if (settings) st = *settings;
    else mrpop_settings_default(&st);
    Engine engine((Context *)c);
    const SolveOptions opt = to_options(options);
    const mrpop_scene none{};
    std::vector<float> cfg((size_t)7 * n_arms, 0.f);
    if (engine.validate(none, n_arms, nullptr, cfg.data(), 1, cfg.data(), st, opt)) return 0;
    return engine.slot_bytes(n_arms, st, opt);
}

// Diagnostics (not part of the public ABI): the hyperspheroid transform an
// optimise pass samples with, for foci f1, f2 (D floats each) and best cost
// cb: tf (D x D row-major) and centre (D).  Host code: needs no device.
extern "C" int mrpop_diag_phs_transform(int D, const float *f1, const float *f2, double cb, float *tf,
                                        float *center) {
    if (D < 1 || D > 35 || !f1 || !f2 || !tf || !center) return MRPOP_ERR_INVALID;
    phs_transform(D, f1, f2, cb, tf, center);
    return MRPOP_OK;
}

// Diagnostics (not part of the public ABI): the lane deadline of a context
// (seconds a planner step may take before its problems end with
// MRPOP_ERR_TIMEOUT).  Tests set it tiny to take the give-up path with
// kernels that still complete.
extern "C" int mrpop_diag_set_lane_deadline(mrpop_ctx *c, double seconds) {
    if (!c || !(seconds >= 0)) return MRPOP_ERR_INVALID;
    ((Context *)c)->lane_deadline_s = seconds;
    return MRPOP_OK;
}

// Diagnostics (not part of the public ABI): the tree size from which dRRT
// expansions run in the five-launch form and connections scan the other
// tree a thread per expansion (0: always, < 0: never), and the least tree
// rows a block of those scans takes.  Both forms give the same results;
// tests compare them.
extern "C" int mrpop_diag_set_expand_split(mrpop_ctx *c, int min_nodes, int slice_rows) {
    if (!c || slice_rows < 1) return MRPOP_ERR_INVALID;
    ((Context *)c)->expand_split_nodes = min_nodes;
    ((Context *)c)->expand_slice_rows = slice_rows;
    return MRPOP_OK;
}

// Diagnostics (not part of the public ABI): the planner's counter-based
// generator, Philox4x32-10, evaluated on the host.
extern "C" void mrpop_diag_philox(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]) {
    const uint4 r = philox4x32_10(make_uint4(counter[0], counter[1], counter[2], counter[3]), make_uint2(key[0], key[1]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

// Diagnostics (not part of the public ABI): one problem's slot, driven round
// by round, with its device arrays readable by name.  The planner tests use
// it to restate roadmap and search rounds on the CPU.
namespace {
struct DiagSlot {
    Context *ctx = nullptr;
    std::unique_ptr<Engine> engine;
    Request r;
    Slot *s = nullptr;
};
}  // namespace

extern "C" int mrpop_diag_slot_open(mrpop_ctx *c, const mrpop_problem *p, const mrpop_settings *settings,
                                    const mrpop_solve_options *options, void **out) {
    if (!c || !p || !out) return MRPOP_ERR_INVALID;
    *out = nullptr;
    Context *ctx = (Context *)c;
    if (hipSetDevice(ctx->device) != hipSuccess) return MRPOP_ERR_HIP;
    auto h = std::make_unique<DiagSlot>();
    h->ctx = ctx;
    if (int rc = copy_problem(*p, h->r)) return ctx->fail(rc, "invalid problem");
    mrpop_settings st;
    if (settings) st = *settings;
    else mrpop_settings_default(&st);
    h->engine = std::make_unique<Engine>(ctx);
    const Request &r = h->r;
    mrpop_scene sc{(int)(r.spheres.size() / 4), (int)(r.capsules.size() / 8), (int)(r.zcaps.size() / 8),
                   (int)(r.cuboids.size() / 15), (int)(r.zcubs.size() / 15),
                   r.spheres.data(), r.capsules.data(), r.zcaps.data(), r.cuboids.data(), r.zcubs.data()};
    if (int rc = h->engine->add(sc, r.n_arms, r.bases.empty() ? nullptr : r.bases.data(), r.start.data(), r.n_goals,
                                r.goals.data(), r.seed, st, to_options(options), &h->s))
        return rc;
    *out = h.release();
    return MRPOP_OK;
}

extern "C" int mrpop_diag_slot_roadmap(void *handle, int rounds) {
    auto *h = (DiagSlot *)handle;
    if (!h || rounds < 0) return MRPOP_ERR_INVALID;
    return h->engine->debug_roadmap(h->s, rounds);
}

extern "C" int mrpop_diag_slot_pass(void *handle, int pass) {
    auto *h = (DiagSlot *)handle;
    if (!h) return MRPOP_ERR_INVALID;
    return h->engine->debug_start_pass(h->s, pass);
}

extern "C" int mrpop_diag_slot_simplify(void *handle, int stage) {
    auto *h = (DiagSlot *)handle;
    if (!h) return MRPOP_ERR_INVALID;
    return h->engine->debug_simplify(h->s, stage);
}

extern "C" int mrpop_diag_slot_search(void *handle, int rounds, int keep_round) {
    auto *h = (DiagSlot *)handle;
    if (!h) return MRPOP_ERR_INVALID;
    return h->engine->debug_search(h->s, rounds, keep_round != 0);
}

// Sizes: info[0..10] = A, D, B, M, rm_cap, rm_K, tr_cap, ce_cap, G, path_cap, sp_cap.
extern "C" int mrpop_diag_slot_info(void *handle, int *info) {
    auto *h = (DiagSlot *)handle;
    if (!h || !info) return MRPOP_ERR_INVALID;
    const SlotDev &v = h->s->dv;
    const int vals[11] = {v.A, v.D, v.B, v.M, v.rm_cap, v.rm_K, v.tr_cap, v.ce_cap, v.G, v.path_cap, v.sp_cap};
    for (int i = 0; i < 11; ++i) info[i] = vals[i];
    return MRPOP_OK;
}

// Copies the named device array (at most `cap` bytes); returns its size in
// bytes, or -1 for an unknown name or a failed copy.  "counters" is the
// slot's block of 64 ints: tr_n[2], ed_count, ce_count, status, best_cost
// (float), path_cost (float), path_n, path_meta[4], sp_count, ..., then the
// round counters at 16..19.
extern "C" long long mrpop_diag_slot_read(void *handle, const char *name, void *out, long long cap) {
    auto *h = (DiagSlot *)handle;
    if (!h || !name) return -1;
    const SlotDev &v = h->s->dv;
    const size_t A = v.A, D = v.D, B = v.B, rc = v.rm_cap, K = v.rm_K, tc = v.tr_cap, cc = v.ce_cap;
    const size_t pc = v.path_cap, sc = v.sp_cap;
    struct F {
        const char *n;
        const void *p;
        size_t bytes;
    };
    const F fields[] = {
        {"rm_q", v.rm_q, 4 * A * rc * 8},      {"rm_nbr", v.rm_nbr, 4 * A * rc * K}, {"rm_deg", v.rm_deg, 4 * A * rc},
        {"rm_n", v.rm_n, 4 * A},               {"tr_q", v.tr_q, 4 * 2 * tc * D},     {"tr_parent", v.tr_parent, 4 * 2 * tc},
        {"tr_cost", v.tr_cost, 4 * 2 * tc},    {"tr_rid", v.tr_rid, 4 * 2 * tc * A}, {"counters", v.tr_n, 4 * 64},
        {"ex_tree", v.ex_tree, 4 * B},         {"ex_state", v.ex_state, 4 * B},      {"ex_near", v.ex_near, 4 * B},
        {"ex_exist", v.ex_exist, 4 * B},       {"ex_index", v.ex_index, 4 * B},      {"ex_dup", v.ex_dup, 4 * B},
        {"ex_cost", v.ex_cost, 4 * B},         {"ex_q", v.ex_q, 4 * B * D},          {"ex_rid", v.ex_rid, 4 * B * A},
        {"ed_b", v.ed_b, 4 * B},               {"ed_coll", v.ed_coll, 4 * B},        {"ed_from", v.ed_from, 4 * B * D},
        {"ed_to", v.ed_to, 4 * B * D},         {"cn_near", v.cn_near, 4 * B},        {"cn_next", v.cn_next, 4 * B},
        {"cn_d", v.cn_d, 4 * B},               {"cn_fail", v.cn_fail, 4 * B},        {"ce_b", v.ce_b, 4 * cc},
        {"ce_coll", v.ce_coll, 4 * cc},        {"ce_from", v.ce_from, 4 * cc * D},   {"ce_to", v.ce_to, 4 * cc * D},
        {"path", v.path, 4 * pc * D},          {"sp_fail", v.sp_fail, 4 * pc * 16},  {"sp_mid", v.sp_mid, 4 * pc * D},
        {"sp_item", v.sp_item, 4 * sc},        {"sp_coll", v.sp_coll, 4 * sc},       {"sp_from", v.sp_from, 4 * sc * D},
        {"sp_to", v.sp_to, 4 * sc * D},        {"sp_cost", v.sp_cost, 4 * pc},       {"sp_parent", v.sp_parent, 4 * pc},
        {"ex_sample", v.ex_sample, 4 * B * D}, {"ex_crand", v.ex_crand, 4 * B},      {"phs_tf", v.phs_tf, 4 * D * D},
        {"rw_key", v.rw_key, 8 * 2 * tc},
    };
    if (std::strcmp(name, "phs_center") == 0) {  // held in the slot descriptor
        const size_t n = std::min<size_t>(4 * D, cap > 0 ? (size_t)cap : 0);
        if (out && n) std::memcpy(out, v.phs_center, n);
        return (long long)(4 * D);
    }
    for (const F &f : fields) {
        if (std::strcmp(f.n, name) != 0) continue;
        const size_t n = std::min<size_t>(f.bytes, cap > 0 ? (size_t)cap : 0);
        if (out && n && hipMemcpy(out, f.p, n, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        return (long long)f.bytes;
    }
    return -1;
}

extern "C" void mrpop_diag_slot_close(void *handle) {
    auto *h = (DiagSlot *)handle;
    if (!h) return;
    if (h->s) h->engine->release(h->s);
    delete h;
}
