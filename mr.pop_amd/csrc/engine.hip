// engine.hip -- multi-problem planning engine (see engine.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include "engine.h"

namespace mrpop {

namespace {

constexpr int kPathCap = 2048;             // waypoints (MAX_PATH_SIZE floats in the reference)
constexpr int kMaxConnectChunks = 64;      // chunks of one connection attempt
constexpr float kFltMax = 3.402823466e38f;
// obstacle records (at most 15 floats), their cull boxes (8), 5 base poses
constexpr size_t kGraphsPerLane = 8;  // captured round sequences kept per lane
constexpr size_t kWorldFloats = (size_t)MRPOP_MAX_OBSTACLES * (15 + 8) + 5 * 12 + 64;

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Entries of the per-round edge table: a power of two of at least 2 B, so
// linear probing stays short with B insertions.
size_t edge_hash_size(int B) {
    size_t h = 64;
    while (h < 2 * (size_t)B) h <<= 1;
    return h;
}


uint64_t ns_since(std::chrono::steady_clock::time_point t) {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t)
        .count();
}

// Sizes of one slot's buffers, in the order they are carved.
struct Layout {
    std::vector<size_t> off;
    size_t total = 0;
    size_t add(size_t bytes) {
        off.push_back(total);
        total += align256(bytes);
        return off.back();
    }
};

}  // namespace

Engine::~Engine() {
    for (Lane &L : lanes_)
        if (L.stream) (void)hipStreamSynchronize(L.stream);
    for (Slot *s : slots_) {
        if (s->mem) (void)hipFree(s->mem);
        delete s;
    }
    for (Lane &L : lanes_) {
        for (void *p : {(void *)L.d_slots, (void *)L.d_ids, (void *)L.d_groups, (void *)L.d_wave, (void *)L.d_report})
            if (p) (void)hipFree(p);
        for (void *p : {(void *)L.h_slots, (void *)L.h_ids, (void *)L.h_groups, (void *)L.h_report})
            if (p) (void)hipHostFree(p);
        drop_graphs(L);
        if (L.done) (void)hipEventDestroy(L.done);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    for (auto &kv : pool_) (void)hipFree(kv.second);
    for (auto &g : graveyard_) (void)hipFree(g.second);  // the lanes' streams have drained above
    for (void *p : {(void *)d_worlds_, (void *)d_aux_slot_, (void *)d_aux_id_})
        if (p) (void)hipFree(p);
    for (void *p : {(void *)h_worlds_, (void *)h_aux_slot_})
        if (p) (void)hipHostFree(p);
}

void Engine::drop_graphs(Lane &L) {
    for (Lane::Graph &g : L.graphs) (void)hipGraphExecDestroy(g.exec);
    L.graphs.clear();
}

bool Engine::idle() const {
    for (const Slot *s : slots_)
        if (s->phase != Phase::Done) return false;
    return true;
}

int Engine::ensure_lanes() {
    if (lanes_[0].stream) return MRPOP_OK;
    for (Lane &L : lanes_) {
        // default priority: only the stream's blocking behaviour is set
        MRPOP_HIP(c_, hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
        MRPOP_HIP(c_, hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
    }
    MRPOP_HIP(c_, hipMalloc(&d_aux_slot_, sizeof(SlotDev)));
    MRPOP_HIP(c_, hipHostMalloc(&h_aux_slot_, sizeof(SlotDev), hipHostMallocDefault));
    MRPOP_HIP(c_, hipMalloc(&d_aux_id_, sizeof(int)));
    MRPOP_HIP(c_, hipMemsetAsync(d_aux_id_, 0, sizeof(int), c_->stream));
    return MRPOP_OK;
}

// Lane tables hold every slot at most once; they are reallocated only when
// no lane is running.
int Engine::ensure_capacity(int n) {
    if (int rc = ensure_lanes()) return rc;
    if (n <= cap_) return MRPOP_OK;
    if (int rc = drain()) return rc;
    const int cap = std::max(n, cap_ ? cap_ * 2 : 16);
    auto realloc_pair = [&](void **d, void **h, size_t bytes, size_t keep) -> int {
        void *nd = nullptr, *nh = nullptr;
        MRPOP_HIP(c_, hipMalloc(&nd, bytes));
        if (h) MRPOP_HIP(c_, hipHostMalloc(&nh, bytes, hipHostMallocDefault));
        if (*d) {
            if (keep) MRPOP_HIP(c_, hipMemcpy(nd, *d, keep, hipMemcpyDeviceToDevice));
            (void)hipFree(*d);
        }
        if (h && *h) {
            if (keep) std::memcpy(nh, *h, keep);
            (void)hipHostFree(*h);
        }
        *d = nd;
        if (h) *h = nh;
        return MRPOP_OK;
    };
    if (int rc = realloc_pair((void **)&d_worlds_, (void **)&h_worlds_, sizeof(dev::WorldDesc) * cap,
                              sizeof(dev::WorldDesc) * cap_))
        return rc;
    // per slot at most 2 * 5 roadmap groups, or dr + cn + sp; twice over for safety
    group_cap_ = (size_t)cap * 32;
    for (Lane &L : lanes_) {
        drop_graphs(L);  // they bake in the buffers reallocated here
        if (int rc = realloc_pair((void **)&L.d_slots, (void **)&L.h_slots, sizeof(SlotDev) * cap, 0)) return rc;
        if (int rc = realloc_pair((void **)&L.d_report, (void **)&L.h_report, sizeof(SlotReport) * cap, 0)) return rc;
        if (int rc = realloc_pair((void **)&L.d_ids, (void **)&L.h_ids, sizeof(int) * cap * 32, 0)) return rc;
        if (int rc = realloc_pair((void **)&L.d_groups, (void **)&L.h_groups, sizeof(dev::EdgeGroup) * group_cap_, 0))
            return rc;
        if (int rc = realloc_pair((void **)&L.d_wave, nullptr, sizeof(unsigned) * (group_cap_ + 1), 0)) return rc;
    }
    cap_ = cap;
    return MRPOP_OK;
}

size_t Engine::slot_bytes(int A, const mrpop_settings &st, const SolveOptions &opt) const {
    const int D = 7 * A;
    const int B = st.num_new_configs;
    const int M = kRoadmapIter * opt.block_per_robot;
    const int rm_cap = opt.roadmap_capacity;
    const int tr_cap = opt.tree_capacity > 0 ? opt.tree_capacity : std::min(st.max_samples, 262144);
    const int Kc = opt.neighbour_candidates, K = opt.neighbour_capacity;
    const size_t ce_cap = (size_t)B * kMaxConnectChunks;
    const size_t sp_cap = (size_t)kPathCap * 40;
    Layout L;
    L.add(sizeof(float) * kWorldFloats);                                    // world
    L.add(sizeof(float) * (1 + kMaxGoals) * D);                             // sg
    L.add(sizeof(float) * A * rm_cap * 8);                                  // rm_q
    L.add(sizeof(int) * (size_t)A * rm_cap * K);                            // rm_nbr
    L.add(sizeof(int) * A * rm_cap);                                        // rm_deg
    L.add(sizeof(int) * A * rm_cap);                                        // rm_uf
    L.add(sizeof(int) * 64);                                                // rm_n, rm_connected, re_count
    L.add(sizeof(float) * A * M * 7);                                       // rs_q
    L.add(sizeof(uint32_t) * A * M);                                        // rs_coll
    L.add(sizeof(int) * A * M);                                             // rs_id
    L.add(sizeof(int) * A * M * 2);                                         // rs_ncand, rs_first
    L.add(sizeof(int) * (size_t)A * M * Kc);                                // rs_cand
    const size_t rl_cap = std::min<size_t>((size_t)rm_cap, (size_t)M * Kc), rl_words = ((size_t)M + 31) / 32;
    L.add(sizeof(int) * (size_t)A * rm_cap);                                // rl_slot
    L.add(sizeof(int) * (size_t)A * rl_cap);                                // rl_touched
    L.add(sizeof(int) * 16);                                                // rs_base, rl_ntouched
    L.add(sizeof(unsigned) * (size_t)A * rl_cap * rl_words);                // rl_bits
    L.add(sizeof(float) * (size_t)A * M * Kc * 7);                          // re_from
    L.add(sizeof(float) * (size_t)A * M * Kc * 7);                          // re_to
    L.add(sizeof(int) * (size_t)A * M * Kc * 2);                            // re_pair
    L.add(sizeof(uint32_t) * (size_t)A * M * Kc);                           // re_coll
    L.add(sizeof(float) * 2 * (size_t)tr_cap * D);                          // tr_q
    L.add(sizeof(int) * 2 * (size_t)tr_cap);                                // tr_parent
    L.add(sizeof(float) * 2 * (size_t)tr_cap);                              // tr_cost
    L.add(sizeof(int) * 2 * (size_t)tr_cap * A);                            // tr_rid
    L.add(sizeof(unsigned long long) * 2 * (size_t)tr_cap);                 // rw_key
    L.add(sizeof(int) * 64);                                                // counters
    L.add(sizeof(int) * B * 6 + sizeof(float) * B);                         // ex_* ints + cost
    L.add(sizeof(float) * (size_t)B * D);                                   // ex_q
    L.add(sizeof(int) * B * A);                                             // ex_rid
    L.add(sizeof(unsigned long long) * edge_hash_size(B));                  // ex_hash
    L.add(sizeof(float) * (size_t)B * (D + 1));                             // ex_sample, ex_crand
    L.add(sizeof(unsigned long long) * B + sizeof(int) * B);                // ex_nnkey, ex_live
    L.add(sizeof(float) * (size_t)B * D * 2 + sizeof(int) * B * 2);         // ed_from/to, ed_b, ed_coll
    L.add(sizeof(unsigned long long) * B + sizeof(int) * B * 3 + sizeof(float) * B);  // cn_key, cn_*
    L.add(sizeof(float) * ce_cap * D * 2);                                  // ce_from/to
    L.add(sizeof(int) * ce_cap * 2);                                        // ce_b, ce_coll
    L.add(sizeof(float) * (size_t)kPathCap * D);                            // path
    L.add(sizeof(float) * (size_t)D * D);                                   // phs_tf
    L.add(sizeof(float) * sp_cap * D * 2);                                  // sp_from/to
    L.add(sizeof(int) * sp_cap * 2);                                        // sp_item, sp_coll
    L.add(sizeof(uint32_t) * kPathCap * 16);                                // sp_fail
    L.add(sizeof(float) * (size_t)kPathCap * D);                            // sp_mid
    L.add(sizeof(float) * kPathCap + sizeof(int) * kPathCap);               // sp_cost, sp_parent
    return L.total;
}

int Engine::validate(const mrpop_scene &scene, int A, const float *bases, const float *start, int n_goals,
                     const float *goals, const mrpop_settings &st, const SolveOptions &opt) const {
    if (A < 2 || A > geom::kMaxArms) return c_->fail(MRPOP_ERR_INVALID, "planning needs 2..5 arms (got %d)", A);
    if (n_goals < 1 || n_goals > kMaxGoals) return c_->fail(MRPOP_ERR_INVALID, "1..%d goals supported", kMaxGoals);
    if (st.granularity < 1 || st.granularity > 64) return c_->fail(MRPOP_ERR_INVALID, "granularity outside 1..64");
    if (st.num_new_configs < 1 || st.num_new_configs > 65535)
        return c_->fail(MRPOP_ERR_INVALID, "num_new_configs outside 1..65535");
    if (!(st.range > 0)) return c_->fail(MRPOP_ERR_INVALID, "range must be positive");
    if (opt.block_per_robot < 1 || opt.roadmap_capacity < 1 + n_goals + 1 || opt.neighbour_capacity < 1 ||
        opt.neighbour_candidates < 1 || opt.neighbour_candidates > 64 || opt.rounds_per_step < 1 ||
        opt.rounds_per_step > 1024 || opt.tree_capacity < 0 ||
        (opt.tree_capacity == 0 && st.max_samples < 2) || opt.probe_samples < 0 || opt.probe_samples > 64 ||
        opt.biased_iter < 0 || (opt.roadmap_cache && opt.roadmap_capacity < 4 * (1 + kMaxGoals)))
        return c_->fail(MRPOP_ERR_INVALID, "bad solve options (neighbour_candidates 1..64, rounds_per_step 1..1024, "
                                           "probe_samples 0..64, biased_iter >= 0, roadmap_capacity >= %d with "
                                           "roadmap_cache)", 4 * (1 + kMaxGoals));
    // the goal tree starts with one root per goal and the start tree with one
    // node; a capacity below n_goals + 1 would not hold the roots
    const long tr_cap = opt.tree_capacity > 0 ? opt.tree_capacity : std::min(st.max_samples, 262144);
    if (tr_cap < n_goals + 1)
        return c_->fail(MRPOP_ERR_INVALID, "tree capacity %ld below n_goals + 1 = %d", tr_cap, n_goals + 1);
    const int counts[5] = {scene.n_spheres, scene.n_capsules, scene.n_z_capsules, scene.n_cuboids, scene.n_z_cuboids};
    long n_obs = 0;
    for (int k = 0; k < 5; ++k) {
        if (counts[k] < 0) return c_->fail(MRPOP_ERR_INVALID, "negative obstacle count");
        n_obs += counts[k];
    }
    if (n_obs > MRPOP_MAX_OBSTACLES) return c_->fail(MRPOP_ERR_INVALID, "too many obstacles");
    // a NaN configuration would compare false against every distance and leave
    // nearest-node searches without a result
    auto finite = [](const float *v, long n) {
        for (long i = 0; i < n; ++i)
            if (!std::isfinite(v[i])) return false;
        return true;
    };
    if (!finite(start, 7L * A) || !finite(goals, 7L * A * n_goals) || (bases && !finite(bases, 12L * A)))
        return c_->fail(MRPOP_ERR_INVALID, "start, goals and base poses must be finite");
    return MRPOP_OK;
}

// Everything an arm's roadmap depends on: the checker resolution and edge
// range, the candidate and list sizes, the node capacity, the arm's base
// pose and every obstacle record.
std::vector<float> roadmap_key(const mrpop_scene &sc, const float *base12, const mrpop_settings &st,
                               const SolveOptions &opt) {
    std::vector<float> k = {(float)st.granularity, st.range, (float)opt.neighbour_candidates,
                            (float)opt.neighbour_capacity, (float)opt.roadmap_capacity};
    k.insert(k.end(), base12, base12 + 12);
    const int counts[5] = {sc.n_spheres, sc.n_capsules, sc.n_z_capsules, sc.n_cuboids, sc.n_z_cuboids};
    const float *recs[5] = {sc.spheres, sc.capsules, sc.z_capsules, sc.cuboids, sc.z_cuboids};
    const int width[5] = {MRPOP_SPHERE_FLOATS, MRPOP_CAPSULE_FLOATS, MRPOP_CAPSULE_FLOATS, MRPOP_CUBOID_FLOATS,
                          MRPOP_CUBOID_FLOATS};
    for (int t = 0; t < 5; ++t) {
        k.push_back((float)counts[t]);
        if (counts[t] > 0) k.insert(k.end(), recs[t], recs[t] + (size_t)counts[t] * width[t]);
    }
    return k;
}

// Copy the cached roadmap of each arm that has one into the slot (its node
// count included); the first round then inserts the query's nodes after them.
int Engine::cache_fetch(Slot *s) {
    if (!s->opt.roadmap_cache) return MRPOP_OK;
    RoadmapCache &C = c_->rm_cache;
    const int K = s->dv.rm_K;
    for (int a = 0; a < s->A; ++a) {
        auto it = C.entries.find(s->rm_key[a]);
        if (it == C.entries.end()) {
            C.misses++;
            continue;
        }
        RoadmapEntry &e = it->second;
        e.used = ++C.clock;
        C.hits++;
        const size_t off = (size_t)a * s->dv.rm_cap;
        const int n = e.n;
        MRPOP_HIP(c_, hipMemcpyAsync(s->dv.rm_q + off * 8, e.q, sizeof(float) * 8 * n, hipMemcpyDeviceToDevice,
                                     c_->stream));
        MRPOP_HIP(c_, hipMemcpyAsync(s->dv.rm_nbr + off * K, e.nbr, sizeof(int) * (size_t)n * K,
                                     hipMemcpyDeviceToDevice, c_->stream));
        MRPOP_HIP(c_, hipMemcpyAsync(s->dv.rm_deg + off, e.deg, sizeof(int) * n, hipMemcpyDeviceToDevice, c_->stream));
        MRPOP_HIP(c_, hipMemcpyAsync(s->dv.rm_uf + off, e.uf, sizeof(int) * n, hipMemcpyDeviceToDevice, c_->stream));
        s->rm_cached[a] = n;
        s->res.roadmap_cached[a] = n;
        // node count, and the id the query's start will get (the tree roots
        // are set up before the insertion round runs)
        MRPOP_HIP(c_, hipMemcpyAsync(s->dv.rm_n + a, &s->rm_cached[a], sizeof(int), hipMemcpyHostToDevice,
                                     c_->stream));
        MRPOP_HIP(c_, hipMemcpyAsync(s->dv.rm_root0 + a, &s->rm_cached[a], sizeof(int), hipMemcpyHostToDevice,
                                     c_->stream));
    }
    MRPOP_HIP(c_, hipStreamSynchronize(c_->stream));
    return MRPOP_OK;
}

// Keep the slot's roadmaps (rm_n nodes per arm) for later queries: only
// arms whose inserted start and goals were collision-free, whose roadmap
// fills at most half the capacity (a query that starts from it can still
// grow it around its own start and goals), and that have more nodes than
// the entry already kept.  The slot is on no lane.
int Engine::cache_store(Slot *s, const int *rm_n, unsigned arms) {
    // roadmaps whose union-find hit its bound are not trusted
    if (!s->opt.roadmap_cache || s->res.events[kEvUnionFind] > 0) return MRPOP_OK;
    RoadmapCache &C = c_->rm_cache;
    int qvalid[geom::kMaxArms] = {};
    MRPOP_HIP(c_, hipMemcpyAsync(qvalid, s->dv.rm_qvalid, sizeof(int) * s->A, hipMemcpyDeviceToHost, c_->stream));
    MRPOP_HIP(c_, hipStreamSynchronize(c_->stream));
    const int K = s->dv.rm_K;
    std::vector<void *> old;
    for (int a = 0; a < s->A; ++a) {
        const int n = rm_n[a];
        if (!((arms >> a) & 1u) || !qvalid[a] || n <= 0 || n > s->dv.rm_cap / 2) continue;
        auto it = C.entries.find(s->rm_key[a]);
        if (it != C.entries.end() && it->second.n >= n) continue;
        RoadmapEntry e;
        e.n = n;
        e.K = K;
        const size_t bq = align256(sizeof(float) * 8 * n), bn = align256(sizeof(int) * (size_t)n * K),
                     bi = align256(sizeof(int) * n);
        e.bytes = bq + bn + 2 * bi;
        if (e.bytes > C.budget) continue;
        if (hipMalloc(&e.mem, e.bytes) != hipSuccess) {
            (void)hipGetLastError();
            continue;  // no room: the roadmap is not kept
        }
        char *m = (char *)e.mem;
        e.q = (float *)m;
        e.nbr = (int *)(m + bq);
        e.deg = (int *)(m + bq + bn);
        e.uf = (int *)(m + bq + bn + bi);
        const size_t off = (size_t)a * s->dv.rm_cap;
        MRPOP_HIP(c_, hipMemcpyAsync(e.q, s->dv.rm_q + off * 8, sizeof(float) * 8 * n, hipMemcpyDeviceToDevice,
                                     c_->stream));
        MRPOP_HIP(c_, hipMemcpyAsync(e.nbr, s->dv.rm_nbr + off * K, sizeof(int) * (size_t)n * K,
                                     hipMemcpyDeviceToDevice, c_->stream));
        MRPOP_HIP(c_, hipMemcpyAsync(e.deg, s->dv.rm_deg + off, sizeof(int) * n, hipMemcpyDeviceToDevice, c_->stream));
        MRPOP_HIP(c_, hipMemcpyAsync(e.uf, s->dv.rm_uf + off, sizeof(int) * n, hipMemcpyDeviceToDevice, c_->stream));
        e.used = ++C.clock;
        if (it != C.entries.end()) {
            C.bytes -= it->second.bytes;
            old.push_back(it->second.mem);
            it->second = e;
        } else {
            C.entries.emplace(s->rm_key[a], e);
        }
        C.bytes += e.bytes;
        C.stored++;
    }
    // least recently used entries leave beyond the budget
    while (C.bytes > C.budget && !C.entries.empty()) {
        auto lru = C.entries.begin();
        for (auto it = C.entries.begin(); it != C.entries.end(); ++it)
            if (it->second.used < lru->second.used) lru = it;
        C.bytes -= lru->second.bytes;
        old.push_back(lru->second.mem);
        C.entries.erase(lru);
    }
    MRPOP_HIP(c_, hipStreamSynchronize(c_->stream));
    for (void *p : old) (void)hipFree(p);
    return MRPOP_OK;
}

int Engine::add(const mrpop_scene &scene, int A, const float *bases, const float *start, int n_goals,
                const float *goals, unsigned long long seed, const mrpop_settings &st, const SolveOptions &opt,
                Slot **out) {
    *out = nullptr;
    if (int rc = validate(scene, A, bases, start, n_goals, goals, st, opt)) return rc;
    const float *b = bases ? bases : &geom::default_bases(A)[0][0];

    auto *s = new Slot();
    s->A = A;
    s->D = 7 * A;
    s->n_goals = n_goals;
    s->st = st;
    s->opt = opt;
    s->seed = seed;
    s->sg.assign(start, start + s->D);
    s->sg.insert(s->sg.end(), goals, goals + (size_t)n_goals * s->D);
    s->t_submit = std::chrono::steady_clock::now();

    // slot id
    int id = -1;
    std::vector<bool> used(cap_, false);
    for (Slot *o : slots_) used[o->id] = true;
    for (int i = 0; i < cap_; ++i)
        if (!used[i]) {
            id = i;
            break;
        }
    if (id < 0) {
        if (int rc = ensure_capacity(cap_ + 1)) {
            delete s;
            return rc;
        }
        std::vector<bool> u2(cap_, false);
        for (Slot *o : slots_) u2[o->id] = true;
        for (int i = 0; i < cap_; ++i)
            if (!u2[i]) {
                id = i;
                break;
            }
    }
    s->id = id;

    // ---- device memory
    const int D = s->D, B = st.num_new_configs, M = kRoadmapIter * opt.block_per_robot;
    const int rm_cap = opt.roadmap_capacity, K = opt.neighbour_capacity, Kc = opt.neighbour_candidates;
    const int tr_cap = opt.tree_capacity > 0 ? opt.tree_capacity : std::min(st.max_samples, 262144);
    const size_t ce_cap = (size_t)B * kMaxConnectChunks, sp_cap = (size_t)kPathCap * 40;
    s->mem_bytes = slot_bytes(A, st, opt);
    // reuse a released block of the same size (problems with the same arm
    // count and settings have identical layouts), else allocate
    auto pooled = pool_.find(s->mem_bytes);
    if (pooled != pool_.end()) {
        s->mem = pooled->second;
        pool_bytes_ -= s->mem_bytes;
        pool_.erase(pooled);
    } else if (hipMalloc(&s->mem, s->mem_bytes) != hipSuccess && (release_pool(), true) &&
               hipMalloc(&s->mem, s->mem_bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete s;
        return c_->fail(MRPOP_ERR_BUDGET, "device memory for a slot (%zu bytes) not available", s->mem_bytes);
    }
    bytes_in_use_ += s->mem_bytes;
    char *base = (char *)s->mem;
    size_t cur = 0;
    auto take = [&](size_t bytes) {
        void *p = base + cur;
        cur += align256(bytes);
        return p;
    };
    SlotDev &v = s->dv;
    std::memset(&v, 0, sizeof v);
    s->wbuf = (float *)take(sizeof(float) * kWorldFloats);
    v.sg = (float *)take(sizeof(float) * (1 + kMaxGoals) * D);
    v.rm_q = (float *)take(sizeof(float) * A * rm_cap * 8);
    v.rm_nbr = (int *)take(sizeof(int) * (size_t)A * rm_cap * K);
    v.rm_deg = (int *)take(sizeof(int) * A * rm_cap);
    v.rm_uf = (int *)take(sizeof(int) * A * rm_cap);
    int *small = (int *)take(sizeof(int) * 64);
    v.rm_n = small;
    v.rm_connected = small + 8;
    v.re_count = small + 16;
    v.rs_nsamp = small + 24;
    v.rs_brounds = small + 32;
    v.rm_root0 = small + 40;
    v.rm_qvalid = small + 48;
    v.rs_q = (float *)take(sizeof(float) * A * M * 7);
    v.rs_coll = (uint32_t *)take(sizeof(uint32_t) * A * M);
    v.rs_id = (int *)take(sizeof(int) * A * M);
    v.rs_ncand = (int *)take(sizeof(int) * A * M * 2);
    v.rs_first = v.rs_ncand + A * M;
    v.rs_cand = (int *)take(sizeof(int) * (size_t)A * M * Kc);
    const size_t rl_cap = std::min<size_t>((size_t)rm_cap, (size_t)M * Kc), rl_words = ((size_t)M + 31) / 32;
    v.rl_cap = (int)rl_cap;
    v.rl_words = (int)rl_words;
    v.rl_slot = (int *)take(sizeof(int) * (size_t)A * rm_cap);
    v.rl_touched = (int *)take(sizeof(int) * (size_t)A * rl_cap);
    v.rs_base = (int *)take(sizeof(int) * 16);
    v.rl_ntouched = v.rs_base + 8;
    v.rl_bits = (unsigned *)take(sizeof(unsigned) * (size_t)A * rl_cap * rl_words);
    v.re_from = (float *)take(sizeof(float) * (size_t)A * M * Kc * 7);
    v.re_to = (float *)take(sizeof(float) * (size_t)A * M * Kc * 7);
    v.re_pair = (int *)take(sizeof(int) * (size_t)A * M * Kc * 2);
    v.re_coll = (uint32_t *)take(sizeof(uint32_t) * (size_t)A * M * Kc);
    v.tr_q = (float *)take(sizeof(float) * 2 * (size_t)tr_cap * D);
    v.tr_parent = (int *)take(sizeof(int) * 2 * (size_t)tr_cap);
    v.tr_cost = (float *)take(sizeof(float) * 2 * (size_t)tr_cap);
    v.tr_rid = (int *)take(sizeof(int) * 2 * (size_t)tr_cap * A);
    v.rw_key = (unsigned long long *)take(sizeof(unsigned long long) * 2 * (size_t)tr_cap);
    int *ctr = (int *)take(sizeof(int) * 64);
    v.tr_n = ctr;
    v.ed_count = ctr + 2;
    v.ce_count = ctr + 3;
    v.status = ctr + 4;
    v.best_cost = (float *)(ctr + 5);
    v.path_cost = (float *)(ctr + 6);
    v.path_n = ctr + 7;
    v.path_meta = ctr + 8;
    v.sp_count = ctr + 12;
    v.ctr = ctr + 16;
    v.events = ctr + 24;
    int *ex = (int *)take(sizeof(int) * B * 6 + sizeof(float) * B);
    v.ex_tree = ex;
    v.ex_state = ex + B;
    v.ex_near = ex + 2 * B;
    v.ex_exist = ex + 3 * B;
    v.ex_index = ex + 4 * B;
    v.ex_dup = ex + 5 * B;
    v.ex_cost = (float *)(ex + 6 * B);
    v.ex_q = (float *)take(sizeof(float) * (size_t)B * D);
    v.ex_rid = (int *)take(sizeof(int) * B * A);
    v.ex_hash = (unsigned long long *)take(sizeof(unsigned long long) * edge_hash_size(B));
    v.ex_hmask = (int)edge_hash_size(B) - 1;
    v.ex_sample = (float *)take(sizeof(float) * (size_t)B * (D + 1));
    v.ex_crand = v.ex_sample + (size_t)B * D;
    v.ex_nnkey = (unsigned long long *)take(sizeof(unsigned long long) * B + sizeof(int) * B);
    v.ex_live = (int *)(v.ex_nnkey + B);
    float *ed = (float *)take(sizeof(float) * (size_t)B * D * 2 + sizeof(int) * B * 2);
    v.ed_from = ed;
    v.ed_to = ed + (size_t)B * D;
    v.ed_b = (int *)(ed + (size_t)B * D * 2);
    v.ed_coll = (uint32_t *)(v.ed_b + B);
    v.cn_key = (unsigned long long *)take(sizeof(unsigned long long) * B + sizeof(int) * B * 3 + sizeof(float) * B);
    int *cn = (int *)(v.cn_key + B);
    v.cn_near = cn;
    v.cn_next = cn + B;
    v.cn_fail = (uint32_t *)(cn + 2 * B);
    v.cn_d = (float *)(cn + 3 * B);
    float *ce = (float *)take(sizeof(float) * ce_cap * D * 2);
    v.ce_from = ce;
    v.ce_to = ce + ce_cap * D;
    int *ceb = (int *)take(sizeof(int) * ce_cap * 2);
    v.ce_b = ceb;
    v.ce_coll = (uint32_t *)(ceb + ce_cap);
    v.path = (float *)take(sizeof(float) * (size_t)kPathCap * D);
    v.phs_tf = (float *)take(sizeof(float) * (size_t)D * D);
    float *sp = (float *)take(sizeof(float) * sp_cap * D * 2);
    v.sp_from = sp;
    v.sp_to = sp + sp_cap * D;
    int *spi = (int *)take(sizeof(int) * sp_cap * 2);
    v.sp_item = spi;
    v.sp_coll = (uint32_t *)(spi + sp_cap);
    v.sp_fail = (uint32_t *)take(sizeof(uint32_t) * kPathCap * 16);
    v.sp_mid = (float *)take(sizeof(float) * (size_t)kPathCap * D);
    float *spc = (float *)take(sizeof(float) * kPathCap + sizeof(int) * kPathCap);
    v.sp_cost = spc;
    v.sp_parent = (int *)(spc + kPathCap);

    v.A = A;
    v.D = D;
    v.n_goals = n_goals;
    v.world = id;
    v.seed = seed;
    v.G = st.granularity;
    v.rm_reassign = st.roadmap_reassign ? 1 : 0;
    v.rm_biased_iter = opt.biased_iter;
    v.rm_stretch = 1;
    // roadmap reuse: the query's nodes are inserted by the first round, and
    // sampled rounds leave room for the nodes of the queries that reuse the
    // roadmap later
    v.rm_insert = opt.roadmap_cache ? 1 : 0;
    v.rm_fill = opt.roadmap_cache ? rm_cap - (1 + kMaxGoals) : rm_cap;
    if (opt.roadmap_cache)
        for (int a = 0; a < A; ++a) s->rm_key[a] = roadmap_key(scene, b + 12 * a, st, opt);
    v.B = B;
    v.M = M;
    v.range = st.range;
    v.tree_ratio = st.tree_ratio;
    v.balance = st.balance;
    v.phs = st.phs;
    v.rm_cap = rm_cap;
    v.rm_K = K;
    v.rm_Kc = Kc;
    v.tr_cap = tr_cap;
    v.ce_cap = (int)ce_cap;
    v.path_cap = kPathCap;
    v.sp_cap = (int)sp_cap;

    // ---- world (same layout as mrpop_world_create)
    std::vector<float> host;
    dev::WorldDesc &w = s->wdesc;
    pack_world(scene, b, A, w, host);
    w.obs = s->wbuf;
    MRPOP_HIP(c_, hipMemcpyAsync(s->wbuf, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, c_->stream));
    MRPOP_HIP(c_, hipMemcpyAsync(v.sg, s->sg.data(), s->sg.size() * sizeof(float), hipMemcpyHostToDevice, c_->stream));
    MRPOP_HIP(c_, hipMemsetAsync(v.rw_key, 0xff, sizeof(unsigned long long) * 2 * (size_t)tr_cap, c_->stream));
    MRPOP_HIP(c_, hipMemsetAsync(ctr, 0, sizeof(int) * 64, c_->stream));
    MRPOP_HIP(c_, hipMemsetAsync(v.ex_hash, 0, sizeof(unsigned long long) * edge_hash_size(B), c_->stream));
    MRPOP_HIP(c_, hipMemsetAsync(v.rl_slot, 0xff, sizeof(int) * (size_t)A * rm_cap, c_->stream));
    MRPOP_HIP(c_, hipMemsetAsync(v.rl_bits, 0, sizeof(unsigned) * (size_t)A * rl_cap * rl_words, c_->stream));
    {
        const float fm = kFltMax;
        MRPOP_HIP(c_, hipMemcpyAsync(v.best_cost, &fm, sizeof(float), hipMemcpyHostToDevice, c_->stream));
    }

    // ---- edge groups (pointers and counts are fixed for the slot's life)
    for (int a = 0; a < A; ++a) {
        s->g_valid[a] = dev::EdgeGroup{v.rs_q + (size_t)a * M * 7, v.rs_q + (size_t)a * M * 7, 7, 1, M, a,
                                       v.rs_nsamp + a,
                                       v.rs_coll + (size_t)a * M, nullptr, id, 0};
        s->g_rm[a] = dev::EdgeGroup{v.re_from + (size_t)a * M * Kc * 7, v.re_to + (size_t)a * M * Kc * 7, 7,
                                    st.granularity, M * Kc, a, v.re_count + a, v.re_coll + (size_t)a * M * Kc, nullptr,
                                    id, 0};
    }
    // groups with a device count carry their capacity in n_edges (a bound for the grid)
    s->g_dr = dev::EdgeGroup{v.ed_from, v.ed_to, D, st.granularity, B, 0, v.ed_count, v.ed_coll, nullptr, id, 0};
    s->g_cn = dev::EdgeGroup{v.ce_from, v.ce_to, D, st.granularity, (int)ce_cap, 0, v.ce_count, v.ce_coll, nullptr,
                             id, 0};
    s->g_sp = dev::EdgeGroup{v.sp_from, v.sp_to, D, st.granularity, (int)sp_cap, 0, v.sp_count, v.sp_coll, nullptr,
                             id, 0};

    h_worlds_[id] = w;
    MRPOP_HIP(c_, hipMemcpyAsync(d_worlds_ + id, &h_worlds_[id], sizeof(dev::WorldDesc), hipMemcpyHostToDevice, c_->stream));
    if (int rc = run_aux(s, plan::launch_roadmap_init)) return rc;
    if (int rc = cache_fetch(s)) return rc;
    slots_.push_back(s);
    s->t_kernel_start = std::chrono::steady_clock::now();
    if (int rc = start_pass(s)) return rc;
    *out = s;
    return MRPOP_OK;
}

void Engine::release_pool() {
    (void)hipGetLastError();
    for (auto &kv : pool_) (void)hipFree(kv.second);
    pool_.clear();
    pool_bytes_ = 0;
}

void Engine::release(Slot *s) {
    auto it = std::find(slots_.begin(), slots_.end(), s);
    if (it == slots_.end()) return;
    if (s->lane >= 0 && lanes_[s->lane].abandoned) {
        // its lane's work may still use the memory and may never finish:
        // keep the block aside until the lane rejoins, instead of waiting
        Lane &L = lanes_[s->lane];
        for (Slot *&t : L.table)
            if (t == s) t = nullptr;
        graveyard_.emplace_back(s->lane, s->mem);
        bytes_in_use_ -= s->mem_bytes;
        slots_.erase(it);
        delete s;
        return;
    }
    if (s->lane >= 0) {
        // its lane may still run kernels on the slot's memory; its report is dropped
        Lane &L = lanes_[s->lane];
        (void)hipStreamSynchronize(L.stream);
        for (Slot *&t : L.table)
            if (t == s) t = nullptr;
    }
    (void)hipStreamSynchronize(c_->stream);
    if (s->mem) {
        bytes_in_use_ -= s->mem_bytes;
        if (pool_bytes_ + s->mem_bytes <= kPoolBytes) {
            pool_.emplace(s->mem_bytes, s->mem);  // kept for the next problem of this size
            pool_bytes_ += s->mem_bytes;
        } else {
            (void)hipFree(s->mem);
        }
    }
    slots_.erase(it);
    delete s;
}

// PHS transform (pop_planner.cu:289-346): Householder rotation taking e_1 to
// the unit start->goal direction, scaled by the hyperspheroid radii
// (c_best / 2, sqrt(c_best^2 - c_min^2) / 2, ...); the centre is the
// midpoint of the foci.  Host code only (no device call).
void phs_transform(int D, const float *f1, const float *f2, double cb, float *tf, float *center) {
    double cmin = 0;
    for (int d = 0; d < D; ++d) cmin += (double)(f2[d] - f1[d]) * (f2[d] - f1[d]);
    cmin = std::sqrt(cmin);
    std::vector<double> rot((size_t)D * D, 0.0);
    if (cmin < 1e-6) {
        for (int d = 0; d < D; ++d) rot[(size_t)d * D + d] = 1.0;
    } else {
        std::vector<double> ax(D), vv(D);
        for (int d = 0; d < D; ++d) ax[d] = (f2[d] - f1[d]) / cmin;
        double vtv = 0;
        for (int d = 0; d < D; ++d) {
            vv[d] = d == 0 ? 1.0 - ax[0] : -ax[d];
            vtv += vv[d] * vv[d];
        }
        const double beta = vtv > 0 ? 2.0 / vtv : 0.0;
        for (int r = 0; r < D; ++r)
            for (int c = 0; c < D; ++c) rot[(size_t)r * D + c] = (r == c ? 1.0 : 0.0) - beta * vv[r] * vv[c];
    }
    const double conj = std::sqrt(std::max(0.0, cb * cb - cmin * cmin));
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) tf[(size_t)r * D + c] = (float)(rot[(size_t)r * D + c] * (c == 0 ? cb / 2 : conj / 2));
    for (int d = 0; d < D; ++d) center[d] = (f1[d] + f2[d]) / 2.0f;
}

// The slot's transform, from its start and goal 0 and its best cost.
void Engine::compute_phs(Slot *s) {
    const int D = s->D;
    std::vector<float> tf((size_t)D * D);
    phs_transform(D, s->sg.data(), s->sg.data() + D, s->best_cost, tf.data(), s->dv.phs_center);
    (void)hipMemcpy(s->dv.phs_tf, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice);
}

// A pass optimises (cost bounds, hyperspheroid samples) only once a solution
// exists.  Pass 0 of the reference repeats until it is solved
// (pop_planner.cu:967-986); here pass 0 can also end unsolved when a tree
// fills up, and the next pass then searches for a first solution again
// instead of optimising against an infinite cost bound.
int Engine::start_pass(Slot *s) {
    s->seeking = !(s->best_cost < kFltMax);
    s->dv.optimize = s->seeking ? 0 : 1;
    s->dv.pass = s->pass;
    s->pass_solved = false;
    s->rm_until_connected = s->seeking;
    s->phase = Phase::Roadmap;
    s->dr_round = 0;
    s->dr_budget = s->seeking ? s->st.max_iters : s->st.optimize_iters;
    if (s->dv.optimize && s->st.phs) compute_phs(s);
    return run_aux(s, plan::launch_tree_reset);
}

// Set-up kernels of one slot that is on no lane, on the context stream.
int Engine::run_aux(Slot *s, void (*launch)(SlotDev *, const int *, int, hipStream_t)) {
    *h_aux_slot_ = s->dv;
    MRPOP_HIP(c_, hipMemcpyAsync(d_aux_slot_, h_aux_slot_, sizeof(SlotDev), hipMemcpyHostToDevice, c_->stream));
    launch(d_aux_slot_, d_aux_id_, 1, c_->stream);
    MRPOP_HIP(c_, hipGetLastError());
    MRPOP_HIP(c_, hipStreamSynchronize(c_->stream));  // the staging slot is reused by the next call
    return MRPOP_OK;
}

int *Engine::upload_ids(Lane &L, const std::vector<Slot *> &list) {
    int *h = L.h_ids + L.id_cursor;
    for (size_t i = 0; i < list.size(); ++i) {
        int k = 0;
        while (L.table[k] != list[i]) ++k;  // entry of the slot in the lane table
        h[i] = k;
    }
    int *d = L.d_ids + L.id_cursor;
    (void)hipMemcpyAsync(d, h, sizeof(int) * list.size(), hipMemcpyHostToDevice, L.stream);
    L.id_cursor += list.size();
    return d;
}

// Copy the edge-group descriptors of `list` (per_slot groups each) to the
// device as one bucket (one launch).  The descriptors
// hold device pointers and counts only, so one copy serves every round of a
// step.
int Engine::stage_groups(Lane &L, int mode, const std::vector<const dev::EdgeGroup *> &groups, int per_slot,
                         const std::vector<Slot *> &list, Staged &out) {
    out.mode = mode;
    out.buckets.clear();
    const size_t begin = L.group_cursor;
    // one launch: SINGLE mode checks one arm per group; the multi-arm modes
    // use the kernel of the arm count when all problems share it, else the
    // kernel that mixes arm counts (check_kernel_mixed)
    int A_all = mode == MRPOP_MODE_SINGLE ? 1 : (list.empty() ? 0 : list[0]->A);
    bool split = false;  // a launch per arm count instead of the mixed kernel
    for (Slot *s : list) {
        if (mode != MRPOP_MODE_SINGLE && s->A != A_all) A_all = 0;
        split |= s->opt.arm_split != 0;
    }
    // probe pass: the largest probe any of the problems asks for; it is used
    // only if some group has more samples per edge than the probe
    int probe = 0, max_G = 0;
    for (Slot *s : list) probe = std::max(probe, s->opt.probe_samples);
    auto bucket = [&](int A_bucket, bool mixed, auto &&take) -> int {
        Staged::Bucket bk{A_bucket, L.group_cursor, 0, 0, 0};
        for (size_t i = 0; i < list.size(); ++i) {
            if (!take(list[i])) continue;
            const int n = mode == MRPOP_MODE_SINGLE ? list[i]->A : 1;
            const int S = mode == MRPOP_MODE_SINGLE ? 64 : dev::tile_samples(list[i]->A, mixed);
            for (int k = 0; k < n && k < per_slot; ++k) {
                if (L.group_cursor >= group_cap_) return c_->fail(MRPOP_ERR_BUDGET, "edge-group staging full");
                const dev::EdgeGroup &g = groups[i][k];
                L.h_groups[L.group_cursor++] = g;
                bk.max_waves += ((size_t)g.n_edges * g.G + S - 1) / S;
                if (probe > 0) bk.probe_waves += ((size_t)g.n_edges * std::min(probe, g.G) + S - 1) / S;
                max_G = std::max(max_G, g.G);
            }
        }
        bk.n = (int)(L.group_cursor - bk.first);
        if (bk.n) out.buckets.push_back(bk);
        return MRPOP_OK;
    };
    if (A_all != 0 || !split) {
        if (int rc = bucket(A_all, A_all == 0, [](Slot *) { return true; })) return rc;
    } else {
        for (int A = 2; A <= geom::kMaxArms; ++A)
            if (int rc = bucket(A, false, [A](Slot *s) { return s->A == A; })) return rc;
    }
    out.probe = max_G > probe ? probe : 0;
    if (L.group_cursor > begin)
        MRPOP_HIP(c_, hipMemcpyAsync(L.d_groups + begin, L.h_groups + begin,
                                     sizeof(dev::EdgeGroup) * (L.group_cursor - begin), hipMemcpyHostToDevice, L.stream));
    return MRPOP_OK;
}

// One batched check launch per bucket.  Every producer of planner
// edges clears their verdict words, so the launches skip the zeroing pass.
// With a probe, a first launch checks `probe` samples per edge and a second
// checks every sample of the edges the first left free: most planner edges
// that collide are found by the probe, and their remaining samples, which
// cannot change the verdict, are not checked.
int Engine::launch_staged(Lane &L, const Staged &st) {
    for (const Staged::Bucket &bk : st.buckets) {
        if (st.probe > 0) {
            if (int rc = launch_check_groups(c_, st.mode, bk.A, d_worlds_, L.d_groups + bk.first, bk.n, L.d_wave,
                                             bk.probe_waves, st.sparse, false, st.probe, L.stream))
                return rc;
        }
        if (int rc = launch_check_groups(c_, st.mode, bk.A, d_worlds_, L.d_groups + bk.first, bk.n, L.d_wave,
                                         bk.max_waves, st.sparse, false, st.probe > 0 ? -1 : 0, L.stream))
            return rc;
    }
    return MRPOP_OK;
}

// Enqueue one step of a lane for the slots in L.rm / L.dr / L.sp.
int Engine::launch_lane(Lane &L) {
    L.table.clear();
    for (auto *lst : {&L.rm, &L.dr, &L.sp}) L.table.insert(L.table.end(), lst->begin(), lst->end());
    if (L.table.empty()) return MRPOP_OK;
    hipStream_t st = L.stream;
    L.id_cursor = 0;
    L.group_cursor = 0;
    for (size_t k = 0; k < L.table.size(); ++k) {
        L.h_slots[k] = L.table[k]->dv;
        L.table[k]->lane = (int)(&L - lanes_);
    }
    MRPOP_HIP(c_, hipMemcpyAsync(L.d_slots, L.h_slots, sizeof(SlotDev) * L.table.size(), hipMemcpyHostToDevice, st));
    SlotDev *ds = L.d_slots;
    // ---- roadmap round (roadmap_interior.cuh:94-339)
    if (!L.rm.empty()) {
        int *d = upload_ids(L, L.rm);
        const int n = (int)L.rm.size();
        int max_am = 0;
        for (Slot *s : L.rm) max_am = std::max(max_am, s->A * s->dv.M);
        plan::launch_rm_sample(ds, d, n, max_am, st);
        std::vector<const dev::EdgeGroup *> gv, gr;
        for (Slot *s : L.rm) {
            gv.push_back(s->g_valid);
            gr.push_back(s->g_rm);
        }
        Staged sv, sr;
        if (int rc = stage_groups(L, MRPOP_MODE_SINGLE, gv, geom::kMaxArms, L.rm, sv)) return rc;
        if (int rc = stage_groups(L, MRPOP_MODE_SINGLE, gr, geom::kMaxArms, L.rm, sr)) return rc;
        if (int rc = launch_staged(L, sv)) return rc;
        plan::launch_rm_assign(ds, d, n, st);
        plan::launch_rm_radius(ds, d, n, max_am, st);
        plan::launch_rm_compact(ds, d, n, max_am, st);
        if (int rc = launch_staged(L, sr)) return rc;
        plan::launch_rm_link(ds, d, n, max_am, st);
        plan::launch_rm_connected(ds, d, n, st);
    }
    // ---- dRRT rounds (pop_planner.cu:350-736).  The round index lives on the
    // device; a slot whose pass is solved or whose launch is out of rounds
    // skips the remaining rounds of the step.
    if (!L.dr.empty()) {
        int *d = upload_ids(L, L.dr);
        const int n = (int)L.dr.size();
        int max_b = 0, rounds = 1 << 30, tree_n = 0;
        bool reset = false;
        for (Slot *s : L.dr) {
            max_b = std::max(max_b, s->dv.B);
            rounds = std::min(rounds, std::max(1, s->opt.rounds_per_step));
            if (s->dv.round_reset) s->tree_n = 0;
            tree_n = std::max(tree_n, s->tree_n);
            reset |= s->dv.round_reset != 0;
            s->dv.round_reset = 0;  // applied by this step's k_dr_begin
        }
        // Large trees: the tree scans of the expansions and connections run
        // a thread per expansion, each block over a slice of the trees, with
        // about 2048 blocks in all and slices of at least expand_slice_rows
        // rows.  The slice count only changes the speed.
        int splits = 0;
        if (c_->expand_split_nodes >= 0 && tree_n >= c_->expand_split_nodes) {
            const int groups = (max_b + 63) / 64 * n;
            const int rows = c_->expand_slice_rows;
            splits = std::max(1, std::min((2048 + groups - 1) / groups, (tree_n + rows - 1) / rows));
        }
        std::vector<const dev::EdgeGroup *> g1, g2;
        for (Slot *s : L.dr) {
            g1.push_back(&s->g_dr);
            g2.push_back(&s->g_cn);
        }
        Staged s1, s2;
        if (int rc = stage_groups(L, MRPOP_MODE_DRRT, g1, 1, L.dr, s1)) return rc;
        if (int rc = stage_groups(L, MRPOP_MODE_FULL, g2, 1, L.dr, s2)) return rc;
        s2.sparse = true;  // connection chunks: sized for 64 per expansion
        if (reset) plan::launch_dr_begin(ds, d, n, st);
        auto enqueue_rounds = [&]() -> int {
            for (int r = 0; r < rounds; ++r) {
                plan::launch_dr_expand(ds, d, n, max_b, splits, st);
                if (int rc = launch_staged(L, s1)) return rc;
                plan::launch_dr_commit(ds, d, n, st);
                plan::launch_dr_connect(ds, d, n, max_b, splits, st);
                if (int rc = launch_staged(L, s2)) return rc;
                plan::launch_dr_finish(ds, d, n, st);
            }
            return MRPOP_OK;
        };
        bool graph = true;
        for (Slot *s : L.dr) graph &= s->opt.graph_rounds != 0;
        if (!graph) {
            if (int rc = enqueue_rounds()) return rc;
        } else {
            // The rounds read every changing value (slot table, ids, group
            // descriptors, counts) from device buffers refreshed above, so a
            // sequence captured once replays for as long as the launch
            // arguments it bakes in stay the same.  It holds kernel nodes only.
            std::vector<uint64_t> key = {(uint64_t)(uintptr_t)ds, (uint64_t)(uintptr_t)d, (uint64_t)n,
                                         (uint64_t)max_b, (uint64_t)rounds, (uint64_t)splits, (uint64_t)(uintptr_t)L.d_wave,
                                         (uint64_t)(uintptr_t)d_worlds_};
            for (const Staged *sg : {&s1, &s2}) {
                key.push_back((uint64_t)sg->mode);
                key.push_back((uint64_t)sg->probe);
                for (const Staged::Bucket &bk : sg->buckets)
                    key.insert(key.end(), {(uint64_t)bk.A, (uint64_t)bk.first, (uint64_t)bk.n, (uint64_t)bk.max_waves,
                                           (uint64_t)bk.probe_waves});
            }
            Lane::Graph *hit = nullptr;
            for (Lane::Graph &g : L.graphs)
                if (g.key == key) hit = &g;
            if (!hit) {
                if (L.graphs.size() >= kGraphsPerLane) {
                    auto lru = std::min_element(L.graphs.begin(), L.graphs.end(),
                                                [](const Lane::Graph &a, const Lane::Graph &b) { return a.used < b.used; });
                    (void)hipGraphExecDestroy(lru->exec);
                    L.graphs.erase(lru);
                }
                MRPOP_HIP(c_, hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
                const int rc = enqueue_rounds();
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(st, &g);
                if (rc || e != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    if (rc) return rc;
                    return c_->fail(MRPOP_ERR_HIP, "round capture: %s", hipGetErrorString(e));
                }
                hipGraphExec_t exec = nullptr;
                const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) return c_->fail(MRPOP_ERR_HIP, "round graph: %s", hipGetErrorString(ei));
                L.graphs.push_back(Lane::Graph{std::move(key), exec, 0});
                hit = &L.graphs.back();
            }
            hit->used = ++L.graph_clock;
            MRPOP_HIP(c_, hipGraphLaunch(hit->exec, st));
        }
    }
    // ---- simplification stage (path_simplification.cu:41-238): slots at the
    // same stage with the same settings share one launch sequence
    if (!L.sp.empty()) {
        std::vector<Slot *> part = L.sp;
        while (!part.empty()) {
            Slot *s0 = part.front();
            std::vector<Slot *> same, rest;
            for (Slot *s : part)
                ((s->sp_stage == s0->sp_stage && s->st.bspline_step == s0->st.bspline_step &&
                  s->st.range == s0->st.range && s->st.bspline_min_change == s0->st.bspline_min_change)
                     ? same
                     : rest)
                    .push_back(s);
            part.swap(rest);
            int *d = upload_ids(L, same);
            const int n = (int)same.size();
            const int stage = s0->sp_stage;
            plan::launch_sp_edges(ds, d, n, stage, s0->st.bspline_step, s0->st.range, s0->st.bspline_min_change, st);
            std::vector<const dev::EdgeGroup *> gs;
            for (Slot *s : same) gs.push_back(&s->g_sp);
            Staged ss;
            if (int rc = stage_groups(L, MRPOP_MODE_FULL, gs, 1, same, ss)) return rc;
            ss.sparse = true;  // simplification chunks: sized for the longest path
            if (int rc = launch_staged(L, ss)) return rc;
            plan::launch_sp_collect(ds, d, n, st);
            if (stage < s0->st.bspline_step) plan::launch_sp_smooth_apply(ds, d, n, st);
            else plan::launch_sp_shortcut(ds, d, n, st);
        }
    }
    // ---- reports
    {
        int *d = upload_ids(L, L.table);
        plan::launch_report(ds, d, (int)L.table.size(), L.d_report, st);
        MRPOP_HIP(c_, hipMemcpyAsync(L.h_report, L.d_report, sizeof(SlotReport) * L.table.size(),
                                     hipMemcpyDeviceToHost, st));
    }
    MRPOP_HIP(c_, hipGetLastError());
    MRPOP_HIP(c_, hipEventRecord(L.done, st));
    L.busy = true;
    L.t_launch = std::chrono::steady_clock::now();
    return MRPOP_OK;
}

// Advance the state machines of a lane's slots from its reports.
int Engine::process_lane(Lane &L) {
    L.busy = false;
    const auto now = std::chrono::steady_clock::now();
    for (size_t k = 0; k < L.table.size(); ++k) {
        Slot *s = L.table[k];
        if (!s) continue;  // released while running
        s->lane = -1;
        if (s->phase == Phase::Done) continue;  // cancelled while running
        const SlotReport &r = L.h_report[k];
        for (int e = 0; e < kEvents; ++e) s->res.events[e] = r.events[e];
        const bool in_rm = std::find(L.rm.begin(), L.rm.end(), s) != L.rm.end();
        const bool in_dr = !in_rm && std::find(L.dr.begin(), L.dr.end(), s) != L.dr.end();
        if (in_rm) {
            s->dv.rm_round++;
            s->res.roadmap_rounds++;
            s->dv.rm_insert = 0;  // the query's nodes are in
            bool all_connected = true, at_cap = false;
            for (int a = 0; a < s->A; ++a) {
                all_connected &= r.rm_connected[a] != 0;
                at_cap |= r.rm_n[a] >= s->dv.rm_fill;
                s->res.roadmap_size[a] = r.rm_n[a];
            }
            if (s->opt.roadmap_cache) {
                // an arm's roadmap is kept when it first connects the query's start and goal
                unsigned now_connected = 0;
                for (int a = 0; a < s->A; ++a)
                    if (r.rm_connected[a] && !((s->rm_kept >> a) & 1u)) now_connected |= 1u << a;
                if (now_connected) {
                    if (int rc = cache_store(s, r.rm_n, now_connected)) return rc;
                    s->rm_kept |= now_connected;
                }
            }
            if (!s->rm_until_connected || all_connected || at_cap) {
                s->rm_until_connected = false;
                s->dv.rm_stretch = 0;  // the first stretch of roadmap rounds is over for good
                s->phase = Phase::Search;
                s->dr_round = 0;
                s->dv.round_reset = 1;  // a fresh dRRT launch
                s->dv.round_limit = s->dr_budget;
            }
        } else if (in_dr) {
            s->dr_round = r.ctr[plan_ctr_round];
            s->res.rounds = r.ctr[plan_ctr_run];
            s->tree_n = std::max(r.tr_n[0], r.tr_n[1]);
            if (r.status > 0) {
                s->pass_solved = true;
                s->best_cost = r.best_cost;
                s->res.start_tree_size = r.path_meta[0];
                s->res.goal_tree_size = r.path_meta[1];
                s->res.iters = r.path_meta[2];
                if (s->st.path_simplify) {
                    s->phase = Phase::Simplify;
                    s->sp_stage = 0;
                } else if (int rc = finish_pass(s, r)) {
                    return rc;
                }
            } else if (r.status < 0) {
                // tree capacity reached: the pass ends unsolved
                if (int rc = finish_pass(s, r)) return rc;
            } else if (s->dr_round >= s->dr_budget) {
                if (s->seeking) {
                    s->phase = Phase::Roadmap;  // one more roadmap round, then a fresh dRRT launch
                } else if (int rc = finish_pass(s, r)) {
                    return rc;
                }
            }
        } else {
            s->sp_stage++;
            if (s->sp_stage > s->st.bspline_step) {
                s->best_cost = r.best_cost;
                if (int rc = finish_pass(s, r)) return rc;
            }
        }
        if (s->phase != Phase::Done && r.events[kEvUnionFind] > 0) {
            // roadmap connectivity can no longer be trusted: end the problem
            s->phase = Phase::Done;
            s->res.status = MRPOP_ERR_INTERNAL;
            s->res.wall_ns = ns_since(s->t_submit);
            s->res.kernel_ns = ns_since(s->t_kernel_start);
        }
        if (s->phase != Phase::Done && s->opt.time_budget_s > 0 &&
            std::chrono::duration<double>(now - s->t_submit).count() > s->opt.time_budget_s) {
            s->phase = Phase::Done;
            s->res.status = MRPOP_ERR_TIMEOUT;
            s->res.wall_ns = ns_since(s->t_submit);
            s->res.kernel_ns = ns_since(s->t_kernel_start);
        }
        if (s->phase == Phase::Done)  // the roadmaps as they ended, for later queries
            if (int rc = cache_store(s, s->res.roadmap_size, ~0u)) return rc;
    }
    L.table.clear();
    L.rm.clear();
    L.dr.clear();
    L.sp.clear();
    return MRPOP_OK;
}

// Wait for a lane's step, at most the context's lane deadline: its kernels
// are bounded, so a step that runs far beyond it is stuck.  Polls the event
// (a blocking wait could not be given up).
int Engine::wait_lane(Lane &L) {
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spin = 0;; ++spin) {
        const hipError_t e = hipEventQuery(L.done);
        if (e == hipSuccess) return MRPOP_OK;
        if (e != hipErrorNotReady) return c_->fail(MRPOP_ERR_HIP, "lane step: %s", hipGetErrorString(e));
        if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > c_->lane_deadline_s)
            return MRPOP_ERR_TIMEOUT;
        if (spin > 64) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

// Wait until one of the busy lanes `cands` completes its step (*done = that
// lane, MRPOP_OK) or one has run past the context's lane deadline since its
// launch (*done = that lane, MRPOP_ERR_TIMEOUT).  Polls the events.
int Engine::wait_any(Lane *const *cands, int n, Lane **done) {
    for (unsigned spin = 0;; ++spin) {
        const auto now = std::chrono::steady_clock::now();
        for (int i = 0; i < n; ++i) {
            Lane &L = *cands[i];
            const hipError_t e = hipEventQuery(L.done);
            if (e == hipSuccess) {
                *done = &L;
                return MRPOP_OK;
            }
            if (e != hipErrorNotReady) return c_->fail(MRPOP_ERR_HIP, "lane step: %s", hipGetErrorString(e));
            if (std::chrono::duration<double>(now - L.t_launch).count() > c_->lane_deadline_s) {
                *done = &L;
                return MRPOP_ERR_TIMEOUT;
            }
        }
        if (spin > 64) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

// A step outlived the deadline: its problems end with MRPOP_ERR_TIMEOUT.
// Their memory stays with the lane until its work completes (release()
// waits for the lane's stream), and the lane takes no new work until then.
void Engine::abandon_lane(Lane &L) {
    L.abandoned = true;
    for (Slot *s : L.table) {
        if (!s || s->phase == Phase::Done) continue;
        s->phase = Phase::Done;
        s->res.status = MRPOP_ERR_TIMEOUT;
        s->res.wall_ns = ns_since(s->t_submit);
        s->res.kernel_ns = ns_since(s->t_kernel_start);
    }
    c_->fail(MRPOP_ERR_TIMEOUT, "a planner step did not complete within %.3g s", c_->lane_deadline_s);
}

int Engine::drain() {
    for (Lane &L : lanes_)
        if (L.busy) {
            const int w = wait_lane(L);
            if (w == MRPOP_ERR_TIMEOUT) {
                abandon_lane(L);
                return w;
            }
            if (w) return w;
            L.abandoned = false;
            if (int rc = process_lane(L)) return rc;
        }
    return MRPOP_OK;
}

int Engine::step() {
    if (int rc = ensure_lanes()) return rc;
    const auto now = std::chrono::steady_clock::now();
    for (Slot *s : slots_) {
        if (s->phase == Phase::Done || s->lane >= 0) continue;
        if (s->opt.time_budget_s > 0 && std::chrono::duration<double>(now - s->t_submit).count() > s->opt.time_budget_s) {
            s->phase = Phase::Done;
            s->res.status = MRPOP_ERR_TIMEOUT;
            s->res.wall_ns = ns_since(s->t_submit);
            s->res.kernel_ns = ns_since(s->t_kernel_start);
        }
    }
    Lane &R = lanes_[0];
    // a lane given up on rejoins once its work has completed (its problems
    // already ended)
    for (Lane &L : lanes_)
        if (L.abandoned && hipEventQuery(L.done) == hipSuccess) {
            L.abandoned = false;
            const int li = (int)(&L - lanes_);
            for (auto g = graveyard_.begin(); g != graveyard_.end();)
                if (g->first == li) {
                    (void)hipFree(g->second);
                    g = graveyard_.erase(g);
                } else {
                    ++g;
                }
            if (int rc = process_lane(L)) return rc;
        }
    if (!R.busy) {
        // slots still looking for their first solution go first; the roadmap
        // rounds of optimising passes wait for a step without them
        for (Slot *s : slots_)
            if (s->lane < 0 && s->phase == Phase::Roadmap && !s->res.solved) R.rm.push_back(s);
        if (R.rm.empty())
            for (Slot *s : slots_)
                if (s->lane < 0 && s->phase == Phase::Roadmap) R.rm.push_back(s);
        if (int rc = launch_lane(R)) return rc;
    }
    // Time to first solution first: while some problem is still searching
    // for its first path, the optimise passes of solved problems join a
    // search lane's launch on one step in kOptimiseShare, so a hard problem's
    // rounds are not slowed by every other problem's optimisation.
    // When fewer problems seek than optimise or simplify, the seekers get
    // search lane 1 to themselves and the others run every step on the last
    // lane, so the rounds of the hard problems that decide a batch's end are
    // not held back by the others' work; otherwise the lanes split the
    // problems by slot id.
    int n_seek = 0, n_rest = 0;
    for (Slot *s : slots_) {
        if (s->phase == Phase::Search && s->seeking) ++n_seek;
        else if (s->phase == Phase::Search || s->phase == Phase::Simplify) ++n_rest;
    }
    const bool by_role = n_seek > 0 && n_seek < n_rest;
    for (int k = 0; k < kSearchLanes; ++k) {
        Lane &S = lanes_[1 + k];
        if (S.busy) continue;
        const bool optimisers = by_role || n_seek == 0 || (search_steps_[k]++ % kOptimiseShare) == 0;
        for (Slot *s : slots_) {
            if (s->lane >= 0) continue;
            const bool seeker = s->phase == Phase::Search && s->seeking;
            const int lane = by_role ? (seeker ? 0 : kSearchLanes - 1) : s->id % kSearchLanes;
            if (lane != k) continue;
            if (s->phase == Phase::Search && (s->seeking || optimisers)) S.dr.push_back(s);
            else if (s->phase == Phase::Simplify) S.sp.push_back(s);
        }
        if (int rc = launch_lane(S)) return rc;
    }
    // wait for a search lane (short steps) when one runs, else for the
    // roadmap lane; lanes that have completed meanwhile are processed too
    Lane *cands[1 + kSearchLanes];
    int n = 0;
    for (int k = 1; k <= kSearchLanes; ++k)
        if (lanes_[k].busy && !lanes_[k].abandoned) cands[n++] = &lanes_[k];
    if (!n && R.busy && !R.abandoned) cands[n++] = &R;
    if (!n) {
        bool abandoned = false;
        for (Lane &L : lanes_) abandoned |= L.abandoned;
        if (abandoned) std::this_thread::sleep_for(std::chrono::microseconds(100));
        return MRPOP_OK;
    }
    Lane *first = nullptr;
    const int w = wait_any(cands, n, &first);
    if (w == MRPOP_ERR_TIMEOUT) {
        abandon_lane(*first);  // its problems end; the others go on
        return MRPOP_OK;
    }
    if (w) return w;
    if (int rc = process_lane(*first)) return rc;
    for (Lane &L : lanes_)
        if (&L != first && L.busy && !L.abandoned && hipEventQuery(L.done) == hipSuccess)
            if (int rc = process_lane(L)) return rc;
    return MRPOP_OK;
}

int Engine::debug_roadmap(Slot *s, int rounds) {
    if (int rc = ensure_lanes()) return rc;
    if (int rc = drain()) return rc;
    Lane &L = lanes_[0];
    for (int r = 0; r < rounds; ++r) {
        s->phase = Phase::Roadmap;
        L.rm.assign(1, s);
        if (int rc = launch_lane(L)) return rc;
        MRPOP_HIP(c_, hipEventSynchronize(L.done));
        if (int rc = process_lane(L)) return rc;
    }
    return MRPOP_OK;
}

// Exactly `rounds` dRRT rounds of one slot in one launch (fewer if the pass
// is solved), without the host state machine: the slot's device state is
// left as the last round's kernels wrote it.  keep_round: the rounds
// continue the previous launch's round count (as the later rounds of one
// launch do) instead of starting a launch at round 0.
int Engine::debug_search(Slot *s, int rounds, bool keep_round) {
    if (rounds < 1) return c_->fail(MRPOP_ERR_INVALID, "rounds must be positive");
    if (int rc = ensure_lanes()) return rc;
    if (int rc = drain()) return rc;
    Lane &L = lanes_[0];
    const int rps = s->opt.rounds_per_step;
    s->opt.rounds_per_step = rounds;
    if (keep_round) {
        int ctr[4];
        MRPOP_HIP(c_, hipMemcpy(ctr, s->dv.ctr, sizeof(ctr), hipMemcpyDeviceToHost));
        ctr[kCtrLimit] = ctr[kCtrRound] + rounds;
        MRPOP_HIP(c_, hipMemcpy(s->dv.ctr, ctr, sizeof(ctr), hipMemcpyHostToDevice));
        s->dv.round_reset = 0;
    } else {
        s->dv.round_reset = 1;
        s->dv.round_limit = rounds;
    }
    s->phase = Phase::Search;
    L.rm.clear();
    L.sp.clear();
    L.dr.assign(1, s);
    const int rc = launch_lane(L);
    s->opt.rounds_per_step = rps;
    L.dr.clear();
    if (rc) return rc;
    MRPOP_HIP(c_, hipEventSynchronize(L.done));
    L.busy = false;
    s->lane = -1;
    return MRPOP_OK;
}

// Start AO-x pass `pass` (trees reset, optimise mode from pass 1 on, PHS from
// the device's best cost) without the host state machine.
int Engine::debug_start_pass(Slot *s, int pass) {
    if (pass < 0) return c_->fail(MRPOP_ERR_INVALID, "pass %d", pass);
    if (int rc = drain()) return rc;
    float best = 0.f;
    MRPOP_HIP(c_, hipMemcpy(&best, s->dv.best_cost, sizeof(float), hipMemcpyDeviceToHost));
    s->best_cost = best;
    s->pass = pass;
    return start_pass(s);
}

// One simplification stage (0 .. bspline_step) of the slot's recorded path,
// without the host state machine.
int Engine::debug_simplify(Slot *s, int stage) {
    if (stage < 0 || stage > s->st.bspline_step) return c_->fail(MRPOP_ERR_INVALID, "stage %d", stage);
    if (int rc = ensure_lanes()) return rc;
    if (int rc = drain()) return rc;
    Lane &L = lanes_[0];
    s->phase = Phase::Simplify;
    s->sp_stage = stage;
    L.rm.clear();
    L.dr.clear();
    L.sp.assign(1, s);
    const int rc = launch_lane(L);
    L.sp.clear();
    if (rc) return rc;
    MRPOP_HIP(c_, hipEventSynchronize(L.done));
    L.busy = false;
    s->lane = -1;
    return MRPOP_OK;
}

int Engine::debug_copy_roadmap(Slot *s, int arm, float *q, int *deg, int *nbr, int *uf, int *connected, int *n,
                               int cap) {
    if (arm < 0 || arm >= s->A) return c_->fail(MRPOP_ERR_INVALID, "arm %d outside the problem's arms", arm);
    int count = 0;
    MRPOP_HIP(c_, hipMemcpy(&count, s->dv.rm_n + arm, sizeof(int), hipMemcpyDeviceToHost));
    *n = count;
    const int m = std::min(count, cap);
    const size_t off = (size_t)arm * s->dv.rm_cap;
    if (q && m) MRPOP_HIP(c_, hipMemcpy(q, s->dv.rm_q + off * 8, sizeof(float) * 8 * m, hipMemcpyDeviceToHost));
    if (deg && m) MRPOP_HIP(c_, hipMemcpy(deg, s->dv.rm_deg + off, sizeof(int) * m, hipMemcpyDeviceToHost));
    if (nbr && m)
        MRPOP_HIP(c_, hipMemcpy(nbr, s->dv.rm_nbr + off * s->dv.rm_K, sizeof(int) * (size_t)m * s->dv.rm_K,
                                hipMemcpyDeviceToHost));
    if (uf && m) MRPOP_HIP(c_, hipMemcpy(uf, s->dv.rm_uf + off, sizeof(int) * m, hipMemcpyDeviceToHost));
    if (connected) MRPOP_HIP(c_, hipMemcpy(connected, s->dv.rm_connected + arm, sizeof(int), hipMemcpyDeviceToHost));
    return MRPOP_OK;
}

// End of an AO-x pass: keep an improving path (pop_planner.cu:1000-1022),
// then start the next pass or finish.
int Engine::finish_pass(Slot *s, const SlotReport &r) {
    if (s->pass_solved && r.path_n > 0) {
        const float cost = r.path_cost;
        if (!s->res.solved || s->res.cost > cost) {
            auto t0 = std::chrono::steady_clock::now();
            std::vector<float> path((size_t)r.path_n * s->D);
            MRPOP_HIP(c_, hipMemcpy(path.data(), s->dv.path, path.size() * sizeof(float), hipMemcpyDeviceToHost));
            s->res.copy_ns += ns_since(t0);
            const uint64_t now_ns = ns_since(s->t_submit);
            if (!s->res.solved) s->res.ttfs_ns = now_ns;
            s->res.improve_ns.push_back(now_ns);
            s->res.solved = true;
            s->res.cost = cost;
            s->res.path = path;
            s->res.seed_paths.push_back(SeedPath{cost, path});
        }
    }
    s->res.passes = s->pass + 1;
    s->pass++;
    if (s->pass >= s->st.rrtc_iter) {
        s->phase = Phase::Done;
        s->res.wall_ns = ns_since(s->t_submit);
        s->res.kernel_ns = ns_since(s->t_kernel_start);
        return MRPOP_OK;
    }
    return start_pass(s);
}

}  // namespace mrpop
