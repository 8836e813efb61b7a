// engine.hip -- multi-problem planning engine (see engine.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.h"

namespace mrpop {

namespace {

constexpr int kRoadmapIterPerBlock = 230;  // ROADMAP_BUILD_ITER (roadmap_interior.cuh:20)
constexpr int kPathCap = 2048;             // waypoints (MAX_PATH_SIZE floats in the reference)
constexpr int kMaxConnectChunks = 64;      // chunks of one connection attempt
constexpr float kFltMax = 3.402823466e38f;

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

bool is_world_axes(const float *rec) {
    static const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int i = 0; i < 9; ++i)
        if (rec[3 + i] != I[i]) return false;
    return true;
}

uint64_t ns_since(std::chrono::steady_clock::time_point t) {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t)
        .count();
}

// Sizes of one slot's buffers, in the order they are carved.
struct Layout {
    size_t off[64];
    int n = 0;
    size_t total = 0;
    size_t add(size_t bytes) {
        off[n++] = total;
        total += align256(bytes);
        return off[n - 1];
    }
};

}  // namespace

Engine::~Engine() {
    for (Slot *s : slots_) {
        if (s->mem) (void)hipFree(s->mem);
        delete s;
    }
    for (void *p : {(void *)d_slots_, (void *)d_worlds_, (void *)d_ids_, (void *)d_groups_, (void *)d_wave_,
                    (void *)d_report_})
        if (p) (void)hipFree(p);
    for (void *p : {(void *)h_slots_, (void *)h_worlds_, (void *)h_ids_, (void *)h_groups_, (void *)h_report_})
        if (p) (void)hipHostFree(p);
}

bool Engine::idle() const {
    for (const Slot *s : slots_)
        if (s->phase != Phase::Done) return false;
    return true;
}

int Engine::ensure_capacity(int n) {
    if (n <= cap_) return MRPOP_OK;
    int cap = std::max(n, cap_ ? cap_ * 2 : 16);
    MRPOP_HIP(c_, hipStreamSynchronize(c_->stream));
    auto grow = [&](void **d, void **h, size_t elem, size_t count, bool host) -> int {
        void *nd = nullptr, *nh = nullptr;
        MRPOP_HIP(c_, hipMalloc(&nd, elem * count));
        if (host) MRPOP_HIP(c_, hipHostMalloc(&nh, elem * count, hipHostMallocDefault));
        if (*d) {
            MRPOP_HIP(c_, hipMemcpy(nd, *d, elem * cap_, hipMemcpyDeviceToDevice));
            (void)hipFree(*d);
        }
        if (host && *h) {
            std::memcpy(nh, *h, elem * cap_);
            (void)hipHostFree(*h);
        }
        *d = nd;
        if (host) *h = nh;
        return MRPOP_OK;
    };
    if (int rc = grow((void **)&d_slots_, (void **)&h_slots_, sizeof(SlotDev), cap, true)) return rc;
    if (int rc = grow((void **)&d_worlds_, (void **)&h_worlds_, sizeof(dev::WorldDesc), cap, true)) return rc;
    if (int rc = grow((void **)&d_report_, (void **)&h_report_, sizeof(SlotReport), cap, true)) return rc;
    // slot-id lists of one step (each launch sequence appends its own list)
    if (d_ids_) (void)hipFree(d_ids_);
    if (h_ids_) (void)hipHostFree(h_ids_);
    MRPOP_HIP(c_, hipMalloc(&d_ids_, sizeof(int) * cap * 32));
    MRPOP_HIP(c_, hipHostMalloc(&h_ids_, sizeof(int) * cap * 32, hipHostMallocDefault));
    // edge groups: per slot at most 2*5 roadmap + dr + cn + sp, twice over for safety
    group_cap_ = (size_t)cap * 32;
    if (d_groups_) (void)hipFree(d_groups_);
    if (h_groups_) (void)hipHostFree(h_groups_);
    if (d_wave_) (void)hipFree(d_wave_);
    MRPOP_HIP(c_, hipMalloc(&d_groups_, sizeof(dev::EdgeGroup) * group_cap_));
    MRPOP_HIP(c_, hipHostMalloc(&h_groups_, sizeof(dev::EdgeGroup) * group_cap_, hipHostMallocDefault));
    MRPOP_HIP(c_, hipMalloc(&d_wave_, sizeof(unsigned) * (group_cap_ + 1)));
    cap_ = cap;
    return MRPOP_OK;
}

size_t Engine::slot_bytes(int A, const mrpop_settings &st, const SolveOptions &opt) const {
    const int D = 7 * A;
    const int B = st.num_new_configs;
    const int M = kRoadmapIterPerBlock * opt.block_per_robot;
    const int rm_cap = opt.roadmap_capacity;
    const int tr_cap = opt.tree_capacity > 0 ? opt.tree_capacity : std::min(st.max_samples, 262144);
    const int Kc = opt.neighbour_candidates, K = opt.neighbour_capacity;
    const size_t ce_cap = (size_t)B * kMaxConnectChunks;
    const size_t sp_cap = (size_t)kPathCap * 40;
    Layout L;
    L.add(sizeof(float) * (MRPOP_MAX_OBSTACLES * 15 + 64));                 // world
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
    L.add(sizeof(float) * (size_t)B * D * 2 + sizeof(int) * B * 2);         // ed_from/to, ed_b, ed_coll
    L.add(sizeof(int) * B * 3 + sizeof(float) * B);                         // cn_*
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

int Engine::add(const mrpop_scene &scene, int A, const float *bases, const float *start, int n_goals,
                const float *goals, unsigned long long seed, const mrpop_settings &st, const SolveOptions &opt,
                Slot **out) {
    *out = nullptr;
    if (A < 2 || A > geom::kMaxArms) return c_->fail(MRPOP_ERR_INVALID, "planning needs 2..5 arms (got %d)", A);
    if (n_goals < 1 || n_goals > kMaxGoals) return c_->fail(MRPOP_ERR_INVALID, "1..%d goals supported", kMaxGoals);
    if (st.granularity < 1 || st.granularity > 64) return c_->fail(MRPOP_ERR_INVALID, "granularity outside 1..64");
    if (st.num_new_configs < 1 || st.num_new_configs > 65535)
        return c_->fail(MRPOP_ERR_INVALID, "num_new_configs outside 1..65535");
    if (!(st.range > 0)) return c_->fail(MRPOP_ERR_INVALID, "range must be positive");
    if (opt.block_per_robot < 1 || opt.roadmap_capacity < 1 + n_goals + 1 || opt.neighbour_capacity < 1 ||
        opt.neighbour_candidates < 1 || opt.neighbour_candidates > 64 || opt.rounds_per_step < 1 ||
        opt.rounds_per_step > 1024)
        return c_->fail(MRPOP_ERR_INVALID, "bad solve options (neighbour_candidates 1..64, rounds_per_step 1..1024)");
    const int counts[5] = {scene.n_spheres, scene.n_capsules, scene.n_z_capsules, scene.n_cuboids, scene.n_z_cuboids};
    long n_obs = 0;
    for (int k = 0; k < 5; ++k) {
        if (counts[k] < 0) return c_->fail(MRPOP_ERR_INVALID, "negative obstacle count");
        n_obs += counts[k];
    }
    if (n_obs > MRPOP_MAX_OBSTACLES) return c_->fail(MRPOP_ERR_INVALID, "too many obstacles");
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
    const int D = s->D, B = st.num_new_configs, M = kRoadmapIterPerBlock * opt.block_per_robot;
    const int rm_cap = opt.roadmap_capacity, K = opt.neighbour_capacity, Kc = opt.neighbour_candidates;
    const int tr_cap = opt.tree_capacity > 0 ? opt.tree_capacity : std::min(st.max_samples, 262144);
    const size_t ce_cap = (size_t)B * kMaxConnectChunks, sp_cap = (size_t)kPathCap * 40;
    s->mem_bytes = slot_bytes(A, st, opt);
    if (hipMalloc(&s->mem, s->mem_bytes) != hipSuccess) {
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
    s->wbuf = (float *)take(sizeof(float) * (MRPOP_MAX_OBSTACLES * 15 + 64));
    v.sg = (float *)take(sizeof(float) * (1 + kMaxGoals) * D);
    v.rm_q = (float *)take(sizeof(float) * A * rm_cap * 8);
    v.rm_nbr = (int *)take(sizeof(int) * (size_t)A * rm_cap * K);
    v.rm_deg = (int *)take(sizeof(int) * A * rm_cap);
    v.rm_uf = (int *)take(sizeof(int) * A * rm_cap);
    int *small = (int *)take(sizeof(int) * 64);
    v.rm_n = small;
    v.rm_connected = small + 8;
    v.re_count = small + 16;
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
    int *ex = (int *)take(sizeof(int) * B * 6 + sizeof(float) * B);
    v.ex_tree = ex;
    v.ex_state = ex + B;
    v.ex_near = ex + 2 * B;
    v.ex_exist = ex + 3 * B;
    v.ex_index = ex + 4 * B;
    v.ex_cost = (float *)(ex + 6 * B);
    v.ex_q = (float *)take(sizeof(float) * (size_t)B * D);
    v.ex_rid = (int *)take(sizeof(int) * B * A);
    float *ed = (float *)take(sizeof(float) * (size_t)B * D * 2 + sizeof(int) * B * 2);
    v.ed_from = ed;
    v.ed_to = ed + (size_t)B * D;
    v.ed_b = (int *)(ed + (size_t)B * D * 2);
    v.ed_coll = (uint32_t *)(v.ed_b + B);
    int *cn = (int *)take(sizeof(int) * B * 3 + sizeof(float) * B);
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

    // ---- world (same record layout as mrpop_world_create)
    std::vector<float> host;
    std::vector<float> aabb, obb;
    for (int i = 0; i < scene.n_cuboids; ++i) {
        const float *r = scene.cuboids + 15 * i;
        if (is_world_axes(r)) aabb.insert(aabb.end(), {r[0], r[1], r[2], r[12], r[13], r[14]});
        else obb.insert(obb.end(), r, r + 15);
    }
    dev::WorldDesc &w = s->wdesc;
    auto put = [&](const float *p, size_t n) { host.insert(host.end(), p, p + n); };
    w.o_sph = (int)host.size();
    w.n_sph = scene.n_spheres;
    if (w.n_sph) put(scene.spheres, 4 * (size_t)w.n_sph);
    w.o_cap = (int)host.size();
    w.n_cap = scene.n_capsules;
    if (w.n_cap) put(scene.capsules, 8 * (size_t)w.n_cap);
    w.o_zcap = (int)host.size();
    w.n_zcap = scene.n_z_capsules;
    if (w.n_zcap) put(scene.z_capsules, 8 * (size_t)w.n_zcap);
    w.o_aabb = (int)host.size();
    w.n_aabb = (int)(aabb.size() / 6);
    put(aabb.data(), aabb.size());
    w.o_obb = (int)host.size();
    w.n_obb = (int)(obb.size() / 15);
    put(obb.data(), obb.size());
    w.o_zcub = (int)host.size();
    w.n_zcub = scene.n_z_cuboids;
    if (w.n_zcub) put(scene.z_cuboids, 15 * (size_t)w.n_zcub);
    w.o_base = (int)host.size();
    put(b, 12 * (size_t)A);
    w.n_arms = A;
    w.obs = s->wbuf;
    MRPOP_HIP(c_, hipMemcpyAsync(s->wbuf, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, c_->stream));
    MRPOP_HIP(c_, hipMemcpyAsync(v.sg, s->sg.data(), s->sg.size() * sizeof(float), hipMemcpyHostToDevice, c_->stream));
    MRPOP_HIP(c_, hipMemsetAsync(v.rw_key, 0xff, sizeof(unsigned long long) * 2 * (size_t)tr_cap, c_->stream));
    MRPOP_HIP(c_, hipMemsetAsync(ctr, 0, sizeof(int) * 64, c_->stream));
    MRPOP_HIP(c_, hipMemsetAsync(v.rl_slot, 0xff, sizeof(int) * (size_t)A * rm_cap, c_->stream));
    MRPOP_HIP(c_, hipMemsetAsync(v.rl_bits, 0, sizeof(unsigned) * (size_t)A * rl_cap * rl_words, c_->stream));
    {
        const float fm = kFltMax;
        MRPOP_HIP(c_, hipMemcpyAsync(v.best_cost, &fm, sizeof(float), hipMemcpyHostToDevice, c_->stream));
    }

    // ---- edge groups (pointers and counts are fixed for the slot's life)
    for (int a = 0; a < A; ++a) {
        s->g_valid[a] = dev::EdgeGroup{v.rs_q + (size_t)a * M * 7, v.rs_q + (size_t)a * M * 7, 7, 1, M, a, nullptr,
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
    h_slots_[id] = v;
    MRPOP_HIP(c_, hipMemcpyAsync(d_worlds_ + id, &h_worlds_[id], sizeof(dev::WorldDesc), hipMemcpyHostToDevice, c_->stream));
    MRPOP_HIP(c_, hipMemcpyAsync(d_slots_ + id, &h_slots_[id], sizeof(SlotDev), hipMemcpyHostToDevice, c_->stream));
    h_ids_[0] = id;
    MRPOP_HIP(c_, hipMemcpyAsync(d_ids_, h_ids_, sizeof(int), hipMemcpyHostToDevice, c_->stream));
    plan::launch_roadmap_init(d_slots_, d_ids_, 1, c_->stream);
    MRPOP_HIP(c_, hipStreamSynchronize(c_->stream));  // staging buffers are reused by the next call
    slots_.push_back(s);
    s->t_kernel_start = std::chrono::steady_clock::now();
    if (int rc = start_pass(s)) return rc;
    *out = s;
    return MRPOP_OK;
}

void Engine::release(Slot *s) {
    auto it = std::find(slots_.begin(), slots_.end(), s);
    if (it == slots_.end()) return;
    (void)hipStreamSynchronize(c_->stream);
    if (s->mem) {
        (void)hipFree(s->mem);
        bytes_in_use_ -= s->mem_bytes;
    }
    slots_.erase(it);
    delete s;
}

// PHS transform of the slot (pop_planner.cu:289-346): Householder rotation
// taking e_1 to the unit start->goal direction, scaled by the hyperspheroid
// radii (c_best / 2, sqrt(c_best^2 - c_min^2) / 2, ...).
void Engine::compute_phs(Slot *s) {
    const int D = s->D;
    const float *f1 = s->sg.data(), *f2 = s->sg.data() + D;
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
    const double cb = s->best_cost;
    const double conj = std::sqrt(std::max(0.0, cb * cb - cmin * cmin));
    std::vector<float> tf((size_t)D * D);
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) tf[(size_t)r * D + c] = (float)(rot[(size_t)r * D + c] * (c == 0 ? cb / 2 : conj / 2));
    for (int d = 0; d < D; ++d) s->dv.phs_center[d] = (f1[d] + f2[d]) / 2.0f;
    (void)hipMemcpy(s->dv.phs_tf, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice);
}

int Engine::start_pass(Slot *s) {
    s->dv.optimize = s->pass > 0 ? 1 : 0;
    s->dv.pass = s->pass;
    s->pass_solved = false;
    s->rm_until_connected = s->pass == 0;
    s->phase = Phase::Roadmap;
    s->dr_round = 0;
    s->dr_budget = s->pass == 0 ? s->st.max_iters : s->st.optimize_iters;
    if (s->dv.optimize && s->st.phs) compute_phs(s);
    h_slots_[s->id] = s->dv;
    MRPOP_HIP(c_, hipMemcpyAsync(d_slots_ + s->id, &h_slots_[s->id], sizeof(SlotDev), hipMemcpyHostToDevice, c_->stream));
    h_ids_[0] = s->id;
    MRPOP_HIP(c_, hipMemcpyAsync(d_ids_, h_ids_, sizeof(int), hipMemcpyHostToDevice, c_->stream));
    plan::launch_tree_reset(d_slots_, d_ids_, 1, c_->stream);
    MRPOP_HIP(c_, hipStreamSynchronize(c_->stream));
    return MRPOP_OK;
}

// Copy the edge-group descriptors of `list` (per_slot groups each) to the
// device, bucketed by arm count for the multi-arm modes.  The descriptors
// hold device pointers and counts only, so one copy serves every round of a
// step.
int Engine::stage_groups(int mode, const std::vector<const dev::EdgeGroup *> &groups, int per_slot,
                         const std::vector<Slot *> &list, size_t &cursor, Staged &out) {
    out.mode = mode;
    out.buckets.clear();
    const size_t begin = cursor;
    for (int A = (mode == MRPOP_MODE_SINGLE ? 1 : 2); A <= (mode == MRPOP_MODE_SINGLE ? 1 : 5); ++A) {
        Staged::Bucket bk{A, cursor, 0, 0};
        const int S = 64 / A;
        for (size_t i = 0; i < list.size(); ++i) {
            if (mode != MRPOP_MODE_SINGLE && list[i]->A != A) continue;
            const int n = mode == MRPOP_MODE_SINGLE ? list[i]->A : 1;
            for (int k = 0; k < n && k < per_slot; ++k) {
                if (cursor >= group_cap_) return c_->fail(MRPOP_ERR_BUDGET, "edge-group staging full");
                const dev::EdgeGroup &g = groups[i][k];
                h_groups_[cursor++] = g;
                bk.max_waves += ((size_t)g.n_edges * g.G + S - 1) / S;
            }
        }
        bk.n = (int)(cursor - bk.first);
        if (bk.n) out.buckets.push_back(bk);
    }
    if (cursor > begin)
        MRPOP_HIP(c_, hipMemcpyAsync(d_groups_ + begin, h_groups_ + begin, sizeof(dev::EdgeGroup) * (cursor - begin),
                                     hipMemcpyHostToDevice, c_->stream));
    return MRPOP_OK;
}

// One batched check launch per arm-count bucket.  Every producer of planner
// edges clears their verdict words, so the launches skip the zeroing pass.
int Engine::launch_staged(const Staged &st) {
    for (const Staged::Bucket &bk : st.buckets)
        if (int rc = launch_check_groups(c_, st.mode, bk.A, d_worlds_, d_groups_ + bk.first, bk.n, d_wave_, bk.max_waves,
                                         false, c_->stream))
            return rc;
    return MRPOP_OK;
}

int Engine::step() {
    std::vector<Slot *> rm, dr, sp;
    for (Slot *s : slots_) {
        if (s->phase == Phase::Done) continue;
        if (s->opt.time_budget_s > 0 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t_submit).count() > s->opt.time_budget_s) {
            s->phase = Phase::Done;
            s->res.status = MRPOP_ERR_TIMEOUT;
            s->res.wall_ns = ns_since(s->t_submit);
            s->res.kernel_ns = ns_since(s->t_kernel_start);
            continue;
        }
        if (s->phase == Phase::Roadmap) rm.push_back(s);
        else if (s->phase == Phase::Search) dr.push_back(s);
        else if (s->phase == Phase::Simplify) sp.push_back(s);
    }
    if (rm.empty() && dr.empty() && sp.empty()) return MRPOP_OK;
    hipStream_t st = c_->stream;
    // per-slot round state to the device
    for (Slot *s : slots_) h_slots_[s->id] = s->dv;
    MRPOP_HIP(c_, hipMemcpyAsync(d_slots_, h_slots_, sizeof(SlotDev) * cap_, hipMemcpyHostToDevice, st));
    size_t cursor = 0, id_cursor = 0;
    // every list gets its own staging range: the copies run after the host has moved on
    auto upload_ids = [&](const std::vector<Slot *> &list, int) -> int * {
        if (id_cursor + list.size() > (size_t)cap_ * 32) return nullptr;
        int *h = h_ids_ + id_cursor;
        for (size_t i = 0; i < list.size(); ++i) h[i] = list[i]->id;
        (void)hipMemcpyAsync(d_ids_ + id_cursor, h, sizeof(int) * list.size(), hipMemcpyHostToDevice, st);
        int *d = d_ids_ + id_cursor;
        id_cursor += list.size();
        return d;
    };
    // ---- roadmap round (roadmap_interior.cuh:94-339)
    if (!rm.empty()) {
        int *d = upload_ids(rm, 0);
        const int n = (int)rm.size();
        int max_am = 0;
        for (Slot *s : rm) max_am = std::max(max_am, s->A * s->dv.M);
        plan::launch_rm_sample(d_slots_, d, n, max_am, st);
        std::vector<const dev::EdgeGroup *> gv, gr;
        for (Slot *s : rm) {
            gv.push_back(s->g_valid);
            gr.push_back(s->g_rm);
        }
        Staged sv, sr;
        if (int rc = stage_groups(MRPOP_MODE_SINGLE, gv, geom::kMaxArms, rm, cursor, sv)) return rc;
        if (int rc = stage_groups(MRPOP_MODE_SINGLE, gr, geom::kMaxArms, rm, cursor, sr)) return rc;
        if (int rc = launch_staged(sv)) return rc;
        plan::launch_rm_assign(d_slots_, d, n, st);
        plan::launch_rm_radius(d_slots_, d, n, max_am, st);
        plan::launch_rm_compact(d_slots_, d, n, st);
        if (int rc = launch_staged(sr)) return rc;
        plan::launch_rm_link(d_slots_, d, n, max_am, st);
        plan::launch_rm_connected(d_slots_, d, n, st);
    }
    // ---- dRRT rounds (pop_planner.cu:350-736).  The round index lives on the
    // device; a slot whose pass is solved or whose launch is out of rounds
    // skips the remaining rounds of the step.
    if (!dr.empty()) {
        int *d = upload_ids(dr, 1);
        const int n = (int)dr.size();
        int max_b = 0, rounds = 1 << 30;
        bool reset = false;
        for (Slot *s : dr) {
            max_b = std::max(max_b, s->dv.B);
            rounds = std::min(rounds, std::max(1, s->opt.rounds_per_step));
            reset |= s->dv.round_reset != 0;
            s->dv.round_reset = 0;  // applied by this step's k_dr_begin
        }
        std::vector<const dev::EdgeGroup *> g1, g2;
        for (Slot *s : dr) {
            g1.push_back(&s->g_dr);
            g2.push_back(&s->g_cn);
        }
        Staged s1, s2;
        if (int rc = stage_groups(MRPOP_MODE_DRRT, g1, 1, dr, cursor, s1)) return rc;
        if (int rc = stage_groups(MRPOP_MODE_FULL, g2, 1, dr, cursor, s2)) return rc;
        if (reset) plan::launch_dr_begin(d_slots_, d, n, st);
        for (int r = 0; r < rounds; ++r) {
            plan::launch_dr_expand(d_slots_, d, n, max_b, st);
            if (int rc = launch_staged(s1)) return rc;
            plan::launch_dr_commit(d_slots_, d, n, st);
            plan::launch_dr_connect(d_slots_, d, n, max_b, st);
            if (int rc = launch_staged(s2)) return rc;
            plan::launch_dr_finish(d_slots_, d, n, st);
        }
    }
    // ---- simplification stage (path_simplification.cu:41-238)
    if (!sp.empty()) {
        // group slots by stage: smoothing stages share one launch sequence, the shortcut another
        std::vector<Slot *> smooth, cut;
        for (Slot *s : sp) (s->sp_stage < s->st.bspline_step ? smooth : cut).push_back(s);
        int which = 2;
        for (auto *lst : {&smooth, &cut}) {
            if (lst->empty()) continue;
            // stages may differ between slots: launch per distinct stage
            std::vector<int> stages;
            for (Slot *s : *lst) stages.push_back(s->sp_stage);
            std::sort(stages.begin(), stages.end());
            stages.erase(std::unique(stages.begin(), stages.end()), stages.end());
            for (int stage : stages) {
                std::vector<Slot *> part;
                for (Slot *s : *lst)
                    if (s->sp_stage == stage) part.push_back(s);
                // settings may differ per slot too: group by (bspline_step, range, min_change)
                while (!part.empty()) {
                    Slot *s0 = part.front();
                    std::vector<Slot *> same, rest;
                    for (Slot *s : part)
                        ((s->st.bspline_step == s0->st.bspline_step && s->st.range == s0->st.range &&
                          s->st.bspline_min_change == s0->st.bspline_min_change)
                             ? same
                             : rest)
                            .push_back(s);
                    part.swap(rest);
                    int *d = upload_ids(same, which++);
                    if (!d) return c_->fail(MRPOP_ERR_BUDGET, "slot-id staging full");
                    const int n = (int)same.size();
                    plan::launch_sp_edges(d_slots_, d, n, stage, s0->st.bspline_step, s0->st.range,
                                          s0->st.bspline_min_change, st);
                    std::vector<const dev::EdgeGroup *> gs;
                    for (Slot *s : same) gs.push_back(&s->g_sp);
                    Staged ss;
                    if (int rc = stage_groups(MRPOP_MODE_FULL, gs, 1, same, cursor, ss)) return rc;
                    if (int rc = launch_staged(ss)) return rc;
                    plan::launch_sp_collect(d_slots_, d, n, st);
                    if (stage < s0->st.bspline_step) plan::launch_sp_smooth_apply(d_slots_, d, n, st);
                    else plan::launch_sp_shortcut(d_slots_, d, n, st);
                }
            }
        }
    }
    // ---- reports
    std::vector<Slot *> all;
    all.insert(all.end(), rm.begin(), rm.end());
    all.insert(all.end(), dr.begin(), dr.end());
    all.insert(all.end(), sp.begin(), sp.end());
    {
        int *d = upload_ids(all, 7);
        plan::launch_report(d_slots_, d, (int)all.size(), d_report_, st);
        MRPOP_HIP(c_, hipMemcpyAsync(h_report_, d_report_, sizeof(SlotReport) * cap_, hipMemcpyDeviceToHost, st));
        MRPOP_HIP(c_, hipStreamSynchronize(st));
    }
    // ---- state machines
    for (Slot *s : rm) {
        const SlotReport &r = h_report_[s->id];
        s->dv.rm_round++;
        s->res.roadmap_rounds++;
        bool all_connected = true, at_cap = false;
        for (int a = 0; a < s->A; ++a) {
            all_connected &= r.rm_connected[a] != 0;
            at_cap |= r.rm_n[a] >= s->dv.rm_cap;
            s->res.roadmap_size[a] = r.rm_n[a];
        }
        if (!s->rm_until_connected || all_connected || at_cap) {
            s->rm_until_connected = false;
            s->phase = Phase::Search;
            s->dr_round = 0;
            s->dv.round_reset = 1;  // a fresh dRRT launch
            s->dv.round_limit = s->dr_budget;
        }
    }
    for (Slot *s : dr) {
        const SlotReport &r = h_report_[s->id];
        s->dr_round = r.ctr[plan_ctr_round];
        s->res.rounds = r.ctr[plan_ctr_run];
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
            if (s->pass == 0) {
                s->phase = Phase::Roadmap;  // one more roadmap round, then a fresh dRRT launch
            } else if (int rc = finish_pass(s, r)) {
                return rc;
            }
        }
    }
    for (Slot *s : sp) {
        const SlotReport &r = h_report_[s->id];
        s->sp_stage++;
        if (s->sp_stage > s->st.bspline_step) {
            s->best_cost = r.best_cost;
            if (int rc = finish_pass(s, r)) return rc;
        }
    }
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
