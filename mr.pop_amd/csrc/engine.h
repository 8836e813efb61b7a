// engine.h -- host engine running many planning problems ("slots") on one
// GPU, one round of every active slot per step.
//
// Reference host flow (src/planning/pop_planner.cu:739-1030): for each AO-x
// pass, reset the trees; in pass 0 grow the per-arm roadmaps until every
// arm's start and goal connect, then alternate dRRT launches with one more
// roadmap round until solved; in later passes (optimize) grow the roadmaps
// one round and run one optimising dRRT launch; simplify the path after a
// solved pass and keep it if it improves the cost.  Here each slot runs that
// flow as a state machine, and a step advances every active slot by one
// round with shared kernel launches.
#pragma once

#include <chrono>
#include <map>
#include <cstdint>
#include <vector>

#include "internal.h"
#include "planner.h"

namespace mrpop {

struct SolveOptions {
    int block_per_robot = 8;       // roadmap samples per arm per round = 230 * block_per_robot
    double time_budget_s = 30.0;   // wall-clock budget of one problem (<= 0: none)
    int roadmap_capacity = 100000; // nodes per arm roadmap (MAX_ROADMAP_SIZE, roadmap_interior.cuh:21)
    int tree_capacity = 0;         // 0: min(settings.max_samples, 262144)
    int neighbour_capacity = 128;  // adjacency entries per roadmap node
    int neighbour_candidates = 64; // radius neighbours examined per new roadmap node
    int rounds_per_step = 8;       // dRRT rounds queued per host step
    int probe_samples = 0;         // > 0: planner edges are checked at this many samples first (see launch_staged)
    int graph_rounds = 0;          // 1: replay a step's dRRT rounds from a captured HIP graph
    int arm_split = 0;             // 1: one check launch per arm count, not the mixed-arm kernel
    int biased_iter = 10000;       // BIASED_ITER (roadmap_interior.cuh:24); 0: no biased roadmap samples
    int roadmap_cache = 0;         // 1: start from, and contribute to, the context's roadmaps of the same scene
};

struct SeedPath {
    float cost;
    std::vector<float> path;  // n x D
};

struct SlotResult {
    bool solved = false;
    float cost = 0.f;
    std::vector<float> path;  // n x D, goal -> start
    std::vector<SeedPath> seed_paths;
    int start_tree_size = 0, goal_tree_size = 0, iters = 0;
    uint64_t wall_ns = 0, kernel_ns = 0, copy_ns = 0;
    uint64_t ttfs_ns = 0;          // submit -> first solution
    std::vector<uint64_t> improve_ns;  // submit -> each improving solution
    int passes = 0, rounds = 0, roadmap_rounds = 0;
    int roadmap_cached[geom::kMaxArms] = {};  // nodes each arm's roadmap started with from the cache
    int roadmap_size[geom::kMaxArms] = {};
    int status = 0;                // MRPOP_OK, MRPOP_ERR_TIMEOUT, MRPOP_ERR_BUDGET, MRPOP_ERR_INTERNAL
    int events[kEvents] = {};      // device event counters (kEv*) at the end of the solve
};

enum class Phase { Roadmap, Search, Simplify, Done };

struct Slot {
    int id = -1;
    int A = 0, D = 0, n_goals = 0;
    mrpop_settings st{};
    SolveOptions opt{};
    unsigned long long seed = 0;
    std::vector<float> sg;  // (1 + n_goals) x D
    // device memory
    void *mem = nullptr;
    size_t mem_bytes = 0;
    float *wbuf = nullptr;
    dev::WorldDesc wdesc{};
    SlotDev dv{};
    dev::EdgeGroup g_valid[geom::kMaxArms], g_rm[geom::kMaxArms], g_dr, g_cn, g_sp;
    // state machine
    Phase phase = Phase::Roadmap;
    int pass = 0;
    bool rm_until_connected = true;
    bool seeking = true;   // the pass looks for a first solution (no optimise mode, no hyperspheroid)
    bool pass_solved = false;
    int dr_round = 0, dr_budget = 0;
    int tree_n = 0;  // larger tree of the dRRT launch at the last report (0 after a reset)
    int sp_stage = 0;
    float best_cost = 3.402823466e38f;
    std::chrono::steady_clock::time_point t_submit, t_kernel_start;
    SlotResult res;
    void *user = nullptr;
    int lane = -1;  // lane running this slot's work, -1 none
    // roadmap reuse: each arm's cache key, and whether the query's nodes
    // (known after the insertion round) allow sharing the roadmaps
    std::vector<float> rm_key[geom::kMaxArms];
    int rm_cached[geom::kMaxArms] = {};  // nodes taken from the cache per arm
    unsigned rm_kept = 0;                // arms whose roadmap was kept when it first connected the query
};

// Per-slot report copied back after each step.
struct SlotReport {
    int status;
    int rm_connected[geom::kMaxArms];
    int rm_n[geom::kMaxArms];
    int tr_n[2];
    int path_n;
    float path_cost;
    float best_cost;
    int path_meta[3];
    int sp_count;
    int ctr[4];  // device round counters (kCtr*)
    int events[kEvents];  // device event counters (kEv*)
};
constexpr int plan_ctr_round = kCtrRound, plan_ctr_run = kCtrRun;

// Prolate-hyperspheroid sampling transform of an optimise pass
// (pop_planner.cu:289-346): tf (D x D, row-major) maps the unit ball onto the
// set {x : |x - f1| + |x - f2| <= cb}, centred at center = (f1 + f2) / 2.
void phs_transform(int D, const float *f1, const float *f2, double cb, float *tf, float *center);

// Work runs on two lanes, each a HIP stream with its own staging: the
// roadmap lane grows the per-arm roadmaps of slots in the Roadmap phase, the
// search lane runs dRRT rounds and simplification stages.  A slot is on at
// most one lane at a time; its state machine advances when that lane's work
// completes, so a long roadmap round does not hold back the searches.
class Engine {
  public:
    explicit Engine(Context *c) : c_(c) {}
    ~Engine();
    // Allocate a slot for a problem; returns MRPOP_OK and the slot in *out.
    int add(const mrpop_scene &scene, int n_arms, const float *bases, const float *start, int n_goals,
            const float *goals, unsigned long long seed, const mrpop_settings &st, const SolveOptions &opt,
            Slot **out);
    // The argument checks of add, without allocating: MRPOP_OK or
    // MRPOP_ERR_INVALID (message in the context).  Sizes from slot_bytes are
    // meaningful only for arguments that pass.
    int validate(const mrpop_scene &scene, int n_arms, const float *bases, const float *start, int n_goals,
                 const float *goals, const mrpop_settings &st, const SolveOptions &opt) const;
    // Release a slot's device memory (waits for its lane if it is running).
    void release(Slot *s);
    // Launch work on idle lanes, wait for one lane and advance its slots.
    int step();
    bool idle() const;
    size_t slot_bytes(int n_arms, const mrpop_settings &st, const SolveOptions &opt) const;
    const std::vector<Slot *> &slots() const { return slots_; }
    size_t device_bytes() const { return bytes_in_use_; }
    // Diagnostics: run `rounds` roadmap rounds of slot s (no search) and copy
    // arm a's roadmap out: nodes [n][8], degrees [n], neighbour lists [n][K].
    int debug_roadmap(Slot *s, int rounds);
    int debug_search(Slot *s, int rounds, bool keep_round);
    int debug_simplify(Slot *s, int stage);
    int debug_start_pass(Slot *s, int pass);
    // uf (optional): union-find parents [n]; connected (optional): the arm's
    // start-goal connectivity flag as the device last computed it.
    int debug_copy_roadmap(Slot *s, int arm, float *q, int *deg, int *nbr, int *uf, int *connected, int *n,
                           int cap);

  private:
    struct Staged {
        int mode = 0;
        struct Bucket {
            int A;
            size_t first;
            int n;
            size_t max_waves;
            size_t probe_waves;  // grid bound of the probe pass
        };
        int probe = 0;  // probe pass samples per edge (0: one pass)
        bool sparse = false;  // counts usually far below the capacity bound (persistent launch)
        std::vector<Bucket> buckets;
    };
    struct Lane {
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        bool busy = false;
        std::vector<Slot *> table;  // slot of each entry of d_slots / the report
        std::vector<Slot *> rm, dr, sp;
        SlotDev *d_slots = nullptr, *h_slots = nullptr;
        int *d_ids = nullptr, *h_ids = nullptr;
        size_t id_cursor = 0;
        dev::EdgeGroup *d_groups = nullptr, *h_groups = nullptr;
        size_t group_cursor = 0;
        unsigned *d_wave = nullptr;
        SlotReport *d_report = nullptr, *h_report = nullptr;
        // dRRT round sequences captured on this lane, keyed by every launch
        // argument they bake in (buffers, counts, grid bounds)
        struct Graph {
            std::vector<uint64_t> key;
            hipGraphExec_t exec = nullptr;
            uint64_t used = 0;
        };
        std::vector<Graph> graphs;
        uint64_t graph_clock = 0;
        bool abandoned = false;  // its work outlived the deadline; its problems were ended
        std::chrono::steady_clock::time_point t_launch;  // when its running step was enqueued
    };
    int wait_lane(Lane &L);
    int wait_any(Lane *const *cands, int n, Lane **done);
    void abandon_lane(Lane &L);
    void drop_graphs(Lane &L);
    int ensure_capacity(int n);
    int ensure_lanes();
    int finish_pass(Slot *s, const SlotReport &r);
    int start_pass(Slot *s);
    void compute_phs(Slot *s);
    int launch_lane(Lane &L);
    int process_lane(Lane &L);
    int drain();
    int *upload_ids(Lane &L, const std::vector<Slot *> &list);
    int stage_groups(Lane &L, int mode, const std::vector<const dev::EdgeGroup *> &groups_by_slot, int per_slot,
                     const std::vector<Slot *> &list, Staged &out);
    int launch_staged(Lane &L, const Staged &st);
    int run_aux(Slot *s, void (*launch)(SlotDev *, const int *, int, hipStream_t));
    void release_pool();
    int cache_fetch(Slot *s);
    int cache_store(Slot *s, const int *rm_n, unsigned arms);
    Context *c_;
    std::vector<Slot *> slots_;
    std::vector<std::pair<int, void *>> graveyard_;  // (lane, block) of slots released while their lane was stuck
    int cap_ = 0;
    // Lane 0 runs roadmap rounds; lanes 1..kSearchLanes run dRRT rounds and
    // simplification, a slot on lane 1 + id % kSearchLanes (or, while fewer
    // problems seek a first solution than optimise, seekers on lane 1 and the
    // rest on the last lane; Engine::step).  A round is a chain of small
    // dependent launches that leave most of the GPU idle, so the search
    // lanes' chains run side by side on their own streams.  Two
    // search lanes: the 64-problem mixed batch plans 18.5 -> 21.3 problems/s,
    // the reference's 50-query fixture batches 9.3 -> 11.3 (4 arms) and 9.2 ->
    // 11.2 (5 arms); three ran slower than two (17.6 / 9.5 / 9.1: with the
    // context and roadmap streams, five streams over the process's four
    // hardware queues).
    static constexpr int kSearchLanes = 2;
    Lane lanes_[1 + kSearchLanes];
    dev::WorldDesc *d_worlds_ = nullptr, *h_worlds_ = nullptr;
    SlotDev *d_aux_slot_ = nullptr, *h_aux_slot_ = nullptr;  // one-slot table for set-up kernels
    int *d_aux_id_ = nullptr;
    size_t group_cap_ = 0;
    size_t bytes_in_use_ = 0;
    // search-lane steps while first solutions are pending; optimise passes
    // run on one of every kOptimiseShare of them
    static constexpr uint64_t kOptimiseShare = 4;
    uint64_t search_steps_[kSearchLanes] = {};
    // released slot blocks by size, reused by later problems (serving)
    static constexpr size_t kPoolBytes = size_t(16) << 30;
    std::multimap<size_t, void *> pool_;
    size_t pool_bytes_ = 0;
};

}  // namespace mrpop
