// planner.h -- device-side state of one planning problem ("slot") and the
// round kernels of the dRRT / AO-x planner (planner.hip).
//
// The reference grows its roadmaps and trees with one thread block per
// expansion stream and inter-block races (src/planning/pop_planner.cu:173-737,
// roadmap_interior.cuh:94-339).  Here every step is round-synchronous: all
// expansions of a round read the trees as they stood at the round's start,
// their edges are checked in one batched launch, and the results are
// committed in expansion order.  Randomness is a counter-based generator
// keyed by (problem seed, stream, round, item), so a problem's result does
// not depend on which other problems share the GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "check_groups.h"
#include "device_geom.h"

namespace mrpop {

constexpr int kMaxGoals = 8;
// Roadmap samples one block of the reference draws per round
// (ROADMAP_BUILD_ITER, roadmap_interior.cuh:20): a round has
// kRoadmapIter * block_per_robot samples per arm.
constexpr int kRoadmapIter = 230;
constexpr float kUnwritten = -9999.0f;  // pop_planner_state.cuh:10

// Slot status codes (device int).
constexpr int kRunning = 0;
// > 0: solved, value = 1 + winning expansion
constexpr int kFailedOverflow = -1;  // tree capacity reached (pop_planner.cu:587-589)

// Device round counters of a slot.  Several dRRT rounds run per host step,
// so the round index lives on the device: k_dr_finish advances it.
enum : int {
    kCtrRound = 0,  // round within the current dRRT launch
    kCtrRng = 1,    // dRRT rounds since the solve began (random stream position)
    kCtrLimit = 2,  // rounds allowed in the current launch
    kCtrRun = 3,    // dRRT rounds run since the solve began
};

// Device event counters of a slot: bounded work that was cut short.  None
// changes a returned path's validity; each is reported with the result.
enum : int {
    kEvUnionFind = 0,     // a union-find walk or union hit its step bound (connectivity may be incomplete)
    kEvEdgeTable = 1,     // the round's duplicate-edge table was full: the expansion checked its own edge
    kEvConnectSkip = 2,   // a connection needed more chunks than the budget per expansion: not attempted
    kEvSimplifyDrop = 3,  // a simplification candidate did not fit the chunk staging: kept as colliding
    kEvents = 4,
};

struct SlotDev {
    // problem
    int A, D, n_goals, world;
    unsigned long long seed;
    // settings (pop_settings.hh)
    int G, B, M;  // granularity, num_new_configs, roadmap samples per arm per round
    float range, tree_ratio;
    int balance, phs;
    // per-pass / per-round state written by the host before each step
    int optimize, rm_round, pass;
    int round_reset;          // 1: k_dr_begin restarts the round count of a dRRT launch
    int round_limit;          // rounds of the current dRRT launch (max_iters / optimize_iters)
    int *ctr;                 // [4] device round counters, see kCtr*
    int *events;              // [kEvents] event counters, see kEv*
    float phs_center[35];
    float *sg;                // [1 + n_goals][D]: start, then goals
    // roadmaps (arm-major)
    int rm_cap, rm_K, rm_Kc;  // node capacity, neighbour-list capacity, candidates per new node
    float *rm_q;              // [A][rm_cap][8]
    int *rm_nbr;              // [A][rm_cap][rm_K]
    int *rm_deg;              // [A][rm_cap]
    int *rm_uf;               // [A][rm_cap]   union-find parent
    int *rm_n;                // [A]
    int *rm_connected;        // [A]
    int *rs_nsamp;            // [A] samples drawn this round (0: the arm sits the round out)
    int rm_reassign;          // roadmap_reassign: arms already connected draw no more samples
    int rm_biased_iter;       // every rm_biased_iter-th sample of a block is steered to its nearest node (0: never)
    int rm_stretch;           // 1 while the first roadmap rounds of the solve run (until the arms connect)
    int *rs_brounds;          // [A] rounds of that stretch in which the arm drew samples
    // roadmap reuse across queries of one scene (roadmaps taken from the
    // context's roadmap cache): the query's start and goals are nodes
    // rm_root0[a] + k of arm a's roadmap (k = 0 start, 1..n_goals goals)
    int *rm_root0;            // [A] (0 for a roadmap built from scratch)
    int rm_insert;            // 1: this round's samples are the query's start and goals (forced in as nodes)
    int rm_fill;              // node ids sampled rounds may take (rm_cap, less a reserve for later queries)
    int *rm_qvalid;           // [A] 1 if the inserted start and goals of the arm are collision-free
    // roadmap round staging
    float *rs_q;              // [A][M][7]
    uint32_t *rs_coll;        // [A][M]
    int *rs_id;               // [A][M]
    int *rs_ncand;            // [A][M]  radius candidates of each sample
    int *rs_first;            // [A][M]  first packed edge of each sample
    int *rs_cand;             // [A][M][Kc] staged candidates
    int *re_count;            // [A]
    int *rs_base;             // [A]  first node id of this round's samples
    // link staging: bitsets over this round's new ids for each older endpoint
    int rl_cap, rl_words;     // claimed bitsets per arm, 32-bit words per bitset
    int *rl_slot;             // [A][rm_cap]  bitset of a node this round (-1 none)
    int *rl_touched;          // [A][rl_cap]  node of each claimed bitset
    int *rl_ntouched;         // [A]
    unsigned *rl_bits;        // [A][rl_cap][rl_words]
    float *re_from, *re_to;   // [A][M*Kc][7]
    int *re_pair;             // [A][M*Kc][2]  (neighbour, new node)
    uint32_t *re_coll;        // [A][M*Kc]
    // trees
    int tr_cap;
    float *tr_q;              // [2][tr_cap][D]
    int *tr_parent;           // [2][tr_cap]
    float *tr_cost;           // [2][tr_cap]
    int *tr_rid;              // [2][tr_cap][A]  roadmap node of each arm
    int *tr_n;                // [2]
    // expansions
    int *ex_tree;             // [B] tree extended by expansion b (kept across rounds)
    int *ex_state;            // [B] 0 idle, 1 insert, 2 rewire
    int *ex_near;             // [B]
    int *ex_exist;            // [B]
    float *ex_cost;           // [B]
    float *ex_q;              // [B][D]
    int *ex_rid;              // [B][A]
    int *ex_index;            // [B] node index after insertion (-1 none)
    int *ex_dup;              // [B] expansion whose identical edge is checked for b (-1 none)
    float *ex_sample;         // [B][D] the round's sample of expansion b
    float *ex_crand;          // [B] its cost bound (optimise passes; FLT_MAX otherwise; -1 if rejected)
    unsigned long long *ex_nnkey;  // [B] five-launch form: (squared distance bits << 32 | index) of the nearest node
    int *ex_live;             // [B] five-launch form: 1 sampled, 2 stepped, 0 ended
    unsigned long long *ex_hash;  // [ex_hmask+1] edges of the round: (round tag << 32) | b
    int ex_hmask;
    unsigned long long *rw_key;  // [tr_cap*2] rewire arbitration (cost bits << 32 | b)
    // extension edges (DRRT check)
    int *ed_count;
    float *ed_from, *ed_to;   // [B][D]
    int *ed_b;                // [B]
    uint32_t *ed_coll;        // [B]
    // connection
    unsigned long long *cn_key;  // [B] split connection scan: (squared distance bits << 32 | index) of the nearest node
    int *cn_near;             // [B]
    int *cn_next;             // [B]
    float *cn_d;              // [B]
    uint32_t *cn_fail;        // [B]
    int ce_cap;
    int *ce_count;
    float *ce_from, *ce_to;   // [ce_cap][D]
    int *ce_b;                // [ce_cap]
    uint32_t *ce_coll;        // [ce_cap]
    // status and result
    int *status;              // kRunning, >0 solved, <0 failed
    float *best_cost;
    float *path_cost;         // cost of the recorded path
    int path_cap;
    float *path;              // [path_cap][D]  goal -> start
    int *path_n;
    int *path_meta;           // [0] start-tree nodes, [1] goal-tree nodes, [2] round of solution
    // PHS transform (row-major D x D)
    float *phs_tf;
    // simplification
    int sp_cap;               // edges
    int *sp_count;
    float *sp_from, *sp_to;   // [sp_cap][D]
    int *sp_item;             // [sp_cap] (candidate index)
    uint32_t *sp_coll;        // [sp_cap]
    uint32_t *sp_fail;        // [path_cap * 16]
    float *sp_mid;            // [path_cap][D]
    float *sp_cost;           // [path_cap]
    int *sp_parent;           // [path_cap]
};

// ----------------------------------------------------------------- RNG ----
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).
__host__ __device__ inline uint4 philox4x32_10(uint4 c, uint2 k) {
    for (int r = 0; r < 10; ++r) {
        const unsigned long long p0 = (unsigned long long)0xD2511F53u * c.x;
        const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c.z;
        const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
        const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

// Uniform in (0, 1] with 24 random bits, as curand_uniform's range.
__host__ __device__ inline float u01(unsigned x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

// Random streams.
enum : unsigned { kStreamRoadmap = 1, kStreamTree = 2, kStreamCostBound = 3, kStreamPhsRadius = 4 };

// Four uniforms for (stream, a, b, chunk) of a slot.
__host__ __device__ inline uint4 slot_random(unsigned long long seed, unsigned stream, unsigned a, unsigned b,
                                             unsigned chunk) {
    return philox4x32_10(make_uint4(a, b, chunk, stream), make_uint2((unsigned)seed, (unsigned)(seed >> 32)));
}

}  // namespace mrpop

namespace mrpop {
namespace plan {
// Kernel launchers (planner.hip); n = number of slot ids in `ids`.
void launch_roadmap_init(SlotDev *d, const int *ids, int n, hipStream_t stream);
void launch_tree_reset(SlotDev *d, const int *ids, int n, hipStream_t stream);
void launch_rm_sample(SlotDev *d, const int *ids, int n, int max_am, hipStream_t stream);
void launch_rm_assign(SlotDev *d, const int *ids, int n, hipStream_t stream);
void launch_rm_radius(SlotDev *d, const int *ids, int n, int max_am, hipStream_t stream);
void launch_rm_compact(SlotDev *d, const int *ids, int n, int max_am, hipStream_t stream);
void launch_rm_link(SlotDev *d, const int *ids, int n, int max_am, hipStream_t stream);
void launch_rm_connected(SlotDev *d, const int *ids, int n, hipStream_t stream);
void launch_dr_begin(SlotDev *d, const int *ids, int n, hipStream_t stream);
// splits > 0: the five-launch form, whose tree scans run a thread per
// expansion over `splits` slices of each tree (large trees); 0: one wave per
// expansion.  Both give the same results.
void launch_dr_expand(SlotDev *d, const int *ids, int n, int max_b, int splits, hipStream_t stream);
void launch_dr_commit(SlotDev *d, const int *ids, int n, hipStream_t stream);
// splits > 0: the nearest-node scans run a thread per expansion over
// `splits` slices of the other tree first (as the five-launch expansion)
void launch_dr_connect(SlotDev *d, const int *ids, int n, int max_b, int splits, hipStream_t stream);
void launch_dr_finish(SlotDev *d, const int *ids, int n, hipStream_t stream);
void launch_sp_edges(SlotDev *d, const int *ids, int n, int stage, int bspline_step, float range, float min_change,
                     hipStream_t stream);
void launch_sp_collect(SlotDev *d, const int *ids, int n, hipStream_t stream);
void launch_sp_smooth_apply(SlotDev *d, const int *ids, int n, hipStream_t stream);
void launch_sp_shortcut(SlotDev *d, const int *ids, int n, hipStream_t stream);
void launch_report(SlotDev *d, const int *ids, int n, void *out, hipStream_t stream);
}  // namespace plan
}  // namespace mrpop
