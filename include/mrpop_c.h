/*
 * mrpop_c.h -- C ABI of the MI355X MR.POP planner library (libmrpop.so).
 *
 * Plain C: pointers, sizes and POD structs only.  Every entry point returns
 * an int status (MRPOP_OK == 0); the text of the last error of a context is
 * available from mrpop_last_error().  Nothing here calls exit().
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   dRRT::solve / PlannerResult        src/planning/Planners.hh:14-32, 57-64
 *   pop_settings                       src/planning/pop_settings.hh:3-32
 *   Environment<float> and shapes      src/collision/environment.hh:11-40,
 *                                      src/collision/shapes.hh:27-185
 *   fkcc_single_buffer{,_mr,_drrt}     src/planning/utils.cuh:447-686
 *   fk / fk_approx (sphere centres)    src/robots/panda_four.cuh:44-131, 289-373
 * The reference has no C ABI of its own (its host API is the C++ template
 * dRRT::solve); these are the entry points its C++ template binds to in this
 * library -- see mr.pop_amd/src/planning/Planners.hh.
 */
#ifndef MRPOP_C_H
#define MRPOP_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MRPOP_OK 0
#define MRPOP_ERR_HIP 1        /* a HIP runtime call failed                 */
#define MRPOP_ERR_INVALID 2    /* bad argument (sizes, arm count, ...)      */
#define MRPOP_ERR_BUDGET 3     /* device-memory budget or capacity exceeded */
#define MRPOP_ERR_TIMEOUT 4    /* wall-clock budget ran out                 */
#define MRPOP_ERR_NO_DEVICE 5  /* no usable GPU                             */
#define MRPOP_ERR_STATE 6      /* call not valid in the object's state      */
#define MRPOP_ERR_INTERNAL 7   /* a device-side step bound was reached (union-find walk); problem ended */

/* ---- collision modes (which reference checker an edge check follows) ---- */
/* Multi-arm, environment + self + arm-arm: fkcc_single_buffer (utils.cuh:447-534). */
#define MRPOP_MODE_FULL 0
/* Multi-arm, self + arm-arm only, no environment: fkcc_single_buffer_drrt (utils.cuh:626-686). */
#define MRPOP_MODE_DRRT 1
/* One arm with its base pose, environment + self: fkcc_single_buffer_mr (utils.cuh:536-624). */
#define MRPOP_MODE_SINGLE 2

/* ---- scene ------------------------------------------------------------- */
/*
 * Flattened Environment<float> (environment.hh:11-40).  Each list is an array
 * of float records:
 *   spheres    4 floats: x y z r
 *   capsules   8 floats: x1 y1 z1 xv yv zv r rdv   (rdv = 1 / |v|^2)
 *   cuboids   15 floats: x y z  a1x a1y a1z  a2x a2y a2z  a3x a3y a3z  r1 r2 r3
 * z_capsules / z_cuboids are the reference's z-aligned variants (same record
 * layout, tested with the z-aligned formulas of utils.cuh:165-228).
 * Environment::cylinders is never tested by the reference checker
 * (utils.cuh:231-260), so it has no slot here; cylinders from problem files
 * are passed as capsules, as the reference's harness does
 * (scripts/evaluate_mr.cu:38-47).
 */
typedef struct mrpop_scene {
    int n_spheres;
    int n_capsules;
    int n_z_capsules;
    int n_cuboids;
    int n_z_cuboids;
    const float *spheres;
    const float *capsules;
    const float *z_capsules;
    const float *cuboids;
    const float *z_cuboids;
} mrpop_scene;

#define MRPOP_SPHERE_FLOATS 4
#define MRPOP_CAPSULE_FLOATS 8
#define MRPOP_CUBOID_FLOATS 15
#define MRPOP_MAX_OBSTACLES 4096
#define MRPOP_MAX_ARMS 5

/* ---- planner settings: pop_settings (pop_settings.hh:3-32), same names,
 *      same defaults (mrpop_settings_default). ----------------------------- */
typedef struct mrpop_settings {
    int max_samples;
    int max_iters;
    int optimize_iters;
    int num_new_configs;
    int granularity;
    float range;
    int balance;
    float tree_ratio;
    int dynamic_domain;        /* accepted; never skips an expansion, as in the reference (radii stay FLT_MAX) */
    float dd_alpha;
    float dd_radius;
    float dd_min_radius;
    int roadmap_reassign;      /* 1: arms whose start and goal are connected draw no more roadmap samples */
    int rrtc_iter;
    int resample_iter;
    int cost_bound_resample;
    int path_simplify;
    int phs;
    int simplify_round;
    int bspline_step;
    float bspline_min_change;
    int perturb_step;
    int perturb_empty_step;
} mrpop_settings;

/* ---- library / context ------------------------------------------------- */
typedef struct mrpop_ctx mrpop_ctx;

/* Library version string. */
const char *mrpop_version(void);

/* Fill *s with the reference's pop_settings defaults (pop_settings.hh:4-31). */
int mrpop_settings_default(mrpop_settings *s);

/* Default world<-base poses (row-major 3x4 per arm) for 2..5 arms.
 * 4 and 5 arms: the reference's tables (panda.cuh:16-73). */
int mrpop_default_bases(int n_arms, float *out_bases /* n_arms*12 */);

/* Create a context on HIP device `device`; owns a stream and device memory. */
int mrpop_create(int device, mrpop_ctx **out);
int mrpop_destroy(mrpop_ctx *ctx);
/* Last error text of the context (never NULL). */
const char *mrpop_last_error(const mrpop_ctx *ctx);
/* The context's HIP stream (hipStream_t) as an opaque pointer. */
void *mrpop_stream(mrpop_ctx *ctx);

/* ---- collision world uploaded to the device ---------------------------- */
typedef struct mrpop_world mrpop_world;

/* Upload a scene plus the arm layout (n_arms in 1..5; bases NULL = default
 * poses of mrpop_default_bases, n_arms*12 floats otherwise). */
int mrpop_world_create(mrpop_ctx *ctx, const mrpop_scene *scene, int n_arms,
                       const float *bases, mrpop_world **out);
int mrpop_world_destroy(mrpop_world *w);

/*
 * Batched edge check (replaces fkcc_single_buffer{,_mr,_drrt}, utils.cuh:447-686,
 * driven by the interpolation of pop_planner.cu:553-560).
 * Edge e goes from from[e*D..] to to[e*D..] with D = 7*n_arms (mode FULL/DRRT)
 * or D = 7 (mode SINGLE, arm `arm` of the world).  Sample k = 1..granularity
 * is q = from + k * ((to - from) / granularity): the start is excluded and
 * the end included, as in the reference.  out_edge[e] = 1 iff any sample
 * collides under the reference's two-phase composite check.
 * out_mask (optional) receives bit (k-1) set for each colliding sample.
 * granularity in 1..64.
 *
 * mrpop_check_edges: host buffers, synchronous.
 * mrpop_check_edges_device: device buffers, enqueued on `stream` (a
 * hipStream_t; NULL = the context stream), no synchronisation, no allocation.
 */
int mrpop_check_edges(mrpop_ctx *ctx, const mrpop_world *w, int mode, int arm, int granularity,
                      int n_edges, const float *from, const float *to, uint8_t *out_edge,
                      uint64_t *out_mask);
int mrpop_check_edges_device(mrpop_ctx *ctx, const mrpop_world *w, int mode, int arm,
                             int granularity, int n_edges, const float *d_from,
                             const float *d_to, uint32_t *d_out_edge, uint64_t *d_out_mask,
                             void *stream);

/*
 * mrpop_check_edges without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture set-up, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and synchronous behaviour as mrpop_check_edges,
 * but the host inputs and outputs are staged through one fixed array in
 * device memory per GPU (MRPOP_NOALLOC_SCRATCH_BYTES, held by each GPU the
 * library runs on until the process ends) instead of context buffers.
 * Staging needs 2*D*4 bytes per edge for the endpoints, 5 per edge for the
 * verdicts and 8 more with out_mask (each part rounded up to 256 bytes);
 * larger calls fail with MRPOP_ERR_BUDGET before anything runs.  Every call
 * on one GPU shares the array: calls that overlap on one GPU (two threads,
 * two contexts) race and must be serialised by the caller.
 */
#define MRPOP_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int mrpop_check_edges_noalloc(mrpop_ctx *ctx, const mrpop_world *w, int mode, int arm, int granularity,
                              int n_edges, const float *from, const float *to, uint8_t *out_edge,
                              uint64_t *out_mask);

/*
 * Sphere centres by forward kinematics (replaces fk / fk_approx,
 * panda_four.cuh:44-131, 289-373, and fk_mr / fk_approx_mr, panda.cuh:598-829).
 * For n_cfg configurations of the world's n_arms arms (D = 7*n_arms floats
 * each): out_coarse[c][arm][11][3], out_fine[c][arm][59][3] (either may be NULL).
 */
int mrpop_sphere_centres(mrpop_ctx *ctx, const mrpop_world *w, int n_cfg, const float *cfgs,
                         float *out_coarse, float *out_fine);

/* ---- planning ----------------------------------------------------------- */
/* One planning query (the inputs of dRRT::solve, Planners.hh:57-64). */
typedef struct mrpop_problem {
    int n_arms;                  /* 2..5 */
    const float *bases;          /* n_arms*12 world<-base (row-major 3x4); NULL = mrpop_default_bases */
    const mrpop_scene *scene;
    const float *start;          /* 7*n_arms */
    int n_goals;                 /* 1..8 */
    const float *goals;          /* n_goals * 7*n_arms */
    uint64_t seed;               /* random stream of this problem */
} mrpop_problem;

/* Resources of one problem.  block_per_robot is dRRT::solve's argument
 * (roadmap samples per arm per round = 230 * block_per_robot, as
 * ROADMAP_BUILD_ITER * block_per_robot in roadmap_interior.cuh:20, 97). */
typedef struct mrpop_solve_options {
    int block_per_robot;       /* default 8 */
    double time_budget_s;      /* wall-clock budget per problem; <= 0: none (default 30) */
    int roadmap_capacity;      /* nodes per arm roadmap (default 100000 = MAX_ROADMAP_SIZE) */
    int tree_capacity;         /* nodes per tree (0: min(max_samples, 262144)) */
    int neighbour_capacity;    /* adjacency entries per roadmap node (default 128) */
    int neighbour_candidates;  /* radius neighbours tried per new roadmap node, 1..64 (default 64) */
    int rounds_per_step;       /* dRRT rounds queued between host checks (default 8) */
    int probe_samples;         /* > 0: planner edges are first checked at this many samples spread along
                                  them, then in full only if still free; same verdicts (default 0) */
    int graph_rounds;          /* 1: replay each step's dRRT rounds from a captured HIP graph (default 0;
                                  same results, no measured gain; not for runs under a kernel tracer) */
    int arm_split;             /* batches of several arm counts: 1 = one check launch per arm count (tiles of
                                  64 samples), 0 = one launch of the mixed-arm kernel (default 0) */
    int biased_iter;           /* roadmap steering (BIASED_ITER, roadmap_interior.cuh:24, 117-165): while the
                                  first roadmap rounds of a solve run, every biased_iter-th sample of each
                                  block of 230 samples per round is moved to within `range` of its nearest
                                  roadmap node (default 10000 as in the reference; 0: never) */
    int roadmap_cache;         /* 1: reuse roadmaps across queries of one scene (no counterpart in the reference,
                                  which rebuilds them per solve, pop_planner.cu:812-830): each arm starts from
                                  the context's roadmap for the same obstacles, base pose, granularity, range and
                                  roadmap sizes, if there is one, with the query's start and goals inserted as
                                  new nodes that take the nearest roadmap nodes within range as candidate
                                  neighbours; roadmaps of queries whose start and goals are collision-free
                                  are kept, while at most half the roadmap capacity, when each arm first
                                  connects them and when the query ends (least recently used ones leave
                                  beyond 8 GiB per context).  Results then depend on the queries that ran before.
                                  Default 0 */
} mrpop_solve_options;

int mrpop_solve_options_default(mrpop_solve_options *o);

/* Result fields of PlannerResult (Planners.hh:14-32) plus timing of the
 * first solution and the status of the solve. */
typedef struct mrpop_result_info {
    int solved;
    float cost;
    int path_length;           /* waypoints of the final path */
    int start_tree_size;
    int goal_tree_size;
    int iters;                 /* dRRT round of the final solution */
    uint64_t wall_ns;          /* submit -> done */
    uint64_t kernel_ns;        /* first kernel -> done */
    uint64_t copy_ns;          /* result copies */
    uint64_t ttfs_ns;          /* submit -> first solution (0 if none) */
    int n_seed_paths;          /* improving solutions recorded (PlannerResult::seed_paths) */
    int passes;                /* AO-x passes run */
    int rounds;                /* dRRT rounds run */
    int roadmap_rounds;
    int roadmap_size[MRPOP_MAX_ARMS];
    int status;                /* MRPOP_OK; MRPOP_ERR_TIMEOUT if the time budget ended it;
                                  MRPOP_ERR_INVALID if its settings, options or configurations
                                  were rejected (granularity outside 1..64, num_new_configs
                                  outside 1..65535, range not positive, a non-finite value):
                                  the problem is not planned, the others of a batch are;
                                  MRPOP_ERR_BUDGET if it needs more device memory than the
                                  whole HBM budget (refused alone, the queue goes on);
                                  MRPOP_ERR_INTERNAL if union_find_bound below is set */
    /* Bounded work cut short during the solve (counts; none affects a returned path's validity): */
    int union_find_bound;      /* union-find walks/unions that reached their step bound (ends the problem) */
    int edge_table_full;       /* expansions whose duplicate-edge lookup found the table full (checked alone) */
    int connect_skipped;       /* connections longer than 64 chunks per expansion, not attempted */
    int simplify_dropped;      /* simplification candidates beyond the chunk staging, kept as colliding */
    int roadmap_cached[MRPOP_MAX_ARMS];  /* nodes each arm's roadmap took from the roadmap cache (0: built anew) */
} mrpop_result_info;

typedef struct mrpop_solution mrpop_solution;

/* dRRT + AO-x (pop_planner.cu:739-1030) on one problem.  settings NULL =
 * pop_settings defaults, options NULL = mrpop_solve_options_default. */
int mrpop_solve(mrpop_ctx *ctx, const mrpop_problem *problem, const mrpop_settings *settings,
                const mrpop_solve_options *options, mrpop_solution **out);
/* n problems planned together (one round of every active problem per step;
 * problems are admitted while their device memory fits).  settings and
 * options are arrays of n (or NULL for defaults); out receives n solutions.
 * Returns MRPOP_ERR_INVALID for a malformed problem (null pointers, arm
 * count outside 2..5, no goal); a problem rejected on admission gets a
 * solution with status MRPOP_ERR_INVALID (or MRPOP_ERR_BUDGET when it alone
 * exceeds the device memory budget) instead, and the others are planned. */
int mrpop_solve_batch(mrpop_ctx *ctx, int n, const mrpop_problem *problems, const mrpop_settings *settings,
                      const mrpop_solve_options *options, mrpop_solution **out);

int mrpop_solution_info(const mrpop_solution *s, mrpop_result_info *info);
/* which = -1: final path; which = k >= 0: k-th improving solution.  Copies at
 * most cap waypoints (7*n_arms floats each, goal -> start); *n_waypoints gets
 * the path length, *cost its cost (either may be NULL). */
int mrpop_solution_path(const mrpop_solution *s, int which, float *out, int cap, int *n_waypoints, float *cost);
/* Times (ns since submit) at which each improving solution was found. */
int mrpop_solution_improvements(const mrpop_solution *s, uint64_t *out_ns, int cap, int *n);
int mrpop_solution_free(mrpop_solution *s);

/* ---- serving: continuous batching over one context --------------------- */
typedef struct mrpop_server mrpop_server;

/* hbm_budget_bytes bounds the device memory of admitted problems (0: 90% of
 * the free memory); max_active bounds their number (0: unbounded). */
int mrpop_server_create(mrpop_ctx *ctx, size_t hbm_budget_bytes, int max_active, mrpop_server **out);
int mrpop_server_destroy(mrpop_server *srv);
/* Queue a problem; returns its ticket. */
int mrpop_server_submit(mrpop_server *srv, const mrpop_problem *problem, const mrpop_settings *settings,
                        const mrpop_solve_options *options, uint64_t *ticket);
/* Admit waiting problems that fit, then run one round of every active
 * problem.  Reports the number of active and waiting problems. */
int mrpop_server_step(mrpop_server *srv, int *n_active, int *n_waiting);
/* state: 0 waiting, 1 active, 2 done, -1 unknown ticket.  When out is not
 * NULL it receives a snapshot of the best solution so far (free it with
 * mrpop_solution_free). */
int mrpop_server_poll(mrpop_server *srv, uint64_t ticket, int *state, mrpop_solution **out);
/* Drop a waiting or active problem. */
int mrpop_server_cancel(mrpop_server *srv, uint64_t ticket);
/* Device bytes held by active problems. */
size_t mrpop_server_bytes(const mrpop_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* MRPOP_C_H */
