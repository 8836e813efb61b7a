// pop_settings.hh -- planner settings, field for field and default for
// default those of the reference (src/planning/pop_settings.hh:3-32), so that
// code written against the reference compiles unchanged.  The C ABI mirror is
// mrpop_settings (include/mrpop_c.h); dRRT::solve converts one to the other.
#pragma once

struct pop_settings {
    int max_samples = 1000000;   // tree capacity bound (the library caps a tree at 262144 nodes)
    int max_iters = 1000;        // rounds of a first-solution dRRT launch
    int optimize_iters = 400000; // rounds of an optimising dRRT launch
    int num_new_configs = 256;   // expansions per round
    int granularity = 16;        // samples per checked edge (1..64)
    float range = 0.5;           // roadmap radius and steering range

    int balance = 2;             // 0 none, 1 or 2: tree balancing strategy
    float tree_ratio = 1.0;

    // accepted; behaves as in the reference, where every node radius starts at
    // FLT_MAX and only finite radii grow (pop_planner.cu:534-537, 591-604,
    // 860-863), so the radius test never skips an expansion
    bool dynamic_domain = false;
    float dd_alpha = 0.0001;
    float dd_radius = 4.0;
    float dd_min_radius = 1.0;

    bool roadmap_reassign = false;  // arms whose start and goal are connected draw no more roadmap samples

    int rrtc_iter = 1;           // AO-x passes
    int resample_iter = 5;
    bool cost_bound_resample = false;
    bool path_simplify = true;
    bool phs = true;             // prolate-hyperspheroid sampling in optimising passes

    int simplify_round = 1;
    int bspline_step = 3;        // smoothing stages before the shortcut stage
    float bspline_min_change = 0.01;
    int perturb_step = 3;
    int perturb_empty_step = 2;
};
