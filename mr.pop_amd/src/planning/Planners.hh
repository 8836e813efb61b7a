// Planners.hh -- the planner's C++ entry point with the signature and result
// type of the reference (src/planning/Planners.hh:14-32, 57-64):
//
//     PlannerResult<Robot> dRRT::solve<Robot>(start, goals, environment,
//                                             settings, num_robots,
//                                             block_per_robot);
//
// Header-only: the template flattens its arguments into the POD structs of
// the C ABI (include/mrpop_c.h) and calls mrpop_solve in libmrpop.so, which
// runs the dRRT + AO-x planner on the GPU.  Behaviour kept from the
// reference:
//   * the path runs goal -> start and keeps the duplicate waypoints that
//     simplification leaves; cost is the simplified path's cost;
//   * seed_paths holds every improving solution in the order found;
//   * errors do not throw: they are printed to stderr and the result says
//     solved == false;
//   * inputs are not modified.
// Differences: any number of goals (1..8) is accepted; 2- and 3-arm teams
// are supported; calls are reentrant (each thread may call solve; the
// process-wide context is shared under a lock); PlannerResult also carries
// the time to the first solution and to each improvement.
#pragma once

#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <mutex>
#include <sstream>
#include <vector>

#include "../../../include/mrpop_c.h"
#include "Robots.hh"
#include "pop_settings.hh"
#include "src/collision/environment.hh"

template <typename Robot>
struct PlannerResult {
    bool solved = false;
    std::vector<typename Robot::Configuration> path;             // goal -> start
    std::vector<typename Robot::Configuration> simplified_path;  // not filled (as in the reference)
    // every improving solution, in the order found (cost non-increasing)
    std::vector<std::pair<float, std::vector<typename Robot::Configuration>>> seed_paths;
    int start_tree_size = 0;
    int goal_tree_size = 0;
    int path_length = 0;
    int iters = 0;
    float cost = 0.0;
    std::size_t wall_ns = 0;    // wall time of the solve
    std::size_t kernel_ns = 0;  // first launch to the end
    std::size_t copy_ns = 0;    // result copies
    // not in the reference
    std::size_t ttfs_ns = 0;                 // time to the first solution (0 if none)
    std::vector<std::size_t> improve_ns;     // time of each seed path
};

template <typename Robot>
inline void print_cfg(typename Robot::Configuration &config) {
    for (int i = 0; i < Robot::dimension; i++) std::cout << config[i] << " ";
    std::cout << "\n";
}

template <typename Robot>
inline void print_cfg_to_ss(typename Robot::Configuration &config, std::stringstream &out) {
    for (int i = 0; i < Robot::dimension; i++) out << config[i] << " ";
    out << "\\n";
}

inline std::size_t get_elapsed_nanoseconds(const std::chrono::time_point<std::chrono::steady_clock> &start) {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - start).count();
}

namespace mrpop_cpp {

// One context per process on HIP device $MRPOP_DEVICE (default 0).
struct SharedContext {
    std::mutex lock;
    mrpop_ctx *ctx = nullptr;
    int status = MRPOP_OK;
    SharedContext() {
        const char *dev = std::getenv("MRPOP_DEVICE");
        status = mrpop_create(dev ? std::atoi(dev) : 0, &ctx);
    }
    ~SharedContext() {
        if (ctx) mrpop_destroy(ctx);
    }
};

inline SharedContext &shared_context() {
    static SharedContext sc;
    return sc;
}

// Random stream of dRRT::solve (the reference seeds its generators with 1234).
inline constexpr uint64_t kSolveSeed = 1234;

inline mrpop_settings to_c(const pop_settings &s) {
    mrpop_settings c{};
    mrpop_settings_default(&c);
    c.max_samples = s.max_samples;
    c.max_iters = s.max_iters;
    c.optimize_iters = s.optimize_iters;
    c.num_new_configs = s.num_new_configs;
    c.granularity = s.granularity;
    c.range = s.range;
    c.balance = s.balance;
    c.tree_ratio = s.tree_ratio;
    c.dynamic_domain = s.dynamic_domain;
    c.dd_alpha = s.dd_alpha;
    c.dd_radius = s.dd_radius;
    c.dd_min_radius = s.dd_min_radius;
    c.roadmap_reassign = s.roadmap_reassign;
    c.rrtc_iter = s.rrtc_iter;
    c.resample_iter = s.resample_iter;
    c.cost_bound_resample = s.cost_bound_resample;
    c.path_simplify = s.path_simplify;
    c.phs = s.phs;
    c.simplify_round = s.simplify_round;
    c.bspline_step = s.bspline_step;
    c.bspline_min_change = s.bspline_min_change;
    c.perturb_step = s.perturb_step;
    c.perturb_empty_step = s.perturb_empty_step;
    return c;
}

// Flattened obstacle records of an environment (cylinders are not tested,
// as in the reference's checker).
struct FlatScene {
    std::vector<float> spheres, capsules, z_capsules, cuboids, z_cuboids;
    mrpop_scene c{};

    explicit FlatScene(const ppln::collision::Environment<float> &env) {
        for (unsigned i = 0; i < env.num_spheres; ++i) {
            const auto &o = env.spheres[i];
            spheres.insert(spheres.end(), {o.x, o.y, o.z, o.r});
        }
        auto caps = [](std::vector<float> &out, const ppln::collision::Capsule<float> *a, unsigned n) {
            for (unsigned i = 0; i < n; ++i)
                out.insert(out.end(), {a[i].x1, a[i].y1, a[i].z1, a[i].xv, a[i].yv, a[i].zv, a[i].r, a[i].rdv});
        };
        caps(capsules, env.capsules, env.num_capsules);
        caps(z_capsules, env.z_aligned_capsules, env.num_z_aligned_capsules);
        auto cubs = [](std::vector<float> &out, const ppln::collision::Cuboid<float> *a, unsigned n) {
            for (unsigned i = 0; i < n; ++i)
                out.insert(out.end(), {a[i].x, a[i].y, a[i].z, a[i].axis_1_x, a[i].axis_1_y, a[i].axis_1_z,
                                       a[i].axis_2_x, a[i].axis_2_y, a[i].axis_2_z, a[i].axis_3_x, a[i].axis_3_y,
                                       a[i].axis_3_z, a[i].axis_1_r, a[i].axis_2_r, a[i].axis_3_r});
        };
        cubs(cuboids, env.cuboids, env.num_cuboids);
        cubs(z_cuboids, env.z_aligned_cuboids, env.num_z_aligned_cuboids);
        c.n_spheres = (int)(spheres.size() / 4);
        c.n_capsules = (int)(capsules.size() / 8);
        c.n_z_capsules = (int)(z_capsules.size() / 8);
        c.n_cuboids = (int)(cuboids.size() / 15);
        c.n_z_cuboids = (int)(z_cuboids.size() / 15);
        c.spheres = spheres.data();
        c.capsules = capsules.data();
        c.z_capsules = z_capsules.data();
        c.cuboids = cuboids.data();
        c.z_cuboids = z_cuboids.data();
    }
};

template <typename Robot>
inline std::vector<typename Robot::Configuration> to_configs(const mrpop_solution *sol, int which) {
    int n = 0;
    float cost = 0;
    std::vector<typename Robot::Configuration> out;
    if (mrpop_solution_path(sol, which, nullptr, 0, &n, &cost) != MRPOP_OK || n <= 0) return out;
    std::vector<float> flat((size_t)n * Robot::dimension);
    mrpop_solution_path(sol, which, flat.data(), n, &n, &cost);
    out.resize(n);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < Robot::dimension; ++d) out[i][d] = flat[(size_t)i * Robot::dimension + d];
    return out;
}

}  // namespace mrpop_cpp

namespace dRRT {

// num_robots must equal Robot::dimension / 7.  block_per_robot sets the
// roadmap samples per arm and round (230 * block_per_robot, as the
// reference's ROADMAP_BUILD_ITER blocks).  Arms use the default base poses
// of mrpop_default_bases (the reference's tables for 4 and 5 arms).
template <typename Robot>
PlannerResult<Robot> solve(typename Robot::Configuration &start, std::vector<typename Robot::Configuration> &goals,
                           ppln::collision::Environment<float> &environment, pop_settings &settings,
                           const int num_robots, const int block_per_robot) {
    PlannerResult<Robot> result;
    auto fail = [&](const char *what, int rc, mrpop_ctx *ctx) {
        std::fprintf(stderr, "dRRT::solve: %s failed (%d): %s\n", what, rc, ctx ? mrpop_last_error(ctx) : "");
        return result;
    };
    if (num_robots * 7 != Robot::dimension) return fail("argument check (num_robots)", MRPOP_ERR_INVALID, nullptr);
    if (goals.empty()) return fail("argument check (no goals)", MRPOP_ERR_INVALID, nullptr);
    auto &sc = mrpop_cpp::shared_context();
    if (sc.status != MRPOP_OK) return fail("mrpop_create", sc.status, nullptr);

    mrpop_cpp::FlatScene scene(environment);
    std::vector<float> flat_goals;
    for (const auto &g : goals) flat_goals.insert(flat_goals.end(), g.begin(), g.end());
    mrpop_problem problem{};
    problem.n_arms = num_robots;
    problem.bases = nullptr;
    problem.scene = &scene.c;
    problem.start = start.data();
    problem.n_goals = (int)goals.size();
    problem.goals = flat_goals.data();
    problem.seed = mrpop_cpp::kSolveSeed;
    const mrpop_settings cs = mrpop_cpp::to_c(settings);
    mrpop_solve_options opt{};
    mrpop_solve_options_default(&opt);
    opt.block_per_robot = block_per_robot > 0 ? block_per_robot : opt.block_per_robot;
    opt.time_budget_s = 0;  // the reference has no time limit

    std::lock_guard<std::mutex> guard(sc.lock);
    mrpop_solution *sol = nullptr;
    const int rc = mrpop_solve(sc.ctx, &problem, &cs, &opt, &sol);
    if (rc != MRPOP_OK) return fail("mrpop_solve", rc, sc.ctx);
    mrpop_result_info info{};
    mrpop_solution_info(sol, &info);
    if (info.status != MRPOP_OK && info.status != MRPOP_ERR_TIMEOUT) {  // settings or values rejected
        mrpop_solution_free(sol);
        return fail("mrpop_solve", info.status, sc.ctx);
    }
    result.solved = info.solved != 0;
    result.cost = info.cost;
    result.start_tree_size = info.start_tree_size;
    result.goal_tree_size = info.goal_tree_size;
    result.iters = info.iters;
    result.wall_ns = info.wall_ns;
    result.kernel_ns = info.kernel_ns;
    result.copy_ns = info.copy_ns;
    result.ttfs_ns = info.ttfs_ns;
    result.path = mrpop_cpp::to_configs<Robot>(sol, -1);
    result.path_length = (int)result.path.size();
    for (int k = 0; k < info.n_seed_paths; ++k) {
        int n = 0;
        float cost = 0;
        mrpop_solution_path(sol, k, nullptr, 0, &n, &cost);
        result.seed_paths.emplace_back(cost, mrpop_cpp::to_configs<Robot>(sol, k));
    }
    int n_imp = 0;
    mrpop_solution_improvements(sol, nullptr, 0, &n_imp);
    std::vector<uint64_t> imp((size_t)n_imp);
    if (n_imp) mrpop_solution_improvements(sol, imp.data(), n_imp, &n_imp);
    result.improve_ns.assign(imp.begin(), imp.end());
    mrpop_solution_free(sol);
    return result;
}

}  // namespace dRRT
