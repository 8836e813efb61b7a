// solve_problem.cpp -- plans one problem through the drop-in C++ entry point
// dRRT::solve<Robot> (src/planning/Planners.hh), the way the reference's
// harness calls it (scripts/evaluate_mr.cu:134-136).
//
//   solve_problem PROBLEM.txt     plan and print the result
//   solve_problem --selftest      host-side checks of the drop-in headers
//
// PROBLEM.txt, whitespace separated:
//   arms A
//   settings num_new_configs granularity range rrtc_iter optimize_iters
//   spheres N     then N records of 4 floats   (x y z r)
//   capsules N    then N records of 8 floats   (x1 y1 z1 xv yv zv r rdv)
//   zcapsules N   then N records of 8 floats
//   cuboids N     then N records of 15 floats  (centre, 3 axes, half extents)
//   zcuboids N    then N records of 15 floats
//   start         7A floats
//   goals M       then M configurations of 7A floats
// Output: "solved S", "cost C", "path N" and N lines of 7A floats
// (goal -> start), then "seeds K" and one "seed <cost>" line per seed path.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "src/collision/factory.hh"
#include "src/planning/Planners.hh"

using namespace ppln::collision;

namespace {

struct Input {
    int arms = 0;
    pop_settings st;
    std::vector<std::vector<float>> rec[5];
    std::vector<float> start;
    std::vector<std::vector<float>> goals;
};

bool read_input(const char *path, Input &in) {
    std::ifstream f(path);
    if (!f) return false;
    std::string key;
    const char *names[5] = {"spheres", "capsules", "zcapsules", "cuboids", "zcuboids"};
    const int width[5] = {4, 8, 8, 15, 15};
    while (f >> key) {
        if (key == "arms") {
            f >> in.arms;
        } else if (key == "settings") {
            f >> in.st.num_new_configs >> in.st.granularity >> in.st.range >> in.st.rrtc_iter >> in.st.optimize_iters;
        } else if (key == "start") {
            in.start.resize(7 * in.arms);
            for (float &v : in.start) f >> v;
        } else if (key == "goals") {
            int m = 0;
            f >> m;
            in.goals.assign(m, std::vector<float>(7 * in.arms));
            for (auto &g : in.goals)
                for (float &v : g) f >> v;
        } else {
            int k = 0;
            while (k < 5 && key != names[k]) ++k;
            if (k == 5) return false;
            int n = 0;
            f >> n;
            in.rec[k].assign(n, std::vector<float>(width[k]));
            for (auto &r : in.rec[k])
                for (float &v : r) f >> v;
        }
    }
    return in.arms >= 2 && in.arms <= 5 && !in.goals.empty() && (bool)f.eof();
}

// An Environment owning its arrays, as the reference's harness builds it.
void fill_environment(const Input &in, Environment<float> &env) {
    env.num_spheres = in.rec[0].size();
    env.spheres = new Sphere<float>[env.num_spheres];
    for (unsigned i = 0; i < env.num_spheres; ++i) {
        const auto &r = in.rec[0][i];
        env.spheres[i] = factory::sphere::flat(r[0], r[1], r[2], r[3]);
    }
    auto caps = [](const std::vector<std::vector<float>> &recs, Capsule<float> *&out, unsigned &n) {
        n = recs.size();
        out = new Capsule<float>[n];
        for (unsigned i = 0; i < n; ++i) {
            const auto &r = recs[i];
            out[i] = Capsule<float>(r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7]);
        }
    };
    caps(in.rec[1], env.capsules, env.num_capsules);
    caps(in.rec[2], env.z_aligned_capsules, env.num_z_aligned_capsules);
    auto cubs = [](const std::vector<std::vector<float>> &recs, Cuboid<float> *&out, unsigned &n) {
        n = recs.size();
        out = new Cuboid<float>[n];
        for (unsigned i = 0; i < n; ++i) {
            const auto &r = recs[i];
            out[i] = Cuboid<float>(r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10], r[11], r[12],
                                   r[13], r[14]);
        }
    };
    cubs(in.rec[3], env.cuboids, env.num_cuboids);
    cubs(in.rec[4], env.z_aligned_cuboids, env.num_z_aligned_cuboids);
}

template <typename Robot>
int plan(const Input &in) {
    Environment<float> env;
    fill_environment(in, env);
    typename Robot::Configuration start;
    std::copy(in.start.begin(), in.start.end(), start.begin());
    std::vector<typename Robot::Configuration> goals(in.goals.size());
    for (size_t g = 0; g < goals.size(); ++g) std::copy(in.goals[g].begin(), in.goals[g].end(), goals[g].begin());
    pop_settings st = in.st;
    auto res = dRRT::solve<Robot>(start, goals, env, st, Robot::dimension / 7, 8);
    std::printf("solved %d\ncost %.9g\npath %d\n", res.solved ? 1 : 0, res.cost, (int)res.path.size());
    for (auto &c : res.path) {
        for (int d = 0; d < Robot::dimension; ++d) std::printf("%.9g ", c[d]);
        std::printf("\n");
    }
    std::printf("seeds %d\n", (int)res.seed_paths.size());
    for (auto &s : res.seed_paths) std::printf("seed %.9g %d\n", s.first, (int)s.second.size());
    std::printf("ttfs_ns %zu\nwall_ns %zu\n", res.ttfs_ns, res.wall_ns);
    return res.solved ? 0 : 3;
}

int fails = 0;
void expect(bool ok, const char *what) {
    if (!ok) {
        std::printf("FAIL %s\n", what);
        ++fails;
    }
}

int selftest() {
    // Euler XYZ: a quarter turn about z maps the cuboid's x axis to world y
    auto c = factory::cuboid::flat(0.1f, 0.2f, 0.3f, 0.f, 0.f, (float)M_PI / 2, 0.1f, 0.2f, 0.3f);
    expect(std::fabs(c.axis_1_x) < 1e-6f && std::fabs(c.axis_1_y - 1.f) < 1e-6f, "cuboid axis 1 after Rz(pi/2)");
    expect(std::fabs(c.axis_2_x + 1.f) < 1e-6f && std::fabs(c.axis_2_y) < 1e-6f, "cuboid axis 2 after Rz(pi/2)");
    // a general rotation has orthonormal, right-handed axes
    auto r = factory::cuboid::array({0, 0, 0}, {0.3f, -1.1f, 2.0f}, {1, 1, 1});
    const float a1[3] = {r.axis_1_x, r.axis_1_y, r.axis_1_z}, a2[3] = {r.axis_2_x, r.axis_2_y, r.axis_2_z},
                a3[3] = {r.axis_3_x, r.axis_3_y, r.axis_3_z};
    expect(std::fabs(dot_3(a1[0], a1[1], a1[2], a2[0], a2[1], a2[2])) < 1e-6f, "axes 1, 2 orthogonal");
    const float cr[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2], a1[0] * a2[1] - a1[1] * a2[0]};
    expect(std::fabs(cr[0] - a3[0]) + std::fabs(cr[1] - a3[1]) + std::fabs(cr[2] - a3[2]) < 1e-5f, "axis 3 = 1 x 2");
    // Rx then Ry then Rz: x axis of Rz(phi) Ry(theta) Rx(rho) is (cz cy, sz cy, -sy)
    expect(std::fabs(r.axis_1_z - std::sin(1.1f)) < 1e-6f, "axis 1 z component is -sin(theta)");
    // segment records
    auto cap = factory::capsule::endpoints::flat(0, 0, 0, 0, 0, 2, 0.1f);
    expect(cap.zv == 2.f && std::fabs(cap.rdv - 0.25f) < 1e-7f, "capsule v and 1/|v|^2");
    auto cyl = factory::cylinder::center::flat(1, 2, 3, 0, 0, 0, 0.1f, 0.4f);
    expect(std::fabs(cyl.z1 - 3.2f) < 1e-6f && std::fabs(cyl.zv + 0.4f) < 1e-6f, "centred cylinder end points");
    // distances from the origin
    expect(std::fabs(factory::sphere::flat(3, 4, 0, 1).min_distance - 4.f) < 1e-6f, "sphere min_distance");
    expect(std::fabs(factory::cuboid::flat(2, 0, 0, 0, 0, 0, 0.5f, 1, 1).min_distance - 1.5f) < 1e-6f,
           "cuboid min_distance");
    expect(factory::cuboid::flat(0, 0, 0, 0, 0, 0, 0.5f, 1, 1).min_distance == 0.f, "origin inside a cuboid");
    // joint scaling round trip and limits
    using R4 = ppln::robots::Panda_four;
    R4::Configuration q{};
    for (int i = 0; i < R4::dimension; ++i) q[i] = (i + 1) / 29.f;
    auto u = q;
    R4::scale_cfg(q.data());
    expect(std::fabs(q[3] - (4 / 29.f * 3.2289f - 3.1416f)) < 1e-6f, "scale_cfg joint 4");
    R4::descale_cfg(q.data());
    expect(std::fabs(q[20] - u[20]) < 1e-6f, "descale_cfg inverts scale_cfg");
    static_assert(R4::dimension == 28 && ppln::robots::Panda_five::dimension == 35, "dimensions");
    static_assert(ppln::robots::Panda_two::dimension == 14 && ppln::robots::Panda_three::dimension == 21, "teams");
    // settings defaults agree with the C ABI's
    pop_settings ps;
    mrpop_settings cs{};
    mrpop_settings_default(&cs);
    expect(ps.num_new_configs == cs.num_new_configs && ps.granularity == cs.granularity && ps.range == cs.range &&
               ps.optimize_iters == cs.optimize_iters && ps.rrtc_iter == cs.rrtc_iter &&
               ps.bspline_step == cs.bspline_step && ps.max_iters == cs.max_iters,
           "pop_settings defaults");
    // flattening keeps the record layout of the C ABI
    Environment<float> env;
    env.num_cuboids = 1;
    env.cuboids = new Cuboid<float>[1]{c};
    env.num_capsules = 1;
    env.capsules = new Capsule<float>[1]{cap};
    env.num_cylinders = 1;
    env.cylinders = new Cylinder<float>[1]{cyl};
    mrpop_cpp::FlatScene fs(env);
    expect(fs.c.n_cuboids == 1 && fs.cuboids[0] == 0.1f && fs.cuboids[4] == c.axis_1_y && fs.cuboids[14] == 0.3f,
           "flat cuboid record");
    expect(fs.c.n_capsules == 1 && fs.capsules[5] == 2.f && fs.capsules[7] == cap.rdv, "flat capsule record");
    expect(fs.c.n_spheres == 0 && fs.c.n_z_cuboids == 0, "cylinders are not passed on");
    std::printf(fails ? "selftest: %d failures\n" : "selftest: ok\n", fails);
    return fails ? 1 : 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc == 2 && std::strcmp(argv[1], "--selftest") == 0) return selftest();
    Input in;
    if (argc != 2 || !read_input(argv[1], in)) {
        std::fprintf(stderr, "usage: solve_problem PROBLEM.txt | --selftest\n");
        return 2;
    }
    switch (in.arms) {
        case 2: return plan<ppln::robots::Panda_two>(in);
        case 3: return plan<ppln::robots::Panda_three>(in);
        case 4: return plan<ppln::robots::Panda_four>(in);
        default: return plan<ppln::robots::Panda_five>(in);
    }
}
