// Robots.hh -- robot types of the planner, with the names, dimensions and
// configuration types of the reference (src/planning/Robots.hh): Panda (one
// arm, 7 DoF) and the multi-arm teams Panda_four (28) and Panda_five (35),
// plus Panda_two (14) and Panda_three (21), which the reference does not
// define.
//
// A configuration is stored in joint units.  scale_cfg maps unit-cube
// coordinates u in (0, 1] to joint values q = u * s_m + s_a (joint limits of
// the Franka Panda, Robots.hh:13-23 of the reference); descale_cfg inverts
// it.  Both are plain host/device functions over all coordinates instead of
// the reference's one-coordinate-per-thread device helpers.
#pragma once

#include <array>
#include <cstddef>

#if defined(__HIPCC__)
#define MRPOP_HD __host__ __device__
#else
#define MRPOP_HD
#endif

namespace ppln::robots {

namespace detail {
// range (upper - lower) and lower limit of the seven Panda joints
inline constexpr float kPandaRange[7] = {5.9342f, 3.6652f, 5.9342f, 3.2289f, 5.9342f, 3.9096f, 5.9342f};
inline constexpr float kPandaLower[7] = {-2.9671f, -1.8326f, -2.9671f, -3.1416f, -2.9671f, -0.0873f, -2.9671f};

template <int Arms>
struct PandaTeam {
    static constexpr int arms = Arms;
    static constexpr int dimension = 7 * Arms;
    using Configuration = std::array<float, 7 * Arms>;

    MRPOP_HD static constexpr float get_s_m(int i) { return kPandaRange[i % 7]; }
    MRPOP_HD static constexpr float get_s_a(int i) { return kPandaLower[i % 7]; }

    MRPOP_HD static void scale_cfg(float *q) {
        for (int i = 0; i < dimension; ++i) q[i] = q[i] * get_s_m(i) + get_s_a(i);
    }
    MRPOP_HD static void descale_cfg(float *q) {
        for (int i = 0; i < dimension; ++i) q[i] = (q[i] - get_s_a(i)) / get_s_m(i);
    }
};
}  // namespace detail

struct Panda : detail::PandaTeam<1> {
    static constexpr auto name = "panda";
};

struct Panda_two : detail::PandaTeam<2> {
    static constexpr auto name = "panda_two";
    using SingleRobot = Panda;
};

struct Panda_three : detail::PandaTeam<3> {
    static constexpr auto name = "panda_three";
    using SingleRobot = Panda;
};

struct Panda_four : detail::PandaTeam<4> {
    static constexpr auto name = "panda_four";
    using SingleRobot = Panda;
};

struct Panda_five : detail::PandaTeam<5> {
    static constexpr auto name = "panda_five";
    using SingleRobot = Panda;
};

}  // namespace ppln::robots
