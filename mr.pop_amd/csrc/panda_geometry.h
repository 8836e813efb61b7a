// Panda link-sphere model, fixed joint transforms and multi-arm base poses.
//
// Numeric data only.  The values are the reference robot model:
//   sphere tables      src/robots/panda.cuh:97-595 (= panda_multi_geometry.cuh:17-516)
//   joint limits       src/planning/Robots.hh:13-23
//   4-arm / 5-arm base src/robots/panda.cuh:16-73, panda_four.cuh:17-41, panda_five.cuh:17-47
// They are kept identical because collision verdicts must match the reference
// bit for bit.  The layout is this library's own: every joint of the Panda
// rotates about its local z axis, so a joint is described by one rigid 3x4
// transform (row-major [R | t]) that precedes the joint rotation Rz(q).
// Joint 0 is the arm base (no rotation); joints 1..7 take q[0..6].
// The 2- and 3-arm base poses have no counterpart in the reference; they are
// this library's layouts for 2- and 3-arm scenes (see kBase2 / kBase3).
#pragma once

namespace mrpop {
namespace geom {

inline constexpr int kJoints = 8;          // base frame + 7 revolute joints
inline constexpr int kDofPerArm = 7;
inline constexpr int kCoarseSpheres = 11;
inline constexpr int kFineSpheres = 59;
inline constexpr int kCoarseSelfRanges = 5;
inline constexpr int kFineSelfRanges = 25;
inline constexpr int kMaxArms = 5;

// Rigid transform preceding joint j (row-major 3x4); entry 0 is the identity
// (joint 0 is the base pose, supplied per arm).
inline constexpr float kJointTf[kJoints][12] = {
    {1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f},
    {1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.333f},
    {1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, -1.0f, 0.0f, 0.0f},
    {1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, -1.0f, -0.316f, 0.0f, 1.0f, 0.0f, 0.0f},
    {1.0f, 0.0f, 0.0f, 0.0825f, 0.0f, 0.0f, -1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f},
    {1.0f, 0.0f, 0.0f, -0.0825f, 0.0f, 0.0f, 1.0f, 0.384f, 0.0f, -1.0f, 0.0f, 0.0f},
    {1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, -1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f},
    {1.0f, 0.0f, 0.0f, 0.088f, 0.0f, 0.0f, -1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f},
};

// coarse spheres: centre (x, y, z) in the frame of joint kCoarseJoint[i], radius r.
inline constexpr float kCoarseSphere[11][4] = {
    {0.0f, 0.0f, 0.05f, 0.08f},
    {-0.001f, -0.039f, -0.085f, 0.154f},
    {0.0f, -0.085f, 0.04f, 0.154f},
    {0.039f, 0.028f, -0.052f, 0.128f},
    {-0.042f, 0.049f, 0.029f, 0.126f},
    {-0.001f, 0.037f, -0.11f, 0.176f},
    {0.042f, 0.014f, 0.0f, 0.095f},
    {0.015f, 0.015f, 0.075f, 0.072f},
    {0.0f, 0.0f, 0.129f, 0.104f},
    {0.054447f, 0.054447f, 0.1984f, 0.024f},
    {-0.054447f, -0.054447f, 0.1984f, 0.024f},
};
inline constexpr int kCoarseJoint[11] = {0, 1, 2, 3, 4, 5, 6, 7, 7, 7, 7};
inline constexpr int kCoarseJointCount[kJoints] = {1, 1, 1, 1, 1, 1, 1, 4};
inline constexpr int kCoarseJointStart[kJoints] = {0, 1, 2, 3, 4, 5, 6, 7};

// fine spheres: centre (x, y, z) in the frame of joint kFineJoint[i], radius r.
inline constexpr float kFineSphere[59][4] = {
    {0.0f, 0.0f, 0.05f, 0.08f},
    {0.0f, -0.08f, 0.0f, 0.06f},
    {0.0f, -0.03f, 0.0f, 0.06f},
    {0.0f, 0.0f, -0.12f, 0.06f},
    {0.0f, 0.0f, -0.17f, 0.06f},
    {0.0f, 0.0f, 0.03f, 0.06f},
    {0.0f, 0.0f, 0.08f, 0.06f},
    {0.0f, -0.12f, 0.0f, 0.06f},
    {0.0f, -0.17f, 0.0f, 0.06f},
    {0.0f, 0.0f, -0.1f, 0.06f},
    {0.0f, 0.0f, -0.06f, 0.05f},
    {0.08f, 0.06f, 0.0f, 0.055f},
    {0.08f, 0.02f, 0.0f, 0.055f},
    {-0.08f, 0.095f, 0.0f, 0.06f},
    {0.0f, 0.0f, 0.02f, 0.055f},
    {0.0f, 0.0f, 0.06f, 0.055f},
    {-0.08f, 0.06f, 0.0f, 0.055f},
    {0.0f, 0.055f, 0.0f, 0.06f},
    {0.0f, 0.075f, 0.0f, 0.06f},
    {0.0f, 0.0f, -0.22f, 0.06f},
    {0.0f, 0.05f, -0.18f, 0.05f},
    {0.01f, 0.08f, -0.14f, 0.025f},
    {0.01f, 0.085f, -0.11f, 0.025f},
    {0.01f, 0.09f, -0.08f, 0.025f},
    {0.01f, 0.095f, -0.05f, 0.025f},
    {-0.01f, 0.08f, -0.14f, 0.025f},
    {-0.01f, 0.085f, -0.11f, 0.025f},
    {-0.01f, 0.09f, -0.08f, 0.025f},
    {-0.01f, 0.095f, -0.05f, 0.025f},
    {0.0f, 0.0f, 0.0f, 0.05f},
    {0.08f, -0.01f, 0.0f, 0.05f},
    {0.08f, 0.035f, 0.0f, 0.052f},
    {0.0f, 0.0f, 0.07f, 0.05f},
    {0.02f, 0.04f, 0.08f, 0.025f},
    {0.04f, 0.02f, 0.08f, 0.025f},
    {0.04f, 0.06f, 0.085f, 0.02f},
    {0.06f, 0.04f, 0.085f, 0.02f},
    {-0.053033f, -0.053033f, 0.117f, 0.028f},
    {-0.03182f, -0.03182f, 0.117f, 0.028f},
    {-0.010607f, -0.010607f, 0.117f, 0.028f},
    {0.010607f, 0.010607f, 0.117f, 0.028f},
    {0.03182f, 0.03182f, 0.117f, 0.028f},
    {0.053033f, 0.053033f, 0.117f, 0.028f},
    {-0.053033f, -0.053033f, 0.137f, 0.026f},
    {-0.03182f, -0.03182f, 0.137f, 0.026f},
    {-0.010607f, -0.010607f, 0.137f, 0.026f},
    {0.010607f, 0.010607f, 0.137f, 0.026f},
    {0.03182f, 0.03182f, 0.137f, 0.026f},
    {0.053033f, 0.053033f, 0.137f, 0.026f},
    {-0.053033f, -0.053033f, 0.157f, 0.024f},
    {-0.03182f, -0.03182f, 0.157f, 0.024f},
    {-0.010607f, -0.010607f, 0.157f, 0.024f},
    {0.010607f, 0.010607f, 0.157f, 0.024f},
    {0.03182f, 0.03182f, 0.157f, 0.024f},
    {0.053033f, 0.053033f, 0.157f, 0.024f},
    {0.056569f, 0.056569f, 0.1874f, 0.012f},
    {0.051619f, 0.051619f, 0.2094f, 0.012f},
    {-0.056569f, -0.056569f, 0.1874f, 0.012f},
    {-0.051619f, -0.051619f, 0.2094f, 0.012f},
};
inline constexpr int kFineJoint[59] = {0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 6, 6, 6, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7};
inline constexpr int kFineJointCount[kJoints] = {1, 4, 4, 4, 4, 12, 3, 27};
inline constexpr int kFineJointStart[kJoints] = {0, 1, 5, 9, 13, 17, 29, 32};

// Intra-arm self-collision pairs: sphere r[0] against spheres r[1]..r[2] (inclusive).
inline constexpr int kCoarseSelfRange[5][3] = {
    {0, 5, 10}, {1, 5, 10}, {2, 5, 5}, {2, 7, 10}, {5, 7, 10}
};

// Intra-arm self-collision pairs: sphere r[0] against spheres r[1]..r[2] (inclusive).
inline constexpr int kFineSelfRange[25][3] = {
    {0, 17, 58}, {1, 17, 58}, {2, 17, 58}, {3, 17, 58}, {4, 17, 58}, {5, 17, 28}, {5, 32, 58}, {6, 17, 28}, {6, 32, 58}, {7, 17, 28}, {7, 32, 58}, {8, 17, 28}, {8, 32, 58}, {17, 32, 58}, {18, 32, 58}, {19, 32, 58}, {20, 32, 58}, {21, 32, 58}, {22, 32, 58}, {23, 32, 58}, {24, 32, 58}, {25, 32, 58}, {26, 32, 58}, {27, 32, 58}, {28, 32, 58}
};

// Joint limits: q = u * kScale[i] + kLow[i] for u in (0, 1].
inline constexpr float kScale[kDofPerArm] = {5.9342f, 3.6652f, 5.9342f, 3.2289f, 5.9342f, 3.9095999999999997f, 5.9342f};
inline constexpr float kLow[kDofPerArm] = {-2.9671f, -1.8326f, -2.9671f, -3.1416f, -2.9671f, -0.0873f, -2.9671f};

// Derived bounding balls (not in the reference; used only to skip tests that
// cannot hit).  Coarse sphere k, grown by kBlockInflate[k] metres, contains
// fine spheres kBlockFirst[k] .. kBlockFirst[k] + kBlockCount[k] - 1 (all on
// the same joint): |p_fine - p_coarse| + r_fine <= r_coarse + inflate.
// The inflation includes a 1e-4 m margin for float rounding.
inline constexpr int kBlockFirst[kCoarseSpheres] = {0, 1, 5, 9, 13, 17, 29, 32, 37, 55, 57};
inline constexpr int kBlockCount[kCoarseSpheres] = {1, 4, 4, 4, 4, 12, 3, 5, 18, 2, 2};
inline constexpr float kBlockInflate[kCoarseSpheres] = {0.000100f, 0.000477f, 0.000100f, 0.000646f, 0.000696f, 0.000483f, 0.000517f, 0.000540f, 0.000156f, 0.000100f, 0.000100f};

// World <- base poses (row-major 3x4) per arm count.
// 2 arms: bases at (-0.4, 0, 0) and (0.4, 0, 0), facing each other (yaw 0 and pi).
inline constexpr float kBase2[2][12] = {
    {1.0f, -0.0f, 0.0f, -0.4f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f},
    {-1.0f, -1.2246467991473532e-16f, 0.0f, 0.4f, 1.2246467991473532e-16f, -1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f},
};
// 3 arms: bases on a 0.6 m circle at 90, 210 and 330 degrees, each facing the centre.
inline constexpr float kBase3[3][12] = {
    {-1.8369701987210297e-16f, 1.0f, 0.0f, 3.6739403974420595e-17f, -1.0f, -1.8369701987210297e-16f, 0.0f, 0.6f, 0.0f, 0.0f, 1.0f, 0.0f},
    {0.866025403784439f, -0.4999999999999993f, 0.0f, -0.5196152422706632f, 0.4999999999999993f, 0.866025403784439f, 0.0f, -0.2999999999999998f, 0.0f, 0.0f, 1.0f, 0.0f},
    {-0.8660254037844388f, -0.4999999999999998f, 0.0f, 0.519615242270663f, 0.4999999999999998f, -0.8660254037844388f, 0.0f, -0.30000000000000027f, 0.0f, 0.0f, 1.0f, 0.0f},
};
// 4 arms: the reference bin-picking layout (corners (+-0.4, +-0.4, 0.15)).
inline constexpr float kBase4[4][12] = {
    {-0.65122966f, -0.75888071f, 0.0f, 0.4f, 0.75888071f, -0.65122966f, 0.0f, -0.4f, 0.0f, 0.0f, 1.0f, 0.15f},
    {0.78999223f, -0.61311685f, 0.0f, -0.4f, 0.61311685f, 0.78999223f, 0.0f, -0.4f, 0.0f, 0.0f, 1.0f, 0.15f},
    {0.78999223f, 0.61311685f, 0.0f, -0.4f, -0.61311685f, 0.78999223f, 0.0f, 0.4f, 0.0f, 0.0f, 1.0f, 0.15f},
    {-0.65122966f, 0.75888071f, 0.0f, 0.4f, -0.75888071f, -0.65122966f, 0.0f, 0.4f, 0.0f, 0.0f, 1.0f, 0.15f},
};
// 5 arms: the reference shelf layout.
inline constexpr float kBase5[5][12] = {
    {0.00079633f, -0.99999968f, 0.0f, 0.6f, 0.99999968f, 0.00079633f, 0.0f, -0.7f, 0.0f, 0.0f, 1.0f, 0.1f},
    {0.00079633f, -0.99999968f, 0.0f, 0.0f, 0.99999968f, 0.00079633f, 0.0f, -0.7f, 0.0f, 0.0f, 1.0f, 0.1f},
    {0.00079633f, 0.99999968f, 0.0f, 0.6f, -0.99999968f, 0.00079633f, 0.0f, 0.7f, 0.0f, 0.0f, 1.0f, 0.1f},
    {0.00079633f, 0.99999968f, 0.0f, 0.0f, -0.99999968f, 0.00079633f, 0.0f, 0.7f, 0.0f, 0.0f, 1.0f, 0.1f},
    {1.0f, 0.0f, 0.0f, -0.6f, 0.0f, 1.0f, 0.0f, -0.2f, 0.0f, 0.0f, 1.0f, 0.1f},
};

// Default base pose table for an arm count (nullptr outside 2..5).
inline constexpr const float (*default_bases(int arms))[12] {
    return arms == 2 ? kBase2 : arms == 3 ? kBase3 : arms == 4 ? kBase4 : arms == 5 ? kBase5 : nullptr;
}

}  // namespace geom
}  // namespace mrpop
