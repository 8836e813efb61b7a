// device_geom.h -- forward kinematics and sphere/obstacle tests for gfx950.
//
// Forward kinematics (reference: panda_four.cuh:44-131, 289-373; utils.cuh:400-445):
//   T_0 = base, T_j = T_{j-1} * FT_j * Rz(q_{j-1}), sphere centre = T_joint * p.
// Written per joint with the fixed transform FT_j known at compile time, so
// the products with its 0 / +-1 entries fold away; frames are rigid 3x4.
//
// Primitive tests follow utils.cuh:100-260: each returns the reference's
// value (squared clearance form); collision iff value < 0.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "panda_geometry.h"

namespace mrpop {
namespace dev {

using namespace mrpop::geom;

struct Frame {
    float m[12];  // row-major [R | t]
};

// k * x for a compile-time constant k: exact for k in {0, 1, -1}.  Callers
// pass k as a constexpr value (a literal in the IR), so the tests fold away;
// a value read from a table in memory would turn them into selects.
__device__ __forceinline__ constexpr float cmul(float k, float x) {
    return k == 0.f ? 0.f : (k == 1.f ? x : (k == -1.f ? -x : k * x));
}
__device__ __forceinline__ constexpr float cfma(float k, float x, float acc) {
    return k == 0.f ? acc : (k == 1.f ? acc + x : (k == -1.f ? acc - x : __builtin_fmaf(k, x, acc)));
}

template <int V> using IntC = std::integral_constant<int, V>;

template <int B, int E, class F> __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(IntC<B>{});
        static_for<B + 1, E>(f);
    }
}

// out = P * FT_J * Rz(q) with cq = cos(q), sq = sin(q).
template <int J>
__device__ __forceinline__ void joint_step(const Frame &P, float cq, float sq, Frame &out) {
    // S = FT * Rz: S[r][0] = F[r][0] c + F[r][1] s; S[r][1] = F[r][1] c - F[r][0] s;
    //              S[r][2] = F[r][2]; S[r][3] = F[r][3].
    float S[6];
    static_for<0, 3>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        constexpr float f0 = kJointTf[J][4 * r + 0], f1 = kJointTf[J][4 * r + 1];
        S[2 * r + 0] = cfma(f1, sq, cmul(f0, cq));
        S[2 * r + 1] = cfma(-f0, sq, cmul(f1, cq));
    });
    static_for<0, 3>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            float v = P.m[4 * r + 0] * S[k];
            v = __builtin_fmaf(P.m[4 * r + 1], S[2 + k], v);
            v = __builtin_fmaf(P.m[4 * r + 2], S[4 + k], v);
            out.m[4 * r + k] = v;
        }
        // column 2: P * F[:,2]; translation: P * F[:,3] + P_t (constants)
        constexpr float f2 = kJointTf[J][2], f6 = kJointTf[J][6], f10 = kJointTf[J][10];
        constexpr float f3 = kJointTf[J][3], f7 = kJointTf[J][7], f11 = kJointTf[J][11];
        out.m[4 * r + 2] = cfma(f10, P.m[4 * r + 2], cfma(f6, P.m[4 * r + 1], cmul(f2, P.m[4 * r + 0])));
        out.m[4 * r + 3] = cfma(f11, P.m[4 * r + 2], cfma(f7, P.m[4 * r + 1], cfma(f3, P.m[4 * r + 0], P.m[4 * r + 3])));
    });
}

// Sine and cosine of a joint angle with the hardware instructions
// (v_sin_f32 / v_cos_f32), as the reference uses the fast __sinf / __cosf
// (utils.cuh:334-335).  Over the joint range their error moves a sphere
// centre by far less than the 1e-5 m tolerance of the checker, and they
// avoid the large-argument branch of the full-range library sincos.
__device__ __forceinline__ void joint_sincos(float q, float &s, float &c) {
    s = __sinf(q);
    c = __cosf(q);
}

// T_J from T_{J-1} and the joint angle q_{J-1}.
template <int J>
__device__ __forceinline__ void joint_frame(const Frame &P, float q, Frame &out) {
    float s, c;
    joint_sincos(q, s, c);
    joint_step<J>(P, c, s, out);
}

// Centre of a sphere with compile-time offset p in frame T.
__device__ __forceinline__ void centre(const Frame &T, float px, float py, float pz, float &x, float &y, float &z) {
    x = __builtin_fmaf(T.m[0], px, __builtin_fmaf(T.m[1], py, __builtin_fmaf(T.m[2], pz, T.m[3])));
    y = __builtin_fmaf(T.m[4], px, __builtin_fmaf(T.m[5], py, __builtin_fmaf(T.m[6], pz, T.m[7])));
    z = __builtin_fmaf(T.m[8], px, __builtin_fmaf(T.m[9], py, __builtin_fmaf(T.m[10], pz, T.m[11])));
}

// ---------------------------------------------------------- obstacles ----
// Device obstacle table: one float buffer, sections in this order.
// Records: sphere 4 (x y z r); capsule 8 (x1 y1 z1 xv yv zv r rdv);
// aabb 6 (x y z hx hy hz) -- cuboids whose axes are exactly the world axes;
// obb 15 (reference cuboid record); zcap 8; zcub 15.
// The arm base poses (n_arms x 12, row-major 3x4) follow at o_base.
struct WorldDesc {
    const float *obs;
    int n_sph, n_cap, n_zcap, n_aabb, n_obb, n_zcub;
    int o_sph, o_cap, o_zcap, o_aabb, o_obb, o_zcub, o_base;  // float offsets into obs
    int n_arms;
    // cull boxes: [lo xyz, hi xyz, f, 0] per obstacle, in the order
    // spheres, capsules, z-capsules, aabbs, obbs, z-cuboids.  Every centre
    // of a sphere of radius r that the obstacle's test reports as colliding
    // lies inside the box grown by r * f on every side (pack_world builds
    // them).
    int o_box;
};

// The world buffer is read-only while a check runs.  Read through the
// constant address space, its wave-uniform reads (obstacle records, cull
// boxes) become scalar loads through the scalar cache; through a generic
// pointer, in a kernel that also writes memory, they are vector loads whose
// latency every obstacle iteration waits for.
using cfloat = const __attribute__((address_space(4))) float;
__device__ __forceinline__ cfloat *world_data(const WorldDesc &w) { return (cfloat *)w.obs; }

// Each test below is q - (a + r)^2 < 0 for a radius-independent term q of
// the obstacle and the sphere centre (squared distance to the obstacle's
// core) and the obstacle's own radius a (0 for cuboids).  The *_q functions
// give q alone, so that one evaluation serves several radii.

// utils.cuh:100-107
__device__ __forceinline__ float sphere_sphere_q(float ax, float ay, float az, float bx, float by, float bz) {
    float dx = ax - bx, dy = ay - by, dz = az - bz;
    return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
}
__device__ __forceinline__ float sphere_sphere(float ax, float ay, float az, float ar, float bx, float by, float bz,
                                               float br) {
    float rs = ar + br;
    return sphere_sphere_q(ax, ay, az, bx, by, bz) - rs * rs;
}

// utils.cuh:148-158
template <class P> __device__ __forceinline__ float capsule_q(P c, float x, float y, float z) {
    float dot = (x - c[0]) * c[3] + (y - c[1]) * c[4] + (z - c[2]) * c[5];
    float t = fmaxf(fminf(dot * c[7], 1.f), 0.f);
    float dx = x - __builtin_fmaf(c[3], t, c[0]);
    float dy = y - __builtin_fmaf(c[4], t, c[1]);
    float dz = z - __builtin_fmaf(c[5], t, c[2]);
    return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
}
template <class P> __device__ __forceinline__ float sphere_capsule(P c, float x, float y, float z, float r) {
    float rs = r + c[6];
    return capsule_q(c, x, y, z) - rs * rs;
}

// utils.cuh:165-174
template <class P> __device__ __forceinline__ float zcapsule_q(P c, float x, float y, float z) {
    float dot = (z - c[2]) * c[5];
    float t = fmaxf(fminf(dot * c[7], 1.f), 0.f);
    float dx = x - c[0], dy = y - c[1], dz = z - __builtin_fmaf(c[5], t, c[2]);
    return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
}
template <class P> __device__ __forceinline__ float sphere_zcapsule(P c, float x, float y, float z, float r) {
    float rs = r + c[6];
    return zcapsule_q(c, x, y, z) - rs * rs;
}

// utils.cuh:185-201 for a cuboid whose axes are the world axes: the axis dot
// products are then exactly the coordinate differences.
template <class P> __device__ __forceinline__ float aabb_q(P b, float x, float y, float z) {
    float a1 = fmaxf(fabsf(x - b[0]) - b[3], 0.f);
    float a2 = fmaxf(fabsf(y - b[1]) - b[4], 0.f);
    float a3 = fmaxf(fabsf(z - b[2]) - b[5], 0.f);
    return __builtin_fmaf(a1, a1, __builtin_fmaf(a2, a2, a3 * a3));
}
template <class P> __device__ __forceinline__ float sphere_aabb(P b, float x, float y, float z, float rsq) {
    return aabb_q(b, x, y, z) - rsq;
}

// utils.cuh:185-201
template <class P> __device__ __forceinline__ float obb_q(P b, float x, float y, float z) {
    float xs = x - b[0], ys = y - b[1], zs = z - b[2];
    float a1 = fmaxf(fabsf(__builtin_fmaf(b[3], xs, __builtin_fmaf(b[4], ys, b[5] * zs))) - b[12], 0.f);
    float a2 = fmaxf(fabsf(__builtin_fmaf(b[6], xs, __builtin_fmaf(b[7], ys, b[8] * zs))) - b[13], 0.f);
    float a3 = fmaxf(fabsf(__builtin_fmaf(b[9], xs, __builtin_fmaf(b[10], ys, b[11] * zs))) - b[14], 0.f);
    return __builtin_fmaf(a1, a1, __builtin_fmaf(a2, a2, a3 * a3));
}
template <class P> __device__ __forceinline__ float sphere_obb(P b, float x, float y, float z, float rsq) {
    return obb_q(b, x, y, z) - rsq;
}

// utils.cuh:208-221
template <class P> __device__ __forceinline__ float zcuboid_q(P b, float x, float y, float z) {
    float xs = x - b[0], ys = y - b[1], zs = z - b[2];
    float a1 = fmaxf(fabsf(__builtin_fmaf(b[3], xs, b[4] * ys)) - b[12], 0.f);
    float a2 = fmaxf(fabsf(__builtin_fmaf(b[6], xs, b[7] * ys)) - b[13], 0.f);
    float a3 = fmaxf(fabsf(zs) - b[14], 0.f);
    return __builtin_fmaf(a1, a1, __builtin_fmaf(a2, a2, a3 * a3));
}
template <class P> __device__ __forceinline__ float sphere_zcuboid(P b, float x, float y, float z, float rsq) {
    return zcuboid_q(b, x, y, z) - rsq;
}

// True iff the sphere touches any obstacle (utils.cuh:231-260).
__device__ __forceinline__ bool env_hit(const WorldDesc &w, float x, float y, float z, float r) {
    const float rsq = r * r;
    bool hit = false;
    for (int i = 0; i < w.n_sph; ++i) {
        cfloat *o = world_data(w) + w.o_sph + 4 * i;
        hit |= sphere_sphere(o[0], o[1], o[2], o[3], x, y, z, r) < 0.f;
    }
    for (int i = 0; i < w.n_cap; ++i) hit |= sphere_capsule(world_data(w) + w.o_cap + 8 * i, x, y, z, r) < 0.f;
    for (int i = 0; i < w.n_zcap; ++i) hit |= sphere_zcapsule(world_data(w) + w.o_zcap + 8 * i, x, y, z, r) < 0.f;
    for (int i = 0; i < w.n_aabb; ++i) hit |= sphere_aabb(world_data(w) + w.o_aabb + 6 * i, x, y, z, rsq) < 0.f;
    for (int i = 0; i < w.n_obb; ++i) hit |= sphere_obb(world_data(w) + w.o_obb + 15 * i, x, y, z, rsq) < 0.f;
    for (int i = 0; i < w.n_zcub; ++i) hit |= sphere_zcuboid(world_data(w) + w.o_zcub + 15 * i, x, y, z, rsq) < 0.f;
    return hit;
}

}  // namespace dev
}  // namespace mrpop
