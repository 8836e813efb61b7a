// check_kernels.h -- batched FK + two-phase collision check of interpolated edges (gfx950).
//
// Replaces the reference's fkcc_single_buffer / _mr / _drrt
// (src/planning/utils.cuh:447-686) with the same per-sample verdict:
//   coarse pass  11 spheres per arm: env test of spheres 1..10 sets the env bit
//                of the sphere's joint (panda_four.cuh:232-276); the 21
//                intra-arm pairs and, for several arms, all 11x11 pairs of each
//                arm pair a < b set the self bit of arm a's joint
//                (panda_four.cuh:134-229);
//   fine pass    59 spheres per arm: spheres 1..58 on env-flagged joints against
//                the environment (panda_four.cuh:456-489); intra-arm pairs
//                gated on the self bit (multi-arm: range i gated by and centred
//                on fine sphere i with the radius of sphere ranges[i][0],
//                panda_four.cuh:390-416; one arm: gated by sphere ranges[i][0],
//                panda.cuh:832-862); inter-arm: sphere i of arm a on a
//                self-flagged joint against all 59 spheres of every arm b > a
//                (panda_four.cuh:419-450).
//   A sample collides iff some fine test does; an edge iff some sample does.
//
// MI355X mapping.
//   check_kernel<A, MODE>: a block of A waves checks a tile of 64 samples;
//   wave a holds arm a of every sample (arm-major), so the obstacle and
//   arm-pair culls of a wave see one arm of 64 neighbouring samples.  FK runs
//   per item in registers with the fixed joint transforms folded in; coarse
//   centres and a bounding ball per arm go to LDS, where the arm-arm pairs
//   split evenly over the A items of a sample (each takes A-1 half-pairs:
//   rows 0-5 or 6-10 of arm a against arm b; rows whose sphere misses arm b's
//   ball are skipped).  Obstacle records are wave-uniform and come through
//   the scalar cache.  The environment tests run only for pairs (sphere,
//   obstacle) where the obstacle's cull box meets the box of that sphere's
//   centres over the wave: the wave computes the 10 boxes from LDS and one
//   lane per obstacle tests them, so a wave of 64 samples typically runs
//   about 8 of the 300 sphere-obstacle tests of a 30-obstacle world (the
//   persistent kernels cull per item with one box around the arm).
//   The fine self, environment and arm-arm tests run in the same kernel,
//   gated exactly as above, but only where they can hit: each coarse sphere
//   grown by a few tenths of a millimetre (kBlockInflate) contains the fine
//   spheres of its block, so a fine pair or obstacle is tested only if its
//   block ball overlaps.  The culling never changes a verdict.  A tile needs
//   three barriers; a persistent grid of resident blocks loops over the
//   tiles of many edge groups (check_loop).
//   check_kernel_mixed<MODE>: blocks of five waves, tiles of 320/A samples of
//   the group's arm count, for planner rounds over problems of several arm
//   counts.
//
// This header holds the device code.  Each kernel instantiation is compiled
// in one object built from check_inst.hip (so the instantiations build in
// parallel); check.hip holds the launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "device_geom.h"
#include "internal.h"

namespace mrpop {
namespace dev {

using u64 = unsigned long long;

}  // namespace dev
}  // namespace mrpop
#include "check_groups.h"
namespace mrpop {
namespace dev {
using EdgeBatch = EdgeGroup;

// ---------------------------------------------------------------------------
// Lane-indexed constant tables (offsets, radii, joints of the fine spheres).
struct SphereTable {
    float off[kFineSpheres][4];
    float r[kFineSpheres];
    unsigned char joint[kFineSpheres];
    float coarse_r[kCoarseSpheres];
    float block_r[kCoarseSpheres];  // coarse radius + inflation: ball holding the block's fine spheres
    unsigned char block_first[kCoarseSpheres];
    unsigned char block_count[kCoarseSpheres];
};
constexpr SphereTable build_sphere_table() {
    SphereTable t{};
    for (int i = 0; i < kFineSpheres; ++i) {
        for (int c = 0; c < 4; ++c) t.off[i][c] = kFineSphere[i][c];
        t.r[i] = kFineSphere[i][3];
        t.joint[i] = (unsigned char)kFineJoint[i];
    }
    for (int i = 0; i < kCoarseSpheres; ++i) {
        t.coarse_r[i] = kCoarseSphere[i][3];
        t.block_r[i] = kCoarseSphere[i][3] + kBlockInflate[i];
        t.block_first[i] = (unsigned char)kBlockFirst[i];
        t.block_count[i] = (unsigned char)kBlockCount[i];
    }
    return t;
}
static __constant__ SphereTable kTab = build_sphere_table();

constexpr bool coarse_joint_is_min_i_7() {
    for (int i = 0; i < kCoarseSpheres; ++i)
        if (kCoarseJoint[i] != (i < 7 ? i : 7)) return false;
    return true;
}
static_assert(coarse_joint_is_min_i_7(), "coarse sphere i belongs to joint min(i, 7)");
constexpr bool blocks_follow_joints() {
    for (int k = 0; k < kCoarseSpheres; ++k)
        for (int i = 0; i < kBlockCount[k]; ++i)
            if (kFineJoint[kBlockFirst[k] + i] != kCoarseJoint[k]) return false;
    return kBlockFirst[5] == 17 && kBlockFirst[10] + kBlockCount[10] == kFineSpheres;
}
static_assert(blocks_follow_joints(), "fine-sphere blocks lie on their coarse sphere's joint");

// Fine intra-arm ranges (src/robots/panda_four.cuh:388-417, panda.cuh:832-862):
// the sphere whose centre is tested, the sphere whose radius is used, the
// joint whose self bit gates the range.
template <int MODE> constexpr int range_pos(int r) {
    return MODE == MRPOP_MODE_SINGLE ? kFineSelfRange[r][0] : r;  // multi-arm: fine sphere r itself
}
constexpr int range_rad(int r) { return kFineSelfRange[r][0]; }
template <int MODE> constexpr int range_gate(int r) { return kFineJoint[range_pos<MODE>(r)]; }

// Ranges gated on joint 5 test spheres on joint 5 against targets on joint
// 7 only, so the edge check evaluates them during the coarse FK (where both
// frames are at hand) and applies the gate afterwards.  Coarse sphere 5
// overlaps the hand's coarse spheres in practically every configuration,
// so the gate is almost always open and nothing is computed in vain.
template <int MODE> constexpr bool gate5_ranges_target_joint7() {
    for (int r = 0; r < kFineSelfRanges; ++r)
        if (range_gate<MODE>(r) == 5)
            for (int t = kFineSelfRange[r][1]; t <= kFineSelfRange[r][2]; ++t)
                if (kFineJoint[t] != 7) return false;
    return true;
}
static_assert(gate5_ranges_target_joint7<MRPOP_MODE_FULL>() && gate5_ranges_target_joint7<MRPOP_MODE_SINGLE>(),
              "joint-5 gated ranges target the hand (joint 7)");

// ------------------------------------------------------------- FK helpers --
__device__ __forceinline__ void load_base(const WorldDesc &w, int arm, Frame &T) {
    cfloat *b = world_data(w) + w.o_base + 12 * arm;
#pragma unroll
    for (int i = 0; i < 12; ++i) T.m[i] = b[i];
}

// (to - from) / G.  For G a power of two the quotient is exact, so it
// equals the product with 1/G (one instruction instead of a correctly
// rounded division); G is uniform over a wave tile.
__device__ __forceinline__ float step_of(float t, float f, int G) {
    if ((G & (G - 1)) == 0) return (t - f) * __builtin_amdgcn_ldexpf(1.0f, -__builtin_ctz(G));
    return (t - f) / (float)G;
}

// Interpolated joint angles of one arm: q = from + k * ((to - from) / G)
// (pop_planner.cu:553-560).
__device__ __forceinline__ void interp7(const EdgeBatch &eb, int e, int k, int off, float q[7]) {
    const float *f = eb.from + (size_t)e * eb.dim + off;
    const float *t = eb.to + (size_t)e * eb.dim + off;
    const float kf = (float)k;
#pragma unroll
    for (int j = 0; j < 7; ++j) q[j] = __builtin_fmaf(kf, step_of(t[j], f[j], eb.G), f[j]);
}

// Joint transforms for a rolled (runtime-depth) FK walk.
struct JointTable {
    float tf[kJoints][12];
};
constexpr JointTable build_joint_table() {
    JointTable t{};
    for (int j = 0; j < kJoints; ++j)
        for (int i = 0; i < 12; ++i) t.tf[j][i] = kJointTf[j][i];
    return t;
}
static __constant__ JointTable kJT = build_joint_table();

// Frame of joint `jt` (runtime 0..7) of an arm at sample (e, k), walked with a
// rolled loop.  Every operation matches joint_step on the same inputs (fma
// with a 0 / +-1 factor is exact), so the frame equals the unrolled one.
__device__ __forceinline__ Frame frame_rt(const WorldDesc &w, const EdgeBatch &eb, int base_arm, int dof_off, int e,
                                       int k, int jt) {
    Frame T;
    load_base(w, base_arm, T);
    const float *f = eb.from + (size_t)e * eb.dim + dof_off;
    const float *t = eb.to + (size_t)e * eb.dim + dof_off;
    const float kf = (float)k;
#pragma unroll 1
    for (int J = 1; J <= jt; ++J) {
        const float qj = __builtin_fmaf(kf, step_of(t[J - 1], f[J - 1], eb.G), f[J - 1]);
        float sq, cq;
        joint_sincos(qj, sq, cq);
        const float *F = kJT.tf[J];
        float S[12];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            S[4 * r + 0] = __builtin_fmaf(F[4 * r + 1], sq, F[4 * r + 0] * cq);
            S[4 * r + 1] = __builtin_fmaf(-F[4 * r + 0], sq, F[4 * r + 1] * cq);
        }
        Frame N;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                float v = T.m[4 * r + 0] * S[c];
                v = __builtin_fmaf(T.m[4 * r + 1], S[4 + c], v);
                v = __builtin_fmaf(T.m[4 * r + 2], S[8 + c], v);
                N.m[4 * r + c] = v;
            }
            N.m[4 * r + 2] = __builtin_fmaf(F[10], T.m[4 * r + 2], __builtin_fmaf(F[6], T.m[4 * r + 1], F[2] * T.m[4 * r + 0]));
            N.m[4 * r + 3] = __builtin_fmaf(F[11], T.m[4 * r + 2],
                                            __builtin_fmaf(F[7], T.m[4 * r + 1], __builtin_fmaf(F[3], T.m[4 * r + 0], T.m[4 * r + 3])));
        }
        T = N;
    }
    return T;
}

__device__ __forceinline__ void centre_rt(const Frame &T, int s, float &x, float &y, float &z) {
    centre(T, kTab.off[s][0], kTab.off[s][1], kTab.off[s][2], x, y, z);
}

// ------------------------------------------------------------ fused check --
// Orders this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A value every lane of the wave holds alike, as a scalar.
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// Calls f(i) for the obstacles i (0 .. n-1) of one type whose bits, from bit
// `first` on, are set in m; for every one of them if all.
template <class F>
__device__ __forceinline__ void for_touched(unsigned long long m, int first, int n, bool all, F &&f) {
    if (all) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    if (n == 0) return;  // first may then be 64
    unsigned long long b = m >> first;
    if (n < 64) b &= (1ull << n) - 1ull;
    while (b) {
        f(__builtin_ctzll(b));
        b &= b - 1ull;
    }
}

// Fine environment tests of block k's spheres (all but sphere 0) in frame T,
// only against obstacles that its bounding ball (centre b, radius R) touches.
// touch: a wave-uniform superset of the obstacles (bit = index in the order
// spheres, capsules, z-capsules, aabbs, obbs, z-cuboids) that the ball can
// touch on a lane of the wave, or all = true to visit every obstacle.
__device__ __forceinline__ bool block_env(const WorldDesc &w, const Frame &T, int k, float bx, float by, float bz,
                                          float R, unsigned long long touch, bool all) {
    bool hit = false;
    const int j0 = max((int)kTab.block_first[k], 1), j1 = kTab.block_first[k] + kTab.block_count[k];
    const float Rsq = R * R;
    int first = 0;
    for_touched(touch, first, w.n_sph, all, [&](int i) {
        cfloat *o = world_data(w) + w.o_sph + 4 * i;
        if (sphere_sphere(o[0], o[1], o[2], o[3], bx, by, bz, R) < 0.f)
            for (int j = j0; j < j1; ++j) {
                float x, y, z;
                centre_rt(T, j, x, y, z);
                hit |= sphere_sphere(o[0], o[1], o[2], o[3], x, y, z, kTab.r[j]) < 0.f;
            }
    });
    first += w.n_sph;
    for_touched(touch, first, w.n_cap, all, [&](int i) {
        cfloat *o = world_data(w) + w.o_cap + 8 * i;
        if (sphere_capsule(o, bx, by, bz, R) < 0.f)
            for (int j = j0; j < j1; ++j) {
                float x, y, z;
                centre_rt(T, j, x, y, z);
                hit |= sphere_capsule(o, x, y, z, kTab.r[j]) < 0.f;
            }
    });
    first += w.n_cap;
    for_touched(touch, first, w.n_zcap, all, [&](int i) {
        cfloat *o = world_data(w) + w.o_zcap + 8 * i;
        if (sphere_zcapsule(o, bx, by, bz, R) < 0.f)
            for (int j = j0; j < j1; ++j) {
                float x, y, z;
                centre_rt(T, j, x, y, z);
                hit |= sphere_zcapsule(o, x, y, z, kTab.r[j]) < 0.f;
            }
    });
    first += w.n_zcap;
    for_touched(touch, first, w.n_aabb, all, [&](int i) {
        cfloat *o = world_data(w) + w.o_aabb + 6 * i;
        if (sphere_aabb(o, bx, by, bz, Rsq) < 0.f)
            for (int j = j0; j < j1; ++j) {
                float x, y, z;
                centre_rt(T, j, x, y, z);
                hit |= sphere_aabb(o, x, y, z, kTab.r[j] * kTab.r[j]) < 0.f;
            }
    });
    first += w.n_aabb;
    for_touched(touch, first, w.n_obb, all, [&](int i) {
        cfloat *o = world_data(w) + w.o_obb + 15 * i;
        if (sphere_obb(o, bx, by, bz, Rsq) < 0.f)
            for (int j = j0; j < j1; ++j) {
                float x, y, z;
                centre_rt(T, j, x, y, z);
                hit |= sphere_obb(o, x, y, z, kTab.r[j] * kTab.r[j]) < 0.f;
            }
    });
    first += w.n_obb;
    for_touched(touch, first, w.n_zcub, all, [&](int i) {
        cfloat *o = world_data(w) + w.o_zcub + 15 * i;
        if (sphere_zcuboid(o, bx, by, bz, Rsq) < 0.f)
            for (int j = j0; j < j1; ++j) {
                float x, y, z;
                centre_rt(T, j, x, y, z);
                hit |= sphere_zcuboid(o, x, y, z, kTab.r[j] * kTab.r[j]) < 0.f;
            }
    });
    return hit;
}

// One FK walk of this lane's arm doing, exactly as the reference gates them:
//   * the fine intra-arm ranges whose gate joint lies in [G0, G1], with
//     bounding-ball culling of the target blocks;
//   * if WITH_ENV, the fine environment tests of every env-flagged joint,
//     with bounding-ball culling of the obstacles.
// Run once for gates on joints 0..4 and once for joint 5 plus the
// environment, so that at most 17 tested centres are live at a time.
template <int MODE, int G0, int G1, bool WITH_ENV, int NT>
__device__ __forceinline__ bool fine_walk(const WorldDesc &w, const Frame &base, const float q[7], unsigned selfbits,
                                          unsigned envbits, const float (*cs)[3][NT], int lane,
                                          const unsigned long long (&touch)[kCoarseSpheres], bool all) {
    bool hit = false;
    float px[kFineSelfRanges], py[kFineSelfRanges], pz[kFineSelfRanges];
    Frame T = base;
    static_for<0, kJoints>([&](auto jc) {
        constexpr int J = decltype(jc)::value;
        if constexpr (J > 0) {
            Frame N;
            joint_frame<J>(T, q[J - 1], N);
            T = N;
        }
        // centres of the ranges' tested spheres lying on joint J
        if constexpr (J >= G0 && J <= G1) {
            if (selfbits & (1u << J)) {
                static_for<0, kFineSelfRanges>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    constexpr int ps = range_pos<MODE>(r);
                    if constexpr (kFineJoint[ps] == J) {
                        constexpr float ox = kFineSphere[ps][0], oy = kFineSphere[ps][1], oz = kFineSphere[ps][2];
                        centre(T, ox, oy, oz, px[r], py[r], pz[r]);
                    }
                });
            }
        }
        // fine environment tests of joint J (sphere 0 is never tested: panda_four.cuh:468)
        if constexpr (WITH_ENV && J > 0) {
            if (envbits & (1u << J)) {
                static_for<1, kCoarseSpheres>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    if constexpr (kCoarseJoint[k] == J) {
                        constexpr float R = kCoarseSphere[k][3] + kBlockInflate[k];
                        hit |= block_env(w, T, k, cs[k][0][lane], cs[k][1][lane], cs[k][2][lane], R,
                                         uniform_u64(touch[k]), all);
                    }
                });
            }
        }
        // target blocks on joint J (targets are fine spheres 17..58: blocks 5..10)
        static_for<5, kCoarseSpheres>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            if constexpr (kCoarseJoint[k] == J) {
                const float bx = cs[k][0][lane], by = cs[k][1][lane], bz = cs[k][2][lane];
                constexpr int b0 = kBlockFirst[k], b1 = kBlockFirst[k] + kBlockCount[k] - 1;
                constexpr float R = kCoarseSphere[k][3] + kBlockInflate[k];
                static_for<0, kFineSelfRanges>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    constexpr int lo = kFineSelfRange[r][1] > b0 ? kFineSelfRange[r][1] : b0;
                    constexpr int hi = kFineSelfRange[r][2] < b1 ? kFineSelfRange[r][2] : b1;
                    constexpr int gj = range_gate<MODE>(r);
                    if constexpr (lo <= hi && gj >= G0 && gj <= G1) {
                        constexpr float rad = kFineSphere[range_rad(r)][3];
                        if (selfbits & (1u << gj)) {
                            const float dx = px[r] - bx, dy = py[r] - by, dz = pz[r] - bz;
                            const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                            if (d2 < (rad + R) * (rad + R)) {
                                for (int t = lo; t <= hi; ++t) {
                                    float x, y, z;
                                    centre_rt(T, t, x, y, z);
                                    hit |= sphere_sphere(px[r], py[r], pz[r], rad, x, y, z, kTab.r[t]) < 0.f;
                                }
                            }
                        }
                    }
                });
            }
        });
    });
    return hit;
}

// Decode half-pair h (0 .. A(A-1)-1) into arm pair (a, b), a < b, and the
// first row i0 (0 or 6) of arm a's coarse spheres it covers.
template <int A> __device__ __forceinline__ void half_pair(int h, int &a, int &b, int &i0) {
    const int p = h >> 1;
    int aa = 0, rem = p;
#pragma unroll
    for (int t = 0; t < A - 1; ++t) {
        const int row = A - 1 - t;
        if (aa == t && rem >= row) {
            rem -= row;
            aa = t + 1;
        }
    }
    a = aa;
    b = aa + 1 + rem;
    i0 = (h & 1) ? 6 : 0;
}

// Largest radius among the environment-tested coarse spheres 1..10.
constexpr float env_radius_max() {
    float r = 0.f;
    for (int s = 1; s < kCoarseSpheres; ++s) r = kCoarseSphere[s][3] > r ? kCoarseSphere[s][3] : r;
    return r;
}
constexpr float kEnvRadiusMax = env_radius_max();

// Largest block radius (coarse radius + inflation) over the coarse spheres,
// and the slack added to arm balls: 0.1 mm, orders of magnitude above the
// float32 rounding of distances of a few metres.
constexpr float max_block_radius() {
    float r = 0.f;
    for (int s = 0; s < kCoarseSpheres; ++s)
        r = kCoarseSphere[s][3] + kBlockInflate[s] > r ? kCoarseSphere[s][3] + kBlockInflate[s] : r;
    return r;
}
constexpr float kMaxBlockRadius = max_block_radius();
constexpr float kBallMargin = 1e-4f;

// Sub-blocks of the fine arm-arm stage: each block's fine spheres in groups
// of at most six that lie close together, with a ball holding each group
// (centre = mean of the offsets, radius = largest |offset - centre| + r,
// plus kBallMargin for float rounding).  Two groups whose balls miss each
// other hold no pair of overlapping spheres, so their pairs are skipped.
// The hand's 18-sphere block is a 6 x 3 grid: its groups are pairs of
// columns (radius 5.3 cm against 10.5 cm for the block); the 12-sphere block
// of joint 5 splits into its two upper spheres, two lower ones and the two
// four-sphere rails.
constexpr int kSubBlocks = 16;
inline constexpr unsigned char kSubIdx[kFineSpheres] = {
    0,                                  // block 0
    1, 2, 3, 4,                         // 1
    5, 6, 7, 8,                         // 2
    9, 10, 11, 12,                      // 3
    13, 14, 15, 16,                     // 4
    17, 18, 19, 20, 21, 22, 25, 26, 23, 24, 27, 28,        // 5: four groups
    29, 30, 31,                         // 6
    32, 33, 34, 35, 36,                 // 7
    37, 38, 43, 44, 49, 50, 39, 40, 45, 46, 51, 52, 41, 42, 47, 48, 53, 54,  // 8: three groups
    55, 56,                             // 9
    57, 58};                            // 10
inline constexpr unsigned char kSubFirst[kSubBlocks] = {0, 1, 5, 9, 13, 17, 19, 21, 25, 29, 32, 37, 43, 49, 55, 57};
inline constexpr unsigned char kSubCount[kSubBlocks] = {1, 4, 4, 4, 4, 2, 2, 4, 4, 3, 5, 6, 6, 6, 2, 2};
inline constexpr unsigned char kBlockSub0[kCoarseSpheres] = {0, 1, 2, 3, 4, 5, 9, 10, 11, 14, 15};  // first group
inline constexpr unsigned char kBlockSubs[kCoarseSpheres] = {1, 1, 1, 1, 1, 4, 1, 1, 3, 1, 1};      // groups

constexpr double csqrt(double x) {
    double r = x > 1.0 ? x : 1.0;
    for (int i = 0; i < 64; ++i) r = 0.5 * (r + x / r);
    return r;
}
struct SubBall {
    float c[kSubBlocks][3];
    float r[kSubBlocks];
    unsigned char idx[kFineSpheres], first[kSubBlocks], count[kSubBlocks];
    unsigned char sub0[kCoarseSpheres], subs[kCoarseSpheres];
};
constexpr SubBall build_sub_balls() {
    SubBall b{};
    for (int i = 0; i < kFineSpheres; ++i) b.idx[i] = kSubIdx[i];
    for (int g = 0; g < kSubBlocks; ++g) {
        b.first[g] = kSubFirst[g];
        b.count[g] = kSubCount[g];
    }
    for (int k = 0; k < kCoarseSpheres; ++k) {
        b.sub0[k] = kBlockSub0[k];
        b.subs[k] = kBlockSubs[k];
    }
    for (int g = 0; g < kSubBlocks; ++g) {
        double c[3] = {0, 0, 0};
        for (int m = 0; m < kSubCount[g]; ++m)
            for (int d = 0; d < 3; ++d) c[d] += kFineSphere[kSubIdx[kSubFirst[g] + m]][d];
        for (int d = 0; d < 3; ++d) c[d] /= kSubCount[g];
        double r = 0;
        for (int m = 0; m < kSubCount[g]; ++m) {
            const float *f = kFineSphere[kSubIdx[kSubFirst[g] + m]];
            const double dx = f[0] - c[0], dy = f[1] - c[1], dz = f[2] - c[2];
            const double e = csqrt(dx * dx + dy * dy + dz * dz) + f[3];
            r = e > r ? e : r;
        }
        for (int d = 0; d < 3; ++d) b.c[g][d] = (float)c[d];
        b.r[g] = (float)(r + kBallMargin);
    }
    return b;
}
static __constant__ SubBall kSub = build_sub_balls();
constexpr bool sub_blocks_partition_blocks() {
    // every block's groups are consecutive and list exactly its fine spheres
    int seen[kFineSpheres] = {};
    for (int k = 0; k < kCoarseSpheres; ++k) {
        int total = 0;
        for (int g = kBlockSub0[k]; g < kBlockSub0[k] + kBlockSubs[k]; ++g) {
            if (g >= kSubBlocks || kSubFirst[g] != kSubFirst[kBlockSub0[k]] + total) return false;
            for (int m = 0; m < kSubCount[g]; ++m) {
                const int i = kSubIdx[kSubFirst[g] + m];
                if (i < kBlockFirst[k] || i >= kBlockFirst[k] + kBlockCount[k]) return false;
                ++seen[i];
            }
            total += kSubCount[g];
        }
        if (total != kBlockCount[k]) return false;
    }
    for (int i = 0; i < kFineSpheres; ++i)
        if (seen[i] != 1) return false;
    return true;
}
static_assert(sub_blocks_partition_blocks(), "fine-sphere groups partition each block");

// Shared memory of one tile of NT items (sample, arm); KH = arm-arm
// half-pairs per item.
template <int KH, int NT> struct TileLds {
    alignas(16) float cs[kCoarseSpheres][3][NT];
    float hull[NT / 64][kCoarseSpheres - 1][6];             // per wave: box of the centres of spheres 1..10
    unsigned long long touch[NT / 64][kCoarseSpheres];      // per wave and block: candidate obstacles 0..63
    float ball[4][NT];                   // per item: its arm's bounding ball (centre, radius)
    unsigned inter_flags[NT];
    unsigned result[NT];
    unsigned long long cand_lo[KH][NT];  // candidate blocks, rows 0..4: bit t*11 + k
    unsigned cand_hi[KH][NT];            // row 5: bit k
    unsigned edge_done[NT];              // per edge of the tile: a sample collides
    unsigned fine[NT];                   // per item: its fine tests found a collision
    int hp_next;                         // next half-pair of the fine arm-arm stage to hand out
};

// Samples per edge in a launch of the given pass.  Pass 0 checks every
// sample k = 1..G; a probe pass P > 0 checks min(P, G) of them spread along
// the edge, the last being k = G; pass -1 checks every sample of the edges
// a probe pass left free.  The edge verdict, the OR over its samples, is
// the same either way.
__device__ __forceinline__ int pass_samples(const EdgeGroup &g, int pass) {
    return pass > 0 ? min(pass, g.G) : g.G;
}

// Tiles of a group: a tile is W waves, 64 W / A samples of all A arms of
// the group (one arm in SINGLE mode).
__device__ __forceinline__ unsigned group_waves(const EdgeGroup &g, int single, int pass, int W) {
    const int S = single ? 64 * W : (64 * W) / (g.dim / 7);
    const unsigned long long samples = (unsigned long long)(g.count ? *g.count : g.n_edges) * pass_samples(g, pass);
    return (unsigned)((samples + S - 1) / S);
}

// One tile of NT / 64 waves: samples lw*S .. lw*S + S - 1 of group eb,
// S = NT / A, thread arm * S + slot holding arm `arm` of sample `slot`.  The
// waves are arm-major, so the environment and arm-pair culls of a wave see
// one arm of 64 nearby samples.
// LEAN (the one-tile-per-block kernels, whose registers are not held across
// a tile loop): evaluate the joint-5 gated fine ranges during the coarse FK
// rather than in the fine walk of phase B, and record per block the
// obstacles its ball touches, so that the fine environment tests visit only
// those.
template <int A, int MODE, int KH, int NT, bool LEAN>
__device__ __forceinline__ void check_tile(TileLds<KH, NT> &sm, const WorldDesc *worlds, const EdgeGroup &eb,
                                           unsigned lw, const FineQueue &fq, int pass) {
    constexpr int S = NT / A;  // samples per tile
    constexpr bool kEnv = MODE != MRPOP_MODE_DRRT;
    constexpr bool kInter = A > 1;
    const int lane = threadIdx.x;  // the thread's own item
    const bool item = lane < S * A;
    // spare threads (NT % A) shadow arm 0 of sample 0 and discard
    const int arm = item ? lane / S : 0;
    const int slot = item ? lane - arm * S : 0;
    const WorldDesc w = worlds[eb.world];
    const int per = pass_samples(eb, pass);
    const long total = (long)(eb.count ? *eb.count : eb.n_edges) * per;
    const long n = (long)lw * S + slot;
    bool active = item && n < total;
    const long nn = active ? n : 0;
    const int e = (int)(nn / per);
    const int j = (int)(nn - (long)e * per);
    const int k = per == eb.G ? j + 1 : ((j + 1) * eb.G) / per;  // sample k of k = 1..G
    if (pass < 0) {
        // the remaining samples of edges still free after the probe pass
        if (active && eb.out_edge[e] != 0u) active = false;
        if (!__syncthreads_or(active)) return;
    }
    const int base_arm = MODE == MRPOP_MODE_SINGLE ? eb.arm : arm;
    const int dof_off = MODE == MRPOP_MODE_SINGLE ? 0 : 7 * arm;

    // ---- A. coarse FK of this lane's arm, and the fine self ranges gated on
    // joint 5 (tested spheres on joint 5, targets on joint 7), evaluated
    // here while both frames pass by; their gate is applied in phase B
    float cx[kCoarseSpheres], cy[kCoarseSpheres], cz[kCoarseSpheres];
    bool self5hit = false;
    {
        float q[7];
        interp7(eb, e, k, dof_off, q);
        Frame T;
        load_base(w, base_arm, T);
        {
            constexpr float ox = kCoarseSphere[0][0], oy = kCoarseSphere[0][1], oz = kCoarseSphere[0][2];
            centre(T, ox, oy, oz, cx[0], cy[0], cz[0]);
        }
        float px[kFineSelfRanges], py[kFineSelfRanges], pz[kFineSelfRanges];
        static_for<1, kJoints>([&](auto jc) {
            constexpr int J = decltype(jc)::value;
            Frame N;
            joint_frame<J>(T, q[J - 1], N);
            T = N;
            static_for<kCoarseJointStart[J], kCoarseJointStart[J] + kCoarseJointCount[J]>([&](auto sc) {
                constexpr int s = decltype(sc)::value;
                constexpr float ox = kCoarseSphere[s][0], oy = kCoarseSphere[s][1], oz = kCoarseSphere[s][2];
                centre(T, ox, oy, oz, cx[s], cy[s], cz[s]);
            });
            if constexpr (LEAN && J == 5)
                static_for<0, kFineSelfRanges>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    if constexpr (range_gate<MODE>(r) == 5) {
                        constexpr int ps = range_pos<MODE>(r);
                        centre(T, kFineSphere[ps][0], kFineSphere[ps][1], kFineSphere[ps][2], px[r], py[r], pz[r]);
                    }
                });
            if constexpr (LEAN && J == 7)
                static_for<0, kFineSelfRanges>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    if constexpr (range_gate<MODE>(r) == 5) {
                        constexpr float rad = kFineSphere[range_rad(r)][3];
                        static_for<7, kCoarseSpheres>([&](auto kc) {  // blocks on joint 7
                            constexpr int kb = decltype(kc)::value;
                            constexpr int b0 = kBlockFirst[kb], b1 = kBlockFirst[kb] + kBlockCount[kb] - 1;
                            constexpr int lo = kFineSelfRange[r][1] > b0 ? kFineSelfRange[r][1] : b0;
                            constexpr int hi = kFineSelfRange[r][2] < b1 ? kFineSelfRange[r][2] : b1;
                            if constexpr (lo <= hi) {
                                constexpr float R = kCoarseSphere[kb][3] + kBlockInflate[kb];
                                const float dx = px[r] - cx[kb], dy = py[r] - cy[kb], dz = pz[r] - cz[kb];
                                const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                                if (d2 < (rad + R) * (rad + R)) {
                                    static_for<lo, hi + 1>([&](auto tc) {
                                        constexpr int t = decltype(tc)::value;
                                        float x, y, z;
                                        centre(T, kFineSphere[t][0], kFineSphere[t][1], kFineSphere[t][2], x, y, z);
                                        self5hit |= sphere_sphere(px[r], py[r], pz[r], rad, x, y, z,
                                                                  kFineSphere[t][3]) < 0.f;
                                    });
                                }
                            }
                        });
                    }
                });
        });
    }
#pragma unroll
    for (int s = 0; s < kCoarseSpheres; ++s) {
        sm.cs[s][0][lane] = cx[s];
        sm.cs[s][1][lane] = cy[s];
        sm.cs[s][2][lane] = cz[s];
    }
    sm.inter_flags[lane] = 0u;
    if constexpr (kInter) {
        // A ball around this arm's coarse block balls: centre = mean of the
        // centres, radius = largest centre distance + largest block radius +
        // a margin far above float rounding.  A sphere whose block ball
        // misses it can neither hit nor be a candidate of any of the arm's
        // spheres in the arm-arm stage.
        float mx = cx[0], my = cy[0], mz = cz[0];
#pragma unroll
        for (int s = 1; s < kCoarseSpheres; ++s) {
            mx += cx[s];
            my += cy[s];
            mz += cz[s];
        }
        constexpr float inv = 1.0f / kCoarseSpheres;
        mx *= inv;
        my *= inv;
        mz *= inv;
        float r2 = 0.f;
#pragma unroll
        for (int s = 0; s < kCoarseSpheres; ++s) {
            const float dx = cx[s] - mx, dy = cy[s] - my, dz = cz[s] - mz;
            r2 = fmaxf(r2, __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)));
        }
        sm.ball[0][lane] = mx;
        sm.ball[1][lane] = my;
        sm.ball[2][lane] = mz;
        sm.ball[3][lane] = sqrtf(r2) + kMaxBlockRadius + kBallMargin;
    }

    // coarse environment: spheres 1..10 -> env bit of their joint (panda_four.cuh:241-261)
    // (the lean kernels run it after barrier (1), below)
    unsigned hitmask = 0;
    if constexpr (kEnv && !LEAN) {
        // Box of this lane's tested centres (spheres 1..10): an obstacle whose
        // cull box, grown by the largest of their radii, misses it cannot hit
        // any of them.
        float lx = cx[1], ly = cy[1], lz = cz[1], hx = cx[1], hy = cy[1], hz = cz[1];
#pragma unroll
        for (int s = 2; s < kCoarseSpheres; ++s) {
            lx = fminf(lx, cx[s]);
            ly = fminf(ly, cy[s]);
            lz = fminf(lz, cz[s]);
            hx = fmaxf(hx, cx[s]);
            hy = fmaxf(hy, cy[s]);
            hz = fmaxf(hz, cz[s]);
        }
        cfloat *box = world_data(w) + w.o_box;
        auto near = [&](int ob) {
            cfloat *bb = box + 8 * ob;
            const float g = kEnvRadiusMax * bb[6];
            return !(hx < bb[0] - g || hy < bb[1] - g || hz < bb[2] - g || lx > bb[3] + g || ly > bb[4] + g ||
                     lz > bb[5] + g);
        };
        // obstacle of radius a (0 for cuboids) against spheres 1..10: the
        // coarse hit is q - (a + r)^2 < 0 for the obstacle's distance term q
        auto test_all = [&](float a, auto &&qf) {
            static_for<1, kCoarseSpheres>([&](auto sc) {
                constexpr int s = decltype(sc)::value;
                constexpr float r = kCoarseSphere[s][3];
                const float q = qf(cx[s], cy[s], cz[s]);
                const float rs = a + r;
                if (q - rs * rs < 0.f) hitmask |= 1u << s;
            });
        };
        int ob = 0;
        for (int i = 0; i < w.n_sph; ++i, ++ob) {
            if (!near(ob)) continue;
            cfloat *o = world_data(w) + w.o_sph + 4 * i;
            const float ox = o[0], oy = o[1], oz = o[2];
            test_all(o[3], [&](float x, float y, float z) { return sphere_sphere_q(ox, oy, oz, x, y, z); });
        }
        for (int i = 0; i < w.n_cap; ++i, ++ob) {
            if (!near(ob)) continue;
            cfloat *o = world_data(w) + w.o_cap + 8 * i;
            test_all(o[6], [&](float x, float y, float z) { return capsule_q(o, x, y, z); });
        }
        for (int i = 0; i < w.n_zcap; ++i, ++ob) {
            if (!near(ob)) continue;
            cfloat *o = world_data(w) + w.o_zcap + 8 * i;
            test_all(o[6], [&](float x, float y, float z) { return zcapsule_q(o, x, y, z); });
        }
        for (int i = 0; i < w.n_aabb; ++i, ++ob) {
            if (!near(ob)) continue;
            cfloat *o = world_data(w) + w.o_aabb + 6 * i;
            test_all(0.f, [&](float x, float y, float z) { return aabb_q(o, x, y, z); });
        }
        for (int i = 0; i < w.n_obb; ++i, ++ob) {
            if (!near(ob)) continue;
            cfloat *o = world_data(w) + w.o_obb + 15 * i;
            test_all(0.f, [&](float x, float y, float z) { return obb_q(o, x, y, z); });
        }
        for (int i = 0; i < w.n_zcub; ++i, ++ob) {
            if (!near(ob)) continue;
            cfloat *o = world_data(w) + w.o_zcub + 15 * i;
            test_all(0.f, [&](float x, float y, float z) { return zcuboid_q(o, x, y, z); });
        }
    }

    // coarse intra-arm pairs -> self bit of the first sphere's joint (panda_four.cuh:144-176)
    unsigned selfbits = 0;
    static_for<0, kCoarseSelfRanges>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        constexpr int s1 = kCoarseSelfRange[r][0];
        constexpr float r1 = kCoarseSphere[s1][3];
        static_for<kCoarseSelfRange[r][1], kCoarseSelfRange[r][2] + 1>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            constexpr float rj = kCoarseSphere[j][3];
            constexpr unsigned bit = 1u << kCoarseJoint[s1];
            if (sphere_sphere(cx[s1], cy[s1], cz[s1], r1, cx[j], cy[j], cz[j], rj) < 0.f) selfbits |= bit;
        });
    });
    // Barriers of a tile: (1) coarse centres, balls and intra-arm self bits
    // published, (2) arm-arm flags complete, (3) fine results published.
    // Every LDS entry is written by its own item before (1), or between (1)
    // and (3) with all readers before (3); so a wave that has passed (3) may
    // start the next tile while others finish this one.
    sm.result[lane] = selfbits;  // intra-arm part of the self bits
    sm.edge_done[lane] = 0u;
    __syncthreads();
    if (lane == 0) sm.hp_next = 0;  // read only after barrier (2)

    // Lean kernels: the coarse environment tests, with this wave's centres in
    // LDS.  (a) Lanes 0..29 take the hull (box) of one coordinate of one of
    // spheres 1..10 over the wave's 64 lanes.  (b) One lane per obstacle:
    // bit s of `cand` is set if the obstacle's cull box, grown by block s's
    // radius, meets sphere s's hull; a sphere or block ball whose centre lies
    // outside it cannot hit or touch the obstacle (the cull box property).
    // (c) The exact tests of each lane's spheres against the candidates, the
    // reference's q - (a + r)^2 < 0.  The candidate masks, per block, are also
    // the obstacles the fine environment tests visit.
    if constexpr (kEnv && LEAN) {
        const int wv = lane >> 6, wl = lane & 63;
        if (wl < 3 * (kCoarseSpheres - 1)) {
            const int s = 1 + wl / 3, c = wl - 3 * (s - 1);
            const float *v = &sm.cs[s][c][64 * wv];
            float lo = v[0], hi = v[0];
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const float4 x = *(const float4 *)(v + 4 * ((m + wl) & 15));  // rotated: spread over the banks
                lo = fminf(lo, fminf(fminf(x.x, x.y), fminf(x.z, x.w)));
                hi = fmaxf(hi, fmaxf(fmaxf(x.x, x.y), fmaxf(x.z, x.w)));
            }
            sm.hull[wv][s - 1][c] = lo;
            sm.hull[wv][s - 1][3 + c] = hi;
        }
        wave_sync();
        const int n_obs = w.n_sph + w.n_cap + w.n_zcap + w.n_aabb + w.n_obb + w.n_zcub;
        const int e_sph = w.n_sph, e_cap = e_sph + w.n_cap, e_zcap = e_cap + w.n_zcap, e_aabb = e_zcap + w.n_aabb,
                  e_obb = e_aabb + w.n_obb;
        cfloat *box = world_data(w) + w.o_box;
        for (int first = 0; first < n_obs; first += 64) {
            unsigned cand = 0;
            if (first + wl < n_obs) {
                cfloat *bb = box + 8 * (first + wl);
                const float b0 = bb[0], b1 = bb[1], b2 = bb[2], b3 = bb[3], b4 = bb[4], b5 = bb[5], f = bb[6];
                static_for<1, kCoarseSpheres>([&](auto sc) {
                    constexpr int s = decltype(sc)::value;
                    constexpr float R = kCoarseSphere[s][3] + kBlockInflate[s];
                    const float g = R * f;
                    const float *h = sm.hull[wv][s - 1];
                    if (!(h[3] < b0 - g || h[4] < b1 - g || h[5] < b2 - g || h[0] > b3 + g || h[1] > b4 + g ||
                          h[2] > b5 + g))
                        cand |= 1u << s;
                });
            }
            if (first == 0)
                static_for<1, kCoarseSpheres>([&](auto sc) {
                    constexpr int s = decltype(sc)::value;
                    sm.touch[wv][s] = __ballot((cand >> s) & 1u);
                });
            // obstacle of radius a (0 for cuboids) against the spheres in sb
            auto tests = [&](unsigned sb, float a, auto &&qf) {
                static_for<1, kCoarseSpheres>([&](auto sc) {
                    constexpr int s = decltype(sc)::value;
                    if (sb & (1u << s)) {
                        constexpr float r = kCoarseSphere[s][3];
                        const float q = qf(cx[s], cy[s], cz[s]);
                        const float rs = a + r;
                        if (q - rs * rs < 0.f) hitmask |= 1u << s;
                    }
                });
            };
            for (unsigned long long u = __ballot(cand != 0u); u; u &= u - 1ull) {
                const int j = __builtin_ctzll(u);
                const unsigned sb = __builtin_amdgcn_readlane(cand, j);
                const int ob = first + j;
                if (ob < e_sph) {
                    cfloat *o = world_data(w) + w.o_sph + 4 * ob;
                    const float ox = o[0], oy = o[1], oz = o[2];
                    tests(sb, o[3], [&](float x, float y, float z) { return sphere_sphere_q(ox, oy, oz, x, y, z); });
                } else if (ob < e_cap) {
                    cfloat *o = world_data(w) + w.o_cap + 8 * (ob - e_sph);
                    tests(sb, o[6], [&](float x, float y, float z) { return capsule_q(o, x, y, z); });
                } else if (ob < e_zcap) {
                    cfloat *o = world_data(w) + w.o_zcap + 8 * (ob - e_cap);
                    tests(sb, o[6], [&](float x, float y, float z) { return zcapsule_q(o, x, y, z); });
                } else if (ob < e_aabb) {
                    cfloat *o = world_data(w) + w.o_aabb + 6 * (ob - e_zcap);
                    tests(sb, 0.f, [&](float x, float y, float z) { return aabb_q(o, x, y, z); });
                } else if (ob < e_obb) {
                    cfloat *o = world_data(w) + w.o_obb + 15 * (ob - e_aabb);
                    tests(sb, 0.f, [&](float x, float y, float z) { return obb_q(o, x, y, z); });
                } else {
                    cfloat *o = world_data(w) + w.o_zcub + 15 * (ob - e_obb);
                    tests(sb, 0.f, [&](float x, float y, float z) { return zcuboid_q(o, x, y, z); });
                }
            }
        }
    }
    unsigned envbits = 0;
    static_for<1, kCoarseSpheres>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr unsigned bit = 1u << kCoarseJoint[s];
        if (hitmask & (1u << s)) envbits |= bit;
    });

    // coarse arm-arm pairs (panda_four.cuh:178-214), A-1 half-pairs per lane.
    // The same distances give the candidate blocks of the fine arm-arm check:
    // blocks p of arm a and k of arm b whose bounding balls overlap.
    if constexpr (kInter) {
#pragma unroll
        for (int m = 0; m < A - 1; ++m) {
            int a, b, i0;
            half_pair<A>(arm + m * A, a, b, i0);
            // rows whose block ball meets arm b's ball; rows no lane of the
            // wave needs are skipped (their tests cannot hit)
            unsigned need = 0;
            {
                const float ux = sm.ball[0][b * S + slot], uy = sm.ball[1][b * S + slot], uz = sm.ball[2][b * S + slot];
                const float ur = sm.ball[3][b * S + slot];
#pragma unroll
                for (int t = 0; t < 6; ++t) {
                    const int i = min(i0 + t, kCoarseSpheres - 1);
                    const float dx = sm.cs[i][0][a * S + slot] - ux, dy = sm.cs[i][1][a * S + slot] - uy,
                                dz = sm.cs[i][2][a * S + slot] - uz;
                    const float lim = ur + kTab.block_r[i];
                    if (__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)) < lim * lim) need |= 1u << t;
                }
            }
            // rows 6..10 are five: row 5 of that half would repeat row 10,
            // whose hit bit row 4 already gives (its candidates are kept
            // only for the half of rows 0..5)
            if (i0 != 0) need &= 0x1fu;
            if (!active) need = 0;
            const unsigned long long rows_any = __ballot(need != 0u);
            if (!rows_any) {
                sm.cand_lo[m][lane] = 0ull;
                sm.cand_hi[m][lane] = 0u;
                continue;
            }
            float bx[kCoarseSpheres], by[kCoarseSpheres], bz[kCoarseSpheres];
#pragma unroll
            for (int j = 0; j < kCoarseSpheres; ++j) {
                bx[j] = sm.cs[j][0][b * S + slot];
                by[j] = sm.cs[j][1][b * S + slot];
                bz[j] = sm.cs[j][2][b * S + slot];
            }
            unsigned mask = 0;
            unsigned long long clo = 0;
            unsigned chi = 0;
            // Rows 0..5 (i0 = 0) or 6..10 (i0 = 6) of arm a against the 11
            // coarse spheres of arm b, one code path per half so that each
            // row's radius is known at compile time: both thresholds of a
            // pair, (r_i + r_j)^2 and the block balls' (R_i + R_j)^2, are the
            // same float values the runtime sums and products give, folded
            // into literals.
            auto rows = [&](auto i0c) {
                constexpr int I0 = decltype(i0c)::value;
                static_for<0, (I0 == 0 ? 6 : 5)>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr int i = I0 + t;
                    if (!__ballot((need >> t) & 1u)) return;
                    const float ax = sm.cs[i][0][a * S + slot], ay = sm.cs[i][1][a * S + slot],
                                az = sm.cs[i][2][a * S + slot];
                    constexpr float rad = kCoarseSphere[i][3], brad = kCoarseSphere[i][3] + kBlockInflate[i];
                    bool hit = false;
                    unsigned cand = 0;
                    static_for<0, kCoarseSpheres>([&](auto jc) {
                        constexpr int j = decltype(jc)::value;
                        constexpr float rj = kCoarseSphere[j][3], Rj = kCoarseSphere[j][3] + kBlockInflate[j];
                        constexpr float rs = rad + rj, Rs = brad + Rj;
                        constexpr float rs2 = rs * rs, Rs2 = Rs * Rs;
                        const float dx = ax - bx[j], dy = ay - by[j], dz = az - bz[j];
                        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                        hit |= d2 < rs2;  // d2 - (r_i + r_j)^2 < 0
                        if (d2 < Rs2) cand |= 1u << j;
                    });
                    if (hit) mask |= 1u << (i < 7 ? i : 7);
                    if constexpr (t < 5)
                        clo |= (unsigned long long)cand << (11 * t);
                    else
                        chi = cand;  // row 5 of the half of rows 0..5
                });
            };
            if (i0 == 0)
                rows(IntC<0>{});
            else
                rows(IntC<6>{});
            sm.cand_lo[m][lane] = active ? clo : 0ull;
            sm.cand_hi[m][lane] = active ? chi : 0u;
            if (mask && item) atomicOr(&sm.inter_flags[a * S + slot], mask);
        }
    }
    __syncthreads();
    if constexpr (kInter) selfbits |= sm.inter_flags[lane];

    // Edge-only verdicts (no per-sample mask, no diagnostics): once one
    // sample of an edge collides, the fine tests left for that edge's lanes
    // cannot change its verdict and are skipped.  A tile holds the samples
    // of at most S consecutive edges; le is this lane's edge within it.
    const bool early_out = !eb.out_mask && !fq.diag;
    const int le = e - (int)(((long)lw * S) / per);
    volatile unsigned *done = sm.edge_done;

    // ---- B. fine intra-arm pairs of this lane's arm
    bool finehit = false;
    // candidate obstacles of the fine environment tests: the lean kernels'
    // per-block masks (worlds of at most 64 obstacles), else every obstacle
    const bool all_obs = !LEAN || w.n_sph + w.n_cap + w.n_zcap + w.n_aabb + w.n_obb + w.n_zcub > 64;
    const unsigned long long(&touch)[kCoarseSpheres] = sm.touch[lane >> 6];
    if (active && (selfbits | envbits)) {
        float q[7];
        interp7(eb, e, k, dof_off, q);
        Frame base;
        load_base(w, base_arm, base);
        if (selfbits & 0x1fu)
            finehit = fine_walk<MODE, 0, 4, false, NT>(w, base, q, selfbits, 0u, sm.cs, lane, touch, true);
        if constexpr (LEAN) {
            if (selfbits & 0x20u) finehit |= self5hit;  // the joint-5 ranges, evaluated in phase A
            // environment only (no self range has a gate in 6..5)
            if (kEnv && envbits)
                finehit |= fine_walk<MODE, 6, 5, kEnv, NT>(w, base, q, 0u, envbits, sm.cs, lane, touch, all_obs);
        } else if ((selfbits & 0x20u) || envbits) {
            finehit |= fine_walk<MODE, 5, 5, kEnv, NT>(w, base, q, selfbits, envbits, sm.cs, lane, touch, all_obs);
        }
    }
    if (active && finehit) done[le] = 1u;

    // ---- C. fine arm-arm pairs for candidate blocks of the half-pairs.
    // With one wave per arm (64 samples per tile) the tile's A (A - 1)
    // half-pairs go to the waves as they come free, each wave taking a
    // half-pair for all 64 samples: a fixed share kept the waves holding
    // close arm pairs busy while the others waited at barrier (3).  Other
    // tilings keep the fixed share: each item takes its arm's A - 1.
    if constexpr (kInter) {
        constexpr bool kShare = S == 64;
        int m_next = 0;
#pragma unroll 1
        while (true) {
            int h;
            if constexpr (kShare) {
                int g = 0;
                if ((lane & 63) == 0) g = atomicAdd(&sm.hp_next, 1);
                h = __shfl(g, 0, 64);
                if (h >= A * (A - 1)) break;
                if (early_out && active && done[le]) continue;
            } else {
                if (m_next >= A - 1 || (early_out && active && done[le])) break;
                h = arm + m_next * A;
                ++m_next;
            }
            // half-pair h = arm w's m-th: its candidates sit in item (w, slot)
            const int m = h / A, owner = h - m * A;
            const int src = kShare ? owner * S + slot : lane;
            int a, b, i0;
            half_pair<A>(h, a, b, i0);
            const unsigned gate = sm.result[a * S + slot] | sm.inter_flags[a * S + slot];  // arm a's self bits
            unsigned long long clo = sm.cand_lo[m][src];
            unsigned chi = sm.cand_hi[m][src];
            // keep rows whose block's joint is self-flagged on arm a
#pragma unroll
            for (int t = 0; t < 5; ++t) {
                const int p = i0 + t;
                if (!((gate >> (p < 7 ? p : 7)) & 1u)) clo &= ~(0x7ffull << (11 * t));
            }
            if (!((gate >> 5) & 1u)) chi = 0;  // row 5 exists only for i0 == 0 (block 5)
            // frames of the last joints walked: consecutive candidate blocks
            // mostly lie on the same joints (the hand's four blocks on 7)
            Frame Ta, Tb;
            int ja = -1, jb = -1;
            while (clo | chi) {
                int t, kb;
                if (clo) {
                    const int bit = __builtin_ctzll(clo);
                    clo &= clo - 1;
                    t = bit / 11;
                    kb = bit - 11 * t;
                } else {
                    kb = __builtin_ctz(chi);
                    chi &= chi - 1;
                    t = 5;
                }
                const int p = i0 + t;
                if ((p < 7 ? p : 7) != ja) {
                    ja = p < 7 ? p : 7;
                    Ta = frame_rt(w, eb, a, 7 * a, e, k, ja);
                }
                if ((kb < 7 ? kb : 7) != jb) {
                    jb = kb < 7 ? kb : 7;
                    Tb = frame_rt(w, eb, b, 7 * b, e, k, jb);
                }
                // groups of block p against groups of block kb; only
                // groups whose balls meet have their sphere pairs tested
                const int ga0 = kSub.sub0[p], ga1 = ga0 + kSub.subs[p];
                const int gb0 = kSub.sub0[kb], gb1 = gb0 + kSub.subs[kb];
                for (int ga = ga0; ga < ga1; ++ga) {
                    float gx, gy, gz;
                    centre(Ta, kSub.c[ga][0], kSub.c[ga][1], kSub.c[ga][2], gx, gy, gz);
                    for (int gb = gb0; gb < gb1; ++gb) {
                        float hx, hy, hz;
                        centre(Tb, kSub.c[gb][0], kSub.c[gb][1], kSub.c[gb][2], hx, hy, hz);
                        if (sphere_sphere(gx, gy, gz, kSub.r[ga], hx, hy, hz, kSub.r[gb]) >= 0.f) continue;
                        for (int m = kSub.first[ga]; m < kSub.first[ga] + kSub.count[ga]; ++m) {
                            const int i = kSub.idx[m];
                            float ax, ay, az;
                            centre_rt(Ta, i, ax, ay, az);
                            const float ar = kTab.r[i];
                            // a sphere that misses group gb's ball misses its spheres
                            if (sphere_sphere(ax, ay, az, ar, hx, hy, hz, kSub.r[gb]) >= 0.f) continue;
                            for (int mb = kSub.first[gb]; mb < kSub.first[gb] + kSub.count[gb]; ++mb) {
                                const int j = kSub.idx[mb];
                                float x, y, z;
                                centre_rt(Tb, j, x, y, z);
                                finehit |= sphere_sphere(ax, ay, az, ar, x, y, z, kTab.r[j]) < 0.f;
                            }
                        }
                    }
                }
                if (finehit) done[le] = 1u;
                if (early_out && done[le]) break;
            }
        }
    }

    // ---- D. per-sample verdict: OR over the sample's lanes
    if (fq.diag && active)
        fq.diag[(size_t)n * 8 + arm] = envbits | (selfbits << 8) | (finehit ? 1u << 16 : 0u) | (1u << 17);
    sm.fine[lane] = finehit ? 1u : 0u;
    __syncthreads();
    bool coll = false;
    if (active && arm == 0) {
#pragma unroll
        for (int a = 0; a < A; ++a) coll |= sm.fine[a * S + slot] != 0u;
        if (fq.diag) fq.diag[(size_t)n * 8 + 7] = (coll ? 1u : 0u) | (1u << 17);
    }
    // The lean kernels write the edge verdict once per edge and wave, from
    // the lowest colliding lane of the edge (lanes holding arm 0 have
    // lane == slot, and an edge's samples occupy consecutive lanes); the
    // others from every colliding sample.
    bool first = true;
    if constexpr (LEAN) {
        const unsigned long long cm = __ballot(coll);
        const int wl = lane & 63;
        const long n0 = (long)lw * S + (lane - wl);        // sample of the wave's lane 0
        const int lo = (int)max((long)e * per - n0, 0L);  // the wave's first lane holding edge e (<= wl if coll)
        const unsigned long long before = ((1ull << wl) - 1ull) & ~((1ull << min(lo, wl)) - 1ull);  // lanes lo .. wl-1
        first = !(cm & before);
    }
    if (coll) {
        if (first) atomicOr(&eb.out_edge[e], 1u);
        if (eb.out_mask) atomicOr(&eb.out_mask[e], 1ull << (k - 1));
    }
}

// The wave tiles of n_groups edge groups (or of the single group g0 when
// groups is null).  Up to 64 groups the wave computes the groups' first
// tiles itself, one group per lane (wave_start is then null); more groups
// come with a prefix from wave_prefix_kernel.  A == 0: groups of different
// arm counts (multi-arm modes), dispatched per tile.  Every wave of the
// block computes the same group lookup.
// PERSISTENT = false: block b checks tile b (blocks beyond the last tile
// return), so the grid must cover every tile.  That is the fast form:
// gfx950 VOP3 instructions take no 32-bit literals, the fine tests' constants
// are materialised in registers, and a loop around the tile lets the
// compiler hoist them all out of it and spill them (168 VGPRs and 48 spilled
// against 138 and none for the 4-arm kernel).  PERSISTENT = true: a grid of
// any size loops over the tiles, for launches whose tile count is only
// known on the device and bounded far above it.
template <int MODE, int W, bool PERSISTENT, class Tile>
__device__ __forceinline__ void check_loop(const EdgeGroup *groups, const EdgeGroup &g0, int n_groups,
                                           const unsigned *wave_start, int pass, Tile &&tile) {
    const int lane = threadIdx.x & 63;
    unsigned lane_start = 0, total_waves;
    if (!groups) {
        total_waves = group_waves(g0, MODE == MRPOP_MODE_SINGLE, pass, W);
    } else if (!wave_start) {
        const unsigned wcount =
            lane < n_groups ? group_waves(groups[lane], MODE == MRPOP_MODE_SINGLE, pass, W) : 0u;
        unsigned incl = wcount;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned y = __shfl_up(incl, off);
            if (lane >= off) incl += y;
        }
        lane_start = incl - wcount;
        total_waves = __shfl(incl, 63);
    } else {
        total_waves = wave_start[n_groups];
    }
    auto run = [&](unsigned wv) {
        int gi = 0;
        unsigned gstart = 0;
        if (groups && !wave_start) {  // last group whose first wave is <= wv (empty groups share starts)
            const unsigned long long m = __ballot(lane < n_groups && lane_start <= wv);
            gi = 63 - __builtin_clzll(m);
            gstart = __shfl(lane_start, gi);
        } else if (groups) {
            int lo = 0, hi = n_groups - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (wave_start[mid] <= wv) lo = mid;
                else hi = mid - 1;
            }
            gi = lo;
            gstart = wave_start[gi];
        }
        const EdgeGroup eb = groups ? groups[gi] : g0;
        tile(eb, wv - gstart);
    };
    if constexpr (PERSISTENT) {
        for (unsigned wv = blockIdx.x; wv < total_waves; wv += gridDim.x) run(wv);
    } else if (blockIdx.x < total_waves) {
        // Tile b on block b.  Blocks go round-robin to the 8 XCDs, so
        // neighbouring tiles, whose edge endpoints share cache lines, run on
        // different L2s and those lines are fetched once per XCD (HBM reads
        // 1.55x the endpoint bytes); giving each XCD consecutive tiles cut
        // that to 1.12x but made the 4-arm check 10 % slower.
        run(blockIdx.x);
    }
}

// A waves per block, one per arm; three waves per SIMD (the LDS of a tile
// allows three blocks per CU).
template <int A, int MODE, bool PERSISTENT>
__global__ __launch_bounds__(64 * A, 3) void check_kernel(const WorldDesc *worlds, const EdgeGroup *groups,
                                                       EdgeGroup g0, int n_groups, const unsigned *wave_start,
                                                       FineQueue fq, int pass) {
    constexpr int KH = A > 1 ? A - 1 : 1;
    __shared__ TileLds<KH, 64 * A> sm;
    check_loop<MODE, A, PERSISTENT>(groups, g0, n_groups, wave_start, pass, [&](const EdgeGroup &eb, unsigned lw) {
        check_tile<A, MODE, KH, 64 * A, !PERSISTENT>(sm, worlds, eb, lw, fq, pass);
    });
}

// Multi-arm groups of 2..5 arms in one launch: each tile runs the code of
// its group's arm count, so a planner round checks all its problems in one
// launch per mode instead of one per arm count.
template <int MODE, bool PERSISTENT>
__global__ __launch_bounds__(64 * kMixedWaves, 3) void check_kernel_mixed(const WorldDesc *worlds,
                                                                       const EdgeGroup *groups, int n_groups,
                                                                       const unsigned *wave_start, int pass) {
    constexpr int NT = 64 * kMixedWaves;
    __shared__ TileLds<4, NT> sm;
    const EdgeGroup none{};
    const FineQueue fq{nullptr};
    check_loop<MODE, kMixedWaves, PERSISTENT>(groups, none, n_groups, wave_start, pass,
                                             [&](const EdgeGroup &eb, unsigned lw) {
        switch (eb.dim / 7) {
            case 2: check_tile<2, MODE, 4, NT, !PERSISTENT>(sm, worlds, eb, lw, fq, pass); break;
            case 3: check_tile<3, MODE, 4, NT, !PERSISTENT>(sm, worlds, eb, lw, fq, pass); break;
            case 4: check_tile<4, MODE, 4, NT, !PERSISTENT>(sm, worlds, eb, lw, fq, pass); break;
            default: check_tile<5, MODE, 4, NT, !PERSISTENT>(sm, worlds, eb, lw, fq, pass); break;
        }
    });
}

}  // namespace dev
}  // namespace mrpop

// ------------------------------------------------------- launch wrappers --
namespace mrpop {
namespace dev {

// persistent = false: blocks must be the tile count (one tile per block).
template <int A, int MODE>
void launch_check_kernel(int blocks, bool persistent, hipStream_t stream, const WorldDesc *worlds,
                         const EdgeGroup *groups, const EdgeGroup &g0, int n_groups, const unsigned *wave_start,
                         const FineQueue &fq, int pass) {
    if (persistent)
        hipLaunchKernelGGL((check_kernel<A, MODE, true>), dim3(blocks), dim3(64 * A), 0, stream, worlds, groups, g0,
                           n_groups, wave_start, fq, pass);
    else
        hipLaunchKernelGGL((check_kernel<A, MODE, false>), dim3(blocks), dim3(64 * A), 0, stream, worlds, groups, g0,
                           n_groups, wave_start, fq, pass);
}

template <int MODE>
void launch_check_kernel_mixed(int blocks, bool persistent, hipStream_t stream, const WorldDesc *worlds,
                               const EdgeGroup *groups, int n_groups, const unsigned *wave_start, int pass) {
    if (persistent)
        hipLaunchKernelGGL((check_kernel_mixed<MODE, true>), dim3(blocks), dim3(64 * kMixedWaves), 0, stream, worlds,
                           groups, n_groups, wave_start, pass);
    else
        hipLaunchKernelGGL((check_kernel_mixed<MODE, false>), dim3(blocks), dim3(64 * kMixedWaves), 0, stream, worlds,
                           groups, n_groups, wave_start, pass);
}

#define MRPOP_CHECK_KERNEL(EXT, A, MODE)                                                                         \
    EXT template void launch_check_kernel<A, MODE>(int, bool, hipStream_t, const WorldDesc *, const EdgeGroup *, \
                                                   const EdgeGroup &, int, const unsigned *, const FineQueue &,  \
                                                   int);
#define MRPOP_CHECK_KERNEL_MIXED(EXT, MODE)                                                                \
    EXT template void launch_check_kernel_mixed<MODE>(int, bool, hipStream_t, const WorldDesc *,             \
                                                      const EdgeGroup *, int, const unsigned *, int);
// every instantiation lives in one object of check_inst.hip
#define MRPOP_CHECK_ALL(EXT)                       \
    MRPOP_CHECK_KERNEL(EXT, 1, MRPOP_MODE_SINGLE)  \
    MRPOP_CHECK_KERNEL(EXT, 2, MRPOP_MODE_FULL)    \
    MRPOP_CHECK_KERNEL(EXT, 3, MRPOP_MODE_FULL)    \
    MRPOP_CHECK_KERNEL(EXT, 4, MRPOP_MODE_FULL)    \
    MRPOP_CHECK_KERNEL(EXT, 5, MRPOP_MODE_FULL)    \
    MRPOP_CHECK_KERNEL(EXT, 2, MRPOP_MODE_DRRT)    \
    MRPOP_CHECK_KERNEL(EXT, 3, MRPOP_MODE_DRRT)    \
    MRPOP_CHECK_KERNEL(EXT, 4, MRPOP_MODE_DRRT)    \
    MRPOP_CHECK_KERNEL(EXT, 5, MRPOP_MODE_DRRT)    \
    MRPOP_CHECK_KERNEL_MIXED(EXT, MRPOP_MODE_FULL) \
    MRPOP_CHECK_KERNEL_MIXED(EXT, MRPOP_MODE_DRRT)
#ifndef MRPOP_CHECK_INST
MRPOP_CHECK_ALL(extern)
#endif

}  // namespace dev
}  // namespace mrpop
