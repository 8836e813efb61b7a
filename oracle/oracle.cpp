// oracle.cpp -- CPU restatement of the reference MR.POP collision checker.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and bench.py's CPU
// baseline may load this library, and only as the checker or as the timed
// CPU baseline; the product never calls it.
//
// Two forms of the same algorithm:
//   * FP64 oracle: every test of the composite verdict is evaluated (no early
//     exit) with its signed clearance, so a sample's verdict can be classified
//     as robust or as "near contact" (some test within eps of 0 could flip it).
//   * FP32 mirror: the reference's float formulas and early exits, one sample
//     after another; this is the CPU baseline that bench.py times.
//
// Semantics followed (reference paths):
//   FK                         panda_four.cuh:44-131 / 289-373, panda.cuh:598-669 / 761-829:
//                              T_0 = base, T_j = T_{j-1} * FT_j * Rz(q_{j-1}), centre = T * p
//   primitive tests            utils.cuh:100-229 (collision iff value < 0)
//   environment loop           utils.cuh:231-260 (spheres, capsules, z-capsules,
//                              cuboids, z-cuboids; cylinders never tested)
//   coarse env                 panda_four.cuh:232-276 / panda.cuh:716-748 (spheres 1..10)
//   coarse self + arm-arm      panda_four.cuh:134-229 / panda.cuh:672-713
//   fine env                   panda_four.cuh:456-489 / panda.cuh:865-888 (spheres 1..58,
//                              gated on the sphere's joint env flag)
//   fine self, multi-arm       panda_four.cuh:376-453 (range i: gate on joint of fine
//                              sphere i, centre of fine sphere i, radius of sphere
//                              ranges[i][0]; inter-arm: sphere i of arm a gated on
//                              arm a's joint flag, against all 59 spheres of arm b > a)
//   fine self, single arm      panda.cuh:832-862 (gate on joint of sphere ranges[i][0])
//   edge interpolation         pop_planner.cu:553-560 (q = from + k*(to-from)/G, k = 1..G)
//   FK+CC drivers              utils.cuh:447-686 (modes FULL, DRRT, SINGLE)
// The per-sample thread re-apportioning of utils.cuh:477-517 is a scheduling
// detail; for a single configuration (every sample the same) it gives every
// sample its share, so the per-configuration verdict is the composite below.
// Along an edge it can leave a sample's fine stages without a thread:
// orc_check_edges_ref_schedule restates it to measure what that changes.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/mrpop_c.h"
#include "../mr.pop_amd/csrc/panda_geometry.h"

namespace {

using namespace mrpop::geom;

constexpr int kModeFull = MRPOP_MODE_FULL;
constexpr int kModeDrrt = MRPOP_MODE_DRRT;
constexpr int kModeSingle = MRPOP_MODE_SINGLE;

// Test switch (orc_set_plain_multi_arm_ranges): when set, the FP64 multi-arm
// fine self ranges are read the single-arm way (gate, centre and radius of
// sphere ranges[r][0]) instead of the reference's multi-arm reading.  Only
// used to find configurations where the reference's reading matters.
int g_plain_multi_arm_ranges = 0;

// ---------------------------------------------------------------- FK ----
// Rigid frame, row-major 3x4 [R | t].
template <class R> struct Frame {
    R m[12];
};

// out = parent * FT * Rz(q); FT is row-major 3x4.
template <class R> void compose(const Frame<R> &p, const float *ft, R c, R s, Frame<R> &out) {
    // step = FT * Rz(q): column 0 = FT_c0*c + FT_c1*s, column 1 = -FT_c0*s + FT_c1*c,
    // column 2 = FT_c2, translation = FT_c3.
    R st[12];
    for (int r = 0; r < 3; ++r) {
        R f0 = ft[4 * r + 0], f1 = ft[4 * r + 1];
        st[4 * r + 0] = f0 * c + f1 * s;
        st[4 * r + 1] = f1 * c - f0 * s;
        st[4 * r + 2] = ft[4 * r + 2];
        st[4 * r + 3] = ft[4 * r + 3];
    }
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 4; ++k) {
            R v = p.m[4 * r + 0] * st[0 + k] + p.m[4 * r + 1] * st[4 + k] + p.m[4 * r + 2] * st[8 + k];
            if (k == 3) v += p.m[4 * r + 3];
            out.m[4 * r + k] = v;
        }
    }
}

// All 8 joint frames of one arm.
template <class R> void arm_frames(const float *base, const R *q, Frame<R> T[kJoints]) {
    for (int i = 0; i < 12; ++i) T[0].m[i] = base[i];
    for (int j = 1; j < kJoints; ++j) {
        R c, s;
        if constexpr (sizeof(R) == sizeof(double)) {
            c = std::cos(q[j - 1]);
            s = std::sin(q[j - 1]);
        } else {
            c = std::cos((float)q[j - 1]);
            s = std::sin((float)q[j - 1]);
        }
        compose(T[j - 1], kJointTf[j], c, s, T[j]);
    }
}

template <class R> void centre(const Frame<R> &T, const float *p, R out[3]) {
    for (int r = 0; r < 3; ++r)
        out[r] = T.m[4 * r + 0] * (R)p[0] + T.m[4 * r + 1] * (R)p[1] + T.m[4 * r + 2] * (R)p[2] + T.m[4 * r + 3];
}

struct ArmSpheres {
    double coarse[kCoarseSpheres][3];
    double fine[kFineSpheres][3];
};

void arm_spheres(const float *base, const double *q, ArmSpheres &out) {
    Frame<double> T[kJoints];
    arm_frames<double>(base, q, T);
    for (int i = 0; i < kCoarseSpheres; ++i) centre(T[kCoarseJoint[i]], kCoarseSphere[i], out.coarse[i]);
    for (int i = 0; i < kFineSpheres; ++i) centre(T[kFineJoint[i]], kFineSphere[i], out.fine[i]);
}

// ------------------------------------------------- FP64 clearances ----
// Signed clearance (metres) of a sphere against each primitive; the
// reference's test "value < 0" holds exactly when this is < 0.
double sphere_sphere_clear(const double *a, double ra, const double *b, double rb) {
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return std::sqrt(dx * dx + dy * dy + dz * dz) - (ra + rb);
}

// utils.cuh:148-158
double sphere_capsule_clear(const float *c, const double *p, double r) {
    double x1 = c[0], y1 = c[1], z1 = c[2], xv = c[3], yv = c[4], zv = c[5], rc = c[6], rdv = c[7];
    double dot = (p[0] - x1) * xv + (p[1] - y1) * yv + (p[2] - z1) * zv;
    double t = std::min(std::max(dot * rdv, 0.0), 1.0);
    double cx = x1 + xv * t, cy = y1 + yv * t, cz = z1 + zv * t;
    double dx = p[0] - cx, dy = p[1] - cy, dz = p[2] - cz;
    return std::sqrt(dx * dx + dy * dy + dz * dz) - (r + rc);
}

// utils.cuh:165-174 (uses only zv of the axis)
double sphere_zcapsule_clear(const float *c, const double *p, double r) {
    double x1 = c[0], y1 = c[1], z1 = c[2], zv = c[5], rc = c[6], rdv = c[7];
    double dot = (p[2] - z1) * zv;
    double t = std::min(std::max(dot * rdv, 0.0), 1.0);
    double dx = p[0] - x1, dy = p[1] - y1, dz = p[2] - (z1 + zv * t);
    return std::sqrt(dx * dx + dy * dy + dz * dz) - (r + rc);
}

// utils.cuh:185-201
double sphere_cuboid_clear(const float *b, const double *p, double r) {
    double xs = p[0] - b[0], ys = p[1] - b[1], zs = p[2] - b[2];
    double s = 0;
    for (int k = 0; k < 3; ++k) {
        double d = std::fabs(b[3 + 3 * k] * xs + b[4 + 3 * k] * ys + b[5 + 3 * k] * zs) - b[12 + k];
        d = std::max(d, 0.0);
        s += d * d;
    }
    return std::sqrt(s) - r;
}

// utils.cuh:208-221 (axes 1 and 2 in the xy plane, axis 3 = z)
double sphere_zcuboid_clear(const float *b, const double *p, double r) {
    double xs = p[0] - b[0], ys = p[1] - b[1], zs = p[2] - b[2];
    double a1 = std::max(std::fabs(b[3] * xs + b[4] * ys) - b[12], 0.0);
    double a2 = std::max(std::fabs(b[6] * xs + b[7] * ys) - b[13], 0.0);
    double a3 = std::max(std::fabs(zs) - b[14], 0.0);
    return std::sqrt(a1 * a1 + a2 * a2 + a3 * a3) - r;
}

// Minimum clearance of a sphere against the whole environment
// (the OR of utils.cuh:231-260 is "min < 0").
double env_clear(const mrpop_scene &s, const double *p, double r) {
    double m = INFINITY;
    for (int i = 0; i < s.n_spheres; ++i) {
        const float *o = s.spheres + 4 * i;
        double c[3] = {o[0], o[1], o[2]};
        m = std::min(m, sphere_sphere_clear(c, o[3], p, r));
    }
    for (int i = 0; i < s.n_capsules; ++i) m = std::min(m, sphere_capsule_clear(s.capsules + 8 * i, p, r));
    for (int i = 0; i < s.n_z_capsules; ++i) m = std::min(m, sphere_zcapsule_clear(s.z_capsules + 8 * i, p, r));
    for (int i = 0; i < s.n_cuboids; ++i) m = std::min(m, sphere_cuboid_clear(s.cuboids + 15 * i, p, r));
    for (int i = 0; i < s.n_z_cuboids; ++i) m = std::min(m, sphere_zcuboid_clear(s.z_cuboids + 15 * i, p, r));
    return m;
}

// Three-valued outcome of one test under a contact tolerance eps.
// hit(bias): bias = 0 nominal, -1 treat |c| < eps as free, +1 as collision.
struct Tri {
    double c;
    bool hit(int bias, double eps) const {
        if (bias == 0) return c < 0;
        if (std::fabs(c) < eps) return bias > 0;
        return c < 0;
    }
};

// Composite per-sample verdict for mode/arms under a bias (see Tri).
// q: n_arms*7 (or 7 for SINGLE).  bases: n_arms*12 (SINGLE: the arm's base).
struct SampleGeo {
    int n_arms;
    ArmSpheres arm[kMaxArms];
};

// Per-sample figures the reference's thread schedule reads (utils.cuh:477-498):
// fine spheres behind the env-flagged and self-flagged joints, summed over the
// arms (panda_four.cuh:216-225, 264-272; panda.cuh:704-711, 739-745, one
// count per flagged joint from the joint-to-sphere-count table), and which of
// the two fine stages reports a collision.
struct Detail {
    int env_spheres = 0, self_spheres = 0;
    bool env_hit = false, self_hit = false;
};

// Fine spheres per joint: the joint-to-sphere-count table of the reference
// (panda_multi_geometry.cuh:362-370, panda.cuh:441-450) is the histogram of
// the sphere-to-joint table.
int joint_sphere_count(int j) {
    int n = 0;
    for (int i = 0; i < kFineSpheres; ++i) n += kFineJoint[i] == j;
    return n;
}

bool composite(const mrpop_scene &scene, int mode, const SampleGeo &g, int bias, double eps,
               double *min_abs_clear, Detail *det = nullptr) {
    const int A = g.n_arms;
    auto note = [&](double c) {
        if (min_abs_clear) *min_abs_clear = std::min(*min_abs_clear, std::fabs(c));
    };
    uint8_t envflag[kMaxArms][kJoints] = {};
    uint8_t selfflag[kMaxArms][kJoints] = {};
    // coarse env: spheres 1..10 (sphere 0 skipped)
    if (mode != kModeDrrt) {
        for (int a = 0; a < A; ++a)
            for (int i = 1; i < kCoarseSpheres; ++i) {
                Tri t{env_clear(scene, g.arm[a].coarse[i], kCoarseSphere[i][3])};
                note(t.c);
                if (t.hit(bias, eps)) envflag[a][kCoarseJoint[i]] = 1;
            }
    }
    // coarse self: intra-arm ranges, flag the joint of the range's first sphere
    for (int a = 0; a < A; ++a)
        for (int r = 0; r < kCoarseSelfRanges; ++r) {
            int s1 = kCoarseSelfRange[r][0];
            for (int j = kCoarseSelfRange[r][1]; j <= kCoarseSelfRange[r][2]; ++j) {
                Tri t{sphere_sphere_clear(g.arm[a].coarse[s1], kCoarseSphere[s1][3], g.arm[a].coarse[j],
                                          kCoarseSphere[j][3])};
                note(t.c);
                if (t.hit(bias, eps)) selfflag[a][kCoarseJoint[s1]] = 1;
            }
        }
    // coarse arm-arm: all 11x11 pairs for a < b, flag arm a's joint
    if (mode != kModeSingle)
        for (int a = 0; a < A; ++a)
            for (int b = a + 1; b < A; ++b)
                for (int i = 0; i < kCoarseSpheres; ++i)
                    for (int j = 0; j < kCoarseSpheres; ++j) {
                        Tri t{sphere_sphere_clear(g.arm[a].coarse[i], kCoarseSphere[i][3], g.arm[b].coarse[j],
                                                  kCoarseSphere[j][3])};
                        note(t.c);
                        if (t.hit(bias, eps)) selfflag[a][kCoarseJoint[i]] = 1;
                    }
    if (det)
        for (int a = 0; a < A; ++a)
            for (int j = 0; j < kJoints; ++j) {
                if (envflag[a][j]) det->env_spheres += joint_sphere_count(j);
                if (selfflag[a][j]) det->self_spheres += joint_sphere_count(j);
            }
    bool hit = false;
    // fine env: spheres 1..58 on env-flagged joints
    if (mode != kModeDrrt)
        for (int a = 0; a < A; ++a)
            for (int i = 1; i < kFineSpheres; ++i) {
                if (!envflag[a][kFineJoint[i]]) continue;
                Tri t{env_clear(scene, g.arm[a].fine[i], kFineSphere[i][3])};
                note(t.c);
                hit |= t.hit(bias, eps);
            }
    const bool env_hit = hit;
    hit = false;
    // fine self
    for (int a = 0; a < A; ++a)
        for (int r = 0; r < kFineSelfRanges; ++r) {
            int s1 = kFineSelfRange[r][0];
            int gate_joint, pos_idx;
            if (mode == kModeSingle || g_plain_multi_arm_ranges) {
                gate_joint = kFineJoint[s1];
                pos_idx = s1;
            } else {
                // multi-arm checker: gate and centre taken from fine sphere r
                gate_joint = kFineJoint[r];
                pos_idx = r;
            }
            if (!selfflag[a][gate_joint]) continue;
            for (int j = kFineSelfRange[r][1]; j <= kFineSelfRange[r][2]; ++j) {
                Tri t{sphere_sphere_clear(g.arm[a].fine[pos_idx], kFineSphere[s1][3], g.arm[a].fine[j],
                                          kFineSphere[j][3])};
                note(t.c);
                hit |= t.hit(bias, eps);
            }
        }
    if (mode != kModeSingle)
        for (int a = 0; a < A; ++a)
            for (int i = 0; i < kFineSpheres; ++i) {
                if (!selfflag[a][kFineJoint[i]]) continue;
                for (int b = a + 1; b < A; ++b)
                    for (int j = 0; j < kFineSpheres; ++j) {
                        Tri t{sphere_sphere_clear(g.arm[a].fine[i], kFineSphere[i][3], g.arm[b].fine[j],
                                                  kFineSphere[j][3])};
                        note(t.c);
                        hit |= t.hit(bias, eps);
                    }
            }
    if (det) {
        det->env_hit = env_hit;
        det->self_hit = hit;
    }
    return env_hit || hit;
}

void sample_geo(int mode, int n_arms, int arm, const float *bases, const double *q, SampleGeo &g) {
    if (mode == kModeSingle) {
        g.n_arms = 1;
        arm_spheres(bases + 12 * arm, q, g.arm[0]);
    } else {
        g.n_arms = n_arms;
        for (int a = 0; a < n_arms; ++a) arm_spheres(bases + 12 * a, q + 7 * a, g.arm[a]);
    }
}

// --------------------------------------------- FP32 mirror (baseline) ----
// The reference's float formulas with its early exits (utils.cuh:100-260).
struct ArmSpheresF {
    float coarse[kCoarseSpheres][3];
    float fine[kFineSpheres][3];
};

inline float ss_f(const float *a, float ra, const float *b, float rb) {
    float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    float rs = ra + rb;
    return dx * dx + dy * dy + dz * dz - rs * rs;
}

bool env_hit_f(const mrpop_scene &s, const float *p, float r) {
    const float rsq = r * r;
    for (int i = 0; i < s.n_spheres; ++i) {
        const float *o = s.spheres + 4 * i;
        if (ss_f(o, o[3], p, r) < 0) return true;
    }
    for (int i = 0; i < s.n_capsules; ++i) {
        const float *c = s.capsules + 8 * i;
        float dot = (p[0] - c[0]) * c[3] + (p[1] - c[1]) * c[4] + (p[2] - c[2]) * c[5];
        float t = std::fmax(std::fmin(dot * c[7], 1.f), 0.f);
        float dx = p[0] - (c[0] + c[3] * t), dy = p[1] - (c[1] + c[4] * t), dz = p[2] - (c[2] + c[5] * t);
        float rs = r + c[6];
        if (dx * dx + dy * dy + dz * dz - rs * rs < 0) return true;
    }
    for (int i = 0; i < s.n_z_capsules; ++i) {
        const float *c = s.z_capsules + 8 * i;
        float dot = (p[2] - c[2]) * c[5];
        float t = std::fmax(std::fmin(dot * c[7], 1.f), 0.f);
        float dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - (c[2] + c[5] * t);
        float rs = r + c[6];
        if (dx * dx + dy * dy + dz * dz - rs * rs < 0) return true;
    }
    for (int i = 0; i < s.n_cuboids; ++i) {
        const float *b = s.cuboids + 15 * i;
        float xs = p[0] - b[0], ys = p[1] - b[1], zs = p[2] - b[2];
        float a1 = std::fmax(0.f, std::fabs(b[3] * xs + b[4] * ys + b[5] * zs) - b[12]);
        float a2 = std::fmax(0.f, std::fabs(b[6] * xs + b[7] * ys + b[8] * zs) - b[13]);
        float a3 = std::fmax(0.f, std::fabs(b[9] * xs + b[10] * ys + b[11] * zs) - b[14]);
        if (a1 * a1 + a2 * a2 + a3 * a3 - rsq < 0) return true;
    }
    for (int i = 0; i < s.n_z_cuboids; ++i) {
        const float *b = s.z_cuboids + 15 * i;
        float xs = p[0] - b[0], ys = p[1] - b[1], zs = p[2] - b[2];
        float a1 = std::fmax(0.f, std::fabs(b[3] * xs + b[4] * ys) - b[12]);
        float a2 = std::fmax(0.f, std::fabs(b[6] * xs + b[7] * ys) - b[13]);
        float a3 = std::fmax(0.f, std::fabs(zs) - b[14]);
        if (a1 * a1 + a2 * a2 + a3 * a3 - rsq < 0) return true;
    }
    return false;
}

void arm_coarse_f(const float *base, const float *q, Frame<float> T[kJoints], ArmSpheresF &s) {
    arm_frames<float>(base, q, T);
    for (int i = 0; i < kCoarseSpheres; ++i) centre(T[kCoarseJoint[i]], kCoarseSphere[i], s.coarse[i]);
}

// One sample with the reference's early exits.  Returns true on collision.
bool sample_f32(const mrpop_scene &scene, int mode, int A, const float *const *arm_base, const float *q) {
    Frame<float> T[kMaxArms][kJoints];
    ArmSpheresF S[kMaxArms];
    uint8_t envflag[kMaxArms][kJoints] = {};
    uint8_t selfflag[kMaxArms][kJoints] = {};
    bool any_env = false, any_self = false;
    for (int a = 0; a < A; ++a) arm_coarse_f(arm_base[a], q + 7 * a, T[a], S[a]);
    if (mode != kModeDrrt)
        for (int a = 0; a < A; ++a)
            for (int i = 1; i < kCoarseSpheres; ++i)
                if (env_hit_f(scene, S[a].coarse[i], kCoarseSphere[i][3])) {
                    envflag[a][kCoarseJoint[i]] = 1;
                    any_env = true;
                }
    for (int a = 0; a < A; ++a) {
        for (int r = 0; r < kCoarseSelfRanges; ++r) {
            int s1 = kCoarseSelfRange[r][0];
            for (int j = kCoarseSelfRange[r][1]; j <= kCoarseSelfRange[r][2]; ++j)
                if (ss_f(S[a].coarse[s1], kCoarseSphere[s1][3], S[a].coarse[j], kCoarseSphere[j][3]) < 0) {
                    selfflag[a][kCoarseJoint[s1]] = 1;
                    any_self = true;
                }
        }
        if (mode != kModeSingle)
            for (int i = 0; i < kCoarseSpheres; ++i)
                for (int b = a + 1; b < A; ++b)
                    for (int j = 0; j < kCoarseSpheres; ++j)
                        if (ss_f(S[a].coarse[i], kCoarseSphere[i][3], S[b].coarse[j], kCoarseSphere[j][3]) < 0) {
                            selfflag[a][kCoarseJoint[i]] = 1;
                            any_self = true;
                        }
    }
    if (!any_env && !any_self) return false;
    for (int a = 0; a < A; ++a)
        for (int i = 0; i < kFineSpheres; ++i) centre(T[a][kFineJoint[i]], kFineSphere[i], S[a].fine[i]);
    if (any_env)
        for (int a = 0; a < A; ++a)
            for (int i = 1; i < kFineSpheres; ++i)
                if (envflag[a][kFineJoint[i]] && env_hit_f(scene, S[a].fine[i], kFineSphere[i][3])) return true;
    if (any_self) {
        for (int a = 0; a < A; ++a) {
            for (int r = 0; r < kFineSelfRanges; ++r) {
                int s1 = kFineSelfRange[r][0];
                int gate = mode == kModeSingle ? kFineJoint[s1] : kFineJoint[r];
                int pos = mode == kModeSingle ? s1 : r;
                if (!selfflag[a][gate]) continue;
                for (int j = kFineSelfRange[r][1]; j <= kFineSelfRange[r][2]; ++j)
                    if (ss_f(S[a].fine[pos], kFineSphere[s1][3], S[a].fine[j], kFineSphere[j][3]) < 0) return true;
            }
            if (mode != kModeSingle)
                for (int i = 0; i < kFineSpheres; ++i) {
                    if (!selfflag[a][kFineJoint[i]]) continue;
                    for (int b = a + 1; b < A; ++b)
                        for (int j = 0; j < kFineSpheres; ++j)
                            if (ss_f(S[a].fine[i], kFineSphere[i][3], S[b].fine[j], kFineSphere[j][3]) < 0) return true;
                }
        }
    }
    return false;
}

int dof(int mode, int n_arms) { return mode == kModeSingle ? 7 : 7 * n_arms; }

}  // namespace

extern "C" {

// Test switch of the FP64 verdicts: 1 = multi-arm fine self ranges read the
// single-arm way (see g_plain_multi_arm_ranges), 0 = the reference's reading.
void orc_set_plain_multi_arm_ranges(int on) { g_plain_multi_arm_ranges = on ? 1 : 0; }

// Sphere centres in FP64 for one configuration of n_arms arms.
// out_coarse: n_arms*11*3, out_fine: n_arms*59*3 (either may be NULL).
void orc_sphere_centres(int n_arms, const float *bases, const double *q, double *out_coarse, double *out_fine) {
    for (int a = 0; a < n_arms; ++a) {
        ArmSpheres s;
        arm_spheres(bases + 12 * a, q + 7 * a, s);
        if (out_coarse) std::memcpy(out_coarse + a * 33, s.coarse, sizeof(s.coarse));
        if (out_fine) std::memcpy(out_fine + a * 177, s.fine, sizeof(s.fine));
    }
}

// Joint frames (8 x 3x4 row-major) of one arm, FP64.
void orc_arm_frames(const float *base, const double *q, double *out /*8*12*/) {
    Frame<double> T[kJoints];
    arm_frames<double>(base, q, T);
    for (int j = 0; j < kJoints; ++j) std::memcpy(out + 12 * j, T[j].m, sizeof(T[j].m));
}

// FP64 signed clearance of one sphere against the scene.
double orc_env_clearance(const mrpop_scene *scene, const double *p, double r) { return env_clear(*scene, p, r); }

// Per-configuration FP64 verdict.  q in double (D = 7*n_arms, or 7 for SINGLE).
// Writes nominal verdict, the verdict with near-contact tests (|clearance| < eps)
// treated as free (lo) and as collision (hi), and the smallest |clearance|.
void orc_check_config(const mrpop_scene *scene, int mode, int n_arms, int arm, const float *bases, const double *q,
                      double eps, int *nominal, int *lo, int *hi, double *min_abs_clear) {
    SampleGeo g;
    sample_geo(mode, n_arms, arm, bases, q, g);
    double m = INFINITY;
    *nominal = composite(*scene, mode, g, 0, eps, &m);
    *lo = composite(*scene, mode, g, -1, eps, nullptr);
    *hi = composite(*scene, mode, g, +1, eps, nullptr);
    if (min_abs_clear) *min_abs_clear = m;
}

// Edge verdicts in FP64.  Sample k = 1..G: q = from + k*(to-from)/G, evaluated
// from the float inputs in double.  out_* are per edge: nominal verdict, lo,
// hi (see orc_check_config); masks get bit k-1 per colliding sample
// (nominal / lo / hi).  Any out pointer may be NULL.
void orc_check_edges(const mrpop_scene *scene, int mode, int n_arms, int arm, const float *bases, int G, int n_edges,
                     const float *from, const float *to, double eps, uint8_t *out_nom, uint8_t *out_lo,
                     uint8_t *out_hi, uint64_t *mask_nom, uint64_t *mask_lo, uint64_t *mask_hi) {
    const int D = dof(mode, n_arms);
#pragma omp parallel for schedule(dynamic, 4)
    for (int e = 0; e < n_edges; ++e) {
        uint64_t mn = 0, ml = 0, mh = 0;
        double q[7 * kMaxArms];
        for (int k = 1; k <= G; ++k) {
            for (int d = 0; d < D; ++d) {
                double f = from[(size_t)e * D + d], t = to[(size_t)e * D + d];
                q[d] = f + k * ((t - f) / G);
            }
            SampleGeo g;
            sample_geo(mode, n_arms, arm, bases, q, g);
            if (composite(*scene, mode, g, 0, eps, nullptr)) mn |= 1ull << (k - 1);
            if (composite(*scene, mode, g, -1, eps, nullptr)) ml |= 1ull << (k - 1);
            if (composite(*scene, mode, g, +1, eps, nullptr)) mh |= 1ull << (k - 1);
        }
        if (out_nom) out_nom[e] = mn != 0;
        if (out_lo) out_lo[e] = ml != 0;
        if (out_hi) out_hi[e] = mh != 0;
        if (mask_nom) mask_nom[e] = mn;
        if (mask_lo) mask_lo[e] = ml;
        if (mask_hi) mask_hi[e] = mh;
    }
}

// The reference's per-edge thread schedule of the fine stages
// (utils.cuh:477-517 for FULL and SINGLE, 653-676 for DRRT): a block has 4*G
// threads; sample i gets (int)((float)n_i / (float)sum(n) * (float)(4*G))
// of them, n_i being its env (or self) sphere count, and the last sample also
// gets whatever the floors leave over.  Thread t works on the sample whose
// running share total first exceeds t, so a sample with a share of 0 has no
// thread and its fine stage never runs.  Division and products are IEEE
// float, as the reference is built without fast math (CMakeLists.txt:9).
void ref_thread_shares(const int *count, int G, int *share) {
    const int threads = 4 * G;
    int total = 0, used = 0;
    for (int i = 0; i < G; ++i) total += count[i];
    for (int i = 0; i < G; ++i) {
        // an all-zero stage is never run (its any-flag is clear); 0/0 would be NaN
        share[i] = total > 0 ? (int)((float)count[i] / (float)total * (float)threads) : 0;
        used += share[i];
    }
    share[G - 1] += threads - used;
}

// FP32 mirror of the reference checker with its early exits: the CPU
// baseline.  Samples of an edge are checked in order k = 1..G and the edge
// stops at its first colliding sample.  n_threads <= 1 runs serially.
void orc_check_edges_f32(const mrpop_scene *scene, int mode, int n_arms, int arm, const float *bases, int G,
                         int n_edges, const float *from, const float *to, uint8_t *out, int n_threads) {
    const int D = dof(mode, n_arms);
    const int A = mode == kModeSingle ? 1 : n_arms;
    const float *arm_base[kMaxArms];
    for (int a = 0; a < A; ++a) arm_base[a] = bases + 12 * (mode == kModeSingle ? arm : a);
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads > 1 ? n_threads : 1) if (n_threads > 1)
    for (int e = 0; e < n_edges; ++e) {
        float delta[7 * kMaxArms], q[7 * kMaxArms];
        const float *f = from + (size_t)e * D;
        for (int d = 0; d < D; ++d) delta[d] = (to[(size_t)e * D + d] - f[d]) / (float)G;
        uint8_t hit = 0;
        for (int k = 1; k <= G && !hit; ++k) {
            for (int d = 0; d < D; ++d) q[d] = f[d] + (float)k * delta[d];
            hit = sample_f32(*scene, mode, A, arm_base, q);
        }
        out[e] = hit;
    }
}

// Edge verdicts (FP64, nominal) under the reference's thread schedule: a
// sample's fine env (self) stage counts only if its env (self) share of the
// block's threads is positive (ref_thread_shares).  out_sched: per-edge
// verdict under the schedule; out_all: every sample checked (the nominal
// verdict of orc_check_edges); mask_dropped: bit k-1 for the colliding
// samples whose colliding stages the schedule left without a thread.
// Any out pointer may be NULL.
void orc_check_edges_ref_schedule(const mrpop_scene *scene, int mode, int n_arms, int arm, const float *bases, int G,
                                  int n_edges, const float *from, const float *to, uint8_t *out_sched,
                                  uint8_t *out_all, uint64_t *mask_dropped) {
    const int D = dof(mode, n_arms);
#pragma omp parallel for schedule(dynamic, 4)
    for (int e = 0; e < n_edges; ++e) {
        Detail det[64];
        int env_n[64], self_n[64], env_share[64], self_share[64];
        double q[7 * kMaxArms];
        bool all = false;
        for (int k = 1; k <= G; ++k) {
            for (int d = 0; d < D; ++d) {
                double f = from[(size_t)e * D + d], t = to[(size_t)e * D + d];
                q[d] = f + k * ((t - f) / G);
            }
            SampleGeo g;
            sample_geo(mode, n_arms, arm, bases, q, g);
            det[k - 1] = Detail{};
            all |= composite(*scene, mode, g, 0, 0.0, nullptr, &det[k - 1]);
            env_n[k - 1] = det[k - 1].env_spheres;
            self_n[k - 1] = det[k - 1].self_spheres;
        }
        ref_thread_shares(env_n, G, env_share);
        ref_thread_shares(self_n, G, self_share);
        bool sched = false;
        uint64_t dropped = 0;
        for (int k = 0; k < G; ++k) {
            const bool env_run = det[k].env_hit && env_share[k] > 0;
            const bool self_run = det[k].self_hit && self_share[k] > 0;
            sched |= env_run || self_run;
            if ((det[k].env_hit || det[k].self_hit) && !env_run && !self_run) dropped |= 1ull << k;
        }
        if (out_sched) out_sched[e] = sched;
        if (out_all) out_all[e] = all;
        if (mask_dropped) mask_dropped[e] = dropped;
    }
}

// The schedule's shares for given per-sample sphere counts (known-answer tests).
void orc_ref_thread_shares(const int *count, int G, int *share) { ref_thread_shares(count, G, share); }

// Geometry tables as compiled into this library (for the table fixture test).
void orc_tables(float *joint_tf /*8*12*/, float *coarse /*11*4*/, int *coarse_joint, float *fine /*59*4*/,
                int *fine_joint, int *coarse_ranges /*5*3*/, int *fine_ranges /*25*3*/, float *scale, float *low,
                float *base4 /*4*12*/, float *base5 /*5*12*/) {
    std::memcpy(joint_tf, kJointTf, sizeof(kJointTf));
    std::memcpy(coarse, kCoarseSphere, sizeof(kCoarseSphere));
    std::memcpy(coarse_joint, kCoarseJoint, sizeof(kCoarseJoint));
    std::memcpy(fine, kFineSphere, sizeof(kFineSphere));
    std::memcpy(fine_joint, kFineJoint, sizeof(kFineJoint));
    std::memcpy(coarse_ranges, kCoarseSelfRange, sizeof(kCoarseSelfRange));
    std::memcpy(fine_ranges, kFineSelfRange, sizeof(kFineSelfRange));
    std::memcpy(scale, kScale, sizeof(kScale));
    std::memcpy(low, kLow, sizeof(kLow));
    std::memcpy(base4, kBase4, sizeof(kBase4));
    std::memcpy(base5, kBase5, sizeof(kBase5));
}

}  // extern "C"
