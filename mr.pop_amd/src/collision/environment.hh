// environment.hh -- a scene as the planner's entry point takes it: the six
// (array, count) pairs of the reference's Environment
// (src/collision/environment.hh:11-40), owning its arrays as the reference's
// does (they are released with delete[]).
//
// The reference's checker never tests `cylinders` (utils.cuh:231-260); its
// harness stores cylinders as capsules (scripts/evaluate_mr.cu:38-47).
// dRRT::solve here keeps both behaviours: `cylinders` are ignored.
#pragma once

#include "shapes.hh"

namespace ppln::collision {

template <typename DataT>
struct Environment {
    Sphere<DataT> *spheres = nullptr;
    unsigned int num_spheres = 0;

    Capsule<DataT> *capsules = nullptr;
    unsigned int num_capsules = 0;

    Capsule<DataT> *z_aligned_capsules = nullptr;
    unsigned int num_z_aligned_capsules = 0;

    Cylinder<DataT> *cylinders = nullptr;
    unsigned int num_cylinders = 0;

    Cuboid<DataT> *cuboids = nullptr;
    unsigned int num_cuboids = 0;

    Cuboid<DataT> *z_aligned_cuboids = nullptr;
    unsigned int num_z_aligned_cuboids = 0;

    Environment() = default;
    Environment(const Environment &) = delete;
    Environment &operator=(const Environment &) = delete;
    ~Environment() {
        delete[] spheres;
        delete[] capsules;
        delete[] z_aligned_capsules;
        delete[] cylinders;
        delete[] cuboids;
        delete[] z_aligned_cuboids;
    }
};

}  // namespace ppln::collision
