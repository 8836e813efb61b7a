// shapes.hh -- obstacle primitives with the field names and constructors of
// the reference (src/collision/shapes.hh), so that scenes built for it build
// unchanged.  min_distance is the distance from the world origin to the
// shape (0 when the origin is inside), computed from the geometry.
//
// Record layouts match the C ABI (include/mrpop_c.h): a cuboid is its centre,
// its three axes and its half extents along them; a cylinder / capsule is
// its first end point, the vector to the second, its radius and
// rdv = 1 / |v|^2.
#pragma once

#include <cmath>
#include <string>

#include "math.hh"

namespace ppln::collision {

template <typename DataT>
struct Shape {
    std::string name;
    DataT min_distance{};

    Shape() = default;
    template <typename OtherDataT>
    explicit Shape(const Shape<OtherDataT> &other) : name(other.name), min_distance(other.min_distance) {}
};

template <typename DataT>
struct Cuboid : public Shape<DataT> {
    DataT x, y, z;
    DataT axis_1_x, axis_1_y, axis_1_z;
    DataT axis_2_x, axis_2_y, axis_2_z;
    DataT axis_3_x, axis_3_y, axis_3_z;
    DataT axis_1_r, axis_2_r, axis_3_r;

    Cuboid() = default;
    explicit Cuboid(DataT x, DataT y, DataT z, DataT axis_1_x, DataT axis_1_y, DataT axis_1_z, DataT axis_2_x,
                    DataT axis_2_y, DataT axis_2_z, DataT axis_3_x, DataT axis_3_y, DataT axis_3_z, DataT axis_1_r,
                    DataT axis_2_r, DataT axis_3_r)
        : x(x), y(y), z(z), axis_1_x(axis_1_x), axis_1_y(axis_1_y), axis_1_z(axis_1_z), axis_2_x(axis_2_x),
          axis_2_y(axis_2_y), axis_2_z(axis_2_z), axis_3_x(axis_3_x), axis_3_y(axis_3_y), axis_3_z(axis_3_z),
          axis_1_r(axis_1_r), axis_2_r(axis_2_r), axis_3_r(axis_3_r) {
        this->min_distance = compute_min_distance();
    }
    template <typename OtherDataT>
    explicit Cuboid(const Cuboid<OtherDataT> &o)
        : Shape<DataT>(o), x(o.x), y(o.y), z(o.z), axis_1_x(o.axis_1_x), axis_1_y(o.axis_1_y),
          axis_1_z(o.axis_1_z), axis_2_x(o.axis_2_x), axis_2_y(o.axis_2_y), axis_2_z(o.axis_2_z),
          axis_3_x(o.axis_3_x), axis_3_y(o.axis_3_y), axis_3_z(o.axis_3_z), axis_1_r(o.axis_1_r),
          axis_2_r(o.axis_2_r), axis_3_r(o.axis_3_r) {}

    // origin to box: the origin's coordinates in the box frame, minus the
    // half extents where they stick out
    DataT compute_min_distance() const {
        const DataT e[3] = {
            std::fabs(dot_3(-x, -y, -z, axis_1_x, axis_1_y, axis_1_z)) - axis_1_r,
            std::fabs(dot_3(-x, -y, -z, axis_2_x, axis_2_y, axis_2_z)) - axis_2_r,
            std::fabs(dot_3(-x, -y, -z, axis_3_x, axis_3_y, axis_3_z)) - axis_3_r,
        };
        DataT s = 0;
        for (DataT v : e) s += v > 0 ? v * v : DataT(0);
        return std::sqrt(s);
    }
};

template <typename DataT>
struct Cylinder : public Shape<DataT> {
    DataT x1, y1, z1;  // first end point
    DataT xv, yv, zv;  // second end point minus the first
    DataT r;
    DataT rdv;         // 1 / (xv^2 + yv^2 + zv^2)

    Cylinder() = default;
    explicit Cylinder(DataT x1, DataT y1, DataT z1, DataT xv, DataT yv, DataT zv, DataT r, DataT rdv)
        : x1(x1), y1(y1), z1(z1), xv(xv), yv(yv), zv(zv), r(r), rdv(rdv) {
        this->min_distance = compute_min_distance();
    }
    template <typename OtherDataT>
    explicit Cylinder(const Cylinder<OtherDataT> &o)
        : Shape<DataT>(o), x1(o.x1), y1(o.y1), z1(o.z1), xv(o.xv), yv(o.yv), zv(o.zv), r(o.r), rdv(o.rdv) {}

    DataT x2() const { return x1 + xv; }
    DataT y2() const { return y1 + yv; }
    DataT z2() const { return z1 + zv; }

    // origin to the segment, less the radius (0 inside)
    DataT compute_min_distance() const {
        const DataT t = clamp(dot_3(-x1, -y1, -z1, xv, yv, zv) * rdv, DataT(0), DataT(1));
        const DataT d = std::sqrt(sql2_3(DataT(0), DataT(0), DataT(0), x1 + xv * t, y1 + yv * t, z1 + zv * t));
        return d > r ? d - r : DataT(0);
    }
};

template <typename DataT>
using Capsule = Cylinder<DataT>;

template <typename DataT>
struct Sphere : public Shape<DataT> {
    DataT x, y, z, r;

    Sphere() = default;
    explicit Sphere(DataT x, DataT y, DataT z, DataT r) : x(x), y(y), z(z), r(r) {
        this->min_distance = std::sqrt(x * x + y * y + z * z) - r;
    }
    template <typename OtherDataT>
    explicit Sphere(const Sphere<OtherDataT> &o) : Shape<DataT>(o), x(o.x), y(o.y), z(o.z), r(o.r) {}
};

}  // namespace ppln::collision
