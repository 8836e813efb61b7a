// factory.hh -- constructors of obstacle primitives with the names and
// argument order of the reference's factory (src/collision/factory.hh):
// cuboid::{flat, array}, cylinder::{endpoints, center}::{flat, array},
// capsule::{endpoints, center}::{flat, array}, sphere::{flat, array}.
// Orientations are Euler XYZ angles (rho about x, theta about y, phi about
// z) composed as R = Rz(phi) * Ry(theta) * Rx(rho); a cuboid's axes are the
// columns of R.  The reference's Eigen overloads (eigen, eigen_rot) are not
// provided: Eigen is not a dependency of this library.
#pragma once

#include <array>
#include <cmath>

#include "shapes.hh"

namespace ppln::collision::factory {

using ConstArrayRef = const std::array<float, 3> &;

namespace detail {
// R = Rz(phi) Ry(theta) Rx(rho), row-major
inline std::array<float, 9> euler_xyz(float rho, float theta, float phi) {
    const double cx = std::cos(rho), sx = std::sin(rho);
    const double cy = std::cos(theta), sy = std::sin(theta);
    const double cz = std::cos(phi), sz = std::sin(phi);
    return {(float)(cz * cy), (float)(cz * sy * sx - sz * cx), (float)(cz * sy * cx + sz * sx),
            (float)(sz * cy), (float)(sz * sy * sx + cz * cx), (float)(sz * sy * cx - cz * sx),
            (float)(-sy),     (float)(cy * sx),                (float)(cy * cx)};
}
}  // namespace detail

namespace cuboid {
// centre, Euler XYZ orientation about the centre, half extents along the
// cuboid's own x, y, z axes
inline auto flat(float center_x, float center_y, float center_z, float rho, float theta, float phi,
                 float half_extent_x, float half_extent_y, float half_extent_z) noexcept -> Cuboid<float> {
    const auto R = detail::euler_xyz(rho, theta, phi);
    return Cuboid<float>(center_x, center_y, center_z, R[0], R[3], R[6], R[1], R[4], R[7], R[2], R[5], R[8],
                         half_extent_x, half_extent_y, half_extent_z);
}
inline auto array(ConstArrayRef center, ConstArrayRef euler_xyz, ConstArrayRef half_extents) noexcept
    -> Cuboid<float> {
    return flat(center[0], center[1], center[2], euler_xyz[0], euler_xyz[1], euler_xyz[2], half_extents[0],
                half_extents[1], half_extents[2]);
}
}  // namespace cuboid

namespace detail {
template <typename T>
inline T segment(float x1, float y1, float z1, float x2, float y2, float z2, float radius) {
    const float xv = x2 - x1, yv = y2 - y1, zv = z2 - z1;
    const float dot = dot_3(xv, yv, zv, xv, yv, zv);
    return T(x1, y1, z1, xv, yv, zv, radius, 1.0f / dot);
}
// end points at +-length/2 along the local z axis of the oriented centre
template <typename T>
inline T centred(float cx, float cy, float cz, float rho, float theta, float phi, float radius, float length) {
    const auto R = euler_xyz(rho, theta, phi);
    const float h = length / 2;
    return segment<T>(cx + R[2] * h, cy + R[5] * h, cz + R[8] * h, cx - R[2] * h, cy - R[5] * h, cz - R[8] * h,
                      radius);
}
}  // namespace detail

namespace cylinder {
namespace endpoints {
inline auto flat(float x1, float y1, float z1, float x2, float y2, float z2, float radius) noexcept
    -> Cylinder<float> {
    return detail::segment<Cylinder<float>>(x1, y1, z1, x2, y2, z2, radius);
}
inline auto array(ConstArrayRef endpoint1, ConstArrayRef endpoint2, float radius) noexcept -> Cylinder<float> {
    return flat(endpoint1[0], endpoint1[1], endpoint1[2], endpoint2[0], endpoint2[1], endpoint2[2], radius);
}
}  // namespace endpoints
namespace center {
inline auto flat(float center_x, float center_y, float center_z, float rho, float theta, float phi, float radius,
                 float length) noexcept -> Cylinder<float> {
    return detail::centred<Cylinder<float>>(center_x, center_y, center_z, rho, theta, phi, radius, length);
}
inline auto array(ConstArrayRef center, ConstArrayRef euler_xyz, float radius, float length) noexcept
    -> Cylinder<float> {
    return flat(center[0], center[1], center[2], euler_xyz[0], euler_xyz[1], euler_xyz[2], radius, length);
}
}  // namespace center
}  // namespace cylinder

namespace capsule {
namespace endpoints {
inline auto flat(float x1, float y1, float z1, float x2, float y2, float z2, float radius) noexcept
    -> Capsule<float> {
    return detail::segment<Capsule<float>>(x1, y1, z1, x2, y2, z2, radius);
}
inline auto array(ConstArrayRef endpoint1, ConstArrayRef endpoint2, float radius) noexcept -> Capsule<float> {
    return flat(endpoint1[0], endpoint1[1], endpoint1[2], endpoint2[0], endpoint2[1], endpoint2[2], radius);
}
}  // namespace endpoints
namespace center {
inline auto flat(float center_x, float center_y, float center_z, float rho, float theta, float phi, float radius,
                 float length) noexcept -> Capsule<float> {
    return detail::centred<Capsule<float>>(center_x, center_y, center_z, rho, theta, phi, radius, length);
}
inline auto array(ConstArrayRef center, ConstArrayRef euler_xyz, float radius, float length) noexcept
    -> Capsule<float> {
    return flat(center[0], center[1], center[2], euler_xyz[0], euler_xyz[1], euler_xyz[2], radius, length);
}
}  // namespace center
}  // namespace capsule

namespace sphere {
inline auto flat(float center_x, float center_y, float center_z, float radius) noexcept -> Sphere<float> {
    return Sphere<float>(center_x, center_y, center_z, radius);
}
inline auto array(ConstArrayRef center, float radius) noexcept -> Sphere<float> {
    return flat(center[0], center[1], center[2], radius);
}
}  // namespace sphere

}  // namespace ppln::collision::factory
