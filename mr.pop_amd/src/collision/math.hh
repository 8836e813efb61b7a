// math.hh -- small vector helpers of the collision headers (the reference's
// src/collision/math.hh offers the same names).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>

namespace ppln::collision {
using Point = std::array<float, 3>;

template <typename DataT>
inline constexpr auto dot_2(const DataT &ax, const DataT &ay, const DataT &bx, const DataT &by) -> DataT {
    return ax * bx + ay * by;
}

template <typename DataT>
inline constexpr auto dot_3(const DataT &ax, const DataT &ay, const DataT &az, const DataT &bx, const DataT &by,
                            const DataT &bz) -> DataT {
    return ax * bx + ay * by + az * bz;
}

// squared distance of two points
template <typename DataT>
inline constexpr auto sql2_3(const DataT &ax, const DataT &ay, const DataT &az, const DataT &bx, const DataT &by,
                             const DataT &bz) -> DataT {
    return dot_3(ax - bx, ay - by, az - bz, ax - bx, ay - by, az - bz);
}

inline constexpr auto clamp(const float &v, const float &lower, const float &upper) -> float {
    return std::max(std::min(v, upper), lower);
}
}  // namespace ppln::collision
