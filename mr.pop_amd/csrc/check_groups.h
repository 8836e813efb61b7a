// check_groups.h -- work description of a batched edge check.
#pragma once

#include <cstdint>

namespace mrpop {
namespace dev {

// One list of edges checked against one world.  Edge e goes from
// from[e*dim ..] to to[e*dim ..]; sample k = 1..G is from + k*(to-from)/G.
struct EdgeGroup {
    const float *from;
    const float *to;
    int dim;                 // floats per configuration (7 for one arm, 7*A otherwise)
    int G;                   // samples per edge
    int n_edges;             // edge count when count == nullptr
    int arm;                 // one-arm mode: the arm whose base pose is used
    const int *count;        // optional device-side edge count
    uint32_t *out_edge;      // 1 = colliding
    unsigned long long *out_mask;  // optional: bit k-1 per colliding sample
    int world;               // index into the worlds array
    int pad_;
};

// A check tile is W waves holding all arms of 64 W / A samples: W = A for
// the kernel of one arm count (64 samples per tile, one wave per arm),
// kMixedWaves for the kernel that mixes arm counts.  SINGLE mode: one wave.
constexpr int kMixedWaves = 5;
constexpr int tile_samples(int A, bool mixed) { return mixed ? 64 * kMixedWaves / A : 64; }

}  // namespace dev
}  // namespace mrpop
