// internal.h -- host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/mrpop_c.h"
#include "device_geom.h"

namespace mrpop {

// Grow-only device buffer owned by a context.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

// Optional outputs of the fused edge-check kernel.
struct FineQueue {
    unsigned *diag;   // per (sample, arm) phase results, 8 words per sample (diagnostics only)
};

// One arm's roadmap kept for later queries in the same scene: nodes
// [n][8], neighbour lists [n][K], degrees and union-find parents [n], in one
// device allocation.
struct RoadmapEntry {
    int n = 0, K = 0;
    void *mem = nullptr;
    float *q = nullptr;
    int *nbr = nullptr, *deg = nullptr, *uf = nullptr;
    size_t bytes = 0;
    uint64_t used = 0;
};

// Roadmaps keyed by everything their nodes and edges depend on (obstacle
// records, the arm's base pose, granularity, range, neighbour candidates
// and capacities); least recently used entries leave beyond `budget` bytes.
struct RoadmapCache {
    std::map<std::vector<float>, RoadmapEntry> entries;
    size_t bytes = 0;
    size_t budget = (size_t)8 << 30;
    uint64_t clock = 0;
    long hits = 0, misses = 0, stored = 0;
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    std::string err;
    DevBuf counters;
    // staging for host-buffer entry points
    DevBuf from, to, out_edge, out_mask, cfg, out_a, out_b;
    std::vector<DevBuf *> owned;
    void *scratch = nullptr;  // this device's static staging array (device address), see capi.hip
    void *diag = nullptr;  // when set, edge checks record per-sample phase results (diagnostics)
    int probe = 0;         // > 0: edge-only checks run a probe pass of this many samples first (diagnostics)
    RoadmapCache rm_cache; // roadmaps shared by the queries of this context (solve option roadmap_cache)
    // a planner step whose kernels have not completed after this many seconds
    // is taken to be stuck: its problems end with MRPOP_ERR_TIMEOUT
    double lane_deadline_s = 120.0;
    // dRRT expansions use the five-launch form (thread-per-expansion tree
    // scans) once a tree of the launch holds this many nodes; < 0: never
    int expand_split_nodes = 1024;
    int expand_slice_rows = 256;  // least tree rows per block of the five-launch scans

    int fail(int code, const char *fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
};

// Uploaded scene + arm layout.
struct World {
    Context *ctx = nullptr;
    dev::WorldDesc desc{};  // desc.obs points at `buf`
    dev::WorldDesc *d_desc = nullptr;  // device copy of desc
    float *buf = nullptr;
    int n_arms = 0;
    float bases[geom::kMaxArms][12];
    int n_obstacles = 0;
    // algorithmic cost of one coarse env test set: sum of per-obstacle flop weights
    double f_obstacles = 0.0;
};

int ensure(Context *c, DevBuf &b, size_t bytes);

// Lay out a scene and the arms' base poses in one float buffer (obstacle
// records by type, world-axis cuboids as centre + half extents, base poses,
// cull boxes) and fill the offsets of d except d.obs.  Returns the number of
// obstacles, or -1 if a count is negative.
long pack_world(const mrpop_scene &s, const float *bases, int n_arms, dev::WorldDesc &d, std::vector<float> &host);

// Enqueue the coarse + fine edge-check kernels (check.hip).
int launch_check(Context *c, const World *w, int mode, int arm, int granularity, int n_edges, const float *d_from,
                 const float *d_to, uint32_t *d_out_edge, uint64_t *d_out_mask, hipStream_t stream);

// Batched check over edge groups in device memory (check.hip).  Every group
// of one launch uses the same mode and arm count A; d_wave_start holds
// n_groups + 1 entries of workspace.  max_waves bounds the persistent grid
// (an upper bound of the waves the groups can need); zero = false when the
// producers of the edges already cleared the verdict words.  A == 0 in a
// multi-arm mode: groups of 2..5 arms (from each group's dim) in one launch.
// pass: 0 checks every sample; P > 0 checks min(P, G) samples per edge
// spread along it (k = floor(j G / P), j = 1..P); -1 then checks every
// sample of the edges still free.  Probe + completion give the verdicts of
// pass 0 and are for groups without a per-sample mask.
// sparse: the groups' counts usually stay far below max_waves (connection
// and simplification chunks, sized for the worst case); their launch uses a
// persistent grid of resident blocks instead of one block per possible tile.
namespace dev {
struct EdgeGroup;
}
int launch_check_groups(Context *c, int mode, int A, const dev::WorldDesc *d_worlds, const dev::EdgeGroup *d_groups,
                        int n_groups, unsigned *d_wave_start, size_t max_waves, bool sparse, bool zero, int pass,
                        hipStream_t stream);

// Sphere centres kernel (check.hip).
int launch_centres(Context *c, const World *w, int n_cfg, const float *d_cfg, float *d_coarse, float *d_fine,
                   hipStream_t stream);

}  // namespace mrpop

#define MRPOP_HIP(ctx, call)                                                                               \
    do {                                                                                                   \
        hipError_t e_ = (call);                                                                            \
        if (e_ != hipSuccess) return (ctx)->fail(MRPOP_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)
