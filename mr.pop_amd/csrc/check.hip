// check.hip -- launchers of the batched FK + two-phase collision check of
// interpolated edges (the kernels are in check_kernels.h, compiled in the
// objects of check_inst.hip), and the sphere-centres kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "check_kernels.h"

namespace mrpop {
namespace dev {

// ---------------------------------------------------------- centres kernel --
// One thread per (configuration, arm): all coarse and fine centres.
__global__ void centres_kernel(WorldDesc w, int n_cfg, const float *cfg, float *out_coarse, float *out_fine) {
    const int A = w.n_arms;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_cfg * A) return;
    const int c = t / A, arm = t - c * A;
    const float *q = cfg + (size_t)c * 7 * A + 7 * arm;
    Frame T;
    load_base(w, arm, T);
    auto emit = [&](const Frame &F, int joint) {
        if (out_coarse)
            for (int s = kCoarseJointStart[joint]; s < kCoarseJointStart[joint] + kCoarseJointCount[joint]; ++s) {
                float x, y, z;
                centre(F, kCoarseSphere[s][0], kCoarseSphere[s][1], kCoarseSphere[s][2], x, y, z);
                float *o = out_coarse + (((size_t)c * A + arm) * kCoarseSpheres + s) * 3;
                o[0] = x;
                o[1] = y;
                o[2] = z;
            }
        if (out_fine)
            for (int s = kFineJointStart[joint]; s < kFineJointStart[joint] + kFineJointCount[joint]; ++s) {
                float x, y, z;
                centre(F, kTab.off[s][0], kTab.off[s][1], kTab.off[s][2], x, y, z);
                float *o = out_fine + (((size_t)c * A + arm) * kFineSpheres + s) * 3;
                o[0] = x;
                o[1] = y;
                o[2] = z;
            }
    };
    emit(T, 0);
    static_for<1, kJoints>([&](auto jc) {
        constexpr int J = decltype(jc)::value;
        Frame N;
        joint_frame<J>(T, q[J - 1], N);
        T = N;
        emit(T, J);
    });
}

}  // namespace dev
}  // namespace mrpop

// ------------------------------------------------------------- launchers --
namespace mrpop {

namespace {

// Clears the verdict words of every edge before a check (the fused kernel
// only ORs into them, from any wave).
__global__ void zero_outputs_kernel(const dev::EdgeGroup *groups, dev::EdgeGroup g0, int n_groups) {
    for (int gi = blockIdx.y; gi < n_groups; gi += gridDim.y) {
        const dev::EdgeGroup &g = groups ? groups[gi] : g0;
        const int n = g.count ? *g.count : g.n_edges;
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
            g.out_edge[e] = 0u;
            if (g.out_mask) g.out_mask[e] = 0ull;
        }
    }
}

// Waves of each group -> exclusive prefix; wave_start[n] = total.
__global__ void wave_prefix_kernel(const dev::EdgeGroup *groups, int n_groups, int single, int pass, int W,
                                   unsigned *wave_start) {
    __shared__ unsigned sh[256];
    const int tid = threadIdx.x;
    unsigned carry = 0;
    for (int base = 0; base < n_groups; base += 256) {
        const int i = base + tid;
        const unsigned v = i < n_groups ? dev::group_waves(groups[i], single, pass, W) : 0u;
        sh[tid] = v;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const unsigned add = tid >= off ? sh[tid - off] : 0u;
            __syncthreads();
            sh[tid] += add;
            __syncthreads();
        }
        if (i < n_groups) wave_start[i] = carry + sh[tid] - v;
        carry += sh[255];
        __syncthreads();
    }
    if (tid == 0) wave_start[n_groups] = carry;
}

template <int A, int MODE>
void enqueue_check(const dev::WorldDesc *d_worlds, const dev::EdgeGroup *d_groups, const dev::EdgeGroup &g0,
                   int n_groups, unsigned *d_wave_start, const FineQueue &aux, int check_blocks, bool persistent,
                   bool zero, int pass, hipStream_t stream) {
    if (zero)
        hipLaunchKernelGGL(zero_outputs_kernel,
                           dim3(d_groups ? 4 : std::max(1, std::min(1024, (g0.n_edges + 255) / 256)),
                                std::min(n_groups, 1024)),
                           dim3(256), 0, stream, d_groups, g0, n_groups);
    // up to 64 groups the check kernel computes the group prefix itself
    if (d_groups && n_groups > 64)
        hipLaunchKernelGGL(wave_prefix_kernel, dim3(1), dim3(256), 0, stream, d_groups, n_groups,
                           MODE == MRPOP_MODE_SINGLE ? 1 : 0, pass, A, d_wave_start);
    if (n_groups <= 64) d_wave_start = nullptr;
    dev::launch_check_kernel<A, MODE>(check_blocks, persistent, stream, d_worlds, d_groups, g0, n_groups, d_wave_start,
                                      aux, pass);
}

// Groups of several arm counts (multi-arm modes) in one launch.
template <int MODE>
void enqueue_mixed(const dev::WorldDesc *d_worlds, const dev::EdgeGroup *d_groups, int n_groups,
                   unsigned *d_wave_start, int check_blocks, bool persistent, bool zero, int pass,
                   hipStream_t stream) {
    const dev::EdgeGroup none{};
    if (zero)
        hipLaunchKernelGGL(zero_outputs_kernel, dim3(4, std::min(n_groups, 1024)), dim3(256), 0, stream, d_groups,
                           none, n_groups);
    if (n_groups > 64)
        hipLaunchKernelGGL(wave_prefix_kernel, dim3(1), dim3(256), 0, stream, d_groups, n_groups, 0, pass,
                           dev::kMixedWaves, d_wave_start);
    dev::launch_check_kernel_mixed<MODE>(check_blocks, persistent, stream, d_worlds, d_groups, n_groups,
                                         n_groups > 64 ? d_wave_start : nullptr, pass);
}

// Blocks of W waves that stay resident at three waves per SIMD: a
// persistent grid needs no more (idle blocks only cost dispatch).
size_t resident_blocks(const Context *c, int W) {
    return (size_t)c->num_cus * std::max(1, 12 / W);
}

// Grouped launches run one tile per block over their bound unless the
// caller marks them sparse or the bound exceeds this many resident grids
// (then a persistent grid of resident blocks loops over the tiles), so that
// blocks without a tile stay few.
constexpr size_t kOneTileGrids = 64;

int dispatch(Context *c, int mode, int A, const dev::WorldDesc *d_worlds, const dev::EdgeGroup *d_groups,
             const dev::EdgeGroup &g0, int n_groups, unsigned *d_wave_start, const FineQueue &aux, int check_blocks,
             bool persistent, bool zero, int pass, hipStream_t stream) {
#define MRPOP_CASE(M, AA)                                                                                         \
    case M * 8 + AA:                                                                                              \
        enqueue_check<AA, M>(d_worlds, d_groups, g0, n_groups, d_wave_start, aux, check_blocks, persistent, zero, pass, \
                             stream);                                                                             \
        break;
    switch (mode * 8 + A) {
        MRPOP_CASE(MRPOP_MODE_SINGLE, 1)
        MRPOP_CASE(MRPOP_MODE_FULL, 2)
        MRPOP_CASE(MRPOP_MODE_FULL, 3)
        MRPOP_CASE(MRPOP_MODE_FULL, 4)
        MRPOP_CASE(MRPOP_MODE_FULL, 5)
        MRPOP_CASE(MRPOP_MODE_DRRT, 2)
        MRPOP_CASE(MRPOP_MODE_DRRT, 3)
        MRPOP_CASE(MRPOP_MODE_DRRT, 4)
        MRPOP_CASE(MRPOP_MODE_DRRT, 5)
        case MRPOP_MODE_FULL * 8 + 0:  // groups of several arm counts
            if (!d_groups) return c->fail(MRPOP_ERR_INVALID, "a mixed-arm check needs edge groups");
            enqueue_mixed<MRPOP_MODE_FULL>(d_worlds, d_groups, n_groups, d_wave_start, check_blocks, persistent, zero,
                                           pass, stream);
            break;
        case MRPOP_MODE_DRRT * 8 + 0:
            if (!d_groups) return c->fail(MRPOP_ERR_INVALID, "a mixed-arm check needs edge groups");
            enqueue_mixed<MRPOP_MODE_DRRT>(d_worlds, d_groups, n_groups, d_wave_start, check_blocks, persistent, zero,
                                           pass, stream);
            break;
        default: return c->fail(MRPOP_ERR_INVALID, "unsupported mode %d with %d arms", mode, A);
    }
#undef MRPOP_CASE
    MRPOP_HIP(c, hipGetLastError());
    return MRPOP_OK;
}

}  // namespace

int launch_check(Context *c, const World *w, int mode, int arm, int granularity, int n_edges, const float *d_from,
                 const float *d_to, uint32_t *d_out_edge, uint64_t *d_out_mask, hipStream_t stream) {
    if (n_edges <= 0) return MRPOP_OK;
    if (granularity < 1 || granularity > 64) return c->fail(MRPOP_ERR_INVALID, "granularity %d outside 1..64", granularity);
    const int A = mode == MRPOP_MODE_SINGLE ? 1 : w->n_arms;
    if (mode == MRPOP_MODE_SINGLE && (arm < 0 || arm >= w->n_arms))
        return c->fail(MRPOP_ERR_INVALID, "arm %d outside the world's %d arms", arm, w->n_arms);
    if (mode != MRPOP_MODE_SINGLE && (A < 2 || A > 5))
        return c->fail(MRPOP_ERR_INVALID, "multi-arm modes need 2..5 arms (world has %d)", A);
    const size_t total = (size_t)n_edges * granularity;
    if (total > 0x7fffffffull) return c->fail(MRPOP_ERR_INVALID, "too many samples in one batch");
    const FineQueue aux{(unsigned *)c->diag};
    dev::EdgeGroup g0{d_from, d_to, mode == MRPOP_MODE_SINGLE ? 7 : 7 * A, granularity, n_edges, arm, nullptr,
                      d_out_edge, (unsigned long long *)d_out_mask, 0, 0};
    const int S = dev::tile_samples(A, false);
    const size_t waves = (total + S - 1) / S;  // tiles
    const int check_blocks = (int)waves;  // one tile per block
    const int probe = d_out_mask || c->diag ? 0 : std::min(c->probe, granularity);
    if (probe > 0 && probe < granularity) {
        const size_t pw = ((size_t)n_edges * probe + S - 1) / S;
        if (int rc = dispatch(c, mode, A, w->d_desc, nullptr, g0, 1, nullptr, aux, (int)pw, false, true, probe, stream))
            return rc;
        return dispatch(c, mode, A, w->d_desc, nullptr, g0, 1, nullptr, aux, check_blocks, false, false, -1, stream);
    }
    return dispatch(c, mode, A, w->d_desc, nullptr, g0, 1, nullptr, aux, check_blocks, false, true, 0, stream);
}

int launch_check_groups(Context *c, int mode, int A, const dev::WorldDesc *d_worlds, const dev::EdgeGroup *d_groups,
                        int n_groups, unsigned *d_wave_start, size_t max_waves, bool sparse, bool zero, int pass,
                        hipStream_t stream) {
    if (n_groups <= 0) return MRPOP_OK;
    if (pass < -1) return c->fail(MRPOP_ERR_INVALID, "check pass %d", pass);
    const FineQueue aux{nullptr};
    dev::EdgeGroup g0{};
    const int W = mode == MRPOP_MODE_SINGLE ? 1 : (A == 0 ? dev::kMixedWaves : A);
    const size_t resident = resident_blocks(c, W);
    const bool persistent = sparse || max_waves > kOneTileGrids * resident;
    const int check_blocks = (int)std::max<size_t>(1, persistent ? resident : max_waves);
    return dispatch(c, mode, A, d_worlds, d_groups, g0, n_groups, d_wave_start, aux, check_blocks, persistent, zero,
                    pass, stream);
}

int launch_centres(Context *c, const World *w, int n_cfg, const float *d_cfg, float *d_coarse, float *d_fine,
                   hipStream_t stream) {
    if (n_cfg <= 0) return MRPOP_OK;
    const int threads = n_cfg * w->n_arms;
    hipLaunchKernelGGL(dev::centres_kernel, dim3((threads + 63) / 64), dim3(64), 0, stream, w->desc, n_cfg, d_cfg,
                       d_coarse, d_fine);
    MRPOP_HIP(c, hipGetLastError());
    return MRPOP_OK;
}

}  // namespace mrpop
