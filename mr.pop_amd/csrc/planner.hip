// planner.hip -- round kernels of the dRRT / AO-x planner on gfx950.
//
// Each kernel handles every active slot (problem) of the engine at once:
// blockIdx.y indexes the slot list.  The reference semantics followed are
// cited per kernel; the reference's inter-block races (first CAS wins,
// unsynchronised tree growth) are replaced by round-synchronous steps whose
// results are committed in expansion order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>

#include "planner.h"

namespace mrpop {
namespace plan {

using namespace mrpop::geom;

__device__ __forceinline__ void wave_argmin(float &d, int &i) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float od = __shfl_xor(d, off);
        const int oi = __shfl_xor(i, off);
        if (od < d || (od == d && oi < i)) {
            d = od;
            i = oi;
        }
    }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ int wave_min_int(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    return v;
}

// Squared distance of two D-vectors, accumulated in a fixed order.
__device__ __forceinline__ float sqdist(const float *a, const float *b, int D) {
    float s = 0.f;
    for (int d = 0; d < D; ++d) {
        const float x = a[d] - b[d];
        s = __builtin_fmaf(x, x, s);
    }
    return s;
}

// Copies a row of D <= 35 floats with every load issued before the first
// store: a plain copy loop waits for each element's load before its store
// (the compiler cannot tell that the rows do not overlap), one memory round
// trip per element.
__device__ __forceinline__ void copy_row(float *dst, const float *src, int D) {
    float v[35];
#pragma unroll
    for (int d = 0; d < 35; ++d)
        if (d < D) v[d] = src[d];
#pragma unroll
    for (int d = 0; d < 35; ++d)
        if (d < D) dst[d] = v[d];
}

__device__ __forceinline__ float *tree_q(const SlotDev &s, int t, int i) {
    return s.tr_q + ((size_t)t * s.tr_cap + i) * s.D;
}

// Tree scans read 64 rows of D floats at a time: the wave copies the chunk
// rows[base .. base + c) (c * D consecutive floats, so the loads coalesce)
// into LDS, and lane i then reads row base + i from there -- the rows lanes
// would otherwise load themselves lie D floats apart, one cache line per
// lane and load.  The chunk never reaches past row base + c - 1 < n <= the
// tree's capacity.  Lane i keeps handling rows i, i + 64, ..., as a
// per-lane strided scan does, so the results (and ties) do not change.
// Staged rows lie S >= D floats apart.  Where lane i reads row i, S is odd
// (wave_stride): with an even stride (D = 14 or 28) the lanes' reads of one
// column would fall on 2 or 4 times fewer LDS banks.  Where every lane reads
// the same row (a broadcast), S = D.  D | 1 <= 35 for D <= 35, so a chunk
// fits the kScanRows * 35 floats of each stage buffer.
constexpr int kScanRows = 64;
__device__ __forceinline__ int wave_stride(int D) { return D | 1; }
__device__ __forceinline__ void stage_rows(const float *rows, int base, int c, int D, int S, float *stage, int lane) {
    // c <= kScanRows rows of D <= 35 floats: at most kPerLane floats a lane.
    // Every load of the chunk is issued before the first store, so a chunk
    // waits for one memory latency, not one per 64 floats.
    constexpr int kPerLane = kScanRows * 35 / 64;
    const float *src = rows + (size_t)base * D;
    const int nf = c * D;
    float v[kPerLane];
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) {
        const int x = lane + 64 * k;
        if (x < nf) v[k] = src[x];
    }
    if (S == D) {
#pragma unroll
        for (int k = 0; k < kPerLane; ++k) {
            const int x = lane + 64 * k;
            if (x < nf) stage[x] = v[k];
        }
        return;
    }
    // element x = r * D + col goes to r * S + col; x advances by 64 = dr * D + dc
    const int dr = 64 / D, dc = 64 - dr * D;
    int r = lane / D, col = lane - r * D;
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) {
        if (lane + 64 * k < nf) stage[r * S + col] = v[k];
        r += dr;
        col += dc;
        if (col >= D) {
            col -= D;
            ++r;
        }
    }
}

// A dRRT round runs for a slot while its pass is unsolved and the launch has
// rounds left (pop_planner.cu:360-364).
__device__ __forceinline__ bool searching(const SlotDev &s) {
    return *s.status == kRunning && s.ctr[kCtrRound] < s.ctr[kCtrLimit];
}

// ============================================================ roadmaps ====
// Roadmap of each arm: node 0 = the arm's start slice, nodes 1..n_goals its
// goal slices (roadmap_interior.cuh:69-88).
//
// With roadmap reuse (rm_insert) the start and goals are not written here:
// the first round inserts them after the nodes copied from the cache.
// (The host then copies any cached roadmap in and sets its node count.)
__global__ void k_roadmap_init(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    if (threadIdx.x < s.A) {
        s.rm_root0[threadIdx.x] = 0;
        s.rm_n[threadIdx.x] = s.rm_insert ? 0 : 1 + s.n_goals;
        s.rm_connected[threadIdx.x] = 0;
        s.rm_qvalid[threadIdx.x] = 0;
        s.re_count[threadIdx.x] = 0;
        s.rs_brounds[threadIdx.x] = 0;
    }
    if (s.rm_insert) return;
    const float *sg = s.sg;
    for (int t = threadIdx.x; t < s.A * (1 + s.n_goals) * 7; t += blockDim.x) {
        const int a = t / ((1 + s.n_goals) * 7);
        const int rem = t - a * (1 + s.n_goals) * 7;
        const int node = rem / 7, j = rem - node * 7;
        s.rm_q[((size_t)a * s.rm_cap + node) * 8 + j] = sg[(size_t)node * s.D + 7 * a + j];
    }
    for (int t = threadIdx.x; t < s.A * (1 + s.n_goals); t += blockDim.x) {
        const int a = t / (1 + s.n_goals), node = t - a * (1 + s.n_goals);
        s.rm_uf[(size_t)a * s.rm_cap + node] = node;
        s.rm_deg[(size_t)a * s.rm_cap + node] = 0;
    }
}

// M uniform samples per arm in the joint limits (roadmap_interior.cuh:105-112).
// They are validated as G = 1 edges q -> q by the edge checker.
//
// Biased steering (roadmap_interior.cuh:117-165): while the first roadmap
// rounds of the solve run, a reference block's iteration counter counts its
// samples, and every BIASED_ITER-th sample is pulled to within `range` of
// its nearest roadmap node, q <- n + (q - n) * min(1, range / |q - n|).
// Sample i of a round is iteration i mod 230 of block i / 230, so its
// counter is 230 * (earlier rounds of the stretch the arm sampled in) +
// (i mod 230) + 1.  The nearest node (squared distance accumulated with one
// fused multiply-add per joint, lowest id among equals) is searched by the
// whole wave over the nodes the arm had at the round's start.
__global__ __launch_bounds__(256) void k_rm_sample(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool live = t < s.A * s.M;
    const int a = live ? t / s.M : 0, i = live ? t - a * s.M : 0;
    // roadmap_reassign (roadmap_interior.cuh:316-327): once an arm's start and
    // goal are connected, its share of the round goes to the arms still
    // unconnected -- here the arm draws no samples and its checks drop out
    const bool skip = live && s.rm_reassign && s.rm_connected[a];
    const bool draw = live && !skip;
    if (s.rm_insert) {
        // insertion round: sample k <= n_goals is the query's node k (start,
        // then goals); the checker validates only those
        if (!live) return;
        if (i == 0) s.rs_nsamp[a] = 1 + s.n_goals;
        if (i <= s.n_goals) {
            float *dst = s.rs_q + ((size_t)a * s.M + i) * 7;
            for (int j = 0; j < 7; ++j) dst[j] = s.sg[(size_t)i * s.D + 7 * a + j];
        }
        s.rs_coll[(size_t)a * s.M + i] = i <= s.n_goals ? 0u : 1u;
        return;
    }
    if (live && i == 0) s.rs_nsamp[a] = skip ? 0 : s.M;
    float q[7] = {};
    if (draw) {
        const uint4 r0 = slot_random(s.seed, kStreamRoadmap, (unsigned)(s.rm_round * kMaxArms + a), (unsigned)i, 0);
        const uint4 r1 = slot_random(s.seed, kStreamRoadmap, (unsigned)(s.rm_round * kMaxArms + a), (unsigned)i, 1);
        const unsigned bits[7] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z};
#pragma unroll
        for (int j = 0; j < 7; ++j) q[j] = __builtin_fmaf(u01(bits[j]), kScale[j], kLow[j]);
    }
    bool biased = false;
    if (draw && s.rm_stretch && s.rm_biased_iter > 0) {
        const long long it = (long long)s.rs_brounds[a] * kRoadmapIter + (i % kRoadmapIter) + 1;
        biased = it % s.rm_biased_iter == 0;
    }
    // every lane of the wave reaches this point (no early exit above)
    unsigned long long pend = __ballot(biased);
    while (pend) {
        const int src = __ffsll((long long)pend) - 1;
        pend &= pend - 1;
        const int sa = __shfl(a, src);
        float sq[7];
#pragma unroll
        for (int j = 0; j < 7; ++j) sq[j] = __shfl(q[j], src);
        const float *rq = s.rm_q + (size_t)sa * s.rm_cap * 8;
        const int n = s.rm_n[sa];
        float bd = FLT_MAX;
        int bi = 0x7fffffff;
        for (int r = lane; r < n; r += 64) {
            const float *p = rq + (size_t)r * 8;
            float d = 0.f;
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const float x = p[j] - sq[j];
                d = __builtin_fmaf(x, x, d);
            }
            if (d < bd) {
                bd = d;
                bi = r;
            }
        }
        wave_argmin(bd, bi);
        if (lane == src) {
            const float *p = rq + (size_t)bi * 8;
            const float sc = fminf(1.0f, s.range / sqrtf(bd));  // a node at the sample itself: 1
#pragma unroll
            for (int j = 0; j < 7; ++j) q[j] = __builtin_fmaf(q[j] - p[j], sc, p[j]);
        }
    }
    if (!live) return;
    if (!draw) {
        s.rs_coll[(size_t)a * s.M + i] = 1u;  // no new nodes
        return;
    }
    float *dst = s.rs_q + ((size_t)a * s.M + i) * 7;
#pragma unroll
    for (int j = 0; j < 7; ++j) dst[j] = q[j];
    s.rs_coll[(size_t)a * s.M + i] = 0u;  // the checker only sets bits
}

// Valid samples become nodes n, n+1, ... in sample order.  One block per (slot, arm).
__global__ void k_rm_assign(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int a = blockIdx.x;
    if (a >= s.A) return;
    __shared__ int sh[256];
    __shared__ int base;
    if (threadIdx.x == 0) {
        base = s.rm_n[a];
        s.re_count[a] = 0;
        s.rs_base[a] = base;     // first id of this round's new nodes
        s.rl_ntouched[a] = 0;
        if (s.rm_stretch && s.rs_nsamp[a] > 0 && !s.rm_insert) s.rs_brounds[a] += 1;  // the round is over for k_rm_sample
        if (s.rm_insert) {
            // the query's nodes go in whatever the checker found (a query
            // assumes a free start and goals); whether they are free decides
            // whether the roadmap may be shared later
            int ok = 1;
            for (int k = 0; k <= s.n_goals; ++k) ok &= s.rs_coll[(size_t)a * s.M + k] == 0u;
            s.rm_qvalid[a] = ok;
            s.rm_root0[a] = base;
        }
    }
    __syncthreads();
    const int limit = s.rm_insert ? s.rm_cap : s.rm_fill;
    const int per = (s.M + 255) / 256;
    const int i0 = threadIdx.x * per;
    int cnt = 0;
    for (int i = i0; i < min(i0 + per, s.M); ++i)
        cnt += s.rs_coll[(size_t)a * s.M + i] == 0u || (s.rm_insert && i <= s.n_goals);
    sh[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    int id = base + sh[threadIdx.x] - cnt;
    for (int i = i0; i < min(i0 + per, s.M); ++i) {
        int out = -1;
        if ((s.rs_coll[(size_t)a * s.M + i] == 0u || (s.rm_insert && i <= s.n_goals)) && id < limit) {
            out = id++;
            const float *q = s.rs_q + ((size_t)a * s.M + i) * 7;
            float *dst = s.rm_q + ((size_t)a * s.rm_cap + out) * 8;
            for (int j = 0; j < 7; ++j) dst[j] = q[j];
            s.rm_uf[(size_t)a * s.rm_cap + out] = out;
            s.rm_deg[(size_t)a * s.rm_cap + out] = 0;
        }
        s.rs_id[(size_t)a * s.M + i] = out;
    }
    __syncthreads();
    if (threadIdx.x == 255) s.rm_n[a] = max(base, min(base + sh[255], limit));
}

// Radius neighbours of each new node among all nodes with a smaller id
// (existing ones and earlier samples of this round), first rm_Kc in id order
// (roadmap_interior.cuh:183-197).  Candidates are staged at fixed rows
// sample*Kc + k; k_rm_compact packs them in sample order.  One wave per
// (slot, arm, sample).
//
// The query's start and goal nodes come first, then the other nodes in id
// order.  In a roadmap built for the query they are nodes 0..n_goals, so
// this is plain id order; in a roadmap taken from the cache they are its
// newest nodes, and every new node still tries them first, as it would in a
// roadmap of its own.
__device__ __forceinline__ void rm_radius_one(SlotDev &s, const int t) {
    const int a = t / s.M, i = t - a * s.M;
    const int id = s.rs_id[(size_t)a * s.M + i];
    const int lane = threadIdx.x;
    int *cnt = s.rs_ncand + (size_t)a * s.M;
    if (id < 0) {
        if (lane == 0) cnt[i] = 0;
        return;
    }
    const float *rq = s.rm_q + (size_t)a * s.rm_cap * 8;
    const float *qn = rq + (size_t)id * 8;
    float q[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) q[j] = qn[j];
    const float r2 = s.range * s.range;
    int *stage = s.rs_cand + ((size_t)a * s.M + i) * s.rm_Kc;
    int found = 0;
    const int root0 = s.rm_root0[a], root1 = root0 + 1 + s.n_goals;  // the query's nodes [root0, root1)
    if (root0 > 0) {
        const int n = root0 + lane;
        bool near = false;
        if (n < root1 && n < id) {
            const float *p = rq + (size_t)n * 8;
            float d = 0.f;
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const float x = p[j] - q[j];
                d = __builtin_fmaf(x, x, d);
            }
            near = d < r2 && d > 0.f;
        }
        const unsigned long long bal = __ballot(near);
        if (near) {
            const int pos = __popcll(bal & ((1ull << lane) - 1ull));
            if (pos < s.rm_Kc) stage[pos] = n;
        }
        found = __popcll(bal);
    }
    auto sqd = [&](int n) {
        const float *p = rq + (size_t)n * 8;
        float d = 0.f;
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            const float x = p[j] - q[j];
            d = __builtin_fmaf(x, x, d);
        }
        return d;
    };
    if (s.rm_insert && root0 > 0) {
        // A query's node inserted into a roadmap from the cache: of the other
        // nodes within range it takes the nearest (by squared distance, then
        // id), not the oldest -- it gets no second chance as the roadmap
        // grows around it the way a start node built with the roadmap does.
        // The (rem)-th smallest distance is found by a radix select on the
        // float bits (non-negative floats order as their bit patterns).
        const int rem = s.rm_Kc - found;
        int in_range = 0;
        for (int n = lane; n < id; n += 64)
            if (n < root0 || n >= root1) {
                const float d = sqd(n);
                in_range += d < r2 && d > 0.f;
            }
        for (int off = 32; off > 0; off >>= 1) in_range += __shfl_xor(in_range, off);
        unsigned thr = 0xffffffffu;  // take keys below thr, and the first `eq` keys equal to it
        int eq = 0;
        if (rem > 0 && in_range > rem) {
            __shared__ int hist[256];
            unsigned prefix = 0, mask = 0;
            int k = rem;  // rank (1-based) of the wanted key among those matching the prefix
            for (int shift = 24; shift >= 0; shift -= 8) {
                for (int b = lane; b < 256; b += 64) hist[b] = 0;
                __syncthreads();
                for (int n = lane; n < id; n += 64) {
                    if (n >= root0 && n < root1) continue;
                    const float d = sqd(n);
                    const unsigned key = __float_as_uint(d);
                    if (d < r2 && d > 0.f && (key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1);
                }
                __syncthreads();
                int digit = 0, below = 0;
                if (lane == 0) {
                    while (below + hist[digit] < k) below += hist[digit++];
                }
                digit = __shfl(digit, 0);
                below = __shfl(below, 0);
                k -= below;
                prefix |= (unsigned)digit << shift;
                mask |= 255u << shift;
                __syncthreads();
            }
            thr = prefix;
            eq = k;
        }
        int taken_eq = 0;
        for (int base = 0; base < id && found < s.rm_Kc; base += 64) {
            const int n = base + lane;
            bool lt = false, is_eq = false;
            if (n < id && (n < root0 || n >= root1)) {
                const float d = sqd(n);
                const unsigned key = __float_as_uint(d);
                lt = d < r2 && d > 0.f && key < thr;
                is_eq = d < r2 && d > 0.f && key == thr;
            }
            const unsigned long long beq = __ballot(is_eq);
            const int eq_pos = taken_eq + __popcll(beq & ((1ull << lane) - 1ull));
            const bool take = lt || (is_eq && eq_pos < eq);
            const unsigned long long bal = __ballot(take);
            if (take) {
                const int pos = found + __popcll(bal & ((1ull << lane) - 1ull));
                if (pos < s.rm_Kc) stage[pos] = n;
            }
            found += __popcll(bal);
            taken_eq += __popcll(beq);
        }
        if (lane == 0) cnt[i] = min(found, s.rm_Kc);
        return;
    }
    for (int base = 0; base < id && found < s.rm_Kc; base += 64) {
        const int n = base + lane;
        bool near = false;
        if (n < id && (root0 == 0 || n < root0 || n >= root1)) {
            const float d = sqd(n);
            near = d < r2 && d > 0.f;  // l2 distance < range; a node repeating v is no neighbour
        }
        const unsigned long long bal = __ballot(near);
        if (near) {
            const int pos = found + __popcll(bal & ((1ull << lane) - 1ull));
            if (pos < s.rm_Kc) stage[pos] = n;
        }
        found += __popcll(bal);
    }
    if (lane == 0) cnt[i] = min(found, s.rm_Kc);
}
__global__ __launch_bounds__(64) void k_rm_radius(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    for (int t = blockIdx.x; t < s.A * s.M; t += gridDim.x) rm_radius_one(s, t);  // a wave per item, striding
}

// First packed edge of each sample: the staged candidates of an arm form
// one edge list neighbour -> new node in sample order (roadmap_interior.cuh:
// 214-234); k_rm_pack writes it.  One block per (slot, arm).
__global__ __launch_bounds__(1024) void k_rm_compact(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int a = blockIdx.x;
    if (a >= s.A) return;
    __shared__ int sh[1024];
    const int tid = threadIdx.x;
    const int per = (s.M + 1023) / 1024;
    const int i0 = tid * per;
    const int *cnt = s.rs_ncand + (size_t)a * s.M;
    int c = 0;
    for (int i = i0; i < min(i0 + per, s.M); ++i) c += cnt[i];
    sh[tid] = c;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int add = tid >= off ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += add;
        __syncthreads();
    }
    int e = sh[tid] - c;
    for (int i = i0; i < min(i0 + per, s.M); ++i) {
        s.rs_first[(size_t)a * s.M + i] = e;
        e += cnt[i];
    }
    if (tid == 1023) s.re_count[a] = sh[1023];
}

// Write the packed edges of one sample: lane k is candidate k.  One wave per
// (slot, arm, sample).
__device__ __forceinline__ void rm_pack_one(SlotDev &s, const int t) {
    const int a = t / s.M, i = t - a * s.M;
    const int k = threadIdx.x;
    if (k >= s.rs_ncand[(size_t)a * s.M + i]) return;
    const float *rq = s.rm_q + (size_t)a * s.rm_cap * 8;
    const size_t e = (size_t)a * s.M * s.rm_Kc + s.rs_first[(size_t)a * s.M + i] + k;
    const int id = s.rs_id[(size_t)a * s.M + i];
    const int n = s.rs_cand[((size_t)a * s.M + i) * s.rm_Kc + k];
    const float *p = rq + (size_t)n * 8, *q = rq + (size_t)id * 8;
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        s.re_from[e * 7 + j] = p[j];
        s.re_to[e * 7 + j] = q[j];
    }
    s.re_pair[e * 2 + 0] = n;
    s.re_pair[e * 2 + 1] = id;
    s.re_coll[e] = 0u;  // the checker only sets bits
}
__global__ __launch_bounds__(64) void k_rm_pack(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    for (int t = blockIdx.x; t < s.A * s.M; t += gridDim.x) rm_pack_one(s, t);  // a wave per item, striding
}

// Parents always have smaller ids than their children, so the walk ends
// within the node count; the step cap only guards against corrupted input
// and is counted as an event when reached.
__device__ __forceinline__ int uf_find(int *uf, int x, int *events) {
    for (int steps = 0; steps < (1 << 24); ++steps) {
        const int p = uf[x];
        if (p == x) return x;
        const int gp = uf[p];
        if (gp != p) atomicCAS(&uf[x], p, gp);  // path halving
        x = p;
    }
    atomicAdd(&events[kEvUnionFind], 1);
    return x;
}

// Lock-free union: the larger root is hooked under the smaller one, so each
// component's root is its smallest node and the partition is deterministic.
// A failed CAS means another union of the same root succeeded, so the
// retries are bounded by the round's edges; the cap is counted if reached.
__device__ __forceinline__ void uf_union(int *uf, int a, int b, int *events) {
    for (int tries = 0; tries < (1 << 20); ++tries) {
        a = uf_find(uf, a, events);
        b = uf_find(uf, b, events);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&uf[a], a, b) == a) return;
    }
    atomicAdd(&events[kEvUnionFind], 1);
}

// Collision-free candidate edges join both adjacency lists (the reference's
// symmetric bitset, roadmap_interior.cuh:245-258) and the union-find.
//
// Lists are filled in id order, so their contents do not depend on the order
// in which threads run: a new node v first lists its older neighbours (its
// own candidates, already in id order), then newer nodes as later rounds
// (or later samples of this round) link to it.  An older node u collects the
// new nodes of this round in a bitset over the round's id range
// [rs_base, rs_base + M); k_rm_link_append then appends them in ascending
// order.  Three launches:
//   k_rm_link_own    one wave per sample: v's own list, union-find, and a
//                    bitset claimed for every older endpoint u;
//   k_rm_link_bits   one wave per sample: set bit (v - base) in u's bitset;
//   k_rm_link_append waves over the claimed bitsets: append, count, clear.
__device__ __forceinline__ void rm_link_own_one(SlotDev &s, const int t) {
    const int a = t / s.M, i = t - a * s.M;
    const int n = s.rs_ncand[(size_t)a * s.M + i];
    if (n == 0) return;
    const int lane = threadIdx.x;
    const size_t e = (size_t)a * s.M * s.rm_Kc + s.rs_first[(size_t)a * s.M + i] + lane;
    bool ok = false;
    int u = 0, v = 0;
    if (lane < n) {
        ok = s.re_coll[e] == 0u;
        u = s.re_pair[e * 2];
        v = s.re_pair[e * 2 + 1];
    }
    const unsigned long long bal = __ballot(ok);
    if (!bal) return;
    int *deg = s.rm_deg + (size_t)a * s.rm_cap;
    int *nbr = s.rm_nbr + (size_t)a * s.rm_cap * s.rm_K;
    if (ok) {
        const int pos = __popcll(bal & ((1ull << lane) - 1ull));
        if (pos < s.rm_K) nbr[(size_t)v * s.rm_K + pos] = u;
        uf_union(s.rm_uf + (size_t)a * s.rm_cap, u, v, s.events);
        int *slot_of = s.rl_slot + (size_t)a * s.rm_cap;
        if (atomicCAS(&slot_of[u], -1, -2) == -1) {  // first claim of u this round
            const int k = atomicAdd(&s.rl_ntouched[a], 1);
            s.rl_touched[(size_t)a * s.rl_cap + k] = u;
            slot_of[u] = k;
        }
    }
    // v is new: nothing else writes its degree before k_rm_link_append
    if (lane == 0) deg[v] = __popcll(bal);
}
__global__ __launch_bounds__(64) void k_rm_link_own(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    for (int t = blockIdx.x; t < s.A * s.M; t += gridDim.x) rm_link_own_one(s, t);  // a wave per item, striding
}

__device__ __forceinline__ void rm_link_bits_one(SlotDev &s, const int t) {
    const int a = t / s.M, i = t - a * s.M;
    const int n = s.rs_ncand[(size_t)a * s.M + i];
    const int lane = threadIdx.x;
    if (lane >= n) return;
    const size_t e = (size_t)a * s.M * s.rm_Kc + s.rs_first[(size_t)a * s.M + i] + lane;
    if (s.re_coll[e]) return;
    const int u = s.re_pair[e * 2], v = s.re_pair[e * 2 + 1];
    const int k = s.rl_slot[(size_t)a * s.rm_cap + u];
    const int j = v - s.rs_base[a];
    atomicOr(&s.rl_bits[((size_t)a * s.rl_cap + k) * s.rl_words + (j >> 5)], 1u << (j & 31));
}
__global__ __launch_bounds__(64) void k_rm_link_bits(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    for (int t = blockIdx.x; t < s.A * s.M; t += gridDim.x) rm_link_bits_one(s, t);  // a wave per item, striding
}

__global__ __launch_bounds__(256) void k_rm_link_append(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.z]];
    const int a = blockIdx.y;
    if (a >= s.A) return;
    const int lane = threadIdx.x & 63;
    const int n_touched = s.rl_ntouched[a];
    const int waves = gridDim.x * (blockDim.x >> 6);
    const int base = s.rs_base[a];
    int *deg = s.rm_deg + (size_t)a * s.rm_cap;
    int *nbr = s.rm_nbr + (size_t)a * s.rm_cap * s.rm_K;
    for (int k = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); k < n_touched; k += waves) {
        const int u = s.rl_touched[(size_t)a * s.rl_cap + k];
        unsigned *bits = s.rl_bits + ((size_t)a * s.rl_cap + k) * s.rl_words;
        int d = deg[u];
        for (int w0 = 0; w0 < s.rl_words; w0 += 64) {
            const int w = w0 + lane;
            const unsigned word = w < s.rl_words ? bits[w] : 0u;
            const int c = __popc(word);
            // exclusive prefix of the set-bit counts over the lanes
            int incl = c;
            for (int off = 1; off < 64; off <<= 1) {
                const int y = __shfl_up(incl, off);
                if (lane >= off) incl += y;
            }
            int pos = d + incl - c;
            unsigned rest = word;
            while (rest) {
                const int b = __ffs(rest) - 1;
                rest &= rest - 1;
                if (pos < s.rm_K) nbr[(size_t)u * s.rm_K + pos] = base + (w << 5) + b;
                ++pos;
            }
            if (w < s.rl_words && word) bits[w] = 0u;  // the pool stays clear between rounds
            d += __shfl(incl, 63);
        }
        if (lane == 0) {
            deg[u] = d;  // every edge counts; the list keeps the first rm_K
            s.rl_slot[(size_t)a * s.rm_cap + u] = -1;
        }
    }
}

// Start and some goal in one component (roadmap_interior.cuh:276-314).
__global__ void k_rm_connected(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int a = threadIdx.x;
    if (a >= s.A) return;
    int *uf = s.rm_uf + (size_t)a * s.rm_cap;
    const int root0 = s.rm_root0[a];
    const int r0 = uf_find(uf, root0, s.events);
    int c = 0;
    for (int g = 0; g < s.n_goals; ++g) c |= uf_find(uf, root0 + 1 + g, s.events) == r0;
    s.rm_connected[a] = c;
}

// ====================================================== tree search ======
// A new dRRT launch restarts its round count (pop_planner.cu:350-364).
__global__ void k_dr_begin(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.x]];
    if (threadIdx.x == 0 && s.round_reset) {
        s.ctr[kCtrRound] = 0;
        s.ctr[kCtrLimit] = s.round_limit;
    }
}

// Trees reset at the start of every AO-x pass (pop_planner.cu:933-950):
// tree 0 = the start, tree 1 = the goals, each its own parent with cost 0.
__global__ void k_tree_reset(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const float *sg = s.sg;
    for (int t = threadIdx.x; t < (1 + s.n_goals) * s.D; t += blockDim.x) {
        const int node = t / s.D, d = t - node * s.D;
        if (node == 0) tree_q(s, 0, 0)[d] = sg[d];
        else tree_q(s, 1, node - 1)[d] = sg[(size_t)node * s.D + d];
    }
    for (int t = threadIdx.x; t < 1 + s.n_goals; t += blockDim.x) {
        const int tr = t == 0 ? 0 : 1, i = t == 0 ? 0 : t - 1;
        s.tr_parent[(size_t)tr * s.tr_cap + i] = i;
        s.tr_cost[(size_t)tr * s.tr_cap + i] = 0.f;
        for (int a = 0; a < s.A; ++a)  // the query's roadmap node t
            s.tr_rid[((size_t)tr * s.tr_cap + i) * s.A + a] = s.rm_root0[a] + t;
    }
    for (int b = threadIdx.x; b < s.B; b += blockDim.x) s.ex_tree[b] = b < s.B / 2 ? 0 : 1;
    if (threadIdx.x == 0) {
        s.tr_n[0] = 1;
        s.tr_n[1] = s.n_goals;
        *s.status = kRunning;
        *s.ed_count = 0;
        *s.ce_count = 0;
    }
}

// Angle at v between (q - v) and (r - v) (utils.cuh:62-95).
__device__ __forceinline__ float edge_angle(const float *v, const float *q, const float *r) {
    float dot = 0.f, vn = 0.f, wn = 0.f;
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const float a = q[j] - v[j], b = r[j] - v[j];
        dot += a * b;
        vn += a * a;
        wn += b * b;
    }
    vn = sqrtf(vn);
    wn = sqrtf(wn);
    if (vn < 1e-8f || wn < 1e-8f) return 0.f;
    const float c = fminf(1.f, fmaxf(-1.f, dot / (vn * wn)));
    return acosf(c);
}

// Enters candidate b's edge (tree ex_tree[b], tree node ex_near[b], roadmap
// nodes ex_rid[b]) in the round's edge table.  Returns -1 if b is the first
// candidate of the round with this edge, else that candidate.  An entry
// belongs to the round whose tag it carries; entries of earlier rounds count
// as free.  It runs in a launch of its own, after the launches that stored
// the candidates' fields, so an entry's owner's fields are read with plain
// loads and no fence (the table entries alone change during the launch).
__device__ int claim_edge(const SlotDev &s, unsigned tag, int b) {
    const int A = s.A;
    const int t = s.ex_tree[b], near = s.ex_near[b];
    const int *rid = s.ex_rid + (size_t)b * A;
    unsigned h = (unsigned)near * 0x9E3779B1u ^ (unsigned)t * 0x85EBCA77u;
    for (int a = 0; a < A; ++a) h = (h ^ (unsigned)rid[a]) * 0xC2B2AE3Du + 0x27D4EB2Fu;
    h ^= h >> 15;
    const unsigned long long mine = ((unsigned long long)tag << 32) | (unsigned)b;
    for (int p = 0; p <= s.ex_hmask; ++p) {
        unsigned long long *e = s.ex_hash + ((h + (unsigned)p) & (unsigned)s.ex_hmask);
        unsigned long long cur = __hip_atomic_load(e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while ((unsigned)(cur >> 32) != tag) {  // free: claim it
            const unsigned long long prev = atomicCAS(e, cur, mine);
            if (prev == cur) return -1;
            cur = prev;
        }
        const int w = (int)(unsigned)cur;
        bool same = s.ex_tree[w] == t && s.ex_near[w] == near;
        for (int a = 0; a < A && same; ++a) same = s.ex_rid[(size_t)w * A + a] == rid[a];
        if (same) return w;
    }
    atomicAdd(&s.events[kEvEdgeTable], 1);  // not reached: the table has more entries than a round has expansions
    return -1;
}

// The round's edges, a thread per expansion: each candidate's edge is
// claimed in the edge table; an expansion of this round with the same edge
// (tree, near, roadmap nodes) has the same verdict, so the edge is checked
// once and later candidates link to its first one (ex_dup).  The first ones
// take edge slots, a wave at a time.
__global__ __launch_bounds__(64) void k_dr_claim(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    if ((int)blockIdx.x * 64 >= s.B || !searching(s)) return;
    const int lane = threadIdx.x;
    const int b = blockIdx.x * 64 + lane;
    const unsigned rng_round = (unsigned)s.ctr[kCtrRng];
    const bool cand = b < s.B && s.ex_state[b] != 0;
    int dup = -1;
    if (cand) dup = claim_edge(s, rng_round + 1u, b);
    if (cand && dup >= 0) s.ex_dup[b] = dup;
    const bool first = cand && dup < 0;
    const unsigned long long m = __ballot(first);
    int base = 0;
    if (lane == 0 && m) base = atomicAdd(s.ed_count, __popcll(m));
    base = __shfl(base, 0);
    if (!first) return;
    const int slot = base + __popcll(m & ((1ull << lane) - 1ull));
    const int D = s.D;
    copy_row(s.ed_from + (size_t)slot * D, tree_q(s, s.ex_tree[b], s.ex_near[b]), D);
    copy_row(s.ed_to + (size_t)slot * D, s.ex_q + (size_t)b * D, D);
    s.ed_b[slot] = b;
    s.ed_coll[slot] = 0u;  // the checker only sets bits
}

// One dRRT expansion (pop_planner.cu:350-622 up to the edge check): pick the
// tree, sample (uniform or PHS), nearest tree node, per-arm roadmap neighbour
// of minimum angle, duplicate test, and the arm-arm edge to check.  The
// stages below run in one wave per expansion (k_dr_expand), or, when the
// trees are large, as five launches in which the two tree scans (nearest
// node, existing node) run a thread per expansion over tree rows shared by
// a block of 64 expansions (k_dr_sample, k_dr_scan<false>, k_dr_step,
// k_dr_scan<true>, k_dr_cand).  Both forms compute the same values in the same order.
struct ExpandLds {
    float q[35];
    float vn[35];
    float nq[35];
    int nrid[kMaxArms];
    float stage[kScanRows * 35];
};

// (1) Tree choice, sample, informed rejection and cost bound.  Writes
// ex_tree, ex_sample, ex_crand (-1: rejected); false if the sample is
// rejected.
__device__ bool expand_sample(SlotDev &s, int b, int lane, ExpandLds &L, int &t_out, float &c_rand_out) {
    const int D = s.D;
    const int round = s.ctr[kCtrRound];
    const unsigned rng_round = (unsigned)s.ctr[kCtrRng];
    float *q = L.q, *nq = L.nq;
    // ---- tree choice (pop_planner.cu:366-376)
    int t = round == 0 ? (b < s.B / 2 ? 0 : 1) : s.ex_tree[b];  // first iteration of a launch
    if (round > 0 && s.balance == 2) {
        const int nt = s.tr_n[t], no = s.tr_n[1 - t];
        const float ratio = (float)abs(nt - no) / (float)nt;
        if (ratio < s.tree_ratio) t = 1 - t;
    }
    if (lane == 0) s.ex_tree[b] = t;
    if (lane == 0) {
        s.ex_state[b] = 0;
        s.ex_index[b] = -1;
        s.ex_dup[b] = -1;
    }
    const int o = 1 - t;
    const float best = *s.best_cost;

    // ---- sample (pop_planner.cu:389-447)
    const uint4 rr = slot_random(s.seed, kStreamTree, rng_round, (unsigned)b, (unsigned)(lane >> 2));
    const unsigned bits = (lane & 3) == 0 ? rr.x : (lane & 3) == 1 ? rr.y : (lane & 3) == 2 ? rr.z : rr.w;
    const float u = u01(bits);  // (0, 1]
    if (!s.phs || !s.optimize) {
        if (lane < D) q[lane] = __builtin_fmaf(u, kScale[lane % 7], kLow[lane % 7]);
    } else {
        // logistic quantile as an approximate normal, then uniform in the unit ball,
        // then the prolate hyperspheroid transform (pop_planner.cu:398-442)
        const float uc = fminf(u, 1.0f - 1.0f / 16777216.0f);  // keep the logit finite
        float z = lane < D ? logf(uc * __frcp_rn(1.0f - uc)) * sqrtf((float)M_PI / 8.0f) : 0.f;
        const float nrm = sqrtf(wave_sum(z * z));
        z = z / nrm;
        const uint4 rad = slot_random(s.seed, kStreamPhsRadius, rng_round, (unsigned)b, 0);
        const float scale = powf(u01(rad.x), 1.0f / (float)D);
        if (lane < D) nq[lane] = z * scale;
        __syncthreads();
        if (lane < D) {
            float v = 0.f;
            for (int c = 0; c < D; ++c) v += s.phs_tf[lane * D + c] * nq[c];
            v += s.phs_center[lane];
            const int j = lane % 7;
            float un = (v - kLow[j]) / kScale[j];
            un = fmaxf(fminf(1.0f, un), 0.0f);
            q[lane] = __builtin_fmaf(un, kScale[j], kLow[j]);
        }
    }
    __syncthreads();

    if (lane < D) s.ex_sample[(size_t)b * D + lane] = q[lane];
    if (lane == 0) s.ex_crand[b] = -1.f;  // stays -1 if the sample is rejected
    // ---- informed rejection (pop_planner.cu:449-456)
    {
        const float *r0 = tree_q(s, 0, 0), *r1 = tree_q(s, 1, 0);
        float d0 = 0.f, d1 = 0.f;
        if (lane < D) {
            const float x0 = r0[lane] - q[lane], x1 = r1[lane] - q[lane];
            d0 = x0 * x0;
            d1 = x1 * x1;
        }
        const float g0 = sqrtf(wave_sum(d0)), g1 = sqrtf(wave_sum(d1));
        if (g0 + g1 >= best) return false;
    }
    // ---- cost bound of the sample (pop_planner.cu:459-469)
    float c_rand = FLT_MAX;
    if (s.optimize) {
        const float *rt = tree_q(s, t, 0), *ro = tree_q(s, o, 0);
        float dt = 0.f, dd = 0.f;
        if (lane < D) {
            const float x = rt[lane] - q[lane], y = ro[lane] - q[lane];
            dt = x * x;
            dd = y * y;
        }
        const float g_hat = sqrtf(wave_sum(dt)), h_hat = sqrtf(wave_sum(dd));
        const float c_range = fmaxf(best - (g_hat + h_hat), 0.0f);
        const uint4 cr = slot_random(s.seed, kStreamCostBound, rng_round, (unsigned)b, 0);
        c_rand = u01(cr.x) * c_range + g_hat;
    }
    if (lane == 0) s.ex_crand[b] = c_rand;
    t_out = t;
    c_rand_out = c_rand;
    return true;
}

// (2) Nearest node of tree t to L.q (pop_planner.cu:472-508), by the wave:
// lane i scans rows i, i + 64, ...; the least squared distance, lowest index
// among equals.
__device__ void expand_nn_wave(SlotDev &s, int t, float c_rand, int lane, ExpandLds &L, int &near, float &dsq) {
    const int D = s.D, S = wave_stride(D);
    const int nt = s.tr_n[t];
    const float *rows = tree_q(s, t, 0);
    float bd = FLT_MAX;
    int bi = 0;
    for (int base = 0; base < nt; base += kScanRows) {
        const int c = min(kScanRows, nt - base);
        __syncthreads();  // the previous chunk has been read
        stage_rows(rows, base, c, D, S, L.stage, lane);
        __syncthreads();
        const int i = base + lane;
        if (lane < c && !(s.optimize && !(s.tr_cost[(size_t)t * s.tr_cap + i] < c_rand))) {
            const float d = sqdist(L.stage + lane * S, L.q, D);
            if (d < bd) {
                bd = d;
                bi = i;
            }
        }
    }
    wave_argmin(bd, bi);
    near = bi;
    dsq = bd;
}

// (3) Per-arm roadmap step of minimum angle from tree node `near` toward L.q
// (utils.cuh:688-759): L.vn <- the node, L.nq / L.nrid <- the new
// composite.  false if some arm has no admissible neighbour.
__device__ bool expand_step(SlotDev &s, int t, int near, float c_rand, int lane, ExpandLds &L) {
    const int D = s.D, A = s.A;
    float *q = L.q, *vn = L.vn, *nq = L.nq;
    if (lane < D) vn[lane] = tree_q(s, t, near)[lane];
    __syncthreads();
    float bound = s.optimize ? c_rand - s.tr_cost[(size_t)t * s.tr_cap + near] : FLT_MAX;
    for (int a = 0; a < A; ++a) {
        const int rid = s.tr_rid[((size_t)t * s.tr_cap + near) * A + a];
        // a degree counts every edge; the list keeps the first rm_K
        const int deg = min(s.rm_deg[(size_t)a * s.rm_cap + rid], s.rm_K);
        const int *nb = s.rm_nbr + ((size_t)a * s.rm_cap + rid) * s.rm_K;
        const float *rq = s.rm_q + (size_t)a * s.rm_cap * 8;
        float bang = 1e3f;
        int bidx = 0x7fffffff;
        for (int j = lane; j < deg; j += 64) {
            const int r = nb[j];
            if (r == rid) continue;
            const float *p = rq + (size_t)r * 8;
            float d2 = 0.f;
            for (int k = 0; k < 7; ++k) {
                const float x = vn[7 * a + k] - p[k];
                d2 = __builtin_fmaf(x, x, d2);
            }
            const float d = sqrtf(d2);
            // a neighbour at the arm's own configuration (a node that repeats
            // it) is no move: every arm must move, as the self-edge excluded
            // above (utils.cuh:708-712)
            if ((s.optimize && d > bound) || d > s.range || d2 == 0.f) continue;
            const float ang = edge_angle(&vn[7 * a], &q[7 * a], p);
            if (ang < bang || (ang == bang && r < bidx)) {
                bang = ang;
                bidx = r;
            }
        }
        wave_argmin(bang, bidx);
        if (bidx == 0x7fffffff) return false;  // this arm has no admissible neighbour
        if (lane < 7) nq[7 * a + lane] = rq[(size_t)bidx * 8 + lane];
        if (lane == 0) L.nrid[a] = bidx;
        __syncthreads();
        float d2 = 0.f;
        for (int k = 0; k < 7; ++k) {
            const float x = nq[7 * a + k] - vn[7 * a + k];
            d2 = __builtin_fmaf(x, x, d2);
        }
        bound -= sqrtf(d2);
    }
    __syncthreads();
    return true;
}

// (4) The lowest index of a node of tree t within squared distance 1e-5 of
// L.nq (pop_planner.cu:526-533), by the wave; 0x7fffffff if none.  Once a
// chunk holds one, later chunks hold only higher indices.
__device__ int expand_exist_wave(SlotDev &s, int t, int lane, ExpandLds &L) {
    const int D = s.D, S = wave_stride(D);
    const int nt = s.tr_n[t];
    const float *rows = tree_q(s, t, 0);
    int exist = 0x7fffffff;
    for (int base = 0; base < nt; base += kScanRows) {
        const int c = min(kScanRows, nt - base);
        __syncthreads();
        stage_rows(rows, base, c, D, S, L.stage, lane);
        __syncthreads();
        if (lane < c && sqdist(L.stage + lane * S, L.nq, D) <= 1e-5f) exist = base + lane;
        if (__ballot(exist != 0x7fffffff)) break;
    }
    return wave_min_int(exist);
}

// (5) The candidate: edge near -> new for the arm-arm check, unless the new
// node exists and is not cheaper, or is above the best cost.  L.nq and
// L.nrid hold the step's result.  k_dr_claim then enters its edge.
__device__ void expand_candidate(SlotDev &s, int b, int t, int near, float dnear, int exist, int lane, ExpandLds &L) {
    const int D = s.D, A = s.A;
    const float best = *s.best_cost;
    const float new_cost = s.tr_cost[(size_t)t * s.tr_cap + near] + dnear;  // (pop_planner.cu:541)
    if (exist != 0x7fffffff) {
        if (!s.optimize || new_cost >= s.tr_cost[(size_t)t * s.tr_cap + exist]) return;
    }
    if (s.optimize && new_cost >= best) return;

    if (lane == 0) {
        s.ex_state[b] = exist != 0x7fffffff ? 2 : 1;
        s.ex_near[b] = near;
        s.ex_exist[b] = exist;
        s.ex_cost[b] = new_cost;
    }
    if (lane < A) s.ex_rid[b * A + lane] = L.nrid[lane];
    if (lane < D) s.ex_q[(size_t)b * D + lane] = L.nq[lane];
}

// All five stages in one wave per expansion.
__global__ __launch_bounds__(64) void k_dr_expand(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int b = blockIdx.x;
    if (b >= s.B || !searching(s)) return;
    const int lane = threadIdx.x;
    __shared__ ExpandLds L;
    int t, near;
    float c_rand, dsq;
    if (!expand_sample(s, b, lane, L, t, c_rand)) return;
    expand_nn_wave(s, t, c_rand, lane, L, near, dsq);
    if (!expand_step(s, t, near, c_rand, lane, L)) return;
    const int exist = expand_exist_wave(s, t, lane, L);
    expand_candidate(s, b, t, near, sqrtf(dsq), exist, lane, L);
}

// ---- the same stages over five launches, for large trees
__global__ __launch_bounds__(64) void k_dr_sample(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int b = blockIdx.x;
    if (b >= s.B || !searching(s)) return;
    __shared__ ExpandLds L;
    int t;
    float c_rand;
    const bool ok = expand_sample(s, b, threadIdx.x, L, t, c_rand);
    if (threadIdx.x == 0) {
        s.ex_live[b] = ok ? 1 : 0;
        s.ex_nnkey[b] = ~0ull;
        s.ex_exist[b] = 0x7fffffff;
    }
}

// Thread-per-expansion scans.  A block holds 64 expansions of one slot and
// one of `splits` contiguous slices of each tree (whole chunks of 64 rows);
// the rows are staged a chunk at a time in LDS and every thread compares
// them with its own query, held in registers, in index order.  Slices merge
// by atomicMin: on (squared distance bits, index) for the nearest node (a
// non-negative float orders as its bits, so the least distance wins and the
// lowest index among equals, as in the wave scan), on the index for the
// existing node.  A squared distance only grows term by term (fma of a
// square, rounded upward or to nearest, never below the partial sum), so a
// row whose first arms already reach the bound cannot pass and its remaining
// terms are skipped.
template <int D, bool EXIST>
__device__ void scan_block(SlotDev &s, int b, int split, int splits, float *stage, float *cstage) {
    const int lane = threadIdx.x;
    const bool live = b < s.B && s.ex_live[b] == (EXIST ? 2 : 1);
    const int t_me = live ? s.ex_tree[b] : -1;
    const float c_rand = live && !EXIST ? s.ex_crand[b] : 0.f;
    const float *src = EXIST ? s.ex_q : s.ex_sample;
    float q[D];
#pragma unroll
    for (int d = 0; d < D; ++d) q[d] = live ? src[(size_t)b * D + d] : 0.f;
    float bd = FLT_MAX;
    int bi = 0x7fffffff;
    for (int t = 0; t < 2; ++t) {
        if (!__syncthreads_or(t_me == t)) continue;
        const int nt = s.tr_n[t];
        const int per = ((nt + kScanRows - 1) / kScanRows + splits - 1) / splits * kScanRows;
        const int lo = split * per, hi = min(nt, lo + per);
        const float *rows = tree_q(s, t, 0);
        const float *cost = s.tr_cost + (size_t)t * s.tr_cap;
        for (int base = lo; base < hi; base += kScanRows) {
            const int c = min(kScanRows, hi - base);
            __syncthreads();  // the previous chunk has been read
            stage_rows(rows, base, c, D, D, stage, lane);
            if (!EXIST && s.optimize && lane < c) cstage[lane] = cost[base + lane];
            __syncthreads();
            if (t_me == t && !(EXIST && bi != 0x7fffffff)) {
                for (int r = 0; r < c; ++r) {
                    if (!EXIST && s.optimize && !(cstage[r] < c_rand)) continue;
                    const float *row = stage + r * D;
                    float acc = 0.f;
                    bool out = false;
#pragma unroll
                    for (int a = 0; a < D / 7; ++a) {
                        if (!out) {
#pragma unroll
                            for (int k = 0; k < 7; ++k) {
                                const float x = row[7 * a + k] - q[7 * a + k];
                                acc = __builtin_fmaf(x, x, acc);
                            }
                            if (a + 1 < D / 7) out = EXIST ? acc > 1e-5f : acc >= bd;
                        }
                    }
                    if (out) continue;
                    if (EXIST) {
                        if (acc <= 1e-5f) {
                            bi = base + r;
                            break;
                        }
                    } else if (acc < bd) {
                        bd = acc;
                        bi = base + r;
                    }
                }
            }
            // the existing-node scan of a slice ends once every expansion of
            // the block on this tree has its node
            if (EXIST && __syncthreads_and(t_me != t || bi != 0x7fffffff)) break;
        }
    }
    if (!live || bi == 0x7fffffff) return;
    if (EXIST)
        atomicMin(s.ex_exist + b, bi);
    else
        atomicMin(s.ex_nnkey + b, ((unsigned long long)__float_as_uint(bd) << 32) | (unsigned)bi);
}

template <bool EXIST>
__global__ __launch_bounds__(64) void k_dr_scan(SlotDev *slots, const int *ids, int splits) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int group = blockIdx.x / splits, split = blockIdx.x % splits;
    const int b0 = group * 64;
    if (b0 >= s.B || !searching(s)) return;
    __shared__ float stage[kScanRows * 35];
    __shared__ float cstage[kScanRows];
    const int b = b0 + threadIdx.x;
    switch (s.D) {
    case 14: scan_block<14, EXIST>(s, b, split, splits, stage, cstage); break;
    case 21: scan_block<21, EXIST>(s, b, split, splits, stage, cstage); break;
    case 28: scan_block<28, EXIST>(s, b, split, splits, stage, cstage); break;
    default: scan_block<35, EXIST>(s, b, split, splits, stage, cstage); break;
    }
}

// The nearest node of expansion b from the merged key: index 0 at distance
// FLT_MAX when no node passed the cost bound, as the wave scan leaves it.
__device__ __forceinline__ void nn_from_key(const SlotDev &s, int b, int &near, float &dsq) {
    const unsigned long long k = s.ex_nnkey[b];
    near = k == ~0ull ? 0 : (int)(unsigned)k;
    dsq = k == ~0ull ? FLT_MAX : __uint_as_float((unsigned)(k >> 32));
}

__global__ __launch_bounds__(64) void k_dr_step(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int b = blockIdx.x;
    if (b >= s.B || !searching(s) || s.ex_live[b] != 1) return;
    const int lane = threadIdx.x;
    __shared__ ExpandLds L;
    const int D = s.D;
    if (lane < D) L.q[lane] = s.ex_sample[(size_t)b * D + lane];
    __syncthreads();
    int near;
    float dsq;
    nn_from_key(s, b, near, dsq);
    const int t = s.ex_tree[b];
    const bool ok = expand_step(s, t, near, s.ex_crand[b], lane, L);
    if (ok) {
        if (lane < s.A) s.ex_rid[b * s.A + lane] = L.nrid[lane];
        if (lane < D) s.ex_q[(size_t)b * D + lane] = L.nq[lane];
    }
    if (lane == 0) s.ex_live[b] = ok ? 2 : 0;
}

__global__ __launch_bounds__(64) void k_dr_cand(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int b = blockIdx.x;
    if (b >= s.B || !searching(s) || s.ex_live[b] != 2) return;
    const int lane = threadIdx.x;
    __shared__ ExpandLds L;
    const int D = s.D, A = s.A;
    int near;
    float dsq;
    nn_from_key(s, b, near, dsq);
    const int t = s.ex_tree[b];
    if (lane < D) L.nq[lane] = s.ex_q[(size_t)b * D + lane];
    if (lane < A) L.nrid[lane] = s.ex_rid[b * A + lane];
    __syncthreads();
    expand_candidate(s, b, t, near, sqrtf(dsq), s.ex_exist[b], lane, L);
}

// Commit the checked expansions in expansion order (pop_planner.cu:571-622):
// free rewires update an existing node (lowest cost wins, then lowest b),
// free inserts get consecutive indices per tree.  One block per slot.
__global__ __launch_bounds__(256) void k_dr_commit(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    if (!searching(s)) return;
    const int tid = threadIdx.x;
    const int ne = *s.ed_count;
    // verdicts back to expansions
    for (int e = tid; e < ne; e += 256) {
        const int b = s.ed_b[e];
        if (s.ed_coll[e]) s.ex_state[b] = 0;
    }
    __syncthreads();
    // an expansion whose edge was checked for another takes that verdict
    // (ex_dup only names expansions that own an edge, so no state read here
    // is written in this loop)
    for (int b = tid; b < s.B; b += 256) {
        const int w = s.ex_dup[b];
        if (w >= 0 && s.ex_state[w] == 0) s.ex_state[b] = 0;
    }
    __syncthreads();
    // rewire arbitration
    for (int b = tid; b < s.B; b += 256) {
        if (s.ex_state[b] == 2) {
            const int t = s.ex_tree[b];
            const unsigned long long key =
                ((unsigned long long)__float_as_uint(s.ex_cost[b]) << 32) | (unsigned)b;
            atomicMin(&s.rw_key[(size_t)t * s.tr_cap + s.ex_exist[b]], key);
        }
    }
    __syncthreads();
    for (int b = tid; b < s.B; b += 256) {
        if (s.ex_state[b] == 2) {
            const int t = s.ex_tree[b], x = s.ex_exist[b];
            const unsigned long long key = s.rw_key[(size_t)t * s.tr_cap + x];
            if ((unsigned)key == (unsigned)b) {
                s.tr_parent[(size_t)t * s.tr_cap + x] = s.ex_near[b];
                s.tr_cost[(size_t)t * s.tr_cap + x] = s.ex_cost[b];
            }
        }
    }
    __syncthreads();
    for (int b = tid; b < s.B; b += 256) {
        if (s.ex_state[b] == 2) s.rw_key[(size_t)s.ex_tree[b] * s.tr_cap + s.ex_exist[b]] = ~0ull;
        s.cn_key[b] = ~0ull;  // the split connection scan merges into it
    }
    // insert: prefix over b per tree
    __shared__ int sh0[256], sh1[256];
    const int per = (s.B + 255) / 256;
    const int b0 = tid * per;
    int c0 = 0, c1 = 0;
    for (int b = b0; b < min(b0 + per, s.B); ++b)
        if (s.ex_state[b] == 1) (s.ex_tree[b] == 0 ? c0 : c1)++;
    sh0[tid] = c0;
    sh1[tid] = c1;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int a0 = tid >= off ? sh0[tid - off] : 0, a1 = tid >= off ? sh1[tid - off] : 0;
        __syncthreads();
        sh0[tid] += a0;
        sh1[tid] += a1;
        __syncthreads();
    }
    const int n0 = s.tr_n[0], n1 = s.tr_n[1];
    int i0 = n0 + sh0[tid] - c0, i1 = n1 + sh1[tid] - c1;
    bool overflow = false;
    for (int b = b0; b < min(b0 + per, s.B); ++b) {
        if (s.ex_state[b] != 1) continue;
        const int t = s.ex_tree[b];
        const int idx = t == 0 ? i0++ : i1++;
        if (idx >= s.tr_cap) {
            overflow = true;
            s.ex_state[b] = 0;
            continue;
        }
        s.ex_index[b] = idx;
        copy_row(tree_q(s, t, idx), s.ex_q + (size_t)b * s.D, s.D);
        for (int a = 0; a < s.A; ++a) s.tr_rid[((size_t)t * s.tr_cap + idx) * s.A + a] = s.ex_rid[b * s.A + a];
        s.tr_parent[(size_t)t * s.tr_cap + idx] = s.ex_near[b];
        s.tr_cost[(size_t)t * s.tr_cap + idx] = s.ex_cost[b];
    }
    if (overflow) atomicExch(s.status, kFailedOverflow);  // (pop_planner.cu:587-589)
    __syncthreads();
    if (tid == 255) {
        s.tr_n[0] = min(n0 + sh0[255], s.tr_cap);
        s.tr_n[1] = min(n1 + sh1[255], s.tr_cap);
        *s.ed_count = 0;
        *s.ce_count = 0;
    }
}

// Thread-per-expansion form of the connection scan below, for large trees:
// a block holds 64 expansions of one slot and one of `splits` contiguous
// slices of the trees; rows are staged a chunk at a time and every thread
// compares them with its own inserted node in index order, keeping the
// least squared distance that passes the cost bound (lowest index among
// equals).  Slices merge by atomicMin on (distance bits, index), so the
// result is the wave scan's.  A partial sum that already reaches the best
// distance cannot pass (the sum only grows), so its row is left early.
template <int D>
__device__ void connect_scan_block(SlotDev &s, int b, int split, int splits, float *stage) {
    const int lane = threadIdx.x;
    const int idx = b < s.B ? s.ex_index[b] : -1;
    const bool live = b < s.B && s.ex_state[b] == 1 && idx >= 0;
    const int t = live ? s.ex_tree[b] : 0;
    const int o_me = live ? 1 - t : -1;
    float q[D];
#pragma unroll
    for (int d = 0; d < D; ++d) q[d] = live ? tree_q(s, t, idx)[d] : 0.f;
    const float ub = live ? *s.best_cost - s.tr_cost[(size_t)t * s.tr_cap + idx] : 0.f;
    float bd = FLT_MAX;
    int bi = 0x7fffffff;
    for (int o = 0; o < 2; ++o) {
        if (!__syncthreads_or(o_me == o)) continue;
        const int no = s.tr_n[o];
        const int per = ((no + kScanRows - 1) / kScanRows + splits - 1) / splits * kScanRows;
        const int lo = split * per, hi = min(no, lo + per);
        const float *rows = tree_q(s, o, 0);
        for (int base = lo; base < hi; base += kScanRows) {
            const int c = min(kScanRows, hi - base);
            __syncthreads();  // the previous chunk has been read
            stage_rows(rows, base, c, D, D, stage, lane);
            __syncthreads();
            if (o_me != o) continue;
            for (int r = 0; r < c; ++r) {
                const float *row = stage + r * D;
                float acc = 0.f;
                bool out = false;
#pragma unroll
                for (int a = 0; a < D / 7; ++a) {
                    if (!out) {
#pragma unroll
                        for (int k = 0; k < 7; ++k) {
                            const float x = row[7 * a + k] - q[7 * a + k];
                            acc = __builtin_fmaf(x, x, acc);
                        }
                        if (a + 1 < D / 7) out = acc >= bd;
                    }
                }
                if (!out && acc < bd && (!s.optimize || sqrtf(acc) < ub)) {
                    bd = acc;
                    bi = base + r;
                }
            }
        }
    }
    if (live && bi != 0x7fffffff)
        atomicMin(s.cn_key + b, ((unsigned long long)__float_as_uint(bd) << 32) | (unsigned)bi);
}

__global__ __launch_bounds__(64) void k_cn_scan(SlotDev *slots, const int *ids, int splits) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int group = blockIdx.x / splits, split = blockIdx.x % splits;
    const int b0 = group * 64;
    if (b0 >= s.B || !searching(s)) return;
    __shared__ float stage[kScanRows * 35];
    const int b = b0 + threadIdx.x;
    switch (s.D) {
    case 14: connect_scan_block<14>(s, b, split, splits, stage); break;
    case 21: connect_scan_block<21>(s, b, split, splits, stage); break;
    case 28: connect_scan_block<28>(s, b, split, splits, stage); break;
    default: connect_scan_block<35>(s, b, split, splits, stage); break;
    }
}

// Nearest node of the other tree for every inserted node, and the straight
// connection split into n_ext = floor(d / range) + 1 chunks of G samples
// (pop_planner.cu:624-703).  One wave per expansion; keyed: the nearest node
// comes from k_cn_scan (cn_key) instead of the wave's own scan.
__global__ __launch_bounds__(64) void k_dr_connect(SlotDev *slots, const int *ids, int keyed) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int b = blockIdx.x;
    if (b >= s.B || !searching(s)) return;
    const int lane = threadIdx.x;
    if (lane == 0) s.cn_near[b] = -1;
    const int idx = s.ex_index[b];
    if (s.ex_state[b] != 1 || idx < 0) return;
    const int t = s.ex_tree[b], o = 1 - t, D = s.D, S = wave_stride(D);
    __shared__ float q[35];
    __shared__ float stage[kScanRows * 35];
    if (lane < D) q[lane] = tree_q(s, t, idx)[lane];
    __syncthreads();
    float bd = FLT_MAX;
    int bi = 0x7fffffff;
    if (keyed) {
        const unsigned long long k = s.cn_key[b];
        if (k != ~0ull) {
            bd = __uint_as_float((unsigned)(k >> 32));
            bi = (int)(unsigned)k;
        }
    } else {
        const float best = *s.best_cost;
        const float ub = best - s.tr_cost[(size_t)t * s.tr_cap + idx];
        const int no = s.tr_n[o];
        const float *rows = tree_q(s, o, 0);
        for (int base = 0; base < no; base += kScanRows) {
            const int c = min(kScanRows, no - base);
            __syncthreads();
            stage_rows(rows, base, c, D, S, stage, lane);
            __syncthreads();
            if (lane < c) {
                const float d = sqdist(stage + lane * S, q, D);
                if (d < bd && (!s.optimize || sqrtf(d) < ub)) {
                    bd = d;
                    bi = base + lane;
                }
            }
        }
        wave_argmin(bd, bi);
    }
    if (bi == 0x7fffffff) return;
    const float d = sqrtf(bd);
    const int next = (int)(d / s.range) + 1;
    if (next > s.ce_cap / s.B) {  // longer than the chunk budget per expansion: no attempt
        if (lane == 0) atomicAdd(&s.events[kEvConnectSkip], 1);
        return;
    }
    int base = 0;
    if (lane == 0) base = atomicAdd(s.ce_count, next);
    base = __shfl(base, 0);
    if (lane == 0) {
        s.cn_near[b] = bi;
        s.cn_next[b] = next;
        s.cn_d[b] = d;
        s.cn_fail[b] = 0;
    }
    if (lane < D) {
        const float *on = tree_q(s, o, bi);
        const float vec = (on[lane] - q[lane]) / (float)next;
        float cur = q[lane];
        for (int c = 0; c < next; ++c) {
            const size_t e = (size_t)(base + c);
            s.ce_from[e * D + lane] = cur;
            cur += vec;  // the reference advances the configuration chunk by chunk
            s.ce_to[e * D + lane] = cur;
        }
    }
    for (int c = lane; c < next; c += 64) {
        s.ce_b[base + c] = b;
        s.ce_coll[base + c] = 0u;  // the checker only sets bits
    }
}

// Connections whose chunks are all free add their end node to tree t; the
// lowest expansion among them records the path (pop_planner.cu:705-731,
// path_processing.cu:10-70).  One block per slot.
__global__ __launch_bounds__(256) void k_dr_finish(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    if (!searching(s)) return;
    const int tid = threadIdx.x;
    const int nce = *s.ce_count;
    for (int e = tid; e < nce; e += 256)
        if (s.ce_coll[e]) s.cn_fail[s.ce_b[e]] = 1;
    __syncthreads();
    // successful connections get consecutive indices per tree in expansion order
    __shared__ int sh0[256], sh1[256];
    __shared__ int s_winner, s_wnode, s_overflow;
    if (tid == 0) {
        s_winner = 0x7fffffff;
        s_overflow = 0;
    }
    const int per = (s.B + 255) / 256;
    const int b0 = tid * per, b1 = min(b0 + per, s.B);
    int c0 = 0, c1 = 0;
    for (int b = b0; b < b1; ++b)
        if (s.cn_near[b] >= 0 && !s.cn_fail[b]) (s.ex_tree[b] == 0 ? c0 : c1)++;
    sh0[tid] = c0;
    sh1[tid] = c1;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int a0 = tid >= off ? sh0[tid - off] : 0, a1 = tid >= off ? sh1[tid - off] : 0;
        __syncthreads();
        sh0[tid] += a0;
        sh1[tid] += a1;
        __syncthreads();
    }
    const int n0 = s.tr_n[0], n1 = s.tr_n[1];
    int i0 = n0 + sh0[tid] - c0, i1 = n1 + sh1[tid] - c1;
    int my_first = 0x7fffffff, my_node = -1;  // lowest successful expansion of this thread
    for (int b = b0; b < b1; ++b) {
        if (s.cn_near[b] < 0 || s.cn_fail[b]) continue;
        const int t = s.ex_tree[b], o = 1 - t;
        const int idx = t == 0 ? i0++ : i1++;
        if (idx >= s.tr_cap) {
            s_overflow = 1;
            continue;
        }
        if (my_first == 0x7fffffff) {
            my_first = b;
            my_node = idx;
        }
        // end configuration: the start advanced by n_ext chunks, as in the chunk edges
        const float *st = tree_q(s, t, s.ex_index[b]), *on = tree_q(s, o, s.cn_near[b]);
        float *dst = tree_q(s, t, idx);
        const int next = s.cn_next[b];
        for (int d = 0; d < s.D; ++d) {
            const float vec = (on[d] - st[d]) / (float)next;
            float cur = st[d];
            for (int c = 0; c < next; ++c) cur += vec;
            dst[d] = cur;
        }
        for (int a = 0; a < s.A; ++a)
            s.tr_rid[((size_t)t * s.tr_cap + idx) * s.A + a] = s.tr_rid[((size_t)o * s.tr_cap + s.cn_near[b]) * s.A + a];
        s.tr_parent[(size_t)t * s.tr_cap + idx] = s.ex_index[b];
        // the reference adds one chunk's length here (pop_planner.cu:709-710)
        s.tr_cost[(size_t)t * s.tr_cap + idx] = s.tr_cost[(size_t)t * s.tr_cap + s.ex_index[b]] + s.cn_d[b] / next;
    }
    if (my_first != 0x7fffffff) atomicMin(&s_winner, my_first);
    __syncthreads();
    if (my_first != 0x7fffffff && my_first == s_winner) s_wnode = my_node;
    __syncthreads();
    if (tid != 0) return;
    s.tr_n[0] = min(n0 + sh0[255], s.tr_cap);
    s.tr_n[1] = min(n1 + sh1[255], s.tr_cap);
    s.ctr[kCtrRound] += 1;
    s.ctr[kCtrRng] += 1;
    s.ctr[kCtrRun] += 1;
    if (s_overflow) {
        *s.status = kFailedOverflow;  // (pop_planner.cu:587-589)
        return;
    }
    const int winner = s_winner == 0x7fffffff ? -1 : s_winner, wnode = s_wnode;
    if (winner < 0) return;
    // ---- trace both trees to their roots; path runs goal -> start
    const int t = s.ex_tree[winner], D = s.D;
    int chain_n[2] = {0, 0};
    // count chains
    const int start_node[2] = {t == 0 ? wnode : s.cn_near[winner], t == 1 ? wnode : s.cn_near[winner]};
    // parent links strictly lower the (round-start) node cost, so chains end at a
    // root; the length check also guards against corrupted links
    for (int tr = 0; tr < 2; ++tr) {
        int cur = start_node[tr], n = 1;
        while (s.tr_parent[(size_t)tr * s.tr_cap + cur] != cur && n <= s.path_cap) {
            cur = s.tr_parent[(size_t)tr * s.tr_cap + cur];
            ++n;
        }
        chain_n[tr] = n;
    }
    const int total = chain_n[0] + chain_n[1];
    if (total > s.path_cap) {
        *s.status = kFailedOverflow;
        return;
    }
    float cost = 0.f;
    // tree 1 reversed first: its root (the goal reached) ... its chain start.
    // The reference writes node 0 of each tree as the root
    // (path_processing.cu:31-33, 48-50); for the start tree and a single goal
    // that is the root, with several goals it would be goal 0 whichever goal
    // the chain reached, so the root reached is written instead.
    {
        int cur = start_node[1], k = chain_n[1] - 1;
        while (true) {
            const int par = s.tr_parent[(size_t)s.tr_cap + cur];
            copy_row(s.path + (size_t)k * D, tree_q(s, 1, cur), D);
            if (par == cur) break;
            cost += sqrtf(sqdist(tree_q(s, 1, cur), tree_q(s, 1, par), D));
            cur = par;
            --k;
        }
    }
    {
        int cur = start_node[0], k = chain_n[1];
        while (true) {
            const int par = s.tr_parent[cur];
            copy_row(s.path + (size_t)k * D, tree_q(s, 0, cur), D);  // the start tree's only root is node 0
            if (par == cur) break;
            cost += sqrtf(sqdist(tree_q(s, 0, cur), tree_q(s, 0, par), D));
            cur = par;
            ++k;
        }
    }
    *s.path_n = total;
    *s.path_cost = cost;
    s.path_meta[0] = s.tr_n[0];
    s.path_meta[1] = s.tr_n[1];
    s.path_meta[2] = s.ctr[kCtrRound];
    *s.best_cost = fminf(*s.best_cost, cost);
    *s.status = winner + 1;
}

}  // namespace plan
}  // namespace mrpop

// ====================================================== simplification ====
// path_simplification.cu:41-238.  Stages 0..bspline_step-1 smooth the odd
// waypoints; the last stage checks every shortcut i -> j (j - i <= 16) and
// runs the shortcut DP.  Chunk edges are allocated with a block-wide prefix
// so the assignment is deterministic; candidates that do not fit are treated
// as colliding (the path stays valid, only less simplified).
namespace mrpop {
namespace plan {

__device__ __forceinline__ float l2(const float *a, const float *b, int D) { return sqrtf(sqdist(a, b, D)); }

// Block-wide exclusive prefix of per-thread counts (blockDim.x == 1024).
__device__ __forceinline__ int block_prefix(int v, int *sh, int &total) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int add = tid >= off ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += add;
        __syncthreads();
    }
    total = sh[1023];
    const int ex = sh[tid] - v;
    __syncthreads();
    return ex;
}

// Writes chunk edges for candidates; fail[item] starts at 0.
//   stage < bspline_step: items are odd waypoints i: edges p[i-1] -> mid and p[i+1] -> mid
//   otherwise:            items are (i, b) pairs, j = i + b + 1, b in 1..15: edge p[i] -> p[j]
__global__ __launch_bounds__(1024) void k_sp_edges(SlotDev *slots, const int *ids, int stage, int bspline_step,
                                                   float range, float min_change) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int N = *s.path_n, D = s.D;
    __shared__ int sh[1024];
    const bool smooth = stage < bspline_step;
    const int n_items = smooth ? N : N * 16;
    int base = 0;
    for (int chunk0 = 0; chunk0 < n_items; chunk0 += 1024) {
        const int item = chunk0 + threadIdx.x;
        int cnt = 0;  // chunks of this item
        int c1 = 0, c2 = 0;
        bool live = false;
        if (item < n_items) {
            s.sp_fail[item] = 0;
            if (smooth) {
                const int i = item;
                if ((i & 1) && i < N - 1) {
                    const float *p0 = s.path + (size_t)(i - 1) * D, *p1 = s.path + (size_t)i * D,
                                *p2 = s.path + (size_t)(i + 1) * D;
                    float *mid = s.sp_mid + (size_t)i * D;
                    float moved = 0.f;
                    for (int d = 0; d < D; ++d) {
                        const float b1 = (p0[d] + p1[d]) / 2.0f, b2 = (p1[d] + p2[d]) / 2.0f;
                        mid[d] = (b1 + b2) / 2.0f;
                        const float x = p1[d] - mid[d];
                        moved = __builtin_fmaf(x, x, moved);
                    }
                    if (sqrtf(moved) >= min_change) {
                        live = true;
                        c1 = (int)ceilf(l2(p0, mid, D) / range);
                        c2 = (int)ceilf(l2(p2, mid, D) / range);
                        cnt = c1 + c2;
                    }
                }
                if (!live) s.sp_fail[i] = 1;  // nothing to apply
            } else {
                const int i = item >> 4, b = item & 15, j = i + b + 1;
                if (b >= 1 && j < N) {
                    live = true;
                    const float d = l2(s.path + (size_t)i * D, s.path + (size_t)j * D, D);
                    c1 = (int)ceilf((d - 0.001f) / range);  // path_simplification.cu:159-162
                    if (c1 < 0) c1 = 0;
                    cnt = c1;
                }
            }
        }
        int total;
        int off = block_prefix(cnt, sh, total);
        off += base;
        if (live && cnt > 0) {
            if (off + cnt > s.sp_cap) {
                s.sp_fail[item] = 1;  // no room: keep the waypoint / no shortcut
                atomicAdd(&s.events[kEvSimplifyDrop], 1);
            } else {
                int e = off;
                for (int side = 0; side < (smooth ? 2 : 1); ++side) {
                    const float *from, *to;
                    int nc;
                    if (smooth) {
                        from = s.path + (size_t)(side == 0 ? item - 1 : item + 1) * D;
                        to = s.sp_mid + (size_t)item * D;
                        nc = side == 0 ? c1 : c2;
                    } else {
                        from = s.path + (size_t)(item >> 4) * D;
                        to = s.path + (size_t)((item >> 4) + (item & 15) + 1) * D;
                        nc = c1;
                    }
                    for (int d = 0; d < D; ++d) {
                        const float delta = (to[d] - from[d]) / (float)nc;
                        float cur = from[d];
                        for (int c = 0; c < nc; ++c) {
                            s.sp_from[(size_t)(e + c) * D + d] = cur;
                            cur += delta;  // the reference advances by delta per chunk
                            s.sp_to[(size_t)(e + c) * D + d] = cur;
                        }
                    }
                    for (int c = 0; c < nc; ++c) {
                        s.sp_item[e + c] = item;
                        s.sp_coll[e + c] = 0u;  // the checker only sets bits
                    }
                    e += nc;
                }
            }
        }
        base += total;
    }
    if (threadIdx.x == 0) *s.sp_count = min(base, s.sp_cap);
}

// Chunk verdicts back to their items.
__global__ void k_sp_collect(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int n = *s.sp_count;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
        if (s.sp_coll[e]) s.sp_fail[s.sp_item[e]] = 1;
}

// Smoothing stage: replace free odd waypoints by their midpoints.
__global__ void k_sp_smooth_apply(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    const int N = *s.path_n, D = s.D;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < N * D; t += gridDim.x * blockDim.x) {
        const int i = t / D;
        if ((i & 1) && i < N - 1 && !s.sp_fail[i]) s.path[t] = s.sp_mid[t];
    }
}

// Shortcut DP over the checked window edges, path rewrite with copies of
// the kept waypoints, and the simplified cost (path_simplification.cu:143-238).
__global__ void k_sp_shortcut(SlotDev *slots, const int *ids) {
    SlotDev &s = slots[ids[blockIdx.y]];
    if (threadIdx.x != 0) return;
    const int N = *s.path_n, D = s.D;
    float *cost = s.sp_cost;
    int *par = s.sp_parent;
    cost[0] = 0.f;
    par[0] = -1;
    for (int t = 1; t < N; ++t) {
        cost[t] = 3.402823466e38f;
        par[t] = -1;
    }
    for (int i = 0; i < N - 1; ++i) {
        for (int b = 0; b < 16; ++b) {
            const int j = i + b + 1;
            if (j >= N) break;
            if (b > 0 && s.sp_fail[i * 16 + b]) continue;
            const float c = l2(s.path + (size_t)i * D, s.path + (size_t)j * D, D) + cost[i];
            if (cost[j] > c) {
                cost[j] = c;
                par[j] = i;
            }
        }
    }
    const float best = *s.best_cost;
    if (cost[N - 1] < best) {
        int child = N - 1, parent = par[child];
        while (parent < N && parent != -1) {
            for (int idx = parent + 1; idx < child; ++idx)
                copy_row(s.path + (size_t)idx * D, s.path + (size_t)parent * D, D);
            child = parent;
            parent = par[child];
        }
    }
    *s.path_cost = cost[N - 1];
    *s.best_cost = fminf(best, cost[N - 1]);
}

// Per-slot report for the host.
struct ReportDev {
    int status;
    int rm_connected[kMaxArms];
    int rm_n[kMaxArms];
    int tr_n[2];
    int path_n;
    float path_cost;
    float best_cost;
    int path_meta[3];
    int sp_count;
    int ctr[4];
    int events[kEvents];
};

__global__ void k_report(SlotDev *slots, const int *ids, int n, ReportDev *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int id = ids[t];
    SlotDev &s = slots[id];
    ReportDev &r = out[id];
    r.status = *s.status;
    for (int a = 0; a < kMaxArms; ++a) {
        r.rm_connected[a] = a < s.A ? s.rm_connected[a] : 1;
        r.rm_n[a] = a < s.A ? s.rm_n[a] : 0;
    }
    r.tr_n[0] = s.tr_n[0];
    r.tr_n[1] = s.tr_n[1];
    r.path_n = *s.path_n;
    r.path_cost = *s.path_cost;
    r.best_cost = *s.best_cost;
    for (int k = 0; k < 3; ++k) r.path_meta[k] = s.path_meta[k];
    r.sp_count = *s.sp_count;
    for (int k = 0; k < 4; ++k) r.ctr[k] = s.ctr[k];
    for (int k = 0; k < kEvents; ++k) r.events[k] = s.events[k];
}

// ------------------------------------------------------------ launchers ----
#define MRPOP_L(kernel, grid, block, ...) hipLaunchKernelGGL(kernel, grid, block, 0, stream, __VA_ARGS__)

void launch_roadmap_init(SlotDev *d, const int *ids, int n, hipStream_t stream) {
    MRPOP_L(k_roadmap_init, dim3(1, n), dim3(256), d, ids);
}
void launch_tree_reset(SlotDev *d, const int *ids, int n, hipStream_t stream) {
    MRPOP_L(k_tree_reset, dim3(1, n), dim3(256), d, ids);
}
void launch_rm_sample(SlotDev *d, const int *ids, int n, int max_am, hipStream_t stream) {
    MRPOP_L(k_rm_sample, dim3((max_am + 255) / 256, n), dim3(256), d, ids);
}
void launch_rm_assign(SlotDev *d, const int *ids, int n, hipStream_t stream) {
    MRPOP_L(k_rm_assign, dim3(kMaxArms, n), dim3(256), d, ids);
}
// Waves of the per-(arm, sample) roadmap kernels per slot: one per item for
// a few slots, at most about 16 K waves over all slots for many (each wave
// then takes every gridDim.x-th item), since a launch of one short wave per
// item of 50 five-arm slots (460 K waves) is bound by wave dispatch.
static int rm_waves(int max_am, int n) { return std::max(64, std::min(max_am, 16384 / std::max(1, n))); }

void launch_rm_radius(SlotDev *d, const int *ids, int n, int max_am, hipStream_t stream) {
    MRPOP_L(k_rm_radius, dim3(rm_waves(max_am, n), n), dim3(64), d, ids);
}
void launch_rm_compact(SlotDev *d, const int *ids, int n, int max_am, hipStream_t stream) {
    MRPOP_L(k_rm_compact, dim3(kMaxArms, n), dim3(1024), d, ids);
    MRPOP_L(k_rm_pack, dim3(max_am, n), dim3(64), d, ids);  // short stores: a wave per item is faster here
}
void launch_rm_link(SlotDev *d, const int *ids, int n, int max_am, hipStream_t stream) {
    MRPOP_L(k_rm_link_own, dim3(rm_waves(max_am, n), n), dim3(64), d, ids);
    MRPOP_L(k_rm_link_bits, dim3(rm_waves(max_am, n), n), dim3(64), d, ids);
    MRPOP_L(k_rm_link_append, dim3(64, kMaxArms, n), dim3(256), d, ids);
}
void launch_rm_connected(SlotDev *d, const int *ids, int n, hipStream_t stream) {
    MRPOP_L(k_rm_connected, dim3(1, n), dim3(64), d, ids);
}
void launch_dr_begin(SlotDev *d, const int *ids, int n, hipStream_t stream) {
    MRPOP_L(k_dr_begin, dim3(n), dim3(64), d, ids);
}
void launch_dr_expand(SlotDev *d, const int *ids, int n, int max_b, int splits, hipStream_t stream) {
    const int groups = (max_b + 63) / 64;
    if (splits <= 0) {
        MRPOP_L(k_dr_expand, dim3(max_b, n), dim3(64), d, ids);
    } else {
        MRPOP_L(k_dr_sample, dim3(max_b, n), dim3(64), d, ids);
        MRPOP_L(k_dr_scan<false>, dim3(groups * splits, n), dim3(64), d, ids, splits);
        MRPOP_L(k_dr_step, dim3(max_b, n), dim3(64), d, ids);
        MRPOP_L(k_dr_scan<true>, dim3(groups * splits, n), dim3(64), d, ids, splits);
        MRPOP_L(k_dr_cand, dim3(max_b, n), dim3(64), d, ids);
    }
    MRPOP_L(k_dr_claim, dim3(groups, n), dim3(64), d, ids);
}
void launch_dr_commit(SlotDev *d, const int *ids, int n, hipStream_t stream) {
    MRPOP_L(k_dr_commit, dim3(1, n), dim3(256), d, ids);
}
void launch_dr_connect(SlotDev *d, const int *ids, int n, int max_b, int splits, hipStream_t stream) {
    if (splits > 0) {
        const int groups = (max_b + 63) / 64;
        MRPOP_L(k_cn_scan, dim3(groups * splits, n), dim3(64), d, ids, splits);
    }
    MRPOP_L(k_dr_connect, dim3(max_b, n), dim3(64), d, ids, splits > 0 ? 1 : 0);
}
void launch_dr_finish(SlotDev *d, const int *ids, int n, hipStream_t stream) {
    MRPOP_L(k_dr_finish, dim3(1, n), dim3(256), d, ids);
}
void launch_sp_edges(SlotDev *d, const int *ids, int n, int stage, int bspline_step, float range, float min_change,
                     hipStream_t stream) {
    MRPOP_L(k_sp_edges, dim3(1, n), dim3(1024), d, ids, stage, bspline_step, range, min_change);
}
void launch_sp_collect(SlotDev *d, const int *ids, int n, hipStream_t stream) {
    MRPOP_L(k_sp_collect, dim3(16, n), dim3(256), d, ids);
}
void launch_sp_smooth_apply(SlotDev *d, const int *ids, int n, hipStream_t stream) {
    MRPOP_L(k_sp_smooth_apply, dim3(16, n), dim3(256), d, ids);
}
void launch_sp_shortcut(SlotDev *d, const int *ids, int n, hipStream_t stream) {
    MRPOP_L(k_sp_shortcut, dim3(1, n), dim3(64), d, ids);
}
void launch_report(SlotDev *d, const int *ids, int n, void *out, hipStream_t stream) {
    MRPOP_L(k_report, dim3((n + 63) / 64), dim3(64), d, ids, n, (ReportDev *)out);
}
#undef MRPOP_L

}  // namespace plan
}  // namespace mrpop
