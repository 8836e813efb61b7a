// capi.hip -- C ABI of libmrpop.so: contexts, uploaded worlds, edge checks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "internal.h"

// Staging array of mrpop_check_edges_noalloc, one copy per device.  Host code
// never uses its name as an address: each device's copy is looked up once
// with hipGetSymbolAddress (scratch_address, called from mrpop_create) and
// that pointer is what kernels, copies and fills receive.
static __device__ __align__(256) unsigned char scratch[MRPOP_NOALLOC_SCRATCH_BYTES];

namespace mrpop {

int ensure(Context *c, DevBuf &b, size_t bytes) {
    if (bytes <= b.bytes) return MRPOP_OK;
    if (b.ptr) {
        MRPOP_HIP(c, hipStreamSynchronize(c->stream));
        MRPOP_HIP(c, hipFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = std::max(bytes, (size_t)4096);
    want = std::max(want, b.bytes * 2);
    MRPOP_HIP(c, hipMalloc(&b.ptr, want));
    b.bytes = want;
    return MRPOP_OK;
}

namespace {

const char *kVersion = "mrpop-mi355x 0.2 (gfx950)";

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

// Device address of this device's scratch array (the device must be
// current); looked up on the first call per device.
int scratch_address(int device, void **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return MRPOP_ERR_INVALID;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MRPOP_ERR_HIP;
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return MRPOP_OK;
}

// Edge verdict words to bytes (0 or 1), on the device.
__global__ void verdict_bytes_kernel(const uint32_t *in, uint8_t *out, int n) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) out[e] = in[e] != 0u;
}

bool is_world_axes(const float *rec) {
    // cuboid record: axes at 3..11; axis-aligned iff they are exactly e_x, e_y, e_z
    static const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int i = 0; i < 9; ++i)
        if (rec[3 + i] != I[i]) return false;
    return true;
}

}  // namespace
}  // namespace mrpop

namespace mrpop {

namespace {

// Box of a segment swept by a ball of radius r.
void seg_box(const float *p, const float *q, float r, float *box) {
    for (int i = 0; i < 3; ++i) {
        box[i] = std::min(p[i], q[i]) - r;
        box[3 + i] = std::max(p[i], q[i]) + r;
    }
}

// Box of {c + d : |A d|_k <= e_k for k = 0..2} where the rows of A are the
// test's axes (sphere_obb / sphere_zcuboid): |d_i| <= sum_k |A^-1[i][k]| e_k.
// *f receives max_i sum_k |A^-1[i][k]|, the growth of the box per unit of
// sphere radius.  A singular A gives an unbounded box (no culling).
void axes_box(const float *c, const double A[3][3], const double e[3], float *box, float *f) {
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                       A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                       A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    if (!(std::fabs(det) > 1e-9)) {
        for (int i = 0; i < 3; ++i) {
            box[i] = -3.0e38f;
            box[3 + i] = 3.0e38f;
        }
        *f = 0.f;
        return;
    }
    double inv[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int i1 = (j + 1) % 3, i2 = (j + 2) % 3, j1 = (i + 1) % 3, j2 = (i + 2) % 3;
            inv[i][j] = (A[i1][j1] * A[i2][j2] - A[i1][j2] * A[i2][j1]) / det;  // adjugate / det
        }
    double fmax = 0;
    for (int i = 0; i < 3; ++i) {
        double ext = 0, row = 0;
        for (int k = 0; k < 3; ++k) {
            ext += std::fabs(inv[i][k]) * e[k];
            row += std::fabs(inv[i][k]);
        }
        fmax = std::max(fmax, row);
        box[i] = (float)(c[i] - ext);
        box[3 + i] = (float)(c[i] + ext);
    }
    *f = (float)(fmax * (1.0 + 1e-6));
}

}  // namespace

long pack_world(const mrpop_scene &s, const float *b, int n_arms, dev::WorldDesc &d, std::vector<float> &host) {
    const int counts[5] = {s.n_spheres, s.n_capsules, s.n_z_capsules, s.n_cuboids, s.n_z_cuboids};
    long total = 0;
    for (int k = 0; k < 5; ++k) {
        if (counts[k] < 0) return -1;
        total += counts[k];
    }
    // split cuboids into world-axis-aligned boxes and general ones
    std::vector<float> aabb, obb;
    for (int i = 0; i < s.n_cuboids; ++i) {
        const float *r = s.cuboids + 15 * i;
        if (is_world_axes(r)) {
            aabb.insert(aabb.end(), {r[0], r[1], r[2], r[12], r[13], r[14]});
        } else {
            obb.insert(obb.end(), r, r + 15);
        }
    }
    host.clear();
    auto put = [&](const float *p, size_t n) { host.insert(host.end(), p, p + n); };
    d.o_sph = (int)host.size();
    d.n_sph = s.n_spheres;
    if (d.n_sph) put(s.spheres, 4 * (size_t)d.n_sph);
    d.o_cap = (int)host.size();
    d.n_cap = s.n_capsules;
    if (d.n_cap) put(s.capsules, 8 * (size_t)d.n_cap);
    d.o_zcap = (int)host.size();
    d.n_zcap = s.n_z_capsules;
    if (d.n_zcap) put(s.z_capsules, 8 * (size_t)d.n_zcap);
    d.o_aabb = (int)host.size();
    d.n_aabb = (int)(aabb.size() / 6);
    put(aabb.data(), aabb.size());
    d.o_obb = (int)host.size();
    d.n_obb = (int)(obb.size() / 15);
    put(obb.data(), obb.size());
    d.o_zcub = (int)host.size();
    d.n_zcub = s.n_z_cuboids;
    if (d.n_zcub) put(s.z_cuboids, 15 * (size_t)d.n_zcub);
    d.o_base = (int)host.size();
    put(b, 12 * (size_t)n_arms);
    d.n_arms = n_arms;
    // cull boxes with a margin far above the tests' rounding
    d.o_box = (int)host.size();
    float box[8];
    auto emit = [&](float f) {
        for (int i = 0; i < 3; ++i) {
            const float m = 1e-4f + 1e-6f * std::max(std::fabs(box[i]), std::fabs(box[3 + i]));
            box[i] -= m;
            box[3 + i] += m;
        }
        box[6] = f;
        box[7] = 0.f;
        put(box, 8);
    };
    for (int i = 0; i < d.n_sph; ++i) {
        const float *o = s.spheres + 4 * i;
        seg_box(o, o, std::fabs(o[3]), box);
        emit(1.f);
    }
    for (int i = 0; i < d.n_cap; ++i) {  // segment p1 .. p1 + v (utils.cuh:148-158 clamps t to [0, 1])
        const float *o = s.capsules + 8 * i;
        const float q[3] = {o[0] + o[3], o[1] + o[4], o[2] + o[5]};
        seg_box(o, q, std::fabs(o[6]), box);
        emit(1.f);
    }
    for (int i = 0; i < d.n_zcap; ++i) {
        const float *o = s.z_capsules + 8 * i;
        const float q[3] = {o[0], o[1], o[2] + o[5]};
        seg_box(o, q, std::fabs(o[6]), box);
        emit(1.f);
    }
    for (size_t i = 0; i < aabb.size() / 6; ++i) {
        const float *o = aabb.data() + 6 * i;
        for (int k = 0; k < 3; ++k) {
            box[k] = o[k] - std::fabs(o[3 + k]);
            box[3 + k] = o[k] + std::fabs(o[3 + k]);
        }
        emit(1.f);
    }
    for (size_t i = 0; i < obb.size() / 15; ++i) {
        const float *o = obb.data() + 15 * i;
        const double A[3][3] = {{o[3], o[4], o[5]}, {o[6], o[7], o[8]}, {o[9], o[10], o[11]}};
        const double e[3] = {std::fabs((double)o[12]), std::fabs((double)o[13]), std::fabs((double)o[14])};
        float f;
        axes_box(o, A, e, box, &f);
        emit(f);
    }
    for (int i = 0; i < d.n_zcub; ++i) {
        const float *o = s.z_cuboids + 15 * i;
        const double A[3][3] = {{o[3], o[4], 0.0}, {o[6], o[7], 0.0}, {0.0, 0.0, 1.0}};
        const double e[3] = {std::fabs((double)o[12]), std::fabs((double)o[13]), std::fabs((double)o[14])};
        float f;
        axes_box(o, A, e, box, &f);
        emit(f);
    }
    return total;
}

}  // namespace mrpop


using namespace mrpop;

struct mrpop_ctx : Context {};
struct mrpop_world : World {};

extern "C" {

const char *mrpop_version(void) { return kVersion; }

int mrpop_settings_default(mrpop_settings *s) {
    if (!s) return MRPOP_ERR_INVALID;
    // pop_settings.hh:4-31
    s->max_samples = 1000000;
    s->max_iters = 1000;
    s->optimize_iters = 400000;
    s->num_new_configs = 256;
    s->granularity = 16;
    s->range = 0.5f;
    s->balance = 2;
    s->tree_ratio = 1.0f;
    s->dynamic_domain = 0;
    s->dd_alpha = 0.0001f;
    s->dd_radius = 4.0f;
    s->dd_min_radius = 1.0f;
    s->roadmap_reassign = 0;
    s->rrtc_iter = 1;
    s->resample_iter = 5;
    s->cost_bound_resample = 0;
    s->path_simplify = 1;
    s->phs = 1;
    s->simplify_round = 1;
    s->bspline_step = 3;
    s->bspline_min_change = 0.01f;
    s->perturb_step = 3;
    s->perturb_empty_step = 2;
    return MRPOP_OK;
}

int mrpop_default_bases(int n_arms, float *out) {
    const float(*b)[12] = geom::default_bases(n_arms);
    if (!b || !out) return MRPOP_ERR_INVALID;
    std::memcpy(out, b, sizeof(float) * 12 * n_arms);
    return MRPOP_OK;
}

int mrpop_create(int device, mrpop_ctx **out) {
    if (!out) return MRPOP_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MRPOP_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return MRPOP_ERR_INVALID;
    auto *c = new (std::nothrow) mrpop_ctx();
    if (!c) return MRPOP_ERR_BUDGET;
    c->device = device;
    int rc = MRPOP_OK;
    do {
        if (hipSetDevice(device) != hipSuccess) { rc = MRPOP_ERR_HIP; break; }
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { rc = MRPOP_ERR_HIP; break; }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess) { rc = MRPOP_ERR_HIP; break; }
        c->num_cus = prop.multiProcessorCount;
        if ((rc = ensure(c, c->counters, 256))) break;
        if (hipMemset(c->counters.ptr, 0, c->counters.bytes) != hipSuccess) { rc = MRPOP_ERR_HIP; break; }
        // this device's static staging array (mrpop_check_edges_noalloc)
        if ((rc = scratch_address(device, &c->scratch))) break;
    } while (0);
    if (rc) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return MRPOP_OK;
}

int mrpop_destroy(mrpop_ctx *c) {
    if (!c) return MRPOP_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (DevBuf *b : {&c->counters, &c->from, &c->to, &c->out_edge, &c->out_mask, &c->cfg, &c->out_a,
                      &c->out_b})
        if (b->ptr) (void)hipFree(b->ptr);
    for (auto &kv : c->rm_cache.entries) (void)hipFree(kv.second.mem);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MRPOP_OK;
}

const char *mrpop_last_error(const mrpop_ctx *c) { return c ? c->err.c_str() : "null context"; }

void *mrpop_stream(mrpop_ctx *c) { return c ? (void *)c->stream : nullptr; }

int mrpop_world_create(mrpop_ctx *c, const mrpop_scene *s, int n_arms, const float *bases, mrpop_world **out) {
    if (!c || !s || !out) return MRPOP_ERR_INVALID;
    *out = nullptr;
    if (n_arms < 1 || n_arms > geom::kMaxArms) return c->fail(MRPOP_ERR_INVALID, "n_arms %d outside 1..5", n_arms);
    const int counts[5] = {s->n_spheres, s->n_capsules, s->n_z_capsules, s->n_cuboids, s->n_z_cuboids};
    long total = 0;
    for (int k = 0; k < 5; ++k) {
        if (counts[k] < 0) return c->fail(MRPOP_ERR_INVALID, "negative obstacle count");
        total += counts[k];
    }
    if (total > MRPOP_MAX_OBSTACLES) return c->fail(MRPOP_ERR_INVALID, "%ld obstacles > %d", total, MRPOP_MAX_OBSTACLES);
    const float *b = bases;
    if (!b) {
        if (n_arms < 2) return c->fail(MRPOP_ERR_INVALID, "a 1-arm world needs explicit base poses");
        b = &geom::default_bases(n_arms)[0][0];
    }
    std::vector<float> host;
    auto *w = new (std::nothrow) mrpop_world();
    if (!w) return MRPOP_ERR_BUDGET;
    dev::WorldDesc &d = w->desc;
    pack_world(*s, b, n_arms, d, host);
    w->ctx = c;
    w->n_arms = n_arms;
    std::memcpy(w->bases, b, sizeof(float) * 12 * n_arms);
    w->n_obstacles = (int)total;
    // canonical flop weights per obstacle: cuboid 37, capsule 30, sphere 12
    w->f_obstacles = 37.0 * (s->n_cuboids + s->n_z_cuboids) + 30.0 * (s->n_capsules + s->n_z_capsules) +
                     12.0 * s->n_spheres;
    if (hipSetDevice(c->device) != hipSuccess ||
        hipMalloc(&w->buf, host.size() * sizeof(float)) != hipSuccess ||
        hipMemcpy(w->buf, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        if (w->buf) (void)hipFree(w->buf);
        delete w;
        return c->fail(MRPOP_ERR_HIP, "world upload failed");
    }
    d.obs = w->buf;
    if (hipMalloc(&w->d_desc, sizeof(dev::WorldDesc)) != hipSuccess ||
        hipMemcpy(w->d_desc, &d, sizeof(dev::WorldDesc), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(w->buf);
        if (w->d_desc) (void)hipFree(w->d_desc);
        delete w;
        return c->fail(MRPOP_ERR_HIP, "world upload failed");
    }
    *out = w;
    return MRPOP_OK;
}

int mrpop_world_destroy(mrpop_world *w) {
    if (!w) return MRPOP_OK;
    // hipFree waits for the device; the context may already be gone
    if (w->buf) (void)hipFree(w->buf);
    if (w->d_desc) (void)hipFree(w->d_desc);
    delete w;
    return MRPOP_OK;
}

int mrpop_check_edges_device(mrpop_ctx *c, const mrpop_world *w, int mode, int arm, int granularity, int n_edges,
                             const float *d_from, const float *d_to, uint32_t *d_out_edge, uint64_t *d_out_mask,
                             void *stream) {
    if (!c || !w || (n_edges > 0 && (!d_from || !d_to || !d_out_edge))) return MRPOP_ERR_INVALID;
    if (n_edges < 0) return c->fail(MRPOP_ERR_INVALID, "negative edge count");
    return launch_check(c, w, mode, arm, granularity, n_edges, d_from, d_to, d_out_edge, d_out_mask,
                        stream ? (hipStream_t)stream : c->stream);
}

int mrpop_check_edges(mrpop_ctx *c, const mrpop_world *w, int mode, int arm, int granularity, int n_edges,
                      const float *from, const float *to, uint8_t *out_edge, uint64_t *out_mask) {
    if (!c || !w || (n_edges > 0 && (!from || !to || !out_edge))) return MRPOP_ERR_INVALID;
    if (n_edges < 0) return c->fail(MRPOP_ERR_INVALID, "negative edge count");
    if (n_edges == 0) return MRPOP_OK;
    const int A = mode == MRPOP_MODE_SINGLE ? 1 : w->n_arms;
    const size_t cfg_bytes = (size_t)n_edges * 7 * A * sizeof(float);
    MRPOP_HIP(c, hipSetDevice(c->device));
    if (int rc = ensure(c, c->from, cfg_bytes)) return rc;
    if (int rc = ensure(c, c->to, cfg_bytes)) return rc;
    if (int rc = ensure(c, c->out_edge, (size_t)n_edges * 4)) return rc;
    if (out_mask)
        if (int rc = ensure(c, c->out_mask, (size_t)n_edges * 8)) return rc;
    MRPOP_HIP(c, hipMemcpyAsync(c->from.ptr, from, cfg_bytes, hipMemcpyHostToDevice, c->stream));
    MRPOP_HIP(c, hipMemcpyAsync(c->to.ptr, to, cfg_bytes, hipMemcpyHostToDevice, c->stream));
    if (int rc = launch_check(c, w, mode, arm, granularity, n_edges, (const float *)c->from.ptr,
                              (const float *)c->to.ptr, (uint32_t *)c->out_edge.ptr,
                              out_mask ? (uint64_t *)c->out_mask.ptr : nullptr, c->stream))
        return rc;
    std::vector<uint32_t> tmp(n_edges);
    MRPOP_HIP(c, hipMemcpyAsync(tmp.data(), c->out_edge.ptr, (size_t)n_edges * 4, hipMemcpyDeviceToHost, c->stream));
    if (out_mask)
        MRPOP_HIP(c, hipMemcpyAsync(out_mask, c->out_mask.ptr, (size_t)n_edges * 8, hipMemcpyDeviceToHost, c->stream));
    MRPOP_HIP(c, hipStreamSynchronize(c->stream));
    for (int e = 0; e < n_edges; ++e) out_edge[e] = tmp[e] != 0;
    return MRPOP_OK;
}

int mrpop_check_edges_noalloc(mrpop_ctx *c, const mrpop_world *w, int mode, int arm, int granularity, int n_edges,
                              const float *from, const float *to, uint8_t *out_edge, uint64_t *out_mask) {
    if (!c || !w || (n_edges > 0 && (!from || !to || !out_edge))) return MRPOP_ERR_INVALID;
    if (n_edges < 0) return c->fail(MRPOP_ERR_INVALID, "negative edge count");
    if (n_edges == 0) return MRPOP_OK;
    if (!c->scratch) return c->fail(MRPOP_ERR_STATE, "the context has no staging array");
    const int A = mode == MRPOP_MODE_SINGLE ? 1 : w->n_arms;
    // layout of the staging array: from, to, verdict words, masks, verdict bytes
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t cfg_bytes = (size_t)n_edges * 7 * A * sizeof(float);
    const size_t o_to = up(cfg_bytes), o_word = o_to + up(cfg_bytes), o_mask = o_word + up((size_t)n_edges * 4);
    const size_t o_byte = o_mask + (out_mask ? up((size_t)n_edges * 8) : 0), total = o_byte + up((size_t)n_edges);
    if (total > MRPOP_NOALLOC_SCRATCH_BYTES)
        return c->fail(MRPOP_ERR_BUDGET, "%d edges need %zu bytes of staging, more than the %zu-byte array", n_edges,
                       total, (size_t)MRPOP_NOALLOC_SCRATCH_BYTES);
    char *base = (char *)c->scratch;  // device address of this device's array
    float *d_from = (float *)base, *d_to = (float *)(base + o_to);
    uint32_t *d_word = (uint32_t *)(base + o_word);
    uint64_t *d_mask = out_mask ? (uint64_t *)(base + o_mask) : nullptr;
    uint8_t *d_byte = (uint8_t *)(base + o_byte);
    MRPOP_HIP(c, hipSetDevice(c->device));
    MRPOP_HIP(c, hipMemcpyAsync(d_from, from, cfg_bytes, hipMemcpyHostToDevice, c->stream));
    MRPOP_HIP(c, hipMemcpyAsync(d_to, to, cfg_bytes, hipMemcpyHostToDevice, c->stream));
    // the check clears the verdict words and masks with a kernel before it ORs into them
    if (int rc = launch_check(c, w, mode, arm, granularity, n_edges, d_from, d_to, d_word, d_mask, c->stream))
        return rc;
    hipLaunchKernelGGL(verdict_bytes_kernel, dim3(std::min(1024, (n_edges + 255) / 256)), dim3(256), 0, c->stream,
                       d_word, d_byte, n_edges);
    MRPOP_HIP(c, hipGetLastError());
    MRPOP_HIP(c, hipMemcpyAsync(out_edge, d_byte, (size_t)n_edges, hipMemcpyDeviceToHost, c->stream));
    if (out_mask)
        MRPOP_HIP(c, hipMemcpyAsync(out_mask, d_mask, (size_t)n_edges * 8, hipMemcpyDeviceToHost, c->stream));
    MRPOP_HIP(c, hipStreamSynchronize(c->stream));
    return MRPOP_OK;
}

int mrpop_sphere_centres(mrpop_ctx *c, const mrpop_world *w, int n_cfg, const float *cfgs, float *out_coarse,
                         float *out_fine) {
    if (!c || !w || n_cfg < 0 || (n_cfg > 0 && !cfgs)) return MRPOP_ERR_INVALID;
    if (n_cfg == 0) return MRPOP_OK;
    const int A = w->n_arms;
    MRPOP_HIP(c, hipSetDevice(c->device));
    const size_t cb = (size_t)n_cfg * 7 * A * sizeof(float);
    const size_t ob = (size_t)n_cfg * A * geom::kCoarseSpheres * 3 * sizeof(float);
    const size_t fb = (size_t)n_cfg * A * geom::kFineSpheres * 3 * sizeof(float);
    if (int rc = ensure(c, c->cfg, cb)) return rc;
    if (int rc = ensure(c, c->out_a, ob)) return rc;
    if (int rc = ensure(c, c->out_b, fb)) return rc;
    MRPOP_HIP(c, hipMemcpyAsync(c->cfg.ptr, cfgs, cb, hipMemcpyHostToDevice, c->stream));
    if (int rc = launch_centres(c, w, n_cfg, (const float *)c->cfg.ptr, (float *)c->out_a.ptr, (float *)c->out_b.ptr,
                                c->stream))
        return rc;
    if (out_coarse) MRPOP_HIP(c, hipMemcpyAsync(out_coarse, c->out_a.ptr, ob, hipMemcpyDeviceToHost, c->stream));
    if (out_fine) MRPOP_HIP(c, hipMemcpyAsync(out_fine, c->out_b.ptr, fb, hipMemcpyDeviceToHost, c->stream));
    MRPOP_HIP(c, hipStreamSynchronize(c->stream));
    return MRPOP_OK;
}

}  // extern "C"

// Diagnostics (not part of the public ABI): run a host-buffer edge check and
// return 8 words per sample of per-phase results of the fused kernel.
extern "C" int mrpop_diag_check_edges(mrpop_ctx *c, const mrpop_world *w, int mode, int arm, int granularity,
                                      int n_edges, const float *from, const float *to, uint8_t *out_edge,
                                      uint32_t *out_diag) {
    if (!c || !w || !out_diag || n_edges <= 0) return MRPOP_ERR_INVALID;
    const size_t words = (size_t)n_edges * granularity * 8;
    void *d = nullptr;
    MRPOP_HIP(c, hipMalloc(&d, words * 4));
    MRPOP_HIP(c, hipMemset(d, 0, words * 4));
    c->diag = d;
    std::vector<uint64_t> mask(n_edges);
    int rc = mrpop_check_edges(c, w, mode, arm, granularity, n_edges, from, to, out_edge, mask.data());
    c->diag = nullptr;
    if (rc == MRPOP_OK) rc = hipMemcpy(out_diag, d, words * 4, hipMemcpyDeviceToHost) == hipSuccess ? MRPOP_OK : MRPOP_ERR_HIP;
    (void)hipFree(d);
    return rc;
}

// Diagnostics (not part of the public ABI): a host-buffer edge check (edge
// verdicts only) run as a probe pass of `probe` samples per edge followed by
// the remaining samples of the edges the probe left free, as the planner
// checks its edges.
extern "C" int mrpop_diag_check_probe(mrpop_ctx *c, const mrpop_world *w, int mode, int arm, int granularity,
                                      int n_edges, const float *from, const float *to, int probe, uint8_t *out_edge) {
    if (!c || !w || probe < 1 || probe > 64) return MRPOP_ERR_INVALID;
    c->probe = probe;
    const int rc = mrpop_check_edges(c, w, mode, arm, granularity, n_edges, from, to, out_edge, nullptr);
    c->probe = 0;
    return rc;
}

// Diagnostics (not part of the public ABI): the cull boxes the checker uses
// for a scene, 8 floats per obstacle in the checker's order (spheres,
// capsules, z-capsules, world-axis cuboids, other cuboids, z-cuboids).
// Returns the number of obstacles, or -1.
extern "C" int mrpop_diag_world_boxes(const mrpop_scene *s, float *out, int cap) {
    if (!s) return -1;
    std::vector<float> host;
    mrpop::dev::WorldDesc d{};
    static const float bases[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    const long n = mrpop::pack_world(*s, bases, 1, d, host);
    if (n < 0) return -1;
    if (out)
        for (long i = 0; i < n && i < cap; ++i)
            for (int k = 0; k < 8; ++k) out[8 * i + k] = host[d.o_box + 8 * i + k];
    return (int)n;
}
