// HIP kernels of the path-tracing hot path and the C ABI of libx3dpt.so (include/x3dpt.h).
//
// Plain form (the drop-in): a first-hit pass caches each pixel's deterministic camera hit
// (pathtracer.cpp:456-472, reused by every sample as in renderer.cpp:121-122), then one
// persistent, grid-strided kernel runs one lane per pixel and loops over that pixel's samples,
// accumulating them in order in registers; a pixel's film value is therefore independent of the
// launch shape and identical to the CPU oracle's.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/x3dpt.h"
#include "objfile.h"
#include "render_common.h"
#include "scene.h"
#include "adjoint.h"
#include "batch.h"
#include "wavefront.h"

namespace x3d {

// ---------------------------------------------------------------------------------------------
// Ray queries.
// ---------------------------------------------------------------------------------------------
struct QueryParams {
    DevScene S;
    const float *rays;
    const float *tmin, *tmax;
    uint32_t n;
    int32_t ref_order;
    x3d_hit *hits;
    uint8_t *occ;
    uint2 *spill;
    uint32_t spill_per_thread;
    Counters *cnt;
};

__global__ __launch_bounds__(kWG) void k_intersect(QueryParams P) {
    __shared__ uint2 lds[kLdsDepth * kWG];
    TravStack<kWG, kLdsDepth> st = make_stack<kWG>(lds, P.spill, P.spill_per_thread);
    uint32_t pops = 0, tests = 0;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < P.n; k += gridDim.x * kWG) {
        const float *r = P.rays + 6ull * k;
        f3 o = mk(r[0], r[1], r[2]), v = mk(r[3], r[4], r[5]);
        ClosestHit h = P.ref_order ? closest_hit<kWG, kLdsDepth, true>(P.S, o, v, st, pops, tests)
                                   : closest_hit<kWG, kLdsDepth, false>(P.S, o, v, st, pops, tests);
        x3d_hit out;
        if (h.prim >= 0) {
            HitRec rec = finish_hit(P.S, h);
            const float4 *tris = static_cast<const float4 *>(P.S.tris);
            out.geo = rec.geo;
            out.tri = static_cast<int32_t>(__float_as_uint(tris[3 * h.prim + 1].w));
            out.prim = h.prim;
            out.t = h.t;
            out.b1 = h.b1;
            out.b2 = h.b2;
            out.vertex[0] = rec.p.x;
            out.vertex[1] = rec.p.y;
            out.vertex[2] = rec.p.z;
            out.normal[0] = rec.n.x;
            out.normal[1] = rec.n.y;
            out.normal[2] = rec.n.z;
        } else {
            out.geo = out.tri = out.prim = -1;
            out.t = __builtin_inff();
            out.b1 = out.b2 = 0.0f;
            out.vertex[0] = out.vertex[1] = out.vertex[2] = 0.0f;
            out.normal[0] = out.normal[1] = out.normal[2] = 0.0f;
        }
        P.hits[k] = out;
    }
    if (P.cnt) {
        flush_counter(&P.cnt->pops, pops);
        flush_counter(&P.cnt->tests, tests);
    }
}

__global__ __launch_bounds__(kWG) void k_occluded(QueryParams P) {
    __shared__ uint2 lds[kLdsDepth * kWG];
    TravStack<kWG, kLdsDepth> st = make_stack<kWG>(lds, P.spill, P.spill_per_thread);
    uint32_t pops = 0, tests = 0;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < P.n; k += gridDim.x * kWG) {
        const float *r = P.rays + 6ull * k;
        f3 o = mk(r[0], r[1], r[2]), v = mk(r[3], r[4], r[5]);
        P.occ[k] = any_hit<kWG, kLdsDepth>(P.S, o, v, P.tmin[k], P.tmax[k], st, pops, tests) ? 1 : 0;
    }
    if (P.cnt) {
        flush_counter(&P.cnt->spops, pops);
        flush_counter(&P.cnt->stests, tests);
    }
}

// ---------------------------------------------------------------------------------------------
// BSDF batches, camera rays, PCG draws.
// ---------------------------------------------------------------------------------------------
struct BsdfParams {
    const DevMaterial *mats;
    int32_t mat;
    uint32_t n;
    const float *nrm, *o, *i;
    float *out_f, *out_i, *out_pdf;
    uint64_t seed;
    int32_t ref_mixture;
    int32_t op; // 0 eval, 1 pdf, 2 sample
};

__global__ void k_bsdf(BsdfParams P) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.n)
        return;
    f3 n = mk(P.nrm[3 * k], P.nrm[3 * k + 1], P.nrm[3 * k + 2]);
    f3 o = mk(P.o[3 * k], P.o[3 * k + 1], P.o[3 * k + 2]);
    if (P.op == 0) {
        f3 i = mk(P.i[3 * k], P.i[3 * k + 1], P.i[3 * k + 2]);
        f3 f = bsdf_eval(P.mats, P.mat, n, o, i);
        P.out_f[3 * k] = f.x;
        P.out_f[3 * k + 1] = f.y;
        P.out_f[3 * k + 2] = f.z;
    } else if (P.op == 1) {
        f3 i = mk(P.i[3 * k], P.i[3 * k + 1], P.i[3 * k + 2]);
        P.out_pdf[k] = bsdf_pdf(P.mats, P.mat, n, o, i);
    } else {
        Pcg g = pcg_for(P.seed, k, 0);
        float d;
        f3 i = bsdf_sample(P.mats, P.mat, n, o, g, &d, P.ref_mixture != 0);
        P.out_i[3 * k] = i.x;
        P.out_i[3 * k + 1] = i.y;
        P.out_i[3 * k + 2] = i.z;
        P.out_pdf[k] = d;
    }
}

__global__ void k_camera_rays(CamParams c, uint32_t w, uint32_t h, float *rays) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= w * h)
        return;
    f3 o, v;
    camera_ray(c, k % w, k / w, w, h, o, v);
    float *r = rays + 6ull * k;
    r[0] = o.x;
    r[1] = o.y;
    r[2] = o.z;
    r[3] = v.x;
    r[4] = v.y;
    r[5] = v.z;
}

__global__ void k_zero_u64(unsigned long long *p, uint32_t n) {
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x)
        p[k] = 0;
}

__global__ void k_pcg_draws(uint64_t seed, uint64_t pixel, uint32_t sample, uint32_t n, uint32_t *out) {
    if (blockIdx.x != 0 || threadIdx.x != 0)
        return;
    Pcg g = pcg_for(seed, pixel, sample);
    for (uint32_t k = 0; k < n; k++)
        out[k] = g.next();
}

// ---------------------------------------------------------------------------------------------
// Rendering.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kWG) void k_first_hit(RenderParams P) {
    __shared__ uint2 lds[kLdsDepth * kWG];
    TravStack<kWG, kLdsDepth> st = make_stack<kWG>(lds, P.spill, P.spill_per_thread);
    uint32_t pops = 0, tests = 0, nq = 0;
    uint32_t npix = P.width * P.height;
    for (uint32_t pix = blockIdx.x * kWG + threadIdx.x; pix < npix; pix += gridDim.x * kWG) {
        f3 o, v;
        pixel_ray(P, pix, o, v);
        nq++;
        ClosestHit h = closest_hit<kWG, kLdsDepth, false>(P.S, o, v, st, pops, tests);
        float4 a = make_float4(0, 0, 0, 0), b = make_float4(0, 0, 0, 0), c = make_float4(0, 0, 0, 0);
        if (h.prim >= 0) {
            HitRec r = finish_hit(P.S, h);
            // compute_first_hit drops hits seen from behind (pathtracer.cpp:463-464).
            if (!(dot(r.n, -v) <= 0.0f)) {
                a = make_float4(r.p.x, r.p.y, r.p.z, r.t);
                b = make_float4(r.n.x, r.n.y, r.n.z, __uint_as_float(static_cast<uint32_t>(r.geo)));
                c = make_float4(__uint_as_float(static_cast<uint32_t>(r.mat)),
                                __uint_as_float(static_cast<uint32_t>(r.light)), 1.0f, 0.0f);
            }
        }
        P.first[3 * pix] = a;
        P.first[3 * pix + 1] = b;
        P.first[3 * pix + 2] = c;
    }
    if (P.cnt) {
        flush_counter(&P.cnt->closest, nq);
        flush_counter(&P.cnt->pops, pops);
        flush_counter(&P.cnt->tests, tests);
    }
}

constexpr uint64_t kGenItems = 1u << 24; // (pixel, sample) items per generation: 256 MiB of sample radiance

// Bidirectional tracer. Work items are the (pixel, sample) pairs of one generation of samples,
// pixel-major within a sample so neighbouring lanes start from neighbouring first hits. A lane
// whose sample is done takes the next unclaimed item from a global counter, so lanes stay busy
// although subpath lengths and connection counts vary widely between samples. Each sample's
// radiance is stored; k_sum_samples then adds every pixel's samples in sample order, as the
// oracle does.
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(4))) void k_render_bdpt(RenderParams P, uint32_t gen_begin, uint32_t gen_samples,
                                                     float4 *samples, uint32_t *next_item) {
    __shared__ uint2 lds[kLdsDepth * kWG];
    TravStack<kWG, kLdsDepth> st = make_stack<kWG>(lds, P.spill, P.spill_per_thread);
    Tracer<kWG, kLdsDepth, false> tr{P.S, st};
    // this wave's block: both subpaths of its 64 lanes (a wave-uniform base)
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * kWG + threadIdx.x) >> 6);
    const PathletRows cam{P.subpaths + static_cast<size_t>(wave) * (2 * kPathletRows * kMaxPathLen * 64u)};
    const PathletRows lig{cam.base + kPathletRows * kMaxPathLen * 64u};
    uint32_t npix = P.width * P.height;
    uint32_t total = npix * gen_samples;
    uint32_t grid_threads = gridDim.x * kWG;
    for (uint32_t q = blockIdx.x * kWG + threadIdx.x; q < total; q = grid_threads + atomicAdd(next_item, 1u)) {
        uint32_t pix = q % npix, s = gen_begin + q / npix;
        f3 o, v;
        pixel_ray(P, pix, o, v);
        bool valid;
        HitRec first = load_first(P, pix, valid);
        Pcg g = pcg_for(P.seed, pix, s);
        f3 L = tr.li_bdpt(first, valid, v, g, P.max_path_len, cam, lig);
        samples[q] = make_float4(L.x, L.y, L.z, 0.0f);
    }
    if (P.cnt) {
        flush_counter(&P.cnt->closest, tr.nclosest);
        flush_counter(&P.cnt->anyhit, tr.nany);
        flush_counter(&P.cnt->pops, tr.pops);
        flush_counter(&P.cnt->tests, tr.tests);
        flush_counter(&P.cnt->spops, tr.spops);
        flush_counter(&P.cnt->stests, tr.stests);
    }
}

// Plain MIS tracer: work items are the (pixel, sample) pairs of a generation, as in
// k_render_bdpt; a lane whose path ends stores its radiance and starts the next unclaimed item
// at once (per-bounce path regeneration), so lanes stay busy while paths of different lengths
// finish. k_sum_samples then adds every pixel's samples in sample order, as the oracle does.
// (Per-pixel lanes that regenerate the pixel's next sample in place measured slower: 559 vs
// 496 ms at 1024^2 x 64 spp and 204 vs 129 ms at 512^2 x 64 spp on the 1.3M-triangle spheres
// scene; the item order keeps neighbouring lanes on neighbouring pixels of one sample.)
__global__ __launch_bounds__(kWG) void k_render_mis_items(RenderParams P, uint32_t gen_begin, uint32_t gen_samples,
                                                          float4 *samples, uint32_t *next_item) {
    __shared__ uint2 lds[kLdsDepth * kWG];
    TravStack<kWG, kLdsDepth> st = make_stack<kWG>(lds, P.spill, P.spill_per_thread);
    Tracer<kWG, kLdsDepth> tr{P.S, st};
    uint32_t npix = P.width * P.height;
    uint32_t total = npix * gen_samples;
    uint32_t grid_threads = gridDim.x * kWG;
    uint32_t q = blockIdx.x * kWG + threadIdx.x;
    PathState ps;
    auto begin = [&](uint32_t item) {
        uint32_t pix = item % npix, s = gen_begin + item / npix;
        f3 o, v;
        pixel_ray(P, pix, o, v);
        bool valid;
        HitRec first = load_first(P, pix, valid);
        tr.mis_begin(ps, first, valid, v, pcg_for(P.seed, pix, s), P.max_depth);
    };
    if (q < total) {
        begin(q);
        for (;;) {
            if (ps.alive)
                tr.mis_bounce(ps, P.max_depth);
            if (!ps.alive) {
                samples[q] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
                q = grid_threads + atomicAdd(next_item, 1u);
                if (q >= total)
                    break;
                begin(q);
            }
        }
    }
    if (P.cnt) {
        flush_counter(&P.cnt->closest, tr.nclosest);
        flush_counter(&P.cnt->anyhit, tr.nany);
        flush_counter(&P.cnt->pops, tr.pops);
        flush_counter(&P.cnt->tests, tr.tests);
        flush_counter(&P.cnt->spops, tr.spops);
        flush_counter(&P.cnt->stests, tr.stests);
    }
}

// Reference-mode tracers over (pixel, sample) items (for the firefly filter, which needs whole
// sample images): item q = pixel q % npix, sample gen_begin + q / npix, radiance to samples[q].
template <int TRACER>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(4))) void k_render_plain_items(RenderParams P, uint32_t gen_begin, uint32_t gen_samples,
                                                           float4 *samples, uint32_t *next_item) {
    __shared__ uint2 lds[kLdsDepth * kWG];
    TravStack<kWG, kLdsDepth> st = make_stack<kWG>(lds, P.spill, P.spill_per_thread);
    Tracer<kWG, kLdsDepth, false> tr{P.S, st};
    const uint32_t npix = P.width * P.height, total = npix * gen_samples;
    for (uint32_t q = blockIdx.x * kWG + threadIdx.x; q < total; q = gridDim.x * kWG + atomicAdd(next_item, 1u)) {
        uint32_t pix = q % npix, s = gen_begin + q / npix;
        f3 o, v;
        pixel_ray(P, pix, o, v);
        bool valid;
        HitRec first = load_first(P, pix, valid);
        Pcg g = pcg_for(P.seed, pix, s);
        f3 L;
        if (TRACER == X3D_PT_UNIDIRECT_LT1)
            L = tr.li_lt1(first, valid, v, g);
        else if (TRACER == X3D_PT_UNIDIRECT)
            L = tr.li_unidirect(first, valid, v, g);
        else if (TRACER == X3D_PT_DIRECT)
            L = tr.li_direct(first, valid, v, g);
        else if (TRACER == X3D_PT_BIDIRECT_LT2)
            L = tr.li_lt2(first, valid, v, g);
        else if (TRACER == X3D_PT_POSITION)
            L = tr.li_position(first, valid);
        else // X3D_PT_NORMAL (normal_tracer, pathtracer.cpp:492-506)
            L = valid ? (first.n + mk(1.0f, 1.0f, 1.0f)) / 2.0f : mk(0.0f, 0.0f, 0.0f);
        samples[q] = make_float4(L.x, L.y, L.z, 0.0f);
    }
    if (P.cnt) {
        flush_counter(&P.cnt->closest, tr.nclosest);
        flush_counter(&P.cnt->anyhit, tr.nany);
        flush_counter(&P.cnt->pops, tr.pops);
        flush_counter(&P.cnt->tests, tr.tests);
        flush_counter(&P.cnt->spops, tr.spops);
        flush_counter(&P.cnt->stests, tr.stests);
    }
}

// Adds samples [gen_begin, gen_begin + gen_samples) of every pixel, in order, to the running sum
// (P.acc, or zero for the first generation of a render starting at sample 0); the last
// generation writes the scaled film.
__global__ __launch_bounds__(256) void k_sum_samples(RenderParams P, uint32_t gen_begin, uint32_t gen_samples,
                                                     const float4 *samples) {
    uint32_t npix = P.width * P.height;
    bool last = gen_begin + gen_samples >= P.s_end;
    float scale = 1.0f / static_cast<float>(P.spp_total);
    for (uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x; pix < npix; pix += gridDim.x * blockDim.x) {
        f3 acc = mk(0.0f, 0.0f, 0.0f);
        if (gen_begin > 0) {
            float4 q = P.acc[pix];
            acc = mk(q.x, q.y, q.z);
        }
        if (P.firefly) {
            const uint32_t x = pix % P.width, y = pix / P.width;
            for (uint32_t k = 0; k < gen_samples; k++)
                acc = acc + xyz(firefly_filtered(samples + static_cast<size_t>(k) * npix, x, y, P.width, P.height));
        } else {
            for (uint32_t k = 0; k < gen_samples; k++)
                acc = acc + xyz(samples[static_cast<size_t>(k) * npix + pix]);
        }
        if (last && P.finalize)
            P.out[pix] = make_float4(scale * acc.x, scale * acc.y, scale * acc.z, scale);
        else
            P.acc[pix] = make_float4(acc.x, acc.y, acc.z, 0.0f);
    }
}

// Sample radiance of a generation (float4 per item) to packed rgb (x3d_trace_rays).
__global__ __launch_bounds__(256) void k_pack_rgb(const float4 *samples, uint64_t n, float *out) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) {
        float4 q = samples[i];
        out[3 * i] = q.x;
        out[3 * i + 1] = q.y;
        out[3 * i + 2] = q.z;
    }
}

// ---------------------------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------------------------
thread_local std::string g_last_error;

int fail(int code, std::string const &msg) {
    g_last_error = msg;
    return code;
}

#define X3D_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(X3D_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

struct DeviceInfo {
    int ok = -1;
    int cus = 0;
    std::string err;
};

DeviceInfo &device_info() {
    static DeviceInfo info;
    static std::once_flag once;
    std::call_once(once, [] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0) {
            info.ok = 0;
            info.err = "no HIP device available (libx3dpt requires an AMD GPU)";
            return;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
            info.ok = 0;
            info.err = "hipGetDeviceProperties failed";
            return;
        }
        info.cus = prop.multiProcessorCount;
        info.ok = 1;
    });
    return info;
}

int require_device() {
    DeviceInfo &d = device_info();
    if (d.ok != 1)
        return fail(X3D_ERR_DEVICE, d.err);
    return X3D_OK;
}

// Persistent grid: enough workgroups to fill every CU at the kernel's occupancy.
template <class K> uint32_t persistent_grid(K kernel, uint64_t items) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kWG, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    uint64_t want = (items + kWG - 1) / kWG;
    uint64_t cap = static_cast<uint64_t>(device_info().cus) * static_cast<uint64_t>(per_cu);
    uint64_t g = want < cap ? want : cap;
    return static_cast<uint32_t>(g < 1 ? 1 : g);
}

uint32_t spill_entries(DevScene const &S) { return spill_entries_for(S, kLdsDepth); }

// Fixed working memory of x3d_render_device_noalloc: a file-scope device array that every GPU the
// library runs on holds (its own copy, at its own address) until the process ends. Host code
// never uses the array's name as an address: the per-device address comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr size_t kScratchBytes = size_t(128) << 20;
constexpr int kMaxScratchDevices = 64;
static __device__ __align__(256) char scratch[kScratchBytes];

int scratch_for_current_device(void **out) {
    static std::mutex mu;
    static void *ptr[kMaxScratchDevices] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxScratchDevices)
        return fail(X3D_ERR_DEVICE, "hipGetDevice failed or device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(X3D_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        ptr[dev] = p;
    }
    *out = ptr[dev];
    return X3D_OK;
}

} // namespace x3d

struct x3d_scene {
    x3d::Scene s;
};

struct x3d_renderer {
    x3d_render_options opt;
    x3d::DevBuf first, acc, out, spill, counters, samples, subpaths;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evm = nullptr;
    x3d::WavefrontEngine *wf = nullptr;
    x3d::BatchEngine *batch = nullptr;
    x3d::WavefrontEngine *batch_wf = nullptr; // jobs of a batch that run on their own
    hipEvent_t bev0 = nullptr, bev1 = nullptr;
    std::vector<x3d::DevBuf> batch_films; // device films of x3d_render_batch
    x3d::AdjointEngine *adj = nullptr;
    x3d::DevBuf adj_dl, adj_grad;          // host-pointer variant of x3d_render_backward
    x3d::DevBuf trace_rays, trace_out;     // host-pointer variant of x3d_trace_rays
    ~x3d_renderer() {
        if (adj)
            x3d::adjoint_destroy(adj);
        if (wf)
            x3d::wf_destroy(wf);
        if (batch)
            x3d::batch_destroy(batch);
        if (batch_wf)
            x3d::wf_destroy(batch_wf);
        for (hipEvent_t e : {bev0, bev1})
            if (e)
                (void)hipEventDestroy(e);
        if (evm)
            (void)hipEventDestroy(evm);
        if (ev0)
            (void)hipEventDestroy(ev0);
        if (ev1)
            (void)hipEventDestroy(ev1);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

using namespace x3d;

namespace {

bool finite3(const float *p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

CamParams cam_params(x3d_camera const &c) {
    CamParams p;
    p.sensor = c.sensor_size;
    p.focal = c.focal_length;
    p.aspect = c.aspect;
    std::memcpy(p.m, c.transform, sizeof p.m);
    return p;
}

int check_scene(const x3d_scene *scene) {
    if (!scene)
        return fail(X3D_ERR_INVALID_ARGUMENT, "scene is null");
    if (!scene->s.committed)
        return fail(X3D_ERR_STATE, "scene is not committed");
    return X3D_OK;
}

int run_queries(const x3d_scene *scene, const float *d_rays, const float *d_tmin, const float *d_tmax,
                uint32_t n, int32_t traversal, x3d_hit *d_hits, uint8_t *d_occ, hipStream_t stream,
                x3d::DevBuf &spill) {
    QueryParams P{};
    P.S = scene->s.dev;
    P.rays = d_rays;
    P.tmin = d_tmin;
    P.tmax = d_tmax;
    P.n = n;
    P.ref_order = traversal == X3D_TRAVERSAL_REFERENCE;
    P.hits = d_hits;
    P.occ = d_occ;
    P.spill_per_thread = spill_entries(P.S);
    uint32_t grid = d_hits ? persistent_grid(k_intersect, n) : persistent_grid(k_occluded, n);
    X3D_HIP(spill.ensure(static_cast<size_t>(grid) * kWG * P.spill_per_thread * sizeof(uint2)));
    P.spill = static_cast<uint2 *>(spill.p);
    if (d_hits)
        hipLaunchKernelGGL(k_intersect, dim3(grid), dim3(kWG), 0, stream, P);
    else
        hipLaunchKernelGGL(k_occluded, dim3(grid), dim3(kWG), 0, stream, P);
    X3D_HIP(hipGetLastError());
    return X3D_OK;
}

} // namespace

namespace {

// Renders into d_out; with d_dl (dLoss/dfilm) it runs the adjoint of the MIS tracer instead and
// writes the material gradient to d_grad.
// Wavefront tuning from the renderer's options (0 / DEFAULT fields keep the defaults).
x3d::WavefrontTuning wf_tuning(x3d_render_options const &o) {
    x3d::WavefrontTuning tune;
    tune.tracer = o.pt_type == X3D_PT_UNIDIRECT_LT1 ? x3d::kWfLt1
                  : o.pt_type == X3D_PT_UNIDIRECT   ? x3d::kWfUnidirect
                                                    : x3d::kWfMis;
    if (o.workgroup_size)
        tune.trace_wg = static_cast<int>(o.workgroup_size);
    if (o.queue_capacity)
        tune.paths_in_flight = o.queue_capacity;
    if (o.stack_depth)
        tune.stack_depth = static_cast<int>(o.stack_depth);
    if (o.sort_key != X3D_SORT_DEFAULT)
        tune.sort_by_material = o.sort_key == X3D_SORT_MATERIAL;
    if (o.ray_order != X3D_RAY_ORDER_DEFAULT)
        tune.sort_rays = o.ray_order == X3D_RAY_ORDER_SORTED;
    tune.profile = o.profile;
    return tune;
}

bool has_wavefront(int32_t pt) {
    return pt == X3D_PT_UNIDIRECT_MIS || pt == X3D_PT_UNIDIRECT_LT1 || pt == X3D_PT_UNIDIRECT ||
           pt == X3D_PT_BIDIRECT_MIS;
}

// The smallest render (w * h * spp samples) X3D_FORM_AUTO runs in the wavefront form, per tracer
// and scene size: where the two forms measured equal. The wavefront form wins back the lanes that
// divergent traversal costs the plain kernels, so the fewer triangles a scene has, the more
// samples it takes to pay for its per-bounce launches and queue tails. Measured crossovers (MIS,
// depth 8): 34 triangles 2^24, 30k 2^23, 82k 2^22, 300k and 1.3M 2^20; lt1: 34 to 82k 2^24, 300k
// and 1.3M 2^22; unidirect: 1.3M 2^23; BDPT (6-vertex subpaths): 82k 2^21. Unidirect below 2^18
// triangles and BDPT below 2^16 take one power of two more, as lt1 and MIS do.
uint64_t auto_wavefront_min(int32_t pt, uint32_t tris) {
    int lg;
    switch (pt) {
    case X3D_PT_UNIDIRECT_MIS: lg = tris < (1u << 14) ? 24 : tris < (1u << 16) ? 23 : tris < (1u << 18) ? 22 : 20; break;
    case X3D_PT_UNIDIRECT_LT1: lg = tris < (1u << 18) ? 24 : 22; break;
    case X3D_PT_UNIDIRECT: lg = tris < (1u << 18) ? 24 : 23; break;
    case X3D_PT_BIDIRECT_MIS: lg = tris < (1u << 16) ? 22 : 21; break;
    default: return ~0ull;
    }
    return 1ull << lg;
}

// The wavefront engine a render of w x h x spp runs with, or null for the plain form.
x3d::WavefrontEngine *form_engine(x3d_renderer *r, const x3d_scene *scene, uint32_t w, uint32_t h, uint32_t spp) {
    if (!r->wf || r->opt.form == X3D_FORM_PLAIN)
        return nullptr;
    if (r->opt.form == X3D_FORM_WAVEFRONT)
        return r->wf;
    const uint64_t npix = static_cast<uint64_t>(w) * h;
    if (npix * spp < auto_wavefront_min(r->opt.pt_type, scene->s.dev.num_prims))
        return nullptr;
    // the bidirectional wavefront needs one sample of every pixel in one generation
    if (r->opt.pt_type == X3D_PT_BIDIRECT_MIS && npix > x3d::bd_max_items(r->opt.max_path_len))
        return nullptr;
    return r->wf;
}

// One render. `wf` (with `wf_tune`) selects the wavefront form of the MIS tracer, null the plain
// kernels; d_dl / d_grad select the adjoint pass.
// Ray-list mode of render_impl (x3d_trace_rays): the caller's rays replace the camera's (width =
// ray count, height = 1), samples [s_first, s_first + spp) are drawn, and every sample's
// radiance is written to d_out (rgb, sample-major) instead of being summed into a film.
struct RayList {
    const float *d_rays; // float[6 * n]
    uint32_t s_first;
    float *d_out; // float[3 * n * spp]
};

// Device memory carved out of one fixed block (x3d_render_device_noalloc): bump allocation in
// 256-byte steps, no HIP allocation.
struct ScratchArena {
    char *base;
    size_t bytes, used;
    static size_t up(size_t n) { return (n + 255) & ~static_cast<size_t>(255); }
    void *take(size_t n) {
        size_t a = up(used);
        if (a > bytes || n > bytes - a)
            return nullptr;
        used = a + n;
        return base + a;
    }
    size_t left() const { return up(used) < bytes ? bytes - up(used) : 0; }
};

int render_impl(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width,
                uint32_t height, uint32_t spp, float *d_out, x3d_render_stats *stats,
                x3d::WavefrontEngine *wf = nullptr, x3d::WavefrontTuning const *wf_tune = nullptr,
                const float4 *d_dl = nullptr, float *d_grad = nullptr, bool firefly = false,
                RayList const *rl = nullptr, ScratchArena *arena = nullptr) {
    auto t0 = std::chrono::steady_clock::now();
    RenderParams P{};
    P.S = scene->s.dev;
    if (camera)
        P.cam = cam_params(*camera);
    P.rays = rl ? rl->d_rays : nullptr;
    P.width = width;
    P.height = height;
    P.spp_total = spp;
    P.seed = r->opt.seed;
    P.max_depth = r->opt.max_depth;
    P.max_path_len = r->opt.max_path_len;
    P.spill_per_thread = spill_entries(P.S);
    P.firefly = firefly ? 1 : 0;
    uint64_t npix = static_cast<uint64_t>(width) * height;
    // Working memory: the renderer's growable buffers, or (arena) pieces of a fixed block.
    auto buffer = [&](DevBuf &b, size_t n, void **p) -> int {
        if (arena) {
            *p = arena->take(n);
            if (!*p)
                return fail(X3D_ERR_INVALID_ARGUMENT, "the film needs more than the " +
                                                          std::to_string(arena->bytes >> 20) +
                                                          " MiB of fixed scratch memory");
            return X3D_OK;
        }
        X3D_HIP(b.ensure(n));
        *p = b.p;
        return X3D_OK;
    };
    void *p_first, *p_acc, *p_cnt, *p_spill;
    // counters: [0] first-hit pass, [1] sample kernels, then the pixel work counter
    if (int rc = buffer(r->counters, 2 * sizeof(Counters) + 64, &p_cnt))
        return rc;
    if (int rc = buffer(r->first, npix * 3 * sizeof(float4), &p_first))
        return rc;
    if (int rc = buffer(r->acc, npix * sizeof(float4), &p_acc))
        return rc;
    P.first = static_cast<float4 *>(p_first);
    P.acc = static_cast<float4 *>(p_acc);
    P.out = reinterpret_cast<float4 *>(d_out);
    P.cnt = static_cast<Counters *>(p_cnt);
    int tracer = r->opt.pt_type;
    uint32_t grid_first = persistent_grid(k_first_hit, npix);
    uint32_t grid_render = 1;
    // The plain MIS tracer runs over (pixel, sample) items. The firefly filter works on whole
    // sample images: every tracer then runs over items.
    const bool mis_items = tracer == X3D_PT_UNIDIRECT_MIS && !wf && !d_dl;
    // The reference-mode tracers run over (pixel, sample) items as well (same films; per-pixel
    // lanes measured 1.3-1.9x slower: lt1 98.2 vs 75.0 ms, direct 16.1 vs 8.7 ms at 1024^2 x 16 spp
    // on the 1.3M-triangle spheres scene).
    switch (tracer) {
    case X3D_PT_UNIDIRECT_MIS:
        grid_render = mis_items ? persistent_grid(k_render_mis_items, npix * static_cast<uint64_t>(spp)) : 1;
        break;
    case X3D_PT_UNIDIRECT_LT1:
        grid_render = persistent_grid(k_render_plain_items<X3D_PT_UNIDIRECT_LT1>, npix * spp);
        break;
    case X3D_PT_UNIDIRECT:
        grid_render = persistent_grid(k_render_plain_items<X3D_PT_UNIDIRECT>, npix * spp);
        break;
    case X3D_PT_DIRECT:
        grid_render = persistent_grid(k_render_plain_items<X3D_PT_DIRECT>, npix * spp);
        break;
    case X3D_PT_BIDIRECT_LT2:
        grid_render = persistent_grid(k_render_plain_items<X3D_PT_BIDIRECT_LT2>, npix * spp);
        break;
    case X3D_PT_POSITION:
        grid_render = persistent_grid(k_render_plain_items<X3D_PT_POSITION>, npix * spp);
        break;
    case X3D_PT_BIDIRECT_MIS:
        // (the wavefront form launches its own kernels)
        grid_render = wf ? 1 : persistent_grid(k_render_bdpt, npix * static_cast<uint64_t>(spp));
        break;
    default:
        grid_render = persistent_grid(k_render_plain_items<X3D_PT_NORMAL>, npix * spp);
        break;
    }
    const size_t spill_block = static_cast<size_t>(kWG) * P.spill_per_thread * sizeof(uint2);
    // the plain bidirectional kernel's subpaths: 2 x kMaxPathLen vertices of 4 float4 per lane
    const bool bd_plain = tracer == X3D_PT_BIDIRECT_MIS && !wf;
    const size_t subpath_block =
        bd_plain ? static_cast<size_t>(kWG) * 2 * kMaxPathLen * kPathletRows * sizeof(float4) : 0;
    if (arena) {
        // Keep room for one sample per generation; the traversal-stack spill (and subpaths) of the
        // resident lanes take at most half of the rest, fewer workgroups where they would not fit
        // (every kernel fetches its work dynamically, so any grid size gives the same film).
        const size_t one_sample = ScratchArena::up(npix * sizeof(float4)) + 512;
        const size_t left = arena->left();
        const size_t cap = left > one_sample ? (left - one_sample) / 2 / (spill_block + subpath_block) : 0;
        if (cap == 0)
            return fail(X3D_ERR_INVALID_ARGUMENT, "the film needs more than the " + std::to_string(arena->bytes >> 20) +
                                                      " MiB of fixed scratch memory");
        grid_first = grid_first < cap ? grid_first : static_cast<uint32_t>(cap);
        grid_render = grid_render < cap ? grid_render : static_cast<uint32_t>(cap);
    }
    uint32_t grid = grid_first > grid_render ? grid_first : grid_render;
    if (int rc = buffer(r->spill, static_cast<size_t>(grid) * spill_block, &p_spill))
        return rc;
    P.spill = static_cast<uint2 *>(p_spill);
    if (bd_plain) {
        void *p_sub;
        if (int rc = buffer(r->subpaths, static_cast<size_t>(grid_render) * subpath_block, &p_sub))
            return rc;
        P.subpaths = static_cast<float4 *>(p_sub);
    }
    // Counters are cleared by a kernel, not a memset, so the sequence stays capturable in a graph.
    uint32_t *next_pixel = reinterpret_cast<uint32_t *>(P.cnt + 2);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, r->stream, reinterpret_cast<unsigned long long *>(P.cnt),
                       static_cast<uint32_t>((2 * sizeof(Counters) + 64) / 8));
    X3D_HIP(hipEventRecord(r->ev0, r->stream));
    hipLaunchKernelGGL(k_first_hit, dim3(grid_first), dim3(kWG), 0, r->stream, P);
    X3D_HIP(hipGetLastError());
    X3D_HIP(hipEventRecord(r->evm, r->stream));
    P.cnt = P.cnt + 1;
    P.s_begin = 0;
    P.s_end = spp;
    P.finalize = 1;
    if (d_dl) {
        if (!r->adj)
            r->adj = x3d::adjoint_create();
        uint32_t user_mats = static_cast<uint32_t>(scene->s.materials.size()) - 1; // fail-safe appended last
        std::string e = x3d::adjoint_render(r->adj, P, d_dl, user_mats, d_grad, r->stream);
        if (!e.empty())
            return fail(e.find("hip") != std::string::npos ? X3D_ERR_DEVICE : X3D_ERR_UNSUPPORTED, e);
        tracer = -1; // launched
    } else if (wf && tracer != X3D_PT_BIDIRECT_MIS) {
        std::string e = x3d::wf_render(wf, P, *wf_tune, r->stream);
        if (!e.empty())
            return fail(X3D_ERR_DEVICE, e);
        tracer = -1; // launched
    }
    int rc = X3D_OK;
    // Generations of whole samples, at most queue_capacity (default kGenItems) (pixel, sample)
    // items each, rendered by `kernel` (or, for the bidirectional tracer's wavefront form, by its
    // kernel sequence) into the sample buffer and summed per pixel in order.
    const bool bd_wf = wf && tracer == X3D_PT_BIDIRECT_MIS;
    auto run_generations = [&](void (*kernel)(RenderParams, uint32_t, uint32_t, float4 *, uint32_t *)) -> int {
        uint64_t items = r->opt.queue_capacity ? r->opt.queue_capacity : kGenItems;
        if (arena && items > arena->left() / sizeof(float4))
            items = arena->left() / sizeof(float4);
        if (bd_wf) {
            const uint64_t cap = x3d::bd_max_items(P.max_path_len);
            if (cap < npix)
                return fail(X3D_ERR_INVALID_ARGUMENT, "the film has more pixels than one bidirectional wavefront "
                                                      "generation holds; use the plain form");
            items = items < cap ? items : cap;
        }
        uint32_t gen = static_cast<uint32_t>(items / npix);
        gen = gen < 1 ? 1 : (gen > spp ? spp : gen);
        void *p_samples;
        if (int rc = buffer(r->samples, static_cast<size_t>(gen) * npix * sizeof(float4), &p_samples))
            return rc;
        float4 *samples = static_cast<float4 *>(p_samples);
        uint32_t g_sum = device_info().cus * 4u;
        const uint32_t s_first = rl ? rl->s_first : 0u;
        for (uint32_t s0 = 0; s0 < spp; s0 += gen) {
            uint32_t cnt = spp - s0 < gen ? spp - s0 : gen;
            if (bd_wf) {
                std::string e = x3d::bd_generation(wf, P, *wf_tune, s_first + s0, cnt, samples, s0 == 0, r->stream);
                if (!e.empty())
                    return fail(X3D_ERR_DEVICE, e);
            } else {
                hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, r->stream,
                                   reinterpret_cast<unsigned long long *>(next_pixel), 1u);
                hipLaunchKernelGGL(kernel, dim3(grid_render), dim3(kWG), 0, r->stream, P, s_first + s0, cnt, samples,
                                   next_pixel);
            }
            if (rl)
                hipLaunchKernelGGL(k_pack_rgb, dim3(g_sum), dim3(256), 0, r->stream,
                                   static_cast<const float4 *>(samples), static_cast<uint64_t>(cnt) * npix,
                                   rl->d_out + 3ull * s0 * npix);
            else
                hipLaunchKernelGGL(k_sum_samples, dim3(g_sum), dim3(256), 0, r->stream, P, s0, cnt,
                                   static_cast<const float4 *>(samples));
        }
        return X3D_OK;
    };
    switch (tracer) {
    case -1: break;
    case X3D_PT_UNIDIRECT_MIS:
        rc = run_generations(k_render_mis_items);
        if (rc)
            return rc;
        break;
    case X3D_PT_UNIDIRECT_LT1:
        rc = run_generations(k_render_plain_items<X3D_PT_UNIDIRECT_LT1>);
        if (rc)
            return rc;
        break;
    case X3D_PT_UNIDIRECT:
        rc = run_generations(k_render_plain_items<X3D_PT_UNIDIRECT>);
        if (rc)
            return rc;
        break;
    case X3D_PT_DIRECT:
        rc = run_generations(k_render_plain_items<X3D_PT_DIRECT>);
        if (rc)
            return rc;
        break;
    case X3D_PT_BIDIRECT_LT2:
        rc = run_generations(k_render_plain_items<X3D_PT_BIDIRECT_LT2>);
        if (rc)
            return rc;
        break;
    case X3D_PT_POSITION:
        rc = run_generations(k_render_plain_items<X3D_PT_POSITION>);
        if (rc)
            return rc;
        break;
    case X3D_PT_BIDIRECT_MIS:
        rc = run_generations(k_render_bdpt);
        if (rc)
            return rc;
        break;
    default:
        rc = run_generations(k_render_plain_items<X3D_PT_NORMAL>);
        if (rc)
            return rc;
        break;
    }
    X3D_HIP(hipGetLastError());
    X3D_HIP(hipEventRecord(r->ev1, r->stream));
    X3D_HIP(hipEventSynchronize(r->ev1));
    if (stats) {
        Counters cc[2];
        X3D_HIP(hipMemcpy(cc, r->counters.p, sizeof cc, hipMemcpyDeviceToHost));
        Counters c = cc[1];
        c.closest += cc[0].closest;
        c.pops += cc[0].pops;
        c.tests += cc[0].tests;
        float ms = 0, ms_first = 0, ms_trace = 0;
        X3D_HIP(hipEventElapsedTime(&ms, r->ev0, r->ev1));
        X3D_HIP(hipEventElapsedTime(&ms_first, r->ev0, r->evm));
        X3D_HIP(hipEventElapsedTime(&ms_trace, r->evm, r->ev1));
        stats->num_samples = npix * spp;
        stats->closest_queries = c.closest;
        stats->anyhit_queries = c.anyhit;
        stats->closest_nodes = c.pops;
        stats->closest_tests = c.tests;
        stats->anyhit_nodes = c.spops;
        stats->anyhit_tests = c.stests;
        stats->first_hit_nodes = cc[0].pops;
        stats->first_hit_tests = cc[0].tests;
        stats->kernel_ms = ms;
        stats->first_hit_ms = ms_first;
        stats->trace_ms = ms_trace;
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->extend_ms = stats->shadow_ms = 0.0;
        stats->trace_launches = 0;
        stats->form = wf ? X3D_FORM_WAVEFRONT : X3D_FORM_PLAIN;
        if (wf && wf_tune->profile)
            x3d::wf_last_timing(wf, &stats->shadow_ms, &stats->extend_ms, &stats->trace_launches);
    }
    return X3D_OK;
}

int check_render_args(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width,
                      uint32_t height, uint32_t spp) {
    if (!r || !camera)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_scene(scene);
    if (rc)
        return rc;
    if (width < 2 || height < 2)
        return fail(X3D_ERR_INVALID_ARGUMENT, "film must be at least 2 x 2 (camera.cpp:38-48 divides by w-1, h-1)");
    if (static_cast<uint64_t>(width) * height > (1ull << 31))
        return fail(X3D_ERR_INVALID_ARGUMENT, "film too large");
    if (spp == 0)
        return fail(X3D_ERR_INVALID_ARGUMENT, "spp must be > 0");
    return X3D_OK;
}

int check_batch(x3d_renderer *r, uint32_t n, x3d_batch_job const *jobs) {
    if (!r || (n && !jobs))
        return fail(X3D_ERR_INVALID_ARGUMENT, "null argument");
    if (r->opt.pt_type != X3D_PT_UNIDIRECT_MIS)
        return fail(X3D_ERR_STATE, "batched rendering needs a UNIDIRECT_MIS renderer");
    for (uint32_t i = 0; i < n; i++) {
        int rc = check_render_args(r, jobs[i].scene, &jobs[i].camera, jobs[i].width, jobs[i].height, jobs[i].spp);
        if (rc)
            return fail(rc, "job " + std::to_string(i) + ": " + g_last_error);
        if (!jobs[i].rgba)
            return fail(X3D_ERR_INVALID_ARGUMENT, "job " + std::to_string(i) + ": rgba is null");
    }
    return X3D_OK;
}

constexpr uint64_t kBatchSplitDefault = 1ull << 22;

// Wavefront form of a batch: each job's first-hit pass, then one multi-job wavefront render in
// which the paths of all jobs share every generation (x3d::wf_render_jobs).
int batch_wavefront(x3d_renderer *r, uint32_t n, x3d_batch_job *jobs, float *const *films, x3d_batch_stats *stats,
                    std::chrono::steady_clock::time_point t0) {
    std::vector<uint64_t> off(n);
    uint64_t px = 0, samples = 0;
    uint32_t spill = 0, grid_first = 1;
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t npix = static_cast<uint64_t>(jobs[i].width) * jobs[i].height;
        off[i] = px;
        px += npix;
        samples += npix * jobs[i].spp;
        const uint32_t sp = spill_entries(jobs[i].scene->s.dev);
        spill = sp > spill ? sp : spill;
        const uint32_t g = persistent_grid(k_first_hit, npix);
        grid_first = g > grid_first ? g : grid_first;
    }
    X3D_HIP(r->first.ensure(px * 3 * sizeof(float4)));
    X3D_HIP(r->acc.ensure(px * sizeof(float4)));
    X3D_HIP(r->spill.ensure(static_cast<size_t>(grid_first) * kWG * spill * sizeof(uint2)));
    Counters *cnt = static_cast<Counters *>(r->counters.p);
    float4 *first = static_cast<float4 *>(r->first.p), *acc = static_cast<float4 *>(r->acc.p);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, r->stream, reinterpret_cast<unsigned long long *>(cnt),
                       static_cast<uint32_t>((2 * sizeof(Counters) + 64) / 8));
    X3D_HIP(hipEventRecord(r->bev0, r->stream));
    std::vector<x3d::WfJob> wj(n);
    for (uint32_t i = 0; i < n; i++) {
        RenderParams P{};
        P.S = jobs[i].scene->s.dev;
        P.cam = cam_params(jobs[i].camera);
        P.width = jobs[i].width;
        P.height = jobs[i].height;
        P.first = first + 3 * off[i];
        P.cnt = cnt;
        P.spill = static_cast<uint2 *>(r->spill.p);
        P.spill_per_thread = spill;
        const uint64_t npix = static_cast<uint64_t>(P.width) * P.height;
        hipLaunchKernelGGL(k_first_hit, dim3(persistent_grid(k_first_hit, npix)), dim3(kWG), 0, r->stream, P);
        x3d::WfJob &J = wj[i];
        J.S = P.S;
        J.cam = P.cam;
        J.first = P.first;
        J.acc = acc + off[i];
        J.out = reinterpret_cast<float4 *>(films[i]);
        J.width = P.width;
        J.height = P.height;
        J.npix = static_cast<uint32_t>(npix);
        J.spp = jobs[i].spp;
    }
    X3D_HIP(hipGetLastError());
    // the multi-job form runs fused: no material or ray sort
    x3d::WavefrontTuning wt = wf_tuning(r->opt);
    wt.tracer = x3d::kWfMis;
    wt.sort_by_material = 0;
    wt.sort_rays = 0;
    wt.profile = 0;
    if (!r->batch_wf)
        r->batch_wf = x3d::wf_create();
    x3d::WfJobsResult res;
    std::string e =
        x3d::wf_render_jobs(r->batch_wf, wj, r->opt.max_depth, r->opt.seed, cnt + 1, wt, r->stream, r->bev0, &res);
    if (!e.empty())
        return fail(X3D_ERR_DEVICE, e);
    X3D_HIP(hipEventRecord(r->bev1, r->stream));
    X3D_HIP(hipEventSynchronize(r->bev1));
    for (uint32_t i = 0; i < n; i++)
        jobs[i].latency_ms = res.job_ms[i];
    if (stats) {
        Counters cc[2];
        X3D_HIP(hipMemcpy(cc, cnt, sizeof cc, hipMemcpyDeviceToHost));
        float ms = 0;
        X3D_HIP(hipEventElapsedTime(&ms, r->bev0, r->bev1));
        *stats = x3d_batch_stats{};
        stats->num_samples = samples;
        stats->closest_queries = cc[0].closest + cc[1].closest;
        stats->anyhit_queries = cc[1].anyhit;
        stats->closest_nodes = cc[0].pops + cc[1].pops;
        stats->closest_tests = cc[0].tests + cc[1].tests;
        stats->anyhit_nodes = cc[1].spops;
        stats->anyhit_tests = cc[1].stests;
        stats->kernel_ms = ms;
        stats->generations = res.generations;
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return X3D_OK;
}

// Runs the batch with films already in device memory (films[i] for job i). Wavefront form: every
// job's paths share the generations of one multi-job wavefront render. Plain form: the jobs below
// the split share one persistent launch (small films cannot fill the GPU on their own); each job
// at or above it then runs alone in the wavefront form, fewest samples first. Either way a job's
// film is bit-identical to rendering it alone.
int batch_impl(x3d_renderer *r, uint32_t n, x3d_batch_job *jobs, float *const *films, x3d_batch_stats *stats) {
    auto t0 = std::chrono::steady_clock::now();
    if (n == 0) {
        if (stats)
            *stats = x3d_batch_stats{};
        return X3D_OK;
    }
    if (!r->batch)
        r->batch = x3d::batch_create();
    if (!r->bev0) {
        X3D_HIP(hipEventCreate(&r->bev0));
        X3D_HIP(hipEventCreate(&r->bev1));
    }
    X3D_HIP(r->counters.ensure(2 * sizeof(Counters) + 64));
    if (r->opt.form != X3D_FORM_PLAIN)
        return batch_wavefront(r, n, jobs, films, stats, t0);
    const uint64_t split = r->opt.batch_split ? r->opt.batch_split : kBatchSplitDefault;
    std::vector<x3d::BatchJob> bj;
    std::vector<uint32_t> shared, alone;
    uint64_t samples = 0, tiles = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint64_t npix = static_cast<uint64_t>(jobs[i].width) * jobs[i].height;
        uint64_t js = npix * jobs[i].spp;
        samples += js;
        if (js >= split) {
            alone.push_back(i);
            continue;
        }
        shared.push_back(i);
        x3d::BatchJob J{};
        J.S = jobs[i].scene->s.dev;
        J.cam = cam_params(jobs[i].camera);
        J.out = reinterpret_cast<float4 *>(films[i]);
        J.width = jobs[i].width;
        J.height = jobs[i].height;
        J.spp = jobs[i].spp;
        bj.push_back(J);
        tiles += (npix + kBatchTilePx - 1) / kBatchTilePx;
    }
    x3d::BatchTuning tune;
    if (r->opt.workgroup_size)
        tune.wg = static_cast<int>(r->opt.workgroup_size);
    if (r->opt.stack_depth)
        tune.stack_depth = static_cast<int>(r->opt.stack_depth);
    tune.sort_by_cost = r->opt.sort_key != X3D_SORT_NONE;
    Counters *cnt = static_cast<Counters *>(r->counters.p);
    Counters total{};
    X3D_HIP(hipEventRecord(r->bev0, r->stream));
    if (!shared.empty()) {
        x3d::BatchResult res;
        std::string e = x3d::batch_render(r->batch, bj, r->opt.max_depth, r->opt.seed, tune, cnt, r->stream, &res);
        if (!e.empty())
            return fail(X3D_ERR_DEVICE, e);
        for (size_t k = 0; k < shared.size(); k++)
            jobs[shared[k]].latency_ms = res.job_ms.empty() ? 0.0 : res.job_ms[k];
        X3D_HIP(hipMemcpy(&total, cnt, sizeof total, hipMemcpyDeviceToHost));
    }
    if (!alone.empty()) {
        // the shared launch's knobs (workgroup size, stack depth, sort key) do not apply here
        x3d::WavefrontTuning wt;
        if (r->opt.queue_capacity)
            wt.paths_in_flight = r->opt.queue_capacity;
        if (!r->batch_wf)
            r->batch_wf = x3d::wf_create();
        std::stable_sort(alone.begin(), alone.end(), [&](uint32_t a, uint32_t b) {
            return static_cast<uint64_t>(jobs[a].width) * jobs[a].height * jobs[a].spp <
                   static_cast<uint64_t>(jobs[b].width) * jobs[b].height * jobs[b].spp;
        });
        for (uint32_t i : alone) {
            x3d_render_stats st{};
            int rc = render_impl(r, jobs[i].scene, &jobs[i].camera, jobs[i].width, jobs[i].height, jobs[i].spp,
                                 films[i], &st, r->batch_wf, &wt);
            if (rc)
                return rc;
            float ms = 0;
            X3D_HIP(hipEventElapsedTime(&ms, r->bev0, r->ev1)); // ev1: the end of that render
            jobs[i].latency_ms = ms;
            total.closest += st.closest_queries;
            total.anyhit += st.anyhit_queries;
            total.pops += st.closest_nodes;
            total.tests += st.closest_tests;
            total.spops += st.anyhit_nodes;
            total.stests += st.anyhit_tests;
        }
    }
    X3D_HIP(hipEventRecord(r->bev1, r->stream));
    X3D_HIP(hipEventSynchronize(r->bev1));
    if (stats) {
        float ms = 0;
        X3D_HIP(hipEventElapsedTime(&ms, r->bev0, r->bev1));
        *stats = x3d_batch_stats{};
        stats->num_samples = samples;
        stats->closest_queries = total.closest;
        stats->anyhit_queries = total.anyhit;
        stats->closest_nodes = total.pops;
        stats->closest_tests = total.tests;
        stats->anyhit_nodes = total.spops;
        stats->anyhit_tests = total.stests;
        stats->kernel_ms = ms;
        stats->tiles = static_cast<uint32_t>(tiles);
        stats->alone_jobs = static_cast<uint32_t>(alone.size());
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return X3D_OK;
}

} // namespace

extern "C" {

const char *x3d_last_error(void) { return g_last_error.c_str(); }

const char *x3d_version(void) { return "x3dpt 0.1 (gfx950)"; }

int x3d_scene_create(x3d_scene **out) {
    if (!out)
        return fail(X3D_ERR_INVALID_ARGUMENT, "out is null");
    *out = new (std::nothrow) x3d_scene();
    return *out ? X3D_OK : fail(X3D_ERR_OUT_OF_MEMORY, "allocation failed");
}

int x3d_scene_destroy(x3d_scene *scene) {
    delete scene;
    return X3D_OK;
}

namespace {

// Material description -> device record (oren_nayar, cook_torr, mat_mixture constructors);
// mixture operands index `existing`.
int make_material(const x3d_material_desc *d, std::vector<DevMaterial> const &existing, DevMaterial &m) {
    m = DevMaterial{};
    m.kind = d->kind;
    for (int c = 0; c < 3; c++)
        m.albedo[c] = d->albedo[c];
    if (!finite3(d->albedo))
        return fail(X3D_ERR_INVALID_ARGUMENT, "albedo must be finite");
    if (d->kind == X3D_MAT_OREN_NAYAR) {
        // oren_nayar constructor (material.cpp:108-119).
        float sigma2 = d->sigma * d->sigma;
        m.on_a = 1 - 0.5f * sigma2 / (sigma2 + 0.33f);
        m.on_b = 0.45f * sigma2 / (sigma2 + 0.09f);
        m.sigma = d->sigma;
    } else if (d->kind == X3D_MAT_COOK_TORR) {
        if (!(d->roughness > 0.0f))
            return fail(X3D_ERR_INVALID_ARGUMENT, "cook_torr roughness must be > 0");
        m.alpha2 = 2 * d->roughness * d->roughness; // material.cpp:379
        m.sample_alpha2 = m.alpha2;
        m.roughness = d->roughness;
        m.ior_re = d->ior_re;
        m.ior_im = d->ior_im;
    } else if (d->kind == X3D_MAT_MIXTURE) {
        int32_t n = static_cast<int32_t>(existing.size());
        if (d->sub0 < 0 || d->sub0 >= n || d->sub1 < 0 || d->sub1 >= n)
            return fail(X3D_ERR_INVALID_ARGUMENT, "mixture operand out of range");
        if (existing[d->sub0].kind == kMatMixture || existing[d->sub1].kind == kMatMixture)
            return fail(X3D_ERR_UNSUPPORTED, "nested mixtures are not supported");
        if (!(d->ratio >= 0.0f && d->ratio <= 1.0f))
            return fail(X3D_ERR_INVALID_ARGUMENT, "mixture ratio must be in [0, 1]");
        m.sub0 = d->sub0;
        m.sub1 = d->sub1;
        m.ratio = d->ratio;
    } else {
        return fail(X3D_ERR_INVALID_ARGUMENT, "unknown material kind");
    }
    return X3D_OK;
}

// Validated mesh arrays -> host mesh. Without a transform the vertices pass through the
// identity transform the reference loads meshes with (pathspace.cpp:15-19, geometry.cpp:76-87),
// which maps -0 to +0 (adding +0 does the same); with one, trimesh::transform's arithmetic.
int make_mesh(const float *verts, const float *normals, uint32_t num_verts, const uint32_t *tris, uint32_t num_tris,
              const float *transform, HostMesh &m) {
    if ((!verts && num_verts) || (!normals && num_verts) || (!tris && num_tris))
        return fail(X3D_ERR_INVALID_ARGUMENT, "null mesh array");
    m.verts.resize(num_verts);
    m.normals.resize(num_verts);
    for (uint32_t i = 0; i < num_verts; i++) {
        if (!finite3(verts + 3 * i) || !finite3(normals + 3 * i))
            return fail(X3D_ERR_INVALID_ARGUMENT, "vertex " + std::to_string(i) + " is not finite");
        m.verts[i] = {verts[3 * i] + 0.0f, verts[3 * i + 1] + 0.0f, verts[3 * i + 2] + 0.0f};
        m.normals[i] = {normals[3 * i] + 0.0f, normals[3 * i + 1] + 0.0f, normals[3 * i + 2] + 0.0f};
    }
    if (transform) {
        for (int k = 0; k < 16; k++)
            if (!std::isfinite(transform[k]))
                return fail(X3D_ERR_INVALID_ARGUMENT, "transform is not finite");
        for (uint32_t i = 0; i < num_verts; i++) {
            m.verts[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
            m.normals[i] = {normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]};
        }
        std::string e = x3d::transform_mesh(transform, m.verts, m.normals);
        if (!e.empty())
            return fail(X3D_ERR_INVALID_ARGUMENT, e);
        for (uint32_t i = 0; i < num_verts; i++)
            if (!finite3(&m.verts[i].x) || !finite3(&m.normals[i].x))
                return fail(X3D_ERR_INVALID_ARGUMENT, "transformed vertex " + std::to_string(i) + " is not finite");
    }
    m.tris.assign(tris, tris + 3ull * num_tris);
    for (uint32_t v : m.tris)
        if (v >= num_verts)
            return fail(X3D_ERR_INVALID_ARGUMENT, "triangle index out of range");
    return X3D_OK;
}

int require_indexed(x3d_scene *scene) {
    if (scene->s.keyed)
        return fail(X3D_ERR_STATE, "the scene takes objects by id (x3d_scene_load_*)");
    if (scene->s.committed)
        return fail(X3D_ERR_STATE, "scene already committed");
    return X3D_OK;
}

int require_keyed(x3d_scene *scene) {
    if (!scene->s.keyed && (!scene->s.meshes.empty() || !scene->s.materials.empty() || !scene->s.lights.empty() ||
                            scene->s.committed))
        return fail(X3D_ERR_STATE, "the scene takes objects by index (x3d_scene_add_*)");
    scene->s.keyed = true;
    return X3D_OK;
}

} // namespace

int x3d_scene_add_material(x3d_scene *scene, const x3d_material_desc *d, int32_t *out_id) {
    if (!scene || !d)
        return fail(X3D_ERR_INVALID_ARGUMENT, "scene or desc is null");
    if (int rc = require_indexed(scene))
        return rc;
    DevMaterial m;
    if (int rc = make_material(d, scene->s.materials, m))
        return rc;
    scene->s.materials.push_back(m);
    if (out_id)
        *out_id = static_cast<int32_t>(scene->s.materials.size()) - 1;
    return X3D_OK;
}

int x3d_scene_add_mesh(x3d_scene *scene, const float *verts, const float *normals, const float *uvs,
                       uint32_t num_verts, const uint32_t *tris, uint32_t num_tris, int32_t material_id,
                       int32_t *out_mesh_id) {
    (void)uvs;
    if (!scene)
        return fail(X3D_ERR_INVALID_ARGUMENT, "scene is null");
    if (int rc = require_indexed(scene))
        return rc;
    if (material_id >= static_cast<int32_t>(scene->s.materials.size()))
        return fail(X3D_ERR_INVALID_ARGUMENT, "material id out of range");
    HostMesh m;
    if (int rc = make_mesh(verts, normals, num_verts, tris, num_tris, nullptr, m))
        return rc;
    m.material = material_id < 0 ? -1 : material_id;
    scene->s.grow_space(m.verts);
    scene->s.meshes.push_back(std::move(m));
    if (out_mesh_id)
        *out_mesh_id = static_cast<int32_t>(scene->s.meshes.size()) - 1;
    return X3D_OK;
}

int x3d_scene_load_geometry(x3d_scene *scene, uint64_t obj_id, const float *verts, const float *normals,
                            const float *uvs, uint32_t num_verts, const uint32_t *tris, uint32_t num_tris,
                            int32_t has_material, uint64_t material_obj_id, const float *transform) {
    (void)uvs;
    if (!scene)
        return fail(X3D_ERR_INVALID_ARGUMENT, "scene is null");
    if (int rc = require_keyed(scene))
        return rc;
    x3d::KeyedGeometry g;
    if (int rc = make_mesh(verts, normals, num_verts, tris, num_tris, transform, g.mesh))
        return rc;
    g.has_material = has_material != 0;
    g.material = material_obj_id;
    // if_path_space::load: the bound grows before the map insert, which keeps an existing entry
    scene->s.grow_space(g.mesh.verts);
    scene->s.k_geos.insert(std::make_pair(obj_id, std::move(g)));
    return X3D_OK;
}

int x3d_scene_load_material(x3d_scene *scene, uint64_t obj_id, const x3d_material_desc *d) {
    if (!scene || !d)
        return fail(X3D_ERR_INVALID_ARGUMENT, "scene or desc is null");
    if (int rc = require_keyed(scene))
        return rc;
    if (d->kind == X3D_MAT_MIXTURE)
        return fail(X3D_ERR_UNSUPPORTED, "mixtures are added by index (x3d_scene_add_material)");
    DevMaterial m;
    if (int rc = make_material(d, {}, m))
        return rc;
    scene->s.k_mats.insert(std::make_pair(obj_id, m));
    return X3D_OK;
}

int x3d_scene_load_light(x3d_scene *scene, uint64_t obj_id, uint64_t geometry_obj_id, const float radiance[3]) {
    if (!scene || !radiance)
        return fail(X3D_ERR_INVALID_ARGUMENT, "scene or radiance is null");
    if (!finite3(radiance) || radiance[0] < 0 || radiance[1] < 0 || radiance[2] < 0)
        return fail(X3D_ERR_INVALID_ARGUMENT, "radiance must be finite and non-negative");
    if (int rc = require_keyed(scene))
        return rc;
    x3d::KeyedLight l;
    l.geometry = geometry_obj_id;
    for (int c = 0; c < 3; c++)
        l.rad[c] = radiance[c];
    scene->s.k_lights.insert(std::make_pair(obj_id, l));
    return X3D_OK;
}

int x3d_scene_load_sky_light(x3d_scene *scene, uint64_t obj_id, const float radiance[3]) {
    if (!scene || !radiance)
        return fail(X3D_ERR_INVALID_ARGUMENT, "scene or radiance is null");
    if (!finite3(radiance) || radiance[0] < 0 || radiance[1] < 0 || radiance[2] < 0)
        return fail(X3D_ERR_INVALID_ARGUMENT, "radiance must be finite and non-negative");
    if (int rc = require_keyed(scene))
        return rc;
    x3d::KeyedLight l;
    l.geometry = 0;
    l.sky = true;
    for (int c = 0; c < 3; c++)
        l.rad[c] = radiance[c];
    scene->s.k_lights.insert(std::make_pair(obj_id, l));
    return X3D_OK;
}

int x3d_scene_unload(x3d_scene *scene, int32_t protocol, uint64_t obj_id) {
    if (!scene)
        return fail(X3D_ERR_INVALID_ARGUMENT, "scene is null");
    if (int rc = require_keyed(scene))
        return rc;
    if (protocol == X3D_OBJ_GEOMETRY)
        scene->s.k_geos.erase(obj_id);
    else if (protocol == X3D_OBJ_MATERIAL)
        scene->s.k_mats.erase(obj_id);
    else if (protocol == X3D_OBJ_LIGHT)
        scene->s.k_lights.erase(obj_id);
    else
        return fail(X3D_ERR_INVALID_ARGUMENT, "unknown object protocol");
    return X3D_OK;
}

int x3d_scene_geometry_obj_id(const x3d_scene *scene, int32_t geo, uint64_t *obj_id) {
    if (!scene || !obj_id)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null argument");
    if (!scene->s.committed)
        return fail(X3D_ERR_STATE, "scene not committed");
    if (geo < 0 || static_cast<size_t>(geo) >= scene->s.meshes.size())
        return fail(X3D_ERR_INVALID_ARGUMENT, "geometry index out of range");
    *obj_id = scene->s.keyed ? scene->s.geo_ids[geo] : static_cast<uint64_t>(geo);
    return X3D_OK;
}

int x3d_transform_mesh(const float transform[16], uint32_t num_verts, const float *verts, const float *normals,
                       float *out_verts, float *out_normals) {
    if (!transform || ((!verts || !normals || !out_verts || !out_normals) && num_verts))
        return fail(X3D_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<HostVec3> v(num_verts), n(num_verts);
    for (uint32_t i = 0; i < num_verts; i++) {
        v[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
        n[i] = {normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]};
    }
    std::string e = x3d::transform_mesh(transform, v, n);
    if (!e.empty())
        return fail(X3D_ERR_INVALID_ARGUMENT, e);
    for (uint32_t i = 0; i < num_verts; i++) {
        out_verts[3 * i] = v[i].x;
        out_verts[3 * i + 1] = v[i].y;
        out_verts[3 * i + 2] = v[i].z;
        out_normals[3 * i] = n[i].x;
        out_normals[3 * i + 1] = n[i].y;
        out_normals[3 * i + 2] = n[i].z;
    }
    return X3D_OK;
}

int x3d_obj_load(const char *path, x3d_obj_mesh *mesh) {
    if (!path || !mesh)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null argument");
    *mesh = x3d_obj_mesh{};
    x3d::ObjMesh m;
    std::string e = x3d::load_obj(path, m);
    if (!e.empty())
        return fail(X3D_ERR_INVALID_ARGUMENT, std::string(path) + ": " + e);
    const size_t nv = m.verts.size() / 3, nt = m.tris.size() / 3;
    if (nv > 0xFFFFFFFFu || nt > 0xFFFFFFFFu)
        return fail(X3D_ERR_INVALID_ARGUMENT, "mesh too large");
    mesh->num_verts = static_cast<uint32_t>(nv);
    mesh->num_tris = static_cast<uint32_t>(nt);
    mesh->verts = new (std::nothrow) float[3 * nv];
    mesh->normals = new (std::nothrow) float[3 * nv];
    mesh->uvs = new (std::nothrow) float[2 * nv];
    mesh->tris = new (std::nothrow) uint32_t[3 * nt + 1];
    if (!mesh->verts || !mesh->normals || !mesh->uvs || !mesh->tris) {
        x3d_obj_free(mesh);
        return fail(X3D_ERR_OUT_OF_MEMORY, "allocation failed");
    }
    std::copy(m.verts.begin(), m.verts.end(), mesh->verts);
    std::copy(m.normals.begin(), m.normals.end(), mesh->normals);
    std::copy(m.uvs.begin(), m.uvs.end(), mesh->uvs);
    std::copy(m.tris.begin(), m.tris.end(), mesh->tris);
    return X3D_OK;
}

int x3d_obj_free(x3d_obj_mesh *mesh) {
    if (!mesh)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null argument");
    delete[] mesh->verts;
    delete[] mesh->normals;
    delete[] mesh->uvs;
    delete[] mesh->tris;
    *mesh = x3d_obj_mesh{};
    return X3D_OK;
}

int x3d_scene_add_obj(x3d_scene *scene, const char *path, int32_t material_id, int32_t *out_mesh_id) {
    if (!scene || !path)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null argument");
    x3d_obj_mesh m{};
    int rc = x3d_obj_load(path, &m);
    if (rc)
        return rc;
    rc = x3d_scene_add_mesh(scene, m.verts, m.normals, m.uvs, m.num_verts, m.tris, m.num_tris, material_id,
                            out_mesh_id);
    x3d_obj_free(&m);
    return rc;
}

int x3d_scene_add_area_light(x3d_scene *scene, int32_t mesh_id, const float radiance[3], int32_t *out_light_id) {
    if (!scene || !radiance)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null argument");
    if (int rc = require_indexed(scene))
        return rc;
    if (mesh_id < 0 || mesh_id >= static_cast<int32_t>(scene->s.meshes.size()))
        return fail(X3D_ERR_INVALID_ARGUMENT, "mesh id out of range");
    if (!finite3(radiance) || radiance[0] < 0 || radiance[1] < 0 || radiance[2] < 0)
        return fail(X3D_ERR_INVALID_ARGUMENT, "radiance must be finite and non-negative");
    HostLight l;
    l.mesh = mesh_id;
    l.rad[0] = radiance[0];
    l.rad[1] = radiance[1];
    l.rad[2] = radiance[2];
    scene->s.lights.push_back(l);
    if (out_light_id)
        *out_light_id = static_cast<int32_t>(scene->s.lights.size()) - 1;
    return X3D_OK;
}

int x3d_scene_add_sky_light(x3d_scene *scene, const float radiance[3], int32_t *out_light_id) {
    if (!scene || !radiance)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null argument");
    if (int rc = require_indexed(scene))
        return rc;
    if (!finite3(radiance) || radiance[0] < 0 || radiance[1] < 0 || radiance[2] < 0)
        return fail(X3D_ERR_INVALID_ARGUMENT, "radiance must be finite and non-negative");
    HostLight l;
    l.mesh = -1;
    for (int c = 0; c < 3; c++)
        l.rad[c] = radiance[c];
    scene->s.lights.push_back(l);
    if (out_light_id)
        *out_light_id = static_cast<int32_t>(scene->s.lights.size()) - 1;
    return X3D_OK;
}

int x3d_scene_commit(x3d_scene *scene) {
    if (!scene)
        return fail(X3D_ERR_INVALID_ARGUMENT, "scene is null");
    int rc = require_device();
    if (rc)
        return rc;
    std::string e = scene->s.commit();
    if (!e.empty())
        return fail(e.find("hip") != std::string::npos ? X3D_ERR_DEVICE : X3D_ERR_INVALID_ARGUMENT, e);
    return X3D_OK;
}

int x3d_scene_set_material_params(x3d_scene *scene, uint32_t first, uint32_t count, const float *albedo,
                                  const float *roughness, int32_t update_sampling) {
    int rc = check_scene(scene);
    if (rc)
        return rc;
    Scene &s = scene->s;
    uint32_t user = static_cast<uint32_t>(s.materials.size()) - 1; // fail-safe appended last
    if (first > user || count > user - first)
        return fail(X3D_ERR_INVALID_ARGUMENT, "material range out of bounds");
    for (uint32_t k = 0; k < count; k++) {
        DevMaterial const &m = s.materials[first + k];
        if (albedo && !finite3(albedo + 3 * k))
            return fail(X3D_ERR_INVALID_ARGUMENT, "albedo must be finite");
        if (roughness && m.kind == kMatCookTorr && !(roughness[k] > 0.0f && std::isfinite(roughness[k])))
            return fail(X3D_ERR_INVALID_ARGUMENT, "cook_torr roughness must be > 0");
    }
    for (uint32_t k = 0; k < count; k++) {
        DevMaterial &m = s.materials[first + k];
        if (m.kind == kMatMixture)
            continue;
        if (albedo)
            for (int c = 0; c < 3; c++)
                m.albedo[c] = albedo[3 * k + c];
        if (roughness && m.kind == kMatCookTorr) {
            m.roughness = roughness[k];
            m.alpha2 = 2 * roughness[k] * roughness[k]; // material.cpp:379
            if (update_sampling)
                m.sample_alpha2 = m.alpha2;
        }
    }
    X3D_HIP(hipMemcpy(s.buf.mats, s.materials.data(), s.materials.size() * sizeof(DevMaterial),
                      hipMemcpyHostToDevice));
    if (s.keyed) {
        // Keep the stored objects in step by id (committed material k has obj id mat_ids[k]), so
        // that a later commit keeps the parameters even after loads or unloads since this
        // commit; a material unloaded since the commit is not brought back.
        for (uint32_t k = 0; k < count; k++) {
            auto it = s.k_mats.find(s.mat_ids[first + k]);
            if (it != s.k_mats.end())
                it->second = s.materials[first + k];
        }
    }
    return X3D_OK;
}

int x3d_scene_get_info(const x3d_scene *scene, x3d_scene_info *info) {
    if (!scene || !info)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null argument");
    Scene const &s = scene->s;
    info->num_meshes = static_cast<uint32_t>(s.meshes.size());
    info->num_tris = static_cast<uint32_t>(s.prims.size());
    info->num_nodes = static_cast<uint32_t>(s.flat.size());
    info->num_interior = s.num_interior;
    info->num_lights = static_cast<uint32_t>(s.lights.size());
    info->num_materials = static_cast<uint32_t>(s.materials.size());
    info->max_leaf_depth = s.max_leaf_depth;
    info->device_bytes = s.buf.bytes;
    info->build_seconds = s.build_seconds;
    info->upload_seconds = s.upload_seconds;
    return X3D_OK;
}

int x3d_scene_dump_bvh(const x3d_scene *scene, float *bounds, uint32_t *meta, int32_t *prims) {
    int rc = check_scene(scene);
    if (rc)
        return rc;
    Scene const &s = scene->s;
    for (size_t n = 0; n < s.flat.size(); n++) {
        FlatNode const &f = s.flat[n];
        if (bounds)
            for (int i = 0; i < 3; i++) {
                bounds[6 * n + i] = f.mn[i];
                bounds[6 * n + 3 + i] = f.mx[i];
            }
        if (meta) {
            meta[4 * n] = f.num_prims;
            meta[4 * n + 1] = f.split_axis;
            meta[4 * n + 2] = f.next_child;
            meta[4 * n + 3] = f.prim_start;
        }
    }
    if (prims)
        for (size_t i = 0; i < s.prims.size(); i++) {
            prims[2 * i] = static_cast<int32_t>(s.prims[i].mesh);
            prims[2 * i + 1] = static_cast<int32_t>(s.prims[i].tri);
        }
    return X3D_OK;
}

int x3d_intersect_device(const x3d_scene *scene, const float *d_rays, uint32_t n, int32_t traversal,
                         x3d_hit *d_hits, void *stream) {
    int rc = check_scene(scene);
    if (rc)
        return rc;
    if (n == 0)
        return X3D_OK;
    if (!d_rays || !d_hits)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null device pointer");
    static thread_local x3d::DevBuf spill;
    return run_queries(scene, d_rays, nullptr, nullptr, n, traversal, d_hits, nullptr,
                       static_cast<hipStream_t>(stream), spill);
}

int x3d_occluded_device(const x3d_scene *scene, const float *d_rays, const float *d_tmin, const float *d_tmax,
                        uint32_t n, uint8_t *d_occ, void *stream) {
    int rc = check_scene(scene);
    if (rc)
        return rc;
    if (n == 0)
        return X3D_OK;
    if (!d_rays || !d_tmin || !d_tmax || !d_occ)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null device pointer");
    static thread_local x3d::DevBuf spill;
    return run_queries(scene, d_rays, d_tmin, d_tmax, n, X3D_TRAVERSAL_CULLED, nullptr, d_occ,
                       static_cast<hipStream_t>(stream), spill);
}

int x3d_intersect(const x3d_scene *scene, const float *rays, uint32_t n, int32_t traversal, x3d_hit *hits) {
    int rc = check_scene(scene);
    if (rc)
        return rc;
    if (n == 0)
        return X3D_OK;
    if (!rays || !hits)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null host pointer");
    if (traversal != X3D_TRAVERSAL_CULLED && traversal != X3D_TRAVERSAL_REFERENCE)
        return fail(X3D_ERR_INVALID_ARGUMENT, "unknown traversal mode");
    x3d::DevBuf dr, dh, spill;
    X3D_HIP(dr.ensure(sizeof(float) * 6ull * n));
    X3D_HIP(dh.ensure(sizeof(x3d_hit) * static_cast<size_t>(n)));
    X3D_HIP(hipMemcpy(dr.p, rays, sizeof(float) * 6ull * n, hipMemcpyHostToDevice));
    rc = run_queries(scene, static_cast<float *>(dr.p), nullptr, nullptr, n, traversal,
                     static_cast<x3d_hit *>(dh.p), nullptr, nullptr, spill);
    if (rc)
        return rc;
    X3D_HIP(hipMemcpy(hits, dh.p, sizeof(x3d_hit) * static_cast<size_t>(n), hipMemcpyDeviceToHost));
    return X3D_OK;
}

int x3d_occluded(const x3d_scene *scene, const float *rays, const float *tmin, const float *tmax, uint32_t n,
                 uint8_t *occ) {
    int rc = check_scene(scene);
    if (rc)
        return rc;
    if (n == 0)
        return X3D_OK;
    if (!rays || !tmin || !tmax || !occ)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null host pointer");
    x3d::DevBuf dr, d0, d1, dout, spill;
    X3D_HIP(dr.ensure(sizeof(float) * 6ull * n));
    X3D_HIP(d0.ensure(sizeof(float) * n));
    X3D_HIP(d1.ensure(sizeof(float) * n));
    X3D_HIP(dout.ensure(n));
    X3D_HIP(hipMemcpy(dr.p, rays, sizeof(float) * 6ull * n, hipMemcpyHostToDevice));
    X3D_HIP(hipMemcpy(d0.p, tmin, sizeof(float) * n, hipMemcpyHostToDevice));
    X3D_HIP(hipMemcpy(d1.p, tmax, sizeof(float) * n, hipMemcpyHostToDevice));
    rc = run_queries(scene, static_cast<float *>(dr.p), static_cast<float *>(d0.p), static_cast<float *>(d1.p), n,
                     X3D_TRAVERSAL_CULLED, nullptr, static_cast<uint8_t *>(dout.p), nullptr, spill);
    if (rc)
        return rc;
    X3D_HIP(hipMemcpy(occ, dout.p, n, hipMemcpyDeviceToHost));
    return X3D_OK;
}

static int bsdf_batch(const x3d_scene *scene, int32_t mat, uint32_t n, const float *nrm, const float *o,
                      const float *i, int op, uint64_t seed, int32_t ref_mixture, float *out_f, float *out_i,
                      float *out_pdf) {
    int rc = check_scene(scene);
    if (rc)
        return rc;
    if (mat < -1 || mat >= static_cast<int32_t>(scene->s.materials.size()) - 1)
        return fail(X3D_ERR_INVALID_ARGUMENT, "material id out of range");
    if (n == 0)
        return X3D_OK;
    if (mat < 0)
        mat = static_cast<int32_t>(scene->s.dev.fail_safe_mat);
    x3d::DevBuf dn, dd, di, df, dI, dp;
    size_t v3 = sizeof(float) * 3ull * n;
    X3D_HIP(dn.ensure(v3));
    X3D_HIP(dd.ensure(v3));
    X3D_HIP(hipMemcpy(dn.p, nrm, v3, hipMemcpyHostToDevice));
    X3D_HIP(hipMemcpy(dd.p, o, v3, hipMemcpyHostToDevice));
    if (op != 2) {
        X3D_HIP(di.ensure(v3));
        X3D_HIP(hipMemcpy(di.p, i, v3, hipMemcpyHostToDevice));
    }
    X3D_HIP(df.ensure(v3));
    X3D_HIP(dI.ensure(v3));
    X3D_HIP(dp.ensure(sizeof(float) * n));
    BsdfParams P{};
    P.mats = scene->s.dev.mats;
    P.mat = mat;
    P.n = n;
    P.nrm = static_cast<float *>(dn.p);
    P.o = static_cast<float *>(dd.p);
    P.i = static_cast<float *>(di.p);
    P.out_f = static_cast<float *>(df.p);
    P.out_i = static_cast<float *>(dI.p);
    P.out_pdf = static_cast<float *>(dp.p);
    P.seed = seed;
    P.ref_mixture = ref_mixture;
    P.op = op;
    hipLaunchKernelGGL(k_bsdf, dim3((n + 255) / 256), dim3(256), 0, nullptr, P);
    X3D_HIP(hipGetLastError());
    if (op == 0)
        X3D_HIP(hipMemcpy(out_f, df.p, v3, hipMemcpyDeviceToHost));
    if (op == 1)
        X3D_HIP(hipMemcpy(out_pdf, dp.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (op == 2) {
        X3D_HIP(hipMemcpy(out_i, dI.p, v3, hipMemcpyDeviceToHost));
        X3D_HIP(hipMemcpy(out_pdf, dp.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    }
    return X3D_OK;
}

int x3d_bsdf_eval(const x3d_scene *scene, int32_t material_id, uint32_t n, const float *normals, const float *o,
                  const float *i, float *f) {
    if (n && (!normals || !o || !i || !f))
        return fail(X3D_ERR_INVALID_ARGUMENT, "null host pointer");
    return bsdf_batch(scene, material_id, n, normals, o, i, 0, 0, 0, f, nullptr, nullptr);
}

int x3d_bsdf_pdf(const x3d_scene *scene, int32_t material_id, uint32_t n, const float *normals, const float *o,
                 const float *i, float *pdf) {
    if (n && (!normals || !o || !i || !pdf))
        return fail(X3D_ERR_INVALID_ARGUMENT, "null host pointer");
    return bsdf_batch(scene, material_id, n, normals, o, i, 1, 0, 0, nullptr, nullptr, pdf);
}

int x3d_bsdf_sample(const x3d_scene *scene, int32_t material_id, uint64_t seed, uint32_t n, const float *normals,
                    const float *o, int32_t reference_mixture, float *i, float *dens) {
    if (n && (!normals || !o || !i || !dens))
        return fail(X3D_ERR_INVALID_ARGUMENT, "null host pointer");
    return bsdf_batch(scene, material_id, n, normals, o, nullptr, 2, seed, reference_mixture, nullptr, i, dens);
}

int x3d_camera_rays(const x3d_camera *camera, uint32_t width, uint32_t height, float *rays) {
    if (!camera || !rays)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null argument");
    if (width < 2 || height < 2)
        return fail(X3D_ERR_INVALID_ARGUMENT, "film must be at least 2 x 2 (camera.cpp:38-48 divides by w-1, h-1)");
    int rc = require_device();
    if (rc)
        return rc;
    size_t bytes = sizeof(float) * 6ull * width * height;
    x3d::DevBuf d;
    X3D_HIP(d.ensure(bytes));
    uint32_t n = width * height;
    hipLaunchKernelGGL(k_camera_rays, dim3((n + 255) / 256), dim3(256), 0, nullptr, cam_params(*camera), width,
                       height, static_cast<float *>(d.p));
    X3D_HIP(hipGetLastError());
    X3D_HIP(hipMemcpy(rays, d.p, bytes, hipMemcpyDeviceToHost));
    return X3D_OK;
}

int x3d_pcg_draws(uint64_t seed, uint64_t pixel, uint32_t sample, uint32_t n, uint32_t *out) {
    if (!out && n)
        return fail(X3D_ERR_INVALID_ARGUMENT, "out is null");
    int rc = require_device();
    if (rc)
        return rc;
    if (n == 0)
        return X3D_OK;
    x3d::DevBuf d;
    X3D_HIP(d.ensure(sizeof(uint32_t) * n));
    hipLaunchKernelGGL(k_pcg_draws, dim3(1), dim3(64), 0, nullptr, seed, pixel, sample, n,
                       static_cast<uint32_t *>(d.p));
    X3D_HIP(hipGetLastError());
    X3D_HIP(hipMemcpy(out, d.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    return X3D_OK;
}

int x3d_render_options_default(int32_t pt_type, x3d_render_options *opts) {
    if (!opts)
        return fail(X3D_ERR_INVALID_ARGUMENT, "opts is null");
    std::memset(opts, 0, sizeof(*opts));
    opts->pt_type = pt_type;
    // The wavefront form wherever it exists: the same film, 3.5x (MIS) / 2.0x (lt1) faster on the
    // 1.3M-triangle spheres workload, 1.9x (BDPT, 6-vertex subpaths) on the 82k-triangle room.
    const bool wf = pt_type == X3D_PT_UNIDIRECT_MIS || pt_type == X3D_PT_UNIDIRECT_LT1 ||
                    pt_type == X3D_PT_UNIDIRECT || pt_type == X3D_PT_BIDIRECT_MIS;
    opts->form = wf ? X3D_FORM_AUTO : X3D_FORM_PLAIN;
    opts->max_depth = 5;
    opts->max_path_len = 8; // bidirect_mis_path_tracer::m_max_path_len (pathtracer.h:201)
    opts->seed = 1;
    return X3D_OK;
}

int x3d_renderer_create(const x3d_render_options *opts, x3d_renderer **out) {
    if (!opts || !out)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null argument");
    switch (opts->pt_type) {
    case X3D_PT_NORMAL:
    case X3D_PT_POSITION:
    case X3D_PT_DIRECT:
    case X3D_PT_UNIDIRECT:
    case X3D_PT_BIDIRECT_LT2:
    case X3D_PT_UNIDIRECT_LT1:
    case X3D_PT_UNIDIRECT_MIS:
    case X3D_PT_BIDIRECT_MIS:
        break;
    default:
        return fail(X3D_ERR_UNSUPPORTED, "tracer type " + std::to_string(opts->pt_type) + " is not implemented");
    }
    if (opts->pt_type == X3D_PT_BIDIRECT_MIS && (opts->max_path_len < 1 || opts->max_path_len > x3d::kMaxPathLen))
        return fail(X3D_ERR_INVALID_ARGUMENT,
                    "max_path_len must be in [1, " + std::to_string(x3d::kMaxPathLen) + "] for BIDIRECT_MIS");
    if (opts->form != X3D_FORM_PLAIN && opts->form != X3D_FORM_WAVEFRONT && opts->form != X3D_FORM_AUTO)
        return fail(X3D_ERR_INVALID_ARGUMENT, "unknown render form " + std::to_string(opts->form));
    if (opts->form == X3D_FORM_WAVEFRONT && opts->pt_type != X3D_PT_UNIDIRECT_MIS &&
        opts->pt_type != X3D_PT_UNIDIRECT_LT1 && opts->pt_type != X3D_PT_UNIDIRECT &&
        opts->pt_type != X3D_PT_BIDIRECT_MIS)
        return fail(X3D_ERR_UNSUPPORTED, "the wavefront form runs the UNIDIRECT_MIS, UNIDIRECT_LT1, UNIDIRECT and "
                                         "BIDIRECT_MIS tracers only");
    if (opts->form != X3D_FORM_PLAIN && opts->pt_type != X3D_PT_UNIDIRECT_MIS &&
        (opts->sort_key == X3D_SORT_MATERIAL || opts->ray_order == X3D_RAY_ORDER_SORTED))
        return fail(X3D_ERR_UNSUPPORTED, "material and ray sorting apply to the wavefront UNIDIRECT_MIS tracer only");
    if (opts->workgroup_size != 0 && opts->workgroup_size != 64 && opts->workgroup_size != 128 &&
        opts->workgroup_size != 256)
        return fail(X3D_ERR_INVALID_ARGUMENT, "workgroup_size must be 0, 64, 128 or 256");
    if (opts->sort_key < X3D_SORT_DEFAULT || opts->sort_key > X3D_SORT_NONE)
        return fail(X3D_ERR_INVALID_ARGUMENT, "unknown sort key");
    if (opts->ray_order < X3D_RAY_ORDER_DEFAULT || opts->ray_order > X3D_RAY_ORDER_APPEND)
        return fail(X3D_ERR_INVALID_ARGUMENT, "unknown ray order");
    if (opts->stack_depth != 0 && opts->stack_depth != 8 && opts->stack_depth != 16)
        return fail(X3D_ERR_INVALID_ARGUMENT, "stack_depth must be 0, 8 or 16");
    if (opts->queue_capacity > x3d::kMaxQueueCapacity)
        return fail(X3D_ERR_INVALID_ARGUMENT, "queue_capacity exceeds " + std::to_string(x3d::kMaxQueueCapacity));
    int rc = require_device();
    if (rc)
        return rc;
    auto r = std::unique_ptr<x3d_renderer>(new (std::nothrow) x3d_renderer());
    if (!r)
        return fail(X3D_ERR_OUT_OF_MEMORY, "allocation failed");
    r->opt = *opts;
    X3D_HIP(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    X3D_HIP(hipEventCreate(&r->ev0));
    X3D_HIP(hipEventCreate(&r->ev1));
    X3D_HIP(hipEventCreate(&r->evm));
    if (opts->form == X3D_FORM_WAVEFRONT || (opts->form == X3D_FORM_AUTO && has_wavefront(opts->pt_type)))
        r->wf = x3d::wf_create();
    *out = r.release();
    return X3D_OK;
}

int x3d_renderer_destroy(x3d_renderer *r) {
    delete r;
    return X3D_OK;
}


int x3d_render_device(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width,
                      uint32_t height, uint32_t spp, float *d_rgba, x3d_render_stats *stats) {
    int rc = check_render_args(r, scene, camera, width, height, spp);
    if (rc)
        return rc;
    if (!d_rgba)
        return fail(X3D_ERR_INVALID_ARGUMENT, "d_rgba is null");
    x3d::WavefrontTuning tune = wf_tuning(r->opt);
    return render_impl(r, scene, camera, width, height, spp, d_rgba, stats, form_engine(r, scene, width, height, spp), &tune);
}

namespace {
int check_trace_args(x3d_renderer *r, const x3d_scene *scene, const void *rays, uint32_t n, uint32_t num_samples,
                     const void *radiance) {
    if (!r || !rays || !radiance)
        return fail(X3D_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_scene(scene);
    if (rc)
        return rc;
    if (n == 0 || num_samples == 0)
        return fail(X3D_ERR_INVALID_ARGUMENT, "n and num_samples must be > 0");
    if (n > (1u << 30) || static_cast<uint64_t>(n) * num_samples > (1ull << 34))
        return fail(X3D_ERR_INVALID_ARGUMENT, "too many rays or samples");
    return X3D_OK;
}
} // namespace

int x3d_trace_rays_device(x3d_renderer *r, const x3d_scene *scene, const float *d_rays, uint32_t n,
                          uint32_t sample_begin, uint32_t num_samples, float *d_radiance, x3d_render_stats *stats) {
    int rc = check_trace_args(r, scene, d_rays, n, num_samples, d_radiance);
    if (rc)
        return rc;
    RayList rl{d_rays, sample_begin, d_radiance};
    return render_impl(r, scene, nullptr, n, 1, num_samples, nullptr, stats, nullptr, nullptr, nullptr, nullptr,
                       false, &rl);
}

int x3d_trace_rays(x3d_renderer *r, const x3d_scene *scene, const float *rays, uint32_t n, uint32_t sample_begin,
                   uint32_t num_samples, float *radiance, x3d_render_stats *stats) {
    int rc = check_trace_args(r, scene, rays, n, num_samples, radiance);
    if (rc)
        return rc;
    for (uint64_t k = 0; k < 6ull * n; k++)
        if (!std::isfinite(rays[k]))
            return fail(X3D_ERR_INVALID_ARGUMENT, "rays must be finite");
    auto t0 = std::chrono::steady_clock::now();
    const size_t out_bytes = 3ull * n * num_samples * sizeof(float);
    X3D_HIP(r->trace_rays.ensure(6ull * n * sizeof(float)));
    X3D_HIP(r->trace_out.ensure(out_bytes));
    X3D_HIP(hipMemcpy(r->trace_rays.p, rays, 6ull * n * sizeof(float), hipMemcpyHostToDevice));
    RayList rl{static_cast<const float *>(r->trace_rays.p), sample_begin, static_cast<float *>(r->trace_out.p)};
    rc = render_impl(r, scene, nullptr, n, 1, num_samples, nullptr, stats, nullptr, nullptr, nullptr, nullptr, false,
                     &rl);
    if (rc)
        return rc;
    X3D_HIP(hipMemcpy(radiance, r->trace_out.p, out_bytes, hipMemcpyDeviceToHost));
    if (stats)
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return X3D_OK;
}

int x3d_render_device_noalloc(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width,
                              uint32_t height, uint32_t spp, float *d_rgba, x3d_render_stats *stats) {
    int rc = check_render_args(r, scene, camera, width, height, spp);
    if (rc)
        return rc;
    if (!d_rgba)
        return fail(X3D_ERR_INVALID_ARGUMENT, "d_rgba is null");
    void *base = nullptr;
    if (int rc2 = scratch_for_current_device(&base))
        return rc2;
    ScratchArena arena{static_cast<char *>(base), kScratchBytes, 0};
    return render_impl(r, scene, camera, width, height, spp, d_rgba, stats, nullptr, nullptr, nullptr, nullptr, false,
                       nullptr, &arena);
}

int x3d_render(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width, uint32_t height,
               uint32_t spp, int32_t firefly_filter, float *rgba, x3d_render_stats *stats) {
    int rc = check_render_args(r, scene, camera, width, height, spp);
    if (rc)
        return rc;

    if (!rgba)
        return fail(X3D_ERR_INVALID_ARGUMENT, "rgba is null");
    uint64_t npix = static_cast<uint64_t>(width) * height;
    X3D_HIP(r->out.ensure(npix * sizeof(float4)));
    auto t0 = std::chrono::steady_clock::now();
    x3d::WavefrontTuning tune = wf_tuning(r->opt);
    rc = render_impl(r, scene, camera, width, height, spp, static_cast<float *>(r->out.p), stats,
                     form_engine(r, scene, width, height, spp), &tune, nullptr, nullptr, firefly_filter != 0);
    if (rc)
        return rc;
    X3D_HIP(hipMemcpy(rgba, r->out.p, npix * sizeof(float4), hipMemcpyDeviceToHost));
    if (stats)
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return X3D_OK;
}

int x3d_renderer_autotune(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width,
                          uint32_t height, uint32_t probe_spp, x3d_autotune_report *report) {
    int rc = check_render_args(r, scene, camera, width, height, probe_spp);
    if (rc)
        return rc;
    if (!r->wf)
        return fail(X3D_ERR_STATE, "autotuning applies to the wavefront form");
    uint64_t npix = static_cast<uint64_t>(width) * height;
    X3D_HIP(r->out.ensure(npix * sizeof(float4)));
    x3d_render_options saved = r->opt;
    r->opt.profile = 0;
    auto time_config = [&](uint32_t wg, uint32_t depth, int32_t sort, uint32_t cap, double *ms) -> int {
        r->opt.workgroup_size = wg;
        r->opt.stack_depth = depth;
        r->opt.sort_key = sort;
        r->opt.queue_capacity = cap;
        x3d_render_stats st;
        double best = 1e300;
        for (int k = 0; k < 3; k++) { // first call captures the graph; best of the other two
            x3d::WavefrontTuning tune = wf_tuning(r->opt);
            int e = render_impl(r, scene, camera, width, height, probe_spp, static_cast<float *>(r->out.p), &st,
                                r->wf, &tune);
            if (e)
                return e;
            if (k > 0 && st.trace_ms < best)
                best = st.trace_ms;
        }
        *ms = best;
        return X3D_OK;
    };
    auto consider = [&](x3d_autotune_report &rep, uint32_t wg, uint32_t d, int32_t so, uint32_t cap) -> int {
        double ms;
        int e = time_config(wg, d, so, cap, &ms);
        if (e)
            return e;
        rep.candidates++;
        if (ms < rep.best_ms) {
            rep.best_ms = ms;
            rep.workgroup_size = wg;
            rep.stack_depth = d;
            rep.sort_key = so;
            rep.queue_capacity = cap;
        }
        return X3D_OK;
    };
    // Stage 1: workgroup size x LDS stack depth x sort key at the default capacity. Stage 2: the
    // queue capacity for the winner. A capacity above the probe's path count is never tried: the
    // probe would clamp it to one generation and could not tell it apart from a smaller one.
    const uint32_t wgs[] = {64, 128, 256};
    const uint32_t depths[] = {8, 16};
    const int32_t sorts_mis[] = {X3D_SORT_MATERIAL, X3D_SORT_NONE};
    // the lt1 / unidirect wavefront runs unsorted only
    const bool mis = r->opt.pt_type == X3D_PT_UNIDIRECT_MIS;
    const std::vector<int32_t> sorts(sorts_mis + (mis ? 0 : 1), sorts_mis + 2);
    const uint32_t caps[] = {1u << 21, 1u << 22, 1u << 23, 1u << 24, 1u << 25, 1u << 26};
    const uint64_t probe_paths = npix * probe_spp;
    x3d_autotune_report rep{};
    rep.best_ms = 1e300;
    rc = time_config(0, 0, X3D_SORT_DEFAULT, 0, &rep.default_ms);
    for (uint32_t wg : wgs)
        for (uint32_t d : depths)
            for (int32_t so : sorts)
                if (!rc)
                    rc = consider(rep, wg, d, so, x3d::kDefaultQueueCapacity);
    const uint32_t wg1 = rep.workgroup_size, d1 = rep.stack_depth;
    const int32_t so1 = rep.sort_key;
    for (uint32_t cap : caps)
        if (!rc && cap != x3d::kDefaultQueueCapacity && cap <= probe_paths)
            rc = consider(rep, wg1, d1, so1, cap);
    if (rc) {
        r->opt = saved;
        return rc;
    }
    // Timing noise between near-equal candidates must not move the configuration: the defaults
    // stay unless the winner is at least 2% faster.
    if (rep.best_ms > 0.98 * rep.default_ms) {
        x3d::WavefrontTuning def;
        rep.workgroup_size = static_cast<uint32_t>(def.trace_wg);
        rep.stack_depth = static_cast<uint32_t>(def.stack_depth);
        rep.sort_key = def.sort_by_material ? X3D_SORT_MATERIAL : X3D_SORT_NONE;
        rep.queue_capacity = def.paths_in_flight;
        rep.best_ms = rep.default_ms;
    }
    r->opt = saved;
    r->opt.workgroup_size = rep.workgroup_size;
    r->opt.stack_depth = rep.stack_depth;
    r->opt.sort_key = rep.sort_key;
    r->opt.queue_capacity = rep.queue_capacity;
    if (report)
        *report = rep;
    return X3D_OK;
}

int x3d_render_batch_device(x3d_renderer *r, uint32_t num_jobs, x3d_batch_job *jobs, x3d_batch_stats *stats) {
    int rc = check_batch(r, num_jobs, jobs);
    if (rc)
        return rc;
    std::vector<float *> films(num_jobs);
    for (uint32_t i = 0; i < num_jobs; i++)
        films[i] = jobs[i].rgba;
    return batch_impl(r, num_jobs, jobs, films.data(), stats);
}

int x3d_render_batch(x3d_renderer *r, uint32_t num_jobs, x3d_batch_job *jobs, x3d_batch_stats *stats) {
    int rc = check_batch(r, num_jobs, jobs);
    if (rc)
        return rc;
    auto t0 = std::chrono::steady_clock::now();
    if (r->batch_films.size() < num_jobs)
        r->batch_films = std::vector<x3d::DevBuf>(num_jobs);
    std::vector<float *> films(num_jobs);
    for (uint32_t i = 0; i < num_jobs; i++) {
        X3D_HIP(r->batch_films[i].ensure(static_cast<size_t>(jobs[i].width) * jobs[i].height * sizeof(float4)));
        films[i] = static_cast<float *>(r->batch_films[i].p);
    }
    rc = batch_impl(r, num_jobs, jobs, films.data(), stats);
    if (rc)
        return rc;
    for (uint32_t i = 0; i < num_jobs; i++)
        X3D_HIP(hipMemcpy(jobs[i].rgba, films[i], static_cast<size_t>(jobs[i].width) * jobs[i].height * sizeof(float4),
                          hipMemcpyDeviceToHost));
    if (stats)
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return X3D_OK;
}

int x3d_renderer_autotune_batch(x3d_renderer *r, uint32_t num_jobs, x3d_batch_job *jobs, x3d_autotune_report *report) {
    int rc = check_batch(r, num_jobs, jobs);
    if (rc)
        return rc;
    x3d_render_options saved = r->opt;
    x3d_autotune_report rep{};
    rep.best_ms = 1e300;
    const uint32_t wgs[] = {64, 128, 256};
    const uint32_t depths[] = {8, 16};
    // the plain form's shared launch: tile order by cost or by job; the wavefront form has no sort
    std::vector<int32_t> sorts = {X3D_SORT_MATERIAL, X3D_SORT_NONE};
    if (r->opt.form != X3D_FORM_PLAIN)
        sorts = {X3D_SORT_DEFAULT};
    auto time_config = [&](uint32_t wg, uint32_t d, int32_t so, double *ms) -> int {
        r->opt.workgroup_size = wg;
        r->opt.stack_depth = d;
        r->opt.sort_key = so;
        double best = 1e300;
        for (int k = 0; k < 3; k++) { // first call captures the graph; best of the other two
            x3d_batch_stats st;
            int e = x3d_render_batch(r, num_jobs, jobs, &st);
            if (e)
                return e;
            if (k > 0 && st.kernel_ms < best)
                best = st.kernel_ms;
        }
        *ms = best;
        return X3D_OK;
    };
    rc = time_config(0, 0, X3D_SORT_DEFAULT, &rep.default_ms);
    for (uint32_t wg : wgs)
        for (uint32_t d : depths)
            for (int32_t so : sorts) {
                double ms = 0;
                if (!rc)
                    rc = time_config(wg, d, so, &ms);
                if (rc)
                    break;
                rep.candidates++;
                if (ms < rep.best_ms) {
                    rep.best_ms = ms;
                    rep.workgroup_size = wg;
                    rep.stack_depth = d;
                    rep.sort_key = so;
                }
            }
    r->opt = saved;
    if (rc)
        return rc;
    r->opt.workgroup_size = rep.workgroup_size;
    r->opt.stack_depth = rep.stack_depth;
    r->opt.sort_key = rep.sort_key;
    if (report)
        *report = rep;
    return X3D_OK;
}

int x3d_render_forward(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width,
                       uint32_t height, uint32_t spp, float *rgba, x3d_render_stats *stats) {
    return x3d_render(r, scene, camera, width, height, spp, 0, rgba, stats);
}

namespace {
int check_backward(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width,
                   uint32_t height, uint32_t spp, const void *dl, const void *grad) {
    int rc = check_render_args(r, scene, camera, width, height, spp);
    if (rc)
        return rc;
    if (!dl || !grad)
        return fail(X3D_ERR_INVALID_ARGUMENT, "dl_dfilm or grad is null");
    if (r->opt.pt_type != X3D_PT_UNIDIRECT_MIS)
        return fail(X3D_ERR_STATE, "the adjoint needs a UNIDIRECT_MIS renderer");
    if (r->opt.max_depth > x3d::kAdjMaxDepth)
        return fail(X3D_ERR_UNSUPPORTED, "the adjoint supports max_depth <= " + std::to_string(x3d::kAdjMaxDepth));
    if (scene->s.materials.size() - 1 > x3d::kAdjMaxMats)
        return fail(X3D_ERR_UNSUPPORTED, "the adjoint supports at most " + std::to_string(x3d::kAdjMaxMats) +
                                             " materials");
    return X3D_OK;
}
} // namespace

int x3d_render_backward_device(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width,
                               uint32_t height, uint32_t spp, const float *d_dl_dfilm, float *d_grad,
                               x3d_render_stats *stats) {
    int rc = check_backward(r, scene, camera, width, height, spp, d_dl_dfilm, d_grad);
    if (rc)
        return rc;
    return render_impl(r, scene, camera, width, height, spp, nullptr, stats, nullptr, nullptr,
                       reinterpret_cast<const float4 *>(d_dl_dfilm), d_grad);
}

int x3d_render_backward(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width,
                        uint32_t height, uint32_t spp, const float *dl_dfilm, float *grad, x3d_render_stats *stats) {
    int rc = check_backward(r, scene, camera, width, height, spp, dl_dfilm, grad);
    if (rc)
        return rc;
    auto t0 = std::chrono::steady_clock::now();
    size_t npix = static_cast<size_t>(width) * height;
    size_t ng = 4 * (scene->s.materials.size() - 1);
    X3D_HIP(r->adj_dl.ensure(npix * sizeof(float4)));
    X3D_HIP(r->adj_grad.ensure(std::max<size_t>(ng, 4) * sizeof(float)));
    X3D_HIP(hipMemcpy(r->adj_dl.p, dl_dfilm, npix * sizeof(float4), hipMemcpyHostToDevice));
    rc = render_impl(r, scene, camera, width, height, spp, nullptr, stats, nullptr, nullptr,
                     static_cast<const float4 *>(r->adj_dl.p), static_cast<float *>(r->adj_grad.p));
    if (rc)
        return rc;
    X3D_HIP(hipMemcpy(grad, r->adj_grad.p, ng * sizeof(float), hipMemcpyDeviceToHost));
    if (stats)
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return X3D_OK;
}

} // extern "C"
