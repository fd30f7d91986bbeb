// Pieces shared by the plain and wavefront render kernels: camera rays, counters, traversal
// stacks and the per-render parameter block.
#pragma once

#include <hip/hip_runtime.h>

#include "integrators.h"

namespace x3d {

constexpr int kWG = 256;      // lanes per workgroup (4 waves)
constexpr int kLdsDepth = 16; // LDS stack entries per lane (16 x 8 B x 256 = 32 KiB per WG)

struct Counters { // device-side totals (uint64)
    unsigned long long closest, anyhit, pops, tests, spops, stests;
};

struct CamParams {
    float sensor, focal, aspect;
    float m[16];
};

// pinhole_camera::sample (camera.cpp:38-48) after pinhole_camera::transform (camera.cpp:63-71).
X3D_DI void camera_ray(CamParams const &c, uint32_t x, uint32_t y, uint32_t w, uint32_t h, f3 &o, f3 &v) {
    float s = c.sensor;
    float dx = (static_cast<float>(x) / static_cast<float>(w - 1) * s - s / 2.0f) * c.aspect;
    float dy = (static_cast<float>(h - 1) - static_cast<float>(y)) / static_cast<float>(h - 1) * s - s / 2.0f;
    float dz = -c.focal;
    float r[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float a = 0.0f;
        a = a + c.m[0 * 4 + i] * dx;
        a = a + c.m[1 * 4 + i] * dy;
        a = a + c.m[2 * 4 + i] * dz;
        a = a + 0.0f * 0.0f;
        r[i] = a;
    }
    o = mk(c.m[12], c.m[13], c.m[14]);
    v = normalize(mk(r[0], r[1], r[2]));
}

// Wave-level sum then one atomic per wave (all lanes of the wave must be converged).
X3D_DI void flush_counter(unsigned long long *dst, uint32_t v) {
    unsigned long long x = v;
    for (int off = 32; off > 0; off >>= 1)
        x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0 && x != 0)
        atomicAdd(dst, x);
}

template <int WG, int LDSD = kLdsDepth>
X3D_DI TravStack<WG, LDSD> make_stack(uint2 *lds, uint2 *spill, uint32_t spill_per_thread) {
    TravStack<WG, LDSD> st;
    st.lds = lds + threadIdx.x;
    uint32_t gtid = blockIdx.x * WG + threadIdx.x;
    st.spill = spill + static_cast<size_t>(gtid) * spill_per_thread;
    st.sp = 0;
    return st;
}

// Stack entries spilled to global memory per lane for a tree of the scene's depth.
inline uint32_t spill_entries_for(DevScene const &S, int lds_depth) {
    return S.max_stack > static_cast<uint32_t>(lds_depth) ? S.max_stack - lds_depth + 1 : 1;
}

// Growable device buffer (host side; contents are not preserved when it grows).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(DevBuf const &) = delete;
    DevBuf &operator=(DevBuf const &) = delete;
    ~DevBuf() {
        if (p)
            (void)hipFree(p);
    }
    hipError_t ensure(size_t n) {
        if (n <= bytes)
            return hipSuccess;
        if (p)
            (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess)
            bytes = n;
        return e;
    }
};

struct RenderParams {
    DevScene S;
    CamParams cam;
    uint32_t width, height;
    uint32_t s_begin, s_end, spp_total;
    uint64_t seed;
    uint32_t max_depth, max_path_len;
    float4 *first; // 3 x float4 per pixel: (p, t), (n, geo), (mat, light, valid, -)
    float4 *acc;   // running rgb sum per pixel
    float4 *out;   // final rgba per pixel (written when finalize != 0)
    uint2 *spill;
    uint32_t spill_per_thread;
    int32_t finalize;
    Counters *cnt;
    int32_t firefly; // firefly filter on the sample images (sampling_task::run, renderer.cpp:46-83)
    const float *rays; // caller rays (origin, direction: 6 floats per "pixel") instead of the
                       // camera's, or null (x3d_trace_rays)
    float4 *subpaths;  // plain bidirectional kernel: both subpaths of every resident lane
                       // (2 x kMaxPathLen x 4 float4 rows per lane)
};

// The primary ray of pixel `pix`: the caller's ray in ray-list mode, else the camera's.
X3D_DI void pixel_ray(RenderParams const &P, uint32_t pix, f3 &o, f3 &v) {
    if (P.rays) {
        const float *r = P.rays + 6u * static_cast<size_t>(pix);
        o = mk(r[0], r[1], r[2]);
        v = mk(r[3], r[4], r[5]);
    } else {
        camera_ray(P.cam, pix % P.width, pix / P.width, P.width, P.height, o, v);
    }
}

// Firefly filter of sampling_task::run (renderer.cpp:46-83) for pixel (x, y) of one sample image
// `img` (rgb per pixel, row-major, width w, height h): an interior pixel whose squared norm
// exceeds each of its 8 neighbours' by more than FireFlyMinDiff = 1.5 (renderer.h:118) is capped,
// component by component, at the mean of the neighbours' squared norms plus 1.5; edge pixels and
// all others are returned unchanged.
X3D_DI float4 firefly_filtered(const float4 *img, uint32_t x, uint32_t y, uint32_t w, uint32_t h) {
    const float4 c = img[x + y * w];
    if (x == 0 || x + 1 >= w || y == 0 || y + 1 >= h)
        return c;
    auto r = [&](uint32_t i, uint32_t j) {
        float4 q = img[i + j * w];
        return dot(mk(q.x, q.y, q.z), mk(q.x, q.y, q.z));
    };
    constexpr float kMinDiff = 1.5f;
    const float r00 = r(x - 1, y - 1), r10 = r(x, y - 1), r20 = r(x + 1, y - 1);
    const float r01 = r(x - 1, y), r11 = r(x, y), r21 = r(x + 1, y);
    const float r02 = r(x - 1, y + 1), r12 = r(x, y + 1), r22 = r(x + 1, y + 1);
    const bool firefly = (r11 - r00) > kMinDiff && (r11 - r10) > kMinDiff && (r11 - r20) > kMinDiff &&
                         (r11 - r01) > kMinDiff && (r11 - r21) > kMinDiff && (r11 - r02) > kMinDiff &&
                         (r11 - r12) > kMinDiff && (r11 - r22) > kMinDiff;
    if (!firefly)
        return c;
    const float cap = 1.0f / 8.0f * (r00 + r10 + r20 + r01 + r21 + r02 + r12 + r22) + kMinDiff;
    return make_float4(cap < c.x ? cap : c.x, cap < c.y ? cap : c.y, cap < c.z ? cap : c.z, c.w);
}

X3D_DI HitRec load_first(RenderParams const &P, uint32_t pix, bool &valid) {
    float4 a = P.first[3 * pix], b = P.first[3 * pix + 1], c = P.first[3 * pix + 2];
    HitRec first;
    first.p = mk(a.x, a.y, a.z);
    first.t = a.w;
    first.n = mk(b.x, b.y, b.z);
    first.geo = static_cast<int32_t>(__float_as_uint(b.w));
    first.mat = static_cast<int32_t>(__float_as_uint(c.x));
    first.light = static_cast<int32_t>(__float_as_uint(c.y));
    valid = c.z != 0.0f;
    return first;
}

} // namespace x3d
