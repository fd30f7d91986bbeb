// Path-replay adjoint of the fixed-depth MIS tracer with respect to material parameters.
//
// Each (pixel, sample) path is traced again from its PCG stream (nothing from the forward pass
// is stored), recording per vertex the directions and the NEE / emission contributions it made.
// With the sampled directions, the sampling densities and the MIS weights held fixed (the
// sampling density of a GGX lobe is a parameter of its own, see x3d_scene_set_material_params),
// the derivative of a contribution is the contribution times the sum of d log f over the BSDF
// evaluations it contains (Vicini, Speierer, Jakob 2021, "Path Replay Backpropagation",
// specialised to BSDF parameters). Walking the recorded vertices backwards, `after` holds the
// contributions made after the BSDF sample of the current vertex, which all contain its f.
//
// Gradients are summed in 64-bit fixed point (2^-32 resolution, |sum| < 2^31): each contribution
// is rounded once to an integer multiple of 2^-32 and added with integer atomics into per-wave
// slots in LDS, then per block and over blocks. Integer addition is associative, so the result
// does not depend on the order of the additions, the lane order of same-address atomics, the
// grid size or the number of CUs: it is the same on every run and every device.

#include "adjoint.h"

#include <algorithm>

namespace x3d {

namespace {

// Scalar factor s of a non-mixture lobe (the lobe is s * albedo) and ds/droughness.
// cook_torr: s = F D G / (4 cos_i cos_o) with u = alpha^2 = 2 r^2:
//   dD/du = D (t - u) / (u (u + t)), t = tan^2 of the half vector's angle to n;
//   dG/du = -G^2 (dLambda_i + dLambda_o), dLambda/du = tan^2 / (4 sqrt(1 + u tan^2));
//   ds/dr = s (dD/du / D + dG/du / G) 4 r.
X3D_DI void lobe_scalar(DevMaterial const &m, f3 n, f3 o, f3 i, float *sv, float *ds_dr) {
    *sv = 0.0f;
    *ds_dr = 0.0f;
    if (m.kind == kMatOrenNayar) {
        float k;
        if (oren_nayar_k(m, n, o, i, &k))
            *sv = (1.0f / kPi) * k;
        return;
    }
    if (m.kind != kMatCookTorr)
        return;
    float cos_o = smax(0.0f, dot(n, o));
    float cos_i = smax(0.0f, dot(n, i));
    if (near0(cos_i) || near0(cos_o))
        return;
    f3 h = i + o;
    if (near0(h))
        return;
    h = normalize(h);
    float u = m.alpha2;
    float f = fresnel(m, i, h);
    float d = ggx_d(u, n, h);
    float g = ggx_g(u, i, o, n);
    *sv = f * d * g / (4.0f * cos_i * cos_o);
    float ct = dot(n, h);
    float t = (1.0f - ct * ct) / (ct * ct);
    auto dlam = [&](f3 w) {
        float c = dot(n, w);
        float t2 = (1.0f - c * c) / (c * c);
        return t2 / (4.0f * __builtin_sqrtf(1.0f + u * t2));
    };
    float dlog = (t - u) / (u * (u + t)) - g * (dlam(i) + dlam(o));
    *ds_dr = *sv * dlog * 4.0f * m.roughness;
}

constexpr double kAdjFixScale = 4294967296.0;      // 2^32
constexpr float kAdjFixMax = 9.2e18f;                // clamp below 2^63

// A contribution in 64-bit fixed point (round to nearest), as the unsigned bit pattern the LDS
// atomics add.
X3D_DI unsigned long long to_fixed(float v) {
    double x = static_cast<double>(v) * kAdjFixScale;
    x = x > kAdjFixMax ? kAdjFixMax : (x < -kAdjFixMax ? -kAdjFixMax : x); // NaN stays NaN
    long long q = x == x ? __double2ll_rn(x) : 0ll;
    return static_cast<unsigned long long>(q);
}

// Adds Q-weighted d log f_c (per channel) of material `mat` at (n, o, i) to the slots
// (4 per material: d albedo rgb, d roughness): f_c = sum over lobes w_l s_l a_{l,c}.
X3D_DI void add_bsdf_grad(DevScene const &S, int32_t mat, f3 n, f3 o, f3 i, f3 Q, unsigned long long *slots) {
    if (mat < 0 || static_cast<uint32_t>(mat) >= S.fail_safe_mat)
        return; // the fail-safe material has no parameters
    f3 f = bsdf_eval(S.mats, mat, n, o, i);
    const float fc[3] = {f.x, f.y, f.z}, qc[3] = {Q.x, Q.y, Q.z};
    auto lobe = [&](int32_t lm, float wgt) {
        DevMaterial const &L = S.mats[lm];
        if (static_cast<uint32_t>(lm) >= kAdjMaxMats)
            return;
        float sv, ds;
        lobe_scalar(L, n, o, i, &sv, &ds);
        float gr = 0.0f;
        for (int c = 0; c < 3; c++) {
            if (fc[c] == 0.0f)
                continue;
            atomicAdd(&slots[4 * lm + c], to_fixed(qc[c] * wgt * sv / fc[c]));
            gr += qc[c] * wgt * L.albedo[c] * ds / fc[c];
        }
        if (L.kind == kMatCookTorr)
            atomicAdd(&slots[4 * lm + 3], to_fixed(gr));
    };
    DevMaterial const &m = S.mats[mat];
    if (m.kind == kMatMixture) {
        lobe(m.sub0, m.ratio);
        lobe(m.sub1, 1.0f - m.ratio);
    } else {
        lobe(mat, 1.0f);
    }
}

constexpr int kAdjWaves = kWG / 64;

// Built for 4 waves per SIMD (128 VGPRs, a few spilled): the compiler's own choice, 3 waves
// without spills, measured 10% slower.
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(4))) void
k_adjoint(RenderParams P, const float4 *dl, uint32_t num_mats, unsigned long long *partial, float4 *records) {
    __shared__ uint2 lds[kLdsDepth * kWG];
    __shared__ unsigned long long slots[kAdjWaves][4 * kAdjMaxMats];
    for (uint32_t i = threadIdx.x; i < kAdjWaves * 4 * kAdjMaxMats; i += kWG)
        (&slots[0][0])[i] = 0ull;
    __syncthreads();
    TravStack<kWG, kLdsDepth> st = make_stack<kWG>(lds, P.spill, P.spill_per_thread);
    Tracer<kWG, kLdsDepth, false> tr{P.S, st};
    unsigned long long *my = slots[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];
    const uint32_t npix = P.width * P.height;
    const float inv = 1.0f / static_cast<float>(P.spp_total);
    const AdjRecord rec = AdjRecord::of_lane(records, (blockIdx.x * kWG + threadIdx.x) >> 6);
    // item q = s * npix + pix, grid-strided; (pix, s) advance by the stride's quotient and
    // remainder instead of dividing a 64-bit q every item
    const uint32_t q0 = blockIdx.x * kWG + threadIdx.x, stride = gridDim.x * kWG;
    const uint32_t ds = stride / npix, dp = stride - ds * npix;
    uint32_t pix = q0 % npix, s = q0 / npix;
    for (; s < P.spp_total; s += ds, pix += dp, s += pix >= npix ? 1u : 0u, pix -= pix >= npix ? npix : 0u) {
        f3 o, v;
        camera_ray(P.cam, pix % P.width, pix / P.width, P.width, P.height, o, v);
        bool valid;
        HitRec first = load_first(P, pix, valid);
        PathState ps;
        tr.mis_begin(ps, first, valid, v, pcg_for(P.seed, pix, s), P.max_depth);
        uint32_t nrec = 0;
        while (ps.alive) {
            nrec = ps.k + 1;
            tr.mis_bounce<true>(ps, P.max_depth, &rec);
        }
        const float4 d = dl[pix];
        const f3 gw = mk(d.x * inv, d.y * inv, d.z * inv);
        f3 tail = mk(0.0f, 0.0f, 0.0f);
        for (int k = static_cast<int>(nrec) - 1; k >= 0; k--) {
            const float4 r0 = rec.row(k, 0), r1 = rec.row(k, 1);
            const int32_t mat = static_cast<int32_t>(__float_as_uint(r0.w));
            const uint32_t flags = __float_as_uint(r1.w);
            const f3 n = xyz(r0), wo = xyz(r1), c_nee = xyz(rec.row(k, 4));
            f3 after = xyz(rec.row(k, 5)) + tail;
            if (flags & kAdjCont)
                add_bsdf_grad(P.S, mat, n, wo, xyz(rec.row(k, 2)), mul(gw, after), my);
            if (flags & kAdjNee)
                add_bsdf_grad(P.S, mat, n, wo, xyz(rec.row(k, 3)), mul(gw, c_nee), my);
            tail = after + c_nee;
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
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 4 * num_mats; i += kWG) {
        unsigned long long acc = 0ull;
        for (int w = 0; w < kAdjWaves; w++)
            acc += slots[w][i];
        partial[static_cast<size_t>(blockIdx.x) * 4 * num_mats + i] = acc;
    }
}

// Fixed-point sums over blocks (exact, wrap-around two's complement) -> float gradient: one
// block per parameter, its threads striding over the per-block partials, then a tree in LDS
// (integer sums: the grouping does not change the result).
__global__ __launch_bounds__(256) void k_adjoint_reduce(const unsigned long long *partial, uint32_t blocks,
                                                        uint32_t n, float *grad) {
    __shared__ unsigned long long red[256];
    const uint32_t i = blockIdx.x;
    unsigned long long acc = 0ull;
    for (uint32_t b = threadIdx.x; b < blocks; b += 256)
        acc += partial[static_cast<size_t>(b) * n + i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        grad[i] = static_cast<float>(static_cast<double>(static_cast<long long>(red[0])) / kAdjFixScale);
}

} // namespace

struct AdjointEngine {
    DevBuf partial, spill, records;
};

AdjointEngine *adjoint_create() { return new AdjointEngine(); }
void adjoint_destroy(AdjointEngine *e) { delete e; }

std::string adjoint_render(AdjointEngine *e, RenderParams const &P, const float4 *d_dl, uint32_t num_mats,
                           float *d_grad, hipStream_t stream) {
    if (num_mats > kAdjMaxMats)
        return "the adjoint supports at most " + std::to_string(kAdjMaxMats) + " materials";
    if (P.max_depth > kAdjMaxDepth)
        return "the adjoint supports max_depth <= " + std::to_string(kAdjMaxDepth);
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess)
        return "hipGetDevice failed";
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_adjoint, kWG, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    uint64_t items = static_cast<uint64_t>(P.width) * P.height * P.spp_total;
    uint64_t want = (items + kWG - 1) / kWG;
    uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(want, static_cast<uint64_t>(cus) * per_cu));
    if (grid < 1)
        grid = 1;
    uint32_t n = 4 * (num_mats ? num_mats : 1);
    if (hipError_t err = e->partial.ensure(sizeof(unsigned long long) * static_cast<size_t>(grid) * n);
        err != hipSuccess)
        return std::string("hipMalloc: ") + hipGetErrorString(err);
    RenderParams Q = P;
    Q.spill_per_thread = spill_entries_for(P.S, kLdsDepth);
    if (hipError_t err = e->spill.ensure(sizeof(uint2) * static_cast<size_t>(grid) * kWG * Q.spill_per_thread);
        err != hipSuccess)
        return std::string("hipMalloc: ") + hipGetErrorString(err);
    Q.spill = static_cast<uint2 *>(e->spill.p);
    // path records: kAdjRows float4 per vertex and resident lane (~150 MB for 256 CUs)
    if (hipError_t err = e->records.ensure(sizeof(float4) * kAdjRows * kAdjMaxDepth * static_cast<size_t>(grid) * kWG);
        err != hipSuccess)
        return std::string("hipMalloc: ") + hipGetErrorString(err);
    hipLaunchKernelGGL(k_adjoint, dim3(grid), dim3(kWG), 0, stream, Q, d_dl, num_mats,
                       static_cast<unsigned long long *>(e->partial.p), static_cast<float4 *>(e->records.p));
    if (num_mats > 0)
        hipLaunchKernelGGL(k_adjoint_reduce, dim3(4 * num_mats), dim3(256), 0, stream,
                           static_cast<const unsigned long long *>(e->partial.p), grid, 4 * num_mats, d_grad);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess)
        return std::string("adjoint launch: ") + hipGetErrorString(err);
    return "";
}

} // namespace x3d
