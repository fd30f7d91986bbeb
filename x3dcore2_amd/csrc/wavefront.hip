// Wavefront form of the fixed-depth MIS tracer, and of the reference-mode lt1 and unidirect
// tracers (Russian roulette, no depth limit below kRefModeMaxDepth) in the fused form.
//
// Paths live in HBM as structure-of-arrays state; each bounce runs as separate kernels:
//   reset      clear the queue counters of this bounce
//   nee_bsdf   per active path: light sample (shadow ray + MIS-weighted contribution queued),
//              BSDF sample (extension ray queued, or the path finishes)
//   shadow     persistent any-hit traversal over the shadow queue; adds unoccluded contributions
//   extend     persistent closest-hit traversal over the extension queue
//   shade      per extension result: throughput update, MIS-weighted emission, next vertex;
//              surviving paths are appended to per-material-kind active lists (material sort)
// The traversal kernels keep every lane busy: a lane whose ray finishes writes its result and
// takes the next ray from its wave's private batch (one atomic per 128 rays on the shared
// queue) without waiting for the rest of its wave. Queue counters live on separate cache lines
// and appends are aggregated per wave. All paths of a generation (a fixed number of samples for every pixel) start
// together, so they all sit at the same depth; one generation's launches are captured once as
// a hipGraph and replayed. Each path performs exactly the operations of the plain kernel in the
// same order (same PCG stream, same float expressions), and each pixel's samples are summed in
// sample order after their generation, so the film is bit-identical to the plain form's.

#include "wavefront.h"

#include <cstring>
#include <vector>

namespace x3d {

constexpr int kKinds = 3; // diffuse (Oren-Nayar / fail-safe), GGX, mixture

// Every queue counter sits alone in its own 256-byte block: counters on one cache line would
// serialize each other's atomics at a single L2 channel.
struct alignas(256) WfCounter {
    uint32_t v;
};

struct WfCtl {
    WfCounter act_count[2][4];
    WfCounter sq_count, sq_head, eq_count, eq_head, fin_count;
    WfCounter eq_count_b, fin_count_b; // second parity of the fused form's double-buffered queues
    uint32_t gen;
    uint32_t pad[2];
};

struct WfArgs {
    DevScene S;
    CamParams cam;
    uint32_t width, height, npix;
    uint32_t s_begin, s_end, spp_total, gen_samples, capacity;
    uint64_t seed;
    uint32_t max_depth;
    int32_t finalize;
    float4 *first, *acc, *out;
    // path state, index = path id = s_local * npix + pixel
    float4 *xp, *xn, *wo, *beta, *L, *wi;
    uint64_t *rng;
    float4 *lbuf;
    // queues
    uint32_t *act[2];     // [kind * capacity + j]
    float4 *sq;           // 3 x float4 per shadow ray: (o, tmax), (d, path), (contribution, -)
    uint32_t *eq, *fin;   // extension / finishing path ids
    ClosestHit *eh;       // extension results
    WfCtl *ctl;
    uint2 *spill;
    uint32_t spill_per_thread;
    Counters *cnt;
    int sort_by_material;
    uint32_t refill_below; // traversal: leave the descent when fewer lanes still seek a leaf
    uint32_t leaf_below;   // traversal: leave the leaf tests when fewer lanes still hold a leaf
    int32_t firefly;       // firefly filter on each generation's sample images
    // Ray reordering: the traversal kernels (and shade) read sq_in / eq_in, which are either the
    // append-order queues or their copies sorted by ray key (queue 0 shadow, 1 extension).
    int sort_rays;
    const float4 *sq_in;
    const uint32_t *eq_in;
    float4 *sq_sorted;
    uint32_t *eq_sorted;
    uint32_t *ray_key[2], *ray_rank[2]; // per queue entry
    uint32_t *bin_mat[2];               // [kSortBlocks][kSortBins] counts, then offsets
    uint32_t *bin_tot[2];               // [kSortBins]
    // Consumer (in) and producer (out) sides of the extension and finishing queues: the same
    // buffers and counters, or, in the fused form, the two parities of a pair double-buffered by
    // bounce (shade reads bounce k's queues while it appends bounce k + 1's).
    const uint32_t *fin_in;
    uint32_t *eq_out, *fin_out;
    WfCounter *eq_cnt_in, *fin_cnt_in, *eq_cnt_out, *fin_cnt_out;
    // Fused form (no material sort, no ray sort): generate and shade also run the next vertex's
    // light sample and BSDF sample (no nee_bsdf launch, no active lists).
    int fused;
    // lt1: the first hit's emission per path, added when the path ends (w = 1 when there is one).
    float4 *le0;
};

// Ray keys: direction octant (3 bits) above the Morton code of the origin's cell in an 8^3 grid
// over the scene box (9 bits). Rays with equal keys start close together and head the same
// general way, so a wave of them visits largely the same nodes.
constexpr uint32_t kSortBins = 4096;
constexpr uint32_t kSortBlocks = 128; // blocks of the count / scatter passes (rows of bin_mat)
constexpr uint32_t kSortWG = 1024;

X3D_DI uint32_t spread3(uint32_t v) { return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4); }

X3D_DI uint32_t ray_key(DevScene const &S, f3 o, f3 d) {
    // fmaxf first: a NaN coordinate lands in cell 0
    uint32_t cx = static_cast<uint32_t>(fminf(fmaxf((o.x - S.box_lo[0]) * S.box_inv[0] * 8.0f, 0.0f), 7.0f));
    uint32_t cy = static_cast<uint32_t>(fminf(fmaxf((o.y - S.box_lo[1]) * S.box_inv[1] * 8.0f, 0.0f), 7.0f));
    uint32_t cz = static_cast<uint32_t>(fminf(fmaxf((o.z - S.box_lo[2]) * S.box_inv[2] * 8.0f, 0.0f), 7.0f));
    uint32_t oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
    return (oct << 9) | spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2);
}

X3D_DI int kind_of(DevScene const &S, int32_t mat) {
    int32_t k = S.mats[mat].kind;
    return k == kMatCookTorr ? 1 : (k == kMatMixture ? 2 : 0);
}

// Block-aggregated appends to NQ queues at once: one LDS atomic per wave and queue, then one
// global atomic per block and queue (a queue counter takes one atomic per 256 items instead of
// one per 64). Every thread of the block must call it the same number of times; `scratch` is
// 2 * NQ words of LDS, zero before the first call. idx[q] is valid where pred[q] is set.
template <int NQ>
X3D_DI void block_append(uint32_t *const counters[NQ], const bool pred[NQ], uint32_t idx[NQ], uint32_t *scratch) {
    uint32_t lane = __lane_id();
    unsigned long long below = (1ull << lane) - 1ull;
    uint32_t woff[NQ], rank[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        unsigned long long mask = __ballot(pred[q]);
        rank[q] = __popcll(mask & below);
        uint32_t cnt = __popcll(mask);
        uint32_t o = 0;
        if (lane == 0 && cnt)
            o = atomicAdd(&scratch[q], cnt);
        woff[q] = __shfl(o, 0, 64);
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        uint32_t tot = scratch[threadIdx.x];
        scratch[NQ + threadIdx.x] = tot ? atomicAdd(counters[threadIdx.x], tot) : 0u;
        scratch[threadIdx.x] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++)
        idx[q] = scratch[NQ + q] + woff[q] + rank[q];
}

// Appends `path` to the active list of its material kind (all kinds in list 0 without the
// material sort). Block-wide: see block_append.
X3D_DI void push_active(WfArgs const &A, int parity, int kind, uint32_t path, bool live, uint32_t *scratch) {
    if (A.sort_by_material) {
        uint32_t *const c[kKinds] = {&A.ctl->act_count[parity][0].v, &A.ctl->act_count[parity][1].v,
                                     &A.ctl->act_count[parity][2].v};
        const bool p[kKinds] = {live && kind == 0, live && kind == 1, live && kind == 2};
        uint32_t idx[kKinds];
        block_append<kKinds>(c, p, idx, scratch);
        if (live)
            A.act[parity][static_cast<size_t>(kind) * A.capacity + (kind == 0 ? idx[0] : (kind == 1 ? idx[1] : idx[2]))] =
                path;
    } else {
        uint32_t *const c[1] = {&A.ctl->act_count[parity][0].v};
        const bool p[1] = {live};
        uint32_t idx[1];
        block_append<1>(c, p, idx, scratch);
        if (live)
            A.act[parity][idx[0]] = path;
    }
}

// Zeroes the block's append scratch (call once at kernel start, before any block_append).
X3D_DI void clear_scratch(uint32_t *scratch, int words) {
    if (threadIdx.x < static_cast<uint32_t>(words))
        scratch[threadIdx.x] = 0;
    __syncthreads();
}

// Work fetch of the persistent traversal kernels: the wave takes kFetchBatch items from
// the global queue at a time into a private cursor [slot.x, slot.y) in LDS and hands them to its
// lanes as they free up, so the shared counter sees one atomic per kFetchBatch rays instead of
// one per refill. Only the lanes that need an item call it (they are the active lanes).
// The cursor is written by one lane and read by the others later: relaxed atomic loads and
// stores keep the compiler from answering those reads from registers.
constexpr uint32_t kFetchBatch = 128;
X3D_DI uint32_t cursor_load(uint32_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT); }
X3D_DI void cursor_store(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
X3D_DI uint32_t wave_fetch_batched(uint32_t *head, uint32_t *slot) {
    unsigned long long mask = __ballot(1);
    uint32_t lane = __lane_id();
    uint32_t leader = __ffsll(static_cast<long long>(mask)) - 1;
    uint32_t rank = __popcll(mask & ((1ull << lane) - 1ull));
    uint32_t cnt = __popcll(mask);
    uint32_t next = cursor_load(slot), end = cursor_load(slot + 1);
    uint32_t rem = end - next;
    if (rem >= cnt) {
        if (lane == leader)
            cursor_store(slot, next + cnt);
        return next + rank;
    }
    uint32_t base = 0;
    if (lane == leader)
        base = atomicAdd(head, kFetchBatch);
    base = __shfl(base, leader, 64);
    if (lane == leader) {
        cursor_store(slot, base + (cnt - rem));
        cursor_store(slot + 1, base + kFetchBatch);
    }
    return rank < rem ? next + rank : base + (rank - rem);
}

__global__ void k_wf_begin_render(WfCtl *ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        ctl->gen = 0;
}

__global__ void k_wf_next_gen(WfCtl *ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        ctl->gen++;
}

// parity < 0: start of a generation (clear both active lists).
__global__ void k_wf_reset(WfCtl *ctl, int parity) {
    if (blockIdx.x != 0)
        return;
    uint32_t t = threadIdx.x;
    if (t < 4) {
        if (parity < 0) {
            ctl->act_count[0][t].v = 0;
            ctl->act_count[1][t].v = 0;
        } else {
            ctl->act_count[parity ^ 1][t].v = 0;
        }
    }
    if (t == 0) {
        ctl->sq_count.v = 0;
        ctl->sq_head.v = 0;
        ctl->eq_count.v = 0;
        ctl->eq_head.v = 0;
        ctl->fin_count.v = 0;
        if (parity < 0) {
            ctl->eq_count_b.v = 0;
            ctl->fin_count_b.v = 0;
        }
    }
}

// Fused form, between a bounce's extension kernel and its shade: the shadow queue and both
// fetch cursors restart, and the next bounce's extension and finishing queues start empty.
__global__ void k_wf_reset_fused(WfCtl *ctl, WfCounter *eq_out, WfCounter *fin_out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->sq_count.v = 0;
        ctl->sq_head.v = 0;
        ctl->eq_head.v = 0;
        eq_out->v = 0;
        fin_out->v = 0;
    }
}

X3D_DI uint32_t gen_count(WfArgs const &A, uint32_t gen) {
    uint32_t first = A.s_begin + gen * A.gen_samples;
    uint32_t left = A.s_end > first ? A.s_end - first : 0;
    return left < A.gen_samples ? left : A.gen_samples;
}

// Next-event estimation and BSDF sampling of a live path at vertex x (Tracer::direct with the MIS
// weight of Tracer::mis_bounce, then the BSDF sample): draws the light sample, then the BSDF
// sample, from the path's PCG state; stores the advanced state and the sampled direction.
struct NeeSample {
    bool has_shadow, ok;
    f3 sc, sdir, edir;
    float sdist;
};

// TR = kWfMis: the light sample's contribution carries the power-heuristic weight and the BSDF is
// sampled with the evaluated mixture densities (Tracer::mis_bounce); kWfLt1: unweighted
// (Tracer::direct as li_lt1 adds it) and the reference's mixture sampling; kWfUnidirect: BSDF
// sample only (no light sample).
template <int TR>
X3D_DI NeeSample nee_sample(WfArgs const &A, uint32_t path, HitRec const &x, f3 wo, f3 beta, Pcg &g) {
    NeeSample r;
    r.has_shadow = false;
    r.sc = r.sdir = mk(0.0f, 0.0f, 0.0f);
    r.sdist = 0.0f;
    if (TR != kWfUnidirect && A.S.num_lights > 0) {
        LightSample ls = sample_light(A.S, g);
        f3 l, illum;
        float r2;
        if (light_arrives(A.S, ls, x.p, x.n, illum, l, r2)) {
            float dist = norm(l);
            f3 wl = -l / dist;
            f3 f = bsdf_eval(A.S.mats, x.mat, x.n, wo, wl);
            f3 d = mul(illum, f) / ls.area_dens;
            if (TR == kWfMis && !ls.sky) {
                float cl = dot(ls.n, -wl);
                float pl = ls.area_dens * r2 / cl;
                float pb = bsdf_pdf(A.S.mats, x.mat, x.n, wo, wl);
                float w = pl * pl / (pl * pl + pb * pb);
                r.sc = mul(beta, d * w);
            } else {
                // lt1 adds it unweighted; the sky, which a BSDF sample never reaches, has MIS
                // weight 1 (d * 1 == d)
                r.sc = mul(beta, d);
            }
            r.sdist = dist;
            r.sdir = wl;
            r.has_shadow = true;
        }
    }
    float dens;
    f3 wi = bsdf_sample(A.S.mats, x.mat, x.n, wo, g, &dens, TR != kWfMis);
    A.rng[path] = g.state;
    // li_mis stops on !(dens > 0); li_lt1 / li_unidirect on dens == 0
    r.ok = TR == kWfMis ? dens > 0.0f : dens != 0.0f;
    if (r.ok)
        A.wi[path] = make_float4(wi.x, wi.y, wi.z, dens);
    r.edir = wi;
    return r;
}

X3D_DI Pcg path_pcg(uint64_t state, uint32_t pix) {
    Pcg g;
    g.state = state;
    g.inc = (static_cast<uint64_t>(pix) << 1u) | 1u;
    return g;
}

// A path's radiance into the generation's sample buffer; lt1 adds its first hit's emission last,
// as li_lt1 does after its loop.
template <int TR> X3D_DI void finish_path(WfArgs const &A, uint32_t path, f3 L) {
    if (TR == kWfLt1) {
        float4 e = A.le0[path];
        if (e.w != 0.0f)
            L = L + xyz(e);
    }
    A.lbuf[path] = make_float4(L.x, L.y, L.z, 0.0f);
}

// Queues a sampled vertex's shadow ray and extension ray (or finishes the path when the BSDF
// sample failed). Block-wide: every thread of the block calls it at one uniform call site
// (block_append synchronizes); `scratch` holds 6 zeroed words.
X3D_DI void push_rays(WfArgs const &A, bool in, uint32_t path, f3 xp3, NeeSample const &r, uint32_t *scratch) {
    uint32_t *const queues[3] = {&A.ctl->sq_count.v, &A.eq_cnt_out->v, &A.fin_cnt_out->v};
    const bool want[3] = {in && r.has_shadow, in && r.ok, in && !r.ok};
    uint32_t idx[3];
    block_append<3>(queues, want, idx, scratch);
    if (want[0]) {
        uint32_t q = idx[0];
        A.sq[3 * q] = make_float4(xp3.x, xp3.y, xp3.z, r.sdist - 1e-3f);
        A.sq[3 * q + 1] = make_float4(r.sdir.x, r.sdir.y, r.sdir.z, __uint_as_float(path));
        A.sq[3 * q + 2] = make_float4(r.sc.x, r.sc.y, r.sc.z, 0.0f);
        if (A.sort_rays)
            A.ray_key[0][q] = ray_key(A.S, xp3, r.sdir);
    }
    if (want[1]) {
        A.eq_out[idx[1]] = path;
        if (A.sort_rays)
            A.ray_key[1][idx[1]] = ray_key(A.S, xp3, r.edir);
    }
    if (want[2])
        A.fin_out[idx[2]] = path;
}

template <bool FUSED, int TR> __global__ __launch_bounds__(256) void k_wf_generate(WfArgs A) {
    __shared__ uint32_t scratch[2 * kKinds];
    clear_scratch(scratch, 2 * kKinds);
    uint32_t gen = A.ctl->gen;
    uint32_t n = gen_count(A, gen) * A.npix;
    // block-uniform trip count (block_append synchronizes the block)
    for (uint32_t b = blockIdx.x * blockDim.x; b < n; b += gridDim.x * blockDim.x) {
        uint32_t id = b + threadIdx.x;
        bool live = false;
        int kind = 0;
        NeeSample r;
        r.has_shadow = r.ok = false;
        f3 xp3 = mk(0.0f, 0.0f, 0.0f);
        if (id < n) {
            uint32_t s_local = id / A.npix, pix = id - s_local * A.npix;
            uint32_t s = A.s_begin + gen * A.gen_samples + s_local;
            f3 o, v;
            camera_ray(A.cam, pix % A.width, pix / A.width, A.width, A.height, o, v);
            RenderParams RP;
            RP.first = A.first;
            bool valid;
            HitRec x = load_first(RP, pix, valid);
            // Tracer::mis_begin / the first vertex of li_lt1, li_unidirect
            Pcg g = pcg_for(A.seed, pix, s);
            f3 L = mk(0.0f, 0.0f, 0.0f);
            f3 wo = -v;
            if (TR == kWfMis) {
                if (valid && x.light >= 0)
                    L = L + light_radiance(A.S, x.light, wo, x.n);
                live = valid && A.max_depth > 0;
            } else if (TR == kWfLt1) {
                live = valid;
                if (valid) {
                    f3 e = x.light >= 0 ? light_radiance(A.S, x.light, wo, x.n) : mk(0.0f, 0.0f, 0.0f);
                    A.le0[id] = make_float4(e.x, e.y, e.z, x.light >= 0 ? 1.0f : 0.0f);
                }
            } else {
                live = valid;
                if (valid && x.light >= 0)
                    L = L + mul(mk(1.0f, 1.0f, 1.0f), light_radiance(A.S, x.light, wo, x.n));
            }
            if (!live) {
                A.lbuf[id] = make_float4(L.x, L.y, L.z, 0.0f);
            } else {
                A.xp[id] = make_float4(x.p.x, x.p.y, x.p.z, __uint_as_float(static_cast<uint32_t>(x.mat)));
                A.xn[id] = make_float4(x.n.x, x.n.y, x.n.z, __uint_as_float(static_cast<uint32_t>(x.light)));
                A.wo[id] = make_float4(wo.x, wo.y, wo.z, __uint_as_float(pix));
                A.beta[id] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(0u));
                A.L[id] = make_float4(L.x, L.y, L.z, 0.0f);
                if (FUSED) {
                    xp3 = x.p;
                    r = nee_sample<TR>(A, id, x, wo, mk(1.0f, 1.0f, 1.0f), g);
                } else {
                    A.rng[id] = g.state;
                    kind = kind_of(A.S, x.mat);
                }
            }
        }
        if (FUSED)
            push_rays(A, live, id, xp3, r, scratch);
        else
            push_active(A, 0, kind, id, live, scratch);
    }
}

__global__ __launch_bounds__(256) void k_wf_nee_bsdf(WfArgs A, int parity) {
    __shared__ uint32_t scratch[6];
    clear_scratch(scratch, 6);
    uint32_t c0 = A.ctl->act_count[parity][0].v, c1 = A.ctl->act_count[parity][1].v,
             c2 = A.ctl->act_count[parity][2].v;
    uint32_t total = c0 + c1 + c2;
    // block-uniform trip count (block_append synchronizes the block)
    for (uint32_t b = blockIdx.x * blockDim.x; b < total; b += gridDim.x * blockDim.x) {
        uint32_t t = b + threadIdx.x;
        bool in = t < total;
        uint32_t path = 0;
        f3 xp3 = mk(0.0f, 0.0f, 0.0f);
        NeeSample r;
        r.has_shadow = r.ok = false;
        if (in) {
            uint32_t kind = t < c0 ? 0u : (t < c0 + c1 ? 1u : 2u);
            uint32_t j = t - (kind == 0 ? 0u : (kind == 1 ? c0 : c0 + c1));
            path = A.act[parity][static_cast<size_t>(kind) * A.capacity + j];
            float4 xp = A.xp[path], xn = A.xn[path], wo4 = A.wo[path], b4 = A.beta[path];
            HitRec x;
            x.p = xyz(xp);
            x.n = xyz(xn);
            x.mat = static_cast<int32_t>(__float_as_uint(xp.w));
            x.light = static_cast<int32_t>(__float_as_uint(xn.w));
            xp3 = x.p;
            Pcg g = path_pcg(A.rng[path], __float_as_uint(wo4.w));
            r = nee_sample<kWfMis>(A, path, x, xyz(wo4), xyz(b4), g);
        }
        // one call site with all lanes active: block_append synchronizes the block
        push_rays(A, in, path, xp3, r, scratch);
    }
}

// Counting sort of both ray queues by key (blockIdx.y = queue), in four passes: per-block key
// counts over contiguous chunks (with each entry's rank among equal keys of its block), column
// totals, per-(block, key) output offsets, scatter. No global atomics.
X3D_DI uint32_t queue_count(WfArgs const &A, int q) { return q == 0 ? A.ctl->sq_count.v : A.eq_cnt_in->v; }

__global__ __launch_bounds__(kSortWG) void k_wf_sort_count(WfArgs A) {
    __shared__ uint32_t hist[kSortBins];
    const int q = blockIdx.y;
    for (uint32_t b = threadIdx.x; b < kSortBins; b += kSortWG)
        hist[b] = 0u;
    __syncthreads();
    const uint32_t n = queue_count(A, q);
    const uint32_t chunk = (n + kSortBlocks - 1) / kSortBlocks;
    const uint32_t begin = blockIdx.x * chunk, end = min(begin + chunk, n);
    const uint32_t *key = A.ray_key[q];
    uint32_t *rank = A.ray_rank[q];
    for (uint32_t i = begin + threadIdx.x; i < end; i += kSortWG)
        rank[i] = atomicAdd(&hist[key[i]], 1u);
    __syncthreads();
    uint32_t *row = A.bin_mat[q] + static_cast<size_t>(blockIdx.x) * kSortBins;
    for (uint32_t b = threadIdx.x; b < kSortBins; b += kSortWG)
        row[b] = hist[b];
}

// One thread per key: total over the count rows.
__global__ __launch_bounds__(256) void k_wf_sort_totals(WfArgs A) {
    const int q = blockIdx.y;
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    const uint32_t *m = A.bin_mat[q];
    uint32_t t = 0;
    for (uint32_t r = 0; r < kSortBlocks; r++)
        t += m[static_cast<size_t>(r) * kSortBins + b];
    A.bin_tot[q][b] = t;
}

// One thread per key: exclusive scan of the totals (every block scans all of them in LDS), then
// the running offsets of that key over the count rows replace the counts.
__global__ __launch_bounds__(256) void k_wf_sort_offsets(WfArgs A) {
    constexpr uint32_t kPer = kSortBins / 256;
    __shared__ uint32_t part[256];
    const int q = blockIdx.y;
    const uint32_t *tot = A.bin_tot[q];
    uint32_t own[kPer], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
        own[k] = tot[threadIdx.x * kPer + k];
        sum += own[k];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan of the 256 partial sums
    for (uint32_t d = 1; d < 256; d <<= 1) {
        uint32_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    // exclusive prefix of key b = blockIdx.x * 256 + threadIdx.x
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    const uint32_t owner = b / kPer;
    uint32_t base = owner ? part[owner - 1] : 0u;
    for (uint32_t k = owner * kPer; k < b; k++)
        base += tot[k];
    // rows in groups of 16: the loads of a group issue together before its stores
    uint32_t *m = A.bin_mat[q];
    for (uint32_t r0 = 0; r0 < kSortBlocks; r0 += 16) {
        uint32_t c[16];
#pragma unroll
        for (uint32_t k = 0; k < 16; k++)
            c[k] = m[static_cast<size_t>(r0 + k) * kSortBins + b];
#pragma unroll
        for (uint32_t k = 0; k < 16; k++) {
            m[static_cast<size_t>(r0 + k) * kSortBins + b] = base;
            base += c[k];
        }
    }
}

__global__ __launch_bounds__(kSortWG) void k_wf_sort_scatter(WfArgs A) {
    __shared__ uint32_t off[kSortBins];
    const int q = blockIdx.y;
    const uint32_t *row = A.bin_mat[q] + static_cast<size_t>(blockIdx.x) * kSortBins;
    for (uint32_t b = threadIdx.x; b < kSortBins; b += kSortWG)
        off[b] = row[b];
    __syncthreads();
    const uint32_t n = queue_count(A, q);
    const uint32_t chunk = (n + kSortBlocks - 1) / kSortBlocks;
    const uint32_t begin = blockIdx.x * chunk, end = min(begin + chunk, n);
    const uint32_t *key = A.ray_key[q], *rank = A.ray_rank[q];
    for (uint32_t i = begin + threadIdx.x; i < end; i += kSortWG) {
        uint32_t dst = off[key[i]] + rank[i];
        if (q == 0) {
            A.sq_sorted[3 * dst] = A.sq[3 * i];
            A.sq_sorted[3 * dst + 1] = A.sq[3 * i + 1];
            A.sq_sorted[3 * dst + 2] = A.sq[3 * i + 2];
        } else {
            A.eq_sorted[dst] = A.eq[i];
        }
    }
}

template <int WG, int LDSD>
__global__ __launch_bounds__(WG) void k_wf_shadow(WfArgs A) {
    __shared__ uint2 lds[LDSD * WG];
    __shared__ uint32_t cursor[2 * (WG / 64)];
    __shared__ float4 top[kTopNodes > 0 ? kTopNodes * kWideRows : 1];
    TravStack<WG, LDSD> st = make_stack<WG, LDSD>(lds, A.spill, A.spill_per_thread);
    DevScene const &S = A.S;
    if (kTopNodes > 0) {
        const uint32_t nt = min(kTopNodes, S.num_nodes4) * kWideRows;
        for (uint32_t q = threadIdx.x; q < nt; q += WG)
            top[q] = static_cast<const float4 *>(S.nodes4)[q];
        __syncthreads();
    }
    const float4 *tris = static_cast<const float4 *>(S.tris);
    const uint32_t n = A.ctl->sq_count.v;
    uint32_t pops = 0, tests = 0;
    uint32_t *slot = cursor + 2 * (threadIdx.x >> 6);
    if (__lane_id() == 0) {
        cursor_store(slot, 0u);
        cursor_store(slot + 1, 0u);
    }
    uint32_t i = wave_fetch_batched(&A.ctl->sq_head.v, slot);
    f3 o = mk(0, 0, 0), v = mk(0, 0, 0), vinv = mk(0, 0, 0);
    float tmax = 0.0f;
    uint32_t ref = kDone, leaf = kDone, cur = 0; // cur: next triangle of the parked leaf
    bool blocked = false, wide = false;
    auto init = [&](uint32_t k) {
        float4 a = A.sq_in[3 * k], b = A.sq_in[3 * k + 1];
        o = xyz(a);
        tmax = a.w;
        v = xyz(b);
        vinv = mk(1.0f / v.x, 1.0f / v.y, 1.0f / v.z);
        wide = !slab_may_nan(o, vinv, S.far_coords != 0);
        blocked = false;
        st.sp = 0;
        ref = wide ? S.root4 : S.root_ref;
        leaf = kDone;
        if (ref != kEmptyRef && !is_interior(ref)) {
            leaf = ref;
            cur = leaf_start(leaf);
            ref = kDone;
        }
    };
    if (i < n)
        init(i);
    while (i < n) {
        while (is_interior(ref)) {
            if (wide) {
                pops += kTreeWidth / 2; // a wide node counts as W/2 binary nodes (same bytes and tests)
                ref = visit_wide<false>(S, ref, o, vinv, 1e-4f, tmax, __builtin_inff(), st, top);
            } else {
                pops++;
                ref = visit_interior<false>(S, ref, o, vinv, 1e-4f, tmax, __builtin_inff(), st, true);
            }
            if (is_leaf_ref(ref) && leaf == kDone) {
                leaf = ref;
                cur = leaf_start(leaf);
                ref = st.pop_next(__builtin_inff());
            }
            // descend until fewer than refill_below lanes still look for their first leaf
            if (__popcll(__ballot(leaf == kDone)) < A.refill_below)
                break;
        }
        // One triangle per lane and iteration: a lane whose leaf is done takes its next stacked
        // leaf at once instead of waiting for lanes with larger leaves.
        while (leaf != kDone) {
            {
                const uint32_t p = cur;
                float4 a = tris[3 * p], b = tris[3 * p + 1], c = tris[3 * p + 2];
                float c1, c2, t0;
                tests++;
                blocked = tri_hit(o, v, xyz(a), xyz(b), xyz(c), 1e-4f, tmax, &c1, &c2, &t0);
            }
            cur++;
            if (blocked) {
                leaf = kDone;
                ref = kDone;
            } else if (cur == leaf_start(leaf) + leaf_count(leaf)) {
                if (is_leaf_ref(ref)) {
                    leaf = ref;
                    cur = leaf_start(leaf);
                    ref = st.pop_next(__builtin_inff());
                } else {
                    leaf = kDone;
                }
            }
            // test leaves until fewer than leaf_below lanes still hold one (the rest keep theirs)
            if (__popcll(__ballot(leaf != kDone)) < A.leaf_below)
                break;
        }
        if (ref == kDone && leaf == kDone) {
            if (!blocked) {
                uint32_t path = __float_as_uint(A.sq_in[3 * i + 1].w);
                float4 c = A.sq_in[3 * i + 2];
                float4 L = A.L[path];
                f3 nl = xyz(L) + xyz(c);
                A.L[path] = make_float4(nl.x, nl.y, nl.z, L.w);
            }
            i = wave_fetch_batched(&A.ctl->sq_head.v, slot);
            if (i < n)
                init(i);
        }
    }
    if (A.cnt) {
        flush_counter(&A.cnt->spops, pops);
        flush_counter(&A.cnt->stests, tests);
    }
}

template <int WG, int LDSD>
__global__ __launch_bounds__(WG) void k_wf_extend(WfArgs A) {
    __shared__ uint2 lds[LDSD * WG];
    __shared__ uint32_t cursor[2 * (WG / 64)];
    __shared__ float4 top[kTopNodes > 0 ? kTopNodes * kWideRows : 1];
    TravStack<WG, LDSD> st = make_stack<WG, LDSD>(lds, A.spill, A.spill_per_thread);
    DevScene const &S = A.S;
    if (kTopNodes > 0) {
        const uint32_t nt = min(kTopNodes, S.num_nodes4) * kWideRows;
        for (uint32_t q = threadIdx.x; q < nt; q += WG)
            top[q] = static_cast<const float4 *>(S.nodes4)[q];
        __syncthreads();
    }
    const float4 *tris = static_cast<const float4 *>(S.tris);
    const uint32_t n = A.eq_cnt_in->v;
    uint32_t pops = 0, tests = 0;
    uint32_t *slot = cursor + 2 * (threadIdx.x >> 6);
    if (__lane_id() == 0) {
        cursor_store(slot, 0u);
        cursor_store(slot + 1, 0u);
    }
    uint32_t i = wave_fetch_batched(&A.ctl->eq_head.v, slot);
    f3 o = mk(0, 0, 0), v = mk(0, 0, 0), vinv = mk(0, 0, 0);
    ClosestHit h{-1, __builtin_inff(), 0.0f, 0.0f};
    uint32_t best_leaf = 0, ref = kDone, leaf = kDone, cur = 0; // cur: next triangle of the parked leaf
    bool wide = false;
    auto init = [&](uint32_t k) {
        uint32_t path = A.eq_in[k];
        o = xyz(A.xp[path]);
        v = xyz(A.wi[path]);
        vinv = mk(1.0f / v.x, 1.0f / v.y, 1.0f / v.z);
        wide = !slab_may_nan(o, vinv, S.far_coords != 0);
        h = ClosestHit{-1, __builtin_inff(), 0.0f, 0.0f};
        best_leaf = 0;
        st.sp = 0;
        ref = wide ? S.root4 : S.root_ref;
        leaf = kDone;
        if (ref != kEmptyRef && !is_interior(ref)) {
            leaf = ref;
            cur = leaf_start(leaf);
            ref = kDone;
        }
    };
    if (i < n)
        init(i);
    while (i < n) {
        while (is_interior(ref)) {
            float cut = h.t + h.t * kCullSlack;
            if (wide) {
                pops += kTreeWidth / 2;
                ref = visit_wide(S, ref, o, vinv, kTMin, kTMax, cut, st, top);
            } else {
                pops++;
                ref = visit_interior<false>(S, ref, o, vinv, kTMin, kTMax, cut, st, true);
            }
            if (is_leaf_ref(ref) && leaf == kDone) {
                leaf = ref;
                cur = leaf_start(leaf);
                ref = st.pop_next(cut);
            }
            // descend until fewer than refill_below lanes still look for their first leaf
            if (__popcll(__ballot(leaf == kDone)) < A.refill_below)
                break;
        }
        // One triangle per lane and iteration: a lane whose leaf is done takes its next stacked
        // leaf at once instead of waiting for lanes with larger leaves.
        while (leaf != kDone) {
            const uint32_t s0 = leaf_start(leaf), p = cur;
            {
                float4 a = tris[3 * p], b = tris[3 * p + 1], c = tris[3 * p + 2];
                float c1, c2, t0;
                tests++;
                if (tri_hit(o, v, xyz(a), xyz(b), xyz(c), kTMin, kTMax, &c1, &c2, &t0) &&
                    (t0 < h.t || (t0 == h.t && s0 > best_leaf))) {
                    h.t = t0;
                    h.b1 = c1;
                    h.b2 = c2;
                    h.prim = static_cast<int32_t>(p);
                    best_leaf = s0;
                }
            }
            cur++;
            if (cur == s0 + leaf_count(leaf)) {
                if (is_leaf_ref(ref)) {
                    leaf = ref;
                    cur = leaf_start(leaf);
                    ref = st.pop_next(h.t + h.t * kCullSlack);
                } else {
                    leaf = kDone;
                }
            }
            // test leaves until fewer than leaf_below lanes still hold one (the rest keep theirs)
            if (__popcll(__ballot(leaf != kDone)) < A.leaf_below)
                break;
        }
        if (ref == kDone && leaf == kDone) {
            A.eh[i] = h;
            i = wave_fetch_batched(&A.ctl->eq_head.v, slot);
            if (i < n)
                init(i);
        }
    }
    if (A.cnt) {
        flush_counter(&A.cnt->pops, pops);
        flush_counter(&A.cnt->tests, tests);
    }
}

template <bool FUSED, int TR>
__global__ __launch_bounds__(256) void k_wf_shade(WfArgs A, int parity, uint32_t bounce) {
    __shared__ uint32_t scratch[2 * kKinds];
    clear_scratch(scratch, 2 * kKinds);
    uint32_t ne = A.eq_cnt_in->v, nf = A.fin_cnt_in->v;
    uint32_t total = ne + nf;
    // block-uniform trip count (block_append synchronizes the block)
    for (uint32_t b = blockIdx.x * blockDim.x; b < total; b += gridDim.x * blockDim.x) {
        uint32_t t = b + threadIdx.x;
        bool live = false;
        int kind = 0;
        uint32_t path = 0;
        NeeSample r;
        r.has_shadow = r.ok = false;
        f3 xp3 = mk(0.0f, 0.0f, 0.0f);
        if (t >= total) {
            // no item: only takes part in the block-wide append below
        } else if (t >= ne) {
            path = A.fin_in[t - ne];
            finish_path<TR>(A, path, xyz(A.L[path]));
        } else {
            path = A.eq_in[t];
            ClosestHit h = A.eh[t];
            // Every load of the item issues at once: the path's state and the hit triangle (a
            // miss reads triangle 0, which every scene with a possible hit has), so one memory
            // round trip instead of three.
            const uint32_t pi = h.prim >= 0 ? static_cast<uint32_t>(h.prim) : 0u;
            const float4 *tris = static_cast<const float4 *>(A.S.tris);
            const float4 *attr = static_cast<const float4 *>(A.S.tri_attr);
            float4 ta = tris[3 * pi], tb = tris[3 * pi + 1], tc = tris[3 * pi + 2];
            float4 na = attr[3 * pi], nb = attr[3 * pi + 1], nc = attr[3 * pi + 2];
            float4 L4 = A.L[path], wi4 = A.wi[path], xn = A.xn[path], wo4 = A.wo[path], b4 = A.beta[path];
            float4 xp4 = A.xp[path];
            f3 L = xyz(L4);
            f3 wi = xyz(wi4);
            HitRec nx;
            bool hit = h.prim >= 0;
            if (hit) {
                nx = finish_hit(h, ta, tb, tc, na, nb, nc);
                hit = !(dot(nx.n, -wi) <= 0.0f);
            }
            if (!hit) {
                finish_path<TR>(A, path, L);
            } else {
                float dens = wi4.w;
                int32_t xmat = static_cast<int32_t>(__float_as_uint(xp4.w));
                f3 xnrm = xyz(xn), wo = xyz(wo4), beta = xyz(b4);
                float cw = dot(xnrm, wi);
                f3 f = bsdf_eval(A.S.mats, xmat, xnrm, wo, wi);
                beta = mul(beta, f) * cw / dens;
                bool go;
                Pcg g = path_pcg(0, __float_as_uint(wo4.w));
                if (TR == kWfMis) {
                    if (nx.light >= 0) {
                        float cl = dot(nx.n, -wi);
                        DevLight const &Ln = A.S.lights[nx.light];
                        float pl = (Ln.prob / Ln.area) * (nx.t * nx.t) / cl;
                        float w = dens * dens / (dens * dens + pl * pl);
                        L = L + mul(beta, light_radiance(A.S, nx.light, -wi, nx.n) * w);
                    }
                    go = bounce + 1 < A.max_depth;
                    if (go && FUSED)
                        g.state = A.rng[path];
                } else {
                    // li_lt1 / li_unidirect at depth bounce + 1: Russian roulette p = 0.5 from depth 2,
                    // then (unidirect) the vertex's emission; the loop stops after kRefModeMaxDepth vertices
                    const uint32_t depth = bounce + 1;
                    go = depth < kRefModeMaxDepth;
                    if (go) {
                        g.state = A.rng[path];
                        if (depth >= 2) {
                            if (g.draw() >= 0.5f)
                                go = false;
                            else
                                beta = beta / 0.5f;
                        }
                    }
                    if (go && TR == kWfUnidirect && nx.light >= 0)
                        L = L + mul(beta, light_radiance(A.S, nx.light, -wi, nx.n));
                }
                if (!go) {
                    finish_path<TR>(A, path, L);
                } else {
                    A.xp[path] =
                        make_float4(nx.p.x, nx.p.y, nx.p.z, __uint_as_float(static_cast<uint32_t>(nx.mat)));
                    A.xn[path] =
                        make_float4(nx.n.x, nx.n.y, nx.n.z, __uint_as_float(static_cast<uint32_t>(nx.light)));
                    A.wo[path] = make_float4(-wi.x, -wi.y, -wi.z, wo4.w);
                    A.beta[path] = make_float4(beta.x, beta.y, beta.z, b4.w);
                    A.L[path] = make_float4(L.x, L.y, L.z, 0.0f);
                    live = true;
                    if (FUSED) {
                        xp3 = nx.p;
                        r = nee_sample<TR>(A, path, nx, -wi, beta, g);
                    } else {
                        kind = kind_of(A.S, nx.mat);
                    }
                }
            }
        }
        if (FUSED)
            push_rays(A, live, path, xp3, r, scratch);
        else
            push_active(A, parity ^ 1, kind, path, live, scratch);
    }
}

__global__ __launch_bounds__(256) void k_wf_accumulate(WfArgs A) {
    uint32_t gen = A.ctl->gen;
    uint32_t cnt = gen_count(A, gen);
    bool last = A.s_begin + gen * A.gen_samples + cnt >= A.s_end;
    float scale = 1.0f / static_cast<float>(A.spp_total);
    for (uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x; pix < A.npix; pix += gridDim.x * blockDim.x) {
        f3 acc = mk(0.0f, 0.0f, 0.0f);
        if (gen > 0 || A.s_begin > 0) {
            float4 q = A.acc[pix];
            acc = mk(q.x, q.y, q.z);
        }
        if (A.firefly) {
            const uint32_t x = pix % A.width, y = pix / A.width;
            for (uint32_t s = 0; s < cnt; s++)
                acc = acc + xyz(firefly_filtered(A.lbuf + static_cast<size_t>(s) * A.npix, x, y, A.width, A.height));
        } else {
            for (uint32_t s = 0; s < cnt; s++)
                acc = acc + xyz(A.lbuf[static_cast<size_t>(s) * A.npix + pix]);
        }
        if (last && A.finalize)
            A.out[pix] = make_float4(scale * acc.x, scale * acc.y, scale * acc.z, scale);
        else
            A.acc[pix] = make_float4(acc.x, acc.y, acc.z, 0.0f);
    }
}

// Ray-query counts (one closest-hit per extension ray, one any-hit per shadow ray).
__global__ void k_wf_count(WfCtl *ctl, WfCounter const *eq_cnt, Counters *cnt) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && cnt) {
        atomicAdd(&cnt->closest, static_cast<unsigned long long>(eq_cnt->v));
        atomicAdd(&cnt->anyhit, static_cast<unsigned long long>(ctl->sq_count.v));
    }
}

// ---------------------------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------------------------
struct WavefrontEngine {
    DevBuf xp, xn, wo, beta, L, wi, rng, lbuf, act0, act1, sq, eq, fin, eh, ctl, spill;
    DevBuf eq_b, fin_b; // the fused form's second queue parity
    DevBuf le0;         // lt1's first-hit emission per path
    DevBuf sq_sorted, eq_sorted, ray_key, ray_rank, bin_mat, bin_tot; // ray reordering
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    WfArgs key{};
    WavefrontTuning key_tune{};
    bool have_graph = false;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    double shadow_ms = 0, extend_ms = 0;
    uint32_t trace_launches = 0;
    ~WavefrontEngine() {
        for (auto &x : ev)
            if (x)
                (void)hipEventDestroy(x);
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
    }
};

WavefrontEngine *wf_create() { return new WavefrontEngine(); }

void wf_last_timing(WavefrontEngine const *e, double *shadow_ms, double *extend_ms, uint32_t *launches) {
    *shadow_ms = e->shadow_ms;
    *extend_ms = e->extend_ms;
    *launches = e->trace_launches;
}
void wf_destroy(WavefrontEngine *e) { delete e; }

namespace {

template <class K> uint32_t grid_for(K kernel, int wg) {
    int per_cu = 0, dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, wg, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return static_cast<uint32_t>(cus * per_cu);
}

using TraceKernel = void (*)(WfArgs);

struct TraceKernels {
    TraceKernel shadow, extend;
    int wg, ldsd;
};

template <int WG, int D> TraceKernels trace_kernels() { return {k_wf_shadow<WG, D>, k_wf_extend<WG, D>, WG, D}; }

// Workgroup sizes 64 / 128 / 256 and LDS stack depths 8 / 16 (deeper entries spill to memory).
TraceKernels pick_trace(int wg, int ldsd) {
    bool shallow = ldsd == 8;
    switch (wg) {
    case 64: return shallow ? trace_kernels<64, 8>() : trace_kernels<64, 16>();
    case 128: return shallow ? trace_kernels<128, 8>() : trace_kernels<128, 16>();
    default: return shallow ? trace_kernels<256, 8>() : trace_kernels<256, 16>();
    }
}

bool same_config(WfArgs const &a, WfArgs const &b, WavefrontTuning const &ta, WavefrontTuning const &tb) {
    return std::memcmp(&a, &b, sizeof(WfArgs)) == 0 && ta.paths_in_flight == tb.paths_in_flight &&
           ta.trace_wg == tb.trace_wg && ta.stack_depth == tb.stack_depth &&
           ta.sort_by_material == tb.sort_by_material && ta.sort_rays == tb.sort_rays &&
           ta.refill_below == tb.refill_below && ta.leaf_below == tb.leaf_below && ta.tracer == tb.tracer;
}

} // namespace

std::string wf_render(WavefrontEngine *e, RenderParams const &P, WavefrontTuning const &tune, hipStream_t stream) {
#define WF_HIP(call)                                                                                 \
    do {                                                                                             \
        hipError_t err_ = (call);                                                                    \
        if (err_ != hipSuccess)                                                                      \
            return std::string(#call) + ": " + hipGetErrorString(err_);                              \
    } while (0)
    uint32_t npix = P.width * P.height;
    if (P.s_end <= P.s_begin)
        return "";
    uint32_t gen_samples = tune.paths_in_flight / npix;
    if (gen_samples < 1)
        gen_samples = 1;
    if (gen_samples > P.s_end - P.s_begin)
        gen_samples = P.s_end - P.s_begin;
    uint32_t cap = gen_samples * npix;
    size_t f4 = sizeof(float4) * static_cast<size_t>(cap);
    WF_HIP(e->xp.ensure(f4));
    WF_HIP(e->xn.ensure(f4));
    WF_HIP(e->wo.ensure(f4));
    WF_HIP(e->beta.ensure(f4));
    WF_HIP(e->L.ensure(f4));
    WF_HIP(e->wi.ensure(f4));
    WF_HIP(e->lbuf.ensure(f4));
    WF_HIP(e->rng.ensure(sizeof(uint64_t) * cap));
    const bool fused = !tune.sort_by_material && !tune.sort_rays;
    if (tune.tracer != kWfMis && !fused)
        return "the lt1 and unidirect tracers run the wavefront form without material or ray sorting only";
    if (tune.tracer == kWfLt1)
        WF_HIP(e->le0.ensure(f4));
    if (fused) {
        WF_HIP(e->eq_b.ensure(sizeof(uint32_t) * cap));
        WF_HIP(e->fin_b.ensure(sizeof(uint32_t) * cap));
    } else {
        WF_HIP(e->act0.ensure(sizeof(uint32_t) * kKinds * static_cast<size_t>(cap)));
        WF_HIP(e->act1.ensure(sizeof(uint32_t) * kKinds * static_cast<size_t>(cap)));
    }
    WF_HIP(e->sq.ensure(3 * f4));
    WF_HIP(e->eq.ensure(sizeof(uint32_t) * cap));
    WF_HIP(e->fin.ensure(sizeof(uint32_t) * cap));
    WF_HIP(e->eh.ensure(sizeof(ClosestHit) * cap));
    WF_HIP(e->ctl.ensure(sizeof(WfCtl)));
    if (tune.sort_rays) {
        WF_HIP(e->sq_sorted.ensure(3 * f4));
        WF_HIP(e->eq_sorted.ensure(sizeof(uint32_t) * cap));
        WF_HIP(e->ray_key.ensure(2 * sizeof(uint32_t) * static_cast<size_t>(cap)));
        WF_HIP(e->ray_rank.ensure(2 * sizeof(uint32_t) * static_cast<size_t>(cap)));
        WF_HIP(e->bin_mat.ensure(2 * sizeof(uint32_t) * static_cast<size_t>(kSortBlocks) * kSortBins));
        WF_HIP(e->bin_tot.ensure(2 * sizeof(uint32_t) * kSortBins));
    }

    TraceKernels tk = pick_trace(tune.trace_wg, tune.stack_depth);
    int wg = tk.wg;
    uint32_t g_shadow = grid_for(tk.shadow, wg), g_extend = grid_for(tk.extend, wg);
    uint32_t trace_threads = (g_shadow > g_extend ? g_shadow : g_extend) * wg;
    uint32_t spill_per_thread = spill_entries_for(P.S, tk.ldsd);
    WF_HIP(e->spill.ensure(sizeof(uint2) * static_cast<size_t>(trace_threads) * spill_per_thread));

    WfArgs A{};
    A.S = P.S;
    A.cam = P.cam;
    A.width = P.width;
    A.height = P.height;
    A.npix = npix;
    A.s_begin = P.s_begin;
    A.s_end = P.s_end;
    A.spp_total = P.spp_total;
    A.gen_samples = gen_samples;
    A.capacity = cap;
    A.seed = P.seed;
    A.max_depth = P.max_depth;
    A.finalize = P.finalize;
    A.first = P.first;
    A.acc = P.acc;
    A.out = P.out;
    A.xp = static_cast<float4 *>(e->xp.p);
    A.xn = static_cast<float4 *>(e->xn.p);
    A.wo = static_cast<float4 *>(e->wo.p);
    A.beta = static_cast<float4 *>(e->beta.p);
    A.L = static_cast<float4 *>(e->L.p);
    A.wi = static_cast<float4 *>(e->wi.p);
    A.rng = static_cast<uint64_t *>(e->rng.p);
    A.lbuf = static_cast<float4 *>(e->lbuf.p);
    A.act[0] = static_cast<uint32_t *>(e->act0.p);
    A.act[1] = static_cast<uint32_t *>(e->act1.p);
    A.sq = static_cast<float4 *>(e->sq.p);
    A.eq = static_cast<uint32_t *>(e->eq.p);
    A.fin = static_cast<uint32_t *>(e->fin.p);
    A.eh = static_cast<ClosestHit *>(e->eh.p);
    A.ctl = static_cast<WfCtl *>(e->ctl.p);
    A.spill = static_cast<uint2 *>(e->spill.p);
    A.spill_per_thread = spill_per_thread;
    A.cnt = P.cnt;
    A.sort_by_material = tune.sort_by_material;
    A.sort_rays = tune.sort_rays;
    A.refill_below = static_cast<uint32_t>(tune.refill_below);
    A.leaf_below = static_cast<uint32_t>(tune.leaf_below);
    A.firefly = P.firefly;
    A.sq_in = A.sq;
    A.eq_in = A.eq;
    A.fin_in = A.fin;
    A.eq_out = A.eq;
    A.fin_out = A.fin;
    A.eq_cnt_in = A.eq_cnt_out = &A.ctl->eq_count;
    A.fin_cnt_in = A.fin_cnt_out = &A.ctl->fin_count;
    A.fused = fused ? 1 : 0;
    A.le0 = tune.tracer == kWfLt1 ? static_cast<float4 *>(e->le0.p) : nullptr;
    if (tune.sort_rays) {
        A.sq_sorted = static_cast<float4 *>(e->sq_sorted.p);
        A.eq_sorted = static_cast<uint32_t *>(e->eq_sorted.p);
        A.sq_in = A.sq_sorted;
        A.eq_in = A.eq_sorted;
        for (int q = 0; q < 2; q++) {
            A.ray_key[q] = static_cast<uint32_t *>(e->ray_key.p) + static_cast<size_t>(q) * cap;
            A.ray_rank[q] = static_cast<uint32_t *>(e->ray_rank.p) + static_cast<size_t>(q) * cap;
            A.bin_mat[q] = static_cast<uint32_t *>(e->bin_mat.p) + static_cast<size_t>(q) * kSortBlocks * kSortBins;
            A.bin_tot[q] = static_cast<uint32_t *>(e->bin_tot.p) + static_cast<size_t>(q) * kSortBins;
        }
    }

    // Grid-strided per-path kernels get exactly the blocks that are resident at once: a larger
    // grid would run its surplus waves as a second, late round.
    using GenKernel = void (*)(WfArgs);
    using ShadeKernel = void (*)(WfArgs, int, uint32_t);
    GenKernel gen_k = k_wf_generate<false, kWfMis>;
    ShadeKernel shade_k = k_wf_shade<false, kWfMis>;
    if (fused) {
        gen_k = tune.tracer == kWfLt1         ? k_wf_generate<true, kWfLt1>
                : tune.tracer == kWfUnidirect ? k_wf_generate<true, kWfUnidirect>
                                              : k_wf_generate<true, kWfMis>;
        shade_k = tune.tracer == kWfLt1         ? k_wf_shade<true, kWfLt1>
                  : tune.tracer == kWfUnidirect ? k_wf_shade<true, kWfUnidirect>
                                                : k_wf_shade<true, kWfMis>;
    }
    // MIS: max_depth bounces; lt1 / unidirect: the reference-mode loop's kRefModeMaxDepth vertices
    // (Russian roulette ends nearly every path long before; bounces with empty queues cost a few
    // microseconds of launch each).
    const uint32_t bounces = tune.tracer == kWfMis ? P.max_depth : kRefModeMaxDepth;
    uint32_t g_gen = grid_for(gen_k, 256), g_nee = grid_for(k_wf_nee_bsdf, 256), g_shade = grid_for(shade_k, 256),
             g_acc = grid_for(k_wf_accumulate, 256);
    // The fused form's queue parities: bounce k consumes parity k & 1 and its shade fills the other.
    auto parity_args = [&](uint32_t k) {
        WfArgs B = A;
        uint32_t *eqs[2] = {static_cast<uint32_t *>(e->eq.p), static_cast<uint32_t *>(e->eq_b.p)};
        uint32_t *fins[2] = {static_cast<uint32_t *>(e->fin.p), static_cast<uint32_t *>(e->fin_b.p)};
        WfCounter *eqc[2] = {&A.ctl->eq_count, &A.ctl->eq_count_b};
        WfCounter *finc[2] = {&A.ctl->fin_count, &A.ctl->fin_count_b};
        const uint32_t p = k & 1u, q = p ^ 1u;
        B.eq = B.eq_out = eqs[q];
        B.eq_in = eqs[p];
        B.fin = B.fin_out = fins[q];
        B.fin_in = fins[p];
        B.eq_cnt_in = eqc[p];
        B.eq_cnt_out = eqc[q];
        B.fin_cnt_in = finc[p];
        B.fin_cnt_out = finc[q];
        return B;
    };
    uint32_t generations = (P.s_end - P.s_begin + gen_samples - 1) / gen_samples;

    bool profile = tune.profile != 0;
    auto trace_bounce = [&](hipStream_t s, WfArgs const &B) {
        if (profile)
            (void)hipEventRecord(e->ev[0], s);
        hipLaunchKernelGGL(tk.shadow, dim3(g_shadow), dim3(wg), 0, s, B);
        if (profile)
            (void)hipEventRecord(e->ev[1], s);
        hipLaunchKernelGGL(tk.extend, dim3(g_extend), dim3(wg), 0, s, B);
        if (profile) {
            (void)hipEventRecord(e->ev[2], s);
            float a = 0, b = 0;
            (void)hipEventSynchronize(e->ev[2]);
            (void)hipEventElapsedTime(&a, e->ev[0], e->ev[1]);
            (void)hipEventElapsedTime(&b, e->ev[1], e->ev[2]);
            e->shadow_ms += a;
            e->extend_ms += b;
            e->trace_launches++;
        }
        hipLaunchKernelGGL(k_wf_count, dim3(1), dim3(64), 0, s, B.ctl, static_cast<const WfCounter *>(B.eq_cnt_in),
                           B.cnt);
    };
    auto record_generation = [&](hipStream_t s) {
        hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, s, A.ctl, -1);
        if (fused) {
            // generate fills bounce 0's queues (parity 0); bounce k's shade fills bounce k + 1's
            WfArgs G = parity_args(1);
            hipLaunchKernelGGL(gen_k, dim3(g_gen), dim3(256), 0, s, G);
            for (uint32_t k = 0; k < bounces; k++) {
                WfArgs B = parity_args(k);
                trace_bounce(s, B);
                hipLaunchKernelGGL(k_wf_reset_fused, dim3(1), dim3(64), 0, s, B.ctl, B.eq_cnt_out, B.fin_cnt_out);
                hipLaunchKernelGGL(shade_k, dim3(g_shade), dim3(256), 0, s, B, 0, k);
            }
        } else {
            hipLaunchKernelGGL(gen_k, dim3(g_gen), dim3(256), 0, s, A);
            for (uint32_t k = 0; k < P.max_depth; k++) {
                int par = static_cast<int>(k & 1u);
                hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, s, A.ctl, par);
                hipLaunchKernelGGL(k_wf_nee_bsdf, dim3(g_nee), dim3(256), 0, s, A, par);
                if (A.sort_rays) {
                    hipLaunchKernelGGL(k_wf_sort_count, dim3(kSortBlocks, 2), dim3(kSortWG), 0, s, A);
                    hipLaunchKernelGGL(k_wf_sort_totals, dim3(kSortBins / 256, 2), dim3(256), 0, s, A);
                    hipLaunchKernelGGL(k_wf_sort_offsets, dim3(kSortBins / 256, 2), dim3(256), 0, s, A);
                    hipLaunchKernelGGL(k_wf_sort_scatter, dim3(kSortBlocks, 2), dim3(kSortWG), 0, s, A);
                }
                trace_bounce(s, A);
                hipLaunchKernelGGL(shade_k, dim3(g_shade), dim3(256), 0, s, A, par, k);
            }
        }
        hipLaunchKernelGGL(k_wf_accumulate, dim3(g_acc), dim3(256), 0, s, A);
        hipLaunchKernelGGL(k_wf_next_gen, dim3(1), dim3(64), 0, s, A.ctl);
    };

    e->shadow_ms = e->extend_ms = 0.0;
    e->trace_launches = 0;
    if (profile && !e->ev[0])
        for (auto &ev : e->ev)
            WF_HIP(hipEventCreate(&ev));
    hipLaunchKernelGGL(k_wf_begin_render, dim3(1), dim3(64), 0, stream, A.ctl);
    WF_HIP(hipGetLastError());
    if (tune.use_graph && !profile) {
        if (!e->have_graph || !same_config(A, e->key, tune, e->key_tune)) {
            if (e->exec)
                WF_HIP(hipGraphExecDestroy(e->exec));
            if (e->graph)
                WF_HIP(hipGraphDestroy(e->graph));
            e->exec = nullptr;
            e->graph = nullptr;
            e->have_graph = false;
            WF_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            record_generation(stream);
            WF_HIP(hipStreamEndCapture(stream, &e->graph));
            WF_HIP(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
            std::memcpy(&e->key, &A, sizeof A);
            e->key_tune = tune;
            e->have_graph = true;
        }
        for (uint32_t g = 0; g < generations; g++)
            WF_HIP(hipGraphLaunch(e->exec, stream));
    } else {
        for (uint32_t g = 0; g < generations; g++)
            record_generation(stream);
    }
    WF_HIP(hipGetLastError());
    return "";
#undef WF_HIP
}

} // namespace x3d
