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

#include <algorithm>
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

// Multi-job renders: a generation holds slices, each samples [s_first, s_first + count) of every
// pixel of one job; the slice's paths are [path_begin, path_begin + count * npix) of the
// generation (sample-major, as in a single-job generation) and its pixels [pix_begin, pix_begin
// + npix) of the generation's accumulation pass. Slices of a generation are stored in order,
// each beginning where the previous one ends.
struct WfSlice {
    uint32_t job, s_first, count, path_begin, pix_begin;
    uint32_t pad[3];
};
struct WfGen {
    uint32_t slice_begin, slice_end; // its slices
    uint32_t paths, pixels;          // totals
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
    // Fused form: path state follows the queues. xp, wi, beta, L, rng above are the state of the
    // paths in the extension queue this bounce consumes (entry k's state at index k: (position,
    // pixel), (direction, density), (throughput after the sample, job), (radiance, path id)), the
    // *_o arrays that of the queue it fills (xn and wo are not used); lfin_in / lfin_out hold
    // (radiance, path id) of the finishing queues. A shadow ray names the entry whose radiance it
    // adds to: an extension entry, or a finishing entry with kWfFinBit set.
    int compact;
    float4 *xp_o, *beta_o, *L_o, *wi_o;
    uint64_t *rng_o;
    float4 *lfin_in, *lfin_out;
    // Multi-job renders (the MULTI kernels): per job scene, camera, first hits and films, and per
    // generation its slices; the fields of one render above (S, cam, width ... first, acc, out)
    // are then unused. A path's job travels with it: in beta.w of its state and in the third
    // float4's w of its shadow ray (0 in a single-job render).
    const WfJob *jobs;
    const WfSlice *slices;
    const WfGen *gens;
};
constexpr uint32_t kWfFinBit = 0x80000000u;

// Slice of generation G that holds index q: a path id, or (by_pix) a pixel of the accumulation pass.
X3D_DI uint32_t find_slice(const WfSlice *sl, WfGen const &G, uint32_t q, bool by_pix) {
    uint32_t lo = G.slice_begin, hi = G.slice_end;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((by_pix ? sl[mid].pix_begin : sl[mid].path_begin) <= q)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

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
    float sdist, dens;
    uint64_t rng;  // the path's PCG state after both samples
    f3 beta_next;  // throughput after the BSDF sample: beta * f * cos / density (ok only)
};

// TR = kWfMis: the light sample's contribution carries the power-heuristic weight and the BSDF is
// sampled with the evaluated mixture densities (Tracer::mis_bounce); kWfLt1: unweighted
// (Tracer::direct as li_lt1 adds it) and the reference's mixture sampling; kWfUnidirect: BSDF
// sample only (no light sample).
template <int TR> X3D_DI NeeSample nee_sample(DevScene const &S, HitRec const &x, f3 wo, f3 beta, Pcg &g) {
    NeeSample r;
    r.has_shadow = false;
    r.sc = r.sdir = mk(0.0f, 0.0f, 0.0f);
    r.sdist = 0.0f;
    if (TR != kWfUnidirect && S.num_lights > 0) {
        LightSample ls = sample_light(S, g);
        f3 l, illum;
        float r2;
        if (light_arrives(S, ls, x.p, x.n, illum, l, r2)) {
            float dist = norm(l);
            f3 wl = -l / dist;
            f3 f = bsdf_eval(S.mats, x.mat, x.n, wo, wl);
            f3 d = mul(illum, f) / ls.area_dens;
            if (TR == kWfMis && !ls.sky) {
                float cl = dot(ls.n, -wl);
                float pl = ls.area_dens * r2 / cl;
                float pb = bsdf_pdf(S.mats, x.mat, x.n, wo, wl);
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
    f3 wi = bsdf_sample(S.mats, x.mat, x.n, wo, g, &dens, TR != kWfMis);
    r.rng = g.state;
    // li_mis stops on !(dens > 0); li_lt1 / li_unidirect on dens == 0
    r.ok = TR == kWfMis ? dens > 0.0f : dens != 0.0f;
    r.dens = dens;
    r.edir = wi;
    // The throughput update the tracers make once the sampled ray has hit (mis_bounce, li_lt1,
    // li_unidirect: beta * f * cos / density), computed here where the vertex is at hand, with the
    // same expression: the path state then needs neither the vertex's normal and material nor
    // its outgoing direction.
    r.beta_next = mk(0.0f, 0.0f, 0.0f);
    if (r.ok) {
        const float cw = dot(x.n, wi);
        const f3 f = bsdf_eval(S.mats, x.mat, x.n, wo, wi);
        r.beta_next = mul(beta, f) * cw / dens;
    }
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
// (block_append synchronizes); `scratch` holds 6 zeroed words. idx[1] / idx[2] receive the
// path's extension / finishing entry. In the compact form the caller stores the path's state at
// that entry; otherwise the queues hold path ids. The shadow ray carries the path's job.
X3D_DI void push_rays(WfArgs const &A, bool in, uint32_t path, f3 xp3, NeeSample const &r, uint32_t *scratch,
                      uint32_t idx[3], uint32_t job) {
    uint32_t *const queues[3] = {&A.ctl->sq_count.v, &A.eq_cnt_out->v, &A.fin_cnt_out->v};
    const bool want[3] = {in && r.has_shadow, in && r.ok, in && !r.ok};
    block_append<3>(queues, want, idx, scratch);
    if (want[0]) {
        uint32_t q = idx[0];
        const uint32_t tgt = !A.compact ? path : (r.ok ? idx[1] : (idx[2] | kWfFinBit));
        A.sq[3 * q] = make_float4(xp3.x, xp3.y, xp3.z, r.sdist - 1e-3f);
        A.sq[3 * q + 1] = make_float4(r.sdir.x, r.sdir.y, r.sdir.z, __uint_as_float(tgt));
        A.sq[3 * q + 2] = make_float4(r.sc.x, r.sc.y, r.sc.z, __uint_as_float(job));
        if (A.sort_rays)
            A.ray_key[0][q] = ray_key(A.S, xp3, r.sdir);
    }
    if (A.compact)
        return;
    if (want[1]) {
        A.eq_out[idx[1]] = path;
        if (A.sort_rays)
            A.ray_key[1][idx[1]] = ray_key(A.S, xp3, r.edir);
    }
    if (want[2])
        A.fin_out[idx[2]] = path;
}

// Compact form: the state of a live path after its vertex's samples, at its new queue entry:
// (vertex position, pixel), (sampled direction, its density), (throughput after the sample, job),
// (radiance, path id), PCG state.
X3D_DI void store_state(WfArgs const &A, uint32_t path, HitRec const &x, uint32_t pix, f3 L, NeeSample const &r,
                        uint32_t const idx[3], uint32_t job) {
    if (r.ok) {
        const uint32_t e = idx[1];
        A.xp_o[e] = make_float4(x.p.x, x.p.y, x.p.z, __uint_as_float(pix));
        A.wi_o[e] = make_float4(r.edir.x, r.edir.y, r.edir.z, r.dens);
        A.beta_o[e] = make_float4(r.beta_next.x, r.beta_next.y, r.beta_next.z, __uint_as_float(job));
        A.L_o[e] = make_float4(L.x, L.y, L.z, __uint_as_float(path));
        A.rng_o[e] = r.rng;
    } else {
        A.lfin_out[idx[2]] = make_float4(L.x, L.y, L.z, __uint_as_float(path));
    }
}

// Tracer::mis_begin / the first vertex of li_lt1, li_unidirect for path `id` (pixel pix, sample s
// of a film width x height of scene S seen through cam, first hits in `first`): the first hit's
// emission and, in the fused form, the vertex's light and BSDF samples. Returns whether the
// path goes on.
template <bool FUSED, int TR>
X3D_DI bool gen_path(WfArgs const &A, DevScene const &S, CamParams const &cam, uint32_t width, uint32_t height,
                     float4 *first, uint32_t id, uint32_t pix, uint32_t s, HitRec &x, f3 &wo, f3 &L, f3 &xp3,
                     NeeSample &r, int &kind) {
    f3 o, v;
    camera_ray(cam, pix % width, pix / width, width, height, o, v);
    RenderParams RP;
    RP.first = first;
    bool valid;
    x = load_first(RP, pix, valid);
    Pcg g = pcg_for(A.seed, pix, s);
    wo = -v;
    bool live;
    if (TR == kWfMis) {
        if (valid && x.light >= 0)
            L = L + light_radiance(S, x.light, wo, x.n);
        live = valid && A.max_depth > 0;
    } else if (TR == kWfLt1) {
        live = valid;
        if (valid) {
            f3 e = x.light >= 0 ? light_radiance(S, x.light, wo, x.n) : mk(0.0f, 0.0f, 0.0f);
            A.le0[id] = make_float4(e.x, e.y, e.z, x.light >= 0 ? 1.0f : 0.0f);
        }
    } else {
        live = valid;
        if (valid && x.light >= 0)
            L = L + mul(mk(1.0f, 1.0f, 1.0f), light_radiance(S, x.light, wo, x.n));
    }
    if (!live) {
        A.lbuf[id] = make_float4(L.x, L.y, L.z, 0.0f);
    } else if (FUSED) {
        // the state is stored at the path's queue entry, after the append
        xp3 = x.p;
        r = nee_sample<TR>(S, x, wo, mk(1.0f, 1.0f, 1.0f), g);
    } else {
        A.xp[id] = make_float4(x.p.x, x.p.y, x.p.z, __uint_as_float(static_cast<uint32_t>(x.mat)));
        A.xn[id] = make_float4(x.n.x, x.n.y, x.n.z, __uint_as_float(static_cast<uint32_t>(x.light)));
        A.wo[id] = make_float4(wo.x, wo.y, wo.z, __uint_as_float(pix));
        A.beta[id] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(0u));
        A.L[id] = make_float4(L.x, L.y, L.z, 0.0f);
        A.rng[id] = g.state;
        kind = kind_of(S, x.mat);
    }
    return live;
}

// MULTI: the generation's paths belong to the jobs of its slices (the fused MIS form only).
template <bool FUSED, int TR, bool MULTI = false> __global__ __launch_bounds__(256) void k_wf_generate(WfArgs A) {
    static_assert(!MULTI || (FUSED && TR == kWfMis), "multi-job generations: fused MIS form");
    __shared__ uint32_t scratch[2 * kKinds];
    clear_scratch(scratch, 2 * kKinds);
    uint32_t gen = A.ctl->gen;
    uint32_t n = MULTI ? A.gens[gen].paths : gen_count(A, gen) * A.npix;
    // block-uniform trip count (block_append synchronizes the block)
    for (uint32_t b = blockIdx.x * blockDim.x; b < n; b += gridDim.x * blockDim.x) {
        uint32_t id = b + threadIdx.x;
        bool live = false;
        int kind = 0;
        NeeSample r;
        r.has_shadow = r.ok = false;
        f3 xp3 = mk(0.0f, 0.0f, 0.0f);
        HitRec x{};
        f3 L = mk(0.0f, 0.0f, 0.0f), wo = L;
        uint32_t pix = 0, job = 0;
        if (id < n) {
            if constexpr (MULTI) {
                const WfSlice sl = A.slices[find_slice(A.slices, A.gens[gen], id, false)];
                WfJob const &J = A.jobs[sl.job];
                const uint32_t local = id - sl.path_begin, npix = J.npix;
                const uint32_t s_local = local / npix;
                pix = local - s_local * npix;
                job = sl.job;
                live = gen_path<FUSED, TR>(A, J.S, J.cam, J.width, J.height, J.first, id, pix, sl.s_first + s_local,
                                           x, wo, L, xp3, r, kind);
            } else {
                uint32_t s_local = id / A.npix;
                pix = id - s_local * A.npix;
                uint32_t s = A.s_begin + gen * A.gen_samples + s_local;
                live = gen_path<FUSED, TR>(A, A.S, A.cam, A.width, A.height, A.first, id, pix, s, x, wo, L, xp3, r,
                                           kind);
            }
        }
        if (FUSED) {
            uint32_t idx[3];
            push_rays(A, live, id, xp3, r, scratch, idx, job);
            if (live)
                store_state(A, id, x, pix, L, r, idx, job);
        } else {
            push_active(A, 0, kind, id, live, scratch);
        }
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
            r = nee_sample<kWfMis>(A.S, x, xyz(wo4), xyz(b4), g);
            A.rng[path] = r.rng;
            if (r.ok)
                A.wi[path] = make_float4(r.edir.x, r.edir.y, r.edir.z, r.dens);
        }
        // one call site with all lanes active: block_append synchronizes the block
        uint32_t idx[3];
        push_rays(A, in, path, xp3, r, scratch, idx, 0u);
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

// Persistent any-hit traversal over the shadow queue. MIS / lt1 rays: 3 x float4 (o, tmax),
// (d, path), (contribution, -), tmin 1e-4; an unoccluded ray adds its contribution to the path's
// radiance. BD (bidirectional connections): 2 x float4 (o, tmax), (d, slot | kBdLongTmin); tmin
// is 1e-3 where the flag is set (subpath connections), else 1e-4 (light samples), and the ray's
// visibility (1 = unoccluded, 0 = occluded) replaces its tmax.
// A multi-job ray's scene: the fields the traversal reads (both trees, triangles, roots).
X3D_DI void use_job_scene(WfArgs const &A, uint32_t job, DevScene &S, const float4 *&tris) {
    DevScene const &J = A.jobs[job].S;
    S.nodes = J.nodes;
    S.nodes4 = J.nodes4;
    S.tris = J.tris;
    S.root_ref = J.root_ref;
    S.root4 = J.root4;
    S.far_coords = J.far_coords;
    tris = static_cast<const float4 *>(J.tris);
}

constexpr uint32_t kBdLongTmin = 0x80000000u;
// MULTI: each ray traverses the scene of its path's job (the rays of one wave may differ).
template <int WG, int LDSD, bool BD = false, bool MULTI = false>
__global__ __launch_bounds__(WG) void k_wf_shadow(WfArgs A) {
    static_assert(!(BD && MULTI), "multi-job renders run the MIS form");
    __shared__ uint2 lds[LDSD * WG];
    __shared__ uint32_t cursor[2 * (WG / 64)];
    __shared__ float4 top[kTopNodes > 0 ? kTopNodes * kWideRows : 1];
    TravStack<WG, LDSD> st = make_stack<WG, LDSD>(lds, A.spill, A.spill_per_thread);
    DevScene S = A.S;
    if (kTopNodes > 0 && !MULTI) {
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
    WideSel sel{};
    float tmin = 1e-4f, tmax = 0.0f;
    uint32_t ref = kDone, leaf = kDone, cur = 0; // cur: next triangle of the parked leaf
    bool blocked = false, wide = false;
    constexpr uint32_t kStride = BD ? 2u : 3u;
    auto init = [&](uint32_t k) {
        float4 a = A.sq_in[kStride * k], b = A.sq_in[kStride * k + 1];
        if constexpr (MULTI)
            use_job_scene(A, __float_as_uint(A.sq_in[kStride * k + 2].w), S, tris);
        o = xyz(a);
        tmax = a.w;
        v = xyz(b);
        if (BD)
            tmin = (__float_as_uint(b.w) & kBdLongTmin) ? 1e-3f : 1e-4f;
        vinv = mk(1.0f / v.x, 1.0f / v.y, 1.0f / v.z);
        wide = !slab_may_nan(o, vinv, S.far_coords != 0);
        sel = wide_sel(vinv);
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
                ref = visit_wide<false>(S, ref, o, vinv, sel, tmin, tmax, st, top);
            } else {
                pops++;
                ref = visit_interior<false>(S, ref, o, vinv, tmin, tmax, __builtin_inff(), st, true);
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
                const float4 *tp = tris + 3 * static_cast<size_t>(p);
                float4 a = ldg(tp), b = ldg(tp + 1), c = ldg(tp + 2);
                float c1, c2, t0;
                tests++;
                blocked = tri_hit(o, v, xyz(a), xyz(b), xyz(c), tmin, tmax, &c1, &c2, &t0);
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
            if (BD) {
                A.sq[2 * static_cast<size_t>(i)].w = blocked ? 0.0f : 1.0f;
            } else if (!blocked) {
                const uint32_t tgt = __float_as_uint(A.sq_in[3 * i + 1].w);
                float4 *const Lp =
                    (A.compact && (tgt & kWfFinBit)) ? A.lfin_in + (tgt & ~kWfFinBit) : A.L + tgt;
                float4 c = A.sq_in[3 * i + 2];
                float4 L = *Lp;
                f3 nl = xyz(L) + xyz(c);
                *Lp = make_float4(nl.x, nl.y, nl.z, L.w);
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

// LDS stack depth of the closest-hit kernel paired with an any-hit depth D, and the resident
// waves per SIMD its registers are sized for (12 entries x 8 B x 64 lanes = 6 KiB per wave:
// six waves fit 160 KiB of LDS per CU, at <= 80 VGPRs).
template <int D> constexpr int kExtendDepth = D == 16 ? 12 : D;
template <int LDSD> constexpr int kExtendWaves = LDSD == 12 ? 6 : 4;

// MULTI: each ray traverses the scene of its path's job (compact form).
template <int WG, int LDSD, bool MULTI = false>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(kExtendWaves<LDSD>))) void k_wf_extend(WfArgs A) {
    __shared__ uint2 lds[LDSD * WG];
    __shared__ uint32_t cursor[2 * (WG / 64)];
    __shared__ float4 top[kTopNodes > 0 ? kTopNodes * kWideRows : 1];
    TravStack<WG, LDSD> st = make_stack<WG, LDSD>(lds, A.spill, A.spill_per_thread);
    DevScene S = A.S;
    if (kTopNodes > 0 && !MULTI) {
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
    WideSel sel{};
    ClosestHit h{-1, __builtin_inff(), 0.0f, 0.0f};
    // cull distance min(tmax, best t (1 + slack)), kept in step with h.t: children that pass the
    // slab test enter at most at tmax, so the tmax bound changes no cull decision
    float cutc = kTMax;
    uint32_t best_leaf = 0, ref = kDone, leaf = kDone, cur = 0; // cur: next triangle of the parked leaf
    bool wide = false;
    auto init = [&](uint32_t k) {
        const uint32_t path = A.compact ? k : A.eq_in[k];
        if constexpr (MULTI)
            use_job_scene(A, __float_as_uint(A.beta[k].w), S, tris);
        o = xyz(A.xp[path]);
        v = xyz(A.wi[path]);
        vinv = mk(1.0f / v.x, 1.0f / v.y, 1.0f / v.z);
        wide = !slab_may_nan(o, vinv, S.far_coords != 0);
        sel = wide_sel(vinv);
        h = ClosestHit{-1, __builtin_inff(), 0.0f, 0.0f};
        cutc = kTMax;
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
            if (wide) {
                pops += kTreeWidth / 2;
                ref = visit_wide(S, ref, o, vinv, sel, kTMin, cutc, st, top);
            } else {
                pops++;
                ref = visit_interior<false>(S, ref, o, vinv, kTMin, kTMax, cutc, st, true);
            }
            if (is_leaf_ref(ref) && leaf == kDone) {
                leaf = ref;
                cur = leaf_start(leaf);
                ref = st.pop_next(cutc);
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
                const float4 *tp = tris + 3 * static_cast<size_t>(p);
                float4 a = ldg(tp), b = ldg(tp + 1), c = ldg(tp + 2);
                float c1, c2, t0;
                tests++;
                if (tri_hit(o, v, xyz(a), xyz(b), xyz(c), kTMin, kTMax, &c1, &c2, &t0) &&
                    (t0 < h.t || (t0 == h.t && s0 > best_leaf))) {
                    h.t = t0;
                    h.b1 = c1;
                    h.b2 = c2;
                    h.prim = static_cast<int32_t>(p);
                    best_leaf = s0;
                    cutc = fminf(kTMax, t0 + t0 * kCullSlack);
                }
            }
            cur++;
            if (cur == s0 + leaf_count(leaf)) {
                if (is_leaf_ref(ref)) {
                    leaf = ref;
                    cur = leaf_start(leaf);
                    ref = st.pop_next(cutc);
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

// One extension entry of a bounce (state at si, hit t) in scene S: the hit vertex, its emission
// with the MIS weight (or lt1 / unidirect's Russian roulette and emission) and, in the fused
// form, its light and BSDF samples. Returns whether the path goes on; a path that ends is
// finished here.
template <bool FUSED, int TR>
X3D_DI bool shade_path(WfArgs const &A, DevScene const &S, uint32_t si, uint32_t t, uint32_t bounce, uint32_t &path,
                       uint32_t &pix, HitRec &nx, f3 &L, f3 &beta, f3 &wi, f3 &xp3, NeeSample &r, int &kind) {
    ClosestHit h = A.eh[t];
    // Every load of the item issues at once: the path's state and the hit triangle (a miss reads
    // triangle 0, which every scene with a possible hit has), so one memory round trip instead of
    // three.
    const uint32_t pi = h.prim >= 0 ? static_cast<uint32_t>(h.prim) : 0u;
    const float4 *tris = static_cast<const float4 *>(S.tris);
    const float4 *attr = static_cast<const float4 *>(S.tri_attr);
    float4 ta = tris[3 * pi], tb = tris[3 * pi + 1], tc = tris[3 * pi + 2];
    float4 na = attr[3 * pi], nb = attr[3 * pi + 1], nc = attr[3 * pi + 2];
    // compact form: (position, pixel), (direction, density), (throughput after the sample, job),
    // (radiance, path id); otherwise the per-path arrays with the vertex's normal and material
    // and its outgoing direction, the throughput before the sample
    float4 L4 = A.L[si], wi4 = A.wi[si], b4 = A.beta[si], xp4 = A.xp[si];
    float4 xn = FUSED ? float4{} : A.xn[si], wo4 = FUSED ? float4{} : A.wo[si];
    path = FUSED ? __float_as_uint(L4.w) : si;
    pix = __float_as_uint(FUSED ? xp4.w : wo4.w);
    L = xyz(L4);
    wi = xyz(wi4);
    bool hit = h.prim >= 0, live = false;
    if (hit) {
        nx = finish_hit(h, ta, tb, tc, na, nb, nc);
        hit = !(dot(nx.n, -wi) <= 0.0f);
    }
    if (!hit) {
        finish_path<TR>(A, path, L);
    } else {
        float dens = wi4.w;
        beta = xyz(b4);
        if (!FUSED) {
            int32_t xmat = static_cast<int32_t>(__float_as_uint(xp4.w));
            f3 xnrm = xyz(xn), wo = xyz(wo4);
            float cw = dot(xnrm, wi);
            f3 f = bsdf_eval(S.mats, xmat, xnrm, wo, wi);
            beta = mul(beta, f) * cw / dens;
        }
        bool go;
        Pcg g = path_pcg(0, pix);
        if (TR == kWfMis) {
            if (nx.light >= 0) {
                float cl = dot(nx.n, -wi);
                DevLight const &Ln = S.lights[nx.light];
                float pl = (Ln.prob / Ln.area) * (nx.t * nx.t) / cl;
                float w = dens * dens / (dens * dens + pl * pl);
                L = L + mul(beta, light_radiance(S, nx.light, -wi, nx.n) * w);
            }
            go = bounce + 1 < A.max_depth;
            if (go && FUSED)
                g.state = A.rng[si];
        } else {
            // li_lt1 / li_unidirect at depth bounce + 1: Russian roulette p = 0.5 from depth 2,
            // then (unidirect) the vertex's emission; the loop stops after kRefModeMaxDepth vertices
            const uint32_t depth = bounce + 1;
            go = depth < kRefModeMaxDepth;
            if (go) {
                g.state = A.rng[si];
                if (depth >= 2) {
                    if (g.draw() >= 0.5f)
                        go = false;
                    else
                        beta = beta / 0.5f;
                }
            }
            if (go && TR == kWfUnidirect && nx.light >= 0)
                L = L + mul(beta, light_radiance(S, nx.light, -wi, nx.n));
        }
        if (!go) {
            finish_path<TR>(A, path, L);
        } else {
            live = true;
            if (FUSED) {
                // the state is stored at the path's new queue entry, after the append
                xp3 = nx.p;
                r = nee_sample<TR>(S, nx, -wi, beta, g);
            } else {
                A.xp[path] = make_float4(nx.p.x, nx.p.y, nx.p.z, __uint_as_float(static_cast<uint32_t>(nx.mat)));
                A.xn[path] = make_float4(nx.n.x, nx.n.y, nx.n.z, __uint_as_float(static_cast<uint32_t>(nx.light)));
                A.wo[path] = make_float4(-wi.x, -wi.y, -wi.z, wo4.w);
                A.beta[path] = make_float4(beta.x, beta.y, beta.z, b4.w);
                A.L[path] = make_float4(L.x, L.y, L.z, 0.0f);
                kind = kind_of(S, nx.mat);
            }
        }
    }
    return live;
}

// MULTI: each entry is shaded in the scene of its path's job (the fused MIS form only).
template <bool FUSED, int TR, bool MULTI = false>
__global__ __launch_bounds__(256) void k_wf_shade(WfArgs A, int parity, uint32_t bounce) {
    static_assert(!MULTI || (FUSED && TR == kWfMis), "multi-job generations: fused MIS form");
    __shared__ uint32_t scratch[2 * kKinds];
    clear_scratch(scratch, 2 * kKinds);
    uint32_t ne = A.eq_cnt_in->v, nf = A.fin_cnt_in->v;
    uint32_t total = ne + nf;
    // block-uniform trip count (block_append synchronizes the block)
    for (uint32_t b = blockIdx.x * blockDim.x; b < total; b += gridDim.x * blockDim.x) {
        uint32_t t = b + threadIdx.x;
        bool live = false;
        int kind = 0;
        uint32_t path = 0, pix = 0, job = 0;
        NeeSample r;
        r.has_shadow = r.ok = false;
        f3 xp3 = mk(0.0f, 0.0f, 0.0f);
        HitRec nx{};
        f3 L = xp3, beta = xp3, wi = xp3;
        if (t >= total) {
            // no item: only takes part in the block-wide append below
        } else if (t >= ne) {
            if (FUSED) {
                // compact form: the finishing entry holds the path's radiance and id
                const float4 lf = A.lfin_in[t - ne];
                finish_path<TR>(A, __float_as_uint(lf.w), xyz(lf));
            } else {
                path = A.fin_in[t - ne];
                finish_path<TR>(A, path, xyz(A.L[path]));
            }
        } else {
            // compact form: entry t's state at index t; otherwise at its path id
            const uint32_t si = FUSED ? t : A.eq_in[t];
            if constexpr (MULTI) {
                job = __float_as_uint(A.beta[si].w);
                live = shade_path<FUSED, TR>(A, A.jobs[job].S, si, t, bounce, path, pix, nx, L, beta, wi, xp3, r,
                                             kind);
            } else {
                live = shade_path<FUSED, TR>(A, A.S, si, t, bounce, path, pix, nx, L, beta, wi, xp3, r, kind);
            }
        }
        if (FUSED) {
            uint32_t idx[3];
            push_rays(A, live, path, xp3, r, scratch, idx, job);
            if (live)
                store_state(A, path, nx, pix, L, r, idx, job);
        } else {
            push_active(A, parity ^ 1, kind, path, live, scratch);
        }
    }
}

// Multi-job form: every pixel of each slice of the generation adds the slice's samples of that
// pixel, in sample order, to its job's running sum (the job's first slice starts from zero, its
// last writes the film), as a single-job render of the job does over its generations.
__global__ __launch_bounds__(256) void k_wf_accumulate_jobs(WfArgs A) {
    const WfGen G = A.gens[A.ctl->gen];
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < G.pixels; q += gridDim.x * blockDim.x) {
        const WfSlice sl = A.slices[find_slice(A.slices, G, q, true)];
        WfJob const &J = A.jobs[sl.job];
        const uint32_t pix = q - sl.pix_begin, npix = J.npix;
        f3 acc = mk(0.0f, 0.0f, 0.0f);
        if (sl.s_first > 0) {
            float4 a = J.acc[pix];
            acc = mk(a.x, a.y, a.z);
        }
        for (uint32_t s = 0; s < sl.count; s++)
            acc = acc + xyz(A.lbuf[sl.path_begin + static_cast<size_t>(s) * npix + pix]);
        if (sl.s_first + sl.count >= J.spp) {
            const float scale = 1.0f / static_cast<float>(J.spp);
            J.out[pix] = make_float4(scale * acc.x, scale * acc.y, scale * acc.z, scale);
        } else {
            J.acc[pix] = make_float4(acc.x, acc.y, acc.z, 0.0f);
        }
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
// Wavefront form of the bidirectional tracer (Tracer::li_bdpt). One generation holds n (pixel,
// sample) items, item j = pixel j % npix, sample s_first + j / npix, as in k_render_bdpt:
//   bd_start    camera vertex 0 from the first-hit cache, first BSDF sample of the camera walk
//   extend      closest hits of the queued walk rays (the MIS form's traversal kernel)
//   bd_shade    per hit: the next camera or light vertex and the walk's next BSDF sample; when
//               a sample's camera walk ends, its emission sample and emission ray (the light
//               walk), as li_bdpt draws them from the same PCG stream
//   bd_connect  per (item, strategy): the shadow ray of each light-sample and subpath
//               connection strategy that passes its geometric test
//   shadow      any-hit traversal of those rays (visibility per ray)
//   bd_contrib  per ray: the value li_bdpt adds for its strategy (the connection's transport
//               and BSDFs, or the light sample's estimate), with li_bdpt's float expressions
//   bd_accumulate per item: li_bdpt's sums over path lengths and strategies, in its order
// Subpath vertices live in HBM, vertex k of every item side by side ([vertex][item]), each vertex
// one 64-byte record of four float4 (position and material, normal, direction to the previous
// vertex, prefix transport), so a walk bounce writes each vertex as one contiguous block although
// items arrive in queue order (four separate row arrays: 46.9 against 42.5 ms per render), instead
// of 1680 B of scratch per lane. Each item draws the same numbers from its PCG stream and
// evaluates the same expressions as li_bdpt, so the sample radiance is bit-identical to the plain
// kernel's.
// ---------------------------------------------------------------------------------------------
constexpr uint32_t kBdCam = 0, kBdLight0 = 1, kBdLight = 2, kBdDone = 3;
constexpr uint32_t kBdDead = 1u << 20;
constexpr uint32_t kBdNoRay = 0xFFFFFFFFu;
X3D_DI uint32_t bd_cam_len(uint32_t s) { return s & 31u; }
X3D_DI uint32_t bd_light_len(uint32_t s) { return (s >> 8) & 31u; }
X3D_DI uint32_t bd_phase(uint32_t s) { return (s >> 16) & 3u; }
X3D_DI uint32_t bd_state(uint32_t cl, uint32_t ll, uint32_t ph) { return cl | (ll << 8) | (ph << 16); }

struct BdArgs {
    DevScene S;
    CamParams cam;
    uint32_t width, height, npix, n;
    uint32_t s_first, max_len, nstrat; // nstrat = (max_len + 1)^2 strategy slots per item
    uint64_t seed;
    const float4 *first; // first-hit cache (RenderParams::first)
    uint32_t *state;     // per item: lengths, phase, dead flag
    uint64_t *rng;       // per item: PCG state (the stream's increment follows from the pixel)
    float4 *ray_o, *ray_d; // per item: the walk's queued ray (d.w = its BSDF density)
    float4 *em0, *em1, *emis; // per item: emission sample (p, area_dens), (n, light | sky << 31),
                              // light_emission of li_bdpt
    float4 *cv, *lv;          // camera / light vertices: vertex k of item j at [4 (k n + j)]: (p, mat), n, v, transport
    uint32_t *eq_in, *eq_out;
    WfCounter *eq_cnt_in, *eq_cnt_out;
    const ClosestHit *eh;
    float4 *sq;          // shadow rays, 2 x float4
    WfCounter *sq_count;
    uint32_t *rix;       // [strategy * n + j]: queue entry of the strategy's shadow ray, or kBdNoRay
    float4 *samples;     // sample radiance [j]
};

struct BdVert {
    f3 p, n, v, t;
    int32_t mat;
};

X3D_DI BdVert bd_load(const float4 *a, uint32_t n, uint32_t k, uint32_t j) {
    size_t q = 4 * (static_cast<size_t>(k) * n + j);
    float4 p = a[q], nn = a[q + 1], v = a[q + 2], t = a[q + 3];
    return BdVert{xyz(p), xyz(nn), xyz(v), xyz(t), static_cast<int32_t>(__float_as_uint(p.w))};
}
X3D_DI void bd_store(float4 *a, uint32_t n, uint32_t k, uint32_t j, f3 p, int32_t mat, f3 nn, f3 v, f3 t) {
    size_t q = 4 * (static_cast<size_t>(k) * n + j);
    a[q] = make_float4(p.x, p.y, p.z, __uint_as_float(static_cast<uint32_t>(mat)));
    a[q + 1] = make_float4(nn.x, nn.y, nn.z, 0.0f);
    a[q + 2] = make_float4(v.x, v.y, v.z, 0.0f);
    a[q + 3] = make_float4(t.x, t.y, t.z, 0.0f);
}

X3D_DI LightSample bd_em(BdArgs const &B, uint32_t j) {
    float4 a = B.em0[j], b = B.em1[j];
    uint32_t w = __float_as_uint(b.w);
    LightSample ls;
    ls.p = xyz(a);
    ls.area_dens = a.w;
    ls.n = xyz(b);
    ls.light = static_cast<int32_t>(w & 0x7FFFFFFFu);
    ls.sky = (w >> 31) != 0;
    return ls;
}

// The walk of item j after its last vertex `last` (of the subpath that `ph` names): the next
// BSDF sample of that subpath (extend_walk's loop; `ended`: the walk already stopped at a miss
// or a back-facing hit), or, once the camera walk is over, the emission sample that starts the
// light walk; once the light walk is over, the item is done. Returns true when a ray is queued
// (ray_o / ray_d written) and stores the new state and PCG state.
X3D_DI bool bd_advance(BdArgs const &B, uint32_t j, Pcg &g, uint32_t cl, uint32_t ll, uint32_t ph, bool ended,
                       BdVert const &last) {
    DevScene const &S = B.S;
    bool queued = false;
    if (ended) {
        // no further BSDF sample on this subpath
    } else if (ph == kBdCam && cl < B.max_len) {
        BdVert const &cur = last;
        float dens;
        f3 i = bsdf_sample(S.mats, cur.mat, cur.n, cur.v, g, &dens, true);
        if (!near0(dens)) {
            B.ray_o[j] = make_float4(cur.p.x, cur.p.y, cur.p.z, 0.0f);
            B.ray_d[j] = make_float4(i.x, i.y, i.z, dens);
            queued = true;
        }
    } else if (ph == kBdLight && ll < B.max_len) {
        BdVert const &cur = last;
        float dens;
        f3 i = bsdf_sample(S.mats, cur.mat, cur.n, cur.v, g, &dens, true);
        if (!near0(dens)) {
            B.ray_o[j] = make_float4(cur.p.x, cur.p.y, cur.p.z, 0.0f);
            B.ray_d[j] = make_float4(i.x, i.y, i.z, dens);
            queued = true;
        }
    }
    if (!queued && ph == kBdCam) {
        // camera walk over: emission sample (li_bdpt after extend_walk<false>)
        LightSample em = sample_light(S, g);
        float e0 = g.draw();
        float e1 = g.draw();
        float w_dens;
        f3 w = emission_dir(em, e0, e1, w_dens);
        float cos_e = dot(em.n, w);
        f3 emitted = cos_e > 0.0f ? light_rad(S, em.light) * cos_e : mk(0.0f, 0.0f, 0.0f);
        f3 light_emission = em.area_dens > 0.0f ? emitted / (w_dens * em.area_dens) : mk(0.0f, 0.0f, 0.0f);
        B.em0[j] = make_float4(em.p.x, em.p.y, em.p.z, em.area_dens);
        B.em1[j] = make_float4(em.n.x, em.n.y, em.n.z,
                               __uint_as_float(static_cast<uint32_t>(em.light) | (em.sky ? 0x80000000u : 0u)));
        B.emis[j] = make_float4(light_emission.x, light_emission.y, light_emission.z, 0.0f);
        B.ray_o[j] = make_float4(em.p.x, em.p.y, em.p.z, 0.0f);
        B.ray_d[j] = make_float4(w.x, w.y, w.z, 0.0f);
        ph = kBdLight0;
        queued = true;
    } else if (!queued) {
        ph = kBdDone;
    }
    B.state[j] = bd_state(cl, ll, ph);
    B.rng[j] = g.state;
    return queued;
}

__global__ __launch_bounds__(256) void k_bd_start(BdArgs B) {
    __shared__ uint32_t scratch[2];
    clear_scratch(scratch, 2);
    DevScene const &S = B.S;
    for (uint32_t b = blockIdx.x * blockDim.x; b < B.n; b += gridDim.x * blockDim.x) {
        const uint32_t j = b + threadIdx.x;
        bool queued = false;
        if (j < B.n) {
            const uint32_t pix = j % B.npix, s = B.s_first + j / B.npix;
            f3 o, v;
            camera_ray(B.cam, pix % B.width, pix / B.width, B.width, B.height, o, v);
            float4 fa = B.first[3 * pix], fb = B.first[3 * pix + 1], fc = B.first[3 * pix + 2];
            const bool valid = fc.z != 0.0f;
            if (!valid || B.max_len == 0 || S.num_lights == 0) {
                B.state[j] = kBdDead;
            } else {
                Pcg g = pcg_for(B.seed, pix, s);
                const BdVert v0{xyz(fa), xyz(fb), -v, mk(1.0f, 1.0f, 1.0f), static_cast<int32_t>(__float_as_uint(fc.x))};
                bd_store(B.cv, B.n, 0, j, v0.p, v0.mat, v0.n, v0.v, v0.t);
                queued = bd_advance(B, j, g, 1, 0, kBdCam, false, v0);
            }
        }
        uint32_t *const c[1] = {&B.eq_cnt_out->v};
        const bool p[1] = {queued};
        uint32_t idx[1];
        block_append<1>(c, p, idx, scratch);
        if (queued)
            B.eq_out[idx[0]] = j;
    }
}

__global__ __launch_bounds__(256) void k_bd_shade(BdArgs B) {
    __shared__ uint32_t scratch[2];
    clear_scratch(scratch, 2);
    DevScene const &S = B.S;
    const uint32_t total = B.eq_cnt_in->v;
    for (uint32_t b = blockIdx.x * blockDim.x; b < total; b += gridDim.x * blockDim.x) {
        const uint32_t t = b + threadIdx.x;
        bool queued = false;
        uint32_t j = 0;
        if (t < total) {
            j = B.eq_in[t];
            const uint32_t st = B.state[j], pix = j % B.npix;
            uint32_t cl = bd_cam_len(st), ll = bd_light_len(st), ph = bd_phase(st);
            Pcg g = path_pcg(B.rng[j], pix);
            const ClosestHit h = B.eh[t];
            const float4 rd = B.ray_d[j];
            const f3 i = xyz(rd);
            HitRec nx;
            bool hit = h.prim >= 0;
            if (hit) {
                nx = finish_hit(S, h);
                hit = dot(nx.n, -i) > 0.0f;
            }
            BdVert nv{};
            if (ph == kBdLight0) {
                // the emission ray: the light walk's first vertex, if it hits a surface facing it
                if (hit) {
                    nv = BdVert{nx.p, nx.n, -i, mk(1.0f, 1.0f, 1.0f), nx.mat};
                    bd_store(B.lv, B.n, 0, j, nv.p, nv.mat, nv.n, nv.v, nv.t);
                    ll = 1;
                    ph = kBdLight;
                } else {
                    ph = kBdDone;
                }
            } else if (hit) {
                const bool cam = ph == kBdCam;
                const uint32_t k = cam ? cl : ll;
                BdVert cur = bd_load(cam ? B.cv : B.lv, B.n, k - 1, j);
                float cos_w = dot(cur.n, i);
                // extend_walk: projected_adjoint_brdf on the camera walk, projected_brdf on the
                // light walk
                f3 f = cam ? bsdf_eval(S.mats, cur.mat, cur.n, cur.v, i) : bsdf_eval(S.mats, cur.mat, cur.n, i, cur.v);
                nv = BdVert{nx.p, nx.n, -i, mul(cur.t, (f * cos_w) / rd.w), nx.mat};
                bd_store(cam ? B.cv : B.lv, B.n, k, j, nv.p, nv.mat, nv.n, nv.v, nv.t);
                if (cam)
                    cl++;
                else
                    ll++;
            }
            if (ph == kBdDone) {
                B.state[j] = bd_state(cl, ll, kBdDone);
            } else {
                // a camera walk that stopped here goes on to its emission sample; a light walk
                // that stopped here is done
                queued = bd_advance(B, j, g, cl, ll, ph, !hit, nv);
            }
        }
        uint32_t *const c[1] = {&B.eq_cnt_out->v};
        const bool p[1] = {queued};
        uint32_t idx[1];
        block_append<1>(c, p, idx, scratch);
        if (queued)
            B.eq_out[idx[0]] = j;
    }
}

// Reserves `cnt` consecutive entries of a queue for each thread of the block: one global atomic
// per block (wave-level inclusive scans, then the waves' totals in LDS). Every thread of the
// block must call it; `scratch` holds WG / 64 + 1 words. Returns the thread's first entry.
template <int WG> X3D_DI uint32_t block_reserve(uint32_t cnt, uint32_t *counter, uint32_t *scratch) {
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    uint32_t incl = cnt;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        uint32_t up = __shfl_up(incl, d, 64);
        if (lane >= d)
            incl += up;
    }
    if (lane == 63)
        scratch[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < WG / 64; w++) {
            uint32_t c = scratch[w];
            scratch[w] = tot;
            tot += c;
        }
        scratch[WG / 64] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    const uint32_t base = scratch[WG / 64] + scratch[wave] + incl - cnt;
    __syncthreads(); // scratch is reused by the next call
    return base;
}

// Shadow rays of item j's strategies, in li_bdpt's tests: light-sample strategies (lp = 0:
// Tracer::direct's ray from camera vertex cp - 1 towards the emission sample,
// [1e-4, dist - 1e-3]) and subpath connections (lj -> cj, [1e-3, dist - 1e-3]) whose cosines
// are both positive. WRITE = false counts them; WRITE = true stores them from queue entry `at`
// on, tagged with their strategy slot index (cp * (max_len + 1) + lp) * n + j, and records in
// rix every strategy's queue entry (kBdNoRay where none was queued).
template <bool WRITE> X3D_DI uint32_t bd_connections(BdArgs const &B, uint32_t j, uint32_t st, uint32_t at) {
    DevScene const &S = B.S;
    const uint32_t side = B.max_len + 1, cam_len = bd_cam_len(st), light_len = bd_light_len(st);
    const LightSample em = bd_em(B, j);
    uint32_t cnt = 0;
    auto emit = [&](f3 o, f3 d, float tmax, uint32_t slot, uint32_t tag) {
        if (WRITE) {
            const size_t e = 2 * static_cast<size_t>(at + cnt);
            B.sq[e] = make_float4(o.x, o.y, o.z, tmax);
            B.sq[e + 1] = make_float4(d.x, d.y, d.z, __uint_as_float(slot | tag));
            B.rix[slot] = at + cnt;
        }
        cnt++;
    };
    // The first kBdRegLight light vertices' positions and normals are read once per item and
    // kept in registers (every camera vertex is joined to each of them); deeper ones are read
    // per camera vertex.
    constexpr uint32_t kBdRegLight = 8;
    f3 rpos[kBdRegLight], rnrm[kBdRegLight];
#pragma unroll
    for (uint32_t k = 0; k < kBdRegLight; k++) {
        rpos[k] = rnrm[k] = mk(0.0f, 0.0f, 0.0f);
        if (k < light_len) {
            const size_t ql = static_cast<size_t>(k) * B.n + j;
            rpos[k] = xyz(B.lv[4 * ql]);
            rnrm[k] = xyz(B.lv[4 * ql + 1]);
        }
    }
    auto connect = [&](f3 cpos, f3 cnrm, f3 lpos, f3 lnrm, uint32_t slot) {
        f3 join = cpos - lpos;
        float dist = norm(join);
        join = join / dist;
        float cos_wo = dot(lnrm, join);
        float cos_wi = dot(cnrm, -join);
        if (cos_wo > 0.0f && cos_wi > 0.0f)
            emit(lpos, join, dist - 1e-3f, slot, kBdLongTmin);
        else if (WRITE)
            B.rix[slot] = kBdNoRay;
    };
    for (uint32_t cp = 1; cp <= cam_len; cp++) {
        const size_t qc = static_cast<size_t>(cp - 1) * B.n + j;
        const f3 cpos = xyz(B.cv[4 * qc]), cnrm = xyz(B.cv[4 * qc + 1]);
        f3 l, illum;
        float r2;
        const uint32_t row = cp * side * B.n + j;
        if (light_arrives(S, em, cpos, cnrm, illum, l, r2)) {
            float dist = norm(l);
            emit(cpos, -l / dist, dist - 1e-3f, row, 0u);
        } else if (WRITE) {
            B.rix[row] = kBdNoRay;
        }
#pragma unroll
        for (uint32_t k = 0; k < kBdRegLight; k++)
            if (k < light_len)
                connect(cpos, cnrm, rpos[k], rnrm[k], row + (k + 1) * B.n);
        for (uint32_t lp = kBdRegLight + 1; lp <= light_len; lp++) {
            const size_t ql = static_cast<size_t>(lp - 1) * B.n + j;
            connect(cpos, cnrm, xyz(B.lv[4 * ql]), xyz(B.lv[4 * ql + 1]), row + lp * B.n);
        }
    }
    return cnt;
}

// Queues every item's shadow rays (counted, one reservation per block, then written).
__global__ __launch_bounds__(256) void k_bd_connect(BdArgs B) {
    __shared__ uint32_t scratch[256 / 64 + 1];
    for (uint32_t b = blockIdx.x * blockDim.x; b < B.n; b += gridDim.x * blockDim.x) {
        const uint32_t j = b + threadIdx.x;
        uint32_t st = kBdDead;
        if (j < B.n)
            st = B.state[j];
        const uint32_t cnt = (st & kBdDead) ? 0u : bd_connections<false>(B, j, st, 0);
        const uint32_t at = block_reserve<256>(cnt, &B.sq_count->v, scratch);
        if (!(st & kBdDead))
            bd_connections<true>(B, j, st, at);
    }
}

// Per shadow ray (one thread each): the value li_bdpt adds to its path length's sum for the
// ray's strategy, written over the ray's first float4 (w = 1: add it, w = 0: add nothing).
// Light sample (lp = 0): mul(tr, cj.transport), tr = Tracer::direct's estimate if the ray is
// unoccluded, else 0 (it is added either way). Connection: mul(tr, cj.transport) with the
// connection's transport tr if unoccluded; an occluded connection adds nothing.
__global__ __launch_bounds__(256) void k_bd_contrib(BdArgs B) {
    DevScene const &S = B.S;
    const uint32_t total = B.sq_count->v, side = B.max_len + 1;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const size_t e = 2 * static_cast<size_t>(i);
        const bool visible = B.sq[e].w != 0.0f;
        const uint32_t slot = __float_as_uint(B.sq[e + 1].w) & ~kBdLongTmin;
        const uint32_t k = slot / B.n, j = slot - k * B.n, cp = k / side, lp = k - cp * side;
        const BdVert cj = bd_load(B.cv, B.n, cp - 1, j);
        f3 val = mk(0.0f, 0.0f, 0.0f);
        float add = 1.0f;
        if (lp == 0) {
            f3 tr = mk(0.0f, 0.0f, 0.0f);
            if (visible) {
                const LightSample em = bd_em(B, j);
                f3 l, illum;
                float r2;
                (void)light_arrives(S, em, cj.p, cj.n, illum, l, r2); // held when the ray was queued
                float dist = norm(l);
                f3 wi = -l / dist;
                f3 f = bsdf_eval(S.mats, cj.mat, cj.n, cj.v, wi);
                tr = mul(illum, f) / em.area_dens;
            }
            val = mul(tr, cj.t);
        } else if (visible) {
            const BdVert lj = bd_load(B.lv, B.n, lp - 1, j);
            f3 join = cj.p - lj.p;
            float dist = norm(join);
            join = join / dist;
            float cos_wo = dot(lj.n, join);
            float cos_wi = dot(cj.n, -join);
            f3 sub = mul(xyz(B.emis[j]), lj.t);
            float geo = cos_wi * cos_wo / (dist * dist);
            f3 fl = bsdf_eval(S.mats, lj.mat, lj.n, join, lj.v);
            f3 fc = bsdf_eval(S.mats, cj.mat, cj.n, cj.v, -join);
            f3 tr = mul(mul(sub, fl), fc) * geo;
            val = mul(tr, cj.t);
        } else {
            add = 0.0f;
        }
        B.sq[e] = make_float4(val.x, val.y, val.z, add);
    }
}

// li_bdpt's transport_all_connectible_subpaths for each item: per path length, the sum over its
// strategies in li_bdpt's order (the values of k_bd_contrib; a light-sample strategy without a
// ray adds mul(0, cj.transport), as li_bdpt does) divided by the strategy count; the sample's
// radiance goes to samples[j].
__global__ __launch_bounds__(256) void k_bd_accumulate(BdArgs B) {
    DevScene const &S = B.S;
    const uint32_t side = B.max_len + 1;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < B.n; j += gridDim.x * blockDim.x) {
        const uint32_t st = B.state[j];
        f3 rad = mk(0.0f, 0.0f, 0.0f);
        if (!(st & kBdDead)) {
            const uint32_t cam_len = bd_cam_len(st), light_len = bd_light_len(st), pix = j % B.npix;
            const int32_t first_light = static_cast<int32_t>(__float_as_uint(B.first[3 * pix + 2].y));
            for (uint32_t plen = 1; plen <= cam_len + light_len + 1; plen++) {
                int cp = static_cast<int>(plen - 1 < cam_len ? plen - 1 : cam_len);
                uint32_t lp = plen - 1 - static_cast<uint32_t>(cp);
                f3 psum = mk(0.0f, 0.0f, 0.0f);
                float wsum = 0.0f;
                for (; cp >= 0 && lp <= light_len; lp++, cp--) {
                    const uint32_t slot = (static_cast<uint32_t>(cp) * side + lp) * B.n + j;
                    if (lp == 0 && cp == 0) {
                        if (first_light >= 0) {
                            const f3 v0 = xyz(B.cv[4 * static_cast<size_t>(j) + 2]), n0 = xyz(B.cv[4 * static_cast<size_t>(j) + 1]);
                            psum = psum + light_radiance(S, first_light, v0, n0);
                        }
                        wsum = wsum + 1.0f;
                    } else if (lp == 0) {
                        const uint32_t r = B.rix[slot];
                        if (r == kBdNoRay) {
                            const f3 t = xyz(B.cv[4 * (static_cast<size_t>(cp - 1) * B.n + j) + 3]);
                            psum = psum + mul(mk(0.0f, 0.0f, 0.0f), t);
                        } else {
                            psum = psum + xyz(B.sq[2 * static_cast<size_t>(r)]);
                        }
                        wsum = wsum + 1.0f;
                    } else if (cp > 0) {
                        const uint32_t r = B.rix[slot];
                        if (r != kBdNoRay) {
                            const float4 c = B.sq[2 * static_cast<size_t>(r)];
                            if (c.w != 0.0f)
                                psum = psum + xyz(c);
                        }
                        wsum = wsum + 1.0f;
                    }
                }
                if (wsum > 0.0f)
                    rad = rad + psum / wsum;
            }
        }
        B.samples[j] = make_float4(rad.x, rad.y, rad.z, 0.0f);
    }
}

// Before each walk bounce's shade: the traversal cursor restarts and the next bounce's queue
// starts empty. Before the connections: the shadow queue and its cursor start empty.
__global__ void k_bd_reset(WfCtl *ctl, WfCounter *eq_out, int shadow) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->eq_head.v = 0;
        if (eq_out)
            eq_out->v = 0;
        if (shadow) {
            ctl->sq_count.v = 0;
            ctl->sq_head.v = 0;
        }
    }
}

// Query counts of the walk bounce (closest hits) or of the connections (any hits).
__global__ void k_bd_count(WfCounter const *c, Counters *cnt, int anyhit) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && cnt)
        atomicAdd(anyhit ? &cnt->anyhit : &cnt->closest, static_cast<unsigned long long>(c->v));
}

// ---------------------------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------------------------
struct WavefrontEngine {
    DevBuf xp, xn, wo, beta, L, wi, rng, lbuf, act0, act1, sq, eq, fin, eh, ctl, spill;
    DevBuf eq_b, fin_b; // the fused form's second queue parity
    // the fused form's second parity of the path state, and the finishing queues' (radiance, id)
    DevBuf xp_b, beta_b, L_b, wi_b, rng_b, lfin_a, lfin_b;
    DevBuf le0;         // lt1's first-hit emission per path
    DevBuf sq_sorted, eq_sorted, ray_key, ray_rank, bin_mat, bin_tot; // ray reordering
    // bidirectional form: per item state, PCG state, walk ray, emission sample, vertices,
    // visibilities
    DevBuf bd_state, bd_rng, bd_ro, bd_rd, bd_em, bd_cv, bd_lv, bd_rix;
    // multi-job renders: job, slice and generation tables; an event at the end of each generation
    DevBuf mj_jobs, mj_slices, mj_gens;
    std::vector<hipEvent_t> gen_ev;
    std::vector<hipGraph_t> graphs;
    std::vector<hipGraphExec_t> execs;
    // segmented generations: per segment the sizes of the queues after it (pinned host memory)
    // and an event after their copy
    uint32_t *pinned = nullptr;
    uint32_t pinned_n = 0;
    std::vector<hipEvent_t> seg_ev;
    WfArgs key{};
    WavefrontTuning key_tune{};
    bool have_graph = false;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    // a bounce's shadow and extension traversals run concurrently: the shadow kernel on a second
    // stream, forked from and joined back into the render's stream
    hipStream_t s2 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    double shadow_ms = 0, extend_ms = 0;
    uint32_t trace_launches = 0;
    ~WavefrontEngine() {
        for (auto &x : ev)
            if (x)
                (void)hipEventDestroy(x);
        for (auto &x : gen_ev)
            (void)hipEventDestroy(x);
        if (fork)
            (void)hipEventDestroy(fork);
        if (join)
            (void)hipEventDestroy(join);
        if (s2)
            (void)hipStreamDestroy(s2);
        for (hipGraphExec_t x : execs)
            (void)hipGraphExecDestroy(x);
        for (hipGraph_t x : graphs)
            (void)hipGraphDestroy(x);
        for (auto &x : seg_ev)
            (void)hipEventDestroy(x);
        if (pinned)
            (void)hipHostFree(pinned);
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
    TraceKernel shadow, extend, bd_shadow;
    int wg;
    int ldsd; // the shallowest LDS stack of the three kernels: the spill slices are sized for it
};

// Stack depth D pairs the any-hit kernels' D LDS entries with kExtendDepth<D> for the
// closest-hit kernel (12 beside 16: 6 instead of 5 resident waves per SIMD for the extension
// rays, whose kernel is VALU-bound with divergent lanes, while more resident shadow rays thrash
// the caches).
template <int WG, int D, bool M> TraceKernels trace_kernels() {
    return {k_wf_shadow<WG, D, false, M>, k_wf_extend<WG, kExtendDepth<D>, M>, k_wf_shadow<WG, D, true>, WG,
            kExtendDepth<D> < D ? kExtendDepth<D> : D};
}

template <bool M> TraceKernels pick_trace_of(int wg, bool shallow) {
    switch (wg) {
    case 64: return shallow ? trace_kernels<64, 8, M>() : trace_kernels<64, 16, M>();
    case 128: return shallow ? trace_kernels<128, 8, M>() : trace_kernels<128, 16, M>();
    default: return shallow ? trace_kernels<256, 8, M>() : trace_kernels<256, 16, M>();
    }
}

// Workgroup sizes 64 / 128 / 256 and LDS stack depths 8 / 16 (deeper entries spill to memory);
// `multi`: the kernels of multi-job renders.
TraceKernels pick_trace(int wg, int ldsd, bool multi = false) {
    return multi ? pick_trace_of<true>(wg, ldsd == 8) : pick_trace_of<false>(wg, ldsd == 8);
}

// A multi-job render's generations (device tables) and the host copy of its jobs.
struct JobsPlan {
    std::vector<WfJob> const *jobs;
    const WfJob *d_jobs;
    const WfSlice *d_slices;
    const WfGen *d_gens;
    uint32_t generations, cap;
};

bool same_config(WfArgs const &a, WfArgs const &b, WavefrontTuning const &ta, WavefrontTuning const &tb) {
    return std::memcmp(&a, &b, sizeof(WfArgs)) == 0 && ta.paths_in_flight == tb.paths_in_flight &&
           ta.trace_wg == tb.trace_wg && ta.stack_depth == tb.stack_depth &&
           ta.sort_by_material == tb.sort_by_material && ta.sort_rays == tb.sort_rays &&
           ta.refill_below == tb.refill_below && ta.leaf_below == tb.leaf_below && ta.tracer == tb.tracer;
}

} // namespace


uint64_t bd_item_bytes(uint32_t max_len) {
    const uint64_t m = max_len, strat = (m + 1) * (m + 1);
    return 4 + 8 + 5 * 16          // state, PCG state, walk ray, emission sample
           + 2 * 4 * 16 * m        // camera and light vertices
           + 2 * 4 + sizeof(ClosestHit) // both queue parities, hit records
           + 2 * 16 * (m + m * m)  // shadow rays (every light-sample and connection strategy)
           + 4 * strat             // queue entry per strategy
           + 16;                   // the sample radiance (render's sample buffer)
}

uint32_t bd_max_items(uint32_t max_len) {
    uint64_t cap = kBdBudgetBytes / bd_item_bytes(max_len);
    const uint64_t strat = static_cast<uint64_t>(max_len + 1) * (max_len + 1);
    const uint64_t slot_cap = (static_cast<uint64_t>(1) << 31) / strat; // slot ids below kBdLongTmin
    cap = cap < slot_cap ? cap : slot_cap;
    return static_cast<uint32_t>(cap < (1u << 26) ? cap : (1u << 26));
}

std::string bd_generation(WavefrontEngine *e, RenderParams const &P, WavefrontTuning const &tune, uint32_t s_first,
                          uint32_t gen_samples, float4 *samples, bool first_gen, hipStream_t stream) {
#define BD_HIP(call)                                                                                 \
    do {                                                                                             \
        hipError_t err_ = (call);                                                                    \
        if (err_ != hipSuccess)                                                                      \
            return std::string(#call) + ": " + hipGetErrorString(err_);                              \
    } while (0)
    const uint32_t npix = P.width * P.height, m = P.max_path_len;
    const uint64_t n64 = static_cast<uint64_t>(gen_samples) * npix;
    if (n64 == 0)
        return "";
    if (m < 1 || m > kMaxPathLen || n64 > bd_max_items(m))
        return "bidirectional generation of " + std::to_string(n64) + " items exceeds its working memory";
    const uint32_t n = static_cast<uint32_t>(n64), side = m + 1, nstrat = side * side;
    const size_t f4 = sizeof(float4) * static_cast<size_t>(n);
    BD_HIP(e->bd_state.ensure(sizeof(uint32_t) * static_cast<size_t>(n)));
    BD_HIP(e->bd_rng.ensure(sizeof(uint64_t) * static_cast<size_t>(n)));
    BD_HIP(e->bd_ro.ensure(f4));
    BD_HIP(e->bd_rd.ensure(f4));
    BD_HIP(e->bd_em.ensure(3 * f4));
    BD_HIP(e->bd_cv.ensure(4 * m * f4));
    BD_HIP(e->bd_lv.ensure(4 * m * f4));
    BD_HIP(e->bd_rix.ensure(sizeof(uint32_t) * static_cast<size_t>(n) * nstrat));
    BD_HIP(e->eq.ensure(sizeof(uint32_t) * static_cast<size_t>(n)));
    BD_HIP(e->eq_b.ensure(sizeof(uint32_t) * static_cast<size_t>(n)));
    BD_HIP(e->eh.ensure(sizeof(ClosestHit) * static_cast<size_t>(n)));
    BD_HIP(e->sq.ensure(2 * f4 * (m + static_cast<size_t>(m) * m)));
    BD_HIP(e->ctl.ensure(sizeof(WfCtl)));
    TraceKernels tk = pick_trace(tune.trace_wg, tune.stack_depth);
    const int wg = tk.wg;
    const uint32_t g_shadow = grid_for(tk.bd_shadow, wg), g_extend = grid_for(tk.extend, wg);
    const uint32_t trace_threads = (g_shadow > g_extend ? g_shadow : g_extend) * wg;
    const uint32_t spill_per_thread = spill_entries_for(P.S, tk.ldsd);
    BD_HIP(e->spill.ensure(sizeof(uint2) * static_cast<size_t>(trace_threads) * spill_per_thread));

    WfCtl *ctl = static_cast<WfCtl *>(e->ctl.p);
    uint32_t *qs[2] = {static_cast<uint32_t *>(e->eq.p), static_cast<uint32_t *>(e->eq_b.p)};
    WfCounter *qc[2] = {&ctl->eq_count, &ctl->eq_count_b};

    BdArgs B{};
    B.S = P.S;
    B.cam = P.cam;
    B.width = P.width;
    B.height = P.height;
    B.npix = npix;
    B.n = n;
    B.s_first = s_first;
    B.max_len = m;
    B.nstrat = nstrat;
    B.seed = P.seed;
    B.first = P.first;
    B.state = static_cast<uint32_t *>(e->bd_state.p);
    B.rng = static_cast<uint64_t *>(e->bd_rng.p);
    B.ray_o = static_cast<float4 *>(e->bd_ro.p);
    B.ray_d = static_cast<float4 *>(e->bd_rd.p);
    B.em0 = static_cast<float4 *>(e->bd_em.p);
    B.em1 = B.em0 + n;
    B.emis = B.em1 + n;
    B.cv = static_cast<float4 *>(e->bd_cv.p);
    B.lv = static_cast<float4 *>(e->bd_lv.p);
    B.eh = static_cast<const ClosestHit *>(e->eh.p);
    B.sq = static_cast<float4 *>(e->sq.p);
    B.sq_count = &ctl->sq_count;
    B.rix = static_cast<uint32_t *>(e->bd_rix.p);
    B.samples = samples;

    WfArgs X{};
    X.S = P.S;
    X.ctl = ctl;
    X.spill = static_cast<uint2 *>(e->spill.p);
    X.spill_per_thread = spill_per_thread;
    X.cnt = P.cnt;
    X.refill_below = static_cast<uint32_t>(tune.refill_below);
    X.leaf_below = static_cast<uint32_t>(tune.leaf_below);
    X.xp = B.ray_o;
    X.wi = B.ray_d;
    X.eh = static_cast<ClosestHit *>(e->eh.p);
    X.sq = B.sq;
    X.sq_in = B.sq;

    const uint32_t g_item = grid_for(k_bd_shade, 256), g_start = grid_for(k_bd_start, 256),
                   g_conn = grid_for(k_bd_connect, 256), g_contrib = grid_for(k_bd_contrib, 256),
                   g_acc = grid_for(k_bd_accumulate, 256);
    const bool profile = tune.profile != 0;
    if (first_gen) {
        e->shadow_ms = e->extend_ms = 0.0;
        e->trace_launches = 0;
    }
    if (profile && !e->ev[0])
        for (auto &ev : e->ev)
            BD_HIP(hipEventCreate(&ev));
    auto timed = [&](TraceKernel k, uint32_t grid, WfArgs const &A, double &acc) {
        if (profile)
            (void)hipEventRecord(e->ev[0], stream);
        hipLaunchKernelGGL(k, dim3(grid), dim3(wg), 0, stream, A);
        if (profile) {
            (void)hipEventRecord(e->ev[1], stream);
            float ms = 0;
            (void)hipEventSynchronize(e->ev[1]);
            (void)hipEventElapsedTime(&ms, e->ev[0], e->ev[1]);
            acc += ms;
        }
    };
    // walks: at most max_len - 1 camera extensions, then the emission ray and max_len - 1 light
    // extensions
    const uint32_t bounces = 2 * m - 1;
    hipLaunchKernelGGL(k_bd_reset, dim3(1), dim3(64), 0, stream, ctl, qc[0], 1);
    B.eq_out = qs[0];
    B.eq_cnt_out = qc[0];
    hipLaunchKernelGGL(k_bd_start, dim3(g_start), dim3(256), 0, stream, B);
    for (uint32_t k = 0; k < bounces; k++) {
        const uint32_t p = k & 1u, q = p ^ 1u;
        X.eq_in = qs[p];
        X.eq_cnt_in = qc[p];
        timed(tk.extend, g_extend, X, e->extend_ms);
        hipLaunchKernelGGL(k_bd_count, dim3(1), dim3(64), 0, stream, static_cast<const WfCounter *>(qc[p]), P.cnt, 0);
        hipLaunchKernelGGL(k_bd_reset, dim3(1), dim3(64), 0, stream, ctl, qc[q], 0);
        B.eq_in = qs[p];
        B.eq_cnt_in = qc[p];
        B.eq_out = qs[q];
        B.eq_cnt_out = qc[q];
        hipLaunchKernelGGL(k_bd_shade, dim3(g_item), dim3(256), 0, stream, B);
    }
    hipLaunchKernelGGL(k_bd_reset, dim3(1), dim3(64), 0, stream, ctl, static_cast<WfCounter *>(nullptr), 1);
    hipLaunchKernelGGL(k_bd_connect, dim3(g_conn), dim3(256), 0, stream, B);
    timed(tk.bd_shadow, g_shadow, X, e->shadow_ms);
    hipLaunchKernelGGL(k_bd_count, dim3(1), dim3(64), 0, stream, static_cast<const WfCounter *>(&ctl->sq_count), P.cnt, 1);
    hipLaunchKernelGGL(k_bd_contrib, dim3(g_contrib), dim3(256), 0, stream, B);
    hipLaunchKernelGGL(k_bd_accumulate, dim3(g_acc), dim3(256), 0, stream, B);
    if (profile)
        e->trace_launches += bounces + 1;
    BD_HIP(hipGetLastError());
    return "";
#undef BD_HIP
}

namespace {

#define WF_HIP(call)                                                                                 \
    do {                                                                                             \
        hipError_t err_ = (call);                                                                    \
        if (err_ != hipSuccess)                                                                      \
            return std::string(#call) + ": " + hipGetErrorString(err_);                              \
    } while (0)

// A single-job render (M null: P's film, samples [P.s_begin, P.s_end)) or a multi-job render
// (M's generations; P then gives only the seed, depth and counters).
std::string wf_run(WavefrontEngine *e, RenderParams const &P, WavefrontTuning const &tune, hipStream_t stream,
                   JobsPlan const *M) {
    const bool multi = M != nullptr;
    uint32_t npix = P.width * P.height;
    if (!multi && P.s_end <= P.s_begin)
        return "";
    uint32_t gen_samples = 0, cap = 0;
    if (multi) {
        cap = M->cap;
    } else {
        gen_samples = tune.paths_in_flight / npix;
        if (gen_samples < 1)
            gen_samples = 1;
        if (gen_samples > P.s_end - P.s_begin)
            gen_samples = P.s_end - P.s_begin;
        cap = gen_samples * npix;
    }
    size_t f4 = sizeof(float4) * static_cast<size_t>(cap);
    const bool fused = !tune.sort_by_material && !tune.sort_rays;
    WF_HIP(e->xp.ensure(f4));
    if (!fused) {
        // the vertex's normal and outgoing direction (the fused form folds them into the
        // throughput when it samples)
        WF_HIP(e->xn.ensure(f4));
        WF_HIP(e->wo.ensure(f4));
    }
    WF_HIP(e->beta.ensure(f4));
    WF_HIP(e->L.ensure(f4));
    WF_HIP(e->wi.ensure(f4));
    WF_HIP(e->lbuf.ensure(f4));
    WF_HIP(e->rng.ensure(sizeof(uint64_t) * cap));
    if (tune.tracer != kWfMis && !fused)
        return "the lt1 and unidirect tracers run the wavefront form without material or ray sorting only";
    if (multi && (!fused || tune.tracer != kWfMis))
        return "multi-job renders run the MIS tracer's wavefront form without material or ray sorting only";
    if (tune.tracer == kWfLt1)
        WF_HIP(e->le0.ensure(f4));
    if (fused) {
        // path state follows the queues: a second parity of every state array
        WF_HIP(e->xp_b.ensure(f4));
        WF_HIP(e->beta_b.ensure(f4));
        WF_HIP(e->L_b.ensure(f4));
        WF_HIP(e->wi_b.ensure(f4));
        WF_HIP(e->rng_b.ensure(sizeof(uint64_t) * cap));
        WF_HIP(e->lfin_a.ensure(f4));
        WF_HIP(e->lfin_b.ensure(f4));
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

    TraceKernels tk = pick_trace(tune.trace_wg, tune.stack_depth, multi);
    int wg = tk.wg;
    uint32_t g_shadow = grid_for(tk.shadow, wg), g_extend = grid_for(tk.extend, wg);
    uint32_t trace_threads = (g_shadow > g_extend ? g_shadow : g_extend) * wg;
    uint32_t spill_per_thread = 0;
    if (multi) {
        for (WfJob const &J : *M->jobs) {
            const uint32_t sp = spill_entries_for(J.S, tk.ldsd);
            spill_per_thread = sp > spill_per_thread ? sp : spill_per_thread;
        }
    } else {
        spill_per_thread = spill_entries_for(P.S, tk.ldsd);
    }
    // one stack-spill slice per thread of each of the two concurrent traversal kernels
    const size_t spill_half = static_cast<size_t>(trace_threads) * spill_per_thread;
    WF_HIP(e->spill.ensure(2 * sizeof(uint2) * spill_half));
    if (!e->s2) {
        WF_HIP(hipStreamCreateWithFlags(&e->s2, hipStreamNonBlocking));
        WF_HIP(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
        WF_HIP(hipEventCreateWithFlags(&e->join, hipEventDisableTiming));
    }

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
    A.compact = fused ? 1 : 0;
    A.le0 = tune.tracer == kWfLt1 ? static_cast<float4 *>(e->le0.p) : nullptr;
    if (multi) {
        A.jobs = M->d_jobs;
        A.slices = M->d_slices;
        A.gens = M->d_gens;
    }
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
    GenKernel acc_k = multi ? k_wf_accumulate_jobs : k_wf_accumulate;
    if (multi) {
        gen_k = k_wf_generate<true, kWfMis, true>;
        shade_k = k_wf_shade<true, kWfMis, true>;
    } else if (fused) {
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
             g_acc = grid_for(acc_k, 256);
    // The fused form's queue parities: bounce k consumes parity k & 1 and its shade fills the other.
    auto parity_args = [&](uint32_t k) {
        WfArgs B = A;
        uint32_t *eqs[2] = {static_cast<uint32_t *>(e->eq.p), static_cast<uint32_t *>(e->eq_b.p)};
        uint32_t *fins[2] = {static_cast<uint32_t *>(e->fin.p), static_cast<uint32_t *>(e->fin_b.p)};
        WfCounter *eqc[2] = {&A.ctl->eq_count, &A.ctl->eq_count_b};
        WfCounter *finc[2] = {&A.ctl->fin_count, &A.ctl->fin_count_b};
        const uint32_t p = k & 1u, q = p ^ 1u;
        auto f4p = [](DevBuf &d) { return static_cast<float4 *>(d.p); };
        float4 *xps[2] = {f4p(e->xp), f4p(e->xp_b)}, *betas[2] = {f4p(e->beta), f4p(e->beta_b)},
               *Ls[2] = {f4p(e->L), f4p(e->L_b)}, *wis[2] = {f4p(e->wi), f4p(e->wi_b)},
               *lfins[2] = {f4p(e->lfin_a), f4p(e->lfin_b)};
        uint64_t *rngs[2] = {static_cast<uint64_t *>(e->rng.p), static_cast<uint64_t *>(e->rng_b.p)};
        B.xp = xps[p];
        B.xn = B.wo = nullptr; // not part of the compact state
        B.beta = betas[p];
        B.L = Ls[p];
        B.wi = wis[p];
        B.rng = rngs[p];
        B.xp_o = xps[q];
        B.beta_o = betas[q];
        B.L_o = Ls[q];
        B.wi_o = wis[q];
        B.rng_o = rngs[q];
        B.lfin_in = lfins[p];
        B.lfin_out = lfins[q];
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
    uint32_t generations = multi ? M->generations : (P.s_end - P.s_begin + gen_samples - 1) / gen_samples;

    bool profile = tune.profile != 0;
    auto trace_bounce = [&](hipStream_t s, WfArgs const &B) {
        WfArgs Bs = B; // the shadow kernel's own stack-spill slices
        Bs.spill = B.spill + spill_half;
        if (!profile) {
            // independent traversals: each fills the other's tail (lanes whose queue ran dry)
            (void)hipEventRecord(e->fork, s);
            (void)hipStreamWaitEvent(e->s2, e->fork, 0);
            hipLaunchKernelGGL(tk.extend, dim3(g_extend), dim3(wg), 0, s, B);
            hipLaunchKernelGGL(tk.shadow, dim3(g_shadow), dim3(wg), 0, e->s2, Bs);
            (void)hipEventRecord(e->join, e->s2);
            (void)hipStreamWaitEvent(s, e->join, 0);
            hipLaunchKernelGGL(k_wf_count, dim3(1), dim3(64), 0, s, B.ctl, static_cast<const WfCounter *>(B.eq_cnt_in),
                               B.cnt);
            return;
        }
        // profiled: one after the other, each timed with events
        (void)hipEventRecord(e->ev[0], s);
        hipLaunchKernelGGL(tk.shadow, dim3(g_shadow), dim3(wg), 0, s, Bs);
        (void)hipEventRecord(e->ev[1], s);
        hipLaunchKernelGGL(tk.extend, dim3(g_extend), dim3(wg), 0, s, B);
        {
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
    // A generation: head (reset, generate), bounces [k0, k1), tail (accumulate, next generation).
    auto record_head = [&](hipStream_t s) {
        hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, s, A.ctl, -1);
        if (fused) {
            // generate fills bounce 0's queues (parity 0); bounce k's shade fills bounce k + 1's
            WfArgs G = parity_args(1);
            hipLaunchKernelGGL(gen_k, dim3(g_gen), dim3(256), 0, s, G);
        } else {
            hipLaunchKernelGGL(gen_k, dim3(g_gen), dim3(256), 0, s, A);
        }
    };
    auto record_bounces = [&](hipStream_t s, uint32_t k0, uint32_t k1) {
        if (fused) {
            for (uint32_t k = k0; k < k1; k++) {
                WfArgs B = parity_args(k);
                trace_bounce(s, B);
                hipLaunchKernelGGL(k_wf_reset_fused, dim3(1), dim3(64), 0, s, B.ctl, B.eq_cnt_out, B.fin_cnt_out);
                hipLaunchKernelGGL(shade_k, dim3(g_shade), dim3(256), 0, s, B, 0, k);
            }
        } else {
            for (uint32_t k = k0; k < k1; k++) {
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
    };
    auto record_tail = [&](hipStream_t s) {
        hipLaunchKernelGGL(acc_k, dim3(g_acc), dim3(256), 0, s, A);
        hipLaunchKernelGGL(k_wf_next_gen, dim3(1), dim3(64), 0, s, A.ctl);
    };
    auto record_generation = [&](hipStream_t s) {
        record_head(s);
        record_bounces(s, 0, bounces);
        record_tail(s);
    };
    // The reference-mode tracers' up to kRefModeMaxDepth bounces (Russian roulette ends nearly
    // every path within about 30): their graph is cut into segments of kSegBounces bounces, and
    // after each segment the host reads the queue sizes the next bounce consumes; once both are
    // empty the remaining segments are not launched (one segment runs ahead of the check, so the
    // GPU never waits for the host).
    constexpr uint32_t kSegBounces = 4;
    const bool segmented = fused && bounces > 2 * kSegBounces;
    const uint32_t nseg = segmented ? (bounces + kSegBounces - 1) / kSegBounces : 0;
    // multi-job renders: the end of each generation (job latencies)
    auto generation_done = [&](uint32_t g) -> hipError_t {
        return multi ? hipEventRecord(e->gen_ev[g], stream) : hipSuccess;
    };

    e->shadow_ms = e->extend_ms = 0.0;
    e->trace_launches = 0;
    if (profile && !e->ev[0])
        for (auto &ev : e->ev)
            WF_HIP(hipEventCreate(&ev));
    hipLaunchKernelGGL(k_wf_begin_render, dim3(1), dim3(64), 0, stream, A.ctl);
    WF_HIP(hipGetLastError());
    if (tune.use_graph && !profile) {
        // one graph per generation, or (segmented) head, segments and tail
        const uint32_t ngraphs = segmented ? nseg + 2 : 1;
        if (!e->have_graph || !same_config(A, e->key, tune, e->key_tune) || e->execs.size() != ngraphs) {
            for (hipGraphExec_t x : e->execs)
                WF_HIP(hipGraphExecDestroy(x));
            for (hipGraph_t x : e->graphs)
                WF_HIP(hipGraphDestroy(x));
            e->execs.clear();
            e->graphs.clear();
            e->have_graph = false;
            for (uint32_t gi = 0; gi < ngraphs; gi++) {
                hipGraph_t gr = nullptr;
                hipGraphExec_t ex = nullptr;
                WF_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                if (!segmented)
                    record_generation(stream);
                else if (gi == 0)
                    record_head(stream);
                else if (gi + 1 == ngraphs)
                    record_tail(stream);
                else
                    record_bounces(stream, (gi - 1) * kSegBounces, std::min(gi * kSegBounces, bounces));
                WF_HIP(hipStreamEndCapture(stream, &gr));
                e->graphs.push_back(gr);
                WF_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
                e->execs.push_back(ex);
            }
            std::memcpy(&e->key, &A, sizeof A);
            e->key_tune = tune;
            e->have_graph = true;
        }
        if (segmented) {
            if (!e->pinned)
                WF_HIP(hipHostMalloc(reinterpret_cast<void **>(&e->pinned), 2 * sizeof(uint32_t) * nseg));
            else if (e->pinned_n < nseg) {
                WF_HIP(hipHostFree(e->pinned));
                e->pinned = nullptr;
                WF_HIP(hipHostMalloc(reinterpret_cast<void **>(&e->pinned), 2 * sizeof(uint32_t) * nseg));
            }
            e->pinned_n = std::max(e->pinned_n, nseg);
            while (e->seg_ev.size() < nseg) {
                hipEvent_t ev;
                WF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
                e->seg_ev.push_back(ev);
            }
        }
        for (uint32_t g = 0; g < generations; g++) {
            if (!segmented) {
                WF_HIP(hipGraphLaunch(e->execs[0], stream));
            } else {
                WF_HIP(hipGraphLaunch(e->execs[0], stream));
                for (uint32_t i = 0; i < nseg; i++) {
                    WF_HIP(hipGraphLaunch(e->execs[1 + i], stream));
                    // the queues the bounce after this segment consumes (parity kend & 1)
                    const uint32_t kend = std::min((i + 1) * kSegBounces, bounces);
                    WfCounter *eqc = (kend & 1u) ? &A.ctl->eq_count_b : &A.ctl->eq_count;
                    WfCounter *finc = (kend & 1u) ? &A.ctl->fin_count_b : &A.ctl->fin_count;
                    WF_HIP(hipMemcpyAsync(e->pinned + 2 * i, &eqc->v, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
                    WF_HIP(hipMemcpyAsync(e->pinned + 2 * i + 1, &finc->v, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                          stream));
                    WF_HIP(hipEventRecord(e->seg_ev[i], stream));
                    if (i > 0) {
                        WF_HIP(hipEventSynchronize(e->seg_ev[i - 1]));
                        if (e->pinned[2 * (i - 1)] == 0 && e->pinned[2 * (i - 1) + 1] == 0)
                            break; // every path of the generation has ended
                    }
                }
                WF_HIP(hipGraphLaunch(e->execs[ngraphs - 1], stream));
            }
            WF_HIP(generation_done(g));
        }
    } else {
        for (uint32_t g = 0; g < generations; g++) {
            record_generation(stream);
            WF_HIP(generation_done(g));
        }
    }
    WF_HIP(hipGetLastError());
    return "";
}

} // namespace

std::string wf_render(WavefrontEngine *e, RenderParams const &P, WavefrontTuning const &tune, hipStream_t stream) {
    return wf_run(e, P, tune, stream, nullptr);
}

std::string wf_render_jobs(WavefrontEngine *e, std::vector<WfJob> const &jobs, uint32_t max_depth, uint64_t seed,
                           Counters *cnt, WavefrontTuning const &tune, hipStream_t stream, hipEvent_t start,
                           WfJobsResult *res) {
    const uint32_t nj = static_cast<uint32_t>(jobs.size());
    if (res) {
        res->generations = 0;
        res->job_ms.assign(nj, 0.0);
    }
    if (nj == 0)
        return "";
    // Pack whole samples of the jobs' pixels into generations of at most `cap` paths, fewest
    // samples first (small jobs finish in the first generation); a job that does not fit the rest
    // of a generation continues in the next one, its samples in order.
    uint64_t cap_paths = tune.paths_in_flight ? tune.paths_in_flight : kDefaultQueueCapacity;
    for (WfJob const &J : jobs) {
        if (J.npix == 0 || J.spp == 0 || J.npix != J.width * J.height)
            return "multi-job render: a job with an empty film or no samples";
        cap_paths = J.npix > cap_paths ? J.npix : cap_paths; // one sample of every film fits
    }
    if (cap_paths >= kWfFinBit)
        return "multi-job render: a film of 2^31 pixels or more";
    std::vector<uint32_t> order(nj);
    for (uint32_t j = 0; j < nj; j++)
        order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return static_cast<uint64_t>(jobs[a].npix) * jobs[a].spp < static_cast<uint64_t>(jobs[b].npix) * jobs[b].spp;
    });
    std::vector<WfSlice> slices;
    std::vector<WfGen> gens;
    std::vector<uint32_t> last_gen(nj, 0);
    WfGen cur{0, 0, 0, 0};
    auto close = [&]() {
        cur.slice_end = static_cast<uint32_t>(slices.size());
        gens.push_back(cur);
        cur = WfGen{cur.slice_end, cur.slice_end, 0, 0};
    };
    for (uint32_t j : order) {
        WfJob const &J = jobs[j];
        uint32_t s = 0;
        while (s < J.spp) {
            const uint64_t fit = (cap_paths - cur.paths) / J.npix;
            if (fit == 0) {
                close();
                continue;
            }
            const uint32_t c = static_cast<uint32_t>(fit < J.spp - s ? fit : J.spp - s);
            WfSlice sl{};
            sl.job = j;
            sl.s_first = s;
            sl.count = c;
            sl.path_begin = cur.paths;
            sl.pix_begin = cur.pixels;
            slices.push_back(sl);
            cur.paths += c * J.npix;
            cur.pixels += J.npix;
            s += c;
            last_gen[j] = static_cast<uint32_t>(gens.size());
        }
    }
    if (cur.paths > 0)
        close();
    uint32_t cap = 0;
    for (WfGen const &G : gens)
        cap = G.paths > cap ? G.paths : cap;
    const uint32_t ng = static_cast<uint32_t>(gens.size());
    WF_HIP(e->mj_jobs.ensure(sizeof(WfJob) * nj));
    WF_HIP(e->mj_slices.ensure(sizeof(WfSlice) * slices.size()));
    WF_HIP(e->mj_gens.ensure(sizeof(WfGen) * ng));
    // synchronous copies of small host tables: the previous render has finished (this function
    // waits for its last generation) and the kernels below are launched after them
    WF_HIP(hipMemcpy(e->mj_jobs.p, jobs.data(), sizeof(WfJob) * nj, hipMemcpyHostToDevice));
    WF_HIP(hipMemcpy(e->mj_slices.p, slices.data(), sizeof(WfSlice) * slices.size(), hipMemcpyHostToDevice));
    WF_HIP(hipMemcpy(e->mj_gens.p, gens.data(), sizeof(WfGen) * ng, hipMemcpyHostToDevice));
    while (e->gen_ev.size() < ng) {
        hipEvent_t ev;
        WF_HIP(hipEventCreate(&ev));
        e->gen_ev.push_back(ev);
    }
    JobsPlan M{&jobs, static_cast<const WfJob *>(e->mj_jobs.p), static_cast<const WfSlice *>(e->mj_slices.p),
               static_cast<const WfGen *>(e->mj_gens.p), ng, cap};
    RenderParams P{};
    P.seed = seed;
    P.max_depth = max_depth;
    P.cnt = cnt;
    WavefrontTuning t = tune;
    t.profile = 0;
    std::string err = wf_run(e, P, t, stream, &M);
    if (!err.empty())
        return err;
    WF_HIP(hipEventSynchronize(e->gen_ev[ng - 1]));
    if (res) {
        res->generations = ng;
        std::vector<double> gen_ms(ng, 0.0);
        for (uint32_t g = 0; g < ng; g++) {
            float ms = 0;
            WF_HIP(hipEventElapsedTime(&ms, start, e->gen_ev[g]));
            gen_ms[g] = ms;
        }
        for (uint32_t j = 0; j < nj; j++)
            res->job_ms[j] = gen_ms[last_gen[j]];
    }
    return "";
}
#undef WF_HIP

} // namespace x3d
