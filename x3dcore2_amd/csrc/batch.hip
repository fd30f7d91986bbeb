// Batched rendering: many independent jobs (scene, camera, film size, spp) share one persistent
// launch of the fixed-depth MIS tracer.
//
// Work is cut into tiles of 64 consecutive pixels of one job. Each wave takes a whole tile from a
// global counter, so the scene of a tile is uniform across the wave (its descriptor is read with
// the tile index made wave-uniform). The wave first traces the tile's 64 camera rays (one per
// lane) and keeps the first hits in LDS; then its lanes work through the tile's (pixel, sample)
// items in chunks of up to kChunk samples per pixel, a lane whose path ends taking the next item
// at once (per-bounce regeneration, no waiting for the slowest pixel); each item's radiance goes
// to the wave's slice of a sample buffer, and lane l then adds pixel l's samples of the chunk in
// sample order. Tiles are handed out most expensive job first, so big jobs do not form the tail.
// Every pixel performs the single-job render's operations in the same order (first hit, samples
// summed in order), so every film is bit-identical to rendering its job alone.

#include "batch.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace x3d {

constexpr uint32_t kTilePx = kBatchTilePx;
constexpr uint32_t kChunk = 64;  // samples per pixel per pass over a tile (sample buffer depth)

struct BatchCtl {
    uint32_t next_tile;
    uint32_t pad;
    unsigned long long start_clock;
};

struct BatchArgs {
    const BatchJob *jobs;
    const uint32_t *tile_job; // job of each tile, in hand-out order
    uint32_t num_jobs, num_tiles, max_depth;
    uint64_t seed;
    uint2 *spill;
    uint32_t spill_per_thread;
    float4 *samples; // kChunk x kTilePx float4 per resident wave
    Counters *cnt;
    BatchCtl *ctl;
    uint32_t *tiles_left;            // per job
    unsigned long long *done_clock;  // per job: device wall clock when its last tile finished
};

__global__ void k_batch_init(BatchArgs A) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        A.ctl->next_tile = 0;
        A.ctl->start_clock = ~0ull;
        Counters z{};
        *A.cnt = z;
    }
    for (; i < A.num_jobs; i += gridDim.x * blockDim.x) {
        A.tiles_left[i] = A.jobs[i].num_tiles;
        A.done_clock[i] = 0;
    }
}

// LDS cells written by one lane and read by others later: relaxed atomics keep the compiler from
// answering the reads from registers.
X3D_DI uint32_t lds_load(uint32_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT); }
X3D_DI void lds_store(uint32_t *p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT); }

template <int WG, int LDSD> __global__ __launch_bounds__(WG) void k_render_batch(BatchArgs A) {
    constexpr int kWaves = WG / 64;
    __shared__ uint2 lds[LDSD * WG];
    // per wave: first hits of the tile's pixels (p.xyz, t | n.xyz, mat | light, geo, valid) and
    // the item cursor
    constexpr uint32_t kHitWords = 11;
    __shared__ uint32_t hits[kWaves][kTilePx * kHitWords];
    __shared__ uint32_t cursor[kWaves];
    TravStack<WG, LDSD> st = make_stack<WG, LDSD>(lds, A.spill, A.spill_per_thread);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t *hit_lds = hits[wave];
    uint32_t *cur = &cursor[wave];
    float4 *buf = A.samples + static_cast<size_t>(blockIdx.x * kWaves + wave) * kChunk * kTilePx;
    if (lane == 0)
        atomicMin(&A.ctl->start_clock, static_cast<unsigned long long>(wall_clock64()));
    uint32_t nclosest = 0, nany = 0, pops = 0, tests = 0, spops = 0, stests = 0;
    for (;;) {
        uint32_t t = 0;
        if (lane == 0)
            t = atomicAdd(&A.ctl->next_tile, 1u);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
        if (t >= A.num_tiles)
            break;
        const uint32_t j = __builtin_amdgcn_readfirstlane(A.tile_job[t]);
        const BatchJob &J = A.jobs[j];
        const DevScene S = J.S;
        Tracer<WG, LDSD> tr{S, st};
        const uint32_t w = J.width, h = J.height, spp = J.spp;
        const uint32_t p0 = (t - J.first_tile) * kTilePx;
        const uint32_t npx = min(kTilePx, w * h - p0);
        const float scale = 1.0f / static_cast<float>(spp);
        // first hits, one pixel per lane; compute_first_hit drops hits seen from behind
        // (pathtracer.cpp:463-464)
        if (lane < npx) {
            uint32_t pix = p0 + lane;
            f3 o, v;
            camera_ray(J.cam, pix % w, pix / w, w, h, o, v);
            HitRec fh{};
            bool valid = tr.trace(o, v, fh) && !(dot(fh.n, -v) <= 0.0f);
            uint32_t *c = hit_lds + kHitWords * lane;
            lds_store(c + 0, __float_as_uint(fh.p.x));
            lds_store(c + 1, __float_as_uint(fh.p.y));
            lds_store(c + 2, __float_as_uint(fh.p.z));
            lds_store(c + 3, __float_as_uint(fh.t));
            lds_store(c + 4, __float_as_uint(fh.n.x));
            lds_store(c + 5, __float_as_uint(fh.n.y));
            lds_store(c + 6, __float_as_uint(fh.n.z));
            lds_store(c + 7, static_cast<uint32_t>(fh.mat));
            lds_store(c + 8, static_cast<uint32_t>(fh.light));
            lds_store(c + 9, static_cast<uint32_t>(fh.geo));
            lds_store(c + 10, valid ? 1u : 0u);
        }
        f3 acc = mk(0.0f, 0.0f, 0.0f); // lane l: running sum of pixel p0 + l
        for (uint32_t s0 = 0; s0 < spp; s0 += kChunk) {
            const uint32_t cs = min(kChunk, spp - s0);
            const uint32_t items = npx * cs; // item q: pixel slot q % npx, sample s0 + q / npx
            if (lane == 0)
                lds_store(cur, 0u);
            bool active = false;
            uint32_t q = 0, slot = 0;
            PathState ps;
            ps.alive = false;
            for (;;) {
                // lanes without an item take the next ones of the tile (rank order)
                const uint64_t need = __ballot(!active);
                const uint32_t base = lds_load(cur);
                if (!active) {
                    uint32_t cand = base + static_cast<uint32_t>(__popcll(need & below));
                    if (cand < items) {
                        q = cand;
                        slot = q % npx;
                        uint32_t s = s0 + q / npx, pix = p0 + slot;
                        f3 o, v;
                        camera_ray(J.cam, pix % w, pix / w, w, h, o, v);
                        uint32_t *c = hit_lds + kHitWords * slot;
                        HitRec first;
                        first.p = mk(__uint_as_float(lds_load(c + 0)), __uint_as_float(lds_load(c + 1)),
                                     __uint_as_float(lds_load(c + 2)));
                        first.t = __uint_as_float(lds_load(c + 3));
                        first.n = mk(__uint_as_float(lds_load(c + 4)), __uint_as_float(lds_load(c + 5)),
                                     __uint_as_float(lds_load(c + 6)));
                        first.mat = static_cast<int32_t>(lds_load(c + 7));
                        first.light = static_cast<int32_t>(lds_load(c + 8));
                        first.geo = static_cast<int32_t>(lds_load(c + 9));
                        bool valid = lds_load(c + 10) != 0u;
                        active = true;
                        tr.mis_begin(ps, first, valid, v, pcg_for(A.seed, pix, s), A.max_depth);
                    }
                }
                if (lane == 0)
                    lds_store(cur, base + static_cast<uint32_t>(__popcll(need)));
                if (__ballot(active) == 0)
                    break;
                if (active) {
                    if (ps.alive)
                        tr.mis_bounce(ps, A.max_depth);
                    if (!ps.alive) {
                        buf[(q / npx) * kTilePx + slot] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
                        active = false;
                    }
                }
            }
            // the chunk's samples, written by any lane, are read back by the lane of their pixel
            __threadfence();
            if (lane < npx)
                for (uint32_t k = 0; k < cs; k++)
                    acc = acc + xyz(buf[k * kTilePx + lane]);
            __threadfence();
        }
        if (lane < npx)
            J.out[p0 + lane] = make_float4(scale * acc.x, scale * acc.y, scale * acc.z, scale);
        nclosest += tr.nclosest;
        nany += tr.nany;
        pops += tr.pops;
        tests += tr.tests;
        spops += tr.spops;
        stests += tr.stests;
        if (lane == 0 && atomicSub(&A.tiles_left[j], 1u) == 1u)
            A.done_clock[j] = static_cast<unsigned long long>(wall_clock64());
    }
    flush_counter(&A.cnt->closest, nclosest);
    flush_counter(&A.cnt->anyhit, nany);
    flush_counter(&A.cnt->pops, pops);
    flush_counter(&A.cnt->tests, tests);
    flush_counter(&A.cnt->spops, spops);
    flush_counter(&A.cnt->stests, stests);
}

// ---------------------------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------------------------
struct BatchEngine {
    DevBuf jobs, tile_job, ctl, tiles_left, done_clock, spill, samples;
    std::vector<BatchJob> h_jobs; // host copies stay alive while the uploads are in flight
    std::vector<uint32_t> h_tile_job;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned char key[sizeof(BatchArgs) + 16] = {};
    bool have_graph = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~BatchEngine() {
        for (auto &x : ev)
            if (x)
                (void)hipEventDestroy(x);
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
    }
};

BatchEngine *batch_create() { return new BatchEngine(); }
void batch_destroy(BatchEngine *e) { delete e; }

namespace {

using BatchKernel = void (*)(BatchArgs);

struct BatchPick {
    BatchKernel k;
    int wg, ldsd;
};

BatchPick pick_batch(int wg, int ldsd) {
    bool shallow = ldsd == 8;
    switch (wg) {
    case 64: return shallow ? BatchPick{k_render_batch<64, 8>, 64, 8} : BatchPick{k_render_batch<64, 16>, 64, 16};
    case 128: return shallow ? BatchPick{k_render_batch<128, 8>, 128, 8} : BatchPick{k_render_batch<128, 16>, 128, 16};
    default: return shallow ? BatchPick{k_render_batch<256, 8>, 256, 8} : BatchPick{k_render_batch<256, 16>, 256, 16};
    }
}

} // namespace

std::string batch_render(BatchEngine *e, std::vector<BatchJob> jobs, uint32_t max_depth, uint64_t seed,
                         BatchTuning const &tune, Counters *cnt, hipStream_t stream, BatchResult *res) {
#define BT_HIP(call)                                                                                 \
    do {                                                                                             \
        hipError_t err_ = (call);                                                                    \
        if (err_ != hipSuccess)                                                                      \
            return std::string(#call) + ": " + hipGetErrorString(err_);                              \
    } while (0)
    const uint32_t n = static_cast<uint32_t>(jobs.size());
    if (n == 0)
        return "";
    const uint32_t tile_px = kTilePx;
    // hand-out order: most expensive first (samples x log2 of the triangle count)
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; i++)
        order[i] = i;
    if (tune.sort_by_cost) {
        auto cost = [&](uint32_t i) {
            BatchJob const &J = jobs[i];
            return static_cast<double>(J.width) * J.height * J.spp * std::log2(2.0 + J.S.num_prims);
        };
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost(a) > cost(b); });
    }
    uint64_t tiles = 0;
    uint32_t spill_per_thread = 1;
    for (uint32_t i : order) {
        BatchJob &J = jobs[i];
        uint64_t npix = static_cast<uint64_t>(J.width) * J.height;
        J.first_tile = static_cast<uint32_t>(tiles);
        J.num_tiles = static_cast<uint32_t>((npix + tile_px - 1) / tile_px);
        tiles += J.num_tiles;
        spill_per_thread = std::max(spill_per_thread, spill_entries_for(J.S, tune.stack_depth));
    }
    if (tiles >= (1ull << 31))
        return "batch has too many tiles";
    e->h_tile_job.resize(tiles);
    for (uint32_t i = 0; i < n; i++)
        std::fill(e->h_tile_job.begin() + jobs[i].first_tile,
                  e->h_tile_job.begin() + jobs[i].first_tile + jobs[i].num_tiles, i);

    BatchPick pk = pick_batch(tune.wg, tune.stack_depth);
    int dev = 0, cus = 0, per_cu = 0;
    BT_HIP(hipGetDevice(&dev));
    BT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pk.k, pk.wg, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    uint64_t waves = static_cast<uint64_t>(cus) * per_cu * (pk.wg / 64);
    uint32_t grid = static_cast<uint32_t>(cus * per_cu);
    if (waves > tiles) // no more waves than tiles
        grid = static_cast<uint32_t>(std::max<uint64_t>(1, (tiles * 64 + pk.wg - 1) / pk.wg));

    BT_HIP(e->jobs.ensure(sizeof(BatchJob) * n));
    BT_HIP(e->tile_job.ensure(sizeof(uint32_t) * tiles));
    BT_HIP(e->ctl.ensure(sizeof(BatchCtl)));
    BT_HIP(e->tiles_left.ensure(sizeof(uint32_t) * n));
    BT_HIP(e->done_clock.ensure(sizeof(unsigned long long) * n));
    BT_HIP(e->spill.ensure(sizeof(uint2) * static_cast<size_t>(grid) * pk.wg * spill_per_thread));
    BT_HIP(e->samples.ensure(sizeof(float4) * static_cast<size_t>(grid) * (pk.wg / 64) * kChunk * kTilePx));
    e->h_jobs = jobs;
    BT_HIP(hipMemcpyAsync(e->jobs.p, e->h_jobs.data(), sizeof(BatchJob) * n, hipMemcpyHostToDevice, stream));
    BT_HIP(hipMemcpyAsync(e->tile_job.p, e->h_tile_job.data(), sizeof(uint32_t) * tiles, hipMemcpyHostToDevice,
                          stream));

    BatchArgs A{};
    A.jobs = static_cast<const BatchJob *>(e->jobs.p);
    A.tile_job = static_cast<const uint32_t *>(e->tile_job.p);
    A.num_jobs = n;
    A.num_tiles = static_cast<uint32_t>(tiles);
    A.max_depth = max_depth;
    A.seed = seed;
    A.spill = static_cast<uint2 *>(e->spill.p);
    A.spill_per_thread = spill_per_thread;
    A.samples = static_cast<float4 *>(e->samples.p);
    A.cnt = cnt;
    A.ctl = static_cast<BatchCtl *>(e->ctl.p);
    A.tiles_left = static_cast<uint32_t *>(e->tiles_left.p);
    A.done_clock = static_cast<unsigned long long *>(e->done_clock.p);

    uint32_t g_init = std::min<uint32_t>(64, (n + 255) / 256 + 1);
    auto record = [&](hipStream_t s) {
        hipLaunchKernelGGL(k_batch_init, dim3(g_init), dim3(256), 0, s, A);
        hipLaunchKernelGGL(pk.k, dim3(grid), dim3(pk.wg), 0, s, A);
    };
    if (!e->ev[0])
        for (auto &x : e->ev)
            BT_HIP(hipEventCreate(&x));
    unsigned char key[sizeof e->key] = {};
    std::memcpy(key, &A, sizeof A);
    std::memcpy(key + sizeof A, &grid, sizeof grid);
    std::memcpy(key + sizeof A + 4, &pk.wg, sizeof pk.wg);
    std::memcpy(key + sizeof A + 8, &pk.ldsd, sizeof pk.ldsd);
    if (tune.use_graph && (!e->have_graph || std::memcmp(key, e->key, sizeof key) != 0)) {
        if (e->exec)
            BT_HIP(hipGraphExecDestroy(e->exec));
        if (e->graph)
            BT_HIP(hipGraphDestroy(e->graph));
        e->exec = nullptr;
        e->graph = nullptr;
        e->have_graph = false;
        BT_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        record(stream);
        BT_HIP(hipStreamEndCapture(stream, &e->graph));
        BT_HIP(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
        std::memcpy(e->key, key, sizeof key);
        e->have_graph = true;
    }
    BT_HIP(hipEventRecord(e->ev[0], stream));
    if (tune.use_graph)
        BT_HIP(hipGraphLaunch(e->exec, stream));
    else
        record(stream);
    BT_HIP(hipGetLastError());
    BT_HIP(hipEventRecord(e->ev[1], stream));
    if (res) {
        BT_HIP(hipEventSynchronize(e->ev[1]));
        float ms = 0;
        BT_HIP(hipEventElapsedTime(&ms, e->ev[0], e->ev[1]));
        res->kernel_ms = ms;
        std::vector<unsigned long long> done(n);
        BatchCtl ctl{};
        BT_HIP(hipMemcpy(done.data(), e->done_clock.p, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost));
        BT_HIP(hipMemcpy(&ctl, e->ctl.p, sizeof ctl, hipMemcpyDeviceToHost));
        int rate_khz = 0;
        BT_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
        res->job_ms.assign(n, 0.0);
        for (uint32_t i = 0; i < n; i++)
            if (rate_khz > 0 && done[i] >= ctl.start_clock)
                res->job_ms[i] = static_cast<double>(done[i] - ctl.start_clock) / rate_khz;
    }
    return "";
#undef BT_HIP
}

} // namespace x3d
