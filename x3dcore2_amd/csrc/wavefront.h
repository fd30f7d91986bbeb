// Wavefront form of the MIS tracer and of the reference-mode lt1 / unidirect tracers (host interface).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "render_common.h"

namespace x3d {

struct WavefrontEngine;

constexpr uint32_t kDefaultQueueCapacity = 1u << 26; // 64M paths: ~17 GiB of path state and queues
constexpr uint32_t kMaxQueueCapacity = 1u << 26;

// Estimator a wavefront render reproduces (bit for bit, against its plain kernel): the MIS tracer,
// unidirect_lt1_path_tracer or unidirect_path_tracer.
enum WfTracer : int { kWfMis = 0, kWfLt1 = 1, kWfUnidirect = 2 };

struct WavefrontTuning {
    int tracer = kWfMis;
    uint32_t paths_in_flight = kDefaultQueueCapacity; // queue capacity (paths of one generation)
    int trace_wg = 64;                   // workgroup size of the traversal kernels (64, 128, 256)
    int stack_depth = 16;                // LDS stack entries per lane of the traversal kernels (8, 16)
    int sort_by_material = 0;            // group active paths by material kind before shading
    int sort_rays = 0;                   // reorder both ray queues by origin cell and direction
    int refill_below = 8;                // traversal kernels leave the descent phase (to test
                                         // leaves and refill finished lanes) once fewer than this
                                         // many lanes of the wave still look for a leaf
    int leaf_below = 8;                  // ... and the leaf phase once fewer than this many
                                         // lanes still hold a leaf
    int use_graph = 1;                   // replay one generation's launches as a hipGraph
    int profile = 0;                     // time the traversal kernels with events (no graph)
};

WavefrontEngine *wf_create();
void wf_destroy(WavefrontEngine *e);

// Event-timed durations of the traversal kernels of the last profiled render.
void wf_last_timing(WavefrontEngine const *e, double *shadow_ms, double *extend_ms, uint32_t *launches);

// Runs the sample part of a render whose first-hit pass has already filled P.first:
// samples [P.s_begin, P.s_end) of every pixel, in generations of `paths_in_flight / pixels`
// samples; writes P.out (P.finalize) or P.acc. Returns an empty string or an error message.
std::string wf_render(WavefrontEngine *e, RenderParams const &P, WavefrontTuning const &tune, hipStream_t stream);

// One job of a multi-job render (a batch whose jobs share every generation): its scene, camera,
// first-hit cache (3 float4 per pixel, filled beforehand by the first-hit pass), running sums
// kept between the job's generations, and film.
struct WfJob {
    DevScene S;
    CamParams cam;
    float4 *first, *acc, *out;
    uint32_t width, height, npix, spp;
};

struct WfJobsResult {
    uint32_t generations = 0;
    std::vector<double> job_ms; // per job: device time from `start` to the end of its last generation
};

// Renders every job with the fixed-depth MIS tracer (max_depth bounces, PCG streams of `seed`)
// in shared generations of up to tune.paths_in_flight paths: each generation holds whole samples
// of several jobs' pixels, a job's samples run in order and may span generations, and the paths
// of all its jobs share the queues and kernels of every bounce. Each film is bit-identical to
// wf_render of that job alone. The fused form only (no material or ray sort). Returns after the
// last generation has finished (res->job_ms is measured from the event `start`, recorded
// earlier on `stream`), with an empty string or an error message. `cnt` (device) receives the
// ray-query counters.
std::string wf_render_jobs(WavefrontEngine *e, std::vector<WfJob> const &jobs, uint32_t max_depth, uint64_t seed,
                           Counters *cnt, WavefrontTuning const &tune, hipStream_t stream, hipEvent_t start,
                           WfJobsResult *res);

// Wavefront form of the bidirectional tracer. Working memory per (pixel, sample) item of a
// generation for subpaths of up to max_len vertices (bd_item_bytes), and the largest generation
// (items) that fits the form's budget.
constexpr uint64_t kBdBudgetBytes = 24ull << 30;
uint64_t bd_item_bytes(uint32_t max_len);
uint32_t bd_max_items(uint32_t max_len);
// One generation: samples [s_first, s_first + gen_samples) of every pixel of a render whose
// first-hit pass has filled P.first; sample radiance of item j (pixel j % npix, sample
// s_first + j / npix) to samples[j], bit-identical to k_render_bdpt's. first_gen restarts the
// profiled traversal timings. Returns an empty string or an error message.
std::string bd_generation(WavefrontEngine *e, RenderParams const &P, WavefrontTuning const &tune, uint32_t s_first,
                          uint32_t gen_samples, float4 *samples, bool first_gen, hipStream_t stream);

} // namespace x3d
