// Batched rendering of many independent jobs (scene, camera, film size, samples per pixel) in
// one persistent launch (host interface). No counterpart in the reference, which renders one
// scene per call (renderer.cpp:96-166).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "render_common.h"

namespace x3d {

struct BatchJob {
    DevScene S;
    CamParams cam;
    float4 *out; // film: rgba per pixel, row-major, row 0 at the top
    uint32_t width, height, spp;
    uint32_t first_tile, num_tiles; // filled by batch_render
    uint32_t pad;
};

struct BatchTuning {
    int wg = 256;          // workgroup size (64, 128, 256)
    int stack_depth = 16;  // LDS stack entries per lane (8, 16)
    int sort_by_cost = 1;  // hand out the tiles of the most expensive jobs first
    int use_graph = 1;     // replay the launch as a hipGraph when the batch is unchanged
};

constexpr uint32_t kBatchTilePx = 64; // pixels per work tile (one per lane of a wave)

struct BatchEngine;

BatchEngine *batch_create();
void batch_destroy(BatchEngine *e);

struct BatchResult {
    double kernel_ms = 0;        // event-timed duration of the batch launch
    std::vector<double> job_ms;  // per job: device time from the launch's start to its last tile
};

// Renders every job with the fixed-depth MIS tracer (max_depth bounces, PCG streams of `seed`);
// each film is bit-identical to the single-job render of the same job. Returns an empty string
// or an error message. `cnt` (device) receives the ray-query counters.
std::string batch_render(BatchEngine *e, std::vector<BatchJob> jobs, uint32_t max_depth, uint64_t seed,
                         BatchTuning const &tune, Counters *cnt, hipStream_t stream, BatchResult *res);

} // namespace x3d
