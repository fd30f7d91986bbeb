// Path-replay adjoint of the fixed-depth MIS tracer with respect to material parameters (host
// interface). No counterpart in the reference, which has no differentiable renderer.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "render_common.h"

namespace x3d {

constexpr uint32_t kAdjMaxMats = 64; // materials with gradient slots (4 floats each)

struct AdjointEngine;

AdjointEngine *adjoint_create();
void adjoint_destroy(AdjointEngine *e);

// Gradient of a film loss, given dLoss/dfilm (d_dl: float4 per pixel, rgb used), with respect to
// (albedo rgb, roughness) of materials [0, num_mats): d_grad receives float[4 * num_mats].
// P describes the render (first-hit buffer already filled, samples [0, spp_total)).
// The sums are reproducible: static work assignment and fixed-order reductions.
std::string adjoint_render(AdjointEngine *e, RenderParams const &P, const float4 *d_dl, uint32_t num_mats,
                           float *d_grad, hipStream_t stream);

} // namespace x3d
