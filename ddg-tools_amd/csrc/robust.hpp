// The dense stage of the step's recovery chain (RS/src/step/direction.rs:259-338).
#pragma once
#include "gmres.hpp"

namespace ddgrs {

// The largest vertex count of the dense stage (RS/src/step/direction.rs:238
// DENSE_FALLBACK_MAX_VERTICES).
constexpr int64_t kDenseFallbackMaxVertices = 6000;

// x = A′⁻¹ b (3 nv + 3, device pointers) for the saddle system with the exact metric
// w_hi B + w_lo B⁰ on the face tree t / c of the surface's current positions and the barycenter rows
// `crow` (device, nv), by a dense LU on the device; returns ‖b − A′ x‖ / ‖b‖.
double dense_saddle_solve(ddgrs_ctx* ctx, ddgrs_surface* s, const Topology& t, const Clusters& c, double s_exp,
                          double w_hi, double w_lo, const double* crow, const double* b, double* x);

}  // namespace ddgrs
