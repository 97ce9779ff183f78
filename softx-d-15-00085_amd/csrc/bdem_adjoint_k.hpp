/*
 * bdem_adjoint_k.hpp - launchers of the adjoint kernels (internal).
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "bdem_device.hpp"

namespace bdem {

constexpr int ADJ_THREADS = 256;
constexpr int ADJ_MOMENTS = 6;   /* sum m, m x, m y, m x^2, m y^2, m x y */

/* One recorded step: its sorted input state, the contact lists the step
 * kernel used (32 sorted positions per particle, full counts), and the
 * permutation of its sort (sorted position -> slot in the previous order). */
struct AdjStep {
  int n;
  const float4 *pos, *vel, *omg;
  const uint32_t* list;
  const uint32_t* nn;
  const int* idx;
};

/* Cotangent of the loss with respect to the moments of one step's output
 * (sum m x, sum m y, sum m x^2, sum m y^2, sum m x y). */
struct LossCoef { float sx, sy, sxx, syy, sxy; };

void launch_adj_moments(int n, const float4* pos, const float4* vel, const float4* omg, double* out, hipStream_t s);
/* Reverse of one recorded step.  lam_out: 3 float4 (x, v, w) per sorted
 * position of the step's output; lam_prev receives the input's cotangent in
 * the previous order.  lamF (2 per particle) and lamIn (3) are scratch.
 * part_rr (one per block) and part_pp (3 per block: Kn, Cn, kinetic
 * friction) accumulate the parameter cotangents. */
void launch_adj_back(const DevParams* dP, const AdjStep& st, const float4* lam_out, float4* lamF, float4* lamIn,
                     float4* lam_prev, const LossCoef& g, double* part_rr, double* part_pp, hipStream_t s);

}  // namespace bdem
