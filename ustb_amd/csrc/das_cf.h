// das_cf.h — the per-pixel arithmetic of the Mallart-Fink coherence factor,
// shared by the postprocess kernel and the transmit kernel's fused form.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/ustbx_das.h"

namespace ustbx {

// S += v, E += abs(v).^2 (coherence_factor.m:126-127), FP32 in index order.
__device__ __forceinline__ void cf_accumulate(float2& S, float& E, float2 v) {
#pragma clang fp contract(off)
  S.x = S.x + v.x;
  S.y = S.y + v.y;
  const float a = hypotf(v.x, v.y);
  E = E + a * a;
}

// The factor from the coherent sum S and the incoherent sum E. MATLAB
// (coherence_factor.m:138-139): abs(S).^2 ./ E .* single(1 ./ M), NaN -> 0.
// Python (python/src/ustb/postprocess/coherence_factor.py:45-46):
// abs(S)^2 / (E * M + 1e-30), NaN -> 0.
__device__ __forceinline__ float coherence_value(float2 S, float E, float M, int flavour) {
#pragma clang fp contract(off)
  const float a = hypotf(S.x, S.y);
  const float a2 = a * a;
  float cf;
  if (flavour == USTBX_CF_PYTHON)
    cf = a2 / (E * M + 1e-30f);
  else
    cf = (a2 / E) * (float)(1.0 / (double)M);
  return cf != cf ? 0.0f : cf;
}

}  // namespace ustbx
