// postprocess.hip — postprocess steps on beamformed data in HBM, layout
// [F][W][N][P] complex64 (MATLAB [P, N, W, F]):
//  - coherent compounding (+postprocess/coherent_compounding.m:41-73), with the
//    moving sum over merged waves and frames;
//  - the active-element count of the coherence factor
//    (+postprocess/coherence_factor.m:82-131);
//  - the Mallart-Fink coherence factor (coherence_factor.m:97-143 and the
//    Python reference's python/src/ustb/postprocess/coherence_factor.py:19-50).
// One lane per output pixel (pixels are the fastest index, so loads and
// stores coalesce); sums run in index order in FP32, as the MATLAB code sums
// single data.
#include <hip/hip_runtime.h>
#include <math.h>

#include "das_cf.h"
#include "das_common.h"
#include "das_status.h"

namespace ustbx {

namespace {

constexpr int kPB = 256;

__device__ __forceinline__ void add2(float2& a, float2 b) {
  a.x = a.x + b.x;
  a.y = a.y + b.y;
}

// out (F, W', N', P) without a window: both sums waves (outer) and channels
// (inner), transmit sums waves, receive sums channels.
__global__ __launch_bounds__(kPB) void compound_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                                       int P, int N, int W, int F, int dim) {
  const size_t PS = (size_t)P;
  const size_t rows = dim == USTBX_DIM_BOTH ? (size_t)F : dim == USTBX_DIM_TX ? (size_t)F * N : (size_t)F * W;
  const size_t total = rows * PS;
  for (size_t i = (size_t)blockIdx.x * kPB + threadIdx.x; i < total; i += (size_t)gridDim.x * kPB) {
    const size_t r = i / PS, p = i - r * PS;
    float2 acc = make_float2(0.0f, 0.0f);
    if (dim == USTBX_DIM_BOTH) {
      const size_t f = r;
      for (int w = 0; w < W; ++w)
        for (int n = 0; n < N; ++n) add2(acc, in[(((f * W + w) * N) + n) * PS + p]);
    } else if (dim == USTBX_DIM_TX) {
      const size_t f = r / N, n = r - f * N;
      for (int w = 0; w < W; ++w) add2(acc, in[(((f * W + w) * N) + n) * PS + p]);
    } else {
      const size_t f = r / W, w = r - f * W;
      for (int n = 0; n < N; ++n) add2(acc, in[(((f * W + w) * N) + n) * PS + p]);
    }
    out[i] = acc;
  }
}

// Moving sum of k consecutive entries along waves and frames merged (wave
// fastest), incomplete ends discarded: out (T, 1, N', P), T = F*W - k + 1;
// both sums the channels of each entry first.
__global__ __launch_bounds__(kPB) void compound_window_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                                              int P, int N, int T, int k, int dim) {
  const size_t PS = (size_t)P;
  const int Nout = dim == USTBX_DIM_BOTH ? 1 : N;
  const size_t total = (size_t)T * Nout * PS;
  for (size_t i = (size_t)blockIdx.x * kPB + threadIdx.x; i < total; i += (size_t)gridDim.x * kPB) {
    const size_t r = i / PS, p = i - r * PS;
    const size_t t = r / Nout, n = r - t * Nout;
    float2 acc = make_float2(0.0f, 0.0f);
    for (int j = 0; j < k; ++j) {
      const size_t e = t + j;  // merged (frame, wave) entry
      if (dim == USTBX_DIM_BOTH) {
        float2 s = make_float2(0.0f, 0.0f);
        for (int c = 0; c < N; ++c) add2(s, in[(e * N + c) * PS + p]);
        add2(acc, s);
      } else {
        add2(acc, in[(e * N + n) * PS + p]);
      }
    }
    out[i] = acc;
  }
}

// M[p]: summed elements whose apodization exceeds the criterium (FP64
// comparison of the FP32 weights; both: the product of transmit and receive
// weights). A NULL apodization is all ones.
__global__ __launch_bounds__(kPB) void active_elements_kernel(const float* __restrict__ ta, const float* __restrict__ ra,
                                                              int P, int N, int W, int dim, double crit,
                                                              float* __restrict__ M) {
  const int p = blockIdx.x * kPB + threadIdx.x;
  if (p >= P) return;
  const size_t PS = (size_t)P;
  int cnt = 0;
  if (dim == USTBX_DIM_RX) {
    for (int n = 0; n < N; ++n) cnt += ((ra ? (double)ra[n * PS + p] : 1.0) > crit) ? 1 : 0;
  } else if (dim == USTBX_DIM_TX) {
    for (int w = 0; w < W; ++w) cnt += ((ta ? (double)ta[w * PS + p] : 1.0) > crit) ? 1 : 0;
  } else {
    for (int w = 0; w < W; ++w) {
      const double a = ta ? (double)ta[w * PS + p] : 1.0;
      for (int n = 0; n < N; ++n) cnt += (a * (ra ? (double)ra[n * PS + p] : 1.0) > crit) ? 1 : 0;
    }
  }
  M[p] = (float)cnt;
}

}  // namespace

namespace {

// image and cf (F, W', N', P): one lane per output pixel, the summed axes in
// index order (waves outer, channels inner for both).
__global__ __launch_bounds__(kPB) void coherence_factor_kernel(const float2* __restrict__ in, float2* __restrict__ image,
                                                               float* __restrict__ cfo, const float* __restrict__ M,
                                                               int P, int N, int W, int F, int dim, int flavour) {
  const size_t PS = (size_t)P;
  const size_t rows = dim == USTBX_DIM_BOTH ? (size_t)F : dim == USTBX_DIM_TX ? (size_t)F * N : (size_t)F * W;
  const float count = (float)(dim == USTBX_DIM_BOTH ? N * W : dim == USTBX_DIM_TX ? W : N);
  const size_t total = rows * PS;
  for (size_t i = (size_t)blockIdx.x * kPB + threadIdx.x; i < total; i += (size_t)gridDim.x * kPB) {
    const size_t r = i / PS, p = i - r * PS;
    float2 S = make_float2(0.0f, 0.0f);
    float E = 0.0f;
    if (dim == USTBX_DIM_BOTH) {
      for (int w = 0; w < W; ++w)
        for (int n = 0; n < N; ++n) cf_accumulate(S, E, in[(((r * W + w) * N) + n) * PS + p]);
    } else if (dim == USTBX_DIM_TX) {
      const size_t f = r / N, n = r - f * N;
      for (int w = 0; w < W; ++w) cf_accumulate(S, E, in[(((f * W + w) * N) + n) * PS + p]);
    } else {
      const size_t f = r / W, w = r - f * W;
      for (int n = 0; n < N; ++n) cf_accumulate(S, E, in[(((f * W + w) * N) + n) * PS + p]);
    }
    const float m = (flavour == USTBX_CF_PYTHON || !M) ? count : M[p];
    const float cf = coherence_value(S, E, m, flavour);
    image[i] = make_float2(cf * S.x, cf * S.y);
    if (cfo) cfo[i] = cf;
  }
}

unsigned grid_for(size_t total) {
  size_t b = (total + kPB - 1) / kPB;
  if (b > 65536) b = 65536;
  return (unsigned)(b ? b : 1);
}

bool dim_ok(int32_t dim) {
  return dim == USTBX_DIM_RX || dim == USTBX_DIM_TX || dim == USTBX_DIM_BOTH;
}

}  // namespace
}  // namespace ustbx

using namespace ustbx;

extern "C" {

int ustbx_coherent_compounding(const void* in, void* out, int32_t P, int32_t N, int32_t W, int32_t F, int32_t dim,
                               int32_t window_size, void* stream, ustbx_status* st) {
  if (!dim_ok(dim)) return set_status(st, USTBX_E_ARG, "Unknown dimension mode; check HELP dimension");
  if (P < 0 || N < 0 || W < 0 || F < 0 || window_size < 0) return set_status(st, USTBX_E_ARG, "negative size");
  if (window_size > W)
    return set_status(st, USTBX_E_ARG,
                      "Cannot have a window size greater that the available number of transmit events.");
  if ((int64_t)P * N * W * F == 0) return ok(st);
  if (!in || !out) return set_status(st, USTBX_E_ARG, "NULL buffer");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const float2* x = reinterpret_cast<const float2*>(in);
  float2* y = reinterpret_cast<float2*>(out);
  if (window_size == 0 || dim == USTBX_DIM_RX) {
    const size_t rows = dim == USTBX_DIM_BOTH ? (size_t)F : dim == USTBX_DIM_TX ? (size_t)F * N : (size_t)F * W;
    hipLaunchKernelGGL(compound_kernel, dim3(grid_for(rows * P)), dim3(kPB), 0, s, x, y, P, N, W, F, dim);
  } else {
    const int T = F * W - window_size + 1;
    const size_t total = (size_t)T * (dim == USTBX_DIM_BOTH ? 1 : N) * P;
    hipLaunchKernelGGL(compound_window_kernel, dim3(grid_for(total)), dim3(kPB), 0, s, x, y, P, N, T, window_size,
                       dim);
  }
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? ok(st) : hip_fail(st, e, "compounding kernel launch");
}

int ustbx_active_elements(const float* tx_apo, const float* rx_apo, int32_t P, int32_t N, int32_t W, int32_t dim,
                          double criterium, float* M, void* stream, ustbx_status* st) {
  if (!dim_ok(dim)) return set_status(st, USTBX_E_ARG, "Unknown dimension mode; check HELP dimension");
  if (P < 0 || N < 0 || W < 0) return set_status(st, USTBX_E_ARG, "negative size");
  if (P == 0) return ok(st);
  if (!M) return set_status(st, USTBX_E_ARG, "NULL output");
  hipLaunchKernelGGL(active_elements_kernel, dim3((unsigned)((P + kPB - 1) / kPB)), dim3(kPB), 0,
                     reinterpret_cast<hipStream_t>(stream), tx_apo, rx_apo, P, N, W, dim, criterium, M);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? ok(st) : hip_fail(st, e, "active-element kernel launch");
}

int ustbx_coherence_factor(const void* in, void* image, float* cf, const float* M, int32_t P, int32_t N, int32_t W,
                           int32_t F, int32_t dim, int32_t flavour, void* stream, ustbx_status* st) {
  if (!dim_ok(dim)) return set_status(st, USTBX_E_ARG, "Unknown dimension mode; check HELP dimension");
  if (flavour != USTBX_CF_MATLAB && flavour != USTBX_CF_PYTHON)
    return set_status(st, USTBX_E_ARG, "flavour must be USTBX_CF_MATLAB or USTBX_CF_PYTHON");
  if (P < 0 || N < 0 || W < 0 || F < 0) return set_status(st, USTBX_E_ARG, "negative size");
  if ((int64_t)P * N * W * F == 0) return ok(st);
  if (!in || !image) return set_status(st, USTBX_E_ARG, "NULL buffer");
  const size_t rows = dim == USTBX_DIM_BOTH ? (size_t)F : dim == USTBX_DIM_TX ? (size_t)F * N : (size_t)F * W;
  hipLaunchKernelGGL(coherence_factor_kernel, dim3(grid_for(rows * P)), dim3(kPB), 0,
                     reinterpret_cast<hipStream_t>(stream), reinterpret_cast<const float2*>(in),
                     reinterpret_cast<float2*>(image), cf, M, P, N, W, F, dim, flavour);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? ok(st) : hip_fail(st, e, "coherence-factor kernel launch");
}

}  // extern "C"
