// demodulate.hip — RF -> IQ on the device: preprocess.fast_demodulation
// (+preprocess/fast_demodulation.m:58-183) and the Python reference's
// FastDemodulation (python/src/ustb/preprocess/fast_demodulation.py:24-85).
//
// Both are: mix each trace by exp(-1j 2 pi fmod t), t = t0 + n / fs (the
// phase in FP64, the factor rounded to FP32 and multiplied in FP32 as MATLAB
// does with single data); low-pass FIR; keep every down-th sample from the
// first. The two filters differ:
//  - causal (MATLAB filter(b, 1, x) with zero initial state): taps at offsets
//    -(K-1) .. 0, samples before the trace are zero, summed as filter's
//    transposed direct form nests them (the last tap innermost);
//  - zero-phase (scipy filtfilt(b, 1, x) with its odd extension of 3K
//    samples): on the cropped output the forward and backward passes never
//    reach the extension's ends, so the result is one FIR with the taps'
//    autocorrelation c[d] = sum_k b[k] b[k+d], d = -(K-1) .. K-1, over the
//    odd-extended signal (2 e[0] - e[-m] before the start, 2 e[L-1] -
//    e[2L-2-m] after the end).
// One block per (trace, run of OB outputs) (grid x: traces, y: runs): the mixed samples of its input
// window are staged in LDS once, then each lane forms one output.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "das_status.h"

namespace ustbx {

namespace {

constexpr int kDB = 256;
constexpr int kMaxTaps = 767;    // passed by value (kernel arguments stay under 4 KB); zero-phase: K <= 384
constexpr int kWindow = 7680;    // staged samples per block (60 KB; with the taps under 64 KB of LDS)

struct DemodArgs {
  const float* rf;
  float2* out;
  int L, Lout, down, ntaps, off0, OB, mode;
  double fs, t0, fmod;
  float taps[kMaxTaps];
};

// mixed sample e[m] of the trace, m in [0, L)
__device__ __forceinline__ float2 mixed(const float* __restrict__ x, int m, const DemodArgs& A) {
#pragma clang fp contract(off)
  const double t = A.t0 + (double)m / A.fs;
  const double ph = ((-2.0 * M_PI) * A.fmod) * t;
  double s, c;
  sincos(ph, &s, &c);
  const float v = x[m];
  return make_float2(v * (float)c, v * (float)s);
}

// e at extended index m (zero / odd extension outside [0, L))
__device__ __forceinline__ float2 extended(const float* __restrict__ x, int m, const DemodArgs& A) {
#pragma clang fp contract(off)
  if (m >= 0 && m < A.L) return mixed(x, m, A);
  if (A.mode == USTBX_FIR_CAUSAL) return make_float2(0.0f, 0.0f);
  const float2 edge = mixed(x, m < 0 ? 0 : A.L - 1, A);
  const float2 refl = mixed(x, m < 0 ? -m : 2 * A.L - 2 - m, A);
  return make_float2(2.0f * edge.x - refl.x, 2.0f * edge.y - refl.y);
}

__global__ __launch_bounds__(kDB) void demodulate_kernel(DemodArgs A) {
#pragma clang fp contract(off)
  extern __shared__ float2 win[];
  __shared__ float tp[kMaxTaps];
  const size_t trace = blockIdx.x;
  const float* __restrict__ x = A.rf + trace * (size_t)A.L;
  const int j0 = (int)blockIdx.y * A.OB;
  const int jn = min(A.OB, A.Lout - j0);
  const int lo = j0 * A.down + A.off0;                         // extended index of win[0]
  const int count = (jn - 1) * A.down + A.ntaps;
  for (int i = threadIdx.x; i < A.ntaps; i += kDB) tp[i] = A.taps[i];
  for (int i = threadIdx.x; i < count; i += kDB) win[i] = extended(x, lo + i, A);
  __syncthreads();
  const int j = threadIdx.x;
  if (j < jn) {
    const float2* w = win + j * A.down;
    float2 acc = make_float2(tp[0] * w[0].x, tp[0] * w[0].y);
    for (int i = 1; i < A.ntaps; ++i) {
      const float t = tp[i];
      acc.x = t * w[i].x + acc.x;
      acc.y = t * w[i].y + acc.y;
    }
    A.out[trace * (size_t)A.Lout + j0 + j] = acc;
  }
}

}  // namespace
}  // namespace ustbx

using namespace ustbx;

extern "C" {

int ustbx_demodulate(const float* rf, void* iq, int32_t L, int64_t traces, double fs, double t0, double fmod,
                     const double* b, int32_t ntaps, int32_t down, int32_t mode, void* stream, ustbx_status* st) {
  if (mode != USTBX_FIR_CAUSAL && mode != USTBX_FIR_ZERO_PHASE)
    return set_status(st, USTBX_E_ARG, "mode must be USTBX_FIR_CAUSAL or USTBX_FIR_ZERO_PHASE");
  if (L < 0 || traces < 0) return set_status(st, USTBX_E_ARG, "negative size");
  if (!b || ntaps < 1) return set_status(st, USTBX_E_ARG, "the filter needs at least one tap");
  if (down < 1) return set_status(st, USTBX_E_ARG, "the decimation factor must be at least 1");
  if (!(fs > 0.0)) return set_status(st, USTBX_E_ARG, "the sampling frequency must be positive");
  const int K = ntaps;
  const int T = mode == USTBX_FIR_CAUSAL ? K : 2 * K - 1;
  if (T > kMaxTaps)
    return set_status(st, USTBX_E_ARG, "at most %d taps", mode == USTBX_FIR_CAUSAL ? kMaxTaps : (kMaxTaps + 1) / 2);
  if (mode == USTBX_FIR_ZERO_PHASE && L <= 3 * K)
    return set_status(st, USTBX_E_ARG, "The length of the input vector x must be greater than padlen, which is %d.",
                      3 * K);
  if (L == 0 || traces == 0) return ok(st);
  if (traces > 0x7FFFFFFF) return set_status(st, USTBX_E_ARG, "too many traces");
  if (!rf || !iq) return set_status(st, USTBX_E_ARG, "NULL buffer");
  if ((int64_t)T + down > kWindow) return set_status(st, USTBX_E_ARG, "decimation factor too large for the filter");
  // taps in FP32 (MATLAB filter runs in the data's single precision); the
  // zero-phase form's autocorrelation is formed in FP64 first
  DemodArgs A;
  if (mode == USTBX_FIR_CAUSAL) {
    for (int i = 0; i < K; ++i) A.taps[i] = (float)b[K - 1 - i];  // offsets -(K-1) .. 0
  } else {
    for (int d = -(K - 1); d <= K - 1; ++d) {
      double c = 0.0;
      for (int k = 0; k < K; ++k)
        if (k + d >= 0 && k + d < K) c += b[k] * b[k + d];
      A.taps[d + K - 1] = (float)c;
    }
  }
  A.rf = rf;
  A.out = reinterpret_cast<float2*>(iq);
  A.L = L;
  A.Lout = (L + down - 1) / down;
  A.down = down;
  A.ntaps = T;
  A.off0 = -(K - 1);
  A.OB = (int)std::min<int64_t>(kDB, (kWindow - T) / down + 1);
  A.mode = mode;
  A.fs = fs;
  A.t0 = t0;
  A.fmod = fmod;
  const int64_t gy = (A.Lout + A.OB - 1) / A.OB;
  if (gy > 65535) return set_status(st, USTBX_E_ARG, "trace too long for the decimation factor");
  const size_t lds = sizeof(float2) * (size_t)((A.OB - 1) * down + T);
  hipLaunchKernelGGL(demodulate_kernel, dim3((unsigned)traces, (unsigned)gy), dim3(kDB), lds,
                     reinterpret_cast<hipStream_t>(stream), A);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(st, e, "demodulate_kernel launch");
  return ok(st);
}

}  // extern "C"
