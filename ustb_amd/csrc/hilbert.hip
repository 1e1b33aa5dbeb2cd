// hilbert.hip — analytic signal of RF channel data on the device, the RF
// branch of +midprocess/das.m:211-214 (MATLAB hilbert) and
// python/src/ustb/midprocess/das.py:144-145 (scipy.signal.hilbert):
// X = FFT(x) along time; Z[0] = X[0], Z[k] = 2 X[k] for 0 < k < L/2,
// Z[L/2] = X[L/2] (L even), Z[k] = 0 above; z = IFFT(Z).
// FFTs are hipFFT's (R2C forward, C2C inverse); plans are cached per
// (device, L, traces).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <stdarg.h>
#include <stdio.h>

#include <map>
#include <mutex>
#include <tuple>

#include "../../include/ustbx_das.h"

namespace {

struct FftPlans {
  hipfftHandle r2c = 0, c2c = 0;
  float2* spec = nullptr;  // [T][L/2+1]
};

std::mutex g_mu;
std::map<std::tuple<int, int, long long>, FftPlans> g_plans;

int fail(ustbx_status* st, int code, const char* fmt, ...) {
  if (st) {
    st->code = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(st->message, sizeof(st->message), fmt, ap);
    va_end(ap);
  }
  return code;
}

__global__ void analytic_weights_kernel(const float2* __restrict__ spec, float2* __restrict__ z, int L,
                                        long long T) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = T * (long long)L;
  if (i >= n) return;
  const long long t = i / L;
  const int k = (int)(i - t * L);
  const int half = L / 2;
  float w;
  if (k == 0) w = 1.0f;
  else if ((L % 2 == 0) && k == half) w = 1.0f;
  else if (k <= (L - 1) / 2) w = 2.0f;
  else w = 0.0f;
  const float s = w / (float)L;
  float2 v = make_float2(0.0f, 0.0f);
  if (k <= half) {
    const float2 x = spec[t * (half + 1) + k];
    v = make_float2(x.x * s, x.y * s);
  }
  z[i] = v;
}

}  // namespace

extern "C" int ustbx_hilbert(const float* rf, void* analytic, int32_t L, int64_t traces, void* stream,
                             ustbx_status* st) {
  if (!rf || !analytic) return fail(st, USTBX_E_ARG, "NULL argument");
  if (L < 1 || traces < 0) return fail(st, USTBX_E_DIMENSIONS, "L must be >= 1 and traces >= 0");
  if (traces == 0) return USTBX_OK;
  if (traces > INT32_MAX) return fail(st, USTBX_E_CAPACITY, "more than 2^31-1 traces");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(st, USTBX_E_GPU, "HIP error: %s", hipGetErrorString(e));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lock(g_mu);
  auto key = std::make_tuple(dev, (int)L, (long long)traces);
  auto it = g_plans.find(key);
  if (it == g_plans.end()) {
    FftPlans p;
    int n[1] = {L};
    if (hipfftPlanMany(&p.r2c, 1, n, nullptr, 1, L, nullptr, 1, L / 2 + 1, HIPFFT_R2C, (int)traces) !=
            HIPFFT_SUCCESS ||
        hipfftPlanMany(&p.c2c, 1, n, nullptr, 1, L, nullptr, 1, L, HIPFFT_C2C, (int)traces) != HIPFFT_SUCCESS)
      return fail(st, USTBX_E_GPU, "hipfftPlanMany failed (L=%d, traces=%lld)", L, (long long)traces);
    e = hipMalloc(&p.spec, sizeof(float2) * (size_t)traces * (L / 2 + 1));
    if (e != hipSuccess) return fail(st, USTBX_E_GPU, "HIP error: %s", hipGetErrorString(e));
    it = g_plans.emplace(key, p).first;
  }
  FftPlans& p = it->second;
  if (hipfftSetStream(p.r2c, s) != HIPFFT_SUCCESS || hipfftSetStream(p.c2c, s) != HIPFFT_SUCCESS)
    return fail(st, USTBX_E_GPU, "hipfftSetStream failed");
  if (hipfftExecR2C(p.r2c, const_cast<float*>(rf), reinterpret_cast<hipfftComplex*>(p.spec)) != HIPFFT_SUCCESS)
    return fail(st, USTBX_E_GPU, "hipfftExecR2C failed");
  const long long total = (long long)traces * L;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  hipLaunchKernelGGL(analytic_weights_kernel, dim3(blocks), dim3(256), 0, s, p.spec,
                     reinterpret_cast<float2*>(analytic), (int)L, (long long)traces);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(st, USTBX_E_GPU, "HIP error: %s", hipGetErrorString(e));
  hipfftComplex* z = reinterpret_cast<hipfftComplex*>(analytic);
  if (hipfftExecC2C(p.c2c, z, z, HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
    return fail(st, USTBX_E_GPU, "hipfftExecC2C failed");
  if (st) {
    st->code = USTBX_OK;
    st->message[0] = 0;
  }
  return USTBX_OK;
}
