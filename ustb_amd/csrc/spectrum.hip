// spectrum.hip — the modulation-frequency estimate of preprocess.fast_demodulation
// (fast_demodulation.m:57-73) when the caller sets none: the power spectrum
// of the RF traces averaged over channels, waves and frames
// (tools.power_spectrum, power_spectrum.m:1-22) on the device, then the -6 dB
// points around its maximum and their centre on the host.
//
// Device: one batched hipFFT R2C of the traces; a reduction kernel sums
// |X[k]|^2 over the traces per bin in FP64 (one lane per bin, the traces in
// order). Host: the two-sided spectrum from the one-sided bins (a real
// signal's spectrum is even), fftshift, normalisation by the maximum, the
// first maximum's index ic, MATLAB interp1 of the frequency at half the
// maximum over pw(1:ic) and over pw(ic:round(L/2)) (interp1 sorts its sample
// points), fc the mean of the two, and modulation_frequency = -fc.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <math.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "../../include/ustbx_das.h"
#include "das_status.h"

namespace {

__global__ __launch_bounds__(256) void power_sum_kernel(const float2* __restrict__ spec, int64_t traces, int bins,
                                                        double* __restrict__ pw) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= bins) return;
  double acc = 0.0;
  for (int64_t t = 0; t < traces; ++t) {
    const float2 v = spec[(size_t)t * bins + k];
    acc += (double)v.x * v.x + (double)v.y * v.y;
  }
  pw[k] = acc;
}

// MATLAB interp1(x, v, xq) for distinct sample points in any order: sorted
// first, then linear; NaN outside their range or on repeated points.
double interp1_sorted(std::vector<double> x, std::vector<double> v, double xq) {
  std::vector<size_t> o(x.size());
  std::iota(o.begin(), o.end(), 0);
  std::stable_sort(o.begin(), o.end(), [&](size_t a, size_t b) { return x[a] < x[b]; });
  std::vector<double> xs(o.size()), vs(o.size());
  for (size_t i = 0; i < o.size(); ++i) {
    xs[i] = x[o[i]];
    vs[i] = v[o[i]];
  }
  for (size_t i = 1; i < xs.size(); ++i)
    if (xs[i] == xs[i - 1]) return NAN;
  if (xs.empty() || xq < xs.front() || xq > xs.back()) return NAN;
  size_t j = (size_t)(std::upper_bound(xs.begin(), xs.end(), xq) - xs.begin());
  if (j >= xs.size()) return vs.back();
  if (j == 0) return vs.front();
  const double t = (xq - xs[j - 1]) / (xs[j] - xs[j - 1]);
  return vs[j - 1] + t * (vs[j] - vs[j - 1]);
}

}  // namespace

extern "C" {

int ustbx_estimate_modulation_frequency(const float* rf, int32_t L, int64_t traces, double fs, double* fmod,
                                        double* pw, void* stream, ustbx_status* st) {
  if (!rf || (!fmod && !pw)) return set_status(st, USTBX_E_ARG, "NULL argument");
  if (L < 4 || traces <= 0 || traces > INT32_MAX) return set_status(st, USTBX_E_DIMENSIONS, "L >= 4 and traces >= 1");
  if (!(fs > 0.0)) return set_status(st, USTBX_E_ARG, "the sampling frequency must be positive");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int bins = L / 2 + 1;
  float2* spec = nullptr;
  double* dpw = nullptr;
  hipfftHandle plan = 0;
  hipError_t e = hipMalloc(&spec, sizeof(float2) * (size_t)bins * (size_t)traces);
  if (e == hipSuccess) e = hipMalloc(&dpw, sizeof(double) * bins);
  auto release = [&]() {
    if (plan) (void)hipfftDestroy(plan);
    (void)hipFree(spec);
    (void)hipFree(dpw);
  };
  if (e != hipSuccess) {
    release();
    return hip_fail(st, e, "hipMalloc (power spectrum)");
  }
  int n[1] = {L};
  if (hipfftPlanMany(&plan, 1, n, nullptr, 1, L, nullptr, 1, bins, HIPFFT_R2C, (int)traces) != HIPFFT_SUCCESS ||
      hipfftSetStream(plan, s) != HIPFFT_SUCCESS ||
      hipfftExecR2C(plan, const_cast<float*>(rf), reinterpret_cast<hipfftComplex*>(spec)) != HIPFFT_SUCCESS) {
    release();
    return set_status(st, USTBX_E_GPU, "hipFFT R2C failed (L=%d, traces=%lld)", L, (long long)traces);
  }
  hipLaunchKernelGGL(power_sum_kernel, dim3((unsigned)((bins + 255) / 256)), dim3(256), 0, s, spec, traces, bins,
                     dpw);
  std::vector<double> half(bins);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(half.data(), dpw, sizeof(double) * bins, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  release();
  if (e != hipSuccess) return hip_fail(st, e, "power spectrum");
  // two-sided |X[k]|^2 (k = 0 .. L-1), fftshift: index j holds bin
  // (j + ceil(L/2)) mod L, frequency -fs/2 + j fs / L
  std::vector<double> spec2(L), fx(L);
  const int sh = (L + 1) / 2;
  double mx = 0.0;
  for (int j = 0; j < L; ++j) {
    const int k = (j + sh) % L;
    spec2[j] = half[k <= L / 2 ? k : L - k] / (double)traces;
    mx = std::max(mx, spec2[j]);
    fx[j] = -fs / 2 + (double)j * fs / L;
  }
  for (int j = 0; j < L; ++j) spec2[j] /= mx;
  if (pw)
    for (int j = 0; j < L; ++j) pw[j] = spec2[j];
  if (fmod) {
    const int ic = (int)(std::max_element(spec2.begin(), spec2.end()) - spec2.begin());  // the first maximum
    const double half_max = spec2[ic] / 2;
    const int mid = (int)floor(L / 2.0 + 0.5);  // round(end/2), 1-based
    const double lo = interp1_sorted(std::vector<double>(spec2.begin(), spec2.begin() + ic + 1),
                                     std::vector<double>(fx.begin(), fx.begin() + ic + 1), half_max);
    const double up = ic < mid ? interp1_sorted(std::vector<double>(spec2.begin() + ic, spec2.begin() + mid),
                                                std::vector<double>(fx.begin() + ic, fx.begin() + mid), half_max)
                               : NAN;
    *fmod = -(lo + up) / 2;
  }
  return ok(st);
}

}  // extern "C"
