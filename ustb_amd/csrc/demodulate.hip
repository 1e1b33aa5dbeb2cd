// demodulate.hip — RF -> IQ on the device: preprocess.fast_demodulation
// (+preprocess/fast_demodulation.m:58-183) and the Python reference's
// FastDemodulation (python/src/ustb/preprocess/fast_demodulation.py:24-85).
//
// Both are: mix each trace by exp(-1j 2 pi fmod t), t = t0 + n / fs (the
// phase in FP64, the factor rounded to FP32 and multiplied in FP32 as MATLAB
// does with single data); low-pass FIR; keep every down-th sample from the
// first. The two filters differ:
//  - causal (MATLAB filter(b, 1, x) with zero initial state): taps at offsets
//    -(K-1) .. 0, samples before the trace are zero;
//  - zero-phase (scipy filtfilt(b, 1, x) with its odd extension of 3K
//    samples): on the cropped output the forward and backward passes never
//    reach the extension's ends, so the result is one FIR with the taps'
//    autocorrelation c[d] = sum_k b[k] b[k+d], d = -(K-1) .. K-1, over the
//    odd-extended signal (2 e[0] - e[-m] before the start, 2 e[L-1] -
//    e[2L-2-m] after the end).
//
// The mixing factors depend on the sample index only: they are computed once
// per (device, L, fs, t0, fmod) into a small device table that later calls
// reuse. One block per (trace, run of up to R * kDB outputs) stages the
// mixed samples of its input window in LDS, split by decimation phase
// (phase r holds window samples r, r + down, ...; one pad slot every R
// entries, so the lanes' reads, R entries apart, hit distinct banks). Each
// lane forms R consecutive outputs (R = 8 without decimation, 4 with): per
// phase, a window of R samples slides along the phase's taps in registers, so
// every staged sample a lane needs is read once. The sum runs phase by phase
// (within a phase in tap order), whatever R is.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "das_status.h"

namespace ustbx {

namespace {

constexpr int kDB = 256;         // threads per block
// window samples a lane loads per staging round: 9 x 256 covers the window of
// 1024 outputs at decimation 2, or 2048 at decimation 1, with up to 241 taps
// in one round, so a block waits for memory once (8: two rounds, 12 % slower
// on 128 x 128 traces of 2048 samples, 31 taps, decimation 2)
constexpr int kSB = 9;
constexpr int kMaxTaps = 767;    // passed by value (kernel arguments stay under 4 KB); zero-phase: K <= 384
constexpr size_t kLdsMax = 64 * 1024;

struct DemodArgs {
  const float* rf;
  const float2* mix;   // [L] mixing factors
  float2* out;
  int L, Lout, down, ntaps, off0, OB, mode;
  int PS;              // LDS entries per phase (padded)
  int TP;              // staged taps, zero beyond ntaps
  float taps[kMaxTaps];
};

// entries of one phase for m window samples (one pad slot per 4) and the
// dynamic LDS of a block of OB outputs
inline int phase_slots(int mr, int R) { return mr + mr / R + 1; }
inline int phase_len(int OB, int down, int T, int R) {
  const int count = (OB - 1) * down + T;
  return (count + down - 1) / down + 2 * R;
}
inline size_t demod_lds(int OB, int down, int T, int R) {
  return sizeof(float2) * (size_t)down * phase_slots(phase_len(OB, down, T, R), R) +
         sizeof(float) * (size_t)(T + (R + 1) * down);
}

__global__ __launch_bounds__(kDB) void mix_table_kernel(float2* __restrict__ tab, int L, double fs, double t0,
                                                        double fmod) {
#pragma clang fp contract(off)
  const int m = blockIdx.x * kDB + threadIdx.x;
  if (m >= L) return;
  const double t = t0 + (double)m / fs;
  const double ph = ((-2.0 * M_PI) * fmod) * t;
  double s, c;
  sincos(ph, &s, &c);
  tab[m] = make_float2((float)c, (float)s);
}

// mixed sample e[m] of the trace, m in [0, L)
__device__ __forceinline__ float2 mixed(const float* __restrict__ x, int m, const DemodArgs& A) {
  const float v = x[m];
  const float2 e = A.mix[m];
  return make_float2(v * e.x, v * e.y);
}

template <int kR>
__device__ __forceinline__ int slot(int m) { return m + m / kR; }

template <int kR>
__global__ __launch_bounds__(kDB) void demodulate_kernel(DemodArgs A) {
  extern __shared__ float2 lds[];
  const int down = A.down, T = A.ntaps, PS = A.PS;
  float* tp = reinterpret_cast<float*>(lds + (size_t)down * PS);
  const size_t trace = blockIdx.x;
  const float* __restrict__ x = A.rf + trace * (size_t)A.L;
  const int j0 = (int)blockIdx.y * A.OB;
  const int jn = min(A.OB, A.Lout - j0);
  const int lo = j0 * down + A.off0;            // extended index of window sample 0
  const int count = (jn - 1) * down + T;        // window samples the outputs read
  const int mr = (count + down - 1) / down + 2 * kR;
  for (int i = threadIdx.x; i < A.TP; i += kDB) tp[i] = i < T ? A.taps[i] : 0.0f;
  // Window sample i (extended index lo + i) goes to phase i mod down, entry
  // i / down. Each lane issues kSB loads per round, all in bounds (an index
  // outside the trace reads its odd-extension reflection, or a sample whose
  // value is not used), then forms and stores the samples.
  const bool zero_phase = A.mode == USTBX_FIR_ZERO_PHASE;
  const int total = mr * down;
  float2 e0 = make_float2(0.0f, 0.0f), e1 = e0;
  if (zero_phase) {
    e0 = mixed(x, 0, A);
    e1 = mixed(x, A.L - 1, A);
  }
  for (int base = 0; base < total; base += kSB * kDB) {
    float xv[kSB];
    float2 mv[kSB];
#pragma unroll
    for (int k = 0; k < kSB; ++k) {
      const int i = base + k * kDB + (int)threadIdx.x;
      const int m = lo + i;
      int src = m < 0 ? (zero_phase ? -m : 0) : (m >= A.L ? (zero_phase ? 2 * A.L - 2 - m : 0) : m);
      // lanes past the window read the window's first sample (cached; unused)
      src = i < total ? min(max(src, 0), A.L - 1) : min(max(lo, 0), A.L - 1);
      xv[k] = x[src];
      mv[k] = A.mix[src];
    }
#pragma unroll
    for (int k = 0; k < kSB; ++k) {
      const int i = base + k * kDB + (int)threadIdx.x;
      if (i < total) {
#pragma clang fp contract(off)
        const int m = lo + i;
        float2 v = make_float2(xv[k] * mv[k].x, xv[k] * mv[k].y);
        if (m < 0 || m >= A.L) {
          const float2 edge = m < 0 ? e0 : e1;
          v = zero_phase ? make_float2(2.0f * edge.x - v.x, 2.0f * edge.y - v.y) : make_float2(0.0f, 0.0f);
        }
        if (i >= count) v = make_float2(0.0f, 0.0f);
        const int q = i / down;
        lds[(size_t)(i - q * down) * PS + slot<kR>(q)] = v;
      }
    }
  }
  __syncthreads();
  const int jt = kR * (int)threadIdx.x;         // first output of this lane (block-local)
  if (jt >= jn) return;
  float2 acc[kR];
#pragma unroll
  for (int k = 0; k < kR; ++k) acc[k] = make_float2(0.0f, 0.0f);
  for (int r = 0; r < down; ++r) {
    const float2* __restrict__ ph = lds + (size_t)r * PS;
    const int Q = (T - r + down - 1) / down;    // taps r, r + down, ... of this phase
    // w[s] holds phase sample jt + q0 + s at the start of each block of kR taps
    float2 w[kR];
#pragma unroll
    for (int k = 0; k < kR; ++k) w[k] = ph[slot<kR>(jt + k)];
    for (int q0 = 0; q0 < Q; q0 += kR) {
#pragma unroll
      for (int u = 0; u < kR; ++u) {
        if (q0 + u >= Q) break;                 // uniform
        const float t = tp[(q0 + u) * down + r];
#pragma unroll
        for (int k = 0; k < kR; ++k) {
          const float2 s = w[(u + k) % kR];
          acc[k].x = fmaf(t, s.x, acc[k].x);
          acc[k].y = fmaf(t, s.y, acc[k].y);
        }
        w[u] = ph[slot<kR>(jt + q0 + u + kR)];
      }
    }
  }
  float2* __restrict__ o = A.out + trace * (size_t)A.Lout + j0 + jt;
#pragma unroll
  for (int k = 0; k < kR; ++k)
    if (jt + k < jn) o[k] = acc[k];
}

struct MixTable {
  float2* tab = nullptr;
  unsigned long long last_use = 0;
};

std::mutex g_mu;
std::map<std::tuple<int, int, double, double, double>, MixTable> g_mix;
unsigned long long g_mix_clock = 0;
// tables kept per device; the least recently used one is released (after the
// device has finished its queued work) to make room for a new one
constexpr int kMaxMixTables = 64;

hipError_t evict_mix(int dev) {
  for (;;) {
    int n = 0;
    auto lru = g_mix.end();
    for (auto it = g_mix.begin(); it != g_mix.end(); ++it) {
      if (std::get<0>(it->first) != dev) continue;
      ++n;
      if (lru == g_mix.end() || it->second.last_use < lru->second.last_use) lru = it;
    }
    if (n < kMaxMixTables) return hipSuccess;
    const hipError_t e = hipDeviceSynchronize();  // no queued work still reads it
    if (e != hipSuccess) return e;
    (void)hipFree(lru->second.tab);
    g_mix.erase(lru);
  }
}

// The mixing-factor table of (device, L, fs, t0, fmod), built on first use on
// the caller's stream (and waited for, so that other streams may read it).
// The caller holds g_mu until it has launched the kernel that reads the
// table, so that no other call releases the table before that launch.
int mix_table(int L, double fs, double t0, double fmod, hipStream_t s, const float2** out, ustbx_status* st) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(st, e, "hipGetDevice");
  const auto key = std::make_tuple(dev, L, fs, t0, fmod);
  auto it = g_mix.find(key);
  if (it != g_mix.end()) {
    it->second.last_use = ++g_mix_clock;
    *out = it->second.tab;
    return ok(st);
  }
  e = evict_mix(dev);
  if (e != hipSuccess) return hip_fail(st, e, "releasing a mixing-factor table");
  float2* tab = nullptr;
  e = hipMalloc(&tab, sizeof(float2) * (size_t)L);
  if (e != hipSuccess) return hip_fail(st, e, "hipMalloc (mixing factors)");
  hipLaunchKernelGGL(mix_table_kernel, dim3((unsigned)((L + kDB - 1) / kDB)), dim3(kDB), 0, s, tab, L, fs, t0, fmod);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(tab);
    return hip_fail(st, e, "mixing factors");
  }
  g_mix.emplace(key, MixTable{tab, ++g_mix_clock});
  *out = tab;
  return ok(st);
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
  // outputs per lane: 8 halves the LDS reads per FMA (the zero-phase form,
  // decimation 1, is VALU-bound: 14 % faster than 4); with decimation the
  // traces have fewer outputs (1024 of 2048 samples at 2), which 4 per lane
  // spreads over every lane of a block
  const int R = down == 1 ? 8 : 4;
  // outputs per block: R per lane, fewer where the window would not fit in LDS
  int OB = R * kDB;
  while (OB > R && demod_lds(OB, down, T, R) > kLdsMax) OB /= 2;
  if (demod_lds(OB, down, T, R) > kLdsMax)
    return set_status(st, USTBX_E_ARG, "decimation factor too large for the filter");
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
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const float2* mix = nullptr;
  std::lock_guard<std::mutex> lock(g_mu);  // held through the launch that reads the table
  const int rc = mix_table(L, fs, t0, fmod, s, &mix, st);
  if (rc != USTBX_OK) return rc;
  A.rf = rf;
  A.mix = mix;
  A.out = reinterpret_cast<float2*>(iq);
  A.L = L;
  A.Lout = (L + down - 1) / down;
  A.down = down;
  A.ntaps = T;
  A.off0 = -(K - 1);
  A.OB = OB;
  A.mode = mode;
  A.PS = phase_slots(phase_len(OB, down, T, R), R);
  A.TP = T + (R + 1) * down;
  const int64_t gy = (A.Lout + A.OB - 1) / A.OB;
  if (gy > 65535) return set_status(st, USTBX_E_ARG, "trace too long for the decimation factor");
  if (R == 8)
    hipLaunchKernelGGL(demodulate_kernel<8>, dim3((unsigned)traces, (unsigned)gy), dim3(kDB),
                       demod_lds(OB, down, T, R), s, A);
  else
    hipLaunchKernelGGL(demodulate_kernel<4>, dim3((unsigned)traces, (unsigned)gy), dim3(kDB),
                       demod_lds(OB, down, T, R), s, A);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(st, e, "demodulate_kernel launch");
  return ok(st);
}

}  // extern "C"
