// hilbert.hip — analytic signal of RF channel data on the device, the RF
// branch of +midprocess/das.m:211-214 (MATLAB hilbert) and
// python/src/ustb/midprocess/das.py:144-145 (scipy.signal.hilbert):
// X = FFT(x) along time; Z[0] = X[0], Z[k] = 2 X[k] for 0 < k < L/2,
// Z[L/2] = X[L/2] (L even), Z[k] = 0 above; z = IFFT(Z).
// L = 512, 1024, 2048 or 4096: one block per two traces computes all of it in
// LDS (one complex signal a + i b; forward and inverse radix-8 Stockham FFTs,
// a last radix-2/4 stage where log2 L is not a multiple of 3, twiddles from a
// per-device table rounded from FP64), so HBM sees the traces read once and
// the analytic signals written once. Other lengths of the form 2^a 3^b 5^c
// 7^d up to 4096: the same with mixed-radix stages. The rest: hipFFT (R2C
// forward, C2C inverse, with the weights in between); plans and their
// spectrum buffers are cached per (device, L, traces, stream).
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
  unsigned long long last_use = 0;
};

std::mutex g_mu;
// per (device, L, traces, stream): a plan's work areas and spectrum buffer are
// never shared by calls that may run at the same time on different streams.
// At most kMaxFftPlans per device are kept (each holds traces x (L/2+1)
// complex values plus hipFFT's work areas); the least recently used one is
// released, after the device has finished its queued work, to make room.
constexpr int kMaxFftPlans = 8;
std::map<std::tuple<int, int, long long, hipStream_t>, FftPlans> g_plans;
unsigned long long g_plan_clock = 0;

void release(FftPlans& p) {
  if (p.r2c) (void)hipfftDestroy(p.r2c);
  if (p.c2c) (void)hipfftDestroy(p.c2c);
  if (p.spec) (void)hipFree(p.spec);
  p = FftPlans{};
}

// Makes room for one more plan on device dev (the current device).
hipError_t evict_plans(int dev) {
  for (;;) {
    int n = 0;
    auto lru = g_plans.end();
    for (auto it = g_plans.begin(); it != g_plans.end(); ++it) {
      if (std::get<0>(it->first) != dev) continue;
      ++n;
      if (lru == g_plans.end() || it->second.last_use < lru->second.last_use) lru = it;
    }
    if (n < kMaxFftPlans) return hipSuccess;
    const hipError_t e = hipDeviceSynchronize();  // no queued work still uses it
    if (e != hipSuccess) return e;
    release(lru->second);
    g_plans.erase(lru);
  }
}

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

// ---- LDS path (L a power of two, 512 .. 4096) ----
__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// (x + iy) * (s i), s = +-1
__device__ __forceinline__ float2 mul_si(float2 a, float s) { return make_float2(-s * a.y, s * a.x); }

// In-register DFTs of size 2, 4, 8: X[q] = sum_r v[r] exp(s 2 pi i r q / R).
__device__ __forceinline__ void dft2(float2* v, float) {
  const float2 a = v[0], b = v[1];
  v[0] = make_float2(a.x + b.x, a.y + b.y);
  v[1] = make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ void dft4v(float2& v0, float2& v1, float2& v2, float2& v3, float s) {
  const float2 t0 = make_float2(v0.x + v2.x, v0.y + v2.y), t1 = make_float2(v0.x - v2.x, v0.y - v2.y);
  const float2 t2 = make_float2(v1.x + v3.x, v1.y + v3.y);
  const float2 t3 = mul_si(make_float2(v1.x - v3.x, v1.y - v3.y), s);
  v0 = make_float2(t0.x + t2.x, t0.y + t2.y);
  v2 = make_float2(t0.x - t2.x, t0.y - t2.y);
  v1 = make_float2(t1.x + t3.x, t1.y + t3.y);
  v3 = make_float2(t1.x - t3.x, t1.y - t3.y);
}
__device__ __forceinline__ void dft4(float2* v, float s) { dft4v(v[0], v[1], v[2], v[3], s); }
__device__ __forceinline__ void dft8(float2* v, float s) {
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4v(e0, e1, e2, e3, s);
  dft4v(o0, o1, o2, o3, s);
  const float c = 0.70710678118654752f;
  // W8^q = exp(s 2 pi i q / 8): 1, c (1 + s i), s i, c (-1 + s i)
  const float2 w1 = cmulf(o1, make_float2(c, s * c));
  const float2 w2 = mul_si(o2, s);
  const float2 w3 = cmulf(o3, make_float2(-c, s * c));
  v[0] = make_float2(e0.x + o0.x, e0.y + o0.y);
  v[4] = make_float2(e0.x - o0.x, e0.y - o0.y);
  v[1] = make_float2(e1.x + w1.x, e1.y + w1.y);
  v[5] = make_float2(e1.x - w1.x, e1.y - w1.y);
  v[2] = make_float2(e2.x + w2.x, e2.y + w2.y);
  v[6] = make_float2(e2.x - w2.x, e2.y - w2.y);
  v[3] = make_float2(e3.x + w3.x, e3.y + w3.y);
  v[7] = make_float2(e3.x - w3.x, e3.y - w3.y);
}

// LDS slot of point e: one pad slot per 8 points, so that the strided writes
// of the first two stages fall on distinct banks.
__device__ __forceinline__ int pad8(int e) { return e + (e >> 3); }

// Twiddle table of one length N (stages after the first, which has none):
// stage Ns = 8 and 64 (radix 8), then Ns = 512 (radix N / 512); entry
// (r - 1) Ns + k of a stage is exp(-2 pi i k r / (Ns R)), so the lanes of a
// wave (consecutive k) read consecutive entries.
constexpr int kTw8 = 0, kTw64 = 7 * 8, kTw512 = kTw64 + 7 * 64;  // 504 entries before the last stage
__host__ __device__ constexpr int twiddle_count(int N) { return kTw512 + (N / 512 - 1) * 512; }

// One Stockham stage of radix R over buf (N points, N / 8 threads, 8 points
// per thread): butterfly j reads j + r N/R, twiddles by exp(s 2 pi i k r /
// (Ns R)), k = j mod Ns, and writes (j - k) R + k + r Ns. All reads happen
// before the barrier, all writes after it.
template <int N, int R>
__device__ __forceinline__ void fft_stage(float2* buf, const float2* tw, int Ns, float s) {
  constexpr int NT = N / 8, ITER = 8 / R, NB = N / R;
  const int tid = threadIdx.x;
  float2 v[ITER][R];
  float2 w[ITER][R];
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int k = (tid + it * NT) & (Ns - 1);
#pragma unroll
    for (int r = 1; r < R; ++r) w[it][r] = Ns > 1 ? tw[(r - 1) * Ns + k] : make_float2(1.0f, 0.0f);
#pragma unroll
    for (int r = 0; r < R; ++r) v[it][r] = buf[pad8(tid + it * NT + r * NB)];
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 t = w[it][r];  // exp(-2 pi i k r / (Ns R)); conjugated for the inverse
        t.y = -s * t.y;
        v[it][r] = cmulf(v[it][r], t);
      }
    }
    if (R == 8) dft8(v[it], s);
    else if (R == 4) dft4(v[it], s);
    else dft2(v[it], s);
  }
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int j = tid + it * NT;
    const int k = j & (Ns - 1);
    const int d = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) buf[pad8(d + r * Ns)] = v[it][r];
  }
  __syncthreads();
}

// s = -1: forward DFT; s = +1: inverse without the 1/N. The twiddles of the
// stages Ns = 8, 64 come from LDS (tws), those of the last stage from the
// device table (twg; consecutive lanes read consecutive entries).
template <int N>
__device__ __forceinline__ void fft_lds(float2* buf, const float2* tws, const float2* __restrict__ twg, float s) {
  fft_stage<N, 8>(buf, tws, 1, s);
  fft_stage<N, 8>(buf, tws + kTw8, 8, s);
  fft_stage<N, 8>(buf, tws + kTw64, 64, s);
  if (N == 1024) fft_stage<N, 2>(buf, twg + kTw512, 512, s);
  if (N == 2048) fft_stage<N, 4>(buf, twg + kTw512, 512, s);
  if (N == 4096) fft_stage<N, 8>(buf, twg + kTw512, 512, s);
}

// Two traces a, b per block. The analytic-signal operator (the weights h on
// the spectrum) is linear over complex input, so w = ifft(h .* fft(a + i b))
// = z_a + i z_b, where z_a = a + i H(a) and z_b = b + i H(b) with H(.) real:
// Re w = a - H(b), Im w = H(a) + b. The outputs' real parts are the inputs
// themselves (MATLAB's come back through the FFT round trip, equal up to
// FP32 rounding); their imaginary parts are H(a) = Im w - b, H(b) = a - Re w.
template <int N>
__global__ __launch_bounds__(N / 8) void hilbert_lds_kernel(const float* __restrict__ rf, float2* __restrict__ z,
                                                            const float2* __restrict__ twg, long long traces) {
  constexpr int NT = N / 8;
  __shared__ float2 buf[N + N / 8];
  __shared__ float2 tws[kTw512];
  const size_t ta = 2 * (size_t)blockIdx.x;
  const bool has_b = (long long)ta + 1 < traces;
  const float* __restrict__ xa = rf + ta * N;
  const float* __restrict__ xb = xa + N;
  // all 16 sample loads (and the twiddles') are issued before the first LDS
  // store of a sample, so the block waits for memory once (storing each pair
  // as it arrived took 20 % longer)
  float av[8], bv[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int i = threadIdx.x + k * NT;
    av[k] = xa[i];
    bv[k] = has_b ? xb[i] : 0.0f;
  }
  for (int i = threadIdx.x; i < kTw512; i += NT) tws[i] = twg[i];
#pragma unroll
  for (int k = 0; k < 8; ++k) buf[pad8(threadIdx.x + k * NT)] = make_float2(av[k], bv[k]);
  __syncthreads();
  fft_lds<N>(buf, tws, twg, -1.0f);
  // MATLAB hilbert's weights with ifft's 1/N (exact: powers of two)
  for (int i = threadIdx.x; i < N; i += NT) {
    const float h = (i == 0 || i == N / 2) ? 1.0f / N : (i < N / 2 ? 2.0f / N : 0.0f);
    const float2 v = buf[pad8(i)];
    buf[pad8(i)] = make_float2(v.x * h, v.y * h);
  }
  __syncthreads();
  fft_lds<N>(buf, tws, twg, 1.0f);
  float2* __restrict__ za = z + ta * N;
  float2* __restrict__ zb = za + N;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int i = threadIdx.x + k * NT;
    const float2 w = buf[pad8(i)];
    za[i] = make_float2(av[k], w.y - bv[k]);
    if (has_b) zb[i] = make_float2(bv[k], av[k] - w.x);
  }
}

// ---- LDS path, other lengths: N = 2^a 3^b 5^c 7^d <= 4096 ----
// Mixed-radix Stockham stages (radix 8 while 8 divides what is left, then 4
// or 2, then 3, 5, 7), kMixT threads per block, each holding up to
// mix_maxb(R) butterflies of a stage; twiddles exp(-2 pi i m / N) from a plain device
// table, copied into LDS at the block's start (read from L2 after each
// barrier, they cost 12 % more time); radix 3, 5, 7 as direct DFTs with
// constant cos/sin. 126 VGPRs and 68 KB of LDS: 2 blocks per CU.
constexpr int kMixT = 512;
constexpr int kMixMax = 4096;
constexpr int kMixStages = 12;

template <int R>
struct RootsR;
template <>
struct RootsR<3> {
  static constexpr float c[3] = {1.0f, -0.5f, -0.5f};
  static constexpr float s[3] = {0.0f, 0.8660254037844387f, -0.8660254037844387f};
};
template <>
struct RootsR<5> {
  static constexpr float c[5] = {1.0f, 0.30901699437494745f, -0.8090169943749475f, -0.8090169943749475f,
                                 0.30901699437494745f};
  static constexpr float s[5] = {0.0f, 0.9510565162951535f, 0.5877852522924731f, -0.5877852522924731f,
                                 -0.9510565162951535f};
};
template <>
struct RootsR<7> {
  static constexpr float c[7] = {1.0f, 0.6234898018587336f, -0.22252093395631434f, -0.900968867902419f,
                                 -0.900968867902419f, -0.22252093395631434f, 0.6234898018587336f};
  static constexpr float s[7] = {0.0f, 0.7818314824680298f, 0.9749279121818236f, 0.43388373911755823f,
                                 -0.43388373911755823f, -0.9749279121818236f, -0.7818314824680298f};
};

// X[q] = sum_r v[r] exp(s 2 pi i r q / R), R = 3, 5, 7.
template <int R>
__device__ __forceinline__ void dft_odd(float2* v, float s) {
  float2 x[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    float re = 0.0f, im = 0.0f;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = (r * q) % R;
      const float c = RootsR<R>::c[k], sn = s * RootsR<R>::s[k];
      re += v[r].x * c - v[r].y * sn;
      im += v[r].x * sn + v[r].y * c;
    }
    x[q] = make_float2(re, im);
  }
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = x[q];
}

// butterflies per thread of a radix-R stage at N <= kMixMax, kMixT threads
__host__ __device__ constexpr int mix_maxb(int R) { return (kMixMax / R + kMixT - 1) / kMixT; }

template <int R>
__device__ __forceinline__ void mix_stage(float2* buf, const float2* __restrict__ tw, int N, int Ns, float s) {
  constexpr int MB = mix_maxb(R);
  const int nb = N / R, step = N / (Ns * R);
  const int tid = threadIdx.x;
  float2 v[MB][R];
#pragma unroll
  for (int k = 0; k < MB; ++k) {
    const int j = tid + k * kMixT;
    if (j < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[k][r] = buf[pad8(j + r * nb)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < MB; ++k) {
    const int j = tid + k * kMixT;
    if (j < nb) {
      const int kk = j % Ns;
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[kk * r * step];  // exp(-2 pi i m / N); conjugated for the inverse
          w.y = -s * w.y;
          v[k][r] = cmulf(v[k][r], w);
        }
      }
      if (R == 8) dft8(v[k], s);
      else if (R == 4) dft4(v[k], s);
      else if (R == 2) dft2(v[k], s);
      else dft_odd<(R == 3 || R == 5 || R == 7) ? R : 3>(v[k], s);
      const int d = (j - kk) * R + kk;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[pad8(d + r * Ns)] = v[k][r];
    }
  }
  __syncthreads();
}

struct MixPlan {
  int n;                     // stages
  int radix[kMixStages];
};

__device__ __forceinline__ void fft_mixed(float2* buf, const float2* __restrict__ tw, int N, const MixPlan& pl,
                                          float s) {
  int Ns = 1;
  for (int i = 0; i < pl.n; ++i) {
    const int R = pl.radix[i];  // uniform
    switch (R) {
      case 8: mix_stage<8>(buf, tw, N, Ns, s); break;
      case 4: mix_stage<4>(buf, tw, N, Ns, s); break;
      case 2: mix_stage<2>(buf, tw, N, Ns, s); break;
      case 3: mix_stage<3>(buf, tw, N, Ns, s); break;
      case 5: mix_stage<5>(buf, tw, N, Ns, s); break;
      default: mix_stage<7>(buf, tw, N, Ns, s); break;
    }
    Ns *= R;
  }
}

// Two traces per block, as hilbert_lds_kernel; the weights for odd N
// (MATLAB hilbert): h[0] = 1, h[k] = 2 for 0 < k < N/2, h[N/2] = 1 for even N.
__global__ __launch_bounds__(kMixT) void hilbert_mixed_kernel(const float* __restrict__ rf, float2* __restrict__ z,
                                                              const float2* __restrict__ tw, long long traces,
                                                              int N, MixPlan pl) {
  __shared__ float2 buf[kMixMax + kMixMax / 8];
  __shared__ float2 twl[kMixMax];  // the twiddle table: no stage waits on L2
  const size_t ta = 2 * (size_t)blockIdx.x;
  const bool has_b = (long long)ta + 1 < traces;
  const float* __restrict__ xa = rf + ta * N;
  const float* __restrict__ xb = xa + N;
  // loads first (at most kMixMax / kMixT per lane; a lane past N reads
  // sample 0), then the LDS stores: the block waits for memory once
  constexpr int kPer = kMixMax / kMixT;
  float av[kPer], bv[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = threadIdx.x + k * kMixT;
    const int j = i < N ? i : 0;
    av[k] = xa[j];
    bv[k] = has_b ? xb[j] : 0.0f;
  }
  float2 tv[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = threadIdx.x + k * kMixT;
    tv[k] = tw[i < N ? i : 0];
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = threadIdx.x + k * kMixT;
    if (i < N) {
      buf[pad8(i)] = make_float2(av[k], bv[k]);
      twl[i] = tv[k];
    }
  }
  __syncthreads();
  fft_mixed(buf, twl, N, pl, -1.0f);
  const float inv = 1.0f / (float)N;
  for (int i = threadIdx.x; i < N; i += kMixT) {
    const float h = (i == 0 || 2 * i == N) ? inv : (2 * i < N ? 2.0f * inv : 0.0f);
    const float2 v = buf[pad8(i)];
    buf[pad8(i)] = make_float2(v.x * h, v.y * h);
  }
  __syncthreads();
  fft_mixed(buf, twl, N, pl, 1.0f);
  // the inputs again (L2-resident by now) for the real parts
  float2* __restrict__ za = z + ta * N;
  float2* __restrict__ zb = za + N;
  for (int i = threadIdx.x; i < N; i += kMixT) {
    const float2 w = buf[pad8(i)];
    const float a = xa[i], bb = has_b ? xb[i] : 0.0f;
    za[i] = make_float2(a, w.y - bb);
    if (has_b) zb[i] = make_float2(bb, a - w.x);
  }
}

// The stage radices of N (8s, then 4 or 2, then 3s, 5s, 7s), or false.
bool mix_plan(int N, MixPlan& pl) {
  if (N < 2 || N > kMixMax) return false;
  pl.n = 0;
  int n = N;
  auto push = [&](int r) {
    if (pl.n >= kMixStages) return false;
    pl.radix[pl.n++] = r;
    n /= r;
    return true;
  };
  while (n % 8 == 0)
    if (!push(8)) return false;
  if (n % 4 == 0 && !push(4)) return false;
  if (n % 2 == 0 && !push(2)) return false;
  for (int p : {3, 5, 7})
    while (n % p == 0)
      if (!push(p)) return false;
  return n == 1;
}

__global__ void plain_twiddle_kernel(float2* __restrict__ tw, int N) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N) return;
  double sv, cv;
  sincospi(-2.0 * (double)m / (double)N, &sv, &cv);
  tw[m] = make_float2((float)cv, (float)sv);
}

std::map<std::pair<int, int>, float2*> g_plain_twiddles;

// exp(-2 pi i m / N), m < N, of the current device: built once (and waited
// for on the caller's stream). Caller holds g_mu.
hipError_t plain_twiddles(int dev, int N, hipStream_t s, const float2** out) {
  const auto key = std::make_pair(dev, N);
  auto it = g_plain_twiddles.find(key);
  if (it != g_plain_twiddles.end()) {
    *out = it->second;
    return hipSuccess;
  }
  float2* tw = nullptr;
  hipError_t e = hipMalloc(&tw, sizeof(float2) * (size_t)N);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(plain_twiddle_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, tw, N);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(tw);
    return e;
  }
  g_plain_twiddles.emplace(key, tw);
  *out = tw;
  return hipSuccess;
}

// Entry i of the twiddle table of length N (FP64, rounded to FP32).
__global__ void twiddle_kernel(float2* __restrict__ tw, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= twiddle_count(N)) return;
  int Ns, R, q;
  if (i < kTw64) {
    Ns = 8, R = 8, q = i - kTw8;
  } else if (i < kTw512) {
    Ns = 64, R = 8, q = i - kTw64;
  } else {
    Ns = 512, R = N / 512, q = i - kTw512;
  }
  const int r = q / Ns + 1, k = q % Ns;
  double sv, cv;
  sincospi(-2.0 * (double)(k * r) / (double)(Ns * R), &sv, &cv);
  tw[i] = make_float2((float)cv, (float)sv);
}

std::map<std::pair<int, int>, float2*> g_twiddles;

// The twiddle table of length N on the current device: built once (and
// waited for on the caller's stream), then shared by every call. Caller
// holds g_mu.
hipError_t twiddles(int dev, int N, hipStream_t s, const float2** out) {
  const auto key = std::make_pair(dev, N);
  auto it = g_twiddles.find(key);
  if (it != g_twiddles.end()) {
    *out = it->second;
    return hipSuccess;
  }
  float2* tw = nullptr;
  const int n = twiddle_count(N);
  hipError_t e = hipMalloc(&tw, sizeof(float2) * (size_t)n);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, tw, N);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(tw);
    return e;
  }
  g_twiddles.emplace(key, tw);
  *out = tw;
  return hipSuccess;
}

template <int N>
hipError_t launch_hilbert_lds(const float* rf, float2* z, long long traces, const float2* tw, hipStream_t s) {
  hipLaunchKernelGGL(hilbert_lds_kernel<N>, dim3((unsigned)((traces + 1) / 2)), dim3(N / 8), 0, s, rf, z, tw,
                     traces);
  return hipGetLastError();
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
  if (L == 512 || L == 1024 || L == 2048 || L == 4096) {
    const float2* tw = nullptr;
    e = twiddles(dev, L, s, &tw);
    if (e == hipSuccess) {
      float2* z = reinterpret_cast<float2*>(analytic);
      switch (L) {
        case 512: e = launch_hilbert_lds<512>(rf, z, traces, tw, s); break;
        case 1024: e = launch_hilbert_lds<1024>(rf, z, traces, tw, s); break;
        case 2048: e = launch_hilbert_lds<2048>(rf, z, traces, tw, s); break;
        default: e = launch_hilbert_lds<4096>(rf, z, traces, tw, s); break;
      }
    }
    if (e != hipSuccess) return fail(st, USTBX_E_GPU, "HIP error: %s", hipGetErrorString(e));
    if (st) {
      st->code = USTBX_OK;
      st->message[0] = 0;
    }
    return USTBX_OK;
  }
  MixPlan pl;
  if (mix_plan(L, pl)) {
    const float2* tw = nullptr;
    e = plain_twiddles(dev, L, s, &tw);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(hilbert_mixed_kernel, dim3((unsigned)((traces + 1) / 2)), dim3(kMixT), 0, s, rf,
                         reinterpret_cast<float2*>(analytic), tw, (long long)traces, (int)L, pl);
      e = hipGetLastError();
    }
    if (e != hipSuccess) return fail(st, USTBX_E_GPU, "HIP error: %s", hipGetErrorString(e));
    if (st) {
      st->code = USTBX_OK;
      st->message[0] = 0;
    }
    return USTBX_OK;
  }
  auto key = std::make_tuple(dev, (int)L, (long long)traces, s);
  auto it = g_plans.find(key);
  if (it == g_plans.end()) {
    e = evict_plans(dev);
    if (e != hipSuccess) return fail(st, USTBX_E_GPU, "HIP error: %s", hipGetErrorString(e));
    FftPlans p;
    int n[1] = {L};
    if (hipfftPlanMany(&p.r2c, 1, n, nullptr, 1, L, nullptr, 1, L / 2 + 1, HIPFFT_R2C, (int)traces) !=
            HIPFFT_SUCCESS ||
        hipfftPlanMany(&p.c2c, 1, n, nullptr, 1, L, nullptr, 1, L, HIPFFT_C2C, (int)traces) != HIPFFT_SUCCESS) {
      release(p);
      return fail(st, USTBX_E_GPU, "hipfftPlanMany failed (L=%d, traces=%lld)", L, (long long)traces);
    }
    e = hipMalloc(&p.spec, sizeof(float2) * (size_t)traces * (L / 2 + 1));
    if (e != hipSuccess) {
      p.spec = nullptr;
      release(p);
      return fail(st, USTBX_E_GPU, "HIP error: %s", hipGetErrorString(e));
    }
    it = g_plans.emplace(key, p).first;
  }
  FftPlans& p = it->second;
  p.last_use = ++g_plan_clock;
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
