// das_geom.hip — dispatch of the geometry kernels and the materialize kernel
// that writes the FP32 matrices das.m would hand to the MEX
// (+midprocess/das.m:95,100,106,206).
#include "das_geom_impl.h"

namespace ustbx {

hipError_t launch_das_geom(const GeomLaunch& G, hipStream_t stream) {
  const int ph = (G.wd == 0.0f) ? PH_RF : (G.separable ? PH_SEP : PH_DASC);
  if (G.variant >= 2 && G.dim == MODE_BOTH) return launch_both_rxtab(G, stream);
  if (G.variant == 2 && G.dim == MODE_TX) return launch_tx_rxtab(G, stream);
  if (G.variant == 1 && G.dim == MODE_BOTH && ph == PH_SEP && G.F == 1)
    return launch_both_shared(G, stream);
  switch (G.dim) {
    case MODE_BOTH: return launch_geom_both(G, ph, G.wave_chunk, G.split, stream);
    case MODE_TX: return launch_geom_tx(G, ph, G.wave_chunk, G.split, stream);
    default: return launch_geom_rx_none(G, ph, G.wave_chunk, stream);
  }
}

// One lane per pixel; rows of the four matrices are pixel-contiguous.
__global__ __launch_bounds__(kBlock) void geom_materialize_kernel(
    GeomLaunch G, float* __restrict__ tx_apo, float* __restrict__ rx_apo,
    float* __restrict__ tx_delay, float* __restrict__ rx_delay) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= G.P) return;
  double px, py, pz, rref;
  pixel_xyz(G.scan, p, px, py, pz, rref);
  const size_t PS = (size_t)G.P;
  for (int w = 0; w < G.W; ++w) {
    const DevWave& wv = G.waves[w];
    if (tx_apo) tx_apo[(size_t)w * PS + p] = tx_apod_f32(wv, G.tx, px, py, pz, w, p, G.P);
    if (tx_delay)
      tx_delay[(size_t)w * PS + p] =
          tx_delay_f32(wv, px, py, pz, G.pw_margin, G.bpow, G.tx_delay_mat, w, p, G.P);
  }
  for (int n = 0; n < G.N; ++n) {
    const double ex = G.ex[n], ey = G.ey[n], ez = G.ez[n];
    if (rx_apo) rx_apo[(size_t)n * PS + p] = rx_apod_f32(G.rx, px, py, pz, ex, ey, ez, n, p, G.P);
    if (rx_delay) rx_delay[(size_t)n * PS + p] = rx_delay_f32(px, py, pz, ex, ey, ez, G.c, G.lens);
  }
}

// Active elements of the coherence factor per pixel (coherence_factor.m:
// 82-131) from the plan's FP32 apodization: receive weights above the
// criterium (dim receive), transmit weights (transmit), or their products
// (both), compared in FP64.
__global__ __launch_bounds__(kBlock) void geom_active_elements_kernel(GeomLaunch G, int dim, double crit,
                                                                      float* __restrict__ M) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= G.P) return;
  double px, py, pz, rref;
  pixel_xyz(G.scan, p, px, py, pz, rref);
  int cnt = 0;
  if (dim == MODE_TX) {
    for (int w = 0; w < G.W; ++w) cnt += ((double)tx_apod_f32(G.waves[w], G.tx, px, py, pz, w, p, G.P) > crit) ? 1 : 0;
  } else {
    for (int n = 0; n < G.N; ++n) {
      const double ra = (double)rx_apod_f32(G.rx, px, py, pz, G.ex[n], G.ey[n], G.ez[n], n, p, G.P);
      if (dim == MODE_RX) {
        cnt += (ra > crit) ? 1 : 0;
      } else {
        for (int w = 0; w < G.W; ++w)
          cnt += ((double)tx_apod_f32(G.waves[w], G.tx, px, py, pz, w, p, G.P) * ra > crit) ? 1 : 0;
      }
    }
  }
  M[p] = (float)cnt;
}

hipError_t launch_geom_active_elements(const GeomLaunch& G, int dim, double crit, float* M, hipStream_t stream) {
  if (G.P <= 0) return hipSuccess;
  hipLaunchKernelGGL(geom_active_elements_kernel, dim3((unsigned)((G.P + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     stream, G, dim, crit, M);
  return hipGetLastError();
}

// Transmit table of a plan: {td, ta, ta*cos(wd td), ta*sin(wd td)} per
// (wave, pixel); grid.y = wave.
__global__ __launch_bounds__(kBlock) void geom_tx_table_kernel(GeomLaunch G,
                                                               float4* __restrict__ table) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  const int w = blockIdx.y;
  if (p >= G.P || w >= G.W) return;
  double px, py, pz, rref;
  pixel_xyz(G.scan, p, px, py, pz, rref);
  const DevWave& wv = G.waves[w];
  const float ta = tx_apod_f32(wv, G.tx, px, py, pz, w, p, G.P);
  const float td = tx_delay_f32(wv, px, py, pz, G.pw_margin, G.bpow, G.tx_delay_mat, w, p, G.P);
  float sv = 0.0f, cv = 1.0f;
  if (G.wd != 0.0f) sincos_wd(G.wd, td, sv, cv);
  table[(size_t)w * G.P + p] = make_float4(td, ta, ta * cv, ta * sv);
}

hipError_t launch_geom_tx_table(const GeomLaunch& G, float4* table, hipStream_t stream) {
  const dim3 grid((unsigned)((G.P + kBlock - 1) / kBlock), (unsigned)G.W);
  hipLaunchKernelGGL(geom_tx_table_kernel, grid, dim3(kBlock), 0, stream, G, table);
  return hipGetLastError();
}

// Terms with nonzero transmit and receive weight (the terms das_c evaluates,
// das_c.cpp:366,371): per pixel (#w: ta != 0) * (#n: ra != 0), summed per
// block in LDS, then one 64-bit atomic per block.
__global__ __launch_bounds__(kBlock) void geom_count_active_kernel(GeomLaunch G,
                                                                   unsigned long long* __restrict__ total) {
  __shared__ unsigned long long part[kBlock];
  const int p = blockIdx.x * kBlock + threadIdx.x;
  unsigned long long v = 0;
  if (p < G.P) {
    double px, py, pz, rref;
    pixel_xyz(G.scan, p, px, py, pz, rref);
    unsigned long long nt = 0, nr = 0;
    for (int w = 0; w < G.W; ++w) nt += tx_apod_f32(G.waves[w], G.tx, px, py, pz, w, p, G.P) != 0.0f;
    for (int n = 0; n < G.N; ++n)
      nr += rx_apod_f32(G.rx, px, py, pz, G.ex[n], G.ey[n], G.ez[n], n, p, G.P) != 0.0f;
    v = nt * nr;
  }
  part[threadIdx.x] = v;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0 && part[0]) atomicAdd(total, part[0]);
}

hipError_t launch_count_active(const GeomLaunch& G, unsigned long long* total, hipStream_t stream) {
  const unsigned grid = (unsigned)((G.P + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(geom_count_active_kernel, dim3(grid), dim3(kBlock), 0, stream, G, total);
  return hipGetLastError();
}

hipError_t launch_geom_materialize(const GeomLaunch& G, float* tx_apo, float* rx_apo,
                                   float* tx_delay, float* rx_delay, hipStream_t stream) {
  const unsigned grid = (unsigned)((G.P + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(geom_materialize_kernel, dim3(grid), dim3(kBlock), 0, stream, G,
                     tx_apo, rx_apo, tx_delay, rx_delay);
  return hipGetLastError();
}

}  // namespace ustbx
