// das_geom_rxtab.hip — the receive-table kernel of one plan (dimension
// both) and the plan-time tables it reads (the tile body is in das_rxtab.h).
//
// Per plan (constant for a probe configuration):
//   * receive table [N][P] float4 {rd, ra, ra*cos(wd rd), ra*sin(wd rd)},
//     the same values the shared-receive kernel computes per tile;
//   * per 64-pixel tile: the lists of channels and of transmit waves with a
//     nonzero weight on some pixel of the tile;
//   * the reference phase per pixel.
// Per run (variant 3, cubic only): the Catmull-Rom (Keys a = -1/2)
// polynomial coefficients of every sample interval, so that a term costs a
// Horner evaluation instead of four weights and four complex products (at
// four times the bytes per sample).
#include "das_rxtab.h"

namespace ustbx {

// Receive table: one lane per (pixel, channel); grid.y = channel. RF plans
// (rx_compact) store float2 {rd, ra}.
__global__ __launch_bounds__(kBlock) void geom_rx_table_kernel(GeomLaunch G, float4* __restrict__ table) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  const int n = blockIdx.y;
  if (p >= G.P || n >= G.N) return;
  double px, py, pz, rref;
  pixel_xyz(G.scan, p, px, py, pz, rref);
  const double ex = G.ex[n], ey = G.ey[n], ez = G.ez[n];
  const float ra = rx_apod_f32(G.rx, px, py, pz, ex, ey, ez, n, p, G.P);
  const float rd = rx_delay_f32(px, py, pz, ex, ey, ez, G.c, G.lens);
  if (G.rx_compact) {
    reinterpret_cast<float2*>(table)[(size_t)n * G.P + p] = make_float2(rd, ra);
    return;
  }
  float sv = 0.0f, cv = 1.0f;
  if (G.wd != 0.0f) sincos_wd_hw(G.wd, rd, sv, cv);
  table[(size_t)n * G.P + p] = make_float4(rd, ra, ra * cv, ra * sv);
}

// Reference phase exp(-i*2*w0*r_ref/c) per pixel (das.m:306-309).
__global__ __launch_bounds__(kBlock) void geom_ref_table_kernel(GeomLaunch G, float2* __restrict__ ref) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= G.P) return;
  double px, py, pz, rref;
  pixel_xyz(G.scan, p, px, py, pz, rref);
  ref[p] = ref_phase_factor(G.w0_ref, G.wd, G.c32, G.ref_single, rref);
}

// Receive entry of (pixel p, channel n) for the plan-time scans: from the
// table, or computed in place with the on-the-fly receive state (the entries
// the tile body will use, so both agree on which weights are nonzero).
__device__ __forceinline__ float4 rx_at(const GeomLaunch& G, const float4* __restrict__ rxt, int n, int p,
                                        double px, double py, double pz) {
  if (G.rx_otf) return rx_otf_entry(G.otf, otf_pixel(G.otf, px, py, pz), n);
  return rx_entry(rxt, G.rx_compact, (size_t)n * G.P + p);
}

// One wave64 per 64-pixel tile: the lists of waves and channels with a
// nonzero weight on some pixel of the tile, and whether a nonzero weight or
// its delay is not finite. The channel list holds first the channels with a
// nonzero weight other than NaN on some pixel (tile_nc[tiles + tile] of
// them), then those whose nonzero weights on the tile are all NaN.
__global__ __launch_bounds__(64) void geom_tile_scan_kernel(GeomLaunch G, const float4* __restrict__ rxt,
                                                            int* __restrict__ tile_nc, int* __restrict__ tile_chl,
                                                            int* __restrict__ tile_waves, int* __restrict__ tile_nw,
                                                            int* __restrict__ tile_bad) {
  const int tile = blockIdx.x;
  const int p = tile_pixel(G.tile_ps, G.scan.n0, tile, (int)threadIdx.x);
  const bool active = p < G.P;
  const size_t PS = (size_t)G.P;
  double px = 0.0, py = 0.0, pz = 0.0, rref = 0.0;
  if (G.rx_otf) pixel_xyz(G.scan, active ? p : 0, px, py, pz, rref);
  bool bad = false;
  int cnt = 0;
  for (int w = 0; w < G.W; ++w) {
    const float4 v = active ? G.tx_table[(size_t)w * PS + p] : make_float4(0.f, 0.f, 0.f, 0.f);
    const bool on = v.y != 0.0f;  // NaN weights count, as in das_c
    if (on && !(isfinite(v.x) && isfinite(v.z) && isfinite(v.w))) bad = true;
    if (__any(on)) {
      if (threadIdx.x == 0) tile_waves[(size_t)tile * G.W + cnt] = w;
      ++cnt;
    }
  }
  int nc = 0;
  for (int n = 0; n < G.N; ++n) {
    const float4 v = active ? rx_at(G, rxt, n, p, px, py, pz) : make_float4(0.f, 0.f, 0.f, 0.f);
    const bool on = v.y != 0.0f;
    if (on && !(isfinite(v.x) && isfinite(v.z) && isfinite(v.w))) bad = true;
    if (__any(on && v.y == v.y)) {
      if (threadIdx.x == 0) tile_chl[(size_t)tile * G.N + nc] = n;
      ++nc;
    }
  }
  const int nfin = nc;
  for (int n = 0; n < G.N; ++n) {
    const float4 v = active ? rx_at(G, rxt, n, p, px, py, pz) : make_float4(0.f, 0.f, 0.f, 0.f);
    const bool on = v.y != 0.0f;
    if (__any(on) && !__any(on && v.y == v.y)) {
      if (threadIdx.x == 0) tile_chl[(size_t)tile * G.N + nc] = n;
      ++nc;
    }
  }
  const bool any_bad = __any(bad);
  if (threadIdx.x == 0) {
    tile_nc[tile] = nc;
    tile_nc[gridDim.x + tile] = nfin;
    tile_nw[tile] = cnt;
    tile_bad[tile] = any_bad ? 1 : 0;
  }
}

// Per tile, the lanes whose das_c sum is NaN through the channels whose
// nonzero weights on the tile are all NaN (the tail of the tile's channel
// list): some term of such a channel with a nonzero transmit weight is in the
// sample window, by the tile body's own index arithmetic. Geometry only, so
// a plan constant. One wave64 per tile, after geom_tile_scan_kernel.
__global__ __launch_bounds__(64) void geom_nan_mask_kernel(GeomLaunch G, const float4* __restrict__ rxt,
                                                           const int* __restrict__ tile_nc,
                                                           const int* __restrict__ tile_chl,
                                                           const int* __restrict__ tile_waves,
                                                           const int* __restrict__ tile_nw,
                                                           unsigned long long* __restrict__ mask) {
  const int tile = blockIdx.x;
  const int p = tile_pixel(G.tile_ps, G.scan.n0, tile, (int)threadIdx.x);
  const bool active = p < G.P;
  const size_t PS = (size_t)G.P;
  double px = 0.0, py = 0.0, pz = 0.0, rref = 0.0;
  if (G.rx_otf) pixel_xyz(G.scan, active ? p : 0, px, py, pz, rref);
  const int nc = tile_nc[tile], nfin = tile_nc[gridDim.x + tile], nw = tile_nw[tile];
  const float lim = (float)(G.interp == INTERP_LINEAR ? G.L - 1 : G.L - 2);
  bool any = false;
  for (int k = nfin; k < nc; ++k) {
    const int n = tile_chl[(size_t)tile * G.N + k];
    const float4 rv = active ? rx_at(G, rxt, n, p, px, py, pz) : make_float4(0.f, 0.f, 0.f, 0.f);
    if (rv.y == 0.0f) continue;
    for (int i = 0; i < nw && !any; ++i) {
      const int w = tile_waves[(size_t)tile * G.W + i];
      const float4 tv = active ? G.tx_table[(size_t)w * PS + p] : make_float4(0.f, 0.f, 0.f, 0.f);
      if (tv.y == 0.0f) continue;
      float d;
      {
#pragma clang fp contract(off)
        d = ((tv.x + rv.x) - G.t0) * G.fs;  // das_c.cpp:372,375
      }
      any = d >= 1.0f && d < lim;
    }
  }
  const unsigned long long m = __ballot(any);
  if (threadIdx.x == 0) mask[tile] = m;
}

// Per tile and channel-list position, whether the channel's terms on the
// tile can run staged (rxtab_channel_staged): every lane of the tile active;
// the channel's receive weight and every transmit weight of the tile's waves
// nonzero on every lane, with finite delays; every term's index in the sample
// window; and every lane's n0 in [n0 of lane 0, that + kMaxSpan], by the tile
// body's own index arithmetic (term_index). Tiles with a non-finite nonzero
// weight or delay (run by the SAFE body) get no bits. One wave64 per tile.
template <int SM>
__global__ __launch_bounds__(64) void geom_stage_mask_kernel(GeomLaunch G, const int* __restrict__ tile_bad,
                                                             unsigned long long* __restrict__ mask) {
  const int tile = blockIdx.x;
  const int lane = (int)threadIdx.x;
  const int p = tile * 64 + lane;
  unsigned long long* out = mask + (size_t)tile * G.stage_words;
  const bool run = tile_bad[tile] == 0 && tile * 64 + 63 < G.P;
  const int nc = run ? G.tile_nc[tile] : 0;
  const int nw = G.tile_nw[tile];
  const int* __restrict__ chl = G.tile_chl + (size_t)tile * G.N;
  const int* __restrict__ wl = G.tile_waves + (size_t)tile * G.W;
  const size_t PS = (size_t)G.P;
  const float lim = (float)(SM == SM_LINEAR ? G.L - 1 : G.L - 2);
  unsigned long long word = 0ull;
  for (int k = 0; k < G.stage_words * 64; ++k) {
    if (k < nc) {
      const float4 rv = rx_entry(G.rx_table, G.rx_compact, (size_t)chl[k] * PS + p);
      bool ok = rv.y != 0.0f && isfinite(rv.x);
      for (int i = 0; i < nw && __all(ok); ++i) {
        const float4 tv = G.tx_table[(size_t)wl[i] * PS + p];
        const float d = term_index(tv.x, rv.x, G.t0, G.fs);
        const int n0 = (int)d;
        const int n00 = __builtin_amdgcn_readfirstlane(n0);
        ok = ok && tv.y != 0.0f && isfinite(tv.x) && d >= 1.0f && d < lim && n0 >= n00 &&
             n0 - n00 <= StageWin<SM>::kMaxSpan;
      }
      if (__all(ok)) word |= 1ull << (k & 63);
    }
    if ((k & 63) == 63) {
      if (lane == 0) out[k >> 6] = word;
      word = 0ull;
    }
  }
}

// Dimension transmit: per tile, bit n (channel n, as the transmit kernel
// indexes them) when channel n's terms on the tile can run staged
// (tx_channel): for every wave of the tile's list, with the transmit kernel's
// own term arithmetic (a zero weight's delay replaced by kFar, term_index,
// ok = the index in the sample window), every ok lane's n0 lies in [nref -
// kTxSlack, nref + kMaxSpan - kTxSlack], nref being the n0 of the first ok
// lane. Tiles with a non-finite nonzero weight or delay get no bits. One
// wave64 per tile.
template <int SM>
__global__ __launch_bounds__(64) void geom_tx_stage_mask_kernel(GeomLaunch G, const int* __restrict__ tile_bad,
                                                                unsigned long long* __restrict__ mask) {
  const int tile = blockIdx.x;
  const int lane = (int)threadIdx.x;
  const int p = tile_pixel(G.tile_ps, G.scan.n0, tile, lane);
  const bool active = p < G.P;
  unsigned long long* out = mask + (size_t)tile * G.stage_words;
  const int nw = G.tile_nw[tile];
  const bool run = tile_bad[tile] == 0;
  const int* __restrict__ wl = G.tile_waves + (size_t)tile * G.W;
  const size_t PS = (size_t)G.P;
  const float lim = (float)(SM == SM_LINEAR ? G.L - 1 : G.L - 2);
  unsigned long long word = 0ull;
  for (int n = 0; n < G.stage_words * 64; ++n) {
    if (run && n < G.N) {
      const float4 rv = active ? rx_entry(G.rx_table, G.rx_compact, (size_t)n * PS + p) : make_float4(0.f, 0.f, 0.f, 0.f);
      const float rd = rv.y != 0.0f ? rv.x : kFar;
      bool good = true;
      for (int i = 0; i < nw && good; ++i) {
        const float4 tv = active ? G.tx_table[(size_t)wl[i] * PS + p] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float d = term_index(tv.y != 0.0f ? tv.x : kFar, rd, G.t0, G.fs);
        const bool ok = d >= 1.0f && d < lim;
        const int n0 = ok ? (int)d : 0;
        const unsigned long long okm = __ballot(ok);
        const unsigned nref = okm ? (unsigned)__shfl(n0, (int)__builtin_ctzll(okm)) : 0u;
        const unsigned rel = (unsigned)n0 - nref + (unsigned)kTxSlack;  // in [0, kMaxSpan] where staged
        good = __all(!ok || rel <= (unsigned)StageWin<SM>::kMaxSpan);
      }
      if (good) word |= 1ull << (n & 63);
    }
    if ((n & 63) == 63) {
      if (lane == 0) out[n >> 6] = word;
      word = 0ull;
    }
  }
}

hipError_t launch_stage_mask(const GeomLaunch& G, const int* tile_bad, unsigned long long* stage_mask,
                             hipStream_t stream) {
  if (!G.stage_words || !G.rx_table || !(G.dim == 3 || G.dim == 2) || G.rx_otf) return hipErrorInvalidValue;
  const unsigned tiles = (unsigned)((G.P + 63) / 64);
  if (G.dim == 2) {
    if (G.interp == INTERP_CUBIC)
      hipLaunchKernelGGL(geom_tx_stage_mask_kernel<SM_CUBIC>, dim3(tiles), dim3(64), 0, stream, G, tile_bad,
                         stage_mask);
    else
      hipLaunchKernelGGL(geom_tx_stage_mask_kernel<SM_LINEAR>, dim3(tiles), dim3(64), 0, stream, G, tile_bad,
                         stage_mask);
    return hipGetLastError();
  }
  if (G.interp == INTERP_CUBIC)
    hipLaunchKernelGGL(geom_stage_mask_kernel<SM_CUBIC>, dim3(tiles), dim3(64), 0, stream, G, tile_bad, stage_mask);
  else
    hipLaunchKernelGGL(geom_stage_mask_kernel<SM_LINEAR>, dim3(tiles), dim3(64), 0, stream, G, tile_bad,
                       stage_mask);
  return hipGetLastError();
}

// Catmull-Rom coefficients of the interval [n, n+1] of every trace:
// p(b) = a0 + a1 b + a2 b^2 + a3 b^3 through samples n-1 .. n+2, the same
// polynomial as Keys' kernel with a = -1/2. Stored per sample as two float4,
// {a0, a1} and {a2, a3} (complex each); zero outside 1 <= n <= L-3, where no
// term reads. grid.x = trace * lblocks + sample block.
__global__ __launch_bounds__(kBlock) void spline_coef_kernel(const float2* __restrict__ ch, float4* __restrict__ coef,
                                                             int L, int lblocks) {
  const int trace = blockIdx.x / lblocks;
  const int n = (blockIdx.x - trace * lblocks) * kBlock + threadIdx.x;
  if (n >= L) return;
  const size_t i = (size_t)trace * L + n;
  float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0;
  if (n >= 1 && n <= L - 3) {
    const float2 s0 = ch[i - 1], s1 = ch[i], s2 = ch[i + 1], s3 = ch[i + 2];
    const double a1x = 0.5 * ((double)s2.x - s0.x), a1y = 0.5 * ((double)s2.y - s0.y);
    const double a2x = (double)s0.x - 2.5 * s1.x + 2.0 * s2.x - 0.5 * s3.x;
    const double a2y = (double)s0.y - 2.5 * s1.y + 2.0 * s2.y - 0.5 * s3.y;
    const double a3x = 0.5 * ((double)s3.x - s0.x) + 1.5 * ((double)s1.x - s2.x);
    const double a3y = 0.5 * ((double)s3.y - s0.y) + 1.5 * ((double)s1.y - s2.y);
    q0 = make_float4(s1.x, s1.y, (float)a1x, (float)a1y);
    q1 = make_float4((float)a2x, (float)a2y, (float)a3x, (float)a3y);
  }
  coef[2 * i] = q0;
  coef[2 * i + 1] = q1;
}

// dimension.both from a plan's resident tables: one block per (tile, frame).
// LDS slots of the staged channels: RB windows per split (none for
// coefficient data, which is never staged).
template <int SM, int WS>
struct StageLds {
  static constexpr int kFloat2 = SM == SM_COEF ? 1 : WS * StageBatch<SM>::RB * kStageWin;
};

template <int SM, int WC, int WS, int RXL>
__global__ __launch_bounds__(64 * WS) void das_both_rxtab_kernel(RxAcq A, const int* __restrict__ order) {
  __shared__ float2 red[WS][64];
  __shared__ float2 win[StageLds<SM, WS>::kFloat2];
  int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  if (order) tile = __builtin_amdgcn_readfirstlane(order[tile]);  // block-uniform
  rxtab_tile<SM, WC, WS, false, RXL, true>(A, tile, (int)blockIdx.y, red, win);
}

// The same where some tiles hold a non-finite nonzero weight or delay (the
// visiting order's entries good .. good + bad - 1): blocks 0 .. bad - 1 run
// those with the SAFE body, dispatched first so that their long channel loops
// overlap the rest; blocks bad8 .. (bad rounded up to whole XCD rounds) take
// the other tiles through the XCD remap; blocks bad .. bad8 - 1 have none.
template <int SM, int WC, int WS, int RXL>
__global__ __launch_bounds__(64 * WS) void das_both_rxtab_mixed_kernel(RxAcq A, const int* __restrict__ order,
                                                                       int good, int bad, int bad8) {
  __shared__ float2 red[WS][64];
  __shared__ float2 win[StageLds<SM, WS>::kFloat2];
  const int b = (int)blockIdx.x;
  if (b < bad8) {
    if (b >= bad) return;  // whole block, before any barrier
    const int tile = __builtin_amdgcn_readfirstlane(order[good + b]);
    rxtab_tile<SM, WC, WS, true, RXL>(A, tile, (int)blockIdx.y, red);
  } else {
    const int tile = __builtin_amdgcn_readfirstlane(order[xcd_tile(b - bad8, good)]);
    rxtab_tile<SM, WC, WS, false, RXL, true>(A, tile, (int)blockIdx.y, red, win);
  }
}

// dimension.both with the receive state on the fly (variant 4): one block per
// (tile, frame). Blocks 0 .. bad - 1 take the tiles with a non-finite nonzero
// weight or delay (the visiting order's entries good ..) with the SAFE body,
// blocks bad .. bad8 - 1 none, the rest the others through the XCD remap.
template <int SM, int WC, int WS>
__global__ __launch_bounds__(64 * WS) void das_both_rxotf_kernel(RxAcq A, const int* __restrict__ order, int good,
                                                                 int bad, int bad8) {
  __shared__ float2 red[WS][64];
  __shared__ float4 rxs[kOtfChunk][64];
  const int b = (int)blockIdx.x;
  if (b < bad8) {
    if (b >= bad) return;  // whole block, before any barrier
    const int tile = __builtin_amdgcn_readfirstlane(order[good + b]);
    rxotf_tile<SM, WC, WS, true>(A, tile, (int)blockIdx.y, red, rxs);
  } else {
    int tile = xcd_tile(b - bad8, good);
    if (order) tile = __builtin_amdgcn_readfirstlane(order[tile]);
    rxotf_tile<SM, WC, WS, false>(A, tile, (int)blockIdx.y, red, rxs);
  }
}

template <int SM, int WC, int WS>
static hipError_t launch_otf_t(const GeomLaunch& G, hipStream_t stream) {
  const int tiles = (G.P + 63) / 64;
  const int good = G.tile_good, bad = tiles - good;
  if (bad > 0 && !G.tile_order) return hipErrorInvalidValue;
  const int bad8 = bad > 0 ? (bad + 7) & ~7 : 0;
  const RxAcq A = rx_acq(G, SM);
  hipLaunchKernelGGL((das_both_rxotf_kernel<SM, WC, WS>), dim3((unsigned)(bad8 + good), (unsigned)G.F),
                     dim3(64 * WS), 0, stream, A, G.tile_order, good, bad, bad8);
  return hipGetLastError();
}

// Plans whose weights are all finite launch the branch-free kernel; the
// others the mixed one (same tile shape for both bodies, so that a batch,
// which runs either body per tile, computes the same sums).
template <int SM, int WC, int WS, int RXL>
static hipError_t launch_rxtab_l(const GeomLaunch& G, hipStream_t stream) {
  const int tiles = (G.P + 63) / 64;
  const int good = G.tile_good, bad = tiles - good;
  const RxAcq A = rx_acq(G, SM);
  if (bad == 0) {
    hipLaunchKernelGGL((das_both_rxtab_kernel<SM, WC, WS, RXL>), dim3((unsigned)tiles, (unsigned)G.F),
                       dim3(64 * WS), 0, stream, A, G.tile_order);
    return hipGetLastError();
  }
  if (!G.tile_order) return hipErrorInvalidValue;
  const int bad8 = (bad + 7) & ~7;
  hipLaunchKernelGGL((das_both_rxtab_mixed_kernel<SM, WC, WS, RXL>), dim3((unsigned)(bad8 + good), (unsigned)G.F),
                     dim3(64 * WS), 0, stream, A, G.tile_order, good, bad, bad8);
  return hipGetLastError();
}

template <int SM, int WC, int WS>
static hipError_t launch_rxtab_t(const GeomLaunch& G, hipStream_t stream) {
  if (SM != SM_COEF && G.rx_otf) return launch_otf_t<SM == SM_COEF ? SM_LINEAR : SM, WC, WS>(G, stream);
  return G.rx_compact ? launch_rxtab_l<SM, WC, WS, RXL_COMPACT>(G, stream)
                      : launch_rxtab_l<SM, WC, WS, RXL_FULL>(G, stream);
}

template <int SM, int WS>
static hipError_t launch_rxtab_wc(const GeomLaunch& G, hipStream_t stream) {
  switch (G.wave_chunk) {
    case 4: return launch_rxtab_t<SM, 4, WS>(G, stream);
    case 5: return launch_rxtab_t<SM, 5, WS>(G, stream);
    case 6: return launch_rxtab_t<SM, 6, WS>(G, stream);
    case 8: return launch_rxtab_t<SM, 8, WS>(G, stream);
    case 10: return launch_rxtab_t<SM, 10, WS>(G, stream);
    case 12: return launch_rxtab_t<SM, 12, WS>(G, stream);
    case 16: return launch_rxtab_t<SM, 16, WS>(G, stream);
    default: return hipErrorInvalidValue;
  }
}

template <int SM>
static hipError_t launch_rxtab_ws(const GeomLaunch& G, hipStream_t stream) {
  switch (G.split) {
    case 4: return launch_rxtab_wc<SM, 4>(G, stream);
    case 5: return launch_rxtab_wc<SM, 5>(G, stream);
    case 8: return launch_rxtab_wc<SM, 8>(G, stream);
    default: return hipErrorInvalidValue;
  }
}

// Variant 2 gathers channel data; variant 3 (cubic only) first rewrites it
// as interpolation coefficients.
hipError_t launch_both_rxtab(const GeomLaunch& G, hipStream_t stream) {
  if (G.interp != INTERP_CUBIC) return launch_rxtab_ws<SM_LINEAR>(G, stream);
  if (G.variant != 3) return launch_rxtab_ws<SM_CUBIC>(G, stream);
  const int lblocks = (G.L + kBlock - 1) / kBlock;
  const long long traces = (long long)G.F * G.W * G.N;
  if (!G.coef || traces * lblocks > 0x7FFFFFFFLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(spline_coef_kernel, dim3((unsigned)(traces * lblocks)), dim3(kBlock), 0, stream,
                     reinterpret_cast<const float2*>(G.ch), G.coef, G.L, lblocks);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_rxtab_ws<SM_COEF>(G, stream);
}

// The element positions rounded to FP32 for the on-the-fly receive state.
__global__ __launch_bounds__(kBlock) void otf_elements_kernel(const double* __restrict__ ex,
                                                              const double* __restrict__ ey,
                                                              const double* __restrict__ ez, float4* __restrict__ e32,
                                                              int N) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n < N) e32[n] = make_float4((float)ex[n], (float)ey[n], (float)ez[n], 0.0f);
}

hipError_t launch_otf_elements(const GeomLaunch& G, float4* e32, hipStream_t stream) {
  hipLaunchKernelGGL(otf_elements_kernel, dim3((unsigned)((G.N + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream,
                     G.ex, G.ey, G.ez, e32, G.N);
  return hipGetLastError();
}

hipError_t launch_rxtab_tables(const GeomLaunch& G, float4* rx_table, float2* ref_table, int* tile_nc,
                               int* tile_chl, int* tile_waves, int* tile_nw, int* tile_bad,
                               unsigned long long* nan_mask, hipStream_t stream) {
  hipError_t e = hipSuccess;
  if (rx_table) {
    const dim3 rgrid((unsigned)((G.P + kBlock - 1) / kBlock), (unsigned)G.N);
    hipLaunchKernelGGL(geom_rx_table_kernel, rgrid, dim3(kBlock), 0, stream, G, rx_table);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  } else if (!G.rx_otf) {
    return hipErrorInvalidValue;
  }
  if (ref_table) {
    hipLaunchKernelGGL(geom_ref_table_kernel, dim3((unsigned)((G.P + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       stream, G, ref_table);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const unsigned tiles = (unsigned)((G.P + 63) / 64);
  hipLaunchKernelGGL(geom_tile_scan_kernel, dim3(tiles), dim3(64), 0, stream, G, rx_table, tile_nc, tile_chl,
                     tile_waves, tile_nw, tile_bad);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(geom_nan_mask_kernel, dim3(tiles), dim3(64), 0, stream, G, rx_table, tile_nc, tile_chl,
                     tile_waves, tile_nw, nan_mask);
  return hipGetLastError();
}

}  // namespace ustbx
