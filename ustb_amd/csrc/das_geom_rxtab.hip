// das_geom_rxtab.hip — the receive-table kernel (dimension both) and the
// plan-time tables it reads.
//
// Per plan (constant for a probe configuration):
//   * receive table [N][P] float4 {rd, ra, ra*cos(wd rd), ra*sin(wd rd)},
//     the same values the shared-receive kernel computes per tile;
//   * per 64-pixel tile: the channel range holding every nonzero receive
//     weight, and the list of transmit waves with a nonzero transmit weight
//     on some pixel of the tile;
//   * the reference phase per pixel.
// Per run: for cubic interpolation, the Catmull-Rom (Keys a = -1/2)
// polynomial coefficients of every sample interval, so that a term costs a
// Horner evaluation instead of four weights and four complex products.
//
// The main kernel holds no FP64 geometry, so it runs at high occupancy, and
// its term loop has no data-dependent branch: a term outside the sample
// window reads through an out-of-range buffer offset, which the hardware
// returns as zeros. That is exact only when every nonzero weight and delay
// is finite (das_c skips a zero-weight or out-of-window term before it
// multiplies, so a NaN weight on a skipped term must not reach the sum);
// the plan checks this when it builds the tables and otherwise keeps the
// exact kernels.
#include "das_geom_impl.h"

namespace ustbx {

namespace {

// Delay that puts every term it enters outside the sample window: finite,
// so that the fraction of the (huge, integral) index stays 0.
constexpr float kFar = -1.0e20f;
// Byte offset beyond any buffer the plan accepts (frames are capped at
// 0x7FFFFE00 bytes), read back as zeros.
constexpr int kOOB = 0x7FFFFF00;

}  // namespace

// Receive table: one lane per (pixel, channel); grid.y = channel.
__global__ __launch_bounds__(kBlock) void geom_rx_table_kernel(GeomLaunch G, float4* __restrict__ table) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  const int n = blockIdx.y;
  if (p >= G.P || n >= G.N) return;
  double px, py, pz, rref;
  pixel_xyz(G.scan, p, px, py, pz, rref);
  const double ex = G.ex[n], ey = G.ey[n], ez = G.ez[n];
  const float ra = rx_apod_f32(G.rx, px, py, pz, ex, ey, ez, n, p, G.P);
  const float rd = rx_delay_f32(px, py, pz, ex, ey, ez, G.c, G.lens);
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
  double sv, cv;
  sincos(-(G.w0_ref * rref), &sv, &cv);
  ref[p] = make_float2((float)cv, (float)sv);
}

// One wave64 per 64-pixel tile: active-wave list, channel range, and whether
// a nonzero weight or its delay is not finite.
__global__ __launch_bounds__(64) void geom_tile_scan_kernel(GeomLaunch G, const float4* __restrict__ rxt,
                                                            int2* __restrict__ tile_ch, int* __restrict__ tile_waves,
                                                            int* __restrict__ tile_nw, int* __restrict__ tile_bad) {
  const int tile = blockIdx.x;
  const int p = tile * 64 + (int)threadIdx.x;
  const bool active = p < G.P;
  const size_t PS = (size_t)G.P;
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
  int lo = G.N, hi = 0;
  for (int n = 0; n < G.N; ++n) {
    const float4 v = active ? rxt[(size_t)n * PS + p] : make_float4(0.f, 0.f, 0.f, 0.f);
    const bool on = v.y != 0.0f;
    if (on && !(isfinite(v.x) && isfinite(v.z) && isfinite(v.w))) bad = true;
    if (__any(on)) {
      lo = min(lo, n);
      hi = n + 1;
    }
  }
  if (lo >= hi) lo = hi = 0;
  const bool any_bad = __any(bad);
  if (threadIdx.x == 0) {
    tile_ch[tile] = make_int2(lo, hi);
    tile_nw[tile] = cnt;
    tile_bad[tile] = any_bad ? 1 : 0;
  }
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

// Sampling modes of the term loop:
//   SM_LINEAR: channel data, one 16-byte gather {x[n0], x[n0+1]};
//   SM_CUBIC:  channel data, two 16-byte gathers {x[n0-1], x[n0]},
//              {x[n0+1], x[n0+2]} and Keys' weights (a = -1/2);
//   SM_COEF:   the Catmull-Rom coefficient buffer, two 16-byte gathers and a
//              Horner evaluation (fewer operations, four times the bytes).
enum { SM_LINEAR = 0, SM_CUBIC = 1, SM_COEF = 2 };

template <int SM>
struct SampleMode {
  static constexpr int kStride = SM == SM_COEF ? 32 : 8;  // bytes per sample
  static constexpr int kLoads = SM == SM_LINEAR ? 1 : 2;
  static constexpr int kLo = SM == SM_CUBIC ? -8 : 0;     // byte offsets of the gathers
  static constexpr int kHi = SM == SM_CUBIC ? 8 : 16;
};

// Keys' cubic weights (a = -1/2) in Horner form.
__device__ __forceinline__ void keys_weights_horner(float b, float& c0, float& c1, float& c2, float& c3) {
  const float b2 = b * b;
  c0 = b * fmaf(b, fmaf(-0.5f, b, 1.0f), -0.5f);
  c1 = fmaf(b2, fmaf(1.5f, b, -2.5f), 1.0f);
  c2 = b * fmaf(b, fmaf(-1.5f, b, 2.0f), 0.5f);
  c3 = b2 * fmaf(0.5f, b, -0.5f);
}

// One channel of one split: part = sum_j T_j * sample_j over the split's
// waves, in batches of RB terms whose gathers are all issued before the
// first is used. Slots j >= nj are skipped when GUARD (a wave-uniform branch
// per batch); without it the slots past the tile's waves hold the far delay
// and read zeros.
constexpr int RB = 5;

template <int SM, int WC, bool GUARD>
__device__ __forceinline__ float2 rxtab_channel(__amdgpu_buffer_rsrc_t rs, const float (&td)[WC],
                                                const float2 (&T)[WC], const int (&wb)[WC], int nj, float rd,
                                                int nb, float t0, float fs, float lo_ok, float lim) {
  using M = SampleMode<SM>;
  float2 part = make_float2(0.0f, 0.0f);
#pragma unroll
  for (int j0 = 0; j0 < WC; j0 += RB) {
    if (GUARD && j0 >= nj) break;
    float4 q[RB][M::kLoads];
    float bb[RB];
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int j = j0 + i;
      if (j < WC) {
        float d;
        {
#pragma clang fp contract(off)
          d = ((td[j] + rd) - t0) * fs;  // das_c.cpp:372,375
        }
        const bool ok = d >= lo_ok && d < lim;
        const int n0 = (int)d;
        bb[i] = __builtin_amdgcn_fractf(d);
        const int off = ok ? n0 * M::kStride + (nb + wb[j]) : kOOB;
        q[i][0] = buf_load16(rs, off + M::kLo);
        if (M::kLoads == 2) q[i][M::kLoads - 1] = buf_load16(rs, off + M::kHi);
      }
    }
    // keep the batch's gathers ahead of their uses (the scheduler would
    // otherwise pair each gather with its use to save registers)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int j = j0 + i;
      if (j < WC) {
        const float b = bb[i];
        float2 sm;
        if (SM == SM_LINEAR) {
          const float4 x = q[i][0];
          sm = make_float2(fmaf(b, x.z - x.x, x.x), fmaf(b, x.w - x.y, x.y));
        } else if (SM == SM_CUBIC) {
          const float4 lo = q[i][0], hi = q[i][M::kLoads - 1];
          float c0, c1, c2, c3;
          keys_weights_horner(b, c0, c1, c2, c3);
          sm = make_float2(fmaf(c3, hi.z, fmaf(c2, hi.x, fmaf(c1, lo.z, c0 * lo.x))),
                           fmaf(c3, hi.w, fmaf(c2, hi.y, fmaf(c1, lo.w, c0 * lo.y))));
        } else {
          const float4 q0 = q[i][0], q1 = q[i][M::kLoads - 1];
          sm = make_float2(fmaf(fmaf(fmaf(q1.z, b, q1.x), b, q0.z), b, q0.x),
                           fmaf(fmaf(fmaf(q1.w, b, q1.y), b, q0.w), b, q0.y));
        }
        cfma(part, T[j], sm);
      }
    }
    // one batch in flight per wave: the batch's sum is complete here (an empty
    // asm that takes it), so no pass can defer this batch's arithmetic past
    // the next batch's gathers
    asm volatile("" : "+v"(part.x), "+v"(part.y));
    __builtin_amdgcn_sched_barrier(0);
  }
  return part;
}

template <int SM, int WC, bool GUARD>
__device__ __forceinline__ void rxtab_channels(__amdgpu_buffer_rsrc_t rs, const float (&td)[WC],
                                               const float2 (&T)[WC], const int (&wb)[WC], int nj,
                                               const float4* __restrict__ rxt, size_t PS, int p, bool active,
                                               int2 cr, int lst, float t0, float fs, float lo_ok, float lim,
                                               float2& acc) {
  float4 rn = make_float4(0.f, 0.f, 0.f, 0.f);
  if (active) rn = rxt[(size_t)cr.x * PS + p];
  for (int n = cr.x; n < cr.y; ++n) {
    const float4 rv = rn;
    if (active && n + 1 < cr.y) rn = rxt[(size_t)(n + 1) * PS + p];
    const bool ron = rv.y != 0.0f;
    const float rd = ron ? rv.x : kFar;
    const float2 R = ron ? make_float2(rv.z, rv.w) : make_float2(0.0f, 0.0f);
    const float2 part = rxtab_channel<SM, WC, GUARD>(rs, td, T, wb, nj, rd, n * lst, t0, fs, lo_ok, lim);
    cfma(acc, R, part);
  }
}

// ---------------------------------------------------------------------------
// dimension.both from the resident tables.
//
// A block is one tile of 64 pixels (lanes) x WS wave-splits. Split s takes
// the tile's active waves k = k0 + s + WS*j (j < WC), with their transmit
// delay and weight T in VGPRs, and runs over the tile's channel range,
// reading the receive row of the next channel ahead. Per term: the sample
// index in das_c's operation order, one or two 16-byte gathers, the
// interpolated sample, then part += T * sample; per channel acc += R * part.
// The WS partial sums of a pixel are added in split order.
// ---------------------------------------------------------------------------
template <int SM, int WC, int WS>
__global__ __launch_bounds__(64 * WS) void das_both_rxtab_kernel(GeomLaunch G) {
  constexpr int TP = 64;
  constexpr int ST = SampleMode<SM>::kStride;
  __shared__ float2 red[WS][TP];
  const int lane = threadIdx.x & (TP - 1);
  const int s = __builtin_amdgcn_readfirstlane((int)threadIdx.x / TP);  // wave-uniform
  int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  if (G.tile_order) tile = G.tile_order[tile];
  const int f = blockIdx.y;
  const int p = tile * TP + lane;
  const bool active = p < G.P;
  const size_t PS = (size_t)G.P;
  const int2 cr = G.tile_ch[tile];
  const int nw = G.tile_nw[tile];
  const int* __restrict__ wl = G.tile_waves + (size_t)tile * G.W;
  const float4* __restrict__ txt = G.tx_table;
  const int lst = G.L * ST;           // bytes per trace
  const int wst = lst * G.N;          // bytes per wave
  const int frame_bytes = wst * G.W;  // < 0x7FFFFE00, checked by the plan
  const char* fbase = (SM == SM_COEF ? (const char*)G.coef : (const char*)G.ch) + (size_t)f * frame_bytes;
  const __amdgpu_buffer_rsrc_t rs = make_rsrc(fbase, frame_bytes);
  const float lim = (float)(SM == SM_LINEAR ? G.L - 1 : G.L - 2);
  const float lo_ok = 1.0f;  // das_c's window (the plan excludes the Python edge rule here)

  float2 acc = make_float2(0.0f, 0.0f);
  for (int k0 = 0; k0 < nw; k0 += WS * WC) {
    // number of this split's waves in the round (wave-uniform)
    const int nj = min(WC, max(0, (nw - k0 - s + WS - 1) / WS));
    if (nj == 0) break;
    float td[WC];
    float2 T[WC];
    int wb[WC];
#pragma unroll
    for (int j = 0; j < WC; ++j) {
      td[j] = kFar;
      T[j] = make_float2(0.0f, 0.0f);
      wb[j] = 0;
      if (j < nj) {
        const int w = wl[k0 + s + WS * j];
        wb[j] = w * wst;
        if (active) {
          const float4 v = txt[(size_t)w * PS + p];
          if (v.y != 0.0f) {
            td[j] = v.x;
            T[j] = make_float2(v.z, v.w);
          }
        }
      }
    }
    if (nj == WC)
      rxtab_channels<SM, WC, false>(rs, td, T, wb, nj, G.rx_table, PS, p, active, cr, lst, G.t0, G.fs, lo_ok,
                                        lim, acc);
    else
      rxtab_channels<SM, WC, true>(rs, td, T, wb, nj, G.rx_table, PS, p, active, cr, lst, G.t0, G.fs, lo_ok,
                                       lim, acc);
  }
  red[s][lane] = acc;
  __syncthreads();
  if (s == 0 && active) {
    float2 v = red[0][lane];
#pragma unroll
    for (int k = 1; k < WS; ++k) cadd(v, red[k][lane]);
    if (G.ref_phase) v = cmul(v, G.ref_table[p]);
    reinterpret_cast<float2*>(G.out)[(size_t)f * PS + p] = v;
  }
}

template <int SM, int WC, int WS>
static hipError_t launch_rxtab_t(const GeomLaunch& G, hipStream_t stream) {
  const dim3 grid((unsigned)((G.P + 63) / 64), (unsigned)G.F);
  hipLaunchKernelGGL((das_both_rxtab_kernel<SM, WC, WS>), grid, dim3(64 * WS), 0, stream, G);
  return hipGetLastError();
}

template <int SM, int WS>
static hipError_t launch_rxtab_wc(const GeomLaunch& G, hipStream_t stream) {
  switch (G.wave_chunk) {
    case 4: return launch_rxtab_t<SM, 4, WS>(G, stream);
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

hipError_t launch_rxtab_tables(const GeomLaunch& G, float4* rx_table, float2* ref_table, int2* tile_ch,
                               int* tile_waves, int* tile_nw, int* tile_bad, hipStream_t stream) {
  const dim3 rgrid((unsigned)((G.P + kBlock - 1) / kBlock), (unsigned)G.N);
  hipLaunchKernelGGL(geom_rx_table_kernel, rgrid, dim3(kBlock), 0, stream, G, rx_table);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (ref_table) {
    hipLaunchKernelGGL(geom_ref_table_kernel, dim3((unsigned)((G.P + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       stream, G, ref_table);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const unsigned tiles = (unsigned)((G.P + 63) / 64);
  hipLaunchKernelGGL(geom_tile_scan_kernel, dim3(tiles), dim3(64), 0, stream, G, rx_table, tile_ch, tile_waves,
                     tile_nw, tile_bad);
  return hipGetLastError();
}

}  // namespace ustbx
