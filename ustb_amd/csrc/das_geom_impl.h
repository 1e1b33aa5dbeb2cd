// das_geom_impl.h — gfx950 delay-and-sum with delays and apodization computed
// on the device: the geometry half of +midprocess/das.m:86-206 fused into the
// beamforming loop of +mex/source/das_c.cpp:355-409.
//
// Per frame only the channel data and the output touch HBM, plus one read of
// the plan's resident transmit table. Delays/apodization are evaluated in FP64
// and rounded to FP32 exactly where das.m applies single(), so every term uses
// the FP32 delay and apodization the matrix path would get from das.m.
//
// Device-resident state (built once per plan): the transmit table [W][P] of
// {tx delay, tx apodization, tx apodization * exp(i wd tx delay)}. The receive
// side (delay, apodization, exp(i wd rx delay)) is computed per (pixel,
// channel, wave chunk), amortised over the WC waves of the chunk.
//
// Tiling. A block is 256 lanes = TP pixels x SPLIT channel groups; one lane
// owns one pixel of one group. Pixels are consecutive along the scan's fastest
// axis (depth), so the 64 lanes of a wave read a short, monotone window of each
// channel trace (coalesced, L1/L2 resident). Blocks map to pixel tiles XCD by
// XCD: the hardware deals blocks round-robin over the 8 XCDs and the remap
// hands each XCD a contiguous run of tiles, so one XCD's L2 serves one band of
// the image.
//
// Receive aperture culling: for a sorted 1-D aperture and a window that is zero
// for ratios above 1/2, each wave visits only the channel range its pixels can
// see (a conservative FP64 bound; the apodization itself is still evaluated
// exactly inside the range, so results do not change).
//
// Loop nest per lane: wave chunk (WC waves; tx values in VGPRs) -> channel
// (rx values) -> wave in chunk (unit work: index, gather, interpolate, rotate,
// accumulate).
//
// Phase. PH_RF: fd == 0, weight = apodization (das_c's exp(i*0) = 1).
// PH_DASC: exp(i*wd*(tx+rx)) per term, das_c's arithmetic (das_c.cpp:372-400).
// PH_SEP: exp(i*wd*tx) per (pixel, wave) and exp(i*wd*rx) per (pixel,
// channel), arguments formed exactly in FP64; BOTH/TX factor the rx weight out
// of the chunk sum: acc += R_n * sum_j T_j * s_nj.
//
// Summation order. PH_RF/PH_DASC with WC = 1 and SPLIT = 1 add terms in
// das_c's order (w outer, n inner); TX and RX keep das_c's order for any WC.
#pragma once
#include "das_common.h"
#include "das_launch.h"

namespace ustbx {

enum { PH_RF = 0, PH_DASC = 1, PH_SEP = 2 };
enum { MODE_NONE = 0, MODE_RX = 1, MODE_TX = 2, MODE_BOTH = 3 };

__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
  const int q = nblocks >> 3, r = nblocks & 7;
  const int k = b & 7, i = b >> 3;
  return k * q + min(k, r) + i;
}

// Interpolated (unweighted) sample at (n0, b).
template <int INTERP>
__device__ __forceinline__ float2 interp_sample(const float2* __restrict__ trace,
                                                int n0, float b) {
  if (INTERP == INTERP_LINEAR) {
    const float4 x = load_pair(trace, n0);
    return make_float2(fmaf(b, x.z - x.x, x.x), fmaf(b, x.w - x.y, x.y));
  } else {
    const float4 lo = load_pair(trace, n0 - 1);
    const float4 hi = load_pair(trace, n0 + 1);
    float c0, c1, c2, c3;
    keys_weights(b, c0, c1, c2, c3);
    return make_float2(fmaf(c3, hi.z, fmaf(c2, hi.x, fmaf(c1, lo.z, c0 * lo.x))),
                       fmaf(c3, hi.w, fmaf(c2, hi.y, fmaf(c1, lo.w, c0 * lo.y))));
  }
}

// das_c-arithmetic contribution of one term for one frame.
template <int INTERP>
__device__ __forceinline__ float2 exact_contrib(const Term& t,
                                                const float2* __restrict__ tr) {
  if (INTERP == INTERP_LINEAR) return dasc_contrib(t, load_pair(tr, t.n0));
  const float2 sm = interp_sample<INTERP>(tr, t.n0, t.b);
  return make_float2(t.wr * sm.x - t.wi * sm.y, t.wr * sm.y + t.wi * sm.x);
}

// reference phase epilogue (identity when the flag is off)
__device__ __forceinline__ float2 ref_apply(float2 v, float2 refp, int on) {
  return on ? cmul(v, refp) : v;
}

// acc += T * s with FMAs
__device__ __forceinline__ void cfma(float2& acc, float2 T, float2 s) {
  acc.x = fmaf(T.x, s.x, fmaf(-T.y, s.y, acc.x));
  acc.y = fmaf(T.x, s.y, fmaf(T.y, s.x, acc.y));
}

// Conservative channel range [lo, hi) whose rx apodization can be nonzero
// (see the culling conditions on GeomLaunch::rx_cull): ratio =
// |f0 * xd / zx| <= 1/2 <=> |xd| <= |zx| / (2 |f0|); elements beyond that
// half-width (plus a relative margin far above FP64 rounding) have a zero
// window. Where the apodization could be NaN (zx == 0) nothing is culled.
__device__ __forceinline__ void rx_range(const GeomLaunch& G, double px, double py,
                                         double pz, int& lo, int& hi) {
  lo = 0;
  hi = G.N;
  if (!G.rx_cull) return;
  const DevRxApod& ap = G.rx;
  double zd, xc;
  if (ap.geo == RXGEO_FLAT) {
    zd = pz - G.ez[0];
    xc = px;
  } else {
    const double dx = px - ap.ox, dy = py - ap.oy, dz = pz - ap.oz;
    zd = sqrt((dx * dx + dy * dy) + dz * dz);
    xc = ap.ox;
  }
  const double zx = aperture_clamp(zd, ap.min0, ap.max0);
  const double h = fabs(zx) / (2.0 * fabs(ap.f0));
  if (!(h > 0.0) || !isfinite(h)) return;
  const double m = h * 1e-9 + 1e-12;
  const double xl = xc - h - m, xh = xc + h + m;
  int a = 0, b = G.N;  // first n with ex[n] >= xl
  while (a < b) {
    const int mid = (a + b) >> 1;
    if (G.ex[mid] < xl) a = mid + 1; else b = mid;
  }
  lo = a;
  b = G.N;  // first n with ex[n] > xh
  while (a < b) {
    const int mid = (a + b) >> 1;
    if (G.ex[mid] <= xh) a = mid + 1; else b = mid;
  }
  hi = a;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

template <int MODE, int INTERP, int PH, int SPLIT, int WC, int FC>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
void das_geom_kernel(GeomLaunch G) {
  static_assert(MODE == MODE_BOTH || MODE == MODE_TX || SPLIT == 1, "split only for BOTH/TX");
  static_assert(MODE == MODE_BOTH || MODE == MODE_TX || FC == 1, "frame chunks only for BOTH/TX");
  constexpr int TP = kBlock / SPLIT;
  constexpr bool EXACT = (PH != PH_SEP);
  constexpr bool IQ = (PH != PH_RF);
  __shared__ float2 red[(SPLIT > 1 && MODE == MODE_BOTH) ? SPLIT * FC * TP : 1];

  const int lp = threadIdx.x % TP;
  const int s = threadIdx.x / TP;
  const int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  const int p = tile * TP + lp;
  const bool active = p < G.P;
  const size_t PS = (size_t)G.P;
  const size_t LN = (size_t)G.L * G.N;
  const size_t LNW = LN * G.W;
  const float2* __restrict__ ch = reinterpret_cast<const float2*>(G.ch);
  float2* __restrict__ out = reinterpret_cast<float2*>(G.out);
  const float4* __restrict__ txt = G.tx_table;

  double px = 0.0, py = 0.0, pz = 0.0, rref = 0.0;
  if (active) pixel_xyz(G.scan, p, px, py, pz, rref);

  // channel range of this lane's group, narrowed to the wave's visible aperture
  const int nper = (G.N + SPLIT - 1) / SPLIT;
  int nb = min(G.N, s * nper);
  int ne = min(G.N, nb + nper);
  {
    int lo = G.N, hi = 0;
    if (active) rx_range(G, px, py, pz, lo, hi);
    lo = wave_min(lo);
    hi = wave_max(hi);
    nb = max(nb, lo);
    ne = min(ne, hi);
    if (MODE == MODE_TX || MODE == MODE_NONE) {
      // every (pixel, channel) output is written: keep the group's full range
      nb = min(G.N, s * nper);
      ne = min(G.N, nb + nper);
    }
  }

  // reference phase exp(-1j*2*w0*r_ref/c) (das.m:306-309), FP64 -> FP32
  float2 refp = make_float2(1.0f, 0.0f);
  if (G.ref_phase && active) refp = ref_phase_factor(G.w0_ref, G.wd, G.c32, G.ref_single, rref);

  // byte strides of the channel data within one wave (< 2 GiB, checked when
  // the plan is created); a frame may be larger: the gathers address each
  // (frame, wave) through a buffer resource of its own
  const int l8 = G.L * 8;
  const int ln8 = G.L * G.N * 8;

  for (int f0 = 0; f0 < G.F; f0 += FC) {
    float2 acc[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) acc[f] = make_float2(0.0f, 0.0f);

    for (int w0 = 0; w0 < G.W; w0 += WC) {
      // ---- transmit chunk from the resident table ----
      float td[WC], ta[WC];
      float2 T[WC];
#pragma unroll
      for (int j = 0; j < WC; ++j) {
        const int w = w0 + j;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (active && w < G.W) v = txt[(size_t)w * PS + p];
        td[j] = v.x;
        ta[j] = v.y;
        T[j] = make_float2(v.z, v.w);
      }
      const bool last_chunk = (w0 + WC >= G.W);
      float2 accw[MODE == MODE_RX ? WC : 1];
#pragma unroll
      for (int j = 0; j < (MODE == MODE_RX ? WC : 1); ++j) accw[j] = make_float2(0.0f, 0.0f);

      for (int n = nb; n < ne; ++n) {
        // ---- receive: weight, then delay and phase where the weight is nonzero ----
        const double ex = G.ex[n], ey = G.ey[n], ez = G.ez[n];
        float ra = 0.0f, rd = 0.0f;
        if (active) ra = rx_apod_f32(G.rx, px, py, pz, ex, ey, ez, n, p, G.P);
        float2 R = make_float2(0.0f, 0.0f);
        if (ra != 0.0f) {
          rd = rx_delay_f32(px, py, pz, ex, ey, ez, G.c, G.lens);
          R = make_float2(ra, 0.0f);
          if (PH == PH_SEP) {
            float sv, cv;
            sincos_wd_hw(G.wd, rd, sv, cv);
            R = make_float2(ra * cv, ra * sv);
          }
        }
        float2 part[FC];
#pragma unroll
        for (int f = 0; f < FC; ++f) part[f] = make_float2(0.0f, 0.0f);
        if (MODE == MODE_TX && EXACT && w0 > 0 && active) {
          // continue das_c's running sum over waves, kept in HBM between chunks
#pragma unroll
          for (int f = 0; f < FC; ++f)
            if (f0 + f < G.F) part[f] = out[((size_t)(f0 + f) * G.N + n) * PS + p];
        }
        float2 val[MODE == MODE_NONE ? WC : 1];
#pragma unroll
        for (int j = 0; j < (MODE == MODE_NONE ? WC : 1); ++j) val[j] = make_float2(0.0f, 0.0f);
        // separable phase: whether some term of this channel is in the window
        // (das_c skips the others before multiplying, so R, which may be NaN,
        // multiplies the partial sum only then)
        bool hit = false;

        if (ra != 0.0f) {
          const int noff = n * l8;
#pragma unroll
          for (int j = 0; j < WC; ++j) {
            const int w = w0 + j;
            // a zero tx apodization skips the wave (das_c.cpp:366); on the
            // separable path it shows as T == 0 (T is NaN for a NaN weight)
            const bool txon = EXACT ? (ta[j] != 0.0f) : (T[j].x != 0.0f || T[j].y != 0.0f);
            if (w < G.W && txon) {
              const float2* trace = ch + (size_t)w * LN + (size_t)n * G.L;
              if (EXACT) {
                const Term t = dasc_term<INTERP, IQ>(td[j], rd, ta[j], ra, G.t0, G.fs,
                                                     G.wd, G.L, G.edge);
                if (t.valid) {
#pragma unroll
                  for (int f = 0; f < FC; ++f) {
                    if (f0 + f < G.F) {
                      const float2 c = exact_contrib<INTERP>(t, trace + (size_t)(f0 + f) * LNW);
                      if (MODE == MODE_BOTH) cadd(acc[f], c);
                      else if (MODE == MODE_TX) cadd(part[f], c);
                      else if (MODE == MODE_RX) cadd(accw[j], c);
                      else val[j] = c;
                    }
                  }
                }
              } else {
                int n0;
                float b;
                bool valid;
                {
#pragma clang fp contract(off)
                  const float c_delay = td[j] + rd;
                  const float denay = (c_delay - G.t0) * G.fs;
                  sample_index<INTERP>(denay, G.L, G.edge, n0, b, valid);
                }
                if (valid) {
                  hit = true;
                  // byte offset of sample n0 of trace n within wave w's data
                  const int off = noff + (n0 << 3);
#pragma unroll
                  for (int f = 0; f < FC; ++f) {
                    if (f0 + f < G.F) {
                      const __amdgpu_buffer_rsrc_t rsw = make_rsrc(ch + (size_t)(f0 + f) * LNW + (size_t)w * LN, ln8);
                      float2 sm;
                      if (INTERP == INTERP_LINEAR) {
                        const float4 x = buf_load16(rsw, off);
                        sm = make_float2(fmaf(b, x.z - x.x, x.x), fmaf(b, x.w - x.y, x.y));
                      } else {
                        const float4 lo = buf_load16(rsw, off - 8);
                        const float4 hi = buf_load16(rsw, off + 8);
                        float c0, c1, c2, c3;
                        keys_weights(b, c0, c1, c2, c3);
                        sm = make_float2(fmaf(c3, hi.z, fmaf(c2, hi.x, fmaf(c1, lo.z, c0 * lo.x))),
                                         fmaf(c3, hi.w, fmaf(c2, hi.y, fmaf(c1, lo.w, c0 * lo.y))));
                      }
                      if (MODE == MODE_BOTH || MODE == MODE_TX) {
                        cfma(part[f], T[j], sm);
                      } else {
                        const float2 v = cmul(cmul(R, T[j]), sm);
                        if (MODE == MODE_RX) cadd(accw[j], v);
                        else val[j] = v;
                      }
                    }
                  }
                }
              }
            }
          }
        }

        // ---- per-channel fold / write ----
        if (MODE == MODE_BOTH && !EXACT) {
          if (hit) {
#pragma unroll
            for (int f = 0; f < FC; ++f) cfma(acc[f], R, part[f]);
          }
        } else if (MODE == MODE_TX && active) {
#pragma unroll
          for (int f = 0; f < FC; ++f) {
            if (f0 + f < G.F) {
              float2* o = out + ((size_t)(f0 + f) * G.N + n) * PS + p;
              float2 v = part[f];
              if (!EXACT) {
                if (hit) v = cmul(R, v);
                if (w0 > 0) {
                  const float2 prev = *o;
                  cadd(v, prev);
                }
              }
              if (last_chunk) v = ref_apply(v, refp, G.ref_phase);
              *o = v;
            }
          }
        } else if (MODE == MODE_NONE && active) {
#pragma unroll
          for (int j = 0; j < WC; ++j) {
            const int w = w0 + j;
            if (w < G.W) out[(((size_t)f0 * G.W + w) * G.N + n) * PS + p] = ref_apply(val[j], refp, G.ref_phase);
          }
        }
      }  // channel

      if (MODE == MODE_RX && active) {
#pragma unroll
        for (int j = 0; j < WC; ++j) {
          const int w = w0 + j;
          if (w < G.W) out[((size_t)f0 * G.W + w) * PS + p] = ref_apply(accw[j], refp, G.ref_phase);
        }
      }
    }  // wave chunk

    if (MODE == MODE_BOTH) {
      if (SPLIT > 1) {
#pragma unroll
        for (int f = 0; f < FC; ++f) red[(s * FC + f) * TP + lp] = acc[f];
        __syncthreads();
        if (s == 0) {
#pragma unroll
          for (int f = 0; f < FC; ++f) {
            float2 v = red[f * TP + lp];
            for (int g = 1; g < SPLIT; ++g) cadd(v, red[(g * FC + f) * TP + lp]);
            acc[f] = v;
          }
        }
        __syncthreads();
      }
      if (active && s == 0) {
#pragma unroll
        for (int f = 0; f < FC; ++f)
          if (f0 + f < G.F) out[(size_t)(f0 + f) * PS + p] = ref_apply(acc[f], refp, G.ref_phase);
      }
    }
  }  // frame chunk
}

// ---------------------------------------------------------------------------
// dimension.both, separable phase, one frame: the benchmark path.
//
// A block is one tile of 64 pixels (lanes) x 4 wave-splits (the block's four
// wave64s). Split s owns transmit waves w = r0 + s + WS*j (j < WC) of each
// round of WS*WC waves (dealt round-robin, so splits stay balanced), with their tx
// delay and tx weight T in VGPRs. The
// receive state of a group of CG channels -- delay rd and weight
// R = ra * exp(i wd rd) per (pixel, channel) -- is computed once per block,
// each wave taking CG/4 channels, and shared through LDS; every split then
// applies it to its own waves. Receive geometry is thus evaluated once per
// (pixel, channel) instead of once per wave chunk. The four partial sums of a
// pixel are added in split order at the end (deterministic).
// ---------------------------------------------------------------------------
template <int INTERP, int WC, int WS>
__global__ __launch_bounds__(64 * WS) __attribute__((amdgpu_waves_per_eu(4)))
void das_both_shared_rx_kernel(GeomLaunch G) {
  constexpr int TP = 64, CG = 16;
  __shared__ float4 rxs[CG][TP];
  __shared__ float2 red[WS][TP];
  const int lane = threadIdx.x & (TP - 1);
  const int s = threadIdx.x / TP;
  const int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  const int p = tile * TP + lane;
  const bool active = p < G.P;
  const size_t PS = (size_t)G.P;
  const float4* __restrict__ txt = G.tx_table;
  const __amdgpu_buffer_rsrc_t rs = make_rsrc(G.ch, G.L * G.N * G.W * 8);
  const int l8 = G.L * 8;
  const int ln8 = G.L * G.N * 8;
  const float lim = (float)(INTERP == INTERP_LINEAR ? G.L - 1 : G.L - 2);
  const float lo_ok = (G.edge == EDGE_PYTHON && INTERP == INTERP_LINEAR) ? 0.0f : 1.0f;

  double px = 0.0, py = 0.0, pz = 0.0, rref = 0.0;
  if (active) pixel_xyz(G.scan, p, px, py, pz, rref);
  int nlo = G.N, nhi = 0;
  if (active) rx_range(G, px, py, pz, nlo, nhi);
  nlo = wave_min(nlo);  // every wave of the block holds the same 64 pixels,
  nhi = wave_max(nhi);  // so the range is uniform across the block

  float2 acc = make_float2(0.0f, 0.0f);
  for (int r0 = 0; r0 < G.W; r0 += WS * WC) {
    float td[WC];
    float2 T[WC];
#pragma unroll
    for (int j = 0; j < WC; ++j) {
      const int w = r0 + s + WS * j;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (active && w < G.W) v = txt[(size_t)w * PS + p];
      td[j] = v.x;
      T[j] = make_float2(v.z, v.w);
    }
    for (int g0 = nlo; g0 < nhi; g0 += CG) {
      // receive state of channels g0 .. g0+CG-1 for this tile
#pragma unroll
      for (int c = s; c < CG; c += WS) {
        const int n = g0 + c;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (active && n < nhi) {
          const double ex = G.ex[n], ey = G.ey[n], ez = G.ez[n];
          const float ra = rx_apod_f32(G.rx, px, py, pz, ex, ey, ez, n, p, G.P);
          if (ra != 0.0f) {
            const float rd = rx_delay_f32(px, py, pz, ex, ey, ez, G.c, G.lens);
            float sv, cv;
            sincos_wd_hw(G.wd, rd, sv, cv);
            v = make_float4(rd, ra * cv, ra * sv, 1.0f);
          }
        }
        rxs[c][lane] = v;
      }
      __syncthreads();
      const int cn = min(CG, nhi - g0);
      for (int c = 0; c < cn; ++c) {
        const float4 rv = rxs[c][lane];
        if (rv.w != 0.0f) {
          const float rd = rv.x;
          const int noff = (g0 + c) * l8;
          float2 part = make_float2(0.0f, 0.0f);
          bool hit = false;  // some term of the channel in the window (R may be NaN)
#pragma unroll
          for (int j = 0; j < WC; ++j) {
            const int w = r0 + s + WS * j;  // waves dealt round-robin over the splits
            if (w < G.W && (T[j].x != 0.0f || T[j].y != 0.0f)) {
              float d;
              {
#pragma clang fp contract(off)
                d = ((td[j] + rd) - G.t0) * G.fs;  // das_c.cpp:372,375
              }
              if (d >= lo_ok && d < lim) {
                hit = true;
                const int n0 = (int)d;
                const float b = __builtin_amdgcn_fractf(d);
                const int off = noff + w * ln8 + (n0 << 3);
                float2 sm;
                if (INTERP == INTERP_LINEAR) {
                  const float4 x = buf_load16(rs, off);
                  sm = make_float2(fmaf(b, x.z - x.x, x.x), fmaf(b, x.w - x.y, x.y));
                } else {
                  const float4 lo = buf_load16(rs, off - 8);
                  const float4 hi = buf_load16(rs, off + 8);
                  float c0, c1, c2, c3;
                  keys_weights(b, c0, c1, c2, c3);
                  sm = make_float2(fmaf(c3, hi.z, fmaf(c2, hi.x, fmaf(c1, lo.z, c0 * lo.x))),
                                   fmaf(c3, hi.w, fmaf(c2, hi.y, fmaf(c1, lo.w, c0 * lo.y))));
                }
                cfma(part, T[j], sm);
              }
            }
          }
          if (hit) cfma(acc, make_float2(rv.y, rv.z), part);
        }
      }
      __syncthreads();
    }
  }
  red[s][lane] = acc;
  __syncthreads();
  if (s == 0 && active) {
    float2 v = red[0][lane];
#pragma unroll
    for (int k = 1; k < WS; ++k) cadd(v, red[k][lane]);
    float2 refp = make_float2(1.0f, 0.0f);
    if (G.ref_phase) refp = ref_phase_factor(G.w0_ref, G.wd, G.c32, G.ref_single, rref);
    reinterpret_cast<float2*>(G.out)[p] = ref_apply(v, refp, G.ref_phase);
  }
}

template <int INTERP, int WC, int WS>
hipError_t launch_both_shared_rx_t(const GeomLaunch& G, hipStream_t stream) {
  const unsigned grid = (unsigned)((G.P + 63) / 64);
  hipLaunchKernelGGL((das_both_shared_rx_kernel<INTERP, WC, WS>), dim3(grid), dim3(64 * WS), 0, stream, G);
  return hipGetLastError();
}

template <int MODE, int INTERP, int PH, int SPLIT, int WC, int FC>
hipError_t launch_geom_t(const GeomLaunch& G, hipStream_t stream) {
  constexpr int TP = kBlock / SPLIT;
  const unsigned grid = (unsigned)((G.P + TP - 1) / TP);
  hipLaunchKernelGGL((das_geom_kernel<MODE, INTERP, PH, SPLIT, WC, FC>), dim3(grid),
                     dim3(kBlock), 0, stream, G);
  return hipGetLastError();
}

// per-mode dispatchers (one translation unit each)
hipError_t launch_both_shared(const GeomLaunch& G, hipStream_t s);
hipError_t launch_both_rxtab(const GeomLaunch& G, hipStream_t s);
hipError_t launch_tx_rxtab(const GeomLaunch& G, hipStream_t s);
hipError_t launch_geom_both(const GeomLaunch& G, int ph, int wc, int split, hipStream_t s);
hipError_t launch_geom_tx(const GeomLaunch& G, int ph, int wc, int split, hipStream_t s);
hipError_t launch_geom_rx_none(const GeomLaunch& G, int ph, int wc, hipStream_t s);

}  // namespace ustbx
