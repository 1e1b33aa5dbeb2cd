// das_geom_rxtab_tx.hip — dimension.transmit from a plan's resident tables
// (variant 2): out[f][n][p] = R(p,n) * sum_w T(p,w) * x_f[w][n](d(p,w,n)),
// the per-channel images das_c keeps when it sums over transmits only
// (das_c.cpp:395-399), with the phase taken as exp(i wd tx) * exp(i wd rx).
//
// A block is one 64-pixel tile x CS wave64s, one channel each: grid.z runs
// over groups of CS channels (slowest), grid.y over chunks of FC frames, so
// the blocks resident together share one channel group and one frame chunk
// and their sample windows fit in L2. Every channel's image is written (a
// channel with no receive weight on the tile gets zeros, das_c's
// zero-initialized output). Per (channel, wave) the sample index is formed
// once (das_c's operation order) and the FC frames' samples are gathered
// together: the tables, the index arithmetic and the transmit weight are
// shared by all frames. Terms outside the window (or with a zero weight) read
// through an out-of-range offset, as in the dimension-both tile body (same
// finiteness condition on the plan).
#include "das_rxtab.h"

namespace ustbx {

template <int SM, int FC, int CS>
__global__ __launch_bounds__(64 * CS) void das_tx_rxtab_kernel(RxAcq A, int F, const int* __restrict__ order) {
  using M = SampleMode<SM>;
  constexpr int TP = 64;
  constexpr int ST = M::kStride;
  const int lane = threadIdx.x & (TP - 1);
  const int s = __builtin_amdgcn_readfirstlane((int)threadIdx.x / TP);
  const int n = (int)blockIdx.z * CS + s;
  if (n >= A.N) return;  // no block barrier below
  int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  if (order) tile = __builtin_amdgcn_readfirstlane(order[tile]);
  const int f0 = (int)blockIdx.y * FC;
  const int nf = min(FC, F - f0);
  const int p = tile * TP + lane;
  const bool active = p < A.P;
  const size_t PS = (size_t)A.P;
  const int nw = __builtin_amdgcn_readfirstlane(A.tile_nw[tile]);
  const int* __restrict__ wl = A.tile_waves + (size_t)tile * A.W;
  const int lst = A.L * ST;   // bytes per trace
  const int wst = lst * A.N;  // bytes per wave
  const int fb = A.frame_bytes;
  // one resource over this block's frames (FC * frame bytes < 2 GiB, checked
  // by the plan); frames past nf fall outside it and read zeros. Offsets are
  // unsigned: kOOB + (FC-1)*fb stays below 2^32 and above the range.
  const __amdgpu_buffer_rsrc_t rs = make_rsrc((const char*)A.data + (size_t)f0 * fb, nf * fb);
  const float lim = (float)(SM == SM_LINEAR ? A.L - 1 : A.L - 2);
  constexpr float lo_ok = 1.0f;
  const float t0 = A.t0, fs = A.fs;

  {
    float4 rv = make_float4(0.f, 0.f, 0.f, 0.f);
    if (active) rv = A.rx[(size_t)n * PS + p];
    const bool ron = rv.y != 0.0f;
    if (!__any(ron)) {  // no receive weight on the tile: zeros
      if (active)
        for (int f = 0; f < nf; ++f) A.out[((size_t)(f0 + f) * A.N + n) * PS + p] = make_float2(0.0f, 0.0f);
      return;
    }
    const float rd = ron ? rv.x : kFar;
    const float2 R = ron ? make_float2(rv.z, rv.w) : make_float2(0.0f, 0.0f);
    const unsigned nb = (unsigned)(n * lst);
    float2 part[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) part[f] = make_float2(0.0f, 0.0f);
    // the next wave's transmit row is in flight during this wave's gathers,
    // its index one further ahead
    int w = nw > 0 ? __builtin_amdgcn_readfirstlane(wl[0]) : 0;
    int w_next = nw > 1 ? __builtin_amdgcn_readfirstlane(wl[1]) : 0;
    float4 vn = make_float4(0.f, 0.f, 0.f, 0.f);
    if (active && nw > 0) vn = A.tx[(size_t)w * PS + p];
    for (int i = 0; i < nw; ++i) {
      const float4 v = vn;
      const int wc = w;
      if (i + 1 < nw) {
        w = w_next;
        if (active) vn = A.tx[(size_t)w * PS + p];
        if (i + 2 < nw) w_next = __builtin_amdgcn_readfirstlane(wl[i + 2]);
      }
      const bool ton = v.y != 0.0f;
      const float td = ton ? v.x : kFar;
      const float2 T = ton ? make_float2(v.z, v.w) : make_float2(0.0f, 0.0f);
      float d;
      {
#pragma clang fp contract(off)
        d = ((td + rd) - t0) * fs;  // das_c.cpp:372,375
      }
      const bool ok = d >= lo_ok && d < lim;
      const int n0 = (int)d;
      const float b = __builtin_amdgcn_fractf(d);
      const unsigned off = ok ? (unsigned)(n0 * ST) + nb + (unsigned)(wc * wst) : (unsigned)kOOB;
      float4 q[FC][M::kLoads];
#pragma unroll
      for (int f = 0; f < FC; ++f) {
        const unsigned o = off + (unsigned)(f * fb);
        q[f][0] = buf_load16(rs, (int)(o + (unsigned)M::kLo));
        if (M::kLoads == 2) q[f][M::kLoads - 1] = buf_load16(rs, (int)(o + (unsigned)M::kHi));
      }
      __builtin_amdgcn_sched_barrier(0);
      float c0 = 0.f, c1 = 0.f, c2 = 0.f, c3 = 0.f;
      if (SM == SM_CUBIC) keys_weights_horner(b, c0, c1, c2, c3);
#pragma unroll
      for (int f = 0; f < FC; ++f) {
        float2 sm;
        if (SM == SM_LINEAR) {
          const float4 x = q[f][0];
          sm = make_float2(fmaf(b, x.z - x.x, x.x), fmaf(b, x.w - x.y, x.y));
        } else {
          const float4 lo = q[f][0], hi = q[f][M::kLoads - 1];
          sm = make_float2(fmaf(c3, hi.z, fmaf(c2, hi.x, fmaf(c1, lo.z, c0 * lo.x))),
                           fmaf(c3, hi.w, fmaf(c2, hi.y, fmaf(c1, lo.w, c0 * lo.y))));
        }
        cfma(part[f], T, sm);
      }
    }
    if (active) {
      const float2 refp = A.ref ? A.ref[p] : make_float2(1.0f, 0.0f);
#pragma unroll
      for (int f = 0; f < FC; ++f) {
        if (f < nf) {
          float2 v = cmul(R, part[f]);
          if (A.ref) v = cmul(v, refp);
          A.out[((size_t)(f0 + f) * A.N + n) * PS + p] = v;
        }
      }
    }
  }
}

template <int SM, int FC, int CS>
static hipError_t launch_tx_t(const GeomLaunch& G, hipStream_t stream) {
  const dim3 grid((unsigned)((G.P + 63) / 64), (unsigned)((G.F + FC - 1) / FC), (unsigned)((G.N + CS - 1) / CS));
  hipLaunchKernelGGL((das_tx_rxtab_kernel<SM, FC, CS>), grid, dim3(64 * CS), 0, stream, rx_acq(G, SM), G.F,
                     G.tile_order);
  return hipGetLastError();
}

template <int SM>
static hipError_t launch_tx_sm(const GeomLaunch& G, hipStream_t stream) {
  const int fc = G.wave_chunk, cs = G.split;
  if (fc == 4) return cs == 4 ? launch_tx_t<SM, 4, 4>(G, stream) : launch_tx_t<SM, 4, 8>(G, stream);
  if (fc == 8) return cs == 4 ? launch_tx_t<SM, 8, 4>(G, stream) : launch_tx_t<SM, 8, 8>(G, stream);
  return hipErrorInvalidValue;
}

// Frames per block (wave_chunk: 4 or 8) and channels per block (split: 4 or 8).
hipError_t launch_tx_rxtab(const GeomLaunch& G, hipStream_t stream) {
  if (G.interp == INTERP_CUBIC) return launch_tx_sm<SM_CUBIC>(G, stream);
  return launch_tx_sm<SM_LINEAR>(G, stream);
}

}  // namespace ustbx
