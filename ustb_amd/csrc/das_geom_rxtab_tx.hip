// das_geom_rxtab_tx.hip — dimension.transmit from a plan's resident tables
// (variant 2): out[f][n][p] = R(p,n) * sum_w T(p,w) * x_f[w][n](d(p,w,n)),
// the per-channel images das_c keeps when it sums over transmits only
// (das_c.cpp:395-399), with the phase taken as exp(i wd tx) * exp(i wd rx).
//
// A block is one 64-pixel tile x CS wave64s, one channel each. The work is
// slabs (a group of CS channels x a chunk of FC frames) x tiles, slab-major,
// and each XCD takes a contiguous share of it (the XCD remap): an XCD owns
// whole slabs and visits a slab's tiles in the plan's order (depth band by
// depth band across all scan lines), so the blocks resident on an XCD read
// one slab's traces over one depth band, a window that fits its L2, and the
// channel data is fetched about once. Every channel's image is written (a
// channel with no receive weight on the tile gets zeros, das_c's
// zero-initialized output). Per (channel, wave) the sample index is formed
// once (das_c's operation order) and the FC frames' samples are gathered
// together: the tables, the index arithmetic and the transmit weight are
// shared by all frames. Terms outside the window (or with a zero weight) read
// through an out-of-range offset, as in the dimension-both tile body (same
// finiteness condition on the plan).
//
// Phase (PH): PH_SEP takes the tables' exp(i wd td) and exp(i wd rd) and
// multiplies the receive factor in after the sum. PH_DASC and PH_RF follow
// das_c's term arithmetic instead (das_c.cpp:372-399): c_delay = td + rd,
// c_apo = ta * ra, weight = c_apo * exp(i fl32(wd * c_delay)) (PH_RF: c_apo),
// das_c's 4-term complex dot, added to the running sum over waves in
// ascending order. The sine and cosine are taken once per (pixel, channel,
// wave) and shared by the block's frames. That is das_c's order for this
// mode, so the result equals das_c's (bit for bit for RF data, linear).
#include "das_cf.h"
#include "das_rxtab.h"

namespace ustbx {

// One channel's images of FC frames f0 .. f0+nf-1 on one 64-pixel tile (a
// wave64, lane = pixel): v[f] is what das_c writes to out[f][n][p] (times the
// reference phase where the plan has it), zeros for a channel with no receive
// weight on the tile. Wave-uniform: tile, n, f0, nf. SAFE (tiles where a
// nonzero weight or its delay is not finite): a term enters the sum only when
// it is in the window, and the separable receive factor multiplies the sum
// only when some term did, as das_c skips the others before multiplying.
// STAGE (branch-free tiles only): the channels that the plan's stage mask
// marks for the tile (bit n of its words) read their samples from LDS. Per
// wave of the tile's list, the reference index nref is the n0 of the first
// lane whose term is in the window (ok); the mask guarantees that every ok
// lane's n0 lies in [nref - kTxSlack, nref + kMaxSpan - kTxSlack]. One
// contiguous 8-byte-per-lane load per frame fills the wave's row win[f][0,
// 64) with samples nref - kBack - kTxSlack .. + 63 of the trace; ok lanes
// read their term's samples back from there, the others from the row's
// zero slots [64, 68) (the gather path reads zeros for them through an
// out-of-range offset). The LDS operations of one wave complete in order, so
// the reads see the writes. The arithmetic is the gather path's, so the two
// agree bit for bit.
constexpr int kTxRow = kStageWin + 4;  // float2 slots per frame row: window + zeros

template <int SM, int FC, int PH, bool SAFE, bool STAGE>
__device__ __forceinline__ void tx_channel(const RxAcq& A, float wd, int tile, int n, int f0, int nf, int lane,
                                           float2 (&v)[FC], float2* __restrict__ win) {
  using M = SampleMode<SM>;
  constexpr int ST = M::kStride;
  const int p = tile_pixel(A.tile_ps, A.tile_n0, tile, lane);
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

#pragma unroll
  for (int f = 0; f < FC; ++f) v[f] = make_float2(0.0f, 0.0f);
  float4 rv = make_float4(0.f, 0.f, 0.f, 0.f);
  if (active) rv = rx_entry(A.rx, A.rx_compact, (size_t)n * PS + p);
  const bool ron = rv.y != 0.0f;
  if (!__any(ron)) return;  // no receive weight on the tile: zeros
  const float rd = ron ? rv.x : kFar;
  const float ra = rv.y;
  const float2 R = ron ? make_float2(rv.z, rv.w) : make_float2(0.0f, 0.0f);
  if (SAFE && !__any(ron && ra == ra)) {
    // every nonzero receive weight on the tile is NaN: das_c's sum is NaN where
    // some term is in the window, else zero; no gathers
    bool any = false;
    for (int i = 0; i < nw; ++i) {
      const int wi = __builtin_amdgcn_readfirstlane(wl[i]);
      const float4 tv = active ? A.tx[(size_t)wi * PS + p] : make_float4(0.f, 0.f, 0.f, 0.f);
      float d;
      {
#pragma clang fp contract(off)
        d = (((tv.y != 0.0f ? tv.x : kFar) + rd) - t0) * fs;
      }
      any = any || (d >= lo_ok && d < lim);
    }
    const float nan = __builtin_nanf("");
#pragma unroll
    for (int f = 0; f < FC; ++f) v[f] = any ? make_float2(nan, nan) : make_float2(0.0f, 0.0f);
    return;
  }
  const unsigned nb = (unsigned)(n * lst);
  float2 part[FC];
#pragma unroll
  for (int f = 0; f < FC; ++f) part[f] = make_float2(0.0f, 0.0f);
  // the next wave's transmit row is in flight during this wave's gathers,
  // its index one further ahead
  int w = nw > 0 ? __builtin_amdgcn_readfirstlane(wl[0]) : 0;
  int w_next = nw > 1 ? __builtin_amdgcn_readfirstlane(wl[1]) : 0;
  bool any = false;  // SAFE: some term of the channel was in the window
  // staged: the plan's stage mask marks channel n on this tile (a scalar
  // load through the constant address space; the mask is read-only here)
  bool stg = false;
  unsigned wbase = 0u, lane8 = 0u;
  if (STAGE && !SAFE && A.stage_mask) {
    const unsigned long long word =
        ((const __attribute__((address_space(4))) unsigned long long*)A.stage_mask)[(size_t)tile * A.stage_words +
                                                                                    (n >> 6)];
    stg = (word >> (n & 63)) & 1ull;
    wbase = (unsigned)(size_t)(__attribute__((address_space(3))) char*)win;  // LDS address
    lane8 = (unsigned)lane * 8u;
    asm volatile("" : "+v"(lane8));  // keeps window addresses one scalar-plus-vector add
  }
  float4 vn = make_float4(0.f, 0.f, 0.f, 0.f);
  if (active && nw > 0) vn = A.tx[(size_t)w * PS + p];
  for (int i = 0; i < nw; ++i) {
    const float4 tv = vn;
    const int wc = w;
    if (i + 1 < nw) {
      w = w_next;
      if (active) vn = A.tx[(size_t)w * PS + p];
      if (i + 2 < nw) w_next = __builtin_amdgcn_readfirstlane(wl[i + 2]);
    }
    const bool ton = tv.y != 0.0f;
    const float td = ton ? tv.x : kFar;
    const float2 T = ton ? make_float2(tv.z, tv.w) : make_float2(0.0f, 0.0f);
    float d, c_delay;
    {
#pragma clang fp contract(off)
      c_delay = td + rd;            // das_c.cpp:372
      d = (c_delay - t0) * fs;      // das_c.cpp:375
    }
    // das_c's weight for this term (zero where either weight is zero)
    float wr = 0.0f, wi = 0.0f;
    if (PH != PH_SEP) {
#pragma clang fp contract(off)
      const bool on = ton && ron;
      const float c_apo = on ? tv.y * ra : 0.0f;  // das_c.cpp:373
      if (PH == PH_DASC) {
        float sv, cv;
        sincosf(wd * (on ? c_delay : 0.0f), &sv, &cv);  // das_c.cpp:383-387
        wr = cv * c_apo;
        wi = sv * c_apo;
      } else {
        wr = c_apo;
      }
    }
    const bool ok = d >= lo_ok && d < lim;
    if (SAFE) any = any || ok;
    const int n0 = (int)d;
    const float b = __builtin_amdgcn_fractf(d);
    // the first gather's offset folded into the wave-uniform part (n0 >= 1
    // wherever ok holds, so the sum stays non-negative)
    float4 q[FC][M::kLoads];
    if (STAGE && stg) {
      using S = StageWin<SM>;
      const unsigned long long okm = __builtin_amdgcn_ballot_w64(ok);
      const unsigned nref =
          okm ? (unsigned)__builtin_amdgcn_readlane(n0, (int)__builtin_ctzll(okm)) : 0u;  // wave-uniform
      // byte offset of the window's first sample (wraps below sample 0: those
      // slots hold zeros or the previous trace, and no lane reads them)
      const unsigned sb = nb + (unsigned)(wc * wst) + (nref - (unsigned)(S::kBack + kTxSlack)) * 8u;
      float2 st[FC];
#pragma unroll
      for (int f = 0; f < FC; ++f) st[f] = buf_load8(rs, (int)(sb + (unsigned)(f * fb) + lane8));
#pragma unroll
      for (int f = 0; f < FC; ++f)
        *reinterpret_cast<lds_f2*>((size_t)(wbase + lane8) + f * kTxRow * 8) = f2v{st[f].x, st[f].y};
      __builtin_amdgcn_wave_barrier();
      // LDS address of the first sample the term reads (n0 - kBack), or of
      // the zero slots
      const unsigned la =
          ok ? lshl3_add((unsigned)n0, wbase + ((unsigned)kTxSlack - nref) * 8u) : wbase + kStageWin * 8u;
#pragma unroll
      for (int f = 0; f < FC; ++f) {
        const f4a8 lo = *reinterpret_cast<const lds_f4a8*>((size_t)la + f * kTxRow * 8);
        q[f][0] = make_float4(lo.x, lo.y, lo.z, lo.w);
        if (M::kLoads == 2) {
          const f4a8 hi = *reinterpret_cast<const lds_f4a8*>((size_t)la + f * kTxRow * 8 + 16);
          q[f][M::kLoads - 1] = make_float4(hi.x, hi.y, hi.z, hi.w);
        }
      }
    } else {
      const unsigned off =
          ok ? (unsigned)(n0 * ST) + (nb + (unsigned)(wc * wst) + (unsigned)M::kLo) : (unsigned)kOOB;
#pragma unroll
      for (int f = 0; f < FC; ++f) {
        const unsigned o = off + (unsigned)(f * fb);
        q[f][0] = buf_load16(rs, (int)o);
        if (M::kLoads == 2) q[f][M::kLoads - 1] = buf_load16(rs, (int)(o + (unsigned)(M::kHi - M::kLo)));
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    float c0 = 0.f, c1 = 0.f, c2 = 0.f, c3 = 0.f;
    if (SM == SM_CUBIC) keys_weights_horner(b, c0, c1, c2, c3);
#pragma unroll
    for (int f = 0; f < FC; ++f) {
      float2 np = part[f];
      if (PH != PH_SEP && SM == SM_LINEAR) {
        Term t;
        t.a = 1.0f - b;
        t.b = b;
        t.wr = wr;
        t.wi = wi;
        cadd(np, dasc_contrib(t, q[f][0]));  // das_c.cpp:389-399
      } else {
        float2 sm;
        if (SM == SM_LINEAR) {
          const float4 x = q[f][0];
          sm = make_float2(fmaf(b, x.z - x.x, x.x), fmaf(b, x.w - x.y, x.y));
        } else {
          const float4 lo = q[f][0], hi = q[f][M::kLoads - 1];
          sm = make_float2(fmaf(c3, hi.z, fmaf(c2, hi.x, fmaf(c1, lo.z, c0 * lo.x))),
                           fmaf(c3, hi.w, fmaf(c2, hi.y, fmaf(c1, lo.w, c0 * lo.y))));
        }
        if (PH == PH_SEP) {
          cfma(np, T, sm);
        } else {
#pragma clang fp contract(off)
          cadd(np, make_float2(wr * sm.x - wi * sm.y, wr * sm.y + wi * sm.x));
        }
      }
      part[f] = (!SAFE || ok) ? np : part[f];
    }
    if (STAGE && stg) __builtin_amdgcn_wave_barrier();  // this wave's reads before the next wave's writes
  }
  const float2 refp = (active && A.ref) ? A.ref[p] : make_float2(1.0f, 0.0f);
#pragma unroll
  for (int f = 0; f < FC; ++f) {
    float2 x = (PH == PH_SEP && (!SAFE || any)) ? cmul(R, part[f]) : part[f];
    if (A.ref) x = cmul(x, refp);
    v[f] = x;
  }
}

// Blocks cover slabs x tiles entries of the visiting order (order[0 ..
// tiles), or the tiles themselves where order is null).
template <int SM, int FC, int CS, int PH, bool SAFE, bool STAGE>
__global__ __launch_bounds__(64 * CS) void das_tx_rxtab_kernel(RxAcq A, int F, float wd,
                                                               const int* __restrict__ order, int tiles) {
  constexpr int TP = 64;
  __shared__ __align__(16) float2 win[STAGE ? CS * FC * kTxRow : 1];
  const int lane = threadIdx.x & (TP - 1);
  const int s = __builtin_amdgcn_readfirstlane((int)threadIdx.x / TP);
  const int chunks = (F + FC - 1) / FC;
  const int item = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  const int slab = item / tiles;
  const int rank = item - slab * tiles;
  const int group = slab / chunks;
  const int n = group * CS + s;
  if (n >= A.N) return;  // no block barrier
  const int tile = order ? __builtin_amdgcn_readfirstlane(order[rank]) : rank;
  const int f0 = (slab - group * chunks) * FC;
  const int nf = min(FC, F - f0);
  if (STAGE && lane < FC * 4) {  // this wave's zero slots, never overwritten
    win[s * FC * kTxRow + (lane >> 2) * kTxRow + kStageWin + (lane & 3)] = make_float2(0.0f, 0.0f);
    __builtin_amdgcn_wave_barrier();
  }
  float2 v[FC];
  tx_channel<SM, FC, PH, SAFE, STAGE>(A, wd, tile, n, f0, nf, lane, v, win + (STAGE ? s * FC * kTxRow : 0));
  const int p = tile_pixel(A.tile_ps, A.tile_n0, tile, lane);
  if (p < A.P) {
    const size_t PS = (size_t)A.P;
#pragma unroll
    for (int f = 0; f < FC; ++f)
      if (f < nf) A.out[((size_t)(f0 + f) * A.N + n) * PS + p] = v[f];
  }
}

// Dimension transmit with the coherence factor over the receive channels,
// first stage: the transmit kernel's blocks and work order (one wave per
// channel, XCDs owning channel-group x frame-chunk slabs), but instead of the
// per-channel images each block stores, per frame and pixel, the sums over
// its CS channels in channel order: S_g = sum v, E_g = sum |v|^2
// (coherence_factor.m:126-127), through LDS. part is [G][F][P] float4
// {S.x, S.y, E, 0}, G = ceil(N / CS).
template <int SM, int FC, int CS, int PH, bool SAFE, bool STAGE>
__global__ __launch_bounds__(64 * CS) void das_tx_cf_part_kernel(RxAcq A, int F, float wd,
                                                                 const int* __restrict__ order, int tiles,
                                                                 float4* __restrict__ part) {
  constexpr int TP = 64;
  __shared__ float2 red[CS][TP];
  __shared__ __align__(16) float2 win[STAGE ? CS * FC * kTxRow : 1];
  const int lane = threadIdx.x & (TP - 1);
  const int s = __builtin_amdgcn_readfirstlane((int)threadIdx.x / TP);
  const int chunks = (F + FC - 1) / FC;
  const int item = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  const int slab = item / tiles;
  const int rank = item - slab * tiles;
  const int group = slab / chunks;
  const int n = group * CS + s;
  const int tile = order ? __builtin_amdgcn_readfirstlane(order[rank]) : rank;
  const int f0 = (slab - group * chunks) * FC;
  const int nf = min(FC, F - f0);
  if (STAGE && lane < FC * 4) {  // this wave's zero slots, never overwritten
    win[s * FC * kTxRow + (lane >> 2) * kTxRow + kStageWin + (lane & 3)] = make_float2(0.0f, 0.0f);
    __builtin_amdgcn_wave_barrier();
  }
  float2 v[FC];
  if (n < A.N) {
    tx_channel<SM, FC, PH, SAFE, STAGE>(A, wd, tile, n, f0, nf, lane, v, win + (STAGE ? s * FC * kTxRow : 0));
  } else {
#pragma unroll
    for (int f = 0; f < FC; ++f) v[f] = make_float2(0.0f, 0.0f);
  }
  const int p = tile_pixel(A.tile_ps, A.tile_n0, tile, lane);
  const size_t PS = (size_t)A.P;
#pragma unroll
  for (int f = 0; f < FC; ++f) {
    red[s][lane] = v[f];
    __syncthreads();
    if (s == 0 && p < A.P && f < nf) {
      float2 S = make_float2(0.0f, 0.0f);
      float E = 0.0f;
#pragma unroll
      for (int k = 0; k < CS; ++k) cf_accumulate(S, E, red[k][lane]);
      part[((size_t)group * F + f0 + f) * PS + p] = make_float4(S.x, S.y, E, 0.0f);
    }
    __syncthreads();
  }
}

// Second stage: S = sum over the groups in order, E likewise, then the
// coherence factor; image and cf [F][P].
__global__ __launch_bounds__(256) void das_tx_cf_finish_kernel(const float4* __restrict__ part, int G, int F, int P,
                                                              int N, float2* __restrict__ image,
                                                              float* __restrict__ cfo, const float* __restrict__ M,
                                                              int flavour) {
  const size_t total = (size_t)F * P;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
    float2 S = make_float2(0.0f, 0.0f);
    float E = 0.0f;
    for (int g = 0; g < G; ++g) {
      const float4 q = part[(size_t)g * total + i];
      S.x = S.x + q.x;
      S.y = S.y + q.y;
      E = E + q.z;
    }
    const size_t p = i % (size_t)P;
    const float m = (flavour == USTBX_CF_PYTHON || !M) ? (float)N : M[p];
    const float cf = coherence_value(S, E, m, flavour);
    image[i] = make_float2(cf * S.x, cf * S.y);
    if (cfo) cfo[i] = cf;
  }
}

// The first tile_good entries of the visiting order take the branch-free
// body; the rest (a nonzero weight or its delay not finite) the SAFE body, in
// a second launch with 8 frames per block and the same channels per block.
template <int SM, int FC, int CS, int PH>
static hipError_t launch_tx_t(const GeomLaunch& G, hipStream_t stream) {
  const int tiles = (G.P + 63) / 64, good = G.tile_good;
  if (good < tiles && !G.tile_order) return hipErrorInvalidValue;
  const int64_t slabs = (int64_t)((G.F + FC - 1) / FC) * ((G.N + CS - 1) / CS);
  if ((int64_t)tiles * slabs > 0x7FFFFFFF) return hipErrorInvalidValue;
  const RxAcq A = rx_acq(G, SM);
  if (good > 0) {
    if (G.stage)
      hipLaunchKernelGGL((das_tx_rxtab_kernel<SM, FC, CS, PH, false, true>), dim3((unsigned)(good * slabs)),
                         dim3(64 * CS), 0, stream, A, G.F, G.wd, G.tile_order, good);
    else
      hipLaunchKernelGGL((das_tx_rxtab_kernel<SM, FC, CS, PH, false, false>), dim3((unsigned)(good * slabs)),
                         dim3(64 * CS), 0, stream, A, G.F, G.wd, G.tile_order, good);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (good < tiles) {
    const int64_t bslabs = (int64_t)((G.F + 7) / 8) * ((G.N + CS - 1) / CS);
    hipLaunchKernelGGL((das_tx_rxtab_kernel<SM, 8, CS, PH, true, false>), dim3((unsigned)((tiles - good) * bslabs)),
                       dim3(64 * CS), 0, stream, A, G.F, G.wd, G.tile_order + good, tiles - good);
    return hipGetLastError();
  }
  return hipSuccess;
}

template <int SM, int PH>
static hipError_t launch_tx_sm(const GeomLaunch& G, hipStream_t stream) {
  const int fc = G.wave_chunk, cs = G.split;
  if (fc == 4) return cs == 4 ? launch_tx_t<SM, 4, 4, PH>(G, stream) : launch_tx_t<SM, 4, 8, PH>(G, stream);
  if (fc == 8) return cs == 4 ? launch_tx_t<SM, 8, 4, PH>(G, stream) : launch_tx_t<SM, 8, 8, PH>(G, stream);
  return hipErrorInvalidValue;
}

template <int SM>
static hipError_t launch_tx_ph(const GeomLaunch& G, hipStream_t stream) {
  if (G.wd == 0.0f) return launch_tx_sm<SM, PH_RF>(G, stream);
  if (G.separable) return launch_tx_sm<SM, PH_SEP>(G, stream);
  return launch_tx_sm<SM, PH_DASC>(G, stream);
}

template <int SM, int FC, int CS, int PH>
static hipError_t launch_cf_t(const GeomLaunch& G, float4* part, hipStream_t stream) {
  const int tiles = (G.P + 63) / 64, good = G.tile_good;
  if (good < tiles && !G.tile_order) return hipErrorInvalidValue;
  const int64_t slabs = (int64_t)((G.F + FC - 1) / FC) * ((G.N + CS - 1) / CS);
  if ((int64_t)tiles * slabs > 0x7FFFFFFF) return hipErrorInvalidValue;
  const RxAcq A = rx_acq(G, SM);
  if (good > 0) {
    if (G.stage)
      hipLaunchKernelGGL((das_tx_cf_part_kernel<SM, FC, CS, PH, false, true>), dim3((unsigned)(good * slabs)),
                         dim3(64 * CS), 0, stream, A, G.F, G.wd, G.tile_order, good, part);
    else
      hipLaunchKernelGGL((das_tx_cf_part_kernel<SM, FC, CS, PH, false, false>), dim3((unsigned)(good * slabs)),
                         dim3(64 * CS), 0, stream, A, G.F, G.wd, G.tile_order, good, part);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (good < tiles) {
    const int64_t bslabs = (int64_t)((G.F + 7) / 8) * ((G.N + CS - 1) / CS);
    hipLaunchKernelGGL((das_tx_cf_part_kernel<SM, 8, CS, PH, true, false>),
                       dim3((unsigned)((tiles - good) * bslabs)), dim3(64 * CS), 0, stream, A, G.F, G.wd,
                       G.tile_order + good, tiles - good, part);
    return hipGetLastError();
  }
  return hipSuccess;
}

template <int SM, int PH>
static hipError_t launch_cf_sh(const GeomLaunch& G, float4* part, hipStream_t stream) {
  const int fc = G.wave_chunk, cs = G.split;
  if (fc == 4) return cs == 4 ? launch_cf_t<SM, 4, 4, PH>(G, part, stream) : launch_cf_t<SM, 4, 8, PH>(G, part, stream);
  if (fc == 8) return cs == 4 ? launch_cf_t<SM, 8, 4, PH>(G, part, stream) : launch_cf_t<SM, 8, 8, PH>(G, part, stream);
  return hipErrorInvalidValue;
}

template <int SM>
static hipError_t launch_cf_ph(const GeomLaunch& G, float4* part, hipStream_t stream) {
  if (G.wd == 0.0f) return launch_cf_sh<SM, PH_RF>(G, part, stream);
  if (G.separable) return launch_cf_sh<SM, PH_SEP>(G, part, stream);
  return launch_cf_sh<SM, PH_DASC>(G, part, stream);
}

// Bytes of the partial sums launch_tx_cf needs: ceil(N / channels per block)
// x F x P float4.
size_t tx_cf_workspace_bytes(const GeomLaunch& G) {
  const int cs = G.split > 0 ? G.split : 1;
  return (size_t)((G.N + cs - 1) / cs) * (size_t)G.F * (size_t)G.P * sizeof(float4);
}

// The fused coherence-factor form: partial sums per channel group (the
// plan's frames and channels per block) into part, then the finishing pass.
hipError_t launch_tx_cf(const GeomLaunch& G, float4* part, float2* image, float* cf, const float* M, int flavour,
                        hipStream_t stream) {
  hipError_t e = G.interp == INTERP_CUBIC ? launch_cf_ph<SM_CUBIC>(G, part, stream)
                                          : launch_cf_ph<SM_LINEAR>(G, part, stream);
  if (e != hipSuccess) return e;
  const int groups = (G.N + G.split - 1) / G.split;
  size_t blocks = ((size_t)G.F * G.P + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(das_tx_cf_finish_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, part, groups, G.F, G.P,
                     G.N, image, cf, M, flavour);
  return hipGetLastError();
}

// Frames per block (wave_chunk: 4 or 8) and channels per block (split: 4 or
// 8); the phase follows the plan (RF data and das_c's phase: das_c's term
// arithmetic; otherwise the separable phase).
hipError_t launch_tx_rxtab(const GeomLaunch& G, hipStream_t stream) {
  if (G.interp == INTERP_CUBIC) return launch_tx_ph<SM_CUBIC>(G, stream);
  return launch_tx_ph<SM_LINEAR>(G, stream);
}

}  // namespace ustbx
