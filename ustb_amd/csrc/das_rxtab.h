// das_rxtab.h — the tile body of the receive-table kernels (dimension both):
// shared by the one-plan kernel (das_geom_rxtab.hip) and the persistent
// batch kernel (das_batch.hip), so both compute a tile with the same
// operations in the same order.
//
// A tile is 64 pixels (lanes) x WS wave-splits (the block's wave64s). Split
// s takes the tile's active waves k = k0 + s + WS*j (j < WC), holding their
// transmit delay and weight T in VGPRs, and runs over the tile's channel
// list, reading the receive row of the next channel ahead. Per term: the
// sample index in das_c's operation order, one or two 16-byte gathers, the
// interpolated sample, then part += T * sample; per channel acc += R * part.
// The WS partial sums of a pixel are added in split order through LDS.
//
// The term loop has no data-dependent branch: a term outside the sample
// window (or with a zero weight, whose delay is replaced by kFar) reads
// through an out-of-range buffer offset, which the hardware returns as zeros.
// That is exact only when every nonzero weight and its delay are finite
// (das_c skips a zero-weight or out-of-window term before it multiplies, so
// a NaN weight on a skipped term must not reach the sum). Tiles where a
// nonzero weight or its delay is not finite (e.g. the NaN apodization of
// pixels at depth 0) take the SAFE form instead: a term enters the sum only
// when it is in the window, and a channel's receive weight multiplies its
// partial sum only when some term of the channel did. Channels whose nonzero
// weights on the tile are all NaN come last in the tile's list and are not
// run: whether one of their terms is in the window (which makes das_c's sum
// NaN) depends on the plan's geometry only, so the plan stores it as a
// per-tile lane mask.
#pragma once
#include "das_geom_impl.h"

namespace ustbx {

// Delay that puts every term it enters outside the sample window: finite,
// so that the fraction of the (huge, integral) index stays 0.
constexpr float kFar = -1.0e20f;
// Byte offset beyond any buffer a plan accepts (frames are capped at
// 0x7FFFFE00 bytes), read back as zeros.
constexpr int kOOB = 0x7FFFFF00;

// Sampling modes of the term loop:
//   SM_LINEAR: channel data, one 16-byte gather {x[n0], x[n0+1]};
//   SM_CUBIC:  channel data, two 16-byte gathers {x[n0-1], x[n0]},
//              {x[n0+1], x[n0+2]} and Keys' weights (a = -1/2);
//   SM_COEF:   Catmull-Rom coefficients (32 bytes per sample), two 16-byte
//              gathers and a Horner evaluation.
enum { SM_LINEAR = 0, SM_CUBIC = 1, SM_COEF = 2 };

template <int SM>
struct SampleMode {
  static constexpr int kStride = SM == SM_COEF ? 32 : 8;  // bytes per sample
  static constexpr int kLoads = SM == SM_LINEAR ? 1 : 2;
  static constexpr int kLo = SM == SM_CUBIC ? -8 : 0;     // byte offsets of the gathers
  static constexpr int kHi = SM == SM_CUBIC ? 8 : 16;
};

// One acquisition as the tile body sees it (device pointers).
struct RxAcq {
  const void* data;        // [F][W][N][L] complex64 (SM_COEF: coefficients)
  float2* out;             // [F][P]
  const float4* tx;        // [W][P] {td, ta, ta*cos(wd td), ta*sin(wd td)}
  const float4* rx;        // [N][P] {rd, ra, ra*cos(wd rd), ra*sin(wd rd)}; rx_compact: float2 {rd, ra}
  const int* tile_nc;      // [tiles] channels with a nonzero receive weight
  const int* tile_ncf;     // [tiles] leading entries of tile_chl with a weight other than NaN
  const unsigned long long* nan_mask;  // [tiles] lanes made NaN by the list's NaN-only tail
  const unsigned long long* stage_mask;  // [tiles][stage_words] staged list positions, or null
  int stage_words;
  const int* tile_chl;     // [tiles][N] their indices, ascending
  const int* tile_nw;      // [tiles] waves with a nonzero transmit weight
  const int* tile_waves;   // [tiles][W] their indices, ascending
  const float2* ref;       // [P] reference phase, or null
  int P, N, W, L;
  int tile_ps, tile_n0;    // the tiles' pixels: tile_pixel(tile_ps, tile_n0, tile, lane)
  float t0, fs;
  int frame_bytes;         // W * N * L * stride (< 0x7FFFFE00)
  int rx_compact;          // RF plans (wd = 0): the receive table holds float2 {rd, ra}
  int rx_otf;              // no receive table: entries computed in the block (rxotf_tile)
  const RxOtf* otf;        // device copy of the on-the-fly receive state's parameters
};

// Receive-table entry idx as {rd, ra, ra*cos(wd rd), ra*sin(wd rd)}: RF plans
// store {rd, ra} only (half the bytes), whose phase factor is (ra, ra * 0).
template <bool COMPACT>
__device__ __forceinline__ float4 rx_entry_t(const float4* __restrict__ rxt, size_t idx) {
  if (COMPACT) {
    const float2 v = reinterpret_cast<const float2*>(rxt)[idx];
    return make_float4(v.x, v.y, v.y, v.y * 0.0f);
  }
  return rxt[idx];
}
__device__ __forceinline__ float4 rx_entry(const float4* __restrict__ rxt, int compact, size_t idx) {
  return compact ? rx_entry_t<true>(rxt, idx) : rx_entry_t<false>(rxt, idx);
}

// Keys' cubic weights (a = -1/2) in Horner form.
__device__ __forceinline__ void keys_weights_horner(float b, float& c0, float& c1, float& c2, float& c3) {
  const float b2 = b * b;
  c0 = b * fmaf(b, fmaf(-0.5f, b, 1.0f), -0.5f);
  c1 = fmaf(b2, fmaf(1.5f, b, -2.5f), 1.0f);
  c2 = b * fmaf(b, fmaf(-1.5f, b, 2.0f), 0.5f);
  c3 = b2 * fmaf(0.5f, b, -0.5f);
}

// das_c's sample index of a term (das_c.cpp:372,375), FMA contraction off:
// the staged tile body and the plan's stage mask form it the same way.
__device__ __forceinline__ float term_index(float td, float rd, float t0, float fs) {
#pragma clang fp contract(off)
  return ((td + rd) - t0) * fs;
}

// Staged channels: every term of the channel on the tile reads its samples
// from one window of kStageWin samples of its trace, loaded by the wave in
// one coalesced 8-byte-per-lane load and read back from LDS. The window
// starts at the first sample lane 0's term reads (n0 - kBack); the plan marks
// a channel staged only when every lane's n0 lies in [n0 of lane 0, that +
// kMaxSpan], so that every sample a term reads is in the window.
constexpr int kStageWin = 64;
template <int SM>
struct StageWin {
  static constexpr int kBack = SM == SM_CUBIC ? 1 : 0;      // samples before n0 a term reads
  static constexpr int kMaxSpan = SM == SM_CUBIC ? 60 : 62;  // kStageWin - (samples a term reads)
};

// The transmit kernel's windows start kTxSlack samples before lane 0's
// first sample, since on tiles of several scan lines lane 0 need not have
// the tile's smallest index.
constexpr int kTxSlack = 4;

__device__ __forceinline__ float2 buf_load8(__amdgpu_buffer_rsrc_t r, int byte_off) {
  typedef unsigned u2v __attribute__((ext_vector_type(2)));
  const u2v v = __builtin_bit_cast(u2v, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0));
  return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}

// One channel of one split: part = sum_j T_j * sample_j over the split's
// waves, in batches of RB terms whose gathers are all issued before the
// first is used. Slots j >= nj are skipped when GUARD (a wave-uniform branch
// per batch); without it the slots past the tile's waves hold the far delay
// and read zeros. RB: 3 terms (6 gathers) in flight for Keys' cubic, 5 for
// the one-gather modes. For cubic the batch sets the kernel's registers: with
// 6 terms the C2 kernel needs 115 VGPRs and holds 4 waves per SIMD, with 3 it
// fits in 96 and holds 5 (C2 2.49 -> 2.38 ms; the staged channels keep their
// own batch of 6, StageBatch). Linear: the best of 3-6 on the 3-D and STA
// classes.
template <int SM>
struct GatherBatch {
  static constexpr int RB = SM == SM_CUBIC ? 3 : 5;
};
// Terms per batch of the staged channels (rxtab_channel_staged): one 8-byte
// window load per term, so batches are cheaper in registers; 6 and 8 measured
// alike on C2, 10 slower.
template <int SM>
struct StageBatch {
  static constexpr int RB = SM == SM_CUBIC ? 6 : 5;
};

template <int SM, int WC, bool GUARD, bool SAFE>
__device__ __forceinline__ float2 rxtab_channel(__amdgpu_buffer_rsrc_t rs, const float (&td)[WC],
                                                const float2 (&T)[WC], const int (&wb)[WC], int nj, float rd,
                                                int nb, float t0, float fs, float lim, bool& any) {
  using M = SampleMode<SM>;
  constexpr int RB = GatherBatch<SM>::RB;
  constexpr float lo_ok = 1.0f;  // das_c's window: 0 < trunc(d)
  float2 part = make_float2(0.0f, 0.0f);
#pragma unroll
  for (int j0 = 0; j0 < WC; j0 += RB) {
    if (GUARD && j0 >= nj) break;
    float4 q[RB][M::kLoads];
    float bb[RB];
    bool okb[RB];
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
        okb[i] = ok;
        const int n0 = (int)d;
        bb[i] = __builtin_amdgcn_fractf(d);
        // the first gather's offset folded into the wave-uniform base, so both
        // gathers address off and off + (kHi - kLo) (an immediate offset)
        const int off = ok ? n0 * M::kStride + (nb + wb[j] + M::kLo) : kOOB;
        q[i][0] = buf_load16(rs, off);
        if (M::kLoads == 2) q[i][M::kLoads - 1] = buf_load16(rs, off + (M::kHi - M::kLo));
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
        if (SAFE) {
          float2 t = part;
          cfma(t, T[j], sm);
          part = okb[i] ? t : part;
          any = any || okb[i];
        } else {
          cfma(part, T[j], sm);
        }
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

// rxtab_channel for a staged channel (every weight nonzero, every term in the
// sample window): the same sum, term for term. Per batch of RB terms the wave
// loads each term's window (lane l: sample base + l), writes the windows to
// its LDS slots win[i][64], and each lane reads its term's samples from there
// (the LDS operations of one wave complete in order, so the reads see the
// writes). Slots past the split's waves (GUARD) load zeros and have T = 0.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) f2v lds_f2;
typedef const __attribute__((address_space(3))) f4a8 lds_f4a8;

// x * 8 + y in one VALU instruction (y wave-uniform, in an SGPR).
__device__ __forceinline__ unsigned lshl3_add(unsigned x, unsigned y) {
  unsigned r;
  asm("v_lshl_add_u32 %0, %1, 3, %2" : "=v"(r) : "v"(x), "s"(y));
  return r;
}

template <int SM, int WC, bool GUARD>
__device__ __forceinline__ float2 rxtab_channel_staged(__amdgpu_buffer_rsrc_t rs, const float (&td)[WC],
                                                       const float2 (&T)[WC], const int (&wb)[WC], int nj,
                                                       float rd, int nb, float t0, float fs,
                                                       float2* __restrict__ win, unsigned lane8) {
  static_assert(SM != SM_COEF, "coefficient data is not staged");
  using S = StageWin<SM>;
  constexpr int RB = StageBatch<SM>::RB;
  const unsigned wbase = (unsigned)(size_t)(__attribute__((address_space(3))) char*)win;  // LDS address
  float2 part = make_float2(0.0f, 0.0f);
#pragma unroll
  for (int j0 = 0; j0 < WC; j0 += RB) {
    if (GUARD && j0 >= nj) break;
    float2 q[RB];
    float bb[RB];
    unsigned la[RB];
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int j = j0 + i;
      if (j < WC) {
        const float d = term_index(td[j], rd, t0, fs);
        const int n0 = (int)d;
        bb[i] = __builtin_amdgcn_fractf(d);
        const unsigned n00 = (unsigned)__builtin_amdgcn_readfirstlane(n0);
        const bool live = !GUARD || j < nj;  // wave-uniform
        // the window's byte offset in the frame (scalar) plus 8 * lane
        const unsigned sb = live ? (unsigned)(nb + wb[j]) + (n00 - (unsigned)S::kBack) * 8u : (unsigned)kOOB;
        q[i] = buf_load8(rs, (int)(sb + lane8));
        // LDS address of the first sample the term reads, in slot i
        la[i] = lshl3_add((unsigned)n0, wbase + (unsigned)(i * kStageWin * 8) - n00 * 8u);
      }
    }
#pragma unroll
    for (int i = 0; i < RB; ++i)
      if (j0 + i < WC) *reinterpret_cast<lds_f2*>((size_t)(wbase + lane8) + i * kStageWin * 8) = f2v{q[i].x, q[i].y};
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int j = j0 + i;
      if (j < WC) {
        const float b = bb[i];
        const lds_f4a8* wp = reinterpret_cast<const lds_f4a8*>((size_t)la[i]);
        float2 sm;
        if (SM == SM_LINEAR) {
          const f4a8 x = wp[0];
          sm = make_float2(fmaf(b, x.z - x.x, x.x), fmaf(b, x.w - x.y, x.y));
        } else {
          const f4a8 lo = wp[0];
          const f4a8 hi = wp[1];
          float c0, c1, c2, c3;
          keys_weights_horner(b, c0, c1, c2, c3);
          sm = make_float2(fmaf(c3, hi.z, fmaf(c2, hi.x, fmaf(c1, lo.z, c0 * lo.x))),
                           fmaf(c3, hi.w, fmaf(c2, hi.y, fmaf(c1, lo.w, c0 * lo.y))));
        }
        cfma(part, T[j], sm);
      }
    }
    asm volatile("" : "+v"(part.x), "+v"(part.y));
    __builtin_amdgcn_wave_barrier();  // this batch's reads before the next batch's writes
  }
  return part;
}

// One channel's fold: acc += R * (its partial sum over the split's waves).
template <int SM, int WC, bool GUARD, bool SAFE>
__device__ __forceinline__ void rxtab_fold(__amdgpu_buffer_rsrc_t rs, const float (&td)[WC], const float2 (&T)[WC],
                                           const int (&wb)[WC], int nj, float4 rv, int nb, float t0, float fs,
                                           float lim, float2& acc) {
  const bool ron = rv.y != 0.0f;
  const float rd = ron ? rv.x : kFar;
  const float2 R = ron ? make_float2(rv.z, rv.w) : make_float2(0.0f, 0.0f);
  bool any = false;
  const float2 part = rxtab_channel<SM, WC, GUARD, SAFE>(rs, td, T, wb, nj, rd, nb, t0, fs, lim, any);
  if (SAFE) {
    float2 t = acc;
    cfma(t, R, part);
    acc = any ? t : acc;
  } else {
    cfma(acc, R, part);
  }
}

template <int SM, int WC, bool GUARD, bool SAFE, bool COMPACT, bool STG>
__device__ __forceinline__ void rxtab_channels(__amdgpu_buffer_rsrc_t rs, const float (&td)[WC],
                                               const float2 (&T)[WC], const int (&wb)[WC], int nj,
                                               const float4* __restrict__ rxt, const int* __restrict__ chl,
                                               int cstride, int nc, size_t PS, int p, bool active, int lst,
                                               float t0, float fs, float lim, float2& acc,
                                               const unsigned long long* __restrict__ smask, int pos0,
                                               float2* __restrict__ win, unsigned lane8) {
  // channels chl[0], chl[cstride], ... (nc of them); the receive row of the
  // next one in flight during this one, its index one further ahead (a
  // scalar load off the critical path). (Four rows in flight for splits of
  // few waves measured slower: the registers cost more occupancy than the
  // latency they hide.)
  float4 rn = make_float4(0.f, 0.f, 0.f, 0.f);
  int n = __builtin_amdgcn_readfirstlane(chl[0]);
  int n_next = nc > 1 ? __builtin_amdgcn_readfirstlane(chl[cstride]) : 0;
  if (active) rn = rx_entry_t<COMPACT>(rxt, (size_t)n * PS + p);
  for (int k = 0; k < nc; ++k) {
    const float4 rv = rn;
    const int nb = n * lst;
    if (k + 1 < nc) {
      n = n_next;
      if (active) rn = rx_entry_t<COMPACT>(rxt, (size_t)n * PS + p);
      if (k + 2 < nc) n_next = __builtin_amdgcn_readfirstlane(chl[(k + 2) * cstride]);
    }
    // a staged channel (list position pos0 + k * cstride) has every weight
    // nonzero: R is its entry's factor, as in rxtab_fold
    if constexpr (STG) {
      // the mask word through the constant address space (a scalar load; the
      // plan's mask is read-only while kernels run)
      const int pos = pos0 + k * cstride;
      if (smask && ((((const __attribute__((address_space(4))) unsigned long long*)smask)[pos >> 6] >> (pos & 63)) &
                    1ull)) {
        const float2 part = rxtab_channel_staged<SM, WC, GUARD>(rs, td, T, wb, nj, rv.x, nb, t0, fs, win, lane8);
        cfma(acc, make_float2(rv.z, rv.w), part);
        continue;
      }
    }
    rxtab_fold<SM, WC, GUARD, SAFE>(rs, td, T, wb, nj, rv, nb, t0, fs, lim, acc);
  }
}

// Receive-table layout of a kernel: float4 only, float2 only (RF plans), or
// either, chosen per acquisition at run time (the batch kernel).
enum { RXL_FULL = 0, RXL_COMPACT = 1, RXL_EITHER = 2 };

// One tile of one frame; every thread of the block calls it (it ends with a
// block barrier and the split-ordered reduction through red). STG: the
// channels A.stage_mask marks run staged, split s through the LDS slots
// win[s * StageBatch<SM>::RB * kStageWin ..] (SAFE tiles never).
template <int SM, int WC, int WS, bool SAFE = false, int RXL = RXL_EITHER, bool STG = false>
__device__ __forceinline__ void rxtab_tile(const RxAcq& A, int tile, int f, float2 (*red)[64],
                                           float2* win = nullptr) {
  constexpr int TP = 64;
  constexpr int ST = SampleMode<SM>::kStride;
  const int lane = threadIdx.x & (TP - 1);
  const int s = __builtin_amdgcn_readfirstlane((int)threadIdx.x / TP);  // wave-uniform
  const int p = tile * TP + lane;
  const bool active = p < A.P;
  const size_t PS = (size_t)A.P;
  // SAFE: the splits run the list's leading channels with a weight other
  // than NaN; the NaN-only tail enters through the plan's lane mask
  const int nc = __builtin_amdgcn_readfirstlane(SAFE ? A.tile_ncf[tile] : A.tile_nc[tile]);
  const int nw = __builtin_amdgcn_readfirstlane(A.tile_nw[tile]);
  const int* __restrict__ wl = A.tile_waves + (size_t)tile * A.W;
  const int* __restrict__ chl = A.tile_chl + (size_t)tile * A.N;
  const float4* __restrict__ txt = A.tx;
  const int lst = A.L * ST;   // bytes per trace
  const int wst = lst * A.N;  // bytes per wave
  const __amdgpu_buffer_rsrc_t rs =
      make_rsrc((const char*)A.data + (size_t)f * A.frame_bytes, A.frame_bytes);
  const float lim = (float)(SM == SM_LINEAR ? A.L - 1 : A.L - 2);
  const int compact = RXL == RXL_EITHER ? __builtin_amdgcn_readfirstlane(A.rx_compact) : (RXL == RXL_COMPACT);

  // The splits share out the tile's waves round-robin, WC per split and round.
  // When the waves need fewer splits than WS in one round (wneed), the block
  // forms cg = WS / wneed channel groups of wneed splits: group g takes the
  // channels chl[g], chl[g + cg], ... and every wave (one round). All of it is
  // wave-uniform; the reduction below sums the splits' partials in order.
  const int wneed = (nw + WC - 1) / WC;
  const int cg = (wneed > 0 && wneed < WS) ? WS / wneed : 1;
  const int wsp = cg > 1 ? wneed : WS;  // splits per channel group
  const int g = s / wsp;
  const int sw = s - g * wsp;           // split within the group
  const int ncg = g < cg ? (nc - g + cg - 1) / cg : 0;
  constexpr bool kStg = STG && !SAFE && SM != SM_COEF;
  const unsigned long long* smask =
      kStg && A.stage_mask ? A.stage_mask + (size_t)tile * (size_t)A.stage_words : nullptr;
  float2* swin = kStg ? win + s * (StageBatch<SM>::RB * kStageWin) : nullptr;
  // 8 * lane, opaque to the compiler so that window addresses stay one
  // scalar-plus-vector add
  unsigned lane8 = (unsigned)lane * 8u;
  asm volatile("" : "+v"(lane8));

  float2 acc = make_float2(0.0f, 0.0f);
  for (int k0 = 0; ncg > 0 && k0 < nw; k0 += wsp * WC) {
    // number of this split's waves in the round (wave-uniform)
    const int nj = min(WC, max(0, (nw - k0 - sw + wsp - 1) / wsp));
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
        const int w = __builtin_amdgcn_readfirstlane(wl[k0 + sw + wsp * j]);
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
    if (nj == WC) {
      if (compact)
        rxtab_channels<SM, WC, false, SAFE, true, kStg>(rs, td, T, wb, nj, A.rx, chl + g, cg, ncg, PS, p, active,
                                                        lst, A.t0, A.fs, lim, acc, smask, g, swin, lane8);
      else
        rxtab_channels<SM, WC, false, SAFE, false, kStg>(rs, td, T, wb, nj, A.rx, chl + g, cg, ncg, PS, p, active,
                                                         lst, A.t0, A.fs, lim, acc, smask, g, swin, lane8);
    } else {
      if (compact)
        rxtab_channels<SM, WC, true, SAFE, true, kStg>(rs, td, T, wb, nj, A.rx, chl + g, cg, ncg, PS, p, active,
                                                       lst, A.t0, A.fs, lim, acc, smask, g, swin, lane8);
      else
        rxtab_channels<SM, WC, true, SAFE, false, kStg>(rs, td, T, wb, nj, A.rx, chl + g, cg, ncg, PS, p, active,
                                                        lst, A.t0, A.fs, lim, acc, smask, g, swin, lane8);
    }
  }
  red[s][lane] = acc;
  __syncthreads();
  if (s == 0 && active) {
    float2 v = red[0][lane];
#pragma unroll
    for (int k = 1; k < WS; ++k) cadd(v, red[k][lane]);
    if (A.ref) v = cmul(v, A.ref[p]);
    if (SAFE && ((A.nan_mask[tile] >> lane) & 1ull)) v = make_float2(__builtin_nanf(""), __builtin_nanf(""));
    A.out[(size_t)f * PS + p] = v;
  }
}

// The same tile as rxtab_tile with the receive state computed in the block
// (RxAcq::rx_otf, no receive table): the tile's channel list is taken kOtfChunk
// entries at a time; the block's waves compute the entries {rd, ra, R} of
// the chunk for the tile's 64 pixels (wave s every WS-th one, rx_otf_entry)
// into LDS, and after a barrier every split folds the chunk's channels of its
// group from there, in list order, as rxtab_channels does from the table. All
// waves run every round and chunk (the barriers); a split with no waves in a
// round only computes. Given the same entries the sums are rxtab_tile's, term
// for term. (Two LDS buffers, filling the next chunk while folding this one
// with one barrier per chunk, measured no faster on the 3-D class.)
constexpr int kOtfChunk = 16;

template <int SM, int WC, int WS, bool SAFE>
__device__ __forceinline__ void rxotf_tile(const RxAcq& A, int tile, int f, float2 (*red)[64],
                                           float4 (*rxs)[64]) {
  constexpr int TP = 64;
  constexpr int CC = kOtfChunk;
  constexpr int ST = SampleMode<SM>::kStride;
  const int lane = threadIdx.x & (TP - 1);
  const int s = __builtin_amdgcn_readfirstlane((int)threadIdx.x / TP);  // wave-uniform
  const int p = tile * TP + lane;
  const bool active = p < A.P;
  const size_t PS = (size_t)A.P;
  const int nc = __builtin_amdgcn_readfirstlane(SAFE ? A.tile_ncf[tile] : A.tile_nc[tile]);
  const int nw = __builtin_amdgcn_readfirstlane(A.tile_nw[tile]);
  const int* __restrict__ wl = A.tile_waves + (size_t)tile * A.W;
  const int* __restrict__ chl = A.tile_chl + (size_t)tile * A.N;
  const float4* __restrict__ txt = A.tx;
  const int lst = A.L * ST;
  const int wst = lst * A.N;
  const __amdgpu_buffer_rsrc_t rs =
      make_rsrc((const char*)A.data + (size_t)f * A.frame_bytes, A.frame_bytes);
  const float lim = (float)(SM == SM_LINEAR ? A.L - 1 : A.L - 2);

  // waves and channel groups as rxtab_tile
  const int wneed = (nw + WC - 1) / WC;
  const int cg = (wneed > 0 && wneed < WS) ? WS / wneed : 1;
  const int wsp = cg > 1 ? wneed : WS;
  const int g = s / wsp;
  const int sw = s - g * wsp;
  const int rounds = nc > 0 ? (nw + wsp * WC - 1) / (wsp * WC) : 0;  // block-uniform

  float2 acc = make_float2(0.0f, 0.0f);
  for (int r = 0; r < rounds; ++r) {
    const int k0 = r * wsp * WC;
    const int nj = g < cg ? min(WC, max(0, (nw - k0 - sw + wsp - 1) / wsp)) : 0;
    float td[WC];
    float2 T[WC];
    int wb[WC];
#pragma unroll
    for (int j = 0; j < WC; ++j) {
      td[j] = kFar;
      T[j] = make_float2(0.0f, 0.0f);
      wb[j] = 0;
      if (j < nj) {
        const int w = __builtin_amdgcn_readfirstlane(wl[k0 + sw + wsp * j]);
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
    for (int c0 = 0; c0 < nc; c0 += CC) {
      const int cn = min(CC, nc - c0);
      {
        // the parameters and the pixel's coordinates are read again for every
        // chunk (scalar loads; the empty asm keeps them from being hoisted),
        // so that none of it stays live across the term loop
        const RxOtf* op = A.otf;
        asm volatile("" : "+s"(op));
        double px = 0.0, py = 0.0, pz = 0.0, rref = 0.0;
        pixel_xyz(op->scan, active ? p : 0, px, py, pz, rref);
        const OtfPixel q = otf_pixel(*op, px, py, pz);
        for (int e = s; e < cn; e += WS) {
          const int n = __builtin_amdgcn_readfirstlane(chl[c0 + e]);
          rxs[e][lane] = active ? rx_otf_entry(*op, q, n) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
      }
      __syncthreads();
      if (nj > 0) {
        // this group's list entries k = g, g + cg, ... within the chunk
        for (int e = ((g - c0 % cg) % cg + cg) % cg; e < cn; e += cg) {
          const int nb = __builtin_amdgcn_readfirstlane(chl[c0 + e]) * lst;
          const float4 rv = rxs[e][lane];
          if (nj == WC)
            rxtab_fold<SM, WC, false, SAFE>(rs, td, T, wb, nj, rv, nb, A.t0, A.fs, lim, acc);
          else
            rxtab_fold<SM, WC, true, SAFE>(rs, td, T, wb, nj, rv, nb, A.t0, A.fs, lim, acc);
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
    if (A.ref) v = cmul(v, A.ref[p]);
    if (SAFE && ((A.nan_mask[tile] >> lane) & 1ull)) v = make_float2(__builtin_nanf(""), __builtin_nanf(""));
    A.out[(size_t)f * PS + p] = v;
  }
}

// Descriptor of a plan's variant-2/3 tables for the tile body.
inline RxAcq rx_acq(const GeomLaunch& G, int sm) {
  RxAcq A;
  A.data = sm == SM_COEF ? (const void*)G.coef : G.ch;
  A.out = reinterpret_cast<float2*>(G.out);
  A.tx = G.tx_table;
  A.rx = G.rx_table;
  A.tile_nc = G.tile_nc;
  A.tile_ncf = G.tile_ncf;
  A.nan_mask = G.nan_mask;
  A.stage_mask = G.stage && sm != SM_COEF ? G.stage_mask : nullptr;
  A.stage_words = G.stage_words;
  A.tile_chl = G.tile_chl;
  A.tile_nw = G.tile_nw;
  A.tile_waves = G.tile_waves;
  A.ref = G.ref_phase ? G.ref_table : nullptr;
  A.P = G.P;
  A.N = G.N;
  A.W = G.W;
  A.L = G.L;
  A.tile_ps = G.tile_ps;
  A.tile_n0 = G.scan.n0;
  A.t0 = G.t0;
  A.fs = G.fs;
  A.frame_bytes = G.L * G.N * G.W * (sm == SM_COEF ? 32 : 8);
  A.rx_compact = G.rx_compact;
  A.rx_otf = G.rx_otf;
  A.otf = G.d_otf;
  return A;
}

}  // namespace ustbx
