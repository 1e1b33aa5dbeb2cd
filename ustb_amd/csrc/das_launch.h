// das_launch.h — host-side launch descriptors shared by the kernel files and
// the C-ABI (das_capi.hip). Plain structs; device pointers only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "das_common.h"

namespace ustbx {

struct MatLaunch {
  const void* ch;
  const float* tx_apo;
  const float* rx_apo;
  const float* tx_delay;
  const float* rx_delay;
  void* out;
  int L, N, W, F, P;
  float fs, t0, wd;
  int dim, interp, edge, split;
  bool py;         // the Python reference's arithmetic (USTBX_FLAG_PYTHON_ARITHMETIC)
  float dt, w0;    // its sample period and angular frequency
};

hipError_t launch_das_matrix(const MatLaunch& m, hipStream_t stream);

// Fills n complex64 values with zeros (a kernel, so it can be graph-captured).
hipError_t launch_zero_fill(void* p, size_t n_complex, hipStream_t stream);

// Geometry (on-the-fly) path: everything resident in device memory.
struct GeomLaunch {
  const void* ch;
  void* out;
  int L, N, W, F, P;
  float fs, t0, wd;
  int dim, interp, edge;
  int separable;       // exp(i wd tx) * exp(i wd rx) instead of exp(i wd (tx+rx))
  int ref_phase;       // multiply by exp(-1j*2*w0*r_ref/c)
  double w0_ref;       // 2*w0/c for the reference phase, in FP64
  int ref_single;      // the reference phase's argument in FP32 (das.m) instead of FP64 (das.py)
  float c32;           // single(c) for it
  DevScan scan;
  const double* ex;    // element positions [N]
  const double* ey;
  const double* ez;
  const DevWave* waves;  // [W]
  DevRxApod rx;
  DevTxApod tx;
  const float* tx_delay_mat;  // [W][P] or null
  double c, lens, pw_margin, bpow;
  int wave_chunk;      // waves per register chunk (autotuned)
  int split;           // channel groups per pixel tile (BOTH/RX)
  // Resident transmit table [W][P] of {tx delay, tx apodization,
  // tx apodization * exp(i wd tx delay)}: constant for a plan, built once.
  const float4* tx_table;
  // Receive aperture culling: elements sorted by x on one line (y, z equal),
  // a window with support ratio <= 1/2 and no tilt, so the channels a pixel
  // can see form one contiguous index range.
  int rx_cull;
  // Kernel variant: 0 = per-lane wave chunks (all modes), 1 = receive state
  // shared through LDS across wave-splits (dimension both, separable phase,
  // one frame), 2 = resident receive table (dimension both with the
  // separable or RF phase; dimension transmit).
  int variant;
  // Variant 2 state. Resident: receive table [N][P] of {rx delay, rx
  // apodization, rx apodization * exp(i wd rx delay)}; per 64-pixel tile the
  // lists of channels ([tiles][N], count in tile_nc) and of waves ([tiles][W],
  // count in tile_nw) with a nonzero weight on some pixel of the tile; the
  // reference phase per pixel (or null). Per run (variant 3): interpolation
  // coefficients [F][W][N][L] of two float4.
  const float4* rx_table;
  int rx_compact;       // RF plans: the receive table holds float2 {rd, ra}
  // Receive state on the fly (variant 4): no receive table; the blocks
  // compute each (pixel, channel) entry from otf (rx_otf_entry).
  int rx_otf;
  RxOtf otf;
  const RxOtf* d_otf;  // its device copy (the tile bodies read it from there)
  const int* tile_nc;
  const int* tile_ncf;  // [tiles] leading entries of the channel list with a weight other than NaN
  const unsigned long long* nan_mask;  // [tiles] lanes whose sum the NaN-only channels make NaN
  // Dimension both (receive table, linear or cubic): per tile, stage_words
  // 64-bit words; bit k set when the channel at list position k can read
  // its samples from a 64-sample LDS window (rxtab_channel_staged), or null.
  const unsigned long long* stage_mask;
  int stage_words;
  const int* tile_chl;
  const int* tile_waves;
  const int* tile_nw;
  const float2* ref_table;
  float4* coef;
  // Variant 2/3 tile visiting order: block b takes tile tile_order[x(b)],
  // x = the XCD-contiguous remap of b; null = x(b) itself. Its first
  // tile_good entries are the tiles whose nonzero weights and their delays
  // are all finite; the order is non-null when tile_good < tiles.
  const int* tile_order;
  int tile_good;
  // Pixels of the variant-2 tiles (tile_pixel): 0 = 64 consecutive pixels;
  // 4 = patches of 16 depths x 4 scan lines (dimension transmit on a linear
  // or sector scan whose axes divide evenly). Fixed when the plan is made.
  int tile_ps;
  // Stage sample windows through LDS (1) or gather every term's samples
  // (0): dimension both through stage_mask, dimension transmit (variant 2)
  // per (channel, wave). Results do not depend on it.
  int stage;
};

// Builds the transmit table of a plan (one lane per (wave, pixel)).
hipError_t launch_geom_tx_table(const GeomLaunch& g, float4* table, hipStream_t stream);
hipError_t launch_geom_unified_tx(const GeomLaunch& G, const DevTxApod& mask_apod, float* out,
                                  hipStream_t stream);
hipError_t launch_geom_active_elements(const GeomLaunch& G, int dim, double crit, float* M, hipStream_t stream);
hipError_t launch_tx_cf(const GeomLaunch& G, float4* part, float2* image, float* cf, const float* M, int flavour,
                        hipStream_t stream);
size_t tx_cf_workspace_bytes(const GeomLaunch& G);

// Builds the variant-2 tables; tile_nc is [2][tiles] (channel counts, then
// the counts of the list's leading channels with a weight other than NaN);
// tile_bad[t] = 1 where a nonzero weight or its delay is not finite.
hipError_t launch_rxtab_tables(const GeomLaunch& g, float4* rx_table, float2* ref_table, int* tile_nc,
                               int* tile_chl, int* tile_waves, int* tile_nw, int* tile_bad,
                               unsigned long long* nan_mask, hipStream_t stream);
// Rounds the element positions to FP32 ([N] float4) for the on-the-fly state.
hipError_t launch_otf_elements(const GeomLaunch& g, float4* e32, hipStream_t stream);
// Fills stage_mask ([tiles][stage_words]) from the variant-2 tables (tile_bad
// as launch_rxtab_tables left it; those tiles get no bits).
hipError_t launch_stage_mask(const GeomLaunch& g, const int* tile_bad, unsigned long long* stage_mask,
                             hipStream_t stream);

hipError_t launch_das_geom(const GeomLaunch& g, hipStream_t stream);

// Adds the number of terms with nonzero transmit and receive weight (per
// frame) to *total (device).
hipError_t launch_count_active(const GeomLaunch& g, unsigned long long* total, hipStream_t stream);

// Writes the geometry's FP32 matrices (das.m's inputs to the MEX).
hipError_t launch_geom_materialize(const GeomLaunch& g, float* tx_apo,
                                   float* rx_apo, float* tx_delay,
                                   float* rx_delay, hipStream_t stream);

}  // namespace ustbx
