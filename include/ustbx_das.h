/*
 * ustbx_das.h — C-ABI of the MI355X (gfx950) delay-and-sum beamformer.
 *
 * Plain C types only: sizes, scalars and pointers. No torch, no HIP types in
 * the signatures (streams travel as `void*`, a hipStream_t; NULL = default).
 *
 * Array layouts follow the reference MEX ABI (MATLAB column-major, first index
 * fastest), i.e. in row-major terms:
 *   channel data  ch [F][W][N][L]   complex64 (interleaved re, im)
 *   tx_apo/tx_delay [W][P]          float32
 *   rx_apo/rx_delay [N][P]          float32
 *   output        out[F][W'][N'][P] complex64, N' = N for NONE/TX else 1,
 *                                   W' = W for NONE/RX else 1
 *
 * Every entry point returns 0 on success or a negative USTBX_E* code; when
 * `st` is not NULL it also receives the code and a message. Nothing longjmps
 * and nothing aborts: a MEX gateway maps the codes to das_c's error
 * identifiers (ustbx_status_id()).
 */
#ifndef USTBX_DAS_H
#define USTBX_DAS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define USTBX_ABI_VERSION 3

/* error codes, with the identifier das_c raises each class with
 * (ustbx_status_id; +mex/source/das_c.cpp:53-252, +mex/source/das_cuda.cu:122).
 * "" = das_c raises it through mexErrMsgTxt, with no identifier. */
#define USTBX_OK 0
#define USTBX_E_NRHS (-1)       /* Toolbox:SRP_SRC:nrhs        (das_c.cpp:53-54) */
#define USTBX_E_NLHS (-2)       /* Toolbox:SRP_SRC:nlhs        (das_c.cpp:55) */
#define USTBX_E_DIMENSIONS (-3) /* Toolbox:SRP_SRC:Dimensions  (das_c.cpp:88-199) */
#define USTBX_E_FLOAT (-4)      /* "" (das_c.cpp:98, 145-236: not single precision) */
#define USTBX_E_SCALAR (-5)     /* "" (das_c.cpp:208, 221, 234: not a scalar) */
#define USTBX_E_GPU (-6)        /* Toolbox:SRP_SRC:GPU         (das_cuda.cu:122) */
#define USTBX_E_ARG (-7)        /* "" (sum mode das_c.cpp:79, real data with fd
                                   das_c.cpp:242; interp, NULL pointers) */
#define USTBX_E_CAPACITY (-8)   /* ustbx:capacity: a request exceeds a stated
                                   capacity (no reference counterpart) */
#define USTBX_E_UNSUPPORTED (-9)/* ustbx:unsupported: feature not available on
                                   the device path (no reference counterpart) */

typedef struct ustbx_status {
  int32_t code;
  char message[256];
} ustbx_status;

/* sum modes: the reference `dimension` enum (dimension.m:27-32) */
#define USTBX_DIM_NONE 0
#define USTBX_DIM_RX 1
#define USTBX_DIM_TX 2
#define USTBX_DIM_BOTH 3

/* interpolation */
#define USTBX_INTERP_LINEAR 0 /* das_c / das_cuda / interp1 'linear' */
#define USTBX_INTERP_CUBIC 1  /* Keys cubic convolution, a = -1/2 (no reference counterpart) */

/* sample-window rule at the trace edges */
#define USTBX_EDGE_DASC 0   /* 0 < trunc(d) < L-1           (das_c.cpp:377,381) */
#define USTBX_EDGE_PYTHON 1 /* 0 <= floor(d) < L-1          (python das.py:233-236) */

/* flags */
#define USTBX_FLAG_DASC_ORDER 1u /* one thread per pixel, das_c summation order
                                    (w outer, n inner) and das_c's phase:
                                    bit-identical to das_c whenever fd == 0;
                                    overrides USTBX_FLAG_SEPARABLE_PHASE */
#define USTBX_FLAG_SEPARABLE_PHASE 2u /* geometry path: exp(i wd (tx+rx)) taken as
                                         exp(i wd tx) * exp(i wd rx) */
#define USTBX_FLAG_REF_PHASE 4u  /* geometry path: multiply by
                                    exp(-1j*2*w0*r_ref/c) (das.m:306-309), the
                                    argument formed in FP64 as python
                                    das.py:169-171 forms it */
#define USTBX_FLAG_REF_PHASE_SINGLE 16u /* with USTBX_FLAG_REF_PHASE: the
   argument formed in single precision as das.m:308 evaluates it,
   fl32(fl32(-2*w0 * single(r_ref)) / single(c)) (MATLAB's products of single
   and double operands are single). The factor is the sine and cosine of the
   argument rounded to FP32 in both flavours. At 55 mm and 5.2 MHz the two
   differ by about 1e-4 (the FP32 rounding of a 2.3e3 rad argument). */
#define USTBX_FLAG_PYTHON_ARITHMETIC 8u /* matrix ABI, with USTBX_EDGE_PYTHON and
   linear interpolation: the Python reference's DAS._beamform arithmetic
   (python das.py:217-250). `fs` holds the sample period dt [s] of the time
   vector and `fd` holds w0 [rad/s]. Index (tau - t0) / dt in FP32; floor and
   fraction exact; interpolation in FP64 rounded to complex64; times
   rx_apo * tx_apo; times exp(1j * fl32(w0 * tau)) in complex64 where w0 != 0.
   Every term with a nonzero (or NaN) apo * weight product is added, in or
   out of the window (out: apo * 0), so NaN weights propagate as in the
   reference. With USTBX_FLAG_DASC_ORDER the sums run in its loop order (w
   outer, n inner). Limits: NaN channel data where apo is zero, and a NaN
   weight whose other weight column is zero at every pixel (the reference
   skips such columns, das.py:218-222), are not reproduced. */

/* ---------------------------------------------------------------------------
 * 1. Matrix ABI: the das_c / das_cuda MEX contract.
 * Replaces: mex.das_c(ch, fs, t0, tx_apo, rx_apo, tx_delay, rx_delay, fd, dim)
 *   (+mex/source/das_c.cpp:46-418, called from +midprocess/das.m:251-259) and
 *   mex.das_cuda(..., dim, gpu_id) (+mex/source/das_cuda.cu:128-306,
 *   das.m:262-271), which only supports dim == BOTH (das_cuda.cu:321-324);
 *   this entry supports all four modes.
 * ------------------------------------------------------------------------- */
typedef struct ustbx_das_desc {
  int32_t L, N, W, F; /* samples, channels, waves, frames */
  int32_t P;          /* pixels */
  float fs;           /* sampling frequency [Hz]   (das_c.cpp:212) */
  float t0;           /* initial time [s]          (das_c.cpp:225) */
  float fd;           /* modulation frequency [Hz] (das_c.cpp:238) */
  int32_t dim;        /* USTBX_DIM_*               (das_c.cpp:78)  */
  int32_t interp;     /* USTBX_INTERP_*            */
  int32_t edge;       /* USTBX_EDGE_*              */
  uint32_t flags;     /* USTBX_FLAG_*              */
  int32_t split;      /* channel groups per pixel tile: 1, 2, 4, 8 (0 = auto: 8) */
} ustbx_das_desc;

/* All pointers are device (HBM) pointers; `out` is fully written (no need to
 * clear it). Asynchronous on `stream`. */
int ustbx_das_run(const ustbx_das_desc* d, const void* ch, const float* tx_apo,
                  const float* rx_apo, const float* tx_delay,
                  const float* rx_delay, void* out, void* stream,
                  ustbx_status* st);

/* Same contract on host buffers (what a MEX gateway holds): stages to the
 * device, runs and copies the result back; synchronous. `device` selects the
 * GPU like das_cuda's gpu_id (das_cuda.cu:161-162); the calling thread's
 * current device is left as it was. */
int ustbx_das_run_host(const ustbx_das_desc* d, const void* ch,
                       const float* tx_apo, const float* rx_apo,
                       const float* tx_delay, const float* rx_delay, void* out,
                       int32_t device, ustbx_status* st);

/* Validation of the das_c MEX argument contract on plain dimensions (what
 * das_c.cpp:53-252 checks on the mxArrays): ndim/dims of each argument, the
 * scalar sizes and the class flags. Returns USTBX_OK or the code das_c would
 * raise; fills P/L/N/W/F of `d`. */
int ustbx_das_check_mex_args(int32_t nrhs, int32_t nlhs, int32_t ch_ndim,
                             const int64_t* ch_dims, int32_t ch_is_double,
                             const int32_t* mat_ndim, const int64_t* mat_dims,
                             const int32_t* mat_is_double,
                             const int64_t* scalar_numel,
                             const int32_t* scalar_is_double, int32_t sum_mode,
                             int32_t ch_is_complex, float fd,
                             ustbx_das_desc* d, ustbx_status* st);

/* Reference error identifier for a code ("Toolbox:SRP_SRC:Dimensions", ...;
 * "" where das_c raises the error with no identifier). */
const char* ustbx_status_id(int32_t code);

/* ---------------------------------------------------------------------------
 * 2. Geometry plans: delays and apodization computed on the device.
 * Replaces the host-side half of +midprocess/das.m:86-206 (tx/rx apodization,
 * receive delay, transmit delay per wavefront) that feeds the MEX, and fuses
 * the reference phase of das.m:306-309. A plan is device-resident state for
 * one probe/scan/sequence configuration, reused across frames.
 * ------------------------------------------------------------------------- */

/* scan kinds */
#define USTBX_SCAN_LINEAR 0   /* ndgrid(z, x, y), z fastest (linear_scan.m:71)    */
#define USTBX_SCAN_SECTOR 1   /* ndgrid(depth, az, el) + sph2cart (sector_scan.m:90-112) */
#define USTBX_SCAN_XYZ 2      /* explicit pixel list */

/* wavefronts (uff.wavefront) */
#define USTBX_WAVE_PLANE 0
#define USTBX_WAVE_SPHERICAL 1
#define USTBX_WAVE_PHOTOACOUSTIC 2

/* spherical_transmit_delay_model (spherical_transmit_delay_model.m:32-37) */
#define USTBX_TXMODEL_SPHERICAL 1
#define USTBX_TXMODEL_UNIFIED 2 /* computed once per plan (tools.calculate_unified_delay_model;
                                   2-D linear and sector scans) */
#define USTBX_TXMODEL_HYBRID 3
#define USTBX_TXMODEL_BLENDED 4

/* windows (uff.window, window.m:24-37) */
#define USTBX_WIN_NONE 0
#define USTBX_WIN_BOXCAR 1
#define USTBX_WIN_HANNING 2
#define USTBX_WIN_HAMMING 3
#define USTBX_WIN_TUKEY25 4
#define USTBX_WIN_TUKEY50 5
#define USTBX_WIN_TUKEY75 6
#define USTBX_WIN_STA 7      /* host: reduces to a per-element vector */
#define USTBX_WIN_SCANLINE 8 /* tx: MLA scan-line blocks (apodization.m:265-306) */
#define USTBX_WIN_TRIANGLE 9

typedef struct ustbx_scan {
  int32_t kind;
  int32_t n0, n1, n2;   /* axis lengths, fastest first; P = n0*n1*n2 */
  const double* a0;     /* LINEAR: z_axis[n0]; SECTOR: depth_axis[n0]; XYZ: x[P] */
  const double* a1;     /* LINEAR: x_axis[n1]; SECTOR: {sin az, cos az}[n1]; XYZ: y[P] */
  const double* a2;     /* LINEAR: y_axis[n2]; SECTOR: {sin el, cos el}[n2]; XYZ: z[P] */
  double origin[3];     /* SECTOR: scan origin (sector_scan.m:101-104) */
  int32_t ref_kind;     /* reference distance: 0 = z (linear_scan.m:128-130),
                           1 = depth rho (sector_scan.m:167-169) */
} ustbx_scan;

typedef struct ustbx_wave {
  int32_t wavefront;    /* USTBX_WAVE_* */
  int32_t _pad;
  double src[3];        /* source x, y, z [m] (point.m:121-144) */
  double src_distance;  /* [m], +inf for plane */
  double azimuth, elevation;             /* source angles [rad] (point.m:28-30) */
  double cos_az, sin_az, cos_el, sin_el; /* cos/sin of those angles */
  double sound_speed;   /* wave.sound_speed (wave.m:55) */
  double delay;         /* wave.delay [s]  (wave.m:54) */
  double origin[3];     /* wave.origin (wave.m:46) */
} ustbx_wave;

/* Probe classes whose apodization geometry differs (apodization.m:377-400,
 * 506-518): flat (linear, matrix, any element list) or curvilinear
 * (uff.curvilinear_array, uff.curvilinear_matrix_array: the receive aperture
 * in azimuth around the origin, default (0, 0, -radius); the STA transmit
 * axis from (0, 0, -radius) to the source). */
#define USTBX_PROBE_FLAT 0
#define USTBX_PROBE_CURVILINEAR 1

typedef struct ustbx_apod {
  int32_t window;       /* USTBX_WIN_*; NONE -> ones */
  int32_t _pad;
  double f_number[2];
  double minimum_aperture[2];
  double maximum_aperture[2];
  double tilt[2];
  double grating_lobe_angle[2];
  int32_t has_origin;   /* rx: apodization.origin set (incidence_aperture) */
  int32_t _pad2;
  double origin[3];
  const float* vector;  /* optional per-element (rx) / per-wave (tx) vector
                           (apodization_vector, window.sta); host pointer */
  const float* matrix;  /* optional explicit matrix ([N][P] rx, [W][P] tx);
                           DEVICE pointer, used as is */
  int32_t mla[2];       /* tx, USTBX_WIN_SCANLINE: scan lines per wave in
                           x / azimuth and y / elevation (apodization.m:47;
                           0 = 1) */
  int32_t mla_overlap[2]; /* lines each wave's block is widened by, as a box
                           filter (apodization.m:48) */
} ustbx_apod;

typedef struct ustbx_geometry {
  int32_t N;                  /* elements */
  int32_t W;                  /* waves    */
  int32_t L;                  /* samples  */
  int32_t F;                  /* frames   */
  const double* elem_x;       /* [N] element positions (probe.m:42) */
  const double* elem_y;
  const double* elem_z;
  ustbx_scan scan;
  const ustbx_wave* waves;    /* [W] host array */
  ustbx_apod rx_apod;
  ustbx_apod tx_apod;
  double sound_speed;         /* channel_data.sound_speed */
  double lens_delay;          /* das.lens_delay (das.m:57) */
  double pw_margin;           /* das.pw_margin (das.m:55) */
  double blending_power;      /* das.blending_power (das.m:58) */
  int32_t tx_model;           /* USTBX_TXMODEL_* for sources with z >= 0 */
  int32_t _pad;
  const float* tx_delay_matrix; /* optional [W][P] DEVICE pointer (unified model) */
  float fs, t0, fd;
  int32_t dim, interp, edge;
  uint32_t flags;
  int32_t probe_kind;         /* USTBX_PROBE_*: which incidence geometry apodization.m uses */
  int32_t _pad2;
  double probe_radius;        /* USTBX_PROBE_CURVILINEAR: radius (uff.curvilinear_array) or
                                 radius_x (uff.curvilinear_matrix_array) [m] */
} ustbx_geometry;

typedef struct ustbx_plan ustbx_plan;

/* Uploads everything the plan needs into device memory on the current device. */
int ustbx_plan_create(const ustbx_geometry* g, ustbx_plan** plan,
                      ustbx_status* st);
/* ch: device [F][W][N][L] complex64; out: device [F][W'][N'][P] complex64. */
int ustbx_plan_run(const ustbx_plan* plan, const void* ch, void* out,
                   void* stream, ustbx_status* st);
/* Tile shape override (0 = keep). Variant 0: waves per register chunk (1, 8,
 * 16) and channel groups per pixel tile (1 or 4; dimension both/transmit).
 * Variants 1-3: waves per wave-split (variant 1: 8, 10, 12, 16; variants 2
 * and 3: 4, 5, 6, 8, 10, 12, 16) and wave-splits per 64-pixel tile (variant
 * 1: 4 or 8; variants 2 and 3: 4, 5 or 8). */
int ustbx_plan_set_tiles(ustbx_plan* plan, int32_t wave_chunk, int32_t split,
                         ustbx_status* st);
/* Kernel variant (results agree to FP32 rounding; no reference counterpart):
 *   0 = per-lane wave chunks (every mode);
 *   1 = receive state shared in LDS by the wave-splits of a 64-pixel tile
 *       (dimension both, IQ, separable phase, one frame);
 *   2 = resident receive table and per-tile wave/channel lists (dimension
 *       both with the separable or RF phase outside das_c's order, or
 *       dimension transmit with any phase; any frame count, receive table
 *       within half the device memory; the default where it applies). The
 *       64-pixel tiles that hold a non-finite nonzero weight or delay (such
 *       as the NaN apodization of pixels at depth 0) are computed with a
 *       per-term window test, in a second launch. For dimension transmit set_tiles'
 *       arguments are frames per block (4 or 8) and channels per block (4 or
 *       8; one wave64 each), and without the separable phase (or with RF
 *       data) it keeps das_c's term arithmetic and summation order;
 *   3 = variant 2 (dimension both) over cubic-interpolation coefficients: the plan holds a
 *       buffer of 32 bytes per channel sample that every run rewrites, so
 *       runs of one variant-3 plan on different streams must not overlap;
 *   4 = variant 2 without the receive table (dimension both, a windowed or
 *       no receive apodization without tilt, flat probes): the blocks compute
 *       their tile's receive delays (das.m's, to the FP32 rounding) and
 *       weights (FP32) in place, 16 channels at a time through LDS. The plan
 *       keeps the transmit table and the tile lists only (the 128 x 128 x 256
 *       volume of a 32 x 32 array: 1.8 GiB against 65.8 GiB); it is the
 *       default where the receive table would exceed half the device memory.
 * USTBX_E_UNSUPPORTED where the variant does not apply. */
int ustbx_plan_set_variant(ustbx_plan* plan, int32_t variant, ustbx_status* st);
/* Tile visiting order of variants 2 and 3 (no reference counterpart):
 * 0 = scan order; 1 = depth band by depth band within each XCD's share of the
 * scan lines (the default; it keeps the sample windows that co-resident
 * blocks read in L2). Results do not depend on it. */
int ustbx_plan_set_order(ustbx_plan* plan, int32_t order, ustbx_status* st);
/* LDS staging of sample windows (no reference counterpart); results do not
 * depend on it. Dimension both on the receive table (variants 2, 3 only for
 * linear or cubic data): 1 (the default) runs every (tile, channel) the plan
 * marked (all 64 pixels' terms in one 64-sample window of each wave's trace,
 * every weight nonzero) from windows loaded with one contiguous load per
 * term and read back from LDS, 0 gathers each term's samples. Dimension
 * transmit on variant 2: 1 runs every (tile, channel) the plan marked (each
 * wave's terms on the tile within one 64-sample window per frame) from LDS,
 * 0 (the default) gathers. Batches and batchers gather every term whatever
 * the setting (their persistent kernel holds fewer registers so). */
int ustbx_plan_set_staging(ustbx_plan* plan, int32_t on, ustbx_status* st);
/* (tile, channel) pairs of a receive-table plan (dimension both or
 * transmit) that the plan marked for staging, and all pairs of the tiles'
 * channel lists (0, 0 otherwise). */
int ustbx_plan_stage_stats(const ustbx_plan* plan, int64_t* staged, int64_t* listed, ustbx_status* st);
/* Current variant (or -1 for NULL) and its tile shape. */
int32_t ustbx_plan_variant(const ustbx_plan* plan, int32_t* wave_chunk, int32_t* split);

/* ---------------------------------------------------------------------------
 * 3. Batches: many acquisitions in one persistent launch (no reference
 *    counterpart; das.m beamforms one acquisition per call).
 * ------------------------------------------------------------------------ */
typedef struct ustbx_batch ustbx_batch;

/* count acquisitions, acquisition i beamformed with plans[i] (a plan may be
 * listed more than once; every plan must have variant-2 state -- see
 * ustbx_plan_set_variant -- and all share one interpolation). Dimension both;
 * each image equals ustbx_plan_run of its plan with the same tile shape, bit
 * for bit. The plans must outlive the batch. */
int ustbx_batch_create(const ustbx_plan* const* plans, int32_t count, ustbx_batch** batch,
                       ustbx_status* st);
/* Device buffers of every acquisition: ch[i] [F][W][N][L] complex64, out[i]
 * [F][P] complex64. Synchronous; a captured graph then uses the new buffers. */
int ustbx_batch_set_io(ustbx_batch* batch, const void* const* ch, void* const* out,
                       ustbx_status* st);
/* Tile shape of the persistent kernel (0 = keep): waves per split (4, 5, 6,
 * 8, 10, 12, 16) and wave-splits per 64-pixel tile (4, 5 or 8). The default
 * follows the wave count below which half of the batch's work lies: up to 32
 * waves 4 splits of 8, else 8 splits covering it in one round. Drops a
 * captured graph. */
int ustbx_batch_set_tiles(ustbx_batch* batch, int32_t wave_chunk, int32_t split, ustbx_status* st);
/* Enqueues the batch on stream: a counter-reset kernel and the persistent
 * kernel; allocates nothing and can be captured in a hipGraph. Launches of
 * one batch must not overlap (they share its claim counters). */
int ustbx_batch_launch(const ustbx_batch* batch, void* stream, ustbx_status* st);
/* The same work through a hipGraph captured on the first call and launched on
 * stream. */
int ustbx_batch_replay(ustbx_batch* batch, void* stream, ustbx_status* st);
/* Acquisition count (or -1 for NULL); tile shape, persistent blocks and work
 * items (tiles x frames). */
int32_t ustbx_batch_shape(const ustbx_batch* batch, int32_t* wave_chunk, int32_t* split, int32_t* blocks,
                          int64_t* items);
void ustbx_batch_destroy(ustbx_batch* batch);

/* Batcher: a different set of acquisitions in every launch (serving loops
 * that batch the frames that are ready). An acquisition's tiles are cut into
 * 8 shares (one per XCD) of ustbx_plan_positions(plan) positions each; a
 * submission computes positions [pos_begin[i], pos_end[i]) of every share of
 * acquisition i, so a large frame can be spread over several submissions and
 * is complete once all positions have run. Tiles are computed whole, so the
 * images equal ustbx_plan_run with the same tile shape bit for bit. */
typedef struct ustbx_batcher ustbx_batcher;

/* Positions per share of a plan with variant-2 state (-1 otherwise). */
int32_t ustbx_plan_positions(const ustbx_plan* plan);
/* interp: the plans' interpolation; capacities: acquisitions and work items
 * (tiles x frames) per submission; tile shape as ustbx_batch_set_tiles. */
int ustbx_batcher_create(int32_t interp, int32_t max_acqs, int64_t max_items, int32_t wave_chunk, int32_t split,
                         ustbx_batcher** batcher, ustbx_status* st);
/* Uploads the submission's descriptors and work items (asynchronously, from
 * pinned staging memory; it first waits until the previous submission's
 * upload has left the staging memory) and launches the persistent kernel on
 * stream. Allocates nothing. Submissions run in submission order: one made on
 * another stream than the previous one waits (hipStreamWaitEvent) for the
 * previous submission's kernel, whose descriptor buffers it rewrites. */
int ustbx_batcher_submit(ustbx_batcher* batcher, int32_t count, const ustbx_plan* const* plans,
                         const void* const* ch, void* const* out, const int32_t* pos_begin,
                         const int32_t* pos_end, void* stream, ustbx_status* st);
void ustbx_batcher_destroy(ustbx_batcher* batcher);
/* Writes the plan's FP32 delay/apodization matrices (what das.m hands to the
 * MEX) into device buffers: tx [W][P], rx [N][P]; any may be NULL. */
int ustbx_plan_materialize(const ustbx_plan* plan, float* tx_apo, float* rx_apo,
                           float* tx_delay, float* rx_delay, void* stream,
                           ustbx_status* st);
/* Number of (pixel, channel, wave, frame) terms with nonzero transmit and
 * receive weight: the terms das_c evaluates (das_c.cpp:366,371). Synchronous. */
int ustbx_plan_active_terms(const ustbx_plan* plan, int64_t* terms, ustbx_status* st);
int64_t ustbx_plan_pixels(const ustbx_plan* plan);
int64_t ustbx_plan_device_bytes(const ustbx_plan* plan);
void ustbx_plan_destroy(ustbx_plan* plan);

/* Analytic signal of real RF traces along time (FFT method), device buffers:
 * rf [traces][L] float32 -> analytic [traces][L] complex64. Replaces the RF
 * branch of +midprocess/das.m:211-214 (MATLAB hilbert) and
 * python/src/ustb/midprocess/das.py:144-145 (scipy.signal.hilbert). */
int ustbx_hilbert(const float* rf, void* analytic, int32_t L, int64_t traces,
                  void* stream, ustbx_status* st);

/* ---------------------------------------------------------------------------
 * 4. Library information.
 * ------------------------------------------------------------------------- */
int32_t ustbx_abi_version(void);
/* Number of HIP devices visible (0 when none); never fails. */
int32_t ustbx_device_count(void);

/* ---------------------------------------------------------------------------
 * 5. Postprocess on beamformed data: device buffers [F][W][N][P] complex64
 *    (MATLAB [P, N, W, F]); dim names the summed dimension (USTBX_DIM_RX,
 *    USTBX_DIM_TX or USTBX_DIM_BOTH).
 * ------------------------------------------------------------------------- */
#define USTBX_CF_MATLAB 0 /* postprocess.coherence_factor (coherence_factor.m:97-143) */
#define USTBX_CF_PYTHON 1 /* python/src/ustb/postprocess/coherence_factor.py:19-50 */

/* postprocess.coherent_compounding (coherent_compounding.m:41-73). Without a
 * window (window_size 0) out is [F][W'][N'][P], the summed axes of size 1.
 * With one (dimension both or transmit) waves and frames are merged, wave
 * fastest, and out[t] sums entries t .. t+window_size-1 (movsum with
 * 'Endpoints' 'discard'): out is [T][1][N'][P], T = W*F - window_size + 1;
 * receive ignores the window. window_size > W is USTBX_E_ARG, the reference's
 * assertion. */
int ustbx_coherent_compounding(const void* in, void* out, int32_t P, int32_t N, int32_t W, int32_t F,
                               int32_t dim, int32_t window_size, void* stream, ustbx_status* st);
/* Active elements per pixel for the coherence factor (coherence_factor.m:
 * 82-131): M[p] = number of summed elements whose apodization exceeds
 * criterium (the reference's active_element_criterium, 0.16 by default); both
 * counts the products tx_apo * rx_apo. tx_apo [W][P] and rx_apo [N][P] are
 * device FP32 (NULL = ones), compared in FP64; M is device [P]. */
int ustbx_active_elements(const float* tx_apo, const float* rx_apo, int32_t P, int32_t N, int32_t W,
                          int32_t dim, double criterium, float* M, void* stream, ustbx_status* st);
/* The Mallart-Fink coherence factor: image = CF * coherent sum, cf = CF
 * (NULL: not written), both [F][W'][N'][P] with the summed axes of size 1.
 * USTBX_CF_MATLAB: CF = |S|^2 / sum|s|^2 * (1/M), M [P] active elements
 * (NULL: every summed element); USTBX_CF_PYTHON: |S|^2 / (sum|s|^2 * M +
 * 1e-30) with M the number of summed elements; NaN -> 0 in both. The
 * reference's dimension fallbacks (coherence_factor.m:62-79) are the
 * caller's (ustb_amd.postprocess.CoherenceFactor applies them). */
int ustbx_coherence_factor(const void* in, void* image, float* cf, const float* M, int32_t P, int32_t N,
                           int32_t W, int32_t F, int32_t dim, int32_t flavour, void* stream, ustbx_status* st);
/* das in dimension transmit with the coherence factor over the receive
 * channels in the kernel's epilogue (no reference counterpart: das.m writes
 * the per-channel images and postprocess.coherence_factor reads them back).
 * Each block of the transmit kernel reduces its channels' images, in channel
 * order, to per-pixel sums S and sum |s|^2 in `workspace` (ceil(N / channels
 * per block) x F x P float4, ustbx_plan_cf_workspace_bytes), a second pass
 * adds the groups in order and applies the factor: image [F][P] complex64,
 * cf [F][P] float (NULL: not written); M and flavour as
 * ustbx_coherence_factor. Equal to ustbx_plan_run followed by
 * ustbx_coherence_factor(dim receive) up to the FP32 rounding of the grouped
 * sums. Needs a dimension-transmit plan on variant 2 and at least two
 * channels. Calls sharing a workspace must not overlap. */
int64_t ustbx_plan_cf_workspace_bytes(const ustbx_plan* plan);
int ustbx_plan_run_cf(const ustbx_plan* plan, const void* ch, void* image, float* cf, const float* M,
                      int32_t flavour, void* workspace, int64_t workspace_bytes, void* stream, ustbx_status* st);
/* ustbx_plan_run_cf without a workspace argument and without allocating: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under hipGraph capture, serving loops, harnesses that forbid
 * hipMalloc). The partial sums go to a device array of the library's own,
 * ustbx_scratch_bytes() (64 MiB) per GPU, held from library load to process
 * end; a plan whose ustbx_plan_cf_workspace_bytes exceeds it is
 * USTBX_E_CAPACITY. The results equal ustbx_plan_run_cf's bit for bit. Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (on
 * different streams, or from different threads) race. */
int64_t ustbx_scratch_bytes(void);
int ustbx_plan_run_cf_noalloc(const ustbx_plan* plan, const void* ch, void* image, float* cf, const float* M,
                              int32_t flavour, void* stream, ustbx_status* st);
/* M[p] from a plan's FP32 apodization (as ustbx_plan_materialize gives it),
 * for ustbx_plan_run_cf and ustbx_coherence_factor. */
int ustbx_plan_active_elements(const ustbx_plan* plan, int32_t dim, double criterium, float* M, void* stream,
                               ustbx_status* st);

/* ---------------------------------------------------------------------------
 * 6. RF -> IQ front end (preprocess.fast_demodulation).
 * ------------------------------------------------------------------------- */
#define USTBX_FIR_CAUSAL 0     /* MATLAB filter(b, 1, x), zero initial state (fast_demodulation.m:149) */
#define USTBX_FIR_ZERO_PHASE 1 /* scipy filtfilt(b, 1, x) with its defaults (fast_demodulation.py:66) */

/* Down-mix, low-pass and decimate RF traces: rf [traces][L] float32 (device;
 * the channel data [F][W][N][L] is traces = F*W*N), iq [traces][ceil(L/down)]
 * complex64 (device). iq[j] is sample j*down of the filtered signal
 * rf(n) * exp(-1j 2 pi fmod (t0 + n/fs)) (fast_demodulation.m:136, 149, 171).
 * b: the FIR taps (host, FP64; the caller designs them: fast_demodulation.m
 * :267-275 or fast_demodulation.py:61-64), at most 767 causal / 384
 * zero-phase; zero-phase needs L > 3 * ntaps, as filtfilt does. The caller
 * updates the time origin (MATLAB: t0 - (ntaps-1)/2/fs) and the sampling
 * frequency (fs / down). No allocation; asynchronous on stream. */
int ustbx_demodulate(const float* rf, void* iq, int32_t L, int64_t traces, double fs, double t0, double fmod,
                     const double* b, int32_t ntaps, int32_t down, int32_t mode, void* stream, ustbx_status* st);

/* The modulation frequency preprocess.fast_demodulation estimates when none
 * is set (fast_demodulation.m:57-73): rf [traces][L] float32 (device; the
 * first min(100, F) frames of the channel data, F*W*N traces in the MEX
 * layout), fs [Hz]. The power spectrum averaged over the traces
 * (tools.power_spectrum, power_spectrum.m:1-22; fftshifted, normalised to its
 * maximum) is written to pw (host [L], NULL: not written); fmod (host, NULL:
 * not written) receives -fc, fc the centre of the -6 dB points around the
 * spectrum's first maximum (MATLAB interp1 over pw(1:ic) and
 * pw(ic:round(L/2))). Synchronous on stream; allocates its FFT workspace. */
int ustbx_estimate_modulation_frequency(const float* rf, int32_t L, int64_t traces, double fs, double* fmod,
                                        double* pw, void* stream, ustbx_status* st);

#ifdef __cplusplus
}
#endif
#endif /* USTBX_DAS_H */
