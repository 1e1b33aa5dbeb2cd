// das_capi.hip — the C-ABI (include/ustbx_das.h): argument validation with the
// reference's error contract, geometry plans (device-resident probe / scan /
// sequence state), host staging for MEX-style callers.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/ustbx_das.h"
#include "das_launch.h"
#include "das_plan.h"
#include "das_status.h"
#include "host_copy.h"

using namespace ustbx;

namespace {

// das_c: wd = 2*(float)PI*fd if |fd| > 1e-6 else 0 (das_c.cpp:32-33,240-244)
float dasc_wd(float fd) {
  volatile float two_pi = 2 * (float)3.14159265359;
  return (fabsf(fd) > 1e-6) ? two_pi * fd : 0.0f;
}

int auto_split(int P) { return P >= (1 << 18) ? 1 : 4; }

// Channel groups of the matrix kernel: 8, the fastest of 1, 2, 4, 8 on C1
// (65 K pixels), C2 (512 K pixels, dimension both: 8.74 ms against 9.98 ms
// with 1 group) and C3 (262 K pixels, transmit: 21.5 ms against 30.2 ms with
// 2); the groups' shorter channel loops keep more blocks' sample windows in
// L2 at once. Dimension transmit and none sum nothing across groups, so
// there the choice does not change the result.
int matrix_auto_split(int /*P*/) { return 8; }

// The shared-receive-state kernel covers dimension both with the separable
// phase on one frame.
bool shared_rx_applies(const GeomLaunch& G) {
  return G.dim == 3 && G.separable && G.F == 1 && G.wd != 0.0f && (int64_t)G.L * G.N * G.W * 8 <= 0x7FFFFE00;
}

// Default tiles of the shared-receive kernel: 8 wave-splits, the smallest
// instantiated waves-per-split covering W in one round (else 8 per round).
void shared_rx_default_tiles(int W, int& wc, int& ws) {
  ws = 8;
  const int need = (W + 7) / 8;
  wc = need <= 8 ? 8 : need <= 10 ? 10 : need <= 12 ? 12 : need <= 16 ? 16 : 8;
}

bool shared_rx_wc_ok(int wc) { return wc == 8 || wc == 10 || wc == 12 || wc == 16; }

// The receive-table kernel covers dimension both in the separable (or RF)
// phase, outside das_c's summation order.
// The receive table may take at most this fraction of the device memory.
constexpr double kRxTableMaxFraction = 0.5;

bool rxtab_wc_ok(int wc) { return wc == 4 || wc == 5 || wc == 6 || wc == 8 || wc == 10 || wc == 12 || wc == 16; }

// Up to 8 waves: 8 splits of 4 waves, most of them in channel groups; up to
// 32 waves: 4 splits of 8 waves (a whole round with enough terms per receive
// row); more: 8 splits with the smallest waves-per-split covering the busiest
// tile in one round (else 16 per round).
void rxtab_default_tiles(int max_nw, int& wc, int& ws);

// Dimension transmit: 8 frames per block (4 when fewer), 8 channels per block
// (4 when fewer).
void rxtab_defaults(const GeomLaunch& G, int max_nw, int& wc, int& ws) {
  if (G.dim == 2) {  // frames per block, channels per block (measured best on 32 frames x 64 channels)
    wc = G.F >= 8 ? 8 : 4;
    ws = G.N >= 8 ? 8 : 4;
    return;
  }
  rxtab_default_tiles(max_nw, wc, ws);
  // Keys' cubic over many waves: 4 splits of 10 waves (C2, 75 waves:
  // 2.49 ms against 2.57 ms with 8 splits)
  if (G.interp == USTBX_INTERP_CUBIC && max_nw > 32) {
    ws = 4;
    wc = 10;
  }
}

void rxtab_default_tiles(int max_nw, int& wc, int& ws) {
  if (max_nw <= 8) {  // 1-2 wave-splits; the others take shares of the channels
    ws = 8;
    wc = 4;
    return;
  }
  if (max_nw <= 16) {
    ws = 4;
    wc = 8;
    return;
  }
  if (max_nw <= 32) {
    // two channel groups of two 16-wave splits (the 3-D class, 25 waves:
    // 52.3 ms against 55.9 ms at 8 x 4)
    ws = 4;
    wc = 16;
    return;
  }
  ws = 8;
  const int need = (max_nw + 7) / 8;
  wc = need <= 4 ? 4 : need <= 6 ? 6 : need <= 8 ? 8 : need <= 10 ? 10 : need <= 12 ? 12 : 16;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host-buffer entry (what a MEX gateway holds): state kept per device across
// calls, so that a call does no device allocation and no device-wide
// synchronisation: a device buffer grown to the largest call so far, two
// pinned staging buffers, a stream of its own (default priority) and an
// event per staging buffer. Inputs go host memory -> staging (the CPU copy,
// split over a few threads) -> device (DMA on the stream), chunk by chunk
// with the two staging buffers alternating, so that the CPU copy of one
// chunk overlaps the DMA of the previous one; the output comes back the same
// way. Calls on one device are serialised.
// ---------------------------------------------------------------------------
namespace {

constexpr int kHostDevices = 64;
constexpr size_t kStageChunk = (size_t)8 << 20;  // bytes per staging buffer

CopyPool& copy_pool() {
  static CopyPool pool(4);  // CPU copy threads, the caller's included
  return pool;
}

struct HostCopy {
  void* dst;        // device
  const void* src;  // host
  size_t bytes;
};

struct HostEntry {
  std::mutex mu;
  hipStream_t stream = nullptr;
  char* dev = nullptr;
  size_t dev_bytes = 0;
  char* stage[2] = {nullptr, nullptr};
  hipEvent_t free_ev[2] = {nullptr, nullptr};  // the staging buffer's last DMA has finished
  bool used[2] = {false, false};

  hipError_t reserve(size_t bytes) {
    hipError_t e = hipSuccess;
    if (!stream && (e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)) != hipSuccess) return e;
    for (int k = 0; k < 2; ++k) {
      if (!stage[k] && (e = hipHostMalloc((void**)&stage[k], kStageChunk, hipHostMallocDefault)) != hipSuccess)
        return e;
      if (!free_ev[k] && (e = hipEventCreateWithFlags(&free_ev[k], hipEventDisableTiming)) != hipSuccess) return e;
    }
    if (bytes > dev_bytes) {
      if (dev) {
        (void)hipStreamSynchronize(stream);
        (void)hipFree(dev);
        dev = nullptr;
        dev_bytes = 0;
      }
      if ((e = hipMalloc(&dev, bytes)) != hipSuccess) return e;
      dev_bytes = bytes;
    }
    return hipSuccess;
  }

  // waits until staging buffer k is free again
  hipError_t acquire(int k) {
    if (!used[k]) return hipSuccess;
    used[k] = false;
    return hipEventSynchronize(free_ev[k]);
  }

  hipError_t upload(const HostCopy* c, int n) {
    int k = 0;
    for (int i = 0; i < n; ++i) {
      for (size_t off = 0; off < c[i].bytes; off += kStageChunk) {
        const size_t len = std::min(kStageChunk, c[i].bytes - off);
        hipError_t e = acquire(k);
        if (e != hipSuccess) return e;
        copy_pool().copy(stage[k], (const char*)c[i].src + off, len);
        e = hipMemcpyAsync((char*)c[i].dst + off, stage[k], len, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipEventRecord(free_ev[k], stream);
        if (e != hipSuccess) return e;
        used[k] = true;
        k ^= 1;
      }
    }
    return hipSuccess;
  }

  // device -> staging (DMA) of chunk c + 1 overlaps staging -> host (CPU) of chunk c
  hipError_t download(void* dst, const void* src, size_t bytes) {
    const size_t chunks = (bytes + kStageChunk - 1) / kStageChunk;
    hipError_t e = hipSuccess;
    for (size_t c = 0; c < chunks && c < 2; ++c) {
      const size_t off = c * kStageChunk, len = std::min(kStageChunk, bytes - off);
      if ((e = acquire((int)c)) != hipSuccess) return e;
      e = hipMemcpyAsync(stage[c], (const char*)src + off, len, hipMemcpyDeviceToHost, stream);
      if (e == hipSuccess) e = hipEventRecord(free_ev[c], stream);
      if (e != hipSuccess) return e;
      used[c] = true;
    }
    for (size_t c = 0; c < chunks; ++c) {
      const int k = (int)(c & 1);
      const size_t off = c * kStageChunk, len = std::min(kStageChunk, bytes - off);
      if ((e = acquire(k)) != hipSuccess) return e;
      copy_pool().copy((char*)dst + off, stage[k], len);
      const size_t n2 = c + 2;
      if (n2 < chunks) {
        const size_t off2 = n2 * kStageChunk, len2 = std::min(kStageChunk, bytes - off2);
        e = hipMemcpyAsync(stage[k], (const char*)src + off2, len2, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipEventRecord(free_ev[k], stream);
        if (e != hipSuccess) return e;
        used[k] = true;
      }
    }
    return hipStreamSynchronize(stream);
  }
};

HostEntry g_host[kHostDevices];

}  // namespace

extern "C" {

int32_t ustbx_abi_version(void) { return USTBX_ABI_VERSION; }

int32_t ustbx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

// The identifier das_c raises each class of error with: nrhs, nlhs and
// Dimensions through mexErrMsgIdAndTxt (das_c.cpp:53-55, 88-199); the
// single-precision, scalar, sum-mode and real-data checks through
// mexErrMsgTxt, with no identifier (das_c.cpp:79, 98, 145-242): "". GPU
// errors take das_cuda's identifier (das_cuda.cu:122). Capacity and
// unsupported-feature errors have no reference counterpart and an identifier
// of this library's own.
const char* ustbx_status_id(int32_t code) {
  switch (code) {
    case USTBX_OK: return "";
    case USTBX_E_NRHS: return "Toolbox:SRP_SRC:nrhs";
    case USTBX_E_NLHS: return "Toolbox:SRP_SRC:nlhs";
    case USTBX_E_DIMENSIONS: return "Toolbox:SRP_SRC:Dimensions";
    case USTBX_E_GPU: return "Toolbox:SRP_SRC:GPU";
    case USTBX_E_CAPACITY: return "ustbx:capacity";
    case USTBX_E_UNSUPPORTED: return "ustbx:unsupported";
    default: return "";  // USTBX_E_FLOAT, USTBX_E_SCALAR, USTBX_E_ARG: mexErrMsgTxt in das_c
  }
}

// The das_c argument checks, in das_c's order and with its messages
// (das_c.cpp:53-244). mat_* index: 0 tx_apo, 1 rx_apo, 2 tx_delay, 3 rx_delay
// (prhs[3..6]); scalar_*: 0 fs, 1 t0, 2 fd (prhs[1], prhs[2], prhs[7]).
int ustbx_das_check_mex_args(int32_t nrhs, int32_t nlhs, int32_t ch_ndim,
                             const int64_t* ch_dims, int32_t ch_is_double,
                             const int32_t* mat_ndim, const int64_t* mat_dims,
                             const int32_t* mat_is_double,
                             const int64_t* scalar_numel,
                             const int32_t* scalar_is_double, int32_t sum_mode,
                             int32_t ch_is_complex, float fd, ustbx_das_desc* d,
                             ustbx_status* st) {
  if (nrhs < 9) return set_status(st, USTBX_E_NRHS, "Too few input arguments");
  if (nrhs > 10) return set_status(st, USTBX_E_NRHS, "Too many input arguments");
  if (nlhs > 1) return set_status(st, USTBX_E_NLHS, "Too many output arguments");
  if (!ch_dims || !mat_ndim || !mat_dims || !mat_is_double || !scalar_numel ||
      !scalar_is_double || !d)
    return set_status(st, USTBX_E_ARG, "NULL argument");
  if (sum_mode < 0 || sum_mode > 3)
    return set_status(st, USTBX_E_ARG,
                      "Unknown sum mode. Available: 0 -> NONE, 1->RX, 2->TX, 3->BOTH");
  if (ch_ndim < 2 || ch_ndim > 4)
    return set_status(st, USTBX_E_DIMENSIONS,
                      "Unknown channel data format. Expected from 2 to 4 dimensions: "
                      "[time, channel, wave, frame]");
  const int64_t L = ch_dims[0], N = ch_dims[1];
  const int64_t W = ch_ndim > 2 ? ch_dims[2] : 1;
  const int64_t F = ch_ndim > 3 ? ch_dims[3] : 1;
  if (ch_is_double) return set_status(st, USTBX_E_FLOAT, "The channel data should be single precision");
  // tx delay (prhs[5])
  if (mat_ndim[2] > 2)
    return set_status(st, USTBX_E_DIMENSIONS,
                      "Unknown transmit delay format. Expected 2 dimensions: [pixel, wave]");
  const int64_t P = mat_dims[2 * 2 + 0];
  int64_t Wc = mat_ndim[2] == 2 ? mat_dims[2 * 2 + 1] : 1;
  if (W != Wc)
    return set_status(st, USTBX_E_DIMENSIONS,
                      "Number of waves in channel data & transmit delays do not match.");
  if (mat_is_double[2]) return set_status(st, USTBX_E_FLOAT, "The transmiy delays should be single precision");
  // rx delay (prhs[6])
  if (mat_ndim[3] > 2)
    return set_status(st, USTBX_E_DIMENSIONS,
                      "Unknown receive delay format. Expected 2 dimensions: [pixel, channel]");
  if (mat_dims[3 * 2 + 0] != P)
    return set_status(st, USTBX_E_DIMENSIONS,
                      "Number of pixels in transmit and receive delays do not match.");
  int64_t Nc = mat_ndim[3] == 2 ? mat_dims[3 * 2 + 1] : 1;
  if (N != Nc)
    return set_status(st, USTBX_E_DIMENSIONS,
                      "Number of channels in channel data & receive delays do not match.");
  if (mat_is_double[3]) return set_status(st, USTBX_E_FLOAT, "The transmiy delays should be single precision");
  // tx apodization (prhs[3])
  if (mat_ndim[0] > 2)
    return set_status(st, USTBX_E_DIMENSIONS,
                      "Unknown transmit apodization format. Expected 2 dimensions: [pixel, wave]");
  if (mat_dims[0] != P)
    return set_status(st, USTBX_E_DIMENSIONS,
                      "Number of pixels in transmit apodization and transmit delays do not match.");
  Wc = mat_ndim[0] == 2 ? mat_dims[1] : 1;
  if (W != Wc)
    return set_status(st, USTBX_E_DIMENSIONS,
                      "Number of waves in channel data & transmit apodization do not match.");
  if (mat_is_double[0]) return set_status(st, USTBX_E_FLOAT, "The transmiy apodization should be single precision");
  // rx apodization (prhs[4])
  if (mat_ndim[1] > 2)
    return set_status(st, USTBX_E_DIMENSIONS,
                      "Unknown receive apodization format. Expected 2 dimensions: [pixel, channel]");
  if (mat_dims[1 * 2 + 0] != P)
    return set_status(st, USTBX_E_DIMENSIONS,
                      "Number of pixels in receive apodization and receive delays do not match.");
  Nc = mat_ndim[1] == 2 ? mat_dims[1 * 2 + 1] : 1;
  if (N != Nc)
    return set_status(st, USTBX_E_DIMENSIONS,
                      "Number of channels in channel data & receive apodization do not match.");
  if (mat_is_double[1]) return set_status(st, USTBX_E_FLOAT, "The transmiy delays should be single precision");
  // scalars
  if (scalar_numel[0] != 1) return set_status(st, USTBX_E_SCALAR, "The sampling frequency should be an escalar");
  if (scalar_is_double[0]) return set_status(st, USTBX_E_FLOAT, "The sampling frequency should be single precision");
  if (scalar_numel[1] != 1) return set_status(st, USTBX_E_SCALAR, "The initial time should be an escalar");
  if (scalar_is_double[1]) return set_status(st, USTBX_E_FLOAT, "The initial time should be single precision");
  if (scalar_numel[2] != 1) return set_status(st, USTBX_E_SCALAR, "Modulation frequency 'fd' should be an escalar");
  if (scalar_is_double[2]) return set_status(st, USTBX_E_FLOAT, "The modulation frequency should be single precision");
  if (fabsf(fd) > 1e-6 && !ch_is_complex)
    return set_status(st, USTBX_E_ARG,
                      "The modulation frequency > 0 but the input data is real. Check inputs.");
  if (L > INT32_MAX || N > INT32_MAX || W > INT32_MAX || F > INT32_MAX || P > INT32_MAX ||
      L < 0 || N < 0 || W < 0 || F < 0 || P < 0)
    return set_status(st, USTBX_E_CAPACITY, "array dimension exceeds 2^31-1");
  d->L = (int32_t)L;
  d->N = (int32_t)N;
  d->W = (int32_t)W;
  d->F = (int32_t)F;
  d->P = (int32_t)P;
  d->dim = sum_mode;
  return ok(st);
}

static int check_desc(const ustbx_das_desc* d, ustbx_status* st) {
  if (!d) return set_status(st, USTBX_E_ARG, "NULL descriptor");
  if (d->dim < 0 || d->dim > 3)
    return set_status(st, USTBX_E_ARG, "Unknown sum mode. Available: 0 -> NONE, 1->RX, 2->TX, 3->BOTH");
  if (d->interp != USTBX_INTERP_LINEAR && d->interp != USTBX_INTERP_CUBIC)
    return set_status(st, USTBX_E_ARG, "unknown interpolation %d", d->interp);
  if (d->edge != USTBX_EDGE_DASC && d->edge != USTBX_EDGE_PYTHON)
    return set_status(st, USTBX_E_ARG, "unknown edge rule %d", d->edge);
  if (d->L < 0 || d->N < 0 || d->W < 0 || d->F < 0 || d->P < 0)
    return set_status(st, USTBX_E_DIMENSIONS, "negative dimension");
  if (d->split < 0) return set_status(st, USTBX_E_ARG, "negative split");
  if ((d->flags & USTBX_FLAG_PYTHON_ARITHMETIC) &&
      (d->edge != USTBX_EDGE_PYTHON || d->interp != USTBX_INTERP_LINEAR || !(d->fs > 0.0f)))
    return set_status(st, USTBX_E_ARG,
                      "the Python arithmetic needs the Python edge rule, linear interpolation and a positive sample "
                      "period");
  return USTBX_OK;
}

static bool empty_problem(const ustbx_das_desc* d) {
  return d->P == 0 || d->F == 0;
}

static size_t out_elems(int dim, size_t P, size_t N, size_t W, size_t F) {
  switch (dim) {
    case 0: return P * N * W * F;
    case 1: return P * W * F;
    case 2: return P * N * F;
    default: return P * F;
  }
}

int ustbx_das_run(const ustbx_das_desc* d, const void* ch, const float* tx_apo,
                  const float* rx_apo, const float* tx_delay, const float* rx_delay,
                  void* out, void* stream, ustbx_status* st) {
  int rc = check_desc(d, st);
  if (rc) return rc;
  if (empty_problem(d)) return ok(st);
  if (!out) return set_status(st, USTBX_E_ARG, "NULL output");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const size_t ne = out_elems(d->dim, d->P, d->N, d->W, d->F);
  if (d->N == 0 || d->W == 0 || d->L < 2) {
    // nothing can be interpolated: das_c returns its zero-initialised array
    hipError_t e = launch_zero_fill(out, ne, s);
    if (e != hipSuccess) return hip_fail(st, e, "zero_fill_kernel launch");
    return ok(st);
  }
  if (!ch || !tx_apo || !rx_apo || !tx_delay || !rx_delay)
    return set_status(st, USTBX_E_ARG, "NULL input array");
  MatLaunch m;
  m.ch = ch;
  m.tx_apo = tx_apo;
  m.rx_apo = rx_apo;
  m.tx_delay = tx_delay;
  m.rx_delay = rx_delay;
  m.out = out;
  m.L = d->L; m.N = d->N; m.W = d->W; m.F = d->F; m.P = d->P;
  m.fs = d->fs;
  m.t0 = d->t0;
  m.wd = dasc_wd(d->fd);
  m.dim = d->dim;
  m.interp = d->interp;
  m.edge = d->edge;
  m.py = (d->flags & USTBX_FLAG_PYTHON_ARITHMETIC) != 0;
  m.dt = m.py ? d->fs : 0.0f;
  m.w0 = m.py ? d->fd : 0.0f;
  if (m.py) m.wd = fabsf(d->fd) > 1.1920929e-7f ? d->fd : 0.0f;  // the phase where |w0| > FP32 eps (das.py:249)
  m.split = (d->flags & USTBX_FLAG_DASC_ORDER) ? 1 : (d->split ? d->split : matrix_auto_split(d->P));
  if (m.split != 1 && m.split != 2 && m.split != 4 && m.split != 8)
    return set_status(st, USTBX_E_ARG, "split must be 1, 2, 4 or 8");
  hipError_t e = launch_das_matrix(m, s);
  if (e != hipSuccess) return hip_fail(st, e, "das_matrix_kernel launch");
  return ok(st);
}

int ustbx_das_run_host(const ustbx_das_desc* d, const void* ch, const float* tx_apo,
                       const float* rx_apo, const float* tx_delay, const float* rx_delay,
                       void* out, int32_t device, ustbx_status* st) {
  int rc = check_desc(d, st);
  if (rc) return rc;
  if (!out) return set_status(st, USTBX_E_ARG, "NULL output");
  const size_t P = d->P, N = d->N, W = d->W, F = d->F, L = d->L;
  const size_t ne = out_elems(d->dim, P, N, W, F);
  if (empty_problem(d)) return ok(st);
  const size_t bytes_ch = L * N * W * F * 8, bytes_tx = P * W * 4, bytes_rx = P * N * 4;
  const size_t bytes_out = ne * 8;
  if ((bytes_ch && !ch) || (bytes_tx && (!tx_apo || !tx_delay)) || (bytes_rx && (!rx_apo || !rx_delay)))
    return set_status(st, USTBX_E_ARG, "NULL input array");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) return set_status(st, USTBX_E_GPU, "no HIP device available");
  if (device < 0 || device >= ndev || device >= kHostDevices)
    return set_status(st, USTBX_E_GPU, "invalid gpu id %d (%d devices)", device, ndev);
  // the caller's current device is restored on return
  int prev = 0;
  e = hipGetDevice(&prev);
  if (e != hipSuccess) return hip_fail(st, e, "hipGetDevice");
  e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(st, e, "hipSetDevice");
  struct Restore {
    int dev;
    ~Restore() { (void)hipSetDevice(dev); }
  } restore{prev};
  HostEntry& H = g_host[device];
  std::lock_guard<std::mutex> lock(H.mu);
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t total = al(bytes_ch) + 2 * al(bytes_tx) + 2 * al(bytes_rx) + al(bytes_out);
  e = H.reserve(total);
  if (e != hipSuccess) return hip_fail(st, e, "host entry buffers");
  char* p = H.dev;
  void* dch = p; p += al(bytes_ch);
  float* dta = (float*)p; p += al(bytes_tx);
  float* dtd = (float*)p; p += al(bytes_tx);
  float* dra = (float*)p; p += al(bytes_rx);
  float* drd = (float*)p; p += al(bytes_rx);
  void* dout = p;
  const HostCopy in[5] = {{dch, ch, bytes_ch}, {dta, tx_apo, bytes_tx}, {dtd, tx_delay, bytes_tx},
                          {dra, rx_apo, bytes_rx}, {drd, rx_delay, bytes_rx}};
  e = H.upload(in, 5);
  if (e != hipSuccess) return hip_fail(st, e, "host-to-device copies");
  const int result = ustbx_das_run(d, dch, dta, dra, dtd, drd, dout, H.stream, st);
  if (result) {
    (void)hipStreamSynchronize(H.stream);
    return result;
  }
  e = H.download(out, dout, bytes_out);
  if (e != hipSuccess) return hip_fail(st, e, "device-to-host copies");
  return ok(st);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Geometry plans
// ---------------------------------------------------------------------------

namespace {

// Tile order of the receive-table kernels. Dimension both (shares = 8):
// blocks go to the 8 XCDs round-robin and each XCD takes a contiguous range of
// tile indices (the XCD remap), i.e. whole scan lines (depth-fastest pixel
// order) with every depth, so the XCDs' work stays balanced. Within its range
// an XCD visits the tiles depth band by depth band across its lines: blocks
// resident together then cover a compact patch of the image, whose delays span
// a short window of every trace, and those windows stay in the XCD's L2.
// Dimension transmit (shares = 1): an XCD owns whole (channel group, frame
// chunk) slabs, each visited depth band by depth band across all lines (with
// patch tiles, tile_ps > 0: band by band across the groups of lines).
// Identity where the scan has no line-of-depths layout in whole tiles.
std::vector<int> band_major_order(const DevScan& sc, int tile_ps, int tiles, int shares) {
  std::vector<int> order(tiles);
  for (int t = 0; t < tiles; ++t) order[t] = t;
  if (!(sc.kind == SCAN_LINEAR || sc.kind == SCAN_SECTOR) || (tile_ps == 0 && sc.n0 % 64 != 0) || tiles < shares)
    return order;
  const int tpl = tile_ps ? sc.n0 >> tile_ps : sc.n0 / 64;  // tiles per line (per line group: patches)
  const int q = tiles / shares, r = tiles % shares;
  for (int k = 0; k < shares; ++k) {
    const int s0 = k * q + std::min(k, r);
    const int cnt = q + (k < r ? 1 : 0);
    std::stable_sort(order.begin() + s0, order.begin() + s0 + cnt, [tpl](int a, int b) {
      return a % tpl != b % tpl ? a % tpl < b % tpl : a / tpl < b / tpl;
    });
  }
  return order;
}

// The on-the-fly receive state covers dimension both with a windowed (or no)
// receive apodization without tilt; variant 3 (coefficients) keeps the table.
bool otf_applies(const ustbx_plan* pl) {
  const GeomLaunch& G = pl->G;
  if (G.dim != 3) return false;
  const DevRxApod& r = G.rx;
  if (r.kind == RXAP_ONES) return true;
  if (r.kind != RXAP_WINDOW || r.tilt != 0 || r.geo == RXGEO_CURV) return false;
  switch (r.window) {
    case WIN_BOXCAR: case WIN_HANNING: case WIN_HAMMING: case WIN_TUKEY25: case WIN_TUKEY50: case WIN_TUKEY75:
    case WIN_TRIANGLE:
      return true;
    default:
      return false;
  }
}

bool rxtab_applies(const ustbx_plan* pl, bool otf) {
  const GeomLaunch& G = pl->G;
  if (!(G.dim == 3 || G.dim == 2)) return false;
  // dimension both needs the separable (or RF) phase outside das_c's order;
  // the transmit kernel has das_c's term arithmetic and summation order too
  if (G.dim == 3 && (pl->dasc_order || !(G.separable || G.wd == 0.0f))) return false;
  if (G.edge == USTBX_EDGE_PYTHON && G.interp == USTBX_INTERP_LINEAR) return false;
  // dimension transmit gathers up to 8 frames through one buffer resource
  if (G.dim == 2 && (int64_t)G.L * G.N * G.W * 8 * 8 > 0x7FFFFE00) return false;
  if (otf && !otf_applies(pl)) return false;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return false;
  if (!otf && (size_t)G.N * (size_t)G.P * sizeof(float4) > (size_t)(kRxTableMaxFraction * (double)total_b))
    return false;
  return (int64_t)G.L * G.N * G.W * 8 <= 0x7FFFFE00;
}

// Whether a new plan takes the on-the-fly receive state: where it applies
// and the receive table would not fit (more than kRxTableMaxFraction of the
// device memory). Where the table fits it is the faster form (the 3-D
// class: 52 ms against 68 ms per frame), at far more memory (65.8 GiB
// against 1.8 GiB); ustbx_plan_set_variant(plan, 4) chooses the small plan.
bool otf_default(const ustbx_plan* pl) {
  const GeomLaunch& G = pl->G;
  if (!otf_applies(pl)) return false;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return false;
  return (size_t)G.N * (size_t)G.P * sizeof(float4) > (size_t)(kRxTableMaxFraction * (double)total_b);
}

// Variant 3 adds a coefficient buffer of 32 bytes per channel sample (cubic).
bool coef_applies(const ustbx_plan* pl) {
  const GeomLaunch& G = pl->G;
  return G.dim == 3 && G.interp == USTBX_INTERP_CUBIC && !G.rx_otf && (int64_t)G.L * G.N * G.W * 32 <= 0x7FFFFE00;
}

void rxtab_free(ustbx_plan* pl) {
  (void)hipFree(pl->rxtab);
  pl->rxtab = nullptr;
  pl->rxtab_bytes = 0;
  pl->rxtab_state = 0;
  GeomLaunch& G = pl->G;
  G.rx_table = nullptr;
  G.tile_nc = G.tile_ncf = nullptr;
  G.nan_mask = nullptr;
  G.stage_mask = nullptr;
  G.stage_words = 0;
  G.tile_chl = G.tile_waves = G.tile_nw = nullptr;
  G.ref_table = nullptr;
  G.tile_order = nullptr;
  pl->tile_order = pl->tile_natural = nullptr;
}

// Builds the variant-2 state (otf: with the receive state on the fly instead
// of the receive table) once, or again in the other receive mode; returns
// false (and frees it) when that form does not apply. Tiles where a nonzero
// weight or its delay is not finite go last in both visiting orders, for the
// kernels' SAFE form.
bool rxtab_build(ustbx_plan* pl, hipError_t& err, bool otf) {
  err = hipSuccess;
  if (pl->rxtab_state > 0 && (pl->G.rx_otf != 0) == otf) return true;
  if (pl->rxtab_state > 0 || (pl->rxtab_state < 0 && (pl->G.rx_otf != 0) != otf)) rxtab_free(pl);
  if (pl->rxtab_state != 0) return false;
  pl->rxtab_state = -1;
  pl->G.rx_otf = otf ? 1 : 0;
  if (!rxtab_applies(pl, otf)) return false;
  GeomLaunch& G = pl->G;
  const size_t P = (size_t)G.P, tiles = (P + 63) / 64;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  G.rx_compact = G.wd == 0.0f ? 1 : 0;  // RF: {rd, ra} (the phase factor is (ra, 0))
  const size_t b_rx = otf ? 0 : al((G.rx_compact ? sizeof(float2) : sizeof(float4)) * (size_t)G.N * P);
  const size_t b_nc = al(sizeof(int) * 2 * tiles);  // channel counts, then counts of the non-NaN ones
  const size_t b_cl = al(sizeof(int) * tiles * (size_t)G.N);
  const size_t b_wl = al(sizeof(int) * tiles * (size_t)G.W);
  const size_t b_nw = al(sizeof(int) * tiles);
  const size_t b_bad = al(sizeof(int) * tiles);
  const size_t b_ref = G.ref_phase ? al(sizeof(float2) * P) : 0;
  const size_t b_ord = al(sizeof(int) * tiles);
  const size_t b_mask = al(sizeof(unsigned long long) * tiles);
  const size_t b_otf = otf ? al(sizeof(RxOtf)) + al(sizeof(float4) * (size_t)G.N) : 0;
  // stage mask: the receive table, dimension both (bits by channel-list
  // position) or transmit (bits by channel)
  const int stage_words = (!otf && (G.dim == 3 || G.dim == 2)) ? (G.N + 63) / 64 : 0;
  const size_t b_smask = al(sizeof(unsigned long long) * tiles * (size_t)stage_words);
  const size_t total = b_rx + b_nc + b_cl + b_wl + b_nw + b_bad + b_ref + 2 * b_ord + b_mask + b_otf + b_smask;
  char* m = nullptr;
  err = hipMalloc(&m, total);
  if (err != hipSuccess) return false;
  size_t off = 0;
  auto take = [&](size_t b) {
    char* q = m + off;
    off += b;
    return q;
  };
  float4* rx = b_rx ? (float4*)take(b_rx) : nullptr;
  int* tnc = (int*)take(b_nc);
  int* tcl = (int*)take(b_cl);
  int* twl = (int*)take(b_wl);
  int* tnw = (int*)take(b_nw);
  int* tbad = (int*)take(b_bad);
  float2* ref = b_ref ? (float2*)take(b_ref) : nullptr;
  int* ord = (int*)take(b_ord);
  int* nat = (int*)take(b_ord);
  unsigned long long* nmask = (unsigned long long*)take(b_mask);
  RxOtf* dotf = b_otf ? (RxOtf*)take(al(sizeof(RxOtf))) : nullptr;
  float4* e32 = b_otf ? (float4*)take(b_otf - al(sizeof(RxOtf))) : nullptr;
  unsigned long long* smask = b_smask ? (unsigned long long*)take(b_smask) : nullptr;
  if (dotf) {
    G.otf.e32 = e32;
    err = launch_otf_elements(G, e32, nullptr);
    if (err == hipSuccess) err = hipMemcpy(dotf, &G.otf, sizeof(RxOtf), hipMemcpyHostToDevice);
  }
  G.d_otf = dotf;
  if (err == hipSuccess) err = launch_rxtab_tables(G, rx, ref, tnc, tcl, twl, tnw, tbad, nmask, nullptr);
  if (err == hipSuccess) err = hipDeviceSynchronize();
  std::vector<int> nw(tiles), nc(tiles), bad(tiles);
  if (err == hipSuccess) err = hipMemcpy(nw.data(), tnw, sizeof(int) * tiles, hipMemcpyDeviceToHost);
  if (err == hipSuccess) err = hipMemcpy(nc.data(), tnc, sizeof(int) * tiles, hipMemcpyDeviceToHost);
  if (err == hipSuccess) err = hipMemcpy(bad.data(), tbad, sizeof(int) * tiles, hipMemcpyDeviceToHost);
  int n_bad = 0, max_nw = 0;
  for (size_t t = 0; t < tiles; ++t) {
    n_bad += bad[t] != 0;
    max_nw = std::max(max_nw, nw[t]);
  }
  if (err != hipSuccess) {
    (void)hipFree(m);
    return false;
  }
  std::vector<int> order = band_major_order(G.scan, G.tile_ps, (int)tiles, G.dim == 2 ? 1 : 8);
  std::vector<int> batch_order = order;  // batches flag the SAFE tiles per item instead
  std::vector<int> natural(tiles);
  for (size_t t = 0; t < tiles; ++t) natural[t] = (int)t;
  auto finite_first = [&](int t) { return bad[t] == 0; };
  std::stable_partition(order.begin(), order.end(), finite_first);
  std::stable_partition(natural.begin(), natural.end(), finite_first);
  err = hipMemcpy(ord, order.data(), sizeof(int) * tiles, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(nat, natural.data(), sizeof(int) * tiles, hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    (void)hipFree(m);
    return false;
  }
  pl->rxtab = m;
  pl->rxtab_bytes = total;
  pl->tile_order = ord;
  pl->tile_natural = nat;
  pl->n_bad = n_bad;
  pl->max_nw = max_nw;
  pl->rxtab_state = 1;
  pl->h_order = std::move(batch_order);
  pl->h_cost.resize(tiles);
  pl->h_bad.resize(tiles);
  for (size_t t = 0; t < tiles; ++t) {
    pl->h_cost[t] = (int64_t)nw[t] * nc[t];
    pl->h_bad[t] = bad[t] != 0;
  }
  G.rx_table = rx;
  G.tile_nc = tnc;
  G.tile_ncf = tnc + tiles;
  G.nan_mask = nmask;
  G.tile_chl = tcl;
  G.tile_waves = twl;
  G.tile_nw = tnw;
  G.ref_table = ref;
  G.tile_order = ord;
  G.tile_good = (int)tiles - n_bad;
  if (smask) {
    G.stage_mask = smask;
    G.stage_words = stage_words;
    err = launch_stage_mask(G, tbad, smask, nullptr);
    if (err == hipSuccess) err = hipDeviceSynchronize();
    if (err != hipSuccess) {
      rxtab_free(pl);
      pl->rxtab_state = -1;
      return false;
    }
  }
  return true;
}

bool coef_build(ustbx_plan* pl, hipError_t& err) {
  err = hipSuccess;
  if (pl->G.coef) return true;
  if (!coef_applies(pl)) return false;
  const GeomLaunch& G = pl->G;
  const size_t bytes = (size_t)32 * G.L * G.N * G.W * (size_t)G.F;
  float4* c = nullptr;
  err = hipMalloc(&c, bytes);
  if (err != hipSuccess) return false;
  pl->G.coef = c;
  pl->coef_bytes = bytes;
  return true;
}

struct Arena {
  std::vector<char> host;
  size_t add(const void* src, size_t bytes) {
    const size_t off = (host.size() + 255) & ~size_t(255);
    host.resize(off + bytes);
    if (bytes) memcpy(host.data() + off, src, bytes);
    return off;
  }
};

// MATLAB's filter2(h, x, 'same') for a box h of k0 x k1 ones / (k0 k1) on an
// n0 x n1 array (column-major): the full correlation with zero padding, its
// central part from offset floor(k/2) on each axis.
std::vector<double> box_filter_same(const std::vector<double>& x, int n0, int n1, int k0, int k1) {
  std::vector<double> y((size_t)n0 * n1, 0.0);
  const double w = 1.0 / ((double)k0 * k1);
  for (int j = 0; j < n1; ++j)
    for (int i = 0; i < n0; ++i) {
      double acc = 0.0;
      // conv2(x, rot90(h, 2), 'same')(i, j) = sum of x over rows
      // i + floor(k0/2) - k0 + 1 .. i + floor(k0/2), likewise columns
      for (int a = i + k0 / 2 - k0 + 1; a <= i + k0 / 2; ++a)
        for (int b = j + k1 / 2 - k1 + 1; b <= j + k1 / 2; ++b)
          if (a >= 0 && a < n0 && b >= 0 && b < n1) acc += x[(size_t)a + (size_t)n0 * b];
      y[(size_t)i + (size_t)n0 * j] = acc * w;
    }
  return y;
}

// Scanline (MLA) transmit apodization (apodization.m:265-306), [W][P] in the
// scan's pixel order: wave i + j * n1 / MLA(1) holds the MLA(1) x MLA(2)
// block of scan lines from ((i-1) MLA(1), (j-1) MLA(2)), smoothed by
// filter2 with a box of MLA_overlap + 1 lines per axis, along all depths.
// Linear scans: lines (x, y); sector scans: (azimuth, elevation).
int scanline_matrix(const ustbx_scan& sc, int W, const int32_t mla[2], const int32_t ov[2],
                    std::vector<float>& out, ustbx_status* st) {
  if (sc.kind != USTBX_SCAN_LINEAR && sc.kind != USTBX_SCAN_SECTOR)
    return set_status(st, USTBX_E_ARG,
                      "The scan class does not support scanline based beamforming. This must be done manually, "
                      "defining several scan and setting the apodization to uff.window.none.");
  const int m1 = mla[0] > 0 ? mla[0] : 1, m2 = mla[1] > 0 ? mla[1] : 1;
  const int k0 = (ov[0] > 0 ? ov[0] : 0) + 1, k1 = (ov[1] > 0 ? ov[1] : 0) + 1;
  const int nz = sc.n0, nx = sc.n1, ny = sc.n2;
  if ((int64_t)W * m1 * m2 != (int64_t)nx * ny || nx % m1 || ny % m2)
    return set_status(st, USTBX_E_DIMENSIONS,
                      "The number of waves in the sequence does not match with the number of scanlines and set MLA.");
  const size_t P = (size_t)nz * nx * ny;
  out.assign((size_t)W * P, 0.0f);
  std::vector<double> a((size_t)nx * ny);
  for (int i = 0; i < nx / m1; ++i)
    for (int j = 0; j < ny / m2; ++j) {
      // circshift of the block at (0, 0) by ((i-1) MLA(1), (j-1) MLA(2))
      std::fill(a.begin(), a.end(), 0.0);
      for (int u = 0; u < m1; ++u)
        for (int v = 0; v < m2; ++v) a[(size_t)((i * m1 + u) % nx) + (size_t)nx * ((j * m2 + v) % ny)] = 1.0;
      const std::vector<double> b = box_filter_same(a, nx, ny, k0, k1);
      float* o = out.data() + (size_t)(i + j * (nx / m1)) * P;
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          const float v = (float)b[(size_t)x + (size_t)nx * y];
          for (int z = 0; z < nz; ++z) o[(size_t)z + (size_t)nz * (x + (size_t)nx * y)] = v;
        }
    }
  return USTBX_OK;
}

int check_apod(const ustbx_apod& a, const char* which, ustbx_status* st) {
  if (a.window < 0 || a.window > 9)
    return set_status(st, USTBX_E_ARG, "%s apodization: unknown window %d", which, a.window);
  return USTBX_OK;
}

}  // namespace

extern "C" {

int ustbx_plan_create(const ustbx_geometry* g, ustbx_plan** plan, ustbx_status* st) {
  if (!g || !plan) return set_status(st, USTBX_E_ARG, "NULL argument");
  *plan = nullptr;
  if (g->N <= 0 || g->W <= 0 || g->L < 2 || g->F <= 0)
    return set_status(st, USTBX_E_DIMENSIONS, "N, W, F must be >= 1 and L >= 2");
  if (!g->elem_x || !g->elem_y || !g->elem_z || !g->waves)
    return set_status(st, USTBX_E_ARG, "NULL element or wave array");
  if (g->dim < 0 || g->dim > 3) return set_status(st, USTBX_E_ARG, "Unknown sum mode");
  if (g->interp != USTBX_INTERP_LINEAR && g->interp != USTBX_INTERP_CUBIC)
    return set_status(st, USTBX_E_ARG, "unknown interpolation");
  if (g->edge != USTBX_EDGE_DASC && g->edge != USTBX_EDGE_PYTHON)
    return set_status(st, USTBX_E_ARG, "unknown edge rule");
  if (g->probe_kind != USTBX_PROBE_FLAT && g->probe_kind != USTBX_PROBE_CURVILINEAR)
    return set_status(st, USTBX_E_ARG, "unknown probe kind");
  if (g->probe_kind == USTBX_PROBE_CURVILINEAR && !(g->probe_radius > 0.0 && isfinite(g->probe_radius)))
    return set_status(st, USTBX_E_ARG, "a curvilinear probe needs a positive radius");
  int rc = check_apod(g->rx_apod, "receive", st);
  if (rc) return rc;
  rc = check_apod(g->tx_apod, "transmit", st);
  if (rc) return rc;
  const ustbx_scan& sc = g->scan;
  if (sc.kind < 0 || sc.kind > 2) return set_status(st, USTBX_E_ARG, "unknown scan kind");
  if (sc.n0 <= 0 || sc.n1 <= 0 || sc.n2 <= 0) return set_status(st, USTBX_E_DIMENSIONS, "empty scan");
  const int64_t P64 = (int64_t)sc.n0 * sc.n1 * sc.n2;
  if (P64 > INT32_MAX) return set_status(st, USTBX_E_CAPACITY, "more than 2^31-1 pixels");
  if (!sc.a0 || !sc.a1 || !sc.a2) return set_status(st, USTBX_E_ARG, "NULL scan axis");
  // gathers use 32-bit byte offsets within a frame, and an offset just below
  // 2^31 as the always-out-of-range position of skipped terms
  // the gathers address one wave's channel data with 32-bit byte offsets;
  // frames of 2 GiB or more run on variant 0 (per-(frame, wave) resources)
  if ((int64_t)g->L * g->N * 8 > 0x7FFFFE00)
    return set_status(st, USTBX_E_CAPACITY, "one wave of channel data (N x L samples) must be below 2 GiB");
  const int P = (int)P64, N = g->N, W = g->W;

  Arena ar;
  size_t off_a0, off_a1, off_a2;
  if (sc.kind == USTBX_SCAN_LINEAR) {
    off_a0 = ar.add(sc.a0, sizeof(double) * sc.n0);
    off_a1 = ar.add(sc.a1, sizeof(double) * sc.n1);
    off_a2 = ar.add(sc.a2, sizeof(double) * sc.n2);
  } else if (sc.kind == USTBX_SCAN_SECTOR) {
    off_a0 = ar.add(sc.a0, sizeof(double) * sc.n0);
    off_a1 = ar.add(sc.a1, sizeof(double) * 2 * sc.n1);
    off_a2 = ar.add(sc.a2, sizeof(double) * 2 * sc.n2);
  } else {
    off_a0 = ar.add(sc.a0, sizeof(double) * P);
    off_a1 = ar.add(sc.a1, sizeof(double) * P);
    off_a2 = ar.add(sc.a2, sizeof(double) * P);
  }
  const size_t off_ex = ar.add(g->elem_x, sizeof(double) * N);
  const size_t off_ey = ar.add(g->elem_y, sizeof(double) * N);
  const size_t off_ez = ar.add(g->elem_z, sizeof(double) * N);

  // ---- receive apodization (apodization.m:323-468) ----
  DevRxApod rx;
  memset(&rx, 0, sizeof(rx));
  const ustbx_apod& ra = g->rx_apod;
  size_t off_rvec = SIZE_MAX;
  std::vector<float> rvec;
  if (ra.matrix) {
    rx.kind = RXAP_MATRIX;
    rx.mat = ra.matrix;
  } else if (ra.vector) {
    rx.kind = RXAP_VECTOR;
    rvec.assign(ra.vector, ra.vector + N);
  } else if (ra.window == USTBX_WIN_NONE) {
    rx.kind = RXAP_ONES;
  } else if (ra.window == USTBX_WIN_STA) {
    // element(s) closest to the origin (apodization.m:345-349)
    if (!ra.has_origin) return set_status(st, USTBX_E_ARG, "origin must be set to use STA apodization");
    std::vector<double> dist(N);
    double mn = INFINITY;
    for (int n = 0; n < N; ++n) {
      const double dx = g->elem_x[n] - ra.origin[0], dy = g->elem_y[n] - ra.origin[1],
                   dz = g->elem_z[n] - ra.origin[2];
      dist[n] = sqrt((dx * dx + dy * dy) + dz * dz);
      mn = fmin(mn, dist[n]);
    }
    rx.kind = RXAP_VECTOR;
    rvec.resize(N);
    for (int n = 0; n < N; ++n) rvec[n] = dist[n] == mn ? 1.0f : 0.0f;
  } else if (ra.window == USTBX_WIN_SCANLINE) {  // a wave apodization only (apodization.m:192-220)
    return set_status(st, USTBX_E_ARG, "Unknown apodization type!");
  } else {
    rx.kind = RXAP_WINDOW;
    rx.window = ra.window;
    if (g->probe_kind == USTBX_PROBE_CURVILINEAR) {
      // the origin defaults to the arc's centre (apodization.m:378-383)
      rx.geo = RXGEO_CURV;
      rx.ox = ra.has_origin ? ra.origin[0] : 0.0;
      rx.oy = ra.has_origin ? ra.origin[1] : 0.0;
      rx.oz = ra.has_origin ? ra.origin[2] : -g->probe_radius;
    } else if (sc.kind == USTBX_SCAN_SECTOR || ra.has_origin) {
      rx.geo = RXGEO_ORIGIN;
      const double* o = ra.has_origin ? ra.origin : sc.origin;
      rx.ox = o[0]; rx.oy = o[1]; rx.oz = o[2];
    } else {
      rx.geo = RXGEO_FLAT;
    }
  }
  if (!rvec.empty()) off_rvec = ar.add(rvec.data(), sizeof(float) * N);
  rx.f0 = ra.f_number[0];
  rx.f1 = ra.f_number[1];
  rx.min0 = ra.minimum_aperture[0] * ra.f_number[0];
  rx.min1 = ra.minimum_aperture[1] * ra.f_number[1];
  rx.max0 = ra.maximum_aperture[0] * ra.f_number[0];
  rx.max1 = ra.maximum_aperture[1] * ra.f_number[1];
  rx.tilt = (fabs(ra.tilt[0]) > 0 ? 1 : 0) | (fabs(ra.tilt[1]) > 0 ? 2 : 0);
  rx.ct0 = cos(ra.tilt[0]); rx.st0 = sin(ra.tilt[0]);
  rx.ct1 = cos(ra.tilt[1]); rx.st1 = sin(ra.tilt[1]);
  rx.win0 = win_eval(ra.window, 0.0);

  // ---- transmit side (das.m:110-206, apodization.m:249-320, 472-555) ----
  const ustbx_apod& ta = g->tx_apod;
  DevTxApod tx;
  memset(&tx, 0, sizeof(tx));
  tx.window = ta.window;
  tx.f0 = ta.f_number[0];
  tx.f1 = ta.f_number[1];
  tx.min0 = ta.minimum_aperture[0] * ta.f_number[0];
  tx.min1 = ta.minimum_aperture[1] * ta.f_number[1];
  tx.max0 = ta.maximum_aperture[0] * ta.f_number[0];
  tx.max1 = ta.maximum_aperture[1] * ta.f_number[1];
  tx.gl0 = 0.5 * ta.grating_lobe_angle[0];
  tx.gl1 = 0.5 * ta.grating_lobe_angle[1];
  tx.mat = ta.matrix;
  std::vector<float> scanline;  // the scanline apodization, computed here and uploaded with the plan
  if (!ta.matrix && ta.window == USTBX_WIN_SCANLINE) {
    rc = scanline_matrix(sc, W, ta.mla, ta.mla_overlap, scanline, st);
    if (rc) return rc;
  }
  if (!ta.matrix && ta.window == USTBX_WIN_STA)
    return set_status(st, USTBX_E_UNSUPPORTED, "Unknown apodization type!");

  std::vector<DevWave> waves(W);
  bool has_unified = false;
  for (int w = 0; w < W; ++w) {
    const ustbx_wave& s = g->waves[w];
    DevWave& d = waves[w];
    memset(&d, 0, sizeof(d));
    d.sx = s.src[0]; d.sy = s.src[1]; d.sz = s.src[2];
    d.sdist = s.src_distance;
    d.ca = s.cos_az; d.sa = s.sin_az; d.ce = s.cos_el; d.se = s.sin_el;
    d.c = s.sound_speed;
    d.delay = s.delay;
    d.ox = s.origin[0]; d.oy = s.origin[1]; d.oz = s.origin[2];
    // delay kind
    if (g->tx_delay_matrix) {
      d.txk = TXK_MATRIX;
    } else if (s.wavefront == USTBX_WAVE_PLANE) {
      d.txk = TXK_PLANE;
    } else if (s.wavefront == USTBX_WAVE_PHOTOACOUSTIC) {
      d.txk = TXK_PHOTO;
    } else if (s.wavefront == USTBX_WAVE_SPHERICAL) {
      if (isinf(s.src_distance)) {
        d.txk = TXK_PLANE;
      } else if (s.src[2] < 0) {
        d.txk = TXK_DIVERGING;
      } else {
        switch (g->tx_model) {
          case USTBX_TXMODEL_SPHERICAL: d.txk = TXK_SPHERICAL; break;
          case USTBX_TXMODEL_HYBRID:
          case USTBX_TXMODEL_BLENDED:
            if (sc.kind == USTBX_SCAN_XYZ)
              return set_status(st, USTBX_E_UNSUPPORTED,
                                "Only linear scan and sector scan in 2D is supported for the "
                                "hybrid spherical transmit delay model.");
            if (g->tx_model == USTBX_TXMODEL_HYBRID)
              d.txk = sc.kind == USTBX_SCAN_LINEAR ? TXK_HYBRID_LIN : TXK_HYBRID_SEC;
            else
              d.txk = sc.kind == USTBX_SCAN_LINEAR ? TXK_BLENDED_LIN : TXK_BLENDED_SEC;
            break;
          case USTBX_TXMODEL_UNIFIED:
            if (!((sc.kind == USTBX_SCAN_LINEAR || sc.kind == USTBX_SCAN_SECTOR) && sc.n2 == 1))
              return set_status(st, USTBX_E_UNSUPPORTED,
                                "Only linear scan and sector scan in 2D is supported for the unified spherical "
                                "transmit delay model.");
            d.txk = TXK_UNIFIED;
            has_unified = true;
            break;
          default:
            return set_status(st, USTBX_E_ARG, "unknown spherical transmit delay model");
        }
      }
    } else {
      return set_status(st, USTBX_E_ARG, "Unknown wavefront. Check available options at uff.wavefront");
    }
    // incidence geometry of a wave with a finite source (apodization.m:489-552),
    // used by the transmit apodization and by the unified model's mask
    const bool plane_like = s.wavefront == USTBX_WAVE_PLANE || isinf(s.src_distance);
    if (!plane_like) {
      double OS[3] = {s.src[0] - s.origin[0], s.src[1] - s.origin[1], s.src[2] - s.origin[2]};
      const double nrm = sqrt((OS[0] * OS[0] + OS[1] * OS[1]) + OS[2] * OS[2]);
      if (nrm <= 2.220446049250313e-16) {  // STA (apodization.m:506-518)
        if (g->probe_kind == USTBX_PROBE_CURVILINEAR) {  // from the arc's centre to the source
          OS[0] = s.src[0]; OS[1] = s.src[1]; OS[2] = s.src[2] + g->probe_radius;
        } else {
          OS[0] = 0.0; OS[1] = 0.0; OS[2] = 1.0;
        }
      }
      const double ns = sqrt((OS[0] * OS[0] + OS[1] * OS[1]) + OS[2] * OS[2]);
      const double zu[3] = {OS[0] / ns, OS[1] / ns, OS[2] / ns};
      const double e1[3] = {1.0, 0.0, 0.0};
      // cross(a, b) = [a2 b3 - a3 b2, a3 b1 - a1 b3, a1 b2 - a2 b1]
      const double yu[3] = {zu[1] * e1[2] - zu[2] * e1[1], zu[2] * e1[0] - zu[0] * e1[2],
                            zu[0] * e1[1] - zu[1] * e1[0]};
      const double xu[3] = {zu[1] * yu[2] - zu[2] * yu[1], zu[2] * yu[0] - zu[0] * yu[2],
                            zu[0] * yu[1] - zu[1] * yu[0]};
      d.zu0 = zu[0]; d.zu1 = zu[1]; d.zu2 = zu[2];
      d.xu0 = xu[0]; d.xu1 = xu[1]; d.xu2 = xu[2];
      d.yu0 = yu[0]; d.yu1 = yu[1]; d.yu2 = yu[2];
      d.os_az = atan2(OS[0], OS[2]);
      d.os_el = atan2(OS[1], OS[2]);
    }
    // apodization kind
    if (ta.matrix || !scanline.empty()) {
      d.apk = TXAP_MATRIX;
    } else if (ta.vector) {
      d.apk = TXAP_CONST;
      d.apo_const = ta.vector[w];
    } else if (ta.window == USTBX_WIN_NONE) {
      d.apk = TXAP_CONST;
      d.apo_const = 1.0f;
    } else if (plane_like) {
      // plane wave: constant per wave (apodization.m:481-487)
      const double tt = tan(s.azimuth - ta.tilt[0]);
      const double tp = tan(s.elevation - ta.tilt[1]);
      const double rt = fabs(ta.f_number[0] * tt), rp = fabs(ta.f_number[1] * tp);
      d.apk = TXAP_CONST;
      d.apo_const = (float)(win_eval(ta.window, rt) * win_eval(ta.window, rp));
    } else {
      d.apk = TXAP_INCIDENCE;
    }
  }
  const size_t off_waves = ar.add(waves.data(), sizeof(DevWave) * W);
  const size_t off_tmat = scanline.empty() ? SIZE_MAX : ar.add(scanline.data(), sizeof(float) * scanline.size());

  // ---- upload ----
  ustbx_plan* pl = new (std::nothrow) ustbx_plan();  // value-initialized
  if (!pl) return set_status(st, USTBX_E_CAPACITY, "host allocation failed");
  hipError_t e = hipGetDevice(&pl->device);
  if (e != hipSuccess) {
    delete pl;
    return hip_fail(st, e, "hipGetDevice");
  }
  pl->dev_bytes = ar.host.size();
  e = hipMalloc(&pl->dev, pl->dev_bytes);
  if (e != hipSuccess) {
    delete pl;
    return hip_fail(st, e, "hipMalloc (plan)");
  }
  e = hipMemcpy(pl->dev, ar.host.data(), pl->dev_bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(pl->dev);
    delete pl;
    return hip_fail(st, e, "hipMemcpy (plan)");
  }
  char* base = pl->dev;
  GeomLaunch& G = pl->G;
  memset(&G, 0, sizeof(G));
  G.L = g->L; G.N = N; G.W = W; G.F = g->F; G.P = P;
  G.fs = g->fs;
  G.t0 = g->t0;
  G.wd = dasc_wd(g->fd);
  G.dim = g->dim;
  G.interp = g->interp;
  G.edge = g->edge;
  // das_c's summation order comes with das_c's phase exp(i wd (tx+rx))
  G.separable = ((g->flags & USTBX_FLAG_SEPARABLE_PHASE) && !(g->flags & USTBX_FLAG_DASC_ORDER)) ? 1 : 0;
  G.ref_phase = ((g->flags & USTBX_FLAG_REF_PHASE) && G.wd != 0.0f) ? 1 : 0;
  // das.m:308: exp(-1j*2*w0*r_ref/c); w0 = 2*pi*single(fd) is the FP32 value
  // das_c calls wd (das.m:88-89); the argument in FP64 (das.py) or in FP32
  // (das.m) as the flags select
  G.w0_ref = 2.0 * (double)G.wd / g->sound_speed;
  G.ref_single = (g->flags & USTBX_FLAG_REF_PHASE_SINGLE) ? 1 : 0;
  G.c32 = (float)g->sound_speed;
  G.scan.kind = sc.kind;
  G.scan.n0 = sc.n0; G.scan.n1 = sc.n1; G.scan.n2 = sc.n2;
  G.scan.a0 = (const double*)(base + off_a0);
  G.scan.a1 = (const double*)(base + off_a1);
  G.scan.a2 = (const double*)(base + off_a2);
  G.scan.ox = sc.origin[0]; G.scan.oy = sc.origin[1]; G.scan.oz = sc.origin[2];
  G.scan.ref_kind = sc.ref_kind;
  G.ex = (const double*)(base + off_ex);
  G.ey = (const double*)(base + off_ey);
  G.ez = (const double*)(base + off_ez);
  G.waves = (const DevWave*)(base + off_waves);
  if (off_rvec != SIZE_MAX) rx.vec = (const float*)(base + off_rvec);
  if (off_tmat != SIZE_MAX) tx.mat = (const float*)(base + off_tmat);
  G.rx = rx;
  G.tx = tx;
  G.tx_delay_mat = g->tx_delay_matrix;
  G.c = g->sound_speed;
  G.lens = g->lens_delay;
  G.pw_margin = g->pw_margin;
  G.bpow = g->blending_power;
  const bool dasc_order = (g->flags & USTBX_FLAG_DASC_ORDER) != 0;
  G.wave_chunk = dasc_order ? 1 : (G.separable ? (W >= 16 ? 16 : 8) : 8);
  G.split = dasc_order ? 1 : ((G.dim == 3 || G.dim == 2) ? auto_split(P) : 1);
  G.variant = 0;
  // LDS-staged sample windows: on for dimension both (where the plan's stage
  // mask marks channels); off for dimension transmit, where its per-(channel,
  // wave) window search measured slower than the gathers
  G.stage = G.dim == 3 ? 1 : 0;
  // transmit-kernel tiles: patches of 16 depths x 4 scan lines where the
  // scan's axes divide evenly (the lanes' samples then span fewer cache
  // lines per gather than along 64 depths: C3 11.0 -> 9.4 ms; tiles along
  // the depth stay for dimension both, whose receive apertures they follow)
  G.tile_ps = (G.dim == 2 && (sc.kind == SCAN_LINEAR || sc.kind == SCAN_SECTOR) && sc.n0 % 16 == 0 &&
               sc.n1 % 4 == 0) ? 4 : 0;
  if (!dasc_order && shared_rx_applies(G)) {
    G.variant = 1;
    shared_rx_default_tiles(W, G.wave_chunk, G.split);
  }

  // receive aperture culling: one line of elements sorted by x, a window that
  // is zero above ratio 1/2, no tilt (see GeomLaunch::rx_cull)
  bool cull = rx.kind == RXAP_WINDOW && rx.tilt == 0 && rx.geo != RXGEO_CURV && N >= 2 &&
              (rx.window == WIN_BOXCAR || rx.window == WIN_HANNING || rx.window == WIN_HAMMING ||
               rx.window == WIN_TUKEY25 || rx.window == WIN_TUKEY50 || rx.window == WIN_TUKEY75 ||
               rx.window == WIN_TRIANGLE) &&
              rx.f0 != 0.0;
  for (int n = 1; cull && n < N; ++n)
    cull = g->elem_x[n] > g->elem_x[n - 1] && g->elem_y[n] == g->elem_y[0] && g->elem_z[n] == g->elem_z[0];
  G.rx_cull = cull ? 1 : 0;

  // unified-model delays (das.m:133-152) into the plan's [W][P] matrix
  if (has_unified) {
    pl->unified_bytes = (size_t)W * (size_t)P * sizeof(float);
    e = hipMalloc(&pl->unified_tx, pl->unified_bytes);
    if (e == hipSuccess) {
      // das.m:136-149: boxcar, no minimum aperture, the transmit apodization's
      // f-number, uff.apodization's other defaults
      DevTxApod mk;
      memset(&mk, 0, sizeof(mk));
      mk.window = USTBX_WIN_BOXCAR;
      mk.f0 = ta.f_number[0];
      mk.f1 = ta.f_number[1];
      mk.min0 = 0.0 * ta.f_number[0];
      mk.min1 = 0.0 * ta.f_number[1];
      mk.max0 = 10.0 * ta.f_number[0];
      mk.max1 = 10.0 * ta.f_number[1];
      mk.gl0 = 0.5 * M_PI;
      mk.gl1 = 0.5 * M_PI;
      G.tx_delay_mat = pl->unified_tx;
      e = launch_geom_unified_tx(G, mk, pl->unified_tx, nullptr);
      if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    if (e != hipSuccess) {
      (void)hipFree(pl->unified_tx);
      (void)hipFree(pl->dev);
      delete pl;
      return hip_fail(st, e, "unified transmit delay model");
    }
  }

  // resident transmit table
  pl->tx_table_bytes = (size_t)W * (size_t)P * sizeof(float4);
  e = hipMalloc(&pl->tx_table, pl->tx_table_bytes);
  if (e != hipSuccess) {
    (void)hipFree(pl->unified_tx);
    (void)hipFree(pl->dev);
    delete pl;
    return hip_fail(st, e, "hipMalloc (transmit table)");
  }
  G.tx_table = pl->tx_table;
  e = launch_geom_tx_table(G, pl->tx_table, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(pl->tx_table);
    (void)hipFree(pl->unified_tx);
    (void)hipFree(pl->dev);
    delete pl;
    return hip_fail(st, e, "transmit table");
  }
  // the receive-table kernel where it applies, with the receive state on the
  // fly where the table would be large
  pl->dasc_order = dasc_order;
  G.otf.scan = G.scan;
  G.otf.ex = G.ex;
  G.otf.ey = G.ey;
  G.otf.ez = G.ez;
  G.otf.inv_c = 1.0 / G.c;
  G.otf.lens = G.lens;
  G.otf.inv_c32 = (float)(1.0 / G.c);
  G.otf.lens32 = (float)G.lens;
  G.otf.wd = G.wd;
  G.otf.compact = G.wd == 0.0f ? 1 : 0;
  {
    RxOtfApod& a = G.otf.ap;
    a.kind = G.rx.kind == RXAP_ONES ? RXAP_ONES : RXAP_WINDOW;
    a.window = G.rx.window;
    a.geo = G.rx.geo;
    a.f0 = G.rx.f0;
    a.f1 = G.rx.f1;
    a.min0 = G.rx.min0;
    a.min1 = G.rx.min1;
    a.max0 = G.rx.max0;
    a.max1 = G.rx.max1;
    a.ox = G.rx.ox;
    a.oy = G.rx.oy;
    a.oz = G.rx.oz;
    a.win0 = (float)G.rx.win0;
    G.otf.f0f = (float)a.f0;
    G.otf.f1f = (float)a.f1;
    G.otf.min0f = (float)a.min0;
    G.otf.min1f = (float)a.min1;
    G.otf.max0f = (float)a.max0;
    G.otf.max1f = (float)a.max1;
    G.otf.oxf = (float)a.ox;
    G.otf.oyf = (float)a.oy;
  }
  if (rxtab_build(pl, e, otf_default(pl))) {
    G.variant = 2;
    rxtab_defaults(G, pl->max_nw, G.wave_chunk, G.split);
  } else if (e != hipSuccess) {
    (void)hipFree(pl->tx_table);
    (void)hipFree(pl->unified_tx);
    (void)hipFree(pl->dev);
    delete pl;
    return hip_fail(st, e, "receive table");
  }
  *plan = pl;
  return ok(st);
}

int ustbx_plan_set_variant(ustbx_plan* plan, int32_t variant, ustbx_status* st) {
  if (!plan) return set_status(st, USTBX_E_ARG, "NULL plan");
  if (variant < 0 || variant > 4) return set_status(st, USTBX_E_ARG, "variant must be 0, 1, 2, 3 or 4");
  if (variant == 4) {
    // variant 2 with the receive state on the fly
    hipError_t e;
    if (!rxtab_build(plan, e, true)) {
      if (e != hipSuccess) return hip_fail(st, e, "receive-state tables");
      return set_status(st, USTBX_E_UNSUPPORTED,
                        "the on-the-fly receive state needs dimension both with the separable (or RF) phase, and a "
                        "windowed receive apodization without tilt");
    }
    if (plan->G.variant != 2) {
      plan->G.variant = 2;
      rxtab_defaults(plan->G, plan->max_nw, plan->G.wave_chunk, plan->G.split);
    }
    return ok(st);
  }
  if (variant == 1 && (plan->dasc_order || !shared_rx_applies(plan->G)))
    return set_status(st, USTBX_E_UNSUPPORTED,
                      "the shared-receive kernel needs dimension both, IQ data, the separable phase and one frame");
  if (variant >= 2) {
    hipError_t e;
    if (!rxtab_build(plan, e, false)) {
      if (e != hipSuccess) return hip_fail(st, e, "receive table");
      return set_status(st, USTBX_E_UNSUPPORTED,
                        "the receive-table kernels need dimension both with the separable (or RF) phase outside das_c's "
                        "order, or dimension transmit, and a receive table within half the device memory");
    }
    if (variant == 3 && !coef_build(plan, e)) {
      if (e != hipSuccess) return hip_fail(st, e, "coefficient buffer");
      return set_status(st, USTBX_E_UNSUPPORTED,
                        "the coefficient variant needs dimension both, cubic interpolation and 32 bytes per sample within 2 GiB "
                        "per frame");
    }
  }
  if (variant == plan->G.variant) return ok(st);
  const bool same_family = variant >= 2 && plan->G.variant >= 2;
  plan->G.variant = variant;
  if (same_family) return ok(st);
  if (variant >= 2) {
    rxtab_defaults(plan->G, plan->max_nw, plan->G.wave_chunk, plan->G.split);
  } else if (variant == 1) {
    shared_rx_default_tiles(plan->G.W, plan->G.wave_chunk, plan->G.split);
  } else {
    plan->G.wave_chunk = plan->G.separable ? 16 : 8;
    plan->G.split = auto_split(plan->G.P);
  }
  return ok(st);
}

int ustbx_plan_set_tiles(ustbx_plan* plan, int32_t wave_chunk, int32_t split, ustbx_status* st) {
  if (!plan) return set_status(st, USTBX_E_ARG, "NULL plan");
  if (plan->G.variant >= 2 && plan->G.dim == 2) {
    if (wave_chunk && wave_chunk != 4 && wave_chunk != 8)
      return set_status(st, USTBX_E_ARG, "frames per block must be 4 or 8");
    if (split && split != 4 && split != 8) return set_status(st, USTBX_E_ARG, "channels per block must be 4 or 8");
    if (wave_chunk) plan->G.wave_chunk = wave_chunk;
    if (split) plan->G.split = split;
    return ok(st);
  }
  if (plan->G.variant >= 2) {
    if (wave_chunk && !rxtab_wc_ok(wave_chunk))
      return set_status(st, USTBX_E_ARG, "waves per split must be 4, 5, 6, 8, 10, 12 or 16");
    if (split && split != 4 && split != 5 && split != 8)
      return set_status(st, USTBX_E_ARG, "wave-splits per block must be 4, 5 or 8");
    if (wave_chunk) plan->G.wave_chunk = wave_chunk;
    if (split) plan->G.split = split;
    return ok(st);
  }
  if (plan->G.variant == 1) {
    if (wave_chunk && !shared_rx_wc_ok(wave_chunk))
      return set_status(st, USTBX_E_ARG, "waves per split must be 8, 10, 12 or 16");
    if (split && split != 4 && split != 8)
      return set_status(st, USTBX_E_ARG, "wave-splits per block must be 4 or 8");
    if (wave_chunk) plan->G.wave_chunk = wave_chunk;
    if (split) plan->G.split = split;
    return ok(st);
  }
  if (wave_chunk) {
    if (wave_chunk != 1 && wave_chunk != 8 && wave_chunk != 16)
      return set_status(st, USTBX_E_ARG, "wave_chunk must be 1, 8 or 16");
    if (wave_chunk == 16 && !plan->G.separable)
      return set_status(st, USTBX_E_ARG, "wave_chunk 16 needs the separable phase");
    if (wave_chunk == 1 && plan->G.separable)
      return set_status(st, USTBX_E_ARG, "wave_chunk 1 needs the das_c phase");
    plan->G.wave_chunk = wave_chunk;
  }
  if (split) {
    if (split != 1 && split != 4) return set_status(st, USTBX_E_ARG, "split must be 1 or 4");
    if (split != 1 && !(plan->G.dim == 3 || plan->G.dim == 2))
      return set_status(st, USTBX_E_ARG, "split applies to dimension both/transmit only");
    plan->G.split = split;
  }
  if (plan->G.wave_chunk == 1) plan->G.split = 1;
  return ok(st);
}

int ustbx_plan_run(const ustbx_plan* plan, const void* ch, void* out, void* stream,
                   ustbx_status* st) {
  if (!plan) return set_status(st, USTBX_E_ARG, "NULL plan");
  if (!ch || !out) return set_status(st, USTBX_E_ARG, "NULL channel data or output");
  GeomLaunch G = plan->G;
  G.ch = ch;
  G.out = out;
  hipError_t e = launch_das_geom(G, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(st, e, "das_geom_kernel launch");
  return ok(st);
}

int64_t ustbx_plan_cf_workspace_bytes(const ustbx_plan* plan) {
  if (!plan || plan->G.dim != USTBX_DIM_TX || plan->G.variant != 2) return 0;
  return (int64_t)tx_cf_workspace_bytes(plan->G);
}

int ustbx_plan_run_cf(const ustbx_plan* plan, const void* ch, void* image, float* cf, const float* M,
                      int32_t flavour, void* workspace, int64_t workspace_bytes, void* stream, ustbx_status* st) {
  if (!plan) return set_status(st, USTBX_E_ARG, "NULL plan");
  if (!ch || !image) return set_status(st, USTBX_E_ARG, "NULL channel data or image");
  if (flavour != USTBX_CF_MATLAB && flavour != USTBX_CF_PYTHON)
    return set_status(st, USTBX_E_ARG, "flavour must be USTBX_CF_MATLAB or USTBX_CF_PYTHON");
  const GeomLaunch& P0 = plan->G;
  if (P0.dim != USTBX_DIM_TX || P0.variant != 2)
    return set_status(st, USTBX_E_UNSUPPORTED,
                      "the fused coherence factor needs a dimension-transmit plan on the receive-table kernel "
                      "(variant 2)");
  if (P0.N < 2) return set_status(st, USTBX_E_ARG, "Not enough channels to compute factor");
  const size_t need = tx_cf_workspace_bytes(P0);
  if (!workspace || workspace_bytes < 0 || (size_t)workspace_bytes < need)
    return set_status(st, USTBX_E_ARG, "workspace of %zu bytes needed (ustbx_plan_cf_workspace_bytes)", need);
  GeomLaunch G = P0;
  G.ch = ch;
  hipError_t e = launch_tx_cf(G, reinterpret_cast<float4*>(workspace), reinterpret_cast<float2*>(image), cf, M,
                              flavour, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(st, e, "das_tx_cf kernels launch");
  return ok(st);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free form of ustbx_plan_run_cf: the per-group partial sums go to
// a file-scope device array instead of a caller workspace. Each GPU holds its
// own copy of the array from library load to process end; host code reaches
// it only through the address hipGetSymbolAddress gives on that device.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices];

// The device address of `scratch` on `device`, looked up once per device with
// that device current (the caller's current device is restored).
hipError_t scratch_address(int device, void** ptr) {
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int prev = 0;
    hipError_t e = hipGetDevice(&prev);
    if (e != hipSuccess) return e;
    if (prev != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    const hipError_t e2 = prev != device ? hipSetDevice(prev) : hipSuccess;
    if (e != hipSuccess) return e;
    if (e2 != hipSuccess) return e2;
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  return hipSuccess;
}

}  // namespace

extern "C" {

int64_t ustbx_scratch_bytes(void) { return (int64_t)kScratchBytes; }

int ustbx_plan_run_cf_noalloc(const ustbx_plan* plan, const void* ch, void* image, float* cf, const float* M,
                              int32_t flavour, void* stream, ustbx_status* st) {
  if (!plan) return set_status(st, USTBX_E_ARG, "NULL plan");
  const GeomLaunch& P0 = plan->G;
  if (P0.dim != USTBX_DIM_TX || P0.variant != 2)
    return set_status(st, USTBX_E_UNSUPPORTED,
                      "the fused coherence factor needs a dimension-transmit plan on the receive-table kernel "
                      "(variant 2)");
  const size_t need = tx_cf_workspace_bytes(P0);
  if (need > kScratchBytes)
    return set_status(st, USTBX_E_CAPACITY,
                      "the partial sums need %zu bytes, more than the %zu-byte device scratch of the allocation-free "
                      "entry; use ustbx_plan_run_cf with a workspace",
                      need, kScratchBytes);
  void* ws = nullptr;
  const hipError_t e = scratch_address(plan->device, &ws);
  if (e != hipSuccess) return hip_fail(st, e, "hipGetSymbolAddress (scratch)");
  return ustbx_plan_run_cf(plan, ch, image, cf, M, flavour, ws, (int64_t)kScratchBytes, stream, st);
}

int ustbx_plan_active_elements(const ustbx_plan* plan, int32_t dim, double criterium, float* M, void* stream,
                               ustbx_status* st) {
  if (!plan) return set_status(st, USTBX_E_ARG, "NULL plan");
  if (!(dim == USTBX_DIM_RX || dim == USTBX_DIM_TX || dim == USTBX_DIM_BOTH))
    return set_status(st, USTBX_E_ARG, "Unknown dimension mode; check HELP dimension");
  if (!M) return set_status(st, USTBX_E_ARG, "NULL output");
  hipError_t e = launch_geom_active_elements(plan->G, dim, criterium, M, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(st, e, "geom_active_elements_kernel launch");
  return ok(st);
}

int ustbx_plan_materialize(const ustbx_plan* plan, float* tx_apo, float* rx_apo,
                           float* tx_delay, float* rx_delay, void* stream, ustbx_status* st) {
  if (!plan) return set_status(st, USTBX_E_ARG, "NULL plan");
  hipError_t e = launch_geom_materialize(plan->G, tx_apo, rx_apo, tx_delay, rx_delay,
                                         reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(st, e, "geom_materialize_kernel launch");
  return ok(st);
}

int ustbx_plan_active_terms(const ustbx_plan* plan, int64_t* terms, ustbx_status* st) {
  if (!plan || !terms) return set_status(st, USTBX_E_ARG, "NULL plan or result");
  unsigned long long* d = nullptr;
  hipError_t e = hipMalloc(&d, sizeof(*d));
  if (e != hipSuccess) return hip_fail(st, e, "hipMalloc (count)");
  unsigned long long h = 0;
  e = hipMemcpy(d, &h, sizeof(h), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_count_active(plan->G, d, nullptr);
  if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(st, e, "geom_count_active_kernel");
  *terms = (int64_t)h * plan->G.F;
  return ok(st);
}

int64_t ustbx_plan_pixels(const ustbx_plan* plan) { return plan ? plan->G.P : 0; }

int64_t ustbx_plan_device_bytes(const ustbx_plan* plan) {
  return plan ? (int64_t)(plan->dev_bytes + plan->tx_table_bytes + plan->unified_bytes + plan->rxtab_bytes +
                          plan->coef_bytes)
              : 0;
}

int ustbx_plan_set_order(ustbx_plan* plan, int32_t order, ustbx_status* st) {
  if (!plan) return set_status(st, USTBX_E_ARG, "NULL plan");
  if (order != 0 && order != 1) return set_status(st, USTBX_E_ARG, "order must be 0 or 1");
  if (plan->rxtab_state <= 0)
    return set_status(st, USTBX_E_UNSUPPORTED, "the tile order applies to the receive-table variants");
  // tiles that take the SAFE body are last in either order, which then
  // stays explicit
  plan->G.tile_order = order ? plan->tile_order : (plan->n_bad > 0 ? plan->tile_natural : nullptr);
  return ok(st);
}

int ustbx_plan_set_staging(ustbx_plan* plan, int32_t on, ustbx_status* st) {
  if (!plan) return set_status(st, USTBX_E_ARG, "NULL plan");
  if (on != 0 && on != 1) return set_status(st, USTBX_E_ARG, "staging must be 0 or 1");
  plan->G.stage = on;
  return ok(st);
}

int ustbx_plan_stage_stats(const ustbx_plan* plan, int64_t* staged, int64_t* listed, ustbx_status* st) {
  if (!plan || !staged || !listed) return set_status(st, USTBX_E_ARG, "NULL argument");
  *staged = 0;
  *listed = 0;
  const GeomLaunch& G = plan->G;
  if (plan->rxtab_state <= 0 || !G.tile_nc) return ok(st);
  const size_t tiles = ((size_t)G.P + 63) / 64;
  std::vector<int> nc(tiles);
  std::vector<unsigned long long> m(tiles * (size_t)G.stage_words);
  hipError_t e = hipSetDevice(plan->device);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(nc.data(), G.tile_nc, sizeof(int) * tiles, hipMemcpyDeviceToHost);
  if (e == hipSuccess && !m.empty())
    e = hipMemcpy(m.data(), G.stage_mask, sizeof(unsigned long long) * m.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(st, e, "stage statistics");
  for (size_t t = 0; t < tiles; ++t) *listed += nc[t];
  for (unsigned long long w : m) *staged += __builtin_popcountll(w);
  return ok(st);
}

int32_t ustbx_plan_variant(const ustbx_plan* plan, int32_t* wave_chunk, int32_t* split) {
  if (!plan) return -1;
  if (wave_chunk) *wave_chunk = plan->G.wave_chunk;
  if (split) *split = plan->G.split;
  return (plan->G.variant == 2 && plan->G.rx_otf) ? 4 : plan->G.variant;
}

void ustbx_plan_destroy(ustbx_plan* plan) {
  if (!plan) return;
  (void)hipFree(plan->G.coef);
  (void)hipFree(plan->rxtab);
  (void)hipFree(plan->tx_table);
  (void)hipFree(plan->unified_tx);
  (void)hipFree(plan->dev);
  delete plan;
}

}  // extern "C"
