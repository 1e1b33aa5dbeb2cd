// das_batch.hip — many acquisitions (different probes, wave counts, scan
// sizes; dimension both) beamformed by one persistent launch.
//
// Work items are (acquisition, 64-pixel tile, frame) of the acquisitions'
// plans (variant-2 tables). They sit in 8 queues, one per XCD: queue k holds,
// for every acquisition, the tiles the one-plan kernel would run on XCD k (a
// balanced share of every depth), in the plan's band-major order, with the
// acquisitions in decreasing order of work. The grid is the number of blocks
// the device holds at once; each block claims items from its own XCD's queue
// with an atomic counter and, when that queue is empty, from the others in
// turn. Every item runs the same tile body as the one-plan kernel (the SAFE
// form on tiles with a non-finite nonzero weight or delay), so each image
// equals the one-plan launch with the same tile shape bit for bit.
//
// A launch is a counter-reset kernel followed by the persistent kernel (no
// memset, no allocation, no host synchronization), so it can be captured in a
// hipGraph; the acquisitions' buffers are read from a device descriptor array
// that set_io rewrites, so a captured graph serves new buffers.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <numeric>
#include <vector>

#include "../../include/ustbx_das.h"
#include "das_plan.h"
#include "das_rxtab.h"
#include "das_status.h"

using namespace ustbx;

namespace {

constexpr int kQueues = 8;

struct BatchLaunch {
  const RxAcq* acqs;
  const int4* items;  // {acquisition, tile, frame, 1 where the tile takes the SAFE body}
  const int* qoff;    // [kQueues + 1] queue offsets into items
  int* ctr;           // [kQueues] claim counters, zeroed before each launch
};

// A descriptor copied word by word through the constant address space, so
// the compiler reads it with scalar loads and keeps it in SGPRs.
__device__ __forceinline__ RxAcq load_constant(const RxAcq* p) {
  static_assert(sizeof(RxAcq) % 4 == 0, "RxAcq is a whole number of words");
  RxAcq a;
  const __attribute__((address_space(4))) uint32_t* src = (const __attribute__((address_space(4))) uint32_t*)p;
  uint32_t* dst = reinterpret_cast<uint32_t*>(&a);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(RxAcq) / 4); ++i) dst[i] = src[i];
  return a;
}

__global__ void batch_reset_kernel(int* ctr) {
  if (threadIdx.x < kQueues) ctr[threadIdx.x] = 0;
}

// OTF: the form for batches with an on-the-fly acquisition (its tile body,
// rxotf_tile, and LDS for its receive entries). The batch kernels gather
// every term (no LDS-staged windows): the bodies a persistent kernel holds
// set its registers for all of them, and without the staged and on-the-fly
// bodies the kernel keeps 78 VGPRs and 6 waves per SIMD (97 and 4 with them:
// C5 served 51 frames/s against 61).
template <int SM, int WC, int WS, bool OTF>
__global__ __launch_bounds__(64 * WS) void das_batch_kernel(BatchLaunch B) {
  __shared__ float2 red[WS][64];
  constexpr int kRxs = OTF ? (int)sizeof(float4) * kOtfChunk * 64 : 16;  // on-the-fly receive entries
  __shared__ __align__(16) char pool[kRxs];
  float4(*rxs)[64] = reinterpret_cast<float4(*)[64]>(pool);
  __shared__ int claimed;
  const int home = (int)(blockIdx.x % kQueues);  // blocks reach the XCDs round-robin
  for (;;) {
    if (threadIdx.x == 0) {
      int it = -1;
      for (int k = 0; k < kQueues && it < 0; ++k) {
        const int q = (home + k) % kQueues;
        const int n = B.qoff[q + 1] - B.qoff[q];
        if (n == 0) continue;
        const int i = atomicAdd(&B.ctr[q], 1);
        if (i < n) it = B.qoff[q] + i;
      }
      claimed = it;
    }
    __syncthreads();
    // block-uniform: the item and its descriptor are read through the
    // constant address space (scalar loads into SGPRs; both are read-only
    // while the kernel runs)
    const int it = __builtin_amdgcn_readfirstlane(claimed);
    if (it < 0) break;  // every queue is empty
    const __attribute__((address_space(4))) int* w = (const __attribute__((address_space(4))) int*)B.items + 4 * it;
    const int acq = w[0], tile = w[1], frame = w[2];
    const __attribute__((address_space(4))) RxAcq* ad = (const __attribute__((address_space(4))) RxAcq*)(B.acqs + acq);
    if (OTF && ad->rx_otf) {  // the receive state computed in the block
      if (w[3])
        rxotf_tile<SM, WC, WS, true>(load_constant(B.acqs + acq), tile, frame, red, rxs);
      else
        rxotf_tile<SM, WC, WS, false>(load_constant(B.acqs + acq), tile, frame, red, rxs);
    } else if (w[3]) {  // a nonzero weight or its delay not finite on this tile
      rxtab_tile<SM, WC, WS, true>(load_constant(B.acqs + acq), tile, frame, red);
    } else {
      rxtab_tile<SM, WC, WS, false>(load_constant(B.acqs + acq), tile, frame, red);
    }
    __syncthreads();  // all threads are past this item before the next claim
  }
}

template <int SM, int WC, int WS, bool OTF>
hipError_t batch_launch_t(const BatchLaunch& B, int nblocks, hipStream_t stream) {
  hipLaunchKernelGGL((das_batch_kernel<SM, WC, WS, OTF>), dim3(nblocks), dim3(64 * WS), 0, stream, B);
  return hipGetLastError();
}

template <int SM, int WC, int WS, bool OTF>
int batch_occupancy_t() {
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, das_batch_kernel<SM, WC, WS, OTF>, 64 * WS, 0) !=
      hipSuccess)
    return 0;
  return occ;
}

// Dispatch over the instantiated (sampling mode, waves per split, splits,
// on-the-fly form): f is a functor with a template <int SM, int WC, int WS,
// bool OTF> call operator.
template <int SM, int WS, bool OTF, typename F>
auto dispatch_wc(int wc, const F& f) {
  switch (wc) {
    case 4: return f.template operator()<SM, 4, WS, OTF>();
    case 5: return f.template operator()<SM, 5, WS, OTF>();
    case 6: return f.template operator()<SM, 6, WS, OTF>();
    case 8: return f.template operator()<SM, 8, WS, OTF>();
    case 10: return f.template operator()<SM, 10, WS, OTF>();
    case 12: return f.template operator()<SM, 12, WS, OTF>();
    default: return f.template operator()<SM, 16, WS, OTF>();
  }
}

template <int SM, bool OTF, typename F>
auto dispatch_ws(int wc, int ws, const F& f) {
  if (ws == 4) return dispatch_wc<SM, 4, OTF>(wc, f);
  if (ws == 5) return dispatch_wc<SM, 5, OTF>(wc, f);
  return dispatch_wc<SM, 8, OTF>(wc, f);
}

template <typename F>
auto dispatch(int sm, int wc, int ws, bool otf, const F& f) {
  if (otf) return sm == SM_LINEAR ? dispatch_ws<SM_LINEAR, true>(wc, ws, f) : dispatch_ws<SM_CUBIC, true>(wc, ws, f);
  return sm == SM_LINEAR ? dispatch_ws<SM_LINEAR, false>(wc, ws, f) : dispatch_ws<SM_CUBIC, false>(wc, ws, f);
}

struct OccupancyFn {
  template <int SM, int WC, int WS, bool OTF>
  int operator()() const {
    return batch_occupancy_t<SM, WC, WS, OTF>();
  }
};

struct LaunchFn {
  const BatchLaunch& B;
  int nblocks;
  hipStream_t stream;
  template <int SM, int WC, int WS, bool OTF>
  hipError_t operator()() const {
    return batch_launch_t<SM, WC, WS, OTF>(B, nblocks, stream);
  }
};

bool wc_ok(int wc) { return wc == 4 || wc == 5 || wc == 6 || wc == 8 || wc == 10 || wc == 12 || wc == 16; }

}  // namespace

struct ustbx_batch {
  int device;
  int count;
  int sm;
  bool otf;  // some plan computes its receive state on the fly
  int wc, ws;
  int nblocks;
  int64_t n_items;
  std::vector<const ustbx_plan*> plans;
  std::vector<RxAcq> h_acqs;
  char* dev;  // descriptors | items | queue offsets | counters
  RxAcq* d_acqs;
  int4* d_items;
  int* d_qoff;
  int* d_ctr;
  hipStream_t capture_stream;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

namespace {

int set_shape(ustbx_batch* b, int wc, int ws, ustbx_status* st) {
  if (!wc_ok(wc)) return set_status(st, USTBX_E_ARG, "waves per split must be 4, 5, 6, 8, 10, 12 or 16");
  if (ws != 4 && ws != 5 && ws != 8) return set_status(st, USTBX_E_ARG, "wave-splits per block must be 4, 5 or 8");
  int cus = 0;
  hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  if (e != hipSuccess) return hip_fail(st, e, "hipDeviceGetAttribute");
  const int occ = dispatch(b->sm, wc, ws, b->otf, OccupancyFn{});
  if (occ <= 0) return set_status(st, USTBX_E_GPU, "the batch kernel does not fit on a compute unit");
  b->wc = wc;
  b->ws = ws;
  b->nblocks = ((cus * occ + kQueues - 1) / kQueues) * kQueues;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  return ok(st);
}

hipError_t launch(const ustbx_batch* b, hipStream_t stream) {
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, stream, b->d_ctr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  BatchLaunch B{b->d_acqs, b->d_items, b->d_qoff, b->d_ctr};
  return dispatch(b->sm, b->wc, b->ws, b->otf, LaunchFn{B, b->nblocks, stream});
}

}  // namespace

extern "C" {

int ustbx_batch_create(const ustbx_plan* const* plans, int32_t count, ustbx_batch** batch, ustbx_status* st) {
  if (!batch) return set_status(st, USTBX_E_ARG, "NULL batch");
  *batch = nullptr;
  if (!plans || count <= 0) return set_status(st, USTBX_E_ARG, "a batch needs at least one plan");
  int device = 0;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return hip_fail(st, e, "hipGetDevice");
  int interp = -1, max_nw = 0;
  for (int i = 0; i < count; ++i) {
    const ustbx_plan* pl = plans[i];
    if (!pl) return set_status(st, USTBX_E_ARG, "NULL plan %d", i);
    if (pl->device != device) return set_status(st, USTBX_E_ARG, "plan %d belongs to another device", i);
    if (pl->rxtab_state != 1 || pl->G.dim != 3)
      return set_status(st, USTBX_E_UNSUPPORTED,
                        "plan %d has no dimension-both receive-table state (select variant 2 first; it needs "
                        "dimension both and the separable or RF phase)",
                        i);
    if (interp >= 0 && pl->G.interp != interp)
      return set_status(st, USTBX_E_ARG, "the plans of a batch share one interpolation");
    interp = pl->G.interp;
    max_nw = std::max(max_nw, pl->max_nw);
  }
  ustbx_batch* b = new (std::nothrow) ustbx_batch();
  if (!b) return set_status(st, USTBX_E_CAPACITY, "host allocation failed");
  b->device = device;
  b->count = count;
  b->sm = interp == USTBX_INTERP_CUBIC ? SM_CUBIC : SM_LINEAR;
  b->otf = false;
  for (int i = 0; i < count; ++i) b->otf = b->otf || plans[i]->G.rx_otf != 0;
  b->plans.assign(plans, plans + count);

  // queues: per acquisition (largest work first) the XCD shares of its tiles
  std::vector<int64_t> work(count, 0);
  for (int i = 0; i < count; ++i) {
    const ustbx_plan* pl = plans[i];
    work[i] = std::accumulate(pl->h_cost.begin(), pl->h_cost.end(), int64_t(0)) * pl->G.F;
  }
  std::vector<int> by_work(count);
  std::iota(by_work.begin(), by_work.end(), 0);
  std::stable_sort(by_work.begin(), by_work.end(), [&](int a, int c) { return work[a] > work[c]; });
  std::vector<std::vector<int4>> q(kQueues);
  for (int i : by_work) {
    const ustbx_plan* pl = plans[i];
    const int tiles = (int)pl->h_order.size();
    const int qq = tiles / kQueues, r = tiles % kQueues;
    for (int k = 0; k < kQueues; ++k) {
      const int s0 = k * qq + std::min(k, r), cnt = qq + (k < r ? 1 : 0);
      for (int f = 0; f < pl->G.F; ++f)
        for (int x = s0; x < s0 + cnt; ++x) {
          const int t = pl->h_order[x];
          q[k].push_back(make_int4(i, t, f, pl->h_bad[t] ? 1 : 0));
        }
    }
  }
  std::vector<int> qoff(kQueues + 1, 0);
  for (int k = 0; k < kQueues; ++k) qoff[k + 1] = qoff[k] + (int)q[k].size();
  b->n_items = qoff[kQueues];
  std::vector<int4> items;
  items.reserve(b->n_items);
  for (int k = 0; k < kQueues; ++k) items.insert(items.end(), q[k].begin(), q[k].end());

  b->h_acqs.resize(count);
  for (int i = 0; i < count; ++i) {
    b->h_acqs[i] = rx_acq(plans[i]->G, b->sm);
    b->h_acqs[i].data = nullptr;
    b->h_acqs[i].out = nullptr;
  }
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  const size_t b_acq = al(sizeof(RxAcq) * count), b_items = al(sizeof(int4) * items.size());
  const size_t b_qoff = al(sizeof(int) * (kQueues + 1)), b_ctr = al(sizeof(int) * kQueues);
  e = hipMalloc(&b->dev, b_acq + b_items + b_qoff + b_ctr);
  if (e == hipSuccess) {
    b->d_acqs = (RxAcq*)b->dev;
    b->d_items = (int4*)(b->dev + b_acq);
    b->d_qoff = (int*)(b->dev + b_acq + b_items);
    b->d_ctr = (int*)(b->dev + b_acq + b_items + b_qoff);
    e = hipMemcpy(b->d_acqs, b->h_acqs.data(), sizeof(RxAcq) * count, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && !items.empty())
    e = hipMemcpy(b->d_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_qoff, qoff.data(), sizeof(int) * (kQueues + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreate(&b->capture_stream);  // default priority
  if (e != hipSuccess) {
    (void)hipFree(b->dev);
    delete b;
    return hip_fail(st, e, "batch setup");
  }
  // default tile shape: 8 splits, waves per split covering (in one round)
  // the wave count below which half of the batch's work lies
  int nw_med = max_nw;
  {
    std::vector<int> by_nw(count);
    std::iota(by_nw.begin(), by_nw.end(), 0);
    std::stable_sort(by_nw.begin(), by_nw.end(), [&](int a, int c) { return plans[a]->max_nw < plans[c]->max_nw; });
    const int64_t total = std::accumulate(work.begin(), work.end(), int64_t(0));
    int64_t acc = 0;
    for (int i : by_nw) {
      acc += work[i];
      if (2 * acc >= total) {
        nw_med = plans[i]->max_nw;
        break;
      }
    }
  }
  int ws = 8, wc = 8;
  if (nw_med <= 32) {
    ws = 4;  // as the one-plan default: 4 splits of 8 waves
  } else {
    const int need = (nw_med + 7) / 8;
    wc = need <= 4 ? 4 : need <= 6 ? 6 : need <= 8 ? 8 : need <= 10 ? 10 : need <= 12 ? 12 : 16;
  }
  const int rc = set_shape(b, wc, ws, st);
  if (rc != USTBX_OK) {
    ustbx_batch_destroy(b);
    return rc;
  }
  *batch = b;
  return ok(st);
}

int ustbx_batch_set_io(ustbx_batch* b, const void* const* ch, void* const* out, ustbx_status* st) {
  if (!b) return set_status(st, USTBX_E_ARG, "NULL batch");
  if (!ch || !out) return set_status(st, USTBX_E_ARG, "NULL buffer arrays");
  for (int i = 0; i < b->count; ++i) {
    if (!ch[i] || !out[i]) return set_status(st, USTBX_E_ARG, "NULL buffer for acquisition %d", i);
    b->h_acqs[i].data = ch[i];
    b->h_acqs[i].out = reinterpret_cast<float2*>(out[i]);
  }
  const hipError_t e = hipMemcpy(b->d_acqs, b->h_acqs.data(), sizeof(RxAcq) * b->count, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(st, e, "batch descriptors");
  return ok(st);
}

int ustbx_batch_set_tiles(ustbx_batch* b, int32_t wave_chunk, int32_t split, ustbx_status* st) {
  if (!b) return set_status(st, USTBX_E_ARG, "NULL batch");
  return set_shape(b, wave_chunk ? wave_chunk : b->wc, split ? split : b->ws, st);
}

int ustbx_batch_launch(const ustbx_batch* b, void* stream, ustbx_status* st) {
  if (!b) return set_status(st, USTBX_E_ARG, "NULL batch");
  if (!b->h_acqs[0].data) return set_status(st, USTBX_E_ARG, "set the batch's buffers first");
  const hipError_t e = launch(b, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(st, e, "das_batch_kernel launch");
  return ok(st);
}

int ustbx_batch_replay(ustbx_batch* b, void* stream, ustbx_status* st) {
  if (!b) return set_status(st, USTBX_E_ARG, "NULL batch");
  if (!b->h_acqs[0].data) return set_status(st, USTBX_E_ARG, "set the batch's buffers first");
  hipError_t e = hipSuccess;
  if (!b->exec) {
    e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(st, e, "hipStreamBeginCapture");
    const hipError_t el = launch(b, b->capture_stream);
    e = hipStreamEndCapture(b->capture_stream, &b->graph);
    if (el != hipSuccess) e = el;
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->graph = nullptr;
      b->exec = nullptr;
      return hip_fail(st, e, "batch graph capture");
    }
  }
  e = hipGraphLaunch(b->exec, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(st, e, "hipGraphLaunch");
  return ok(st);
}

int32_t ustbx_batch_shape(const ustbx_batch* b, int32_t* wave_chunk, int32_t* split, int32_t* blocks,
                          int64_t* items) {
  if (!b) return -1;
  if (wave_chunk) *wave_chunk = b->wc;
  if (split) *split = b->ws;
  if (blocks) *blocks = b->nblocks;
  if (items) *items = b->n_items;
  return b->count;
}

void ustbx_batch_destroy(ustbx_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  (void)hipFree(b->dev);
  delete b;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batcher: a new set of acquisitions (and slices of them) every submission,
// for serving loops that batch whatever frames are ready. Each acquisition's
// tiles are cut into the same 8 per-XCD shares as above; a submission takes
// positions [begin, end) of every share, so a large frame can be spread over
// several submissions (its tiles are independent: each is computed whole).
// ---------------------------------------------------------------------------
struct ustbx_batcher {
  int device;
  int sm;
  int wc, ws;
  int nblocks[2];  // the kernel without and with the on-the-fly form
  int max_acqs;
  int64_t max_items;
  char* dev;         // queue offsets | descriptors | items | counters
  char* staging[2];  // pinned host copies of the first three, used in turn
  size_t b_qoff, b_acq;
  hipEvent_t staged[2];  // that buffer's last uploads have left it
  bool pending[2];
  int cur;               // the buffer the next submission fills
  hipEvent_t launched;   // the last submission's kernel (it reads dev) is done
  bool has_launched;
};

extern "C" {

int32_t ustbx_plan_positions(const ustbx_plan* plan) {
  if (!plan || plan->rxtab_state != 1 || plan->G.dim != 3) return -1;
  const int tiles = (int)plan->h_order.size();
  return (tiles + kQueues - 1) / kQueues;
}

int ustbx_batcher_create(int32_t interp, int32_t max_acqs, int64_t max_items, int32_t wave_chunk, int32_t split,
                         ustbx_batcher** out, ustbx_status* st) {
  if (!out) return set_status(st, USTBX_E_ARG, "NULL batcher");
  *out = nullptr;
  if (interp != USTBX_INTERP_LINEAR && interp != USTBX_INTERP_CUBIC)
    return set_status(st, USTBX_E_ARG, "unknown interpolation");
  if (max_acqs <= 0 || max_items <= 0) return set_status(st, USTBX_E_ARG, "capacities must be positive");
  if (!wc_ok(wave_chunk)) return set_status(st, USTBX_E_ARG, "waves per split must be 4, 5, 6, 8, 10, 12 or 16");
  if (split != 4 && split != 5 && split != 8) return set_status(st, USTBX_E_ARG, "wave-splits must be 4, 5 or 8");
  ustbx_batcher* b = new (std::nothrow) ustbx_batcher();
  if (!b) return set_status(st, USTBX_E_CAPACITY, "host allocation failed");
  hipError_t e = hipGetDevice(&b->device);
  b->sm = interp == USTBX_INTERP_CUBIC ? SM_CUBIC : SM_LINEAR;
  b->wc = wave_chunk;
  b->ws = split;
  b->max_acqs = max_acqs;
  b->max_items = max_items;
  int cus = 0;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  for (int o = 0; o < 2; ++o) {
    const int occ = dispatch(b->sm, b->wc, b->ws, o == 1, OccupancyFn{});
    b->nblocks[o] = ((cus * std::max(occ, 1) + kQueues - 1) / kQueues) * kQueues;
  }
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  b->b_qoff = al(sizeof(int) * (kQueues + 1));
  b->b_acq = al(sizeof(RxAcq) * (size_t)max_acqs);
  const size_t b_items = al(sizeof(int4) * (size_t)max_items), b_ctr = al(sizeof(int) * kQueues);
  if (e == hipSuccess) e = hipMalloc(&b->dev, b->b_qoff + b->b_acq + b_items + b_ctr);
  for (int i = 0; i < 2; ++i) {
    if (e == hipSuccess) e = hipHostMalloc(&b->staging[i], b->b_qoff + b->b_acq + b_items, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->staged[i], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->launched, hipEventDisableTiming);
  if (e != hipSuccess) {
    ustbx_batcher_destroy(b);
    return hip_fail(st, e, "batcher setup");
  }
  *out = b;
  return ok(st);
}

int ustbx_batcher_submit(ustbx_batcher* b, int32_t count, const ustbx_plan* const* plans, const void* const* ch,
                         void* const* out, const int32_t* pos_begin, const int32_t* pos_end, void* stream,
                         ustbx_status* st) {
  if (!b) return set_status(st, USTBX_E_ARG, "NULL batcher");
  if (count <= 0) return ok(st);
  if (count > b->max_acqs) return set_status(st, USTBX_E_CAPACITY, "more acquisitions than the batcher holds");
  if (!plans || !ch || !out || !pos_begin || !pos_end) return set_status(st, USTBX_E_ARG, "NULL argument array");
  hipError_t e = hipSuccess;
  // two staging buffers in turn: this one's last uploads were queued two
  // submissions ago, so with at most one submission in flight the host does
  // not wait here
  const int c = b->cur;
  if (b->pending[c]) {
    e = hipEventSynchronize(b->staged[c]);  // this staging buffer is free again
    if (e != hipSuccess) return hip_fail(st, e, "batcher staging");
    b->pending[c] = false;
  }
  char* stg = b->staging[c];
  int* qoff = reinterpret_cast<int*>(stg);
  RxAcq* acqs = reinterpret_cast<RxAcq*>(stg + b->b_qoff);
  int4* items = reinterpret_cast<int4*>(stg + b->b_qoff + b->b_acq);
  // count first, so that capacity errors leave nothing half-written
  int64_t n_items = 0;
  for (int i = 0; i < count; ++i) {
    const ustbx_plan* pl = plans[i];
    if (!pl || pl->rxtab_state != 1 || pl->G.dim != 3)
      return set_status(st, USTBX_E_UNSUPPORTED, "acquisition %d: plan without dimension-both receive-table state",
                        i);
    if (pl->device != b->device) return set_status(st, USTBX_E_ARG, "acquisition %d: plan of another device", i);
    if ((pl->G.interp == USTBX_INTERP_CUBIC) != (b->sm == SM_CUBIC))
      return set_status(st, USTBX_E_ARG, "acquisition %d: interpolation differs from the batcher's", i);
    if (!ch[i] || !out[i]) return set_status(st, USTBX_E_ARG, "acquisition %d: NULL buffer", i);
    const int tiles = (int)pl->h_order.size(), q = tiles / kQueues, r = tiles % kQueues;
    const int pb = std::max(0, pos_begin[i]);
    for (int k = 0; k < kQueues; ++k) {
      const int cnt = q + (k < r ? 1 : 0);
      const int pe = std::min(cnt, pos_end[i]);
      if (pe > pb) n_items += (int64_t)(pe - pb) * pl->G.F;
    }
  }
  if (n_items > b->max_items) return set_status(st, USTBX_E_CAPACITY, "more work items than the batcher holds");
  int64_t at = 0;
  for (int k = 0; k < kQueues; ++k) {
    qoff[k] = (int)at;
    for (int i = 0; i < count; ++i) {
      const ustbx_plan* pl = plans[i];
      const int tiles = (int)pl->h_order.size(), q = tiles / kQueues, r = tiles % kQueues;
      const int s0 = k * q + std::min(k, r), cnt = q + (k < r ? 1 : 0);
      const int pb = std::max(0, pos_begin[i]), pe = std::min(cnt, pos_end[i]);
      for (int f = 0; f < pl->G.F; ++f)
        for (int x = pb; x < pe; ++x) {
          const int t = pl->h_order[s0 + x];
          items[at++] = make_int4(i, t, f, pl->h_bad[t] ? 1 : 0);
        }
    }
  }
  qoff[kQueues] = (int)at;
  for (int i = 0; i < count; ++i) {
    acqs[i] = rx_acq(plans[i]->G, b->sm);
    acqs[i].data = ch[i];
    acqs[i].out = reinterpret_cast<float2*>(out[i]);
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // the device buffers are shared by all submissions: on another stream than
  // the previous one, the upload must wait for that submission's kernel
  if (b->has_launched) {
    e = hipStreamWaitEvent(s, b->launched, 0);
    if (e != hipSuccess) return hip_fail(st, e, "batcher stream order");
  }
  e = hipMemcpyAsync(b->dev, stg, b->b_qoff + sizeof(RxAcq) * count, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && at > 0)
    e = hipMemcpyAsync(b->dev + b->b_qoff + b->b_acq, items, sizeof(int4) * at, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipEventRecord(b->staged[c], s);
  if (e != hipSuccess) return hip_fail(st, e, "batcher upload");
  b->pending[c] = true;
  b->cur = 1 - c;
  if (at == 0) {
    e = hipEventRecord(b->launched, s);
    if (e != hipSuccess) return hip_fail(st, e, "batcher stream order");
    b->has_launched = true;
    return ok(st);
  }
  int* d_ctr = reinterpret_cast<int*>(b->dev + b->b_qoff + b->b_acq + ((sizeof(int4) * b->max_items + 255) & ~size_t(255)));
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, d_ctr);
  e = hipGetLastError();
  if (e == hipSuccess) {
    BatchLaunch B{reinterpret_cast<const RxAcq*>(b->dev + b->b_qoff),
                  reinterpret_cast<const int4*>(b->dev + b->b_qoff + b->b_acq), reinterpret_cast<const int*>(b->dev),
                  d_ctr};
    bool otf = false;
    for (int i = 0; i < count; ++i) otf = otf || plans[i]->G.rx_otf != 0;
    e = dispatch(b->sm, b->wc, b->ws, otf, LaunchFn{B, b->nblocks[otf ? 1 : 0], s});
  }
  if (e == hipSuccess) e = hipEventRecord(b->launched, s);
  if (e != hipSuccess) return hip_fail(st, e, "das_batch_kernel launch");
  b->has_launched = true;
  return ok(st);
}

void ustbx_batcher_destroy(ustbx_batcher* b) {
  if (!b) return;
  if (b->launched) {
    if (b->has_launched) (void)hipEventSynchronize(b->launched);
    (void)hipEventDestroy(b->launched);
  }
  for (int i = 0; i < 2; ++i) {
    if (b->pending[i]) (void)hipEventSynchronize(b->staged[i]);
    if (b->staged[i]) (void)hipEventDestroy(b->staged[i]);
    if (b->staging[i]) (void)hipHostFree(b->staging[i]);
  }
  (void)hipFree(b->dev);
  delete b;
}

}  // extern "C"
