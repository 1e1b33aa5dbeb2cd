// das_unified.hip — the unified (Nguyen-Prager) transmit delay model of
// focused waves, tools.calculate_unified_delay_model
// (+tools/calculate_unified_delay_model.m:14-100) as das.m calls it
// (das.m:133-152), evaluated once per plan into the plan's [W][P] transmit
// delay matrix.
//
// One block per (wave, scan line) of a 2-D linear or sector scan (pixels
// depth-fastest, so a line is N_depth consecutive pixels). Per line:
//  1. the mask is das.m's boxcar transmit apodization with no minimum
//     aperture and the transmit apodization's f-number (das.m:136-149);
//  2. the quick fix (lines 34-37): the largest masked delay of the bottom row
//     over all lines stands in for the bottom pixel of lines whose bottom
//     pixel is outside the mask, and the bottom row counts as masked;
//  3. idx_a = the first pixel with the largest negative masked delay, idx_b
//     = the first with the smallest positive one (lines 52-61);
//  4. outside the mask the delay is |z_a - z| / |z_a - z_b| * d(idx_b) +
//     |z_b - z| / |z_b - z_a| * d(idx_a), each ratio formed as the reference
//     does (lines 66-72); inside it is the spherical delay; infinities are
//     zeroed and the source distance is added (lines 93-96).
// Because the bottom row is set before the per-line loop, the reference's two
// elseif branches (lines 74-90) cannot run and have no counterpart here.
// FP64 throughout with contraction off, rounded to FP32 after das.m's
// / sound_speed - delay (das.m:203-206).
#include "das_common.h"
#include "das_launch.h"

namespace ustbx {

namespace {

constexpr int kUB = 256;

// (value, index) over the block, keeping the first index among equal values
// (MATLAB's max/min); every thread calls it.
template <bool MAX>
__device__ void block_arg(double& v, int& i, double* sv, int* si) {
  const int t = threadIdx.x;
  sv[t] = v;
  si[t] = i;
  __syncthreads();
  for (int s = kUB / 2; s > 0; s >>= 1) {
    if (t < s) {
      const double a = sv[t], b = sv[t + s];
      const int ia = si[t], ib = si[t + s];
      const bool take = MAX ? (b > a || (b == a && ib < ia)) : (b < a || (b == a && ib < ia));
      if (take) {
        sv[t] = b;
        si[t] = ib;
      }
    }
    __syncthreads();
  }
  v = sv[0];
  i = si[0];
  __syncthreads();
}

// (-1)^(z < src.z) * |src - x| (das.m:120)
__device__ __forceinline__ double signed_distance(const DevWave& w, double x, double y, double z) {
#pragma clang fp contract(off)
  const double dx = w.sx - x, dy = w.sy - y, dz = w.sz - z;
  const double sgn = (z < w.sz) ? -1.0 : 1.0;
  return sgn * sqrt((dx * dx + dy * dy) + dz * dz);
}

// das.m's boxcar mask (NaN counts as outside, das.m:150)
__device__ __forceinline__ bool in_mask(const DevWave& w, const DevTxApod& mk, double x, double y, double z) {
  const float a = tx_apod_f32(w, mk, x, y, z, 0, 0, 0);
  return a == a && a != 0.0f;
}

__global__ __launch_bounds__(kUB) void geom_unified_tx_kernel(GeomLaunch G, DevTxApod mk,
                                                              float* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double sv[kUB];
  __shared__ int si[kUB];
  const int line = blockIdx.x;
  const int wi = blockIdx.y;
  DevWave w = G.waves[wi];
  if (w.txk != TXK_UNIFIED) return;  // the whole block
  w.apk = TXAP_INCIDENCE;
  const int nd = G.scan.n0, nl = G.scan.n1;
  double x, y, z, rr;

  // the quick fix's value: max over lines of tx_delay(end,:) .* mask(end,:)
  double maxval = -INFINITY;
  int unused = 0;
  for (int l = threadIdx.x; l < nl; l += kUB) {
    pixel_xyz(G.scan, (nd - 1) + nd * l, x, y, z, rr);
    const double d = signed_distance(w, x, y, z);
    const double v = in_mask(w, mk, x, y, z) ? d : d * 0.0;
    if (v > maxval) maxval = v;
  }
  block_arg<true>(maxval, unused, sv, si);

  // idx_a / idx_b of this line
  const int p0 = nd * line;
  double va = -INFINITY, vb = INFINITY;
  int ia = 0x7FFFFFFF, ib = 0x7FFFFFFF;
  for (int iz = threadIdx.x; iz < nd; iz += kUB) {
    pixel_xyz(G.scan, p0 + iz, x, y, z, rr);
    const double d = signed_distance(w, x, y, z);
    const bool m = in_mask(w, mk, x, y, z);
    double r = m ? d : d * 0.0;
    if (iz == nd - 1 && !m) r = maxval;
    const double pa = r < 0.0 ? r : -INFINITY;
    const double pb = r > 0.0 ? r : INFINITY;
    if (pa > va || (pa == va && iz < ia)) {
      va = pa;
      ia = iz;
    }
    if (pb < vb || (pb == vb && iz < ib)) {
      vb = pb;
      ib = iz;
    }
  }
  block_arg<true>(va, ia, sv, si);
  block_arg<false>(vb, ib, sv, si);

  double xa, ya, za, xb, yb, zb;
  pixel_xyz(G.scan, p0 + ia, xa, ya, za, rr);
  pixel_xyz(G.scan, p0 + ib, xb, yb, zb, rr);
  const double da = signed_distance(w, xa, ya, za);
  const double db = signed_distance(w, xb, yb, zb);
  const double ex = xb - xa, ez = zb - za;
  const double nrm = sqrt(ex * ex + ez * ez);  // norm(pos_b - pos_a)
  const double w1b = sqrt((za - zb) * (za - zb)) / nrm;
  const double w2a = sqrt((zb - za) * (zb - za)) / nrm;
  for (int iz = threadIdx.x; iz < nd; iz += kUB) {
    pixel_xyz(G.scan, p0 + iz, x, y, z, rr);
    const double d = signed_distance(w, x, y, z);
    double v;
    if (iz == nd - 1 || in_mask(w, mk, x, y, z)) {
      v = d;
    } else {
      const double w1 = (sqrt((za - z) * (za - z)) / nrm) / w1b;
      const double w2 = (sqrt((zb - z) * (zb - z)) / nrm) / w2a;
      v = w1 * db + w2 * da;
    }
    if (isinf(v)) v = 0.0;
    v = v + w.sdist;
    out[(size_t)wi * G.P + p0 + iz] = (float)(v / w.c - w.delay);
  }
}

}  // namespace

// grid = (scan lines, waves); waves of another kind return at once.
hipError_t launch_geom_unified_tx(const GeomLaunch& G, const DevTxApod& mask_apod, float* out,
                                  hipStream_t stream) {
  if (G.scan.n1 <= 0 || G.W <= 0) return hipSuccess;
  hipLaunchKernelGGL(geom_unified_tx_kernel, dim3((unsigned)G.scan.n1, (unsigned)G.W), dim3(kUB), 0, stream, G,
                     mask_apod, out);
  return hipGetLastError();
}

}  // namespace ustbx
