// das_common.h — device helpers shared by the gfx950 delay-and-sum kernels.
//
// Two kinds of arithmetic live here:
//  * das_c term arithmetic (+mex/source/das_c.cpp:372-400), evaluated in the
//    same order with FMA contraction switched off, so a term computed here is
//    the term das_c computes (up to the last ulp of sin/cos when fd != 0);
//  * FP64 geometry (+midprocess/das.m:103-206, +uff/apodization.m:172-555),
//    evaluated in the reference's operation order and rounded to FP32 at the
//    point where das.m applies single(): the device-side delays/apodization
//    equal the matrices das.m hands to the MEX.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ustbx {

constexpr int kBlock = 256;  // threads per block (4 wave64)

// ---------------------------------------------------------------------------
// Sample window rules
// ---------------------------------------------------------------------------
enum { EDGE_DASC = 0, EDGE_PYTHON = 1 };
enum { INTERP_LINEAR = 0, INTERP_CUBIC = 1 };

// One (pixel, channel, wave) term: interpolation index, fraction and the
// complex weight (apodization x phase rotation).
struct Term {
  int n0;      // first sample index (linear: n0, n0+1; cubic: n0-1 .. n0+2)
  float a, b;  // 1-b, b  (b = fractional sample)
  float wr, wi;
  bool valid;
};

// Fractional sample index d -> (n0, b, valid) following the reference rule.
// das_c: n0 = int(d) (truncation), b = d - floor(d), valid iff 0 < n0 < L-1
// (das_c.cpp:375-381). For finite d, "0 < trunc(d) < L-1" is "1 <= d < L-1",
// which is evaluated without the out-of-range float->int conversion.
// Python: n0 = floor(d), valid iff 0 <= n0 < L-1 (python das.py:233-236).
// Cubic (no reference counterpart): n0 = floor(d), valid iff 1 <= n0 <= L-3.
template <int INTERP>
__device__ __forceinline__ void sample_index(float denay, int L, int edge,
                                             int& n0, float& b, bool& valid) {
#pragma clang fp contract(off)
  const float n0f = floorf(denay);
  b = denay - n0f;
  if (INTERP == INTERP_LINEAR) {
    if (edge == EDGE_DASC)
      valid = (denay >= 1.0f) && (denay < (float)(L - 1));
    else
      valid = (denay >= 0.0f) && (denay < (float)(L - 1));
  } else {
    valid = (denay >= 1.0f) && (denay < (float)(L - 2));
  }
  n0 = valid ? (int)n0f : 0;
}

// das_c term: c_delay = tx + rx; c_apo = ta*ra; denay = (c_delay - t0)*Fs;
// weight = exp(i*wd*c_delay)*c_apo (das_c.cpp:372-387).
template <int INTERP, bool IQ>
__device__ __forceinline__ Term dasc_term(float td, float rd, float ta, float ra,
                                          float t0, float fs, float wd, int L,
                                          int edge) {
#pragma clang fp contract(off)
  Term t;
  const float c_delay = td + rd;
  const float c_apo = ta * ra;
  const float denay = (c_delay - t0) * fs;
  sample_index<INTERP>(denay, L, edge, t.n0, t.b, t.valid);
  t.a = 1.0f - t.b;
  if (IQ) {
    const float phase = wd * c_delay;
    float s, c;
    sincosf(phase, &s, &c);
    t.wr = c * c_apo;
    t.wi = s * c_apo;
  } else {
    // exp(i*0) = (1, 0): weight = (c_apo, 0)
    t.wr = c_apo;
    t.wi = 0.0f;
  }
  return t;
}

// exp(i*wd*t) for FP32 wd and t with the argument formed exactly: the
// product of two 24-bit mantissas fits FP64's 53 bits, the reduction by 2*pi
// is done in FP64 (|k| < 2^20 keeps its error below 1e-9 rad), and sin/cos are
// taken of the reduced FP32 argument in [-pi, pi].
__device__ __forceinline__ void sincos_wd(float wd, float t, float& s, float& c) {
#pragma clang fp contract(off)
  const double ph = (double)wd * (double)t;
  const double k = rint(ph * 0.15915494309189535);  // 1/(2*pi)
  const float r = (float)(ph - k * 6.283185307179586);
  sincosf(r, &s, &c);
}

// Same exact argument, reduced to revolutions in FP64; sin/cos from the
// hardware transcendental unit (v_sin_f32 / v_cos_f32 take revolutions).
__device__ __forceinline__ void sincos_wd_hw(float wd, float t, float& s, float& c) {
#pragma clang fp contract(off)
  const double u = ((double)wd * (double)t) * 0.15915494309189535;
  const float r = (float)(u - rint(u));
  s = __builtin_amdgcn_sinf(r);
  c = __builtin_amdgcn_cosf(r);
}

typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

// Raw buffer resource over a byte range (< 2 GiB): gathers then take an SGPR
// base plus a 32-bit VGPR byte offset, range-checked by the hardware (reads
// past num_records return zero).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, 0x00020000);
}

__device__ __forceinline__ float4 buf_load16(__amdgpu_buffer_rsrc_t r, int byte_off) {
  const f4v v = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, 0));
  return make_float4(v.x, v.y, v.z, v.w);
}

// Minimum (MAX = false) or maximum of v over the 64 lanes of a wave, every
// lane active: an inclusive scan within each row of 16 lanes (DPP row_shr 1,
// 2, 4, 8), then row_bcast15 and row_bcast31 carry the row results up, which
// leaves the wave's result in lane 63; returned wave-uniform.
template <bool MAX>
__device__ __forceinline__ unsigned wave_reduce_u32(unsigned v) {
  constexpr int kId = MAX ? 0 : -1;  // identity: 0 or 0xFFFFFFFF
  auto step = [&](unsigned t) { v = MAX ? (v > t ? v : t) : (v < t ? v : t); };
  step((unsigned)__builtin_amdgcn_update_dpp(kId, (int)v, 0x111, 0xf, 0xf, false));  // row_shr:1
  step((unsigned)__builtin_amdgcn_update_dpp(kId, (int)v, 0x112, 0xf, 0xf, false));  // row_shr:2
  step((unsigned)__builtin_amdgcn_update_dpp(kId, (int)v, 0x114, 0xf, 0xf, false));  // row_shr:4
  step((unsigned)__builtin_amdgcn_update_dpp(kId, (int)v, 0x118, 0xf, 0xf, false));  // row_shr:8
  step((unsigned)__builtin_amdgcn_update_dpp(kId, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  step((unsigned)__builtin_amdgcn_update_dpp(kId, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// x0 = X[n0], x1 = X[n0+1] as one 16-byte load (8-byte aligned).
__device__ __forceinline__ float4 load_pair(const float2* __restrict__ trace, int n0) {
  const f4a8 v = *reinterpret_cast<const f4a8*>(trace + n0);
  return make_float4(v.x, v.y, v.z, v.w);
}

// das_c's 4-term complex dot (das_c.cpp:389-399), summed from zero in order.
__device__ __forceinline__ float2 dasc_contrib(const Term& t, float4 x) {
#pragma clang fp contract(off)
  const float ar = t.a * t.wr, ai = t.a * t.wi, br = t.b * t.wr, bi = t.b * t.wi;
  float re = ar * x.x;
  re = re + (-ai) * x.y;
  re = re + br * x.z;
  re = re + (-bi) * x.w;
  float im = ai * x.x;
  im = im + ar * x.y;
  im = im + bi * x.z;
  im = im + br * x.w;
  return make_float2(re, im);
}

// Keys cubic convolution kernel, a = -1/2 (Catmull-Rom). Weights for samples
// n0-1, n0, n0+1, n0+2 at fractional offset b in [0, 1). Reproduces
// polynomials of degree <= 2 exactly.
__device__ __forceinline__ void keys_weights(float b, float& c0, float& c1,
                                             float& c2, float& c3) {
  const float b2 = b * b, b3 = b2 * b;
  c0 = -0.5f * b3 + b2 - 0.5f * b;
  c1 = 1.5f * b3 - 2.5f * b2 + 1.0f;
  c2 = -1.5f * b3 + 2.0f * b2 + 0.5f * b;
  c3 = 0.5f * b3 - 0.5f * b2;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// a += b without contraction (das_c's `*currpixPtr += complex(re, im)`)
__device__ __forceinline__ void cadd(float2& a, float2 b) {
#pragma clang fp contract(off)
  a.x = a.x + b.x;
  a.y = a.y + b.y;
}

// ---------------------------------------------------------------------------
// Windows (apodization.m:172-186), FP64, including MATLAB's NaN behaviour
// (a logical mask times NaN is NaN).
// ---------------------------------------------------------------------------
enum {
  WIN_NONE = 0, WIN_BOXCAR = 1, WIN_HANNING = 2, WIN_HAMMING = 3,
  WIN_TUKEY25 = 4, WIN_TUKEY50 = 5, WIN_TUKEY75 = 6, WIN_STA = 7,
  WIN_SCANLINE = 8, WIN_TRIANGLE = 9
};

// cos(x) in FP64 for the bounded arguments of the window functions
// (|x| <= 2*pi; NaN and +-inf give NaN, as cos does). Quadrant reduction with
// pi/2 split in two doubles (hi = nearest double to pi/2, lo = remainder),
// then the Taylor series of sin and cos on |r| <= pi/4 up to the r^17 / r^16
// terms (first omitted term < 3e-18).
__host__ __device__ __forceinline__ double cos_bounded(double x) {
#ifdef __HIP_DEVICE_COMPILE__
  const double q = rint(x * 0.6366197723675814);  // 2/pi
  double r = fma(-q, 1.5707963267948966, x);
  r = fma(-q, 6.123233995736766e-17, r);
  const double r2 = r * r;
  // cos: sum (-1)^k r^2k / (2k)!, k = 0..8
  double c = 4.779477332387385e-14;               //  1/16!
  c = fma(c, r2, -1.1470745597729725e-11);        // -1/14!
  c = fma(c, r2, 2.08767569878681e-09);           //  1/12!
  c = fma(c, r2, -2.755731922398589e-07);         // -1/10!
  c = fma(c, r2, 2.48015873015873e-05);           //  1/8!
  c = fma(c, r2, -0.001388888888888889);          // -1/6!
  c = fma(c, r2, 0.041666666666666664);           //  1/4!
  c = fma(c, r2, -0.5);                           // -1/2!
  c = fma(c, r2, 1.0);
  // sin: sum (-1)^k r^(2k+1) / (2k+1)!, k = 0..8
  double s = 2.8114572543455206e-15;              //  1/17!
  s = fma(s, r2, -7.647163731819816e-13);         // -1/15!
  s = fma(s, r2, 1.6059043836821613e-10);         //  1/13!
  s = fma(s, r2, -2.505210838544172e-08);         // -1/11!
  s = fma(s, r2, 2.7557319223985893e-06);         //  1/9!
  s = fma(s, r2, -0.0001984126984126984);         // -1/7!
  s = fma(s, r2, 0.008333333333333333);           //  1/5!
  s = fma(s, r2, -0.16666666666666666);           // -1/3!
  s = r * fma(s, r2, 1.0);
  const int k = (int)fmod(fabs(q), 4.0);          // cos is even in q
  double v = (k == 0) ? c : (k == 1) ? -s : (k == 2) ? -c : s;
  if (q < 0.0 && (k == 1 || k == 3)) v = -v;      // sin term flips with the sign of q
  return (x - x == 0.0) ? v : x - x;              // non-finite x: NaN
#else
  return cos(x);
#endif
}

__host__ __device__ __forceinline__ double win_tukey(double r, double roll) {
#pragma clang fp contract(off)
  const double th = 0.5 * (1.0 - roll);
  const double flat = (r <= th) ? 1.0 : 0.0;
  const double taper = (((r > th) ? 1.0 : 0.0) * ((r < 0.5) ? 1.0 : 0.0)) * 0.5;
  // cos is only needed where the taper mask is set; elsewhere the product is
  // 0*finite = 0 -- except for non-finite r, where MATLAB's 0*NaN = NaN.
  double cv;
  if (taper != 0.0 || !isfinite(r))
    cv = cos_bounded(((2.0 * M_PI) / roll) * ((r - roll / 2.0) - 0.5));
  else
    cv = 0.0;
  return flat + taper * (1.0 + cv);
}

__host__ __device__ __forceinline__ double win_eval(int win, double r) {
#pragma clang fp contract(off)
  const double m = (r <= 0.5) ? 1.0 : 0.0;
  switch (win) {
    case WIN_BOXCAR:
      return m;
    case WIN_HANNING: {
      const double cv = (m != 0.0 || !isfinite(r)) ? cos_bounded((2.0 * M_PI) * r) : 0.0;
      return m * (0.5 + 0.5 * cv);
    }
    case WIN_HAMMING: {
      const double cv = (m != 0.0 || !isfinite(r)) ? cos_bounded((2.0 * M_PI) * r) : 0.0;
      return m * (0.53836 + 0.46164 * cv);
    }
    case WIN_TUKEY25: return win_tukey(r, 0.25);
    case WIN_TUKEY50: return win_tukey(r, 0.50);
    case WIN_TUKEY75: return win_tukey(r, 0.75);
    case WIN_TRIANGLE: return m * (1.0 - 2.0 * r);
    default: return 1.0;
  }
}

// MATLAB sign(): -1, 0 or 1 (NaN -> NaN).
__host__ __device__ __forceinline__ double msign(double v) {
  return v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : (v == 0.0 ? 0.0 : v));
}

// Minimum/maximum aperture clamp of incidence_aperture / incidence_wave
// (apodization.m:453-462, 534-543).
__host__ __device__ __forceinline__ double aperture_clamp(double z_dist, double min_af,
                                                 double max_af) {
#pragma clang fp contract(off)
  double zd = z_dist;
  if (fabs(z_dist) <= min_af) zd = msign(zd) * min_af;
  if (fabs(z_dist) >= max_af) zd = msign(zd) * max_af;
  return zd;
}

// ---------------------------------------------------------------------------
// Device geometry descriptors
// ---------------------------------------------------------------------------
enum { SCAN_LINEAR = 0, SCAN_SECTOR = 1, SCAN_XYZ = 2 };

struct DevScan {
  int kind, n0, n1, n2;
  const double* a0;
  const double* a1;
  const double* a2;
  double ox, oy, oz;
  int ref_kind;
};

// Pixel coordinates in the reference's pixel order.
//  LINEAR: [Z, X, Y] = ndgrid(z_axis, x_axis, y_axis)      (linear_scan.m:71-81)
//  SECTOR: [rho, theta, phi] = ndgrid(depth, az, el);
//          [Z, Y, X] = sph2cart(phi, theta, rho) + origin   (sector_scan.m:90-112)
//          i.e. X = rho*sin(theta), Y = (rho*cos(theta))*sin(phi),
//               Z = (rho*cos(theta))*cos(phi)
__device__ __forceinline__ void pixel_xyz(const DevScan& s, int p, double& x,
                                          double& y, double& z, double& rref) {
#pragma clang fp contract(off)
  if (s.kind == SCAN_LINEAR) {
    const int i0 = p % s.n0;
    const int r = p / s.n0;
    const int i1 = r % s.n1;
    const int i2 = r / s.n1;
    z = s.a0[i0];
    x = s.a1[i1];
    y = s.a2[i2];
    rref = z;
  } else if (s.kind == SCAN_SECTOR) {
    const int i0 = p % s.n0;
    const int r = p / s.n0;
    const int i1 = r % s.n1;
    const int i2 = r / s.n1;
    const double rho = s.a0[i0];
    const double st = s.a1[2 * i1], ct = s.a1[2 * i1 + 1];
    const double sp = s.a2[2 * i2], cp = s.a2[2 * i2 + 1];
    const double rc = rho * ct;
    x = rho * st + s.ox;
    y = rc * sp + s.oy;
    z = rc * cp + s.oz;
    rref = rho;
  } else {
    x = s.a0[p];
    y = s.a1[p];
    z = s.a2[p];
    rref = z;
  }
  if (s.ref_kind == 0) rref = z;
}

// Pixel of lane `lane` (0..63) of 64-pixel tile `tile`. ps = 0: pixels
// tile*64 .. tile*64+63 in the scan order. ps > 0 (a structured scan whose
// fastest axis length n0 is a multiple of 2^ps and whose line count
// n1*n2 is a multiple of 64 >> ps, n1 itself too): the tile is a patch of
// 2^ps consecutive depths by 64 >> ps consecutive scan lines, lanes
// depth-fastest; tiles run down the depth first (n0 >> ps of them per
// group of lines), then across the line groups.
__device__ __forceinline__ int tile_pixel(int ps, int n0, int tile, int lane) {
  if (ps == 0) return tile * 64 + lane;
  const int ntz = n0 >> ps;
  const int grp = tile / ntz, tz = tile - grp * ntz;
  return ((tz << ps) | (lane & ((1 << ps) - 1))) + n0 * ((grp << (6 - ps)) + (lane >> ps));
}

// Receive apodization parameters (one uff.apodization with probe set).
enum { RXAP_ONES = 0, RXAP_WINDOW = 1, RXAP_VECTOR = 2, RXAP_MATRIX = 3 };
enum { RXGEO_FLAT = 0, RXGEO_ORIGIN = 1, RXGEO_CURV = 2 };

struct DevRxApod {
  int kind;      // RXAP_*
  int window;    // WIN_*
  int geo;       // RXGEO_FLAT: x_dist = focus - element; RXGEO_ORIGIN: centred on origin;
                 // RXGEO_CURV: curvilinear probe, azimuth around the origin
  int tilt;      // rotate_points needed
  double f0, f1;
  double min0, min1;  // minimum_aperture .* f_number
  double max0, max1;  // maximum_aperture .* f_number
  double ct0, st0, ct1, st1;  // cos/sin of tilt
  double ox, oy, oz;
  double win0;          // window(0): the y factor of a 1-D aperture (yd == 0)
  const float* vec;     // [N] (RXAP_VECTOR)
  const float* mat;     // [N][P] (RXAP_MATRIX)
};

// tools.rotate_points (rotate_points.m:11-25)
__device__ __forceinline__ void rotate_points(double& x, double& y, double& z,
                                              const DevRxApod& ap) {
#pragma clang fp contract(off)
  if (ap.tilt & 1) {
    const double xp = x * ap.ct0 - z * ap.st0;
    const double zp = x * ap.st0 + z * ap.ct0;
    x = xp;
    z = zp;
  }
  if (ap.tilt & 2) {
    const double yp = y * ap.ct1 - z * ap.st1;
    const double zp = y * ap.st1 + z * ap.ct1;
    y = yp;
    z = zp;
  }
}

// Reference phase exp(-1j*2*w0*r/c) (das.m:306-309). single = 0: the
// argument -(2*w0/c)*r in FP64, as python das.py:169-171 forms it; single = 1:
// das.m's single-precision argument fl32(fl32(-2*w0 * fl32(r)) / fl32(c)).
// The factor is the FP64 sine and cosine of the argument rounded to FP32.
__device__ __forceinline__ float2 ref_phase_factor(double w0_ref, float wd, float c32, int single, double r) {
  double arg;
  if (single) {
#pragma clang fp contract(off)
    const float a = (-2.0f * wd) * (float)r;
    arg = (double)(a / c32);
  } else {
    arg = -(w0_ref * r);
  }
  double sv, cv;
  sincos(arg, &sv, &cv);
  return make_float2((float)cv, (float)sv);
}

// receive_delay = single(sqrt(xm.^2+ym.^2+zm.^2)/c + lens_delay)  (das.m:103-106)
__device__ __forceinline__ float rx_delay_f32(double px, double py, double pz,
                                              double ex, double ey, double ez,
                                              double c, double lens) {
#pragma clang fp contract(off)
  const double xm = ex - px, ym = ey - py, zm = ez - pz;
  return (float)(sqrt((xm * xm + ym * ym) + zm * zm) / c + lens);
}

// rx apodization for (pixel, element) (apodization.m:323-367, 370-468).
__device__ __forceinline__ float rx_apod_f32(const DevRxApod& ap, double px,
                                             double py, double pz, double ex,
                                             double ey, double ez, int n, int p,
                                             int P) {
#pragma clang fp contract(off)
  if (ap.kind == RXAP_ONES) return 1.0f;
  if (ap.kind == RXAP_VECTOR) return ap.vec[n];
  if (ap.kind == RXAP_MATRIX) return ap.mat[(size_t)n * P + p];
  double xd, yd, zd;
  if (ap.geo == RXGEO_FLAT) {
    xd = px - ex;
    yd = py - ey;
    zd = pz - ez;
  } else if (ap.geo == RXGEO_CURV) {
    // curvilinear probes (apodization.m:388-398): the lateral distance is the
    // azimuth between pixel and element seen from the origin times origin.z,
    // the axial one the pixel's distance from the origin in (x, z) minus
    // origin.z
    const double el_az = atan2(ex - ap.ox, ez - ap.oz);
    const double px_az = atan2(px - ap.ox, pz - ap.oz);
    const double dx = px - ap.ox, dz = pz - ap.oz;
    xd = ap.oz * (px_az - el_az);
    yd = ap.oy - ey;
    zd = sqrt(dx * dx + dz * dz) - ap.oz;
  } else {
    const double dx = px - ap.ox, dy = py - ap.oy, dz = pz - ap.oz;
    xd = ex - ap.ox;
    yd = ey - ap.oy;
    zd = sqrt((dx * dx + dy * dy) + dz * dz);
  }
  rotate_points(xd, yd, zd, ap);
  const double zx = aperture_clamp(zd, ap.min0, ap.max0);
  const double zy = aperture_clamp(zd, ap.min1, ap.max1);
  const double rt = fabs(ap.f0 * (xd / zx));
  // yd == 0 with a nonzero finite zy gives ratio 0 exactly: window(0)
  const double wp = (yd == 0.0 && fabs(zy) > 0.0 && fabs(zy) < INFINITY)
                        ? ap.win0
                        : win_eval(ap.window, fabs(ap.f1 * (yd / zy)));
  return (float)(win_eval(ap.window, rt) * wp);
}

// ---------------------------------------------------------------------------
// Receive state computed on the fly (no [N][P] table): per (pixel, element)
// the delay das.m forms (FP64, rounded to FP32: rx_delay_f32, the table's
// value bit for bit) and the receive weight from the same formulas as
// rx_apod_f32 (aperture clamps and ratios in FP64, the pixel- or
// origin-centred aperture; no tilt), the window evaluated in FP32 with the
// hardware cosine. Weights agree with rx_apod_f32 to FP32 rounding.
// ---------------------------------------------------------------------------
struct RxOtfApod {
  int kind;    // RXAP_ONES or RXAP_WINDOW
  int window;  // WIN_*
  int geo;     // RXGEO_*
  float win0;  // window(0)
  double f0, f1;
  double min0, min1, max0, max1;  // minimum / maximum aperture .* f_number
  double ox, oy, oz;
};

// The window of ratio r in FP32 (apodization.m:172-186); v_cos_f32 takes
// revolutions. Non-finite r: 0 for the boxcar, NaN for the others (MATLAB's
// logical mask times NaN), as win_eval.
__device__ __forceinline__ float win_eval_f32(int win, float r) {
  if (!(r <= 3.402823466e38f)) return win == WIN_BOXCAR ? 0.0f : __builtin_nanf("");
  const bool m = r <= 0.5f;
  switch (win) {
    case WIN_BOXCAR: return m ? 1.0f : 0.0f;
    case WIN_HANNING: return m ? 0.5f + 0.5f * __builtin_amdgcn_cosf(r) : 0.0f;
    case WIN_HAMMING: return m ? 0.53836f + 0.46164f * __builtin_amdgcn_cosf(r) : 0.0f;
    case WIN_TUKEY25:
    case WIN_TUKEY50:
    case WIN_TUKEY75: {
      const float roll = win == WIN_TUKEY25 ? 0.25f : (win == WIN_TUKEY50 ? 0.5f : 0.75f);
      if (r <= 0.5f * (1.0f - roll)) return 1.0f;
      if (!(r < 0.5f)) return 0.0f;
      // 0.5 (1 + cos(2 pi / roll (r - roll/2 - 1/2)))
      return 0.5f * (1.0f + __builtin_amdgcn_cosf((r - (0.5f + 0.5f * roll)) / roll));
    }
    case WIN_TRIANGLE: return m ? 1.0f - 2.0f * r : 0.0f;
    default: return 1.0f;
  }
}

__device__ __forceinline__ float aperture_clamp_f32(float z, float min_af, float max_af) {
  float zd = z;
  const float sg = z > 0.0f ? 1.0f : (z < 0.0f ? -1.0f : (z == 0.0f ? 0.0f : z));  // MATLAB sign()
  if (fabsf(z) <= min_af) zd = sg * min_af;
  if (fabsf(z) >= max_af) zd = sg * max_af;
  return zd;
}

// |f * (x / z)|: in FP32 (the hardware reciprocal), except within 1e-5 of
// 1/2 for the windows that jump there (boxcar, Hamming), where the FP64
// quotient rx_apod_f32 forms decides the side (a ratio exactly on the edge
// occurs for aligned pixel and element grids).
__device__ __forceinline__ float otf_ratio(int window, double f, double x, double z) {
  const float r = fabsf((float)f * ((float)x * __builtin_amdgcn_rcpf((float)z)));
  if ((window == WIN_BOXCAR || window == WIN_HAMMING) && fabsf(r - 0.5f) < 1e-5f) {
#pragma clang fp contract(off)
    const double r64 = fabs(f * (x / z));
    return r64 <= 0.5 ? fminf(r, 0.5f) : fmaxf(r, nextafterf(0.5f, 1.0f));
  }
  return r;
}

// Receive weight of distances (xd, yd, zd) (apodization.m:415-467 without
// the tilt): the clamps in FP64, the ratios by otf_ratio, the window in FP32.
__device__ __forceinline__ float rx_apod_otf(const RxOtfApod& a, double xd, double yd, double zd) {
#pragma clang fp contract(off)
  if (a.kind == RXAP_ONES) return 1.0f;
  const double zx = aperture_clamp(zd, a.min0, a.max0);
  const double zy = aperture_clamp(zd, a.min1, a.max1);
  const float rt = otf_ratio(a.window, a.f0, xd, zx);
  const float wp = (yd == 0.0 && fabs(zy) > 0.0 && fabs(zy) < INFINITY) ? a.win0
                                                                         : win_eval_f32(a.window, otf_ratio(a.window, a.f1, yd, zy));
  return win_eval_f32(a.window, rt) * wp;
}

// Everything the on-the-fly receive state needs (device pointers).
struct RxOtf {
  DevScan scan;
  const double* ex;
  const double* ey;
  const double* ez;
  const float4* e32;  // [N] {ex, ey, ez, 0} rounded to FP32
  double inv_c, lens;
  float inv_c32, lens32;  // 1/c and the lens delay rounded to FP32
  float wd;     // das_c's wd (the separable receive factor)
  int compact;  // RF (wd == 0): the factor is (ra, 0), as the compact table
  RxOtfApod ap;
  float f0f, f1f, min0f, min1f, max0f, max1f;  // the aperture parameters rounded to FP32
  float oxf, oyf;
};

// A pixel as the on-the-fly entries read it: its coordinates rounded to
// FP32 and, for an origin-centred aperture, its distance from the origin
// (FP64, rounded).
struct OtfPixel {
  double px, py, pz;
  float x, y, z, zo;
};

__device__ __forceinline__ OtfPixel otf_pixel(const RxOtf& o, double px, double py, double pz) {
#pragma clang fp contract(off)
  OtfPixel q;
  q.px = px;
  q.py = py;
  q.pz = pz;
  q.x = (float)px;
  q.y = (float)py;
  q.z = (float)pz;
  q.zo = 0.0f;
  if (o.ap.geo != RXGEO_FLAT) {
    const double dx = px - o.ap.ox, dy = py - o.ap.oy, dz = pz - o.ap.oz;
    q.zo = (float)sqrt((dx * dx + dy * dy) + dz * dz);
  }
  return q;
}

// |f * (x / z)| in FP32 with the hardware reciprocal; near is set where the
// ratio is within 1e-5 of 1/2 for a window that jumps there (boxcar, Hamming),
// where FP32 rounding could put it on the other side of the edge than das.m's
// FP64 quotient.
__device__ __forceinline__ float otf_ratio32(int window, float f, float x, float z, bool& near) {
  const float r = fabsf(f * (x * __builtin_amdgcn_rcpf(z)));
  near = near || ((window == WIN_BOXCAR || window == WIN_HAMMING) && fabsf(r - 0.5f) < 1e-5f);
  return r;
}

// das.m's receive delay single(sqrt(xm^2+ym^2+zm^2)/c + lens) (das.m:103-106)
// without an FP64 square root or division: the FP64 sum of squares, its FP32
// square root, one Newton step in FP64 (relative error about 1e-14 against
// the correctly rounded root), times 1/c. The FP32 result equals das.m's
// except where its FP64 value lies within about 1e-14 (relative) of an FP32
// rounding boundary, where it may differ by one ulp.
__device__ __forceinline__ float otf_rx_delay(const RxOtf& o, const OtfPixel& q, int n) {
  const double xm = o.ex[n] - q.px, ym = o.ey[n] - q.py, zm = o.ez[n] - q.pz;
  const double S = fma(xm, xm, fma(ym, ym, zm * zm));
  const float s32 = __builtin_amdgcn_sqrtf((float)S);
  const double s0 = (double)s32;
  const double r = s32 > 0.0f ? fma(fma(-s0, s0, S), (double)(0.5f * __builtin_amdgcn_rcpf(s32)), s0) : 0.0;
  return (float)fma(r, o.inv_c, o.lens);
}

// {rd, ra, ra*cos(wd rd), ra*sin(wd rd)} of (pixel q, element n), the
// receive-table entry computed in place: the delay of otf_rx_delay and the
// receive weight of apodization.m:415-467 (no tilt) with FP32 coordinate
// differences, clamps, ratios and window; where a ratio lies within 1e-5 of a
// window's jump at 1/2 the weight is formed in FP64 instead (rx_apod_otf), so
// that it falls on das.m's side of the edge. The phase factor reduces wd*rd
// in FP64, as the table.
__device__ __forceinline__ float4 rx_otf_entry(const RxOtf& o, const OtfPixel& q, int n) {
  const float4 e = o.e32[n];
  const float xm = e.x - q.x, ym = e.y - q.y, zm = e.z - q.z;
  const float rd = otf_rx_delay(o, q, n);
  float ra = 1.0f;
  if (o.ap.kind != RXAP_ONES) {
    float xd, yd, zd;
    if (o.ap.geo == RXGEO_FLAT) {
      xd = -xm;
      yd = -ym;
      zd = -zm;
    } else {
      xd = e.x - o.oxf;
      yd = e.y - o.oyf;
      zd = q.zo;
    }
    bool near = false;
    const float zx = aperture_clamp_f32(zd, o.min0f, o.max0f);
    const float zy = aperture_clamp_f32(zd, o.min1f, o.max1f);
    const float rt = otf_ratio32(o.ap.window, o.f0f, xd, zx, near);
    float wp = o.ap.win0;
    if (!(yd == 0.0f && fabsf(zy) > 0.0f && fabsf(zy) < INFINITY))
      wp = win_eval_f32(o.ap.window, otf_ratio32(o.ap.window, o.f1f, yd, zy, near));
    ra = win_eval_f32(o.ap.window, rt) * wp;
    if (near) {
#pragma clang fp contract(off)
      const double px = q.px, py = q.py, pz = q.pz;
      const double ex = o.ex[n], ey = o.ey[n], ez = o.ez[n];
      double xd64, yd64, zd64;
      if (o.ap.geo == RXGEO_FLAT) {
        xd64 = px - ex;
        yd64 = py - ey;
        zd64 = pz - ez;
      } else {
        const double dx = px - o.ap.ox, dy = py - o.ap.oy, dz = pz - o.ap.oz;
        xd64 = ex - o.ap.ox;
        yd64 = ey - o.ap.oy;
        zd64 = sqrt((dx * dx + dy * dy) + dz * dz);
      }
      ra = rx_apod_otf(o.ap, xd64, yd64, zd64);
    }
  }
  if (o.compact) return make_float4(rd, ra, ra, ra * 0.0f);
  float sv, cv;
  sincos_wd_hw(o.wd, rd, sv, cv);
  return make_float4(rd, ra, ra * cv, ra * sv);
}

// Transmit side. Kinds of tx delay per wave (das.m:110-204).
enum {
  TXK_PLANE = 0,        // z*cos(az)*cos(el) + x*sin(az)*cos(el) + y*sin(el)
  TXK_PHOTO = 1,        // 0
  TXK_DIVERGING = 2,    // (-1)^(z<sz)*|s-x| - |s.distance|
  TXK_SPHERICAL = 3,    // (-1)^(z<sz)*|s-x| + s.distance
  TXK_HYBRID_LIN = 4,   // spherical outside +-pw_margin of focus depth, linear-scan plane patch
  TXK_HYBRID_SEC = 5,   // same with the sector-scan plane patch
  TXK_BLENDED_LIN = 6,
  TXK_BLENDED_SEC = 7,
  TXK_MATRIX = 8,       // explicit [W][P] matrix given by the caller
  TXK_UNIFIED = 9       // unified model, precomputed into the plan's [W][P] matrix
};
enum { TXAP_CONST = 0, TXAP_INCIDENCE = 1, TXAP_MATRIX = 2 };

struct DevWave {
  int txk;
  int apk;
  float apo_const;
  int _pad;
  double sx, sy, sz, sdist;
  double ca, sa, ce, se;
  double c, delay;
  double ox, oy, oz;
  double zu0, zu1, zu2, xu0, xu1, xu2, yu0, yu1, yu2;
  double os_az, os_el;  // atan2(OS_x, OS_z), atan2(OS_y, OS_z)
};

struct DevTxApod {
  int window;
  int _pad;
  double f0, f1, min0, min1, max0, max1;
  double gl0, gl1;  // 0.5 * grating_lobe_angle
  const float* mat; // [W][P] (TXAP_MATRIX)
};

__device__ __forceinline__ double plane_delay(const DevWave& w, double x,
                                              double y, double z) {
#pragma clang fp contract(off)
  return ((z * w.ca) * w.ce + (x * w.sa) * w.ce) + y * w.se;
}

// transmit_delay(:,w) before and after the /sound_speed - delay step (das.m:110-206).
__device__ __forceinline__ float tx_delay_f32(const DevWave& w, double x,
                                              double y, double z,
                                              double pw_margin, double bpow,
                                              const float* mat, int wi, int p,
                                              int P) {
#pragma clang fp contract(off)
  double v = 0.0;
  switch (w.txk) {
    case TXK_PLANE:
      v = plane_delay(w, x, y, z);
      break;
    case TXK_PHOTO:
      v = 0.0;
      break;
    case TXK_MATRIX:
    case TXK_UNIFIED:
      return mat[(size_t)wi * P + p];
    default: {
      const double dx = w.sx - x, dy = w.sy - y, dz = w.sz - z;
      const double sgn = (z < w.sz) ? -1.0 : 1.0;
      const double d = sgn * sqrt((dx * dx + dy * dy) + dz * dz);
      if (w.txk == TXK_DIVERGING) {
        v = d - fabs(w.sdist);
      } else if (w.txk == TXK_SPHERICAL) {
        v = d + w.sdist;
      } else {
        double pd;
        if (w.txk == TXK_HYBRID_LIN || w.txk == TXK_BLENDED_LIN)
          pd = sgn * sqrt(dz * dz) + w.sdist;
        else
          pd = plane_delay(w, x, y, z);
        if (w.txk == TXK_HYBRID_LIN || w.txk == TXK_HYBRID_SEC) {
          const bool mask = (z < (w.sz + pw_margin)) && (z > (w.sz - pw_margin));
          v = mask ? pd : d + w.sdist;
        } else {
          const double rr = sqrt((x * x + y * y) + z * z);
          const double nd = fmin(fabs(w.sdist - rr) / w.sdist, 1.0);
          const double np = pow(nd, bpow);
          v = (d + w.sdist) * np + pd * (1.0 - np);
        }
      }
    }
  }
  return (float)(v / w.c - w.delay);
}

// transmit apodization (apodization.m:249-320, 472-555).
__device__ __forceinline__ float tx_apod_f32(const DevWave& w, const DevTxApod& ap,
                                             double x, double y, double z,
                                             int wi, int p, int P) {
#pragma clang fp contract(off)
  if (w.apk == TXAP_CONST) return w.apo_const;
  if (w.apk == TXAP_MATRIX) return ap.mat[(size_t)wi * P + p];
  const double spx = x - w.sx, spy = y - w.sy, spz = z - w.sz;
  const double opx = x - w.ox, opy = y - w.oy, opz = z - w.oz;
  const double zd = (spx * w.zu0 + spy * w.zu1) + spz * w.zu2;
  const double xd = (spx * w.xu0 + spy * w.xu1) + spz * w.xu2;
  const double yd = (spx * w.yu0 + spy * w.yu1) + spz * w.yu2;
  double zx = aperture_clamp(zd, ap.min0, ap.max0);
  double zy = aperture_clamp(zd, ap.min1, ap.max1);
  const double eps = 2.220446049250313e-16;
  if (fabs(atan2(opx, opz) - w.os_az) > ap.gl0) zx = eps;
  if (fabs(atan2(opy, opz) - w.os_el) > ap.gl1) zy = eps;
  const double rt = fabs(ap.f0 * (xd / zx));
  const double rp = fabs(ap.f1 * (yd / zy));
  return (float)(win_eval(ap.window, rt) * win_eval(ap.window, rp));
}

}  // namespace ustbx
