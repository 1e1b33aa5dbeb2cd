// das_matrix.hip — gfx950 kernel for the das_c / das_cuda matrix contract.
//
// Inputs are the reference MEX arrays (+mex/source/das_c.cpp:11-29): channel
// data [F][W][N][L] complex64, tx apodization/delay [W][P], rx
// apodization/delay [N][P], all in HBM. One lane owns one pixel (pixels are
// the fastest index of every matrix, so each wave64 reads 256 contiguous
// bytes per matrix row and writes 512 contiguous bytes of output).
//
// A block is 256 lanes = TP pixels x SPLIT channel groups. SPLIT > 1 gives the
// chip enough waves when P is small (C1: 65,536 pixels = 1,024 wave64 at
// SPLIT = 1); partial sums of the groups are added in a fixed order through
// LDS, so results are deterministic. SPLIT = 1 reproduces das_c's summation
// order exactly (w outer, n inner, das_c.cpp:361-404).
//
// Per-term arithmetic is das_c's (das_common.h: dasc_term / dasc_contrib),
// with FMA contraction off, so for fd == 0 (no sin/cos) and SPLIT = 1 the
// output is bit-identical to das_c. PY instead follows the Python
// reference's DAS._beamform term by term (py_term / py_contrib).
#include "das_common.h"
#include "das_launch.h"

namespace ustbx {

struct MatArgs {
  const float2* __restrict__ ch;
  const float* __restrict__ tx_apo;
  const float* __restrict__ rx_apo;
  const float* __restrict__ tx_delay;
  const float* __restrict__ rx_delay;
  float2* __restrict__ out;
  int L, N, W, F, P;
  float fs, t0, wd;
  int edge;
  float dt, w0;  // PY: sample period, angular frequency
  // PY: per wave / channel, whether its weight column holds a nonzero (NaN
  // included): the reference skips only all-zero columns (das.py:221-225)
  const int* __restrict__ tx_any;
  const int* __restrict__ rx_any;
};

// The Python reference's term (python das.py:223-250) in NumPy's dtypes:
// delay = rx + tx and apo = rx_apo * tx_apo in float32; the index
// (delay - t0) / dt in float32 (the Python scalars cast to float32); floor and
// fraction exact (NumPy promotes float32 - int32 to float64); the
// interpolation x0 (1 - b) + x1 b in float64, stored as complex64; times apo;
// times exp(1j * w0 * delay) in complex64, whose argument is the float32
// product and whose cos/sin are taken in FP64 and rounded.
struct PyTerm {
  int n0;
  bool valid;
  double b;
  float apo, c, s;
};

template <bool IQ>
__device__ __forceinline__ PyTerm py_term(float td, float rd, float ta, float ra, float t0, float dt, float w0,
                                          int L) {
#pragma clang fp contract(off)
  PyTerm t;
  const float delay = rd + td;                  // das.py:224
  t.apo = ra * ta;                              // das.py:223
  const float idx = (delay - t0) / dt;          // das.py:227
  const float fl = floorf(idx);                 // das.py:228
  t.valid = fl >= 0.0f && fl < (float)(L - 1);  // das.py:231 (NaN: false)
  t.n0 = t.valid ? (int)fl : 0;
  t.b = (double)idx - (double)fl;               // das.py:229
  t.c = 1.0f;
  t.s = 0.0f;
  if (IQ) {
    double sd, cd;
    sincos((double)(w0 * delay), &sd, &cd);     // das.py:250
    t.c = (float)cd;
    t.s = (float)sd;
  }
  return t;
}

// apo * interpolated sample (zero outside the window, das.py:240-247), times
// the phase factor.
template <bool IQ>
__device__ __forceinline__ float2 py_contrib(const PyTerm& t, const float2* __restrict__ trace) {
#pragma clang fp contract(off)
  float ivr = 0.0f, ivi = 0.0f;
  if (t.valid) {
    const float4 x = load_pair(trace, t.n0);
    const double a = 1.0 - t.b;
    ivr = (float)((double)x.x * a + (double)x.z * t.b);
    ivi = (float)((double)x.y * a + (double)x.w * t.b);
  }
  const float vr = t.apo * ivr, vi = t.apo * ivi;
  if (!IQ) return make_float2(vr, vi);
  return make_float2(t.c * vr - t.s * vi, t.c * vi + t.s * vr);
}

// Gather and weight one term for one frame.
template <int INTERP>
__device__ __forceinline__ float2 mat_contrib(const Term& t,
                                              const float2* __restrict__ trace) {
  if (INTERP == INTERP_LINEAR) {
    return dasc_contrib(t, load_pair(trace, t.n0));
  } else {
    const float4 lo = load_pair(trace, t.n0 - 1);
    const float4 hi = load_pair(trace, t.n0 + 1);
    float c0, c1, c2, c3;
    keys_weights(t.b, c0, c1, c2, c3);
    const float sr = ((c0 * lo.x + c1 * lo.z) + c2 * hi.x) + c3 * hi.z;
    const float si = ((c0 * lo.y + c1 * lo.w) + c2 * hi.y) + c3 * hi.w;
    return make_float2(t.wr * sr - t.wi * si, t.wr * si + t.wi * sr);
  }
}

// Unit loop over channels [nb, ne) for fixed (pixel, wave), U channels per
// step: rx loads of the U channels are issued first, then the gathers, then
// the accumulation in channel order (das_c's order within the group).
template <int INTERP, bool IQ, bool PY, int FC, int U>
__device__ __forceinline__ void sum_channels(const MatArgs& A, int p, int w,
                                             float ta, float td, int nb, int ne,
                                             int f0, float2 (&acc)[FC]) {
  const size_t LN = (size_t)A.L * A.N;
  const size_t LNW = LN * A.W;
  for (int n = nb; n < ne; n += U) {
    float ra[U], rd[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int nn = n + u;
      ra[u] = 0.0f;
      rd[u] = 0.0f;
      if (nn < ne) {
        ra[u] = A.rx_apo[(size_t)nn * A.P + p];
        rd[u] = A.rx_delay[(size_t)nn * A.P + p];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (PY) {
        if (n + u < ne && A.rx_any[n + u]) {
          const PyTerm t = py_term<IQ>(td, rd[u], ta, ra[u], A.t0, A.dt, A.w0, A.L);
          {  // every pixel of a column with a nonzero weight (0 * NaN reaches the sum)
            const float2* trace = A.ch + (size_t)w * LN + (size_t)(n + u) * A.L;
#pragma unroll
            for (int f = 0; f < FC; ++f)
              if (f0 + f < A.F) cadd(acc[f], py_contrib<IQ>(t, trace + (size_t)(f0 + f) * LNW));
          }
        }
      } else if (ra[u] != 0.0f) {
        const Term t = dasc_term<INTERP, IQ>(td, rd[u], ta, ra[u], A.t0, A.fs,
                                             A.wd, A.L, A.edge);
        if (t.valid) {
          const float2* trace = A.ch + (size_t)w * LN + (size_t)(n + u) * A.L;
#pragma unroll
          for (int f = 0; f < FC; ++f) {
            if (f0 + f < A.F) {
              cadd(acc[f], mat_contrib<INTERP>(t, trace + (size_t)(f0 + f) * LNW));
            }
          }
        }
      }
    }
  }
}

// Deterministic cross-group reduction: group 0 returns
// ((g0 + g1) + g2) + ... ; every lane of the block must call it.
template <int SPLIT, int FC>
__device__ __forceinline__ void reduce_groups(float2 (&acc)[FC], int lp, int s,
                                              float2* red) {
  constexpr int TP = kBlock / SPLIT;
  if (SPLIT == 1) return;
#pragma unroll
  for (int f = 0; f < FC; ++f) red[(s * FC + f) * TP + lp] = acc[f];
  __syncthreads();
  if (s == 0) {
#pragma unroll
    for (int f = 0; f < FC; ++f) {
      float2 v = red[f * TP + lp];
      for (int g = 1; g < SPLIT; ++g) {
        cadd(v, red[(g * FC + f) * TP + lp]);
      }
      acc[f] = v;
    }
  }
  __syncthreads();
}

template <int MODE, int INTERP, bool IQ, bool PY, int SPLIT, int FC>
__global__ __launch_bounds__(kBlock) void das_matrix_kernel(MatArgs A) {
  constexpr int TP = kBlock / SPLIT;
  constexpr int U = (FC == 1) ? 4 : 1;
  __shared__ float2 red[(SPLIT > 1 && (MODE == 1 || MODE == 3)) ? SPLIT * FC * TP : 1];
  const int lp = threadIdx.x % TP;
  const int s = threadIdx.x / TP;
  const int p = blockIdx.x * TP + lp;
  const bool active = p < A.P;
  const int nper = (A.N + SPLIT - 1) / SPLIT;
  const int nb = min(A.N, s * nper);
  const int ne = min(A.N, nb + nper);
  const size_t PS = (size_t)A.P;

  for (int f0 = 0; f0 < A.F; f0 += FC) {
    if (MODE == 3) {  // BOTH: out[f][p]
      float2 acc[FC];
#pragma unroll
      for (int f = 0; f < FC; ++f) acc[f] = make_float2(0.0f, 0.0f);
      if (active) {
        for (int w = 0; w < A.W; ++w) {
          const float ta = A.tx_apo[(size_t)w * PS + p];
          if (PY ? A.tx_any[w] != 0 : ta != 0.0f) {
            const float td = A.tx_delay[(size_t)w * PS + p];
            sum_channels<INTERP, IQ, PY, FC, U>(A, p, w, ta, td, nb, ne, f0, acc);
          }
        }
      }
      reduce_groups<SPLIT, FC>(acc, lp, s, red);
      if (active && s == 0) {
#pragma unroll
        for (int f = 0; f < FC; ++f)
          if (f0 + f < A.F) A.out[(size_t)(f0 + f) * PS + p] = acc[f];
      }
    } else if (MODE == 1) {  // RX: sum over channels, out[f][w][p]
      for (int w = 0; w < A.W; ++w) {
        float2 acc[FC];
#pragma unroll
        for (int f = 0; f < FC; ++f) acc[f] = make_float2(0.0f, 0.0f);
        if (active) {
          const float ta = A.tx_apo[(size_t)w * PS + p];
          if (PY ? A.tx_any[w] != 0 : ta != 0.0f) {
            const float td = A.tx_delay[(size_t)w * PS + p];
            sum_channels<INTERP, IQ, PY, FC, U>(A, p, w, ta, td, nb, ne, f0, acc);
          }
        }
        reduce_groups<SPLIT, FC>(acc, lp, s, red);
        if (active && s == 0) {
#pragma unroll
          for (int f = 0; f < FC; ++f)
            if (f0 + f < A.F)
              A.out[((size_t)(f0 + f) * A.W + w) * PS + p] = acc[f];
        }
      }
    } else if (MODE == 2) {  // TX: sum over waves, out[f][n][p]; each (p, n) owned by one lane
      if (active) {
        const size_t LN = (size_t)A.L * A.N;
        const size_t LNW = LN * A.W;
        for (int n = nb; n < ne; ++n) {
          float2 acc[FC];
#pragma unroll
          for (int f = 0; f < FC; ++f) acc[f] = make_float2(0.0f, 0.0f);
          const float ra = A.rx_apo[(size_t)n * PS + p];
          if (PY) {
            const float rd = A.rx_delay[(size_t)n * PS + p];
            const float2* trace = A.ch + (size_t)n * A.L;
            for (int w = 0; w < A.W; ++w) {
              const PyTerm t = py_term<IQ>(A.tx_delay[(size_t)w * PS + p], rd, A.tx_apo[(size_t)w * PS + p], ra,
                                           A.t0, A.dt, A.w0, A.L);
              if (A.rx_any[n] && A.tx_any[w]) {
#pragma unroll
                for (int f = 0; f < FC; ++f)
                  if (f0 + f < A.F)
                    cadd(acc[f], py_contrib<IQ>(t, trace + (size_t)w * LN + (size_t)(f0 + f) * LNW));
              }
            }
          } else if (ra != 0.0f) {
            const float rd = A.rx_delay[(size_t)n * PS + p];
            for (int w = 0; w < A.W; ++w) {
              const float ta = A.tx_apo[(size_t)w * PS + p];
              if (ta != 0.0f) {
                const float td = A.tx_delay[(size_t)w * PS + p];
                const Term t = dasc_term<INTERP, IQ>(td, rd, ta, ra, A.t0, A.fs,
                                                     A.wd, A.L, A.edge);
                if (t.valid) {
                  const float2* trace = A.ch + (size_t)w * LN + (size_t)n * A.L;
#pragma unroll
                  for (int f = 0; f < FC; ++f) {
                    if (f0 + f < A.F) {
                      cadd(acc[f], mat_contrib<INTERP>(t, trace + (size_t)(f0 + f) * LNW));
                    }
                  }
                }
              }
            }
          }
#pragma unroll
          for (int f = 0; f < FC; ++f)
            if (f0 + f < A.F) A.out[((size_t)(f0 + f) * A.N + n) * PS + p] = acc[f];
        }
      }
    } else {  // NONE: out[f][w][n][p]
      if (active) {
        const size_t LN = (size_t)A.L * A.N;
        const size_t LNW = LN * A.W;
        for (int w = 0; w < A.W; ++w) {
          const float ta = A.tx_apo[(size_t)w * PS + p];
          const float td = A.tx_delay[(size_t)w * PS + p];
          for (int n = nb; n < ne; ++n) {
            float2 v[FC];
#pragma unroll
            for (int f = 0; f < FC; ++f) v[f] = make_float2(0.0f, 0.0f);
            const float ra = A.rx_apo[(size_t)n * PS + p];
            if (PY) {
              const PyTerm t = py_term<IQ>(td, A.rx_delay[(size_t)n * PS + p], ta, ra, A.t0, A.dt, A.w0, A.L);
              if (A.rx_any[n] && A.tx_any[w]) {
                const float2* trace = A.ch + (size_t)w * LN + (size_t)n * A.L;
#pragma unroll
                for (int f = 0; f < FC; ++f)
                  if (f0 + f < A.F) v[f] = py_contrib<IQ>(t, trace + (size_t)(f0 + f) * LNW);
              }
            } else if (ta != 0.0f && ra != 0.0f) {
              const float rd = A.rx_delay[(size_t)n * PS + p];
              const Term t = dasc_term<INTERP, IQ>(td, rd, ta, ra, A.t0, A.fs,
                                                   A.wd, A.L, A.edge);
              if (t.valid) {
                const float2* trace = A.ch + (size_t)w * LN + (size_t)n * A.L;
#pragma unroll
                for (int f = 0; f < FC; ++f)
                  if (f0 + f < A.F) v[f] = mat_contrib<INTERP>(t, trace + (size_t)(f0 + f) * LNW);
              }
            }
#pragma unroll
            for (int f = 0; f < FC; ++f)
              if (f0 + f < A.F)
                A.out[(((size_t)(f0 + f) * A.W + w) * A.N + n) * PS + p] = v[f];
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// launcher
// ---------------------------------------------------------------------------
template <int MODE, int INTERP, bool IQ, bool PY, int SPLIT, int FC>
static hipError_t launch_mat(const MatArgs& A, hipStream_t stream) {
  constexpr int TP = kBlock / SPLIT;
  const unsigned grid = (unsigned)((A.P + TP - 1) / TP);
  hipLaunchKernelGGL((das_matrix_kernel<MODE, INTERP, IQ, PY, SPLIT, FC>), dim3(grid),
                     dim3(kBlock), 0, stream, A);
  return hipGetLastError();
}

template <int MODE, int INTERP, bool IQ, bool PY, int SPLIT>
static hipError_t pick_fc(const MatArgs& A, hipStream_t s) {
  if (A.F == 1) return launch_mat<MODE, INTERP, IQ, PY, SPLIT, 1>(A, s);
  return launch_mat<MODE, INTERP, IQ, PY, SPLIT, 4>(A, s);
}

template <int MODE, int INTERP, bool IQ, bool PY>
static hipError_t pick_split(const MatArgs& A, int split, hipStream_t s) {
  switch (split) {
    case 1: return pick_fc<MODE, INTERP, IQ, PY, 1>(A, s);
    case 2: return pick_fc<MODE, INTERP, IQ, PY, 2>(A, s);
    case 4: return pick_fc<MODE, INTERP, IQ, PY, 4>(A, s);
    default: return pick_fc<MODE, INTERP, IQ, PY, 8>(A, s);
  }
}

template <int MODE, int INTERP, bool PY>
static hipError_t pick_iq(const MatArgs& A, int split, hipStream_t s) {
  if (A.wd != 0.0f) return pick_split<MODE, INTERP, true, PY>(A, split, s);
  return pick_split<MODE, INTERP, false, PY>(A, split, s);
}

template <int MODE>
static hipError_t pick_interp(const MatArgs& A, int interp, bool py, int split, hipStream_t s) {
  if (interp == INTERP_CUBIC) return pick_iq<MODE, INTERP_CUBIC, false>(A, split, s);
  if (py) return pick_iq<MODE, INTERP_LINEAR, true>(A, split, s);
  return pick_iq<MODE, INTERP_LINEAR, false>(A, split, s);
}

__global__ __launch_bounds__(kBlock) void zero_fill_kernel(float2* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    p[i] = make_float2(0.0f, 0.0f);
}

hipError_t launch_zero_fill(void* p, size_t n_complex, hipStream_t stream) {
  if (n_complex == 0) return hipSuccess;
  size_t blocks = (n_complex + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(zero_fill_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, stream,
                     reinterpret_cast<float2*>(p), n_complex);
  return hipGetLastError();
}

// PY: flags[c] = 1 where weight column c (W transmit columns, then N
// receive ones, each P long) holds a nonzero, NaN included (np.any). One
// block per column.
__global__ __launch_bounds__(kBlock) void py_columns_kernel(const float* __restrict__ tx_apo,
                                                            const float* __restrict__ rx_apo, int W, int P,
                                                            int* __restrict__ flags) {
  const int c = blockIdx.x;
  const float* col = c < W ? tx_apo + (size_t)c * P : rx_apo + (size_t)(c - W) * P;
  bool any = false;
  for (int p = threadIdx.x; p < P && !any; p += kBlock) any = col[p] != 0.0f;
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) flags[c] = any ? 1 : 0;
}

static hipError_t dispatch_dim(const MatLaunch& m, MatArgs& A, hipStream_t stream) {
  switch (m.dim) {
    case 0: return pick_interp<0>(A, m.interp, m.py, m.split, stream);
    case 1: return pick_interp<1>(A, m.interp, m.py, m.split, stream);
    case 2: return pick_interp<2>(A, m.interp, m.py, m.split, stream);
    default: return pick_interp<3>(A, m.interp, m.py, m.split, stream);
  }
}

hipError_t launch_das_matrix(const MatLaunch& m, hipStream_t stream) {
  MatArgs A;
  A.ch = reinterpret_cast<const float2*>(m.ch);
  A.tx_apo = m.tx_apo;
  A.rx_apo = m.rx_apo;
  A.tx_delay = m.tx_delay;
  A.rx_delay = m.rx_delay;
  A.out = reinterpret_cast<float2*>(m.out);
  A.L = m.L; A.N = m.N; A.W = m.W; A.F = m.F; A.P = m.P;
  A.fs = m.fs; A.t0 = m.t0; A.wd = m.wd; A.edge = m.edge;
  A.dt = m.dt; A.w0 = m.w0;
  A.tx_any = A.rx_any = nullptr;
  if (!m.py) return dispatch_dim(m, A, stream);
  // the column flags live in stream-ordered memory for this launch only
  int* flags = nullptr;
  hipError_t e = hipMallocAsync((void**)&flags, sizeof(int) * (size_t)(m.W + m.N), stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(py_columns_kernel, dim3((unsigned)(m.W + m.N)), dim3(kBlock), 0, stream, m.tx_apo, m.rx_apo,
                     m.W, m.P, flags);
  e = hipGetLastError();
  if (e == hipSuccess) {
    A.tx_any = flags;
    A.rx_any = flags + m.W;
    e = dispatch_dim(m, A, stream);
  }
  const hipError_t e2 = hipFreeAsync(flags, stream);
  return e != hipSuccess ? e : e2;
}

}  // namespace ustbx
