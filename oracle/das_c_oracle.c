/*
 * das_c_oracle.c — CPU restatement of the reference das_c MEX beamformer.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the HIP path is compared
 * against and the CPU baseline timed beside it; the product never links it.
 *
 * Follows +mex/source/das_c.cpp of the reference:
 *   - wd = 2*(float)PI*fd when |fd| > 1e-6, else 0            (das_c.cpp:32-33,240-244)
 *   - output strides RX_F/W_F/F_F per sum mode                 (das_c.cpp:324-353)
 *   - pixel-parallel loop, w outer, n inner, frames innermost  (das_c.cpp:355-409)
 *   - skip when tx or rx apodization == 0                      (das_c.cpp:366,371)
 *   - c_delay = tx + rx; denay = (c_delay - t0)*Fs; n0 = (int)denay;
 *     b = denay - floor(denay); a = 1 - b; valid iff 0 < n0 < L-1 (das_c.cpp:372-381)
 *   - weight = exp(i*wd*c_delay)*c_apo; 4-term complex dot summed from zero in
 *     order, then added to the output                          (das_c.cpp:385-400)
 * Built with -ffp-contract=off (the reference MEX is built by `mex` with g++
 * defaults for x86-64, which emit no FMA).
 *
 * Array layout (MATLAB column-major, first index fastest):
 *   ch [L,N,W,F] complex float interleaved; tx_* [P,W]; rx_* [N... ] = [P,N];
 *   out [P, N', W', F], N' = N for NONE/TX else 1, W' = W for NONE/RX else 1.
 * The caller passes a zero-filled `out` (mxCreateNumericArray zero-fills,
 * das_c.cpp:297).
 *
 * Parallelism: the pixel range is split over `nthreads` POSIX threads, each
 * pixel owned by one thread (das_c.cpp:358 uses tbb::parallel_for(0, P)).
 * Summation order per pixel does not depend on the thread count.
 */
#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define ORACLE_NONE 0
#define ORACLE_RX 1
#define ORACLE_TX 2
#define ORACLE_BOTH 3

typedef struct {
  const float* ch;
  int L, N, W, F, P;
  float Fs, t0, wd;
  const float* tx_apo;
  const float* rx_apo;
  const float* tx_delay;
  const float* rx_delay;
  float* out;
  size_t RX_F, W_F, F_F; /* output strides in complex elements */
  int p_begin, p_end;
} oracle_job;

static void pixel_range(const oracle_job* j) {
  const size_t P = (size_t)j->P, L = (size_t)j->L, N = (size_t)j->N, W = (size_t)j->W;
  const size_t LN = L * N, LNW = LN * W;
  for (int pp = j->p_begin; pp < j->p_end; ++pp) {
    for (int w = 0; w < j->W; ++w) {
      const float c_tx_delay = j->tx_delay[P * (size_t)w + pp];
      const float c_tx_apo = j->tx_apo[P * (size_t)w + pp];
      if (c_tx_apo != 0.0f) {
        for (int rx = 0; rx < j->N; ++rx) {
          const float c_rx_delay = j->rx_delay[P * (size_t)rx + pp];
          const float c_rx_apo = j->rx_apo[P * (size_t)rx + pp];
          if (c_rx_apo != 0.0f) {
            const float c_delay = c_tx_delay + c_rx_delay;
            const float c_apo = c_tx_apo * c_rx_apo;
            const float denay = (c_delay - j->t0) * j->Fs;
            const float n0f = floorf(denay);
            /* int(denay): x86 cvttss2si gives INT_MIN for NaN and out-of-range
               values, which the 0 < n0 test then rejects */
            int n0;
            if (!(denay < 2147483648.0f && denay > -2147483648.0f)) n0 = INT32_MIN;
            else n0 = (int)denay;
            const float b = denay - n0f;
            const float a = 1.0f - b;
            if (n0 > 0 && n0 < (j->L - 1)) {
              const float* mem = j->ch + 2 * ((size_t)L * rx + LN * w + (size_t)n0);
              float* pix = j->out + 2 * (j->RX_F * rx + j->W_F * w + (size_t)pp);
              const float phase = j->wd * c_delay;
              const float complex e = cexpf(CMPLXF(0.0f, phase));
              const float wr = crealf(e) * c_apo, wi = cimagf(e) * c_apo;
              const float weightr[4] = {a * wr, -a * wi, b * wr, -b * wi};
              const float weighti[4] = {a * wi, a * wr, b * wi, b * wr};
              for (int f = 0; f < j->F; ++f) {
                float re = 0.0f, im = 0.0f;
                for (int vv = 0; vv < 4; ++vv) {
                  re += weightr[vv] * mem[vv];
                  im += weighti[vv] * mem[vv];
                }
                pix[0] += re;
                pix[1] += im;
                mem += 2 * LNW;
                pix += 2 * j->F_F;
              }
            }
          }
        }
      }
    }
  }
}

static void* worker(void* arg) {
  pixel_range((const oracle_job*)arg);
  return NULL;
}

/* Returns 0 on success, -1 on a bad sum mode. */
int oracle_das_c(const float* ch, int L, int N, int W, int F, int P, float Fs, float t0,
                 float fd, int sum_mode, const float* tx_apo, const float* rx_apo,
                 const float* tx_delay, const float* rx_delay, float* out, int nthreads) {
  float wd = 0.0f;
  if (fabsf(fd) > 1e-6) wd = 2 * (float)3.14159265359 * fd;
  const size_t sP = (size_t)P, sN = (size_t)N, sW = (size_t)W;
  size_t RX_F, W_F, F_F;
  switch (sum_mode) {
    case ORACLE_NONE: RX_F = sP; W_F = sP * sN; F_F = sP * sN * sW; break;
    case ORACLE_RX: RX_F = 0; W_F = sP; F_F = sP * sW; break;
    case ORACLE_TX: RX_F = sP; W_F = 0; F_F = sP * sN; break;
    case ORACLE_BOTH: RX_F = 0; W_F = 0; F_F = sP; break;
    default: return -1;
  }
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (nthreads > P) nthreads = P > 0 ? P : 1;
  oracle_job jobs[256];
  pthread_t th[256];
  for (int t = 0; t < nthreads; ++t) {
    oracle_job* j = &jobs[t];
    j->ch = ch; j->L = L; j->N = N; j->W = W; j->F = F; j->P = P;
    j->Fs = Fs; j->t0 = t0; j->wd = wd;
    j->tx_apo = tx_apo; j->rx_apo = rx_apo; j->tx_delay = tx_delay; j->rx_delay = rx_delay;
    j->out = out; j->RX_F = RX_F; j->W_F = W_F; j->F_F = F_F;
    j->p_begin = (int)(((int64_t)P * t) / nthreads);
    j->p_end = (int)(((int64_t)P * (t + 1)) / nthreads);
  }
  if (nthreads == 1) {
    pixel_range(&jobs[0]);
    return 0;
  }
  int started = 0;
  for (int t = 0; t < nthreads; ++t) {
    if (pthread_create(&th[t], NULL, worker, &jobs[t]) != 0) break;
    started++;
  }
  for (int t = started; t < nthreads; ++t) pixel_range(&jobs[t]);
  for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
  return 0;
}
