/*
 * ustbx_synth.h — seeded synthetic channel data for benchmarks and tests.
 *
 * No counterpart in the reference's MEX interface: this is the input
 * generator of the harness, modelled on the reference's point-scatterer
 * simulator (@fresnel/fresnel.m:87-137) and pulse (+uff/pulse.m:76-79).
 * Host code; deterministic for any thread count.
 */
#ifndef USTBX_SYNTH_H
#define USTBX_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ustbx_synth_desc {
  int32_t L, N, W, F; /* samples, elements, waves, frames */
  int32_t S;          /* scatterers */
  int32_t iq;         /* 1: complex64 IQ at f0; 0: float32 RF */
  int32_t half;       /* taps each side of the echo centre */
  int32_t threads;
  double fs, t0, f0, bandwidth, c;
} ustbx_synth_desc;

/* positions [F][S][3], amplitudes [F][S], tx_delays [F][W][S] (seconds);
 * out: [F][W][N][L] complex64 (iq) or float32. Returns 0 or -1 (bad args). */
int ustbx_synth_echoes(const ustbx_synth_desc* d, const double* elem_x, const double* elem_y,
                       const double* elem_z, const double* positions, const double* amplitudes,
                       const double* tx_delays, void* out);

#ifdef __cplusplus
}
#endif
#endif
