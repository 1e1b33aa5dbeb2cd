// synth_host.hip — host-side synthetic channel data (point scatterers and
// speckle), used by the benchmark and tests to build seeded inputs of the
// configurations' sizes. Host code only (no kernels).
//
// Echo model after the reference's point-scatterer simulator
// (@fresnel/fresnel.m:87-137) and pulse (+uff/pulse.m:76-79):
//   env(t) = exp(-1.7886 (t bw f0)^2), tau = tx_delay(s, w) + |x_s - x_n| / c
//   RF:  sum_s a_s env(t - tau) cos(2 pi f0 (t - tau))
//   IQ:  sum_s a_s env(t - tau) exp(-1j 2 pi f0 tau)   (down-mixed at f0)
// Each trace (f, w, n) is produced by one thread, scatterers in index order,
// so the output does not depend on the thread count. Along a trace the
// Gaussian and the carrier are advanced by recurrences in FP64.
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/ustbx_synth.h"

namespace {

struct SynthJob {
  const ustbx_synth_desc* d;
  const double* ex;
  const double* ey;
  const double* ez;
  const double* pos;    // [F][S][3]
  const double* amp;    // [F][S]
  const double* txd;    // [F][W][S]
  void* out;
  int64_t t_begin, t_end;  // trace range [F*W*N]
};

void synth_trace(const SynthJob& j, int64_t tr, std::vector<double>& re, std::vector<double>& im) {
  const ustbx_synth_desc& d = *j.d;
  const int L = d.L, N = d.N, W = d.W, S = d.S;
  const int n = (int)(tr % N);
  const int w = (int)((tr / N) % W);
  const int f = (int)(tr / ((int64_t)N * W));
  std::fill(re.begin(), re.end(), 0.0);
  std::fill(im.begin(), im.end(), 0.0);
  const double dt = 1.0 / d.fs;
  const double alpha = 1.7886 * (d.bandwidth * d.f0) * (d.bandwidth * d.f0);
  const double q = exp(-2.0 * alpha * dt * dt);
  const double wc = 2.0 * M_PI * d.f0;
  const double crot = cos(wc * dt), srot = sin(wc * dt);
  const double* P = j.pos + (size_t)f * S * 3;
  const double* A = j.amp + (size_t)f * S;
  const double* T = j.txd + ((size_t)f * W + w) * S;
  for (int s = 0; s < S; ++s) {
    const double dx = P[3 * s] - j.ex[n], dy = P[3 * s + 1] - j.ey[n], dz = P[3 * s + 2] - j.ez[n];
    const double tau = T[s] + sqrt(dx * dx + dy * dy + dz * dz) / d.c;
    const double centre = (tau - d.t0) * d.fs;
    int k0 = (int)floor(centre) - d.half + 1;
    int k1 = k0 + 2 * d.half;  // exclusive
    if (k1 <= 0 || k0 >= L) continue;
    // start at k0 (even if < 0: advance the recurrences, write only in range)
    double t = d.t0 + k0 * dt - tau;
    double g = exp(-alpha * t * t);
    double m = exp(-alpha * (2.0 * t * dt + dt * dt));
    const double a = A[s];
    if (d.iq) {
      const double pr = cos(-wc * tau) * a, pi = sin(-wc * tau) * a;
      for (int k = k0; k < k1; ++k) {
        if (k >= 0 && k < L) {
          re[k] += pr * g;
          im[k] += pi * g;
        }
        g *= m;
        m *= q;
      }
    } else {
      double c = cos(wc * t), sn = sin(wc * t);
      for (int k = k0; k < k1; ++k) {
        if (k >= 0 && k < L) re[k] += a * g * c;
        g *= m;
        m *= q;
        const double c2 = c * crot - sn * srot;
        sn = sn * crot + c * srot;
        c = c2;
      }
    }
  }
  if (d.iq) {
    float* o = reinterpret_cast<float*>(j.out) + (size_t)tr * L * 2;
    for (int k = 0; k < L; ++k) {
      o[2 * k] = (float)re[k];
      o[2 * k + 1] = (float)im[k];
    }
  } else {
    float* o = reinterpret_cast<float*>(j.out) + (size_t)tr * L;
    for (int k = 0; k < L; ++k) o[k] = (float)re[k];
  }
}

void* synth_worker(void* arg) {
  const SynthJob& j = *reinterpret_cast<SynthJob*>(arg);
  std::vector<double> re(j.d->L), im(j.d->L);
  for (int64_t tr = j.t_begin; tr < j.t_end; ++tr) synth_trace(j, tr, re, im);
  return nullptr;
}

}  // namespace

extern "C" int ustbx_synth_echoes(const ustbx_synth_desc* d, const double* elem_xyz_x,
                                  const double* elem_xyz_y, const double* elem_xyz_z,
                                  const double* positions, const double* amplitudes,
                                  const double* tx_delays, void* out) {
  if (!d || !elem_xyz_x || !elem_xyz_y || !elem_xyz_z || !positions || !amplitudes || !tx_delays || !out)
    return -1;
  if (d->L <= 0 || d->N <= 0 || d->W <= 0 || d->F <= 0 || d->S < 0 || d->half <= 0) return -1;
  const int64_t traces = (int64_t)d->F * d->W * d->N;
  int nt = d->threads < 1 ? 1 : (d->threads > 64 ? 64 : d->threads);
  if (nt > traces) nt = (int)traces;
  std::vector<SynthJob> jobs(nt);
  std::vector<pthread_t> th(nt);
  for (int t = 0; t < nt; ++t) {
    jobs[t] = SynthJob{d, elem_xyz_x, elem_xyz_y, elem_xyz_z, positions, amplitudes, tx_delays, out,
                       traces * t / nt, traces * (t + 1) / nt};
  }
  int started = 0;
  for (int t = 1; t < nt; ++t) {
    if (pthread_create(&th[t], nullptr, synth_worker, &jobs[t]) != 0) break;
    started = t;
  }
  synth_worker(&jobs[0]);
  for (int t = started + 1; t < nt; ++t) synth_worker(&jobs[t]);
  for (int t = 1; t <= started; ++t) pthread_join(th[t], nullptr);
  return 0;
}
