// das_geom_rxnone.hip — instantiations of the geometry kernel for dimension.receive and dimension.none
// (one translation unit per mode family so the build compiles them in parallel).
#include "das_geom_impl.h"

namespace ustbx {

template <int MODE, int INTERP, int PH>
static hipError_t rn_wc(const GeomLaunch& G, int wc, hipStream_t s) {
  if (wc <= 1) return launch_geom_t<MODE, INTERP, PH, 1, 1, 1>(G, s);
  return launch_geom_t<MODE, INTERP, PH, 1, 8, 1>(G, s);
}

template <int MODE, int INTERP>
static hipError_t rn_ph(const GeomLaunch& G, int ph, int wc, hipStream_t s) {
  if (ph == PH_RF) return rn_wc<MODE, INTERP, PH_RF>(G, wc, s);
  if (ph == PH_DASC) return rn_wc<MODE, INTERP, PH_DASC>(G, wc, s);
  return rn_wc<MODE, INTERP, PH_SEP>(G, wc, s);
}

template <int MODE>
static hipError_t rn_interp(const GeomLaunch& G, int ph, int wc, hipStream_t s) {
  if (G.interp == INTERP_CUBIC) return rn_ph<MODE, INTERP_CUBIC>(G, ph, wc, s);
  return rn_ph<MODE, INTERP_LINEAR>(G, ph, wc, s);
}

hipError_t launch_geom_rx_none(const GeomLaunch& G, int ph, int wc, hipStream_t s) {
  if (G.dim == MODE_RX) return rn_interp<MODE_RX>(G, ph, wc, s);
  return rn_interp<MODE_NONE>(G, ph, wc, s);
}

}  // namespace ustbx
