// das_geom_tx.hip — instantiations of the geometry kernel for dimension.transmit
// (one translation unit per mode family so the build compiles them in parallel).
#include "das_geom_impl.h"

namespace ustbx {

template <int INTERP, int PH, int FC>
static hipError_t tx_tiles(const GeomLaunch& G, int wc, int split, hipStream_t s) {
  if (PH == PH_SEP) {
    if (wc >= 16) return split >= 4 ? launch_geom_t<MODE_TX, INTERP, PH, 4, 16, FC>(G, s)
                                    : launch_geom_t<MODE_TX, INTERP, PH, 1, 16, FC>(G, s);
    return split >= 4 ? launch_geom_t<MODE_TX, INTERP, PH, 4, 8, FC>(G, s)
                      : launch_geom_t<MODE_TX, INTERP, PH, 1, 8, FC>(G, s);
  }
  if (wc <= 1) return launch_geom_t<MODE_TX, INTERP, PH, 1, 1, FC>(G, s);
  return split >= 4 ? launch_geom_t<MODE_TX, INTERP, PH, 4, 8, FC>(G, s)
                    : launch_geom_t<MODE_TX, INTERP, PH, 1, 8, FC>(G, s);
}

template <int INTERP, int PH>
static hipError_t tx_fc(const GeomLaunch& G, int wc, int split, hipStream_t s) {
  if (G.F == 1) return tx_tiles<INTERP, PH, 1>(G, wc, split, s);
  return tx_tiles<INTERP, PH, 8>(G, wc, split, s);
}

template <int INTERP>
static hipError_t tx_ph(const GeomLaunch& G, int ph, int wc, int split, hipStream_t s) {
  if (ph == PH_RF) return tx_fc<INTERP, PH_RF>(G, wc, split, s);
  if (ph == PH_DASC) return tx_fc<INTERP, PH_DASC>(G, wc, split, s);
  return tx_fc<INTERP, PH_SEP>(G, wc, split, s);
}

hipError_t launch_geom_tx(const GeomLaunch& G, int ph, int wc, int split, hipStream_t s) {
  if (G.interp == INTERP_CUBIC) return tx_ph<INTERP_CUBIC>(G, ph, wc, split, s);
  return tx_ph<INTERP_LINEAR>(G, ph, wc, split, s);
}

}  // namespace ustbx
