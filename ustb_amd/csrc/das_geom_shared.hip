// das_geom_shared.hip — instantiations of the shared-receive-state kernel
// (dimension both, separable phase, one frame). WS = wave-splits per block
// (block = 64 pixels x WS wave64s), WC = transmit waves per split per round.
#include "das_geom_impl.h"

namespace ustbx {

template <int INTERP>
static hipError_t shared_tiles(const GeomLaunch& G, hipStream_t s) {
  if (G.split == 8) {
    switch (G.wave_chunk) {
      case 8: return launch_both_shared_rx_t<INTERP, 8, 8>(G, s);
      case 10: return launch_both_shared_rx_t<INTERP, 10, 8>(G, s);
      case 12: return launch_both_shared_rx_t<INTERP, 12, 8>(G, s);
      default: return launch_both_shared_rx_t<INTERP, 16, 8>(G, s);
    }
  }
  switch (G.wave_chunk) {
    case 8: return launch_both_shared_rx_t<INTERP, 8, 4>(G, s);
    case 10: return launch_both_shared_rx_t<INTERP, 10, 4>(G, s);
    case 12: return launch_both_shared_rx_t<INTERP, 12, 4>(G, s);
    default: return launch_both_shared_rx_t<INTERP, 16, 4>(G, s);
  }
}

hipError_t launch_both_shared(const GeomLaunch& G, hipStream_t s) {
  if (G.interp == INTERP_CUBIC) return shared_tiles<INTERP_CUBIC>(G, s);
  return shared_tiles<INTERP_LINEAR>(G, s);
}

}  // namespace ustbx
