// Launch-shape selection for the dense kernels.
#include "internal.hpp"

namespace sctl_amd {

DenseShape pick_dense_shape(int ker, int precision, int64_t nt, int64_t ns) {
  (void)ker;
  (void)ns;
  (void)precision;
  DenseShape s;
  // Enough target blocks to fill the device with TPL targets per lane,
  // otherwise one target per lane and the source range split over the grid.
  s.tpl = nt >= 256LL * 2 * 2048 ? 2 : 1;
  return s;
}

}  // namespace sctl_amd
