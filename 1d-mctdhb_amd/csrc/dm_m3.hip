// dm_m3.hip - fused density-matrix kernel instantiated for M = 3 orbitals.
#include "density_impl.hpp"

namespace mctdhb {
template struct DmFused<3>;
}  // namespace mctdhb
