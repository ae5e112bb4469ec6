// dm_m4.hip - fused density-matrix kernel instantiated for M = 4 orbitals.
#include "density_impl.hpp"

namespace mctdhb {
template struct DmFused<4>;
}  // namespace mctdhb
