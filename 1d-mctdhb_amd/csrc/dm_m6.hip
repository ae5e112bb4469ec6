// dm_m6.hip - fused density-matrix kernel instantiated for M = 6 orbitals.
#include "density_impl.hpp"

namespace mctdhb {
template struct DmFused<6>;
}  // namespace mctdhb
