// dm_m7.hip - fused density-matrix kernel instantiated for M = 7 orbitals.
#include "density_impl.hpp"

namespace mctdhb {
template struct DmFused<7>;
}  // namespace mctdhb
