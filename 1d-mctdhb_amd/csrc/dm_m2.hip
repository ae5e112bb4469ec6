// dm_m2.hip - fused density-matrix kernel instantiated for M = 2 orbitals.
#include "density_impl.hpp"

namespace mctdhb {
template struct DmFused<2>;
}  // namespace mctdhb
