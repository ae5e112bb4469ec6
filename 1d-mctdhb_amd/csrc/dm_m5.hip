// dm_m5.hip - fused density-matrix kernel instantiated for M = 5 orbitals.
#include "density_impl.hpp"

namespace mctdhb {
template struct DmFused<5>;
}  // namespace mctdhb
