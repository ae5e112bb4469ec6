// hc_m3.hip - fused H.C kernels instantiated for M = 3 orbitals.
#include "hamiltonian_pair.hpp"

namespace mctdhb {
template struct HcFused<3>;
}  // namespace mctdhb
