// hc_m7.hip - fused H.C kernels instantiated for M = 7 orbitals.
#include "hamiltonian_pair.hpp"

namespace mctdhb {
template struct HcFused<7>;
}  // namespace mctdhb
