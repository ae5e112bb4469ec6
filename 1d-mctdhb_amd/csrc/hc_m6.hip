// hc_m6.hip - fused H.C kernels instantiated for M = 6 orbitals.
#include "hamiltonian_pair.hpp"

namespace mctdhb {
template struct HcFused<6>;
}  // namespace mctdhb
