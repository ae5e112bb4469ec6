// hc_m4.hip - fused H.C kernels instantiated for M = 4 orbitals.
#include "hamiltonian_pair.hpp"

namespace mctdhb {
template struct HcFused<4>;
}  // namespace mctdhb
