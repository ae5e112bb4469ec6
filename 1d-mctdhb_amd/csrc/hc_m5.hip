// hc_m5.hip - fused H.C kernels instantiated for M = 5 orbitals.
#include "hamiltonian_pair.hpp"

namespace mctdhb {
template struct HcFused<5>;
}  // namespace mctdhb
