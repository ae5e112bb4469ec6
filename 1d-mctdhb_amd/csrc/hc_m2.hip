// hc_m2.hip - fused H.C kernels instantiated for M = 2 orbitals.
#include "hamiltonian_pair.hpp"

namespace mctdhb {
template struct HcFused<2>;
}  // namespace mctdhb
