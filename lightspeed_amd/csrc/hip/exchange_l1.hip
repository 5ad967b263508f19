// Instantiations of the K build kernels with bra row angular momentum L1 = 1.
#include "exchange_impl.hpp"

namespace lsjk {
template void launch_k_ket_l1<1>(int, int, int, int, const KArgs&, const KGroups&, const KUnitG*, int, hipStream_t);
}  // namespace lsjk
