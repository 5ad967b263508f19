// Instantiations of the exchange kernels with bra row angular momentum L1 = 1.
#include "exchange_impl.hpp"

namespace lsjk {
template void launch_k_quartets_l1<1>(int, int, int, int, const KArgs&, const KUnit*, int, hipStream_t);
}  // namespace lsjk
