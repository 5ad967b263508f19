// Instantiations of the exchange-gradient kernels with L1 = 1, L2 = 2.
#include "exchange_impl.hpp"

namespace lsjk {
template void launch_k_grad_l12<1, 2>(int, int, const KArgs&, const KUnit*, int, hipStream_t);
}  // namespace lsjk
