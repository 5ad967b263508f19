// Instantiations of the exchange-gradient kernels with L1 = 2, L2 = 1.
#include "exchange_impl.hpp"

namespace lsjk {
template void launch_k_grad_l12<2, 1>(int, int, const KArgs&, const KUnit*, int, hipStream_t);
}  // namespace lsjk
