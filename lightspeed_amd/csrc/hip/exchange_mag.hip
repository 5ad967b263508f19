// Training data of the learned screen: the unit magnitude kernels (k_unit_mag, exchange_impl.hpp)
// of the classes with s and p functions only.
#include "exchange_impl.hpp"

namespace lsjk {

// Unit magnitudes of one class with s and p functions only (instantiated in exchange_mag.hip).
void launch_k_unit_mag_sp(int L1, int L2, int L3, int L4, const KArgs& a, const KGroups& kg,
                                 const KUnitG* units, int nunits, float smin, float* out, hipStream_t s)
{
    if (nunits <= 0) return;
    dim3 grid((unsigned)((nunits + kKWaves - 1) / kKWaves)), block(64 * kKWaves);
#define CASE(A, B, C, D)                                                                                   \
    if (L1 == A && L2 == B && L3 == C && L4 == D) {                                                        \
        hipLaunchKernelGGL((k_unit_mag<A, B, C, D>), grid, block, 0, s, a, kg, units, nunits, smin, out); \
        return;                                                                                            \
    }
#define C2(A, B) CASE(A, B, 0, 0) CASE(A, B, 0, 1) CASE(A, B, 1, 0) CASE(A, B, 1, 1)
    C2(0, 0) C2(0, 1) C2(1, 0) C2(1, 1)
#undef C2
#undef CASE
    throw std::runtime_error("lsjk: unit magnitudes are instantiated for s and p classes only");
}

}  // namespace lsjk
