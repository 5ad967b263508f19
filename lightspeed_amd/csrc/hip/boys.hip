// Boys function F_n(T) = int_0^1 t^(2n) exp(-T t^2) dt, n = 0..nboys (nboys <= 16), in FP64: the
// role of the reference's Boys::interpolate_F (src/lightspeed/boys.hpp:49-114; Python
// Boys.compute, src/cpp/quad/boys.cpp:9-49), by a method of its own:
//   * T < kBoysTc: F_nboys from its power series
//       F_n(T) = exp(-T) sum_{k>=0} (2T)^k / ((2n+1)(2n+3)...(2n+2k+1))
//     (all terms positive; summed until a term no longer changes the sum), then the downward
//     recursion F_{n-1} = (2T F_n + exp(-T)) / (2n-1), stable for every T;
//   * T >= kBoysTc: F_0 = sqrt(pi/T) erf(sqrt(T)) / 2 and the upward recursion
//     F_{n+1} = ((2n+1) F_n - exp(-T)) / (2T), stable while 2n+1 < 2T (here 2 nboys + 1 <= 33 < 70);
//   * T = 0: F_n = 1/(2n+1).
// Reference: T. Helgaker, P. Jorgensen, J. Olsen, Molecular Electronic-Structure Theory (Wiley,
// 2000), section 9.8 (series, recursions and asymptotics of the Boys function).
#include "internal.hpp"
#include "kernels.hpp"

namespace lsjk {

namespace {

constexpr double kBoysTc = 35.0;

__device__ void boys_eval(int nboys, double T, double* F)
{
    if (T == 0.0) {
        for (int k = 0; k <= nboys; k++) F[k] = 1.0 / (2 * k + 1);
        return;
    }
    const double eT = exp(-T);
    if (T >= kBoysTc) {
        F[0] = 0.5 * sqrt(M_PI / T) * erf(sqrt(T));
        const double h = 0.5 / T;
        for (int k = 1; k <= nboys; k++) F[k] = h * ((2 * k - 1) * F[k - 1] - eT);
        return;
    }
    double term = 1.0 / (2 * nboys + 1), sum = term;
    for (int k = 0; k < 400; k++) {
        term *= 2.0 * T / (2 * nboys + 2 * k + 3);
        const double s = sum + term;
        if (s == sum) break;
        sum = s;
    }
    F[nboys] = eT * sum;
    for (int k = nboys; k > 0; k--) F[k - 1] = (2.0 * T * F[k] + eT) / (2 * k - 1);
}

__global__ void boys_kernel(int nboys, int npoint, const double* __restrict__ T, double* __restrict__ F)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoint) return;
    double f[kBoysMax + 1];
    boys_eval(nboys, T[i], f);
    for (int k = 0; k <= nboys; k++) F[(size_t)i * (nboys + 1) + k] = f[k];
}

}  // namespace

void launch_boys(int nboys, int npoint, const double* T, double* F, hipStream_t s)
{
    if (npoint <= 0) return;
    hipLaunchKernelGGL(boys_kernel, dim3((unsigned)((npoint + 255) / 256)), dim3(256), 0, s, nboys, npoint, T, F);
}

}  // namespace lsjk
