"""Boys function F_n(T) on the GPU (lsjk_boys / Boys.compute, the reference's Boys::compute,
src/lightspeed/boys.hpp:170-181) against the closed form F_n(T) = 1F1(n + 1/2; n + 3/2; -T) / (2n + 1)
evaluated by mpmath at 30 digits, n = 0..16, across both branches (series + downward recursion
below T = 35, erf + upward recursion above) and the T = 0 limit; relative error <= 1e-13."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ref(n, T):
    import mpmath
    mpmath.mp.dps = 30
    return float(mpmath.hyp1f1(n + 0.5, n + 1.5, -mpmath.mpf(T)) / (2 * n + 1))


def test_boys_against_closed_form(product_backend):
    ls = product_backend.ls
    T = np.array([0.0, 1e-14, 1e-6, 0.01, 0.3, 1.0, 2.5, 7.0, 12.0, 20.0, 27.5, 34.99, 35.0, 35.01, 42.0, 60.0,
                  150.0, 1e3, 1e5])
    F = np.array(ls.Boys.compute(16, ls.Tensor.array(T)).np)
    assert F.shape == (len(T), 17)
    worst = 0.0
    for i, t in enumerate(T):
        for n in range(17):
            ref = _ref(n, t)
            worst = max(worst, abs(F[i, n] - ref) / ref)
    print("max relative error %.2e" % worst, flush=True)
    assert worst <= 1e-13
    # the low orders alone
    F0 = np.array(ls.Boys.compute(0, ls.Tensor.array(T)).np)
    assert np.abs(F0[:, 0] - F[:, 0]).max() <= 1e-15 * np.abs(F[:, 0]).max()
    with pytest.raises(RuntimeError):
        ls.Boys.compute(17, ls.Tensor.array(T))
