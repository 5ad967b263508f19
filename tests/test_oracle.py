"""The CPU oracle against the reference's own stored J/K results (CPU only).

These pin the oracle (oracle/jk_oracle.c) before it is used to judge the HIP path: a dropped
term, a wrong sign, a transposed operand or a wrong screening predicate shows up as a
mismatch against fixtures the reference computed itself.
"""
import numpy as np
import pytest

import fixture_cases as FC


@pytest.mark.parametrize("basname", FC.BASES)
@pytest.mark.parametrize("symmetric", [True, False])
def test_prop_fixtures(oracle_backend, basname, symmetric):
    for key, val, ref, tol in FC.prop_cases(oracle_backend, basname, symmetric):
        assert np.abs(val - ref).max() < tol, key


@pytest.mark.parametrize("basname", FC.BASES)
@pytest.mark.parametrize("symmetric", [True, False])
def test_ewald_fixtures(oracle_backend, basname, symmetric):
    for key, val, ref, tol in FC.ewald_cases(oracle_backend, basname, symmetric):
        assert np.abs(val - ref).max() < tol, key


@pytest.mark.parametrize("combo", FC.MIX2_COMBOS)
def test_mix2_fixtures(oracle_backend, combo):
    for key, val, ref, tol in FC.mix2_cases(oracle_backend, *combo):
        assert np.abs(val - ref).max() < tol, key


@pytest.mark.parametrize("rot", range(4))
def test_mix4_fixtures(oracle_backend, rot):
    for key, val, ref, tol in FC.mix4_cases(oracle_backend, rot):
        assert np.abs(val - ref).max() < tol, key
