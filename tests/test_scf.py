"""RHF harness (one-electron integrals + J/K in an SCF) against the reference's known answer.

tests/coverage/rhf.py:607-646 of the reference quotes the PSI4 RHF energy of H2O/STO-3G
(tests/golden/reference/h2o.xyz) as -74.9420799265908784 Eh with tolerance 1e-11.  The harness
reaches it with the CPU oracle's integrals (pins the oracle's S/T/V/J/K chain and the driver) and
with the HIP integrals (gpu).  The mixed FP32/FP64 route is checked for |dE_SCF| <= 1e-6 Eh over
10 iterations against the FP64 route, from the core guess and, for (H2O)256, from the superposition
of the waters' own densities.
"""
import numpy as np
import pytest

import fixture_cases as FC
from lightspeed_amd import scf

E_REF = -74.9420799265908784


class OracleIntegrals:
    """The same interface as scf.Integrals, from the CPU oracle (test infrastructure)."""

    def __init__(self, O, molecule, basis_name):
        from lightspeed_amd import loaders

        class _B:
            pass
        b = _B()
        b.primitives = loaders.basis_records_from_gbs_file(molecule, basis_name)[1]
        self.O = O
        self.b = O.OBasis(b)
        self.pl = O.OPairList(self.b, self.b, True, 1.0e-14)

    def overlap(self):
        return self.O.overlap(self.pl)

    def kinetic(self):
        return self.O.kinetic(self.pl)

    def potential(self, xyzc):
        return self.O.potential(FC.COULOMB, self.pl, xyzc)

    def coulomb(self, D, thresp, thredp):
        return self.O.coulomb(FC.COULOMB, self.pl, self.pl, D, thresp)

    def exchange(self, D, thresp, thredp):
        return self.O.exchange(FC.COULOMB, self.pl, self.pl, D, True, thresp)


def test_rhf_h2o_sto3g_oracle(oracle_backend):
    mol = FC.molecule("h2o")
    hf = scf.RHF(mol, OracleIntegrals(oracle_backend.O, mol, "sto-3g"))
    E = hf.run()
    assert hf.converged
    assert abs(E - E_REF) < 1.0e-11, E  # the reference's tolerance (tests/coverage/rhf.py:609)


def test_fragment_guess_oracle(oracle_backend):
    """The superposition-of-fragments guess: block-diagonal with each block the converged density of
    that water alone (idempotent in its own overlap metric, trace = 5 pairs), and RHF from it reaches
    the same (H2O)3/STO-3G energy as RHF from the core guess."""
    from lightspeed_amd import workloads as W
    O = oracle_backend.O
    mol = W.water_cluster(3, 6.0, 3)

    def make(frag):
        return OracleIntegrals(O, frag, "sto-3g")
    D0 = scf.fragment_guess(mol, 3, make)
    assert D0.shape == (21, 21)
    assert np.abs(D0[:7, 7:]).max() == 0.0 and np.abs(D0[7:14, 14:]).max() == 0.0
    ints = make(mol)
    S = ints.overlap()
    for m in range(3):
        b = slice(7 * m, 7 * m + 7)
        Dm, Sm = D0[b, b], S[b, b]
        assert abs(np.trace(Dm @ Sm) - 5.0) < 1.0e-10
        assert np.abs(Dm @ Sm @ Dm - Dm).max() < 1.0e-8
    ref = scf.RHF(mol, ints)
    Eref = ref.run()
    hf = scf.RHF(mol, ints, D0=D0)
    E = hf.run()
    assert ref.converged and hf.converged
    assert abs(E - Eref) < 1.0e-9, (E, Eref)
    # the first energy from the fragment guess is within the interaction energy of the answer
    assert abs(hf.history[0][1] - E) < 0.05
    assert hf.iterations <= ref.iterations


def test_nuclear_repulsion_closed_form():
    xyzZ = np.array([[0.0, 0.0, 0.0, 8.0], [0.0, 0.0, 2.0, 1.0], [0.0, 1.5, 0.0, 1.0]])
    ref = 8.0 / 2.0 + 8.0 / 1.5 + 1.0 / np.sqrt(4.0 + 2.25)
    assert abs(scf.nuclear_repulsion(xyzZ) - ref) < 1e-14


@pytest.mark.gpu
def test_rhf_h2o_sto3g_gpu(product_backend):
    ls = product_backend.ls
    mol = FC.molecule("h2o")
    bas = ls.Basis.from_gbs_file(mol, "sto-3g")
    hf = scf.RHF(mol, scf.Integrals(bas))
    E = hf.run()
    print("RHF H2O/STO-3G: E = %.12f after %d iterations" % (E, hf.iterations), flush=True)
    assert hf.converged
    assert abs(E - E_REF) < 1.0e-11, E  # the reference's tolerance (tests/coverage/rhf.py:609)


@pytest.mark.gpu
def test_rhf_mixed_route_energy_error(product_backend):
    """(H2O)10/6-31G, 10 iterations from the core guess: the mixed route (thredp = 1e-6) stays within
    1e-6 Eh of the FP64 route at every iteration."""
    ls = product_backend.ls
    from lightspeed_amd import workloads as W
    mol, bname, _ = W.config("h2o10_631g")
    bas = ls.Basis.from_gbs_file(mol, bname)
    ints = scf.Integrals(bas, 1.0e-12)
    e64 = scf.RHF(mol, ints, thresp=1.0e-12, thredp=1.0e-12, fixed_iterations=10)
    e64.run()
    emx = scf.RHF(mol, ints, thresp=1.0e-12, thredp=1.0e-6, fixed_iterations=10)
    emx.run()
    dE = max(abs(a[1] - b[1]) for a, b in zip(e64.history, emx.history))
    print("(H2O)10/6-31G: E10 = %.10f, max |dE| mixed vs fp64 over 10 iterations = %.3e" % (e64.energy, dE),
          flush=True)
    assert dE <= 1.0e-6


@pytest.mark.gpu
def test_rhf_h2o64_mixed_route_energy_error(product_backend):
    """(H2O)64/6-31G** (1600 functions, the benchmark's molecule), 10 iterations from the core guess:
    the mixed route (thredp = 1e-6) stays within 1e-6 Eh of the FP64 route at every iteration."""
    ls = product_backend.ls
    from lightspeed_amd import workloads as W
    mol, bname, _ = W.config("h2o64_631gss")
    bas = ls.Basis.from_gbs_file(mol, bname)
    ints = scf.Integrals(bas, 1.0e-12)
    e64 = scf.RHF(mol, ints, thresp=1.0e-12, thredp=1.0e-12, fixed_iterations=10)
    e64.run()
    print("(H2O)64/6-31G** FP64 route: E10 = %.10f" % e64.energy, flush=True)
    emx = scf.RHF(mol, ints, thresp=1.0e-12, thredp=1.0e-6, fixed_iterations=10)
    emx.run()
    dE = max(abs(a[1] - b[1]) for a, b in zip(e64.history, emx.history))
    print("(H2O)64/6-31G**: max |dE| mixed vs fp64 over 10 iterations = %.3e" % dE, flush=True)
    assert np.isfinite(e64.energy) and e64.energy < -4000.0  # 64 waters, about -76 Eh each
    assert dE <= 1.0e-6


@pytest.mark.gpu
def test_rhf_h2o256_mixed_route_energy_error(product_backend):
    """(H2O)256/6-31G* (4864 functions, the largest workload), 10 iterations from the superposition of
    the 256 waters' own densities (the core guess of a cluster this size wanders far before DIIS
    takes hold): the mixed route (thredp = 1e-6) stays within 1e-6 Eh of the FP64 route at every
    iteration (dense SCF algebra on the host, J and K through IntBox; one progress line per
    iteration)."""
    ls = product_backend.ls
    from lightspeed_amd import workloads as W

    def say(m):
        print(m, flush=True)

    mol, bname, _ = W.config("h2o256_631gs")
    bas = ls.Basis.from_gbs_file(mol, bname)
    D0 = scf.fragment_guess(mol, 3, lambda frag: scf.Integrals(ls.Basis.from_gbs_file(frag, bname), 1.0e-12))
    say("(H2O)256/6-31G*: fragment guess built from 256 water SCFs")
    ints = scf.Integrals(bas, 1.0e-12)
    e64 = scf.RHF(mol, ints, thresp=1.0e-12, thredp=1.0e-12, fixed_iterations=10, log=say, D0=D0)
    e64.run()
    say("(H2O)256/6-31G* FP64 route: E10 = %.10f" % e64.energy)
    emx = scf.RHF(mol, ints, thresp=1.0e-12, thredp=1.0e-6, fixed_iterations=10, log=say, D0=D0)
    emx.run()
    dE = max(abs(a[1] - b[1]) for a, b in zip(e64.history, emx.history))
    say("(H2O)256/6-31G*: max |dE| mixed vs fp64 over 10 iterations = %.3e" % dE)
    assert np.isfinite(e64.energy) and -76.1 * 256 < e64.energy < -75.9 * 256  # about -76.0 Eh a water
    assert e64.history[-1][2] < 1.0e-2  # the commutator has come down: the SCF is converging
    assert dE <= 1.0e-6


@pytest.mark.gpu
def test_rhf_h2o256_core_guess_mixed_route_energy_error(product_backend):
    """(H2O)256/6-31G*, 10 iterations from the core-Hamiltonian guess.  Plain DIIS from that guess
    swings by 10^4 Eh between iterations (charge sloshing between the waters), which turns any
    difference in J/K into a different trajectory; with the density damped by 0.5 until
    max|FDS - SDF| <= 0.1 (DIIS after that) the trajectory is smooth, and the mixed route stays
    within 1e-6 Eh of the FP64 route at every iteration."""
    ls = product_backend.ls
    from lightspeed_amd import workloads as W

    def say(m):
        print(m, flush=True)

    mol, bname, _ = W.config("h2o256_631gs")
    bas = ls.Basis.from_gbs_file(mol, bname)
    ints = scf.Integrals(bas, 1.0e-12)
    kw = dict(fixed_iterations=10, log=say, damping=0.5, diis_start=0.1)
    e64 = scf.RHF(mol, ints, thresp=1.0e-12, thredp=1.0e-12, **kw)
    e64.run()
    say("(H2O)256/6-31G* core guess, FP64 route: E10 = %.10f" % e64.energy)
    emx = scf.RHF(mol, ints, thresp=1.0e-12, thredp=1.0e-6, **kw)
    emx.run()
    dE = max(abs(a[1] - b[1]) for a, b in zip(e64.history, emx.history))
    say("(H2O)256/6-31G* core guess: max |dE| mixed vs fp64 over 10 iterations = %.3e" % dE)
    assert all(np.isfinite(h[1]) for h in e64.history + emx.history)
    assert dE <= 1.0e-6
