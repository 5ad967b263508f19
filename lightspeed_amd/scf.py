"""Minimal closed-shell RHF driver on top of the HIP integrals: the harness that turns J/K errors
into SCF energy errors.

It follows the structure of the reference's RHF (est/rhf.py:596-723; the coverage test
tests/coverage/rhf.py:607-646 quotes the H2O/STO-3G energy -74.9420799265908784): one-electron
Hamiltonian H = T + V from the nuclear charges, symmetric orthogonalisation S^-1/2, a core-Hamiltonian
guess (or a given starting density, e.g. `fragment_guess`'s superposition of fragment densities), Fock matrix F = H + 2 J[D] - K[D] with D = C_occ C_occ^T, energy
E = tr[D (H + F)] + E_nuc, and DIIS extrapolation on the commutator FDS - SDF (Pulay, Chem. Phys.
Lett. 73, 393 (1980)).  Dense linear algebra runs on the host (numpy).

The integrals come from an `Integrals` object; the default one calls lightspeed_amd.IntBox (the
HIP path).  `thredp` > `thresp` selects the mixed FP32/FP64 route of IntBox::coulomb/exchange.
"""
import math

import numpy as np


class Integrals:
    """S, T, V, J, K of one basis through lightspeed_amd.IntBox (HIP)."""

    def __init__(self, basis, threpq=1.0e-14):
        import lightspeed_amd as ls
        self.ls = ls
        self.basis = basis
        self.res = ls.ResourceList.build()
        self.ewald = ls.Ewald([1.0], [-1.0])
        self.pairs = ls.PairList.build_schwarz(basis, basis, True, threpq)

    def overlap(self):
        return np.array(self.ls.IntBox.overlap(self.res, self.pairs).np)

    def kinetic(self):
        return np.array(self.ls.IntBox.kinetic(self.res, self.pairs).np)

    def potential(self, xyzc):
        xyzc = self.ls.Tensor.array(np.ascontiguousarray(xyzc, dtype=float))
        return np.array(self.ls.IntBox.potential(self.res, self.ewald, self.pairs, xyzc).np)

    def coulomb(self, D, thresp, thredp):
        Dt = self.ls.Tensor.array(np.ascontiguousarray(D))
        return np.array(self.ls.IntBox.coulomb(self.res, self.ewald, self.pairs, self.pairs, Dt, thresp, thredp).np)

    def exchange(self, D, thresp, thredp):
        Dt = self.ls.Tensor.array(np.ascontiguousarray(D))
        return np.array(
            self.ls.IntBox.exchange(self.res, self.ewald, self.pairs, self.pairs, Dt, True, thresp, thredp).np)


class ScreenedIntegrals(Integrals):
    """Integrals whose exchange runs on the learned route (lightspeed_amd.screen.Screen); J stays on
    the FP64 route."""

    def __init__(self, basis, screen, threpq=1.0e-14):
        super().__init__(basis, threpq)
        self.screen = screen
        n = basis.ncart
        self.dD = self.ls.DeviceArray([n, n])
        self.dK = self.ls.DeviceArray([n, n])
        if basis.has_pure:
            raise ValueError("ScreenedIntegrals: Cartesian bases only")

    def coulomb(self, D, thresp, thredp):
        return super().coulomb(D, thresp, thresp)

    def exchange(self, D, thresp, thredp):
        self.dD.upload(np.ascontiguousarray(D))
        self.dK.zero()
        self.screen.exchange_device(self.ewald, self.pairs, self.pairs, self.dD.ptr, True, thresp, thredp, self.dK.ptr)
        return self.dK.download()


def nuclear_repulsion(xyzZ):
    """sum_{A<B} Z_A Z_B / R_AB (bohr)."""
    xyzZ = np.asarray(xyzZ, dtype=float)
    e = 0.0
    for a in range(len(xyzZ)):
        for b in range(a):
            r = np.linalg.norm(xyzZ[a, :3] - xyzZ[b, :3])
            e += xyzZ[a, 3] * xyzZ[b, 3] / r
    return e


class RHF:
    """Closed-shell RHF.  After run(): energy, converged, iterations, C, eps, D, history."""

    def __init__(self, molecule, integrals, charge=0, thresp=1.0e-14, thredp=1.0e-14, maxiter=100, e_conv=1.0e-12,
                 g_conv=1.0e-9, diis_max=8, fixed_iterations=None, log=None, D0=None, level_shift=0.0,
                 diis_start=None, damping=0.0):
        self.molecule = molecule
        self.ints = integrals
        nel = int(round(sum(a.Z for a in molecule.atoms))) - charge
        if nel % 2:
            raise ValueError("RHF: odd number of electrons")
        self.nocc = nel // 2
        self.thresp, self.thredp = thresp, thredp
        self.maxiter, self.e_conv, self.g_conv, self.diis_max = maxiter, e_conv, g_conv, diis_max
        self.fixed_iterations = fixed_iterations
        self.history = []
        self.log = log  # optional callable: one progress line per iteration
        self.D0 = D0  # optional starting density (nao x nao); None = core-Hamiltonian guess
        # level_shift sigma > 0 raises the virtual orbitals by sigma (F + sigma (S - S D S), Saunders &
        # Hillier, Int. J. Quantum Chem. 7, 699 (1973)) while max|FDS - SDF| > diis_start; below it
        # (or from the first iteration when diis_start is None) DIIS extrapolates without the shift
        self.level_shift, self.diis_start = level_shift, diis_start
        # damping a in [0, 1): while DIIS is off, the next density is (1 - a) D_new + a D (the
        # remedy for charge sloshing between the fragments of a cluster started from the core guess)
        self.damping = damping

    def run(self):
        xyzZ = self.molecule.xyzZ
        S = self.ints.overlap()
        H = self.ints.kinetic() + self.ints.potential(xyzZ)
        self.S, self.H = S, H
        self.Enuc = nuclear_repulsion(xyzZ)
        s, U = np.linalg.eigh(S)
        X = (U / np.sqrt(s)) @ U.T

        def diagonalise(F):
            eps, Cp = np.linalg.eigh(X.T @ F @ X)
            return eps, X @ Cp

        eps, C = diagonalise(H)
        D = C[:, :self.nocc] @ C[:, :self.nocc].T
        if self.D0 is not None:
            if self.D0.shape != D.shape:
                raise ValueError("RHF: D0 has shape %s, the basis needs %s" % (self.D0.shape, D.shape))
            D = np.array(self.D0, dtype=float)
        Fs, Es = [], []
        E_old = 0.0
        self.converged = False
        niter = self.fixed_iterations if self.fixed_iterations is not None else self.maxiter
        for it in range(1, niter + 1):
            J = self.ints.coulomb(D, self.thresp, self.thredp)
            K = self.ints.exchange(D, self.thresp, self.thredp)
            F = H + 2.0 * J - K
            E = float(np.sum(D * (H + F))) + self.Enuc
            G = F @ D @ S - S @ D @ F
            gnorm = float(np.max(np.abs(G)))
            self.history.append((it, E, gnorm))
            if self.log:
                self.log("RHF iteration %d: E = %.10f  max|FDS - SDF| = %.2e" % (it, E, gnorm))
            if self.fixed_iterations is None and it > 1 and abs(E - E_old) < self.e_conv and gnorm < self.g_conv:
                self.converged = True
                break
            E_old = E
            # DIIS on the orthogonal-basis commutator
            Fs.append(F)
            Es.append(X.T @ G @ X)
            if len(Fs) > self.diis_max:
                Fs.pop(0)
                Es.pop(0)
            use_diis = self.diis_start is None or gnorm <= self.diis_start
            if not use_diis and self.level_shift > 0.0:
                F = F + self.level_shift * (S - S @ D @ S)
            elif len(Fs) >= 2:
                n = len(Fs)
                B = -np.ones((n + 1, n + 1))
                B[n, n] = 0.0
                for i in range(n):
                    for j in range(n):
                        B[i, j] = float(np.sum(Es[i] * Es[j]))
                rhs = np.zeros(n + 1)
                rhs[n] = -1.0
                try:
                    c = np.linalg.solve(B, rhs)[:n]
                    F = sum(ci * Fi for ci, Fi in zip(c, Fs))
                except np.linalg.LinAlgError:
                    pass
            eps, C = diagonalise(F)
            Dn = C[:, :self.nocc] @ C[:, :self.nocc].T
            D = Dn if use_diis or self.damping == 0.0 else (1.0 - self.damping) * Dn + self.damping * D
        self.energy = self.history[-1][1]
        self.iterations = len(self.history)
        self.C, self.eps, self.D = C, eps, D
        return self.energy



def fragment_guess(molecule, natom_fragment, make_integrals, thre=1.0e-12, e_conv=1.0e-10, g_conv=1.0e-7,
                   maxiter=60):
    """Starting density for a molecule made of neutral closed-shell fragments of `natom_fragment`
    consecutive atoms each (a water cluster: O, H, H per water): D = diag(D_1, ..., D_n) with D_m the
    converged RHF density of fragment m alone (its own nuclei and electrons).

    The basis functions of fragment m are a consecutive block of the molecule's, because the basis
    loader emits the shells atom by atom in atom order.  `make_integrals(fragment)` returns an
    `Integrals`-like object for the fragment in the molecule's basis set.

    For a large cluster the core-Hamiltonian guess puts the electrons in the unscreened potential of
    every nucleus and DIIS from there wanders for many iterations; this guess starts one interaction
    energy away from the answer."""
    from .loaders import Molecule
    if molecule.natom % natom_fragment:
        raise ValueError("fragment_guess: %d atoms is not a whole number of %d-atom fragments" %
                         (molecule.natom, natom_fragment))
    blocks = []
    for a0 in range(0, molecule.natom, natom_fragment):
        atoms = [a._replace(atomIdx=i) for i, a in enumerate(molecule.atoms[a0:a0 + natom_fragment])]
        frag = Molecule("%s_fragment_%d" % (molecule.name, a0 // natom_fragment), atoms)
        hf = RHF(frag, make_integrals(frag), thresp=thre, thredp=thre, e_conv=e_conv, g_conv=g_conv,
                 maxiter=maxiter)
        hf.run()
        if not hf.converged:
            raise RuntimeError("fragment_guess: fragment %d did not converge" % (a0 // natom_fragment))
        blocks.append(hf.D)
    n = sum(b.shape[0] for b in blocks)
    D = np.zeros((n, n))
    o = 0
    for b in blocks:
        k = b.shape[0]
        D[o:o + k, o:o + k] = b
        o += k
    return D
