"""Minimal closed-shell RHF driver on top of the HIP integrals: the harness that turns J/K errors
into SCF energy errors.

It follows the structure of the reference's RHF (est/rhf.py:596-723; the coverage test
tests/coverage/rhf.py:607-646 quotes the H2O/STO-3G energy -74.9420799265908784): one-electron
Hamiltonian H = T + V from the nuclear charges, symmetric orthogonalisation S^-1/2, a core-Hamiltonian
guess, Fock matrix F = H + 2 J[D] - K[D] with D = C_occ C_occ^T, energy
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
                 g_conv=1.0e-9, diis_max=8, fixed_iterations=None, log=None):
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
            if len(Fs) >= 2:
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
            D = C[:, :self.nocc] @ C[:, :self.nocc].T
        self.energy = self.history[-1][1]
        self.iterations = len(self.history)
        self.C, self.eps, self.D = C, eps, D
        return self.energy

