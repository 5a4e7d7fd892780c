"""Molecules, Gaussian basis sets and AO integrals for the real-molecule front end.

The reference takes its AO integrals from PySCF through an HDF5 file
(scripts/generate.py:14-45 writes S, h, eri, dm, Enuc, nelec, aolabels;
src/read.rs:17-40 reads them).  Neither PySCF nor h5py is available here, so
this module computes the same quantities itself for light-atom molecules:
contracted Cartesian Gaussians of angular momentum s and p (d is accepted by
the formulas but no basis set here uses it), with overlap, kinetic, nuclear
attraction and two-electron repulsion integrals by the McMurchie-Davidson
scheme (Hermite expansion coefficients E_t^{ij} and Hermite Coulomb integrals
R_tuv built on the Boys function F_n).

Conventions follow the reference: ERIs are chemists' (ab|cd) stored
[a, b, c, d]; lengths are converted from Angstrom with PySCF's Bohr radius;
the AO label of every basis function records its atom (src/basis/atoms.rs
groups AOs per atom for spin and spatial biasing).

Energies computed from these integrals depend only on the span of each atom's
contracted functions, so the per-function normalisation convention (each
contracted function here has unit norm) does not affect any comparison with
the reference's fixture energies.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.special

BOHR = 0.52917721092  # Angstrom per Bohr (PySCF's value)

# Basis sets as {element: [(l, [(exponent, coefficient), ...]), ...]}, the
# published contractions (coefficients refer to normalised primitives).
BASIS_SETS = {
    "sto-3g": {
        "H": [(0, [(3.42525091, 0.15432897), (0.62391373, 0.53532814), (0.16885540, 0.44463454)])],
        # Hehre, Stewart, Pople, J. Chem. Phys. 51, 2657 (1969): 1s and the shared-exponent 2sp shell
        "F": [(0, [(166.6791300, 0.15432897), (30.3608120, 0.53532814), (8.2168207, 0.44463454)]),
              (0, [(6.4648032, -0.09996723), (1.5022812, 0.39951283), (0.4885885, 0.70011547)]),
              (1, [(6.4648032, 0.15591627), (1.5022812, 0.60768372), (0.4885885, 0.39195739)])],
    },
    "3-21g": {
        "H": [(0, [(5.4471780, 0.1562850), (0.8245470, 0.9046910)]),
              (0, [(0.1831920, 1.0)])],
    },
    "6-31g": {
        "H": [(0, [(18.7311370, 0.03349460), (2.8253937, 0.23472695), (0.6401217, 0.81375733)]),
              (0, [(0.1612778, 1.0)])],
        # Dill and Pople, J. Chem. Phys. 62, 2921 (1975): 1s, 2sp (shared exponents), 3sp
        "Li": [(0, [(642.4189150, 0.0021426), (96.7985150, 0.0162089), (22.0911210, 0.0773156),
                    (6.2010703, 0.2457860), (1.9351177, 0.4701890), (0.6367358, 0.3454708)]),
               (0, [(2.3249184, -0.0350917), (0.6324306, -0.1912328), (0.0790534, 1.0839878)]),
               (1, [(2.3249184, 0.0089415), (0.6324306, 0.1410095), (0.0790534, 0.9453637)]),
               (0, [(0.0359620, 1.0)]),
               (1, [(0.0359620, 1.0)])],
    },
    "cc-pvdz": {
        "H": [(0, [(13.01, 0.019685), (1.962, 0.137977), (0.4446, 0.478148), (0.122, 0.50124)]),
              (0, [(0.122, 1.0)]),
              (1, [(0.727, 1.0)])],
    },
}

CHARGES = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9}


def _cart(l):
    """Cartesian components of angular momentum l in PySCF order (x before y before z)."""
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def _dfact(n):
    return 1.0 if n <= 0 else float(np.prod(np.arange(n, 0, -2)))


def _prim_norm(alpha, lmn):
    i, k, m = lmn
    L = i + k + m
    num = (2.0 * alpha / math.pi) ** 0.75 * (4.0 * alpha) ** (L / 2.0)
    return num / math.sqrt(_dfact(2 * i - 1) * _dfact(2 * k - 1) * _dfact(2 * m - 1))


@dataclass
class Molecule:
    """Atoms (symbol, xyz) and a basis set name; geometry in `unit`."""
    atoms: list
    basis: str
    unit: str = "Ang"
    charge: int = 0
    spin: int = 0  # n_alpha - n_beta
    coords: np.ndarray = field(init=False)
    Z: np.ndarray = field(init=False)

    def __post_init__(self):
        scale = 1.0 / BOHR if self.unit.lower().startswith("ang") else 1.0
        self.coords = np.array([a[1] for a in self.atoms], dtype=float) * scale
        self.Z = np.array([CHARGES[a[0]] for a in self.atoms], dtype=float)
        bs = BASIS_SETS[self.basis.lower()]
        # primitive Cartesian functions: centre, exponent, lmn, column of the contraction
        prims, cols, labels = [], [], []
        nbf = 0
        for ia, (sym, _) in enumerate(self.atoms):
            for l, shell in bs[sym]:
                for lmn in _cart(l):
                    for alpha, c in shell:
                        prims.append((ia, alpha, lmn))
                        cols.append((nbf, c * _prim_norm(alpha, lmn)))
                    labels.append(ia)
                    nbf += 1
        self.nbf = nbf
        self.ao_atom = np.array(labels)
        self.p_centre = np.array([self.coords[p[0]] for p in prims])
        self.p_alpha = np.array([p[1] for p in prims])
        self.p_lmn = np.array([p[2] for p in prims], dtype=int)
        C = np.zeros((len(prims), nbf))
        for ip, (ib, c) in enumerate(cols):
            C[ip, ib] = c
        self.contraction = C
        ne = int(round(self.Z.sum())) - self.charge
        self.nelec = ((ne + self.spin) // 2, (ne - self.spin) // 2)

    def enuc(self):
        e = 0.0
        for i in range(len(self.Z)):
            for j in range(i):
                e += self.Z[i] * self.Z[j] / np.linalg.norm(self.coords[i] - self.coords[j])
        return e

    def atomao(self):
        """AO indices of each atom (src/basis/atoms.rs atomao_for_labels)."""
        return [list(np.nonzero(self.ao_atom == a)[0]) for a in range(len(self.atoms))]


def boys(nmax, T):
    """F_n(T) for n = 0..nmax, shape (nmax+1, *T.shape), from the regularised
    lower incomplete gamma function; a Taylor series near T = 0."""
    T = np.asarray(T, dtype=float)
    out = np.empty((nmax + 1,) + T.shape)
    small = T < 1e-10
    Ts = np.where(small, 1.0, T)
    for n in range(nmax + 1):
        a = n + 0.5
        val = scipy.special.gamma(a) * scipy.special.gammainc(a, Ts) / (2.0 * Ts ** a)
        ser = 1.0 / (2 * n + 1) - T / (2 * n + 3) + T * T / (2.0 * (2 * n + 5))
        out[n] = np.where(small, ser, val)
    return out


def _hermite_e(imax, jmax, XPA, XPB, p, K):
    """E[i][j][t] arrays over pairs (McMurchie-Davidson recurrences)."""
    E = [[[None] * (imax + jmax + 2) for _ in range(jmax + 1)] for _ in range(imax + 1)]
    zero = np.zeros_like(p)
    inv2p = 0.5 / p

    def get(i, j, t):
        if t < 0 or t > i + j or i < 0 or j < 0:
            return zero
        return E[i][j][t]

    E[0][0][0] = K
    for i in range(imax + 1):
        for j in range(jmax + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    E[i][j][t] = inv2p * get(i - 1, j, t - 1) + XPA * get(i - 1, j, t) + (t + 1) * get(i - 1, j, t + 1)
                else:
                    E[i][j][t] = inv2p * get(i, j - 1, t - 1) + XPB * get(i, j - 1, t) + (t + 1) * get(i, j - 1, t + 1)
    return get


class _Pairs:
    """All ordered primitive pairs with their Gaussian-product data."""

    def __init__(self, mol, extra_j=0):
        n = len(mol.p_alpha)
        ia, ib = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        self.ia, self.ib = ia.ravel(), ib.ravel()
        a, b = mol.p_alpha[self.ia], mol.p_alpha[self.ib]
        A, B = mol.p_centre[self.ia], mol.p_centre[self.ib]
        self.p = a + b
        self.b = b
        self.P = (a[:, None] * A + b[:, None] * B) / self.p[:, None]
        self.la, self.lb = mol.p_lmn[self.ia], mol.p_lmn[self.ib]
        lmax = int(mol.p_lmn.max())
        self.lmax = lmax
        self.E = []
        for d in range(3):
            XAB = A[:, d] - B[:, d]
            K = np.exp(-a * b / self.p * XAB * XAB)
            self.E.append(_hermite_e(lmax, lmax + extra_j, self.P[:, d] - A[:, d], self.P[:, d] - B[:, d], self.p, K))

    def e1(self, d, ia, jb, t):
        """E_t for direction d with per-pair angular numbers ia, jb (arrays)."""
        out = np.zeros_like(self.p)
        for i in np.unique(ia):
            for j in np.unique(jb):
                sel = (ia == i) & (jb == j)
                if sel.any():
                    out[sel] = self.E[d](int(i), int(j), t)[sel]
        return out

    def hermite(self):
        """H[pair, t, u, v] = E_t^x E_u^y E_v^z."""
        L = 2 * self.lmax + 1
        Ed = [np.stack([self.e1(d, self.la[:, d], self.lb[:, d], t) for t in range(L)], axis=1) for d in range(3)]
        return np.einsum("pt,pu,pv->ptuv", Ed[0], Ed[1], Ed[2])


def _hermite_r(L, alpha, XPC, YPC, ZPC):
    """R_tuv (t+u+v <= L) over combinations, by the R^n recurrences."""
    T = alpha * (XPC * XPC + YPC * YPC + ZPC * ZPC)
    F = boys(L, T)
    memo = {}

    def R(n, t, u, v):
        if t < 0 or u < 0 or v < 0:
            return 0.0
        key = (n, t, u, v)
        if key in memo:
            return memo[key]
        if t == u == v == 0:
            val = (-2.0 * alpha) ** n * F[n]
        elif t > 0:
            val = (t - 1) * R(n + 1, t - 2, u, v) + XPC * R(n + 1, t - 1, u, v)
        elif u > 0:
            val = (u - 1) * R(n + 1, t, u - 2, v) + YPC * R(n + 1, t, u - 1, v)
        else:
            val = (v - 1) * R(n + 1, t, u, v - 2) + ZPC * R(n + 1, t, u, v - 1)
        memo[key] = val
        return val

    return lambda t, u, v: R(0, t, u, v)


def _contract2(M, C):
    return C.T @ M @ C


def integrals(mol: Molecule):
    """AO integrals (S, T, V, eri) over the contracted functions, each
    contracted function normalised to one.  eri[a, b, c, d] = (ab|cd)."""
    pr = _Pairs(mol, extra_j=2)
    n = len(mol.p_alpha)
    sq = np.sqrt(np.pi / pr.p)
    # overlap and kinetic
    Sd = [pr.e1(d, pr.la[:, d], pr.lb[:, d], 0) * sq for d in range(3)]
    Sprim = (Sd[0] * Sd[1] * Sd[2]).reshape(n, n)
    Td = []
    for d in range(3):
        j = pr.lb[:, d]
        s_p2 = pr.e1(d, pr.la[:, d], j + 2, 0) * sq
        s_m2 = np.where(j >= 2, pr.e1(d, pr.la[:, d], np.maximum(j - 2, 0), 0) * sq, 0.0)
        Td.append(-2.0 * pr.b ** 2 * s_p2 + pr.b * (2 * j + 1) * Sd[d] - 0.5 * j * (j - 1) * s_m2)
    Tprim = (Td[0] * Sd[1] * Sd[2] + Sd[0] * Td[1] * Sd[2] + Sd[0] * Sd[1] * Td[2]).reshape(n, n)
    # nuclear attraction
    H = pr.hermite()  # [pair, t, u, v]
    L = 2 * pr.lmax
    Vprim = np.zeros(n * n)
    for c, Zc in enumerate(mol.Z):
        PC = pr.P - mol.coords[c]
        R = _hermite_r(L, pr.p, PC[:, 0], PC[:, 1], PC[:, 2])
        acc = np.zeros_like(pr.p)
        for t in range(L + 1):
            for u in range(L + 1 - t):
                for v in range(L + 1 - t - u):
                    acc += H[:, t, u, v] * R(t, u, v)
        Vprim -= Zc * 2.0 * np.pi / pr.p * acc
    Vprim = Vprim.reshape(n, n)
    # two-electron repulsion over pair combinations
    npair = len(pr.p)
    p, q = pr.p[:, None], pr.p[None, :]
    alpha = p * q / (p + q)
    PQ = pr.P[:, None, :] - pr.P[None, :, :]
    R = _hermite_r(2 * L, alpha, PQ[..., 0], PQ[..., 1], PQ[..., 2])
    tuv = [(t, u, v) for t in range(L + 1) for u in range(L + 1 - t) for v in range(L + 1 - t - u)]
    acc = np.zeros((npair, npair))
    for (t2, u2, v2) in tuv:
        h2 = H[:, t2, u2, v2] * (-1.0) ** (t2 + u2 + v2)
        if not h2.any():
            continue
        inner = np.zeros((npair, npair))
        for (t1, u1, v1) in tuv:
            h1 = H[:, t1, u1, v1]
            if not h1.any():
                continue
            inner += h1[:, None] * R(t1 + t2, u1 + u2, v1 + v2)
        acc += inner * h2[None, :]
    acc *= 2.0 * np.pi ** 2.5 / (p * q * np.sqrt(p + q))
    eprim = acc.reshape(n, n, n, n)
    C = mol.contraction
    norm = 1.0 / np.sqrt(np.einsum("ia,ij,ja->a", C, Sprim, C))
    Cn = C * norm[None, :]
    S = _contract2(Sprim, Cn)
    T = _contract2(Tprim, Cn)
    V = _contract2(Vprim, Cn)
    eri = np.tensordot(eprim, Cn, axes=([0], [0]))        # b c d A
    eri = np.tensordot(eri, Cn, axes=([0], [0]))          # c d A B
    eri = np.tensordot(eri, Cn, axes=([0], [0]))          # d A B C
    eri = np.tensordot(eri, Cn, axes=([0], [0]))          # A B C D
    return S, T, V, np.ascontiguousarray(eri)
