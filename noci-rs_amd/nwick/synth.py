"""Seeded synthetic nonorthogonal-reference systems.

A system is an AO problem (overlap S, core Hamiltonian h, dense chemists'
ERIs ``eri[a,b,c,d] = (ab|cd)`` built from seeded Cholesky-like vectors) plus
a set of reference determinants, each an orbital-rotated copy of a common
orthonormal MO set: ``C = C0 @ expm(kappa)`` with ``kappa`` antisymmetric in
its occupied-virtual blocks.  MO coefficients are ``n_bas x nmo`` row-major
with occupied columns first, the convention the reference requires
(src/scf/occupation.rs:75-100, src/nonorthogonalwicks/build.rs:1008-1031).

Optionally some references get ``orth_count`` occupied orbitals of one spin
rotated to be exactly orthogonal (in the S metric) to the occupied spaces of
every other reference, so those reference pairs have zero singular values in
their occupied-overlap matrix (m > 0).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.linalg


@dataclass
class System:
    n_bas: int
    nmo: int
    nocc_a: int
    nocc_b: int
    S: np.ndarray
    h: np.ndarray
    eri: np.ndarray | None
    chol: np.ndarray | None
    enuc: float
    refs: list = field(default_factory=list)  # list of (Ca, Cb)

    @property
    def nref(self) -> int:
        return len(self.refs)

    def eri_asym(self) -> np.ndarray:
        """(ab||cd) = (ab|cd) - (ac|bd), as src/read.rs:17-20 antisymmetrise."""
        return self.eri - self.eri.transpose(0, 2, 1, 3)


def _sym_normal(rng, n, sigma):
    a = rng.normal(0.0, sigma, size=(n, n))
    iu = np.triu_indices(n)
    out = np.zeros((n, n))
    out[iu] = a[iu]
    return out + np.triu(out, 1).T


def _s_orthonormalise(C, S, fixed=0):
    """Gram-Schmidt columns of C in the S metric, keeping the first `fixed`
    columns' span order (they are assumed orthonormal already)."""
    C = C.copy()
    n = C.shape[1]
    for k in range(fixed, n):
        v = C[:, k]
        for _ in range(2):
            proj = C[:, :k].T @ (S @ v)
            v = v - C[:, :k] @ proj
        nrm = np.sqrt(v @ S @ v)
        C[:, k] = v / nrm
    return C


def make_system(n_bas: int, nocc_a: int, nocc_b: int, nref: int, seed: int,
                n_aux: int | None = None, orthonormal_ao: bool = True,
                h_sigma: float = 0.1, chol_sigma: float = 0.05,
                kappa_sigma: float = 0.3, orth_refs=(), orth_count: int = 1,
                orth_spin: str = "alpha", dense_eri: bool = True,
                enuc: float = 0.0, orth_against=None) -> System:
    """Build one seeded synthetic system.

    h: symmetric N(0, h_sigma) entries plus diag linspace(-2, 0).
    ERI: (pq|rs) = sum_K L_Kpq L_Krs with symmetric L_K ~ N(0, chol_sigma),
         n_aux = 3 n_bas unless given.
    C0: eigenvectors of h in the S metric (ascending), occupied first.
    Each reference: C = C0 expm(kappa), kappa occ-virt antisymmetric with
    N(0, kappa_sigma) entries, drawn independently per spin.
    orth_refs: references whose first orth_count occupied orbitals (of
    orth_spin) are made S-orthogonal to the occupied spaces of the references
    in orth_against (default: every other reference).
    """
    rng = np.random.default_rng(seed)
    n = n_bas
    if n_aux is None:
        n_aux = 3 * n
    h = _sym_normal(rng, n, h_sigma) + np.diag(np.linspace(-2.0, 0.0, n))
    chol = np.empty((n_aux, n, n))
    for k in range(n_aux):
        chol[k] = _sym_normal(rng, n, chol_sigma)
    if orthonormal_ao:
        S = np.eye(n)
    else:
        a = rng.normal(0.0, 0.08, size=(n, n))
        S = np.eye(n) + 0.5 * (a + a.T)
        w = np.linalg.eigvalsh(S)
        if w.min() < 0.3:
            S += (0.3 - w.min()) * np.eye(n)
        d = 1.0 / np.sqrt(np.diag(S))
        S = S * d[:, None] * d[None, :]
    eri = None
    if dense_eri:
        L2 = chol.reshape(n_aux, n * n)
        eri = (L2.T @ L2).reshape(n, n, n, n)
    # C0: S-orthonormal eigenvectors of h
    w, C0 = scipy.linalg.eigh(h, S)
    refs = []
    for r in range(nref):
        pair = []
        for nocc in (nocc_a, nocc_b):
            nv = n - nocc
            kap = np.zeros((n, n))
            blk = rng.normal(0.0, kappa_sigma, size=(nv, nocc))
            kap[nocc:, :nocc] = blk
            kap[:nocc, nocc:] = -blk.T
            pair.append(C0 @ scipy.linalg.expm(kap))
        refs.append(pair)
    spin_idx = 0 if orth_spin == "alpha" else 1
    nocc_s = (nocc_a, nocc_b)[spin_idx]
    for r in orth_refs:
        against = range(nref) if orth_against is None else orth_against
        others = [refs[x][spin_idx][:, :nocc_s] for x in against if x != r]
        C = refs[r][spin_idx].copy()
        Q = _s_orthonormalise(np.concatenate(others, axis=1), S) if others else np.zeros((n, 0))
        occ = C[:, :nocc_s].copy()
        # replace the first orth_count occupied orbitals by directions S-orthogonal
        # to every other reference's occupied space and to the kept orbitals
        kept = occ[:, orth_count:]
        basis = np.concatenate([Q, kept], axis=1)
        basis = _s_orthonormalise(basis, S)
        new = []
        for k in range(orth_count):
            v = occ[:, k] + 0.3 * rng.normal(size=n)
            for _ in range(3):
                v = v - basis @ (basis.T @ (S @ v))
                for u in new:
                    v = v - u * (u @ S @ v)
            v = v / np.sqrt(v @ S @ v)
            new.append(v)
        occ_new = np.concatenate([np.stack(new, axis=1), kept], axis=1)
        full = np.concatenate([occ_new, C[:, nocc_s:]], axis=1)
        full = _s_orthonormalise(full, S, fixed=0)
        refs[r][spin_idx] = full
    sysm = System(n_bas=n, nmo=n, nocc_a=nocc_a, nocc_b=nocc_b, S=S, h=h, eri=eri,
                  chol=chol, enuc=enuc, refs=[(np.ascontiguousarray(a), np.ascontiguousarray(b)) for a, b in refs])
    return sysm


# Workload presets named by what they are.
def wick_batch_n100(seed: int = 2024, dense_eri: bool = True) -> System:
    """n_bas = nmo = 100, 10+10 electrons, 8 references, orthonormal AO
    basis, n_aux = 300, one reference (10% of 8, rounded up) with one alpha
    occupied orbital exactly orthogonal to every other reference."""
    return make_system(100, 10, 10, 8, seed, n_aux=300, orth_refs=(7,), orth_count=1,
                       dense_eri=dense_eri)


# Uneven multi-system batch: (n, references, excitations per reference) with
# n_bas = 5 n and n_occ = n / 2 per spin, n = 4..10 (n_bas 20..50).
C3_SYSTEMS = [(4, 4, 1024), (4, 16, 64), (6, 8, 512), (6, 12, 256),
              (8, 16, 256), (8, 6, 768), (10, 8, 512), (10, 4, 1024)]


def capped_counts(nmo: int, nocc: int, per_ref: int) -> dict:
    """The default excitation mix (10% alpha singles, 10% beta singles, doubles
    split over aa/bb/ab) with every type capped at the number that exists; the
    remainder goes to alpha-beta doubles."""
    from math import comb
    v = nmo - nocc
    pool = {"a": nocc * v, "b": nocc * v, "aa": comb(nocc, 2) * comb(v, 2), "bb": comb(nocc, 2) * comb(v, 2),
            "ab": (nocc * v) ** 2}
    ns = per_ref // 10
    nd = per_ref - 2 * ns
    want = {"a": ns, "b": ns, "aa": nd // 3, "bb": nd // 3}
    c = {k: min(w, pool[k]) for k, w in want.items()}
    c["ab"] = min(per_ref - sum(c.values()), pool["ab"])
    return c


def batch_systems(seed: int = 3030, systems=C3_SYSTEMS):
    """Seeded systems and determinant lists of the uneven batch.  References
    of systems with at least 6 references get one exactly orthogonal
    occupied alpha orbital on the last reference (m > 0 pairs) when the
    basis has room for it."""
    from . import basis
    out = []
    for k, (n, nref, per) in enumerate(systems):
        nb, no = 5 * n, n // 2
        orth = (nref - 1,) if nref >= 6 and nref * no <= nb else ()
        sysm = make_system(nb, no, no, nref, seed=seed + k, orth_refs=orth, orthonormal_ao=True)
        dets = basis.excited_basis(nref, nb, no, no, per_ref=per, seed=seed + k, include_refs=True,
                                   counts=capped_counts(nb, no, per))
        out.append((sysm, dets))
    return out
