"""Reference determinants for real molecules: RHF/UHF with DIIS, MOM and biased guesses.

Restates the reference's SCF state generation (src/scf/cycle.rs:220-380
scf_cycle; src/scf/diis.rs Diis; src/scf/select.rs aufbau_indices/mom_select;
src/basis/bias.rs bias_spin/bias_spatial/biased_density_guess;
src/basis/mom.rs generate_states_mom; src/scf/occupation.rs:75-110
occ_first; src/basis/duplicate.rs:69 mark_duplicate_noci_states) so the
reference NOCI fixtures can be driven through the GPU engine.  The AO data
comes from molecule.integrals in place of PySCF.

This is host set-up work on matrices of a few tens of AOs, not part of the
accelerated path.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .molecule import Molecule, integrals
from .synth import System


@dataclass
class AoData:
    """AoData (src/lib.rs:52-76) for one geometry."""
    S: np.ndarray
    X: np.ndarray          # Loewdin S^-1/2
    h: np.ndarray
    eri: np.ndarray        # chemists' (ab|cd) stored [a, b, c, d]
    enuc: float
    nelec: tuple
    atomao: list
    dm: np.ndarray | None = None  # initial guess density (both spins)

    @property
    def n(self):
        return self.S.shape[0]


def ao_data(mol: Molecule) -> AoData:
    S, T, V, eri = integrals(mol)
    w, U = np.linalg.eigh(S)
    X = (U / np.sqrt(w)) @ U.T
    ao = AoData(S=S, X=X, h=T + V, eri=eri, enuc=mol.enuc(), nelec=mol.nelec, atomao=mol.atomao())
    ao.dm = atomic_guess(mol, S, T + V)
    return ao


def fock(h, eri, da, db):
    """Spin Fock matrices F_s = h + J[Da + Db] - K[D_s] (src/scf/kernels.rs:46)."""
    J = np.einsum("abcd,cd->ab", eri, da + db)
    Ka = np.einsum("acbd,cd->ab", eri, da)
    Kb = np.einsum("acbd,cd->ab", eri, db)
    return h + J - Ka, h + J - Kb


def energy(h, enuc, da, db, fa, fb):
    return 0.5 * (np.sum((h + fa) * da) + np.sum((h + fb) * db)) + enuc


class Diis:
    """Pulay DIIS over both spins with a shared coefficient vector (diis.rs)."""

    def __init__(self, space=8):
        self.m = space
        self.f, self.e = [], []

    def push(self, fa, fb, da, db, S, X):
        err = [X.T @ (F @ D @ S - S @ D @ F) @ X for F, D in ((fa, da), (fb, db))]
        self.f.append((fa, fb))
        self.e.append(err)
        if len(self.f) > self.m:
            self.f.pop(0)
            self.e.pop(0)

    def last_error(self):
        ea, eb = self.e[-1]
        return float(np.sqrt(np.sum(ea * ea) + np.sum(eb * eb)))

    def extrapolate(self):
        m = len(self.f)
        if m < 2:
            return None
        B = np.zeros((m + 1, m + 1))
        B[0, 1:] = B[1:, 0] = 1.0
        for i in range(m):
            for j in range(i + 1):
                bij = sum(np.sum(self.e[i][s] * self.e[j][s]) for s in (0, 1))
                B[i + 1, j + 1] = B[j + 1, i + 1] = bij
        w, V = np.linalg.eigh(B)
        keep = np.abs(w) > 1e-13
        if not keep.any():
            return None
        c = (V[:, keep] * (V[0, keep] / w[keep])).sum(axis=1)[1:]
        if abs(c.sum()) > 0:
            c = c / c.sum()
        return (sum(ci * f[0] for ci, f in zip(c, self.f)), sum(ci * f[1] for ci, f in zip(c, self.f)))


def _occupy(e, C, S, nocc, c_old):
    """Aufbau on the first cycle, maximum-overlap selection afterwards
    (select.rs aufbau_indices / mom_select: score = sum_i |(C_old^T S C)_ij|)."""
    if c_old is None:
        idx = np.argsort(e, kind="stable")[:nocc]
    else:
        p = np.abs(c_old.T @ S @ C).sum(axis=0)
        idx = np.argsort(-p, kind="stable")[:nocc]
    return idx


@dataclass
class SCFState:
    label: str
    e: float
    ca: np.ndarray  # all MOs, occupied columns first
    cb: np.ndarray
    na: int
    nb: int
    da: np.ndarray
    db: np.ndarray
    noci: bool = True
    converged: bool = True
    iterations: int = 0


def scf_cycle(ao: AoData, da, db, label="", mom=True, e_tol=1e-12, fds_tol=1e-8, max_cycle=10000, diis_space=8):
    """One UHF solve from densities (da, db) (cycle.rs:220-380)."""
    na, nb = ao.nelec
    e = np.inf
    diis = Diis(diis_space)
    old = [None, None]
    fa, fb = fock(ao.h, ao.eri, da, db)
    for it in range(int(max_cycle)):
        diis.push(fa, fb, da, db, ao.S, ao.X)
        ext = diis.extrapolate()
        fa_use, fb_use = ext if ext is not None else (fa, fb)
        out = []
        for sp, (F, nocc) in enumerate(((fa_use, na), (fb_use, nb))):
            w, V = np.linalg.eigh(ao.X.T @ F @ ao.X)
            C = ao.X @ V
            idx = _occupy(w, C, ao.S, nocc, old[sp] if mom else None)
            old[sp] = C[:, idx]
            out.append((w, C, idx))
        da = old[0] @ old[0].T
        db = old[1] @ old[1].T
        fa, fb = fock(ao.h, ao.eri, da, db)
        e_new = energy(ao.h, ao.enuc, da, db, fa, fb)
        err = diis.last_error()
        if abs(e_new - e) < e_tol and err < fds_tol:
            cols = []
            for (w, C, idx) in out:
                rest = [k for k in range(C.shape[1]) if k not in set(idx.tolist())]
                cols.append(C[:, sorted(idx.tolist()) + rest])   # occ_first (occupation.rs:75-110)
            return SCFState(label, e_new, cols[0], cols[1], na, nb, da, db, iterations=it + 1)
        e = e_new
    raise RuntimeError(f"SCF '{label}' did not converge in {max_cycle} cycles")


def atomic_guess(mol: Molecule, S, h):
    """Initial density (both spins) as a superposition of neutral-atom
    densities: each atom's core Hamiltonian in its own AO block, with its
    electrons filled into the lowest atomic orbitals and shared equally over
    degenerate ones.  This plays the role of PySCF's RHF initial guess
    (rhf.get_init_guess(), the `dm` of scripts/generate.py:18): symmetric
    under the molecule's point group and block-diagonal over atoms, so the
    spin and spatial biases alone decide which broken-symmetry solution each
    state reaches."""
    n = S.shape[0]
    dm = np.zeros((n, n))
    for a, idx in enumerate(mol.atomao()):
        ix = np.ix_(idx, idx)
        w, U = np.linalg.eigh(S[ix])
        X = (U / np.sqrt(w)) @ U.T
        e, V = np.linalg.eigh(X.T @ h[ix] @ X)
        C = X @ V
        ne = float(mol.Z[a])
        occ = np.zeros(len(idx))
        k = 0
        while ne > 1e-12 and k < len(idx):
            deg = [j for j in range(k, len(idx)) if abs(e[j] - e[k]) < 1e-6]
            put = min(ne, 2.0 * len(deg))
            occ[deg] = put / len(deg)
            ne -= put
            k = deg[-1] + 1
        dm[ix] = (C * occ) @ C.T
    return dm


def biased_guess(ao: AoData, spin_bias=None, spatial_bias=None):
    """biased_density_guess (bias.rs:108): half the RHF density per spin,
    AO blocks of each atom scaled by 1 +- pol."""
    da = ao.dm * 0.5
    db = ao.dm * 0.5
    for kind, bias in (("spin", spin_bias), ("spatial", spatial_bias)):
        if not bias:
            continue
        pol, pattern = bias["pol"], bias["pattern"]
        for a, sgn in enumerate(pattern):
            if sgn == 0:
                continue
            idx = np.ix_(ao.atomao[a], ao.atomao[a])
            up, dn = 1.0 + pol, 1.0 - pol
            if kind == "spin":
                fa, fb = (up, dn) if sgn > 0 else (dn, up)
            else:
                fa = fb = up if sgn > 0 else dn
            da[idx] *= fa
            db[idx] *= fb
    return da, db


def _density_distance(a: SCFState, b: SCFState, S):
    d = 0.0
    for Da, Db in ((a.da, b.da), (a.db, b.db)):
        M = (Da - Db) @ S
        d += float(np.trace(M @ M))
    return d


def reference_states(ao: AoData, recipes, d_tol=1e-4, **scf_opts):
    """generate_states_mom (mom.rs:91-118) for ground-state recipes:
    recipes are dicts {label, noci, spin_bias?, spatial_bias?}."""
    states = []
    for r in recipes:
        da, db = biased_guess(ao, r.get("spin_bias"), r.get("spatial_bias"))
        st = scf_cycle(ao, da, db, r.get("label", ""), **scf_opts)
        st.noci = r.get("noci", True)
        states.append(st)
    for i, si in enumerate(states):
        if not si.noci:
            continue
        for sj in states[:i]:
            if sj.noci and _density_distance(sj, si, ao.S) < d_tol:
                si.noci = False
                break
    return states


def to_system(ao: AoData, states) -> System:
    """The NOCI reference basis as a System for the engine and the oracle."""
    refs = [(s.ca, s.cb) for s in states if s.noci]
    na, nb = ao.nelec
    return System(n_bas=ao.n, nmo=ao.n, nocc_a=na, nocc_b=nb, S=ao.S, h=ao.h, eri=ao.eri, chol=None,
                  enuc=ao.enuc, refs=refs)
