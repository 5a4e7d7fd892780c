"""Brute-force determinant-space H and S for tiny systems (TEST INFRASTRUCTURE).

Independent of the Wick and generalised Slater-Condon routes: each
nonorthogonal determinant is expanded in the determinants of an orthonormal
spin-orbital basis (Loewdin-orthonormalised AOs), and H is applied by second
quantisation, H = sum h_pq a+_p a_q + 1/2 sum <pq|rs> a+_p a+_q a_s a_r,
on occupation bitstrings.  Only for n_bas <= ~8.
"""
from __future__ import annotations

from itertools import combinations

import numpy as np


def _popcount(x):
    return bin(x).count("1")


def _sign_annihilate(det, p):
    return -1 if _popcount(det & ((1 << p) - 1)) & 1 else 1


def _apply_h(det, hso, gso, nso):
    """Return {det': amplitude} of H|det> for one occupation bitstring."""
    out = {}
    occ = [p for p in range(nso) if det >> p & 1]
    # one-body
    for q in occ:
        s1 = _sign_annihilate(det, q)
        d1 = det & ~(1 << q)
        for p in range(nso):
            if d1 >> p & 1:
                continue
            v = hso[p, q]
            if v == 0.0:
                continue
            s2 = _sign_annihilate(d1, p)
            d2 = d1 | (1 << p)
            out[d2] = out.get(d2, 0.0) + s1 * s2 * v
    # two-body: 1/2 <pq|rs> a+p a+q a_s a_r
    for r in occ:
        sr = _sign_annihilate(det, r)
        d1 = det & ~(1 << r)
        for s in occ:
            if s == r:
                continue
            ss = _sign_annihilate(d1, s)
            d2 = d1 & ~(1 << s)
            for q in range(nso):
                if d2 >> q & 1:
                    continue
                sq = _sign_annihilate(d2, q)
                d3 = d2 | (1 << q)
                for p in range(nso):
                    if d3 >> p & 1:
                        continue
                    v = gso[p, q, r, s]
                    if v == 0.0:
                        continue
                    sp = _sign_annihilate(d3, p)
                    d4 = d3 | (1 << p)
                    out[d4] = out.get(d4, 0.0) + 0.5 * sr * ss * sq * sp * v
    return out


class BruteForce:
    def __init__(self, S, h, eri, nalpha, nbeta, enuc=0.0):
        n = S.shape[0]
        w, U = np.linalg.eigh(S)
        X = U @ np.diag(w ** -0.5) @ U.T  # S^-1/2: orthonormal basis chi' = chi X
        self.X = X
        self.Xinv = np.linalg.inv(X)
        ho = X.T @ h @ X
        eo = np.einsum("abcd,ap,bq,cr,ds->pqrs", eri, X, X, X, X)
        nso = 2 * n
        hso = np.zeros((nso, nso))
        gso = np.zeros((nso, nso, nso, nso))
        for sp in range(2):
            o = sp * n
            hso[o:o + n, o:o + n] = ho
        # <pq|rs> = (pr|qs) with spin(p)=spin(r), spin(q)=spin(s)
        for s1 in range(2):
            for s2 in range(2):
                a, b = s1 * n, s2 * n
                gso[a:a + n, b:b + n, a:a + n, b:b + n] = eo.transpose(0, 2, 1, 3)
        self.n, self.nso = n, nso
        self.enuc = enuc
        self.strings_a = list(combinations(range(n), nalpha))
        self.strings_b = list(combinations(range(n), nbeta))
        dets = []
        for ia in self.strings_a:
            for ib in self.strings_b:
                d = 0
                for p in ia:
                    d |= 1 << p
                for p in ib:
                    d |= 1 << (n + p)
                dets.append(d)
        self.dets = dets
        self.index = {d: i for i, d in enumerate(dets)}
        N = len(dets)
        H = np.zeros((N, N))
        for j, d in enumerate(dets):
            for d2, v in _apply_h(d, hso, gso, nso).items():
                H[self.index[d2], j] += v
        self.H = H + enuc * np.eye(N)

    def vector(self, Ca, Cb, occa_cols, occb_cols):
        """Coefficients of a determinant built from MO columns (creation order =
        the given column order, alpha operators first)."""
        Ma = self.Xinv @ Ca[:, list(occa_cols)]  # coefficients in orthonormal basis
        Mb = self.Xinv @ Cb[:, list(occb_cols)]
        ca = np.array([np.linalg.det(Ma[list(s), :]) for s in self.strings_a])
        cb = np.array([np.linalg.det(Mb[list(s), :]) for s in self.strings_b])
        return np.kron(ca, cb)

    def hs(self, va, vb):
        return float(va @ self.H @ vb), float(va @ vb)
