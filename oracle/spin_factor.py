"""Spin-factorised one-body / overlap operator, CPU restatement.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the engine's
nw_factor_* path.  The product never uses it.

Follows src/noci/factorise (noci-rs):
- spin_factorise: SpinFactorisation::new (mod.rs:75-93) -- per parent the
  unique alpha and beta excitation strings and every determinant's
  (parent, a, b) (assign_aids / assign_bids / assign_spin_ids :185-278,
  build_parent_spin_spaces :315-410).  The reference numbers the strings in
  the order of its sort key (occupation, hole mask, particle mask, phase);
  the engine numbers them by (rank, holes, particles, phase), which this
  restatement uses too.  Only the numbering differs: the components and the
  operator are the same.
- ordered_parent_pair (mod.rs:163-183): the bra of a parent pair's factors is
  the parent whose first determinant comes first.
- factor_table: build_spin_one_body_factor_row (onebody.rs:1300-1569), one
  spin's {S, F} between a target and a source string (oracle Slab.fs_spin).
- apply: apply_one_body (onebody.rs:290-345) as the sum over parent pairs of
  F^a S^b + S^a F^b + lambda S^a S^b, evaluated densely per block.
"""
from __future__ import annotations

import numpy as np


def _bits(v: int):
    out, k = [], 0
    while v:
        if v & 1:
            out.append(k)
        v >>= 1
        k += 1
    return out


def spin_key(det, spin):
    h, p = (det.ha, det.pa) if spin == 0 else (det.hb, det.pb)
    ph = det.pha if spin == 0 else det.phb
    hs, ps = _bits(h), _bits(p)
    return (len(hs), tuple(hs + [0] * (4 - len(hs))), tuple(ps + [0] * (4 - len(ps))), 1 if ph < 0 else 0)


class SpinFactorisation:
    """Per parent: unique alpha / beta strings (a representative determinant
    each), first determinant, and the (det, a, b) entries in determinant order."""

    def __init__(self, dets, nref):
        self.nref = nref
        self.aid = np.zeros(len(dets), dtype=np.int64)
        self.bid = np.zeros(len(dets), dtype=np.int64)
        self.reps = [[[], []] for _ in range(nref)]
        self.first = [None] * nref
        self.members = [[] for _ in range(nref)]
        for i, d in enumerate(dets):
            self.members[d.parent].append(i)
            if self.first[d.parent] is None:
                self.first[d.parent] = i
        for p in range(nref):
            for spin, ids in ((0, self.aid), (1, self.bid)):
                keyed = sorted((spin_key(dets[i], spin), i) for i in self.members[p])
                last = None
                for key, i in keyed:
                    if key != last:
                        self.reps[p][spin].append(i)
                        last = key
                    ids[i] = len(self.reps[p][spin]) - 1

    def ncomp(self, p, spin):
        return len(self.reps[p][spin])

    def target_left(self, q, p):
        return self.first[q] <= self.first[p]


def factor_table(slab, dets, fac, q, p, spin, mode=0):
    """[source component, target component, {S, F}] for target parent q, source parent p."""
    tl = fac.target_left(q, p)
    tr, sr = fac.reps[q][spin], fac.reps[p][spin]
    out = np.zeros((len(sr), len(tr), 2))
    for si, s in enumerate(sr):
        for ti, t in enumerate(tr):
            bra, ket = (dets[t], dets[s]) if tl else (dets[s], dets[t])
            out[si, ti] = slab.fs_spin(bra, ket, spin, mode)
    return out


def tables(slab, dets, fac, mode=0):
    return {(q, p, sp): factor_table(slab, dets, fac, q, p, sp, mode)
            for q in range(fac.nref) for p in range(fac.nref) for sp in (0, 1)
            if fac.members[q] and fac.members[p]}


def block(fac, tabs, q, p, lam):
    """Dense (F + lam S) and S blocks rows = dets of q, cols = dets of p."""
    rows, cols = fac.members[q], fac.members[p]
    A = tabs[(q, p, 0)][np.ix_(fac.aid[cols], fac.aid[rows])].transpose(1, 0, 2)
    B = tabs[(q, p, 1)][np.ix_(fac.bid[cols], fac.bid[rows])].transpose(1, 0, 2)
    Sa, Fa, Sb, Fb = A[..., 0], A[..., 1], B[..., 0], B[..., 1]
    return Fa * Sb + Sa * Fb + lam * Sa * Sb, Sa * Sb


def apply(fac, tabs, lam, x):
    """y = (F + lam S) x and ys = S x through the per-spin factors."""
    y = np.zeros(len(x))
    ys = np.zeros(len(x))
    for q in range(fac.nref):
        for p in range(fac.nref):
            if not fac.members[q] or not fac.members[p]:
                continue
            M, S = block(fac, tabs, q, p, lam)
            xs = x[fac.members[p]]
            y[fac.members[q]] += M @ xs
            ys[fac.members[q]] += S @ xs
    return y, ys
