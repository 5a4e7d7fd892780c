"""Excited-determinant bases and their C-ABI records.

Restates the parts of src/basis/excitation.rs that fix the determinant-pair
index lists bit-exactly: generation order (excitation.rs:276-369), the
fermionic excitation phase (excitation.rs:79-116) and the excitation masks
relative to the parent reference (excitation.rs:232-246).
"""
from __future__ import annotations

from itertools import combinations

import numpy as np

# Mirrors `struct nw_det` in include/nwick.h.
DET_DTYPE = np.dtype([
    ("parent", "<u4"), ("flags", "<u4"),
    ("pha", "<f8"), ("phb", "<f8"),
    ("oa", "<u8", 2), ("ob", "<u8", 2),
    ("ha", "<u8", 2), ("pa", "<u8", 2),
    ("hb", "<u8", 2), ("pb", "<u8", 2),
], align=True)
assert DET_DTYPE.itemsize == 120


def excitation_phase_bits(occ: int, holes: int, parts: int) -> float:
    """excitation.rs:79-116: remove holes highest first, add particles lowest
    first, each contributing the parity of occupied orbitals below it."""
    odd = 0
    while holes:
        i = holes.bit_length() - 1
        odd ^= bin(occ & ((1 << i) - 1)).count("1") & 1
        occ &= ~(1 << i)
        holes &= ~(1 << i)
    while parts:
        a = (parts & -parts).bit_length() - 1
        odd ^= bin(occ & ((1 << a) - 1)).count("1") & 1
        parts &= parts - 1
        if parts:
            occ |= 1 << a
    return -1.0 if odd else 1.0


def _mask(idx) -> int:
    m = 0
    for i in idx:
        m |= 1 << int(i)
    return m


def _split(v: int):
    return (v & 0xFFFFFFFFFFFFFFFF, (v >> 64) & 0xFFFFFFFFFFFFFFFF)


class Det:
    __slots__ = ("parent", "oa", "ob", "ha", "pa", "hb", "pb", "pha", "phb")

    def __init__(self, parent, ref_oa, ref_ob, ah=(), ap=(), bh=(), bp=()):
        self.parent = parent
        self.ha, self.pa, self.hb, self.pb = _mask(ah), _mask(ap), _mask(bh), _mask(bp)
        self.oa = (ref_oa & ~self.ha) | self.pa
        self.ob = (ref_ob & ~self.hb) | self.pb
        self.pha = excitation_phase_bits(ref_oa, self.ha, self.pa)
        self.phb = excitation_phase_bits(ref_ob, self.hb, self.pb)

    @property
    def rank_a(self):
        return bin(self.ha).count("1")

    @property
    def rank_b(self):
        return bin(self.hb).count("1")


def to_records(dets) -> np.ndarray:
    out = np.zeros(len(dets), dtype=DET_DTYPE)
    for i, d in enumerate(dets):
        out["parent"][i] = d.parent
        out["pha"][i] = d.pha
        out["phb"][i] = d.phb
        for f in ("oa", "ob", "ha", "pa", "hb", "pb"):
            out[f][i] = _split(getattr(d, f))
    return out


def generation_key(nocc_a, nocc_b, ah, ap, bh, bp):
    """Sort key reproducing generate_excited_basis order within one reference:
    order k ascending, k_alpha ascending, then lexicographic combinations of
    (alpha holes, alpha parts, beta holes, beta parts)."""
    ka = len(ah)
    kb = len(bh)
    return (ka + kb, ka, tuple(ah), tuple(ap), tuple(bh), tuple(bp))


def all_excitations(nmo, nocc_a, nocc_b, orders=(1, 2)):
    """generate_excited_basis (excitation.rs:276) for one occupied-first reference."""
    occ_a, occ_b = range(nocc_a), range(nocc_b)
    vir_a, vir_b = range(nocc_a, nmo), range(nocc_b, nmo)
    out = []
    for k in sorted(set(orders)):
        for ka in range(k + 1):
            kb = k - ka
            for ah in combinations(occ_a, ka):
                for ap in combinations(vir_a, ka):
                    for bh in combinations(occ_b, kb):
                        for bp in combinations(vir_b, kb):
                            out.append((ah, ap, bh, bp))
    return out


def random_excitations(rng, nmo, nocc_a, nocc_b, counts):
    """Sample unique excitations per type; counts = dict with keys
    'a', 'b' (singles), 'aa', 'bb', 'ab' (doubles).  Returned in generation order."""
    seen = set()
    out = []
    occ_a, occ_b = np.arange(nocc_a), np.arange(nocc_b)
    vir_a, vir_b = np.arange(nocc_a, nmo), np.arange(nocc_b, nmo)

    def pick(pool, k):
        return tuple(sorted(int(x) for x in rng.choice(pool, size=k, replace=False)))

    for typ, cnt in counts.items():
        got = 0
        tries = 0
        while got < cnt:
            tries += 1
            if tries > 100 * cnt + 1000:
                raise ValueError(f"cannot draw {cnt} unique {typ} excitations")
            ka = typ.count("a")
            kb = typ.count("b")
            ah = pick(occ_a, ka) if ka else ()
            ap = pick(vir_a, ka) if ka else ()
            bh = pick(occ_b, kb) if kb else ()
            bp = pick(vir_b, kb) if kb else ()
            key = (ah, ap, bh, bp)
            if key in seen:
                continue
            seen.add(key)
            out.append(key)
            got += 1
    out.sort(key=lambda e: generation_key(nocc_a, nocc_b, *e))
    return out


def excited_basis(nref, nmo, nocc_a, nocc_b, per_ref=None, seed=0, include_refs=False,
                  counts=None, orders=(1, 2)):
    """Determinants for every reference, grouped by parent in reference order.

    per_ref=None enumerates all excitations of the given orders; otherwise
    `counts` (or a default mix of 10% alpha singles, 10% beta singles and
    80% doubles split over aa/bb/ab) of random unique excitations per reference."""
    rng = np.random.default_rng(seed)
    ref_oa = (1 << nocc_a) - 1
    ref_ob = (1 << nocc_b) - 1
    dets = []
    for r in range(nref):
        if include_refs:
            dets.append(Det(r, ref_oa, ref_ob))
        if per_ref is None:
            exc = all_excitations(nmo, nocc_a, nocc_b, orders)
        else:
            c = counts
            if c is None:
                ns = per_ref // 10
                nd = per_ref - 2 * ns
                c = {"a": ns, "b": ns, "aa": nd // 3, "bb": nd // 3, "ab": nd - 2 * (nd // 3)}
            exc = random_excitations(rng, nmo, nocc_a, nocc_b, c)
        for ah, ap, bh, bp in exc:
            dets.append(Det(r, ref_oa, ref_ob, ah, ap, bh, bp))
    return dets


def ref_masks(nocc_a, nocc_b):
    return _split((1 << nocc_a) - 1), _split((1 << nocc_b) - 1)
