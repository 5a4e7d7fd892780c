"""The oracle's reference-arithmetic m > 0 path, pinned reading by reading.

The oracle reproduces the reference's m > 0 evaluator
(src/nonorthogonalwicks/eval/preparehamiltonianoverlap.rs:20550-20966) as the
exact extended-Wick element plus five literal readings, each behind its own
switch (RD_* in oracle/nwick_oracle.c; mode 0 = all of them, mode 1 = none).
Here:

- no reading applied equals the brute-force second-quantised expansion;
- each reading applied alone changes H by the closed form written from the
  Rust text in tests/m_readings.py (contraction ranks 1 to 3 per spin,
  m = 1 and m = 2 reference pairs); the skip readings also on slab blocks
  with one contraction entry set to zero, which makes a chosen minor or
  distribution vanish; the
  two index-order readings change nothing (their closed form is zero by a
  symmetry of the distribution sum), and planted index mistakes show that
  the check would see a wrong one;
- the first-minor skip also fires on unmodified blocks (rank-deficient D1)
  and changes H by its closed form there;
- every reading on equals mode 0 bit for bit, and every combination of
  readings matches a numpy restatement of the whole element from the Rust
  text, which with no reading equals brute force.
"""
import copy

import numpy as np
import pytest

import m_readings as R
import oracle as O
from brute import BruteForce
from nwick import basis, synth


def _cols(mask):
    return [i for i in range(128) if mask >> i & 1]


def _system(orth_count, seed):
    n, na, nb = (8, 3, 3) if orth_count == 1 else (8, 3, 2)
    sysm = synth.make_system(n, na, nb, 2, seed=seed, orth_refs=(1,), orth_count=orth_count,
                             orthonormal_ao=False, n_aux=2 * n, enuc=1.7)
    slab = O.Slab(sysm, tol=1e-8)
    assert slab.m[(0, 1)] == (orth_count, 0)
    dets = basis.excited_basis(2, n, na, nb, per_ref=None, include_refs=True)
    return sysm, slab, dets


def _pairs(dets, la_wanted, count, seed):
    """(bra from parent 0, ket from parent 1) with alpha contraction rank in
    la_wanted, a spread of beta ranks."""
    rng = np.random.default_rng(seed)
    bras = [d for d in dets if d.parent == 0]
    kets = [d for d in dets if d.parent == 1]
    out = []
    while len(out) < count:
        a, b = bras[rng.integers(len(bras))], kets[rng.integers(len(kets))]
        if a.rank_a + b.rank_a in la_wanted:
            out.append((a, b))
    return out


@pytest.fixture(scope="module", params=[1, 2], ids=["m1", "m2"])
def case(request):
    return _system(request.param, seed=70 + request.param)


def test_no_reading_equals_brute_force(case):
    sysm, slab, dets = case
    bf = BruteForce(sysm.S, sysm.h, sysm.eri, sysm.nocc_a, sysm.nocc_b)
    for a, b in _pairs(dets, (0, 1, 2, 3, 4), 40, seed=1):
        va = bf.vector(*sysm.refs[a.parent], _cols(a.oa), _cols(a.ob))
        vb = bf.vector(*sysm.refs[b.parent], _cols(b.oa), _cols(b.ob))
        hb, sb = bf.hs(va, vb)
        h, s = slab.hs_readings(a, b, 0)
        assert abs(h - (hb + sysm.enuc * sb)) < 1e-10 and abs(s - sb) < 1e-10
        if a.parent != b.parent:
            hn, sn = R.element(slab.pair_view(0, 1), sysm.nmo, a, b, 0, sysm.enuc, slab.tol)
            assert abs(hn - h) < 1e-10 and abs(sn - s) < 1e-10


@pytest.mark.parametrize("la", [1, 2])
def test_v0_reading_closed_form(case, la):
    sysm, slab, dets = case
    view = slab.pair_view(0, 1)
    worst, seen = 0.0, 0.0
    for a, b in _pairs(dets, (la,), 30, seed=10 * la + 1):
        h0, s0 = slab.hs_readings(a, b, 0)
        h1, s1 = slab.hs_readings(a, b, O.RD_V0)
        assert s1 == s0
        want = R.same_spin_change(view, sysm.nmo, a, b, "v0")
        worst = max(worst, abs((h1 - h0) - want))
        seen = max(seen, abs(want))
    assert worst < 1e-11
    # the reading matters on these elements (a closed form of zero pins nothing)
    assert seen > 1e-4


@pytest.mark.parametrize("which,flag,plant", [("vidx", O.RD_VIDX, O.RD_PLANT_VT),
                                              ("jslot", O.RD_JSLOT, O.RD_PLANT_JCOL)])
@pytest.mark.parametrize("la", [2, 3])
def test_index_order_readings_are_inert(case, which, flag, plant, la):
    """The index orders v[m1][m2][mk] and jslot(m1, m2, mk, mj) change no
    element: exchanging the operator assignment m2 with the replaced column's
    assignment mk maps the distributions onto themselves, and the cofactors
    of column k do not involve column k, so the sum over distributions is the
    same as with the stored order.  The closed form of the change is zero;
    the oracle's change is zero to rounding, and a planted index mistake in
    the oracle (an untransposed one-column read, exchanged J column labels) moves
    H by far more, so the check would see one."""
    sysm, slab, dets = case
    view = slab.pair_view(0, 1)
    planted = 0.0
    for a, b in _pairs(dets, (la,), 25, seed=10 * la + flag):
        h0, _ = slab.hs_readings(a, b, 0)
        h1, _ = slab.hs_readings(a, b, flag)
        want = R.same_spin_change(view, sysm.nmo, a, b, which)
        scale = max(1.0, abs(h0))
        assert abs(want) < 1e-13 * scale
        assert abs((h1 - h0) - want) < 1e-12 * scale
        hp, _ = slab.hs_readings(a, b, flag | plant)
        planted = max(planted, abs(hp - h0))
    assert planted > 1e-4


def test_enuc_reading(case):
    sysm, slab, dets = case
    for a, b in _pairs(dets, (0, 1, 2), 20, seed=5):
        h0, s0 = slab.hs_readings(a, b, 0)
        h1, _ = slab.hs_readings(a, b, O.RD_ENUC)
        assert abs((h1 - h0) + sysm.enuc * s0) < 1e-12


@pytest.mark.parametrize("la", [2, 3, 4])
def test_minor_skip_reading_on_generic_input(case, la):
    """Reading `minor` fires on unmodified blocks: the m-rank D1 has rows
    that vanish to rounding (a ket hole that is not a zero-overlap orbital
    meets the rank-m co-density), so first minors built on D1 columns are
    ~1e-16 and the reference drops their two-column terms, which do not
    shrink with the minor.  The change equals the closed form."""
    sysm, slab, dets = case
    view = slab.pair_view(0, 1)
    seen = 0.0
    for a, b in _pairs(dets, (la,), 20, seed=40 + la):
        h0, _ = slab.hs_readings(a, b, 0)
        h1, _ = slab.hs_readings(a, b, O.RD_MINOR)
        want = R.minor_skip_change(view, sysm.nmo, a, b, slab.tol)
        assert abs((h1 - h0) - want) < 1e-11 * max(1.0, abs(h0))
        seen = max(seen, abs(want))
    if la <= 3:
        assert seen > 1e-4


READING_SETS = [0, O.RD_V0, O.RD_VIDX, O.RD_JSLOT, O.RD_DIST, O.RD_MINOR, O.RD_ENUC,
                O.RD_JSLOT | O.RD_MINOR, O.RD_VIDX | O.RD_DIST, O.RD_ALL] + \
               [O.RD_ALL ^ f for f in (O.RD_V0, O.RD_VIDX, O.RD_JSLOT, O.RD_DIST, O.RD_MINOR, O.RD_ENUC)]


def test_reading_combinations_vs_restatement(case):
    """Every reading, alone and in combination (mode 0 = all of them, bit for
    bit), against the numpy restatement of the whole m > 0 element written
    from the Rust text: the skips interact with the index orders (a dropped
    minor carries the J slot of the order in force)."""
    sysm, slab, dets = case
    view = slab.pair_view(0, 1)
    for a, b in _pairs(dets, (1, 2, 3, 4), 6, seed=6):
        assert slab.hs_readings(a, b, O.RD_ALL) == slab.hs(a, b, 0)
        assert slab.hs_readings(a, b, 0) == slab.hs(a, b, 1)
        for rd in READING_SETS:
            h, s = slab.hs_readings(a, b, rd)
            hn, sn = R.element(view, sysm.nmo, a, b, rd, sysm.enuc, slab.tol)
            assert abs(h - hn) < 1e-11 * max(1.0, abs(hn)) and abs(s - sn) < 1e-12, rd


def _zero_entry(slab, x, w, spin, which, r, c):
    """Copy of the oracle slab with one X0 / Y0 entry of pair (x, w) set to zero."""
    out = copy.deepcopy(slab)
    off = out.off[(x, w)]
    o = (R.O_BB if spin else 0) + (R.O_X if which == "X" else R.O_Y)
    out.slab[out.base[(x, w)] + off[o] + r * slab.nmo + c] = 0.0
    return out


def test_minor_skip_reading():
    """An alpha double on the bra against the ket reference (L_alpha = 2) with
    X0[rows[1], cols[0]] = 0: the reference skips the vanishing first minor."""
    sysm, slab, dets = _system(1, seed=71)
    n, na = sysm.nmo, sysm.nocc_a
    bras = [d for d in dets if d.parent == 0 and d.rank_a == 2 and d.rank_b <= 1][:6]
    kets = [d for d in dets if d.parent == 1 and d.rank_a == 0 and d.rank_b <= 1][:3]
    for a in bras:
        rows = [p - na for p in R.bits_of(a.pa)]
        cols = R.bits_of(a.ha)
        mod = _zero_entry(slab, 0, 1, 0, "X", rows[1], cols[0])
        for b in kets:
            h0, _ = mod.hs_readings(a, b, 0)
            h1, _ = mod.hs_readings(a, b, O.RD_MINOR)
            # the planted zero's own skipped term, and the whole change
            assert abs(R.minor_drop_change_l2(mod.pair_view(0, 1), n, a, b)) > 1e-7
            want = R.minor_skip_change(mod.pair_view(0, 1), n, a, b, slab.tol)
            assert abs((h1 - h0) - want) < 1e-11


def test_distribution_skip_reading():
    """An alpha single on the bra against the ket reference (L_alpha = 1,
    m_alpha = 1) with X0[rows[0], cols[0]] = 0: the reference skips the
    singular distributions in the same-spin and different-spin sums and keeps
    their overlap."""
    sysm, slab, dets = _system(1, seed=71)
    n, na = sysm.nmo, sysm.nocc_a
    bras = [d for d in dets if d.parent == 0 and d.rank_a == 1 and d.rank_b <= 1][:6]
    kets = [d for d in dets if d.parent == 1 and d.rank_a == 0 and d.rank_b <= 1][:3]
    for a in bras:
        r0, c0 = R.bits_of(a.pa)[0] - na, R.bits_of(a.ha)[0]
        mod = _zero_entry(slab, 0, 1, 0, "X", r0, c0)
        for b in kets:
            h0, s0 = mod.hs_readings(a, b, 0)
            h1, s1 = mod.hs_readings(a, b, O.RD_DIST)
            want = R.dist_drop_change_l1(mod.pair_view(0, 1), n, a, b)
            assert s1 == s0
            assert abs(want) > 1e-7
            assert abs((h1 - h0) - want) < 1e-11
            assert slab.hs_readings(a, b, O.RD_DIST) == slab.hs_readings(a, b, 0)
