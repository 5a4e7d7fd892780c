"""Generalised Slater-Condon route on the GPU (nw_hs_naive_batched) and the
wicks.compare cross-check (nw_hs_compare), through the C ABI.

The oracle's generalised Slater-Condon route (ow_hs_naive, a restatement of
src/noci/naive.rs and src/noci/hs.rs:372-402) is pinned against brute force in
tests/test_oracle.py; the GPU route is compared with it per element at 1e-10
(H in Ha and S), on pairs with no, one and two zero-overlap orbitals, and at
n_bas = 36 where the integral contractions run on MFMA.  The compare entry
point's discrepancy is checked against the same sum formed from the oracle.
"""
import numpy as np
import pytest

import nwick
import oracle as O
from nwick import basis, synth

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _case(n, na, nb, nref, seed, orth_refs=(), orth_count=1, orthonormal=False, per_ref=12):
    sysm = synth.make_system(n, na, nb, nref, seed=seed, orth_refs=orth_refs, orth_count=orth_count,
                             orthonormal_ao=orthonormal, n_aux=2 * n)
    dets = basis.excited_basis(nref, n, na, nb, per_ref=per_ref, seed=seed, include_refs=True)
    return sysm, dets


def _oracle_naive(sysm, dets, req):
    ea = sysm.eri_asym()
    out = [O.hs_naive(sysm, dets[a], dets[b], eri_asym=ea) for a, b in req]
    return np.array([h for h, _ in out]), np.array([s for _, s in out])


@pytest.mark.parametrize("orth_count,orthonormal", [(0, False), (1, False), (2, True)])
def test_naive_batched_vs_oracle(orth_count, orthonormal):
    """Pairs with m = 0 (every pair), 1 and 2 zero-overlap alpha orbitals
    (reference 2 rotated against references 0 and 1)."""
    orth = (2,) if orth_count else ()
    sysm, dets = _case(10, 3, 2, 3, seed=31 + orth_count, orth_refs=orth, orth_count=max(1, orth_count),
                       orthonormal=orthonormal)
    rec = basis.to_records(dets)
    rng = np.random.default_rng(5)
    req = rng.integers(0, len(dets), size=(300, 2)).astype(np.uint64)
    print(f"naive: n=10 3+2 e, orth_count {orth_count}, ndet {len(dets)}, {len(req)} requests", flush=True)
    eng = nwick.Engine.from_system(sysm)
    h, s = eng.hs_naive_batched(rec, req)
    ho, so = _oracle_naive(sysm, dets, req)
    assert np.max(np.abs(h - ho)) < TOL and np.max(np.abs(s - so)) < TOL
    if orth_count:
        # pairs across the rotated reference that keep a zero-overlap orbital:
        # S = 0 exactly, H from the P_i co-densities
        cross = [k for k, (a, b) in enumerate(req) if dets[a].parent != 2 and dets[b].parent == 2]
        zero = [k for k in cross if so[k] == 0.0]
        assert zero and np.max(np.abs(s[zero])) == 0.0
        assert np.max(np.abs(h[zero])) > 1e-6
    eng.close()


def test_naive_batched_mfma_size():
    """n_bas = 36 (n^2 = 1296: the contractions take the MFMA product kernel), m = 1 pairs included."""
    sysm, dets = _case(36, 3, 2, 2, seed=41, orth_refs=(1,), orth_count=1, per_ref=6)
    rec = basis.to_records(dets)
    rng = np.random.default_rng(6)
    req = rng.integers(0, len(dets), size=(40, 2)).astype(np.uint64)
    eng = nwick.Engine.from_system(sysm)
    h, s = eng.hs_naive_batched(rec, req)
    ho, so = _oracle_naive(sysm, dets, req)
    print(f"naive n=36: max|dH| {np.max(np.abs(h - ho)):.2e}", flush=True)
    assert np.max(np.abs(h - ho)) < TOL and np.max(np.abs(s - so)) < TOL
    eng.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_compare_matches_oracle_discrepancy(mode):
    """wicks.compare: H, S equal nw_hs_matrix's; with m > 0 pairs the
    reference arithmetic (mode 0) departs from the naive route and the
    corrected one (mode 1) does not.  total and max are the sums over the
    upper triangle of |H_wick - H_naive| + |S_wick - S_naive| and the largest
    single difference, formed here from the oracle's elements."""
    sysm, dets = _case(10, 3, 2, 3, seed=33, orth_refs=(2,), orth_count=1, per_ref=8)
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    eng.build(tol=1e-8)
    H, S, tot, mx = eng.hs_compare(rec, mode=mode)
    Hm, Sm = eng.hs_matrix(rec, mode=mode)
    assert np.array_equal(H, Hm) and np.array_equal(S, Sm)
    n = len(dets)
    req = np.array([(i, j) for i in range(n) for j in range(i, n)], dtype=np.uint64)
    hn, sn = eng.hs_naive_batched(rec, req)
    dh = np.abs(hn - H[req[:, 0].astype(int), req[:, 1].astype(int)])
    ds = np.abs(sn - S[req[:, 0].astype(int), req[:, 1].astype(int)])
    print(f"compare mode {mode}: total {tot:.3e} max {mx:.3e}", flush=True)
    assert tot == pytest.approx(float(np.sum(dh + ds)), rel=1e-12, abs=1e-14)
    assert mx == pytest.approx(float(max(dh.max(), ds.max())), rel=1e-12, abs=1e-14)
    # the oracle's Wick elements against its naive ones give the same discrepancy
    slab = O.Slab(sysm, tol=1e-8)
    ea = sysm.eri_asym()
    sub = req[:: max(1, len(req) // 150)]
    do = []
    for a, b in sub:
        hw, sw = slab.hs(dets[int(a)], dets[int(b)], mode)
        hq, sq = O.hs_naive(sysm, dets[int(a)], dets[int(b)], eri_asym=ea)
        do.append(max(abs(hw - hq), abs(sw - sq)))
    dg = np.maximum(dh, ds)[:: max(1, len(req) // 150)]
    assert np.max(np.abs(dg - np.array(do))) < 1e-9
    if mode == 1:
        assert mx < 1e-9
    else:
        assert mx > 1e-6  # the literal m > 0 arithmetic differs from the naive route
    eng.close()


def test_naive_refuses_wrong_occupation_count():
    sysm, dets = _case(8, 2, 2, 2, seed=3, per_ref=4)
    rec = basis.to_records(dets)
    assert dets[0].oa == 0b11  # the first reference: alpha orbitals 0 and 1
    rec["oa"][0, 0] |= np.uint64(1 << 7)  # a third alpha electron
    eng = nwick.Engine.from_system(sysm)
    with pytest.raises(nwick.NwickError) as e:
        eng.hs_naive_batched(rec, np.array([[1, 0]], dtype=np.uint64))
    assert e.value.code == 1
    eng.close()
