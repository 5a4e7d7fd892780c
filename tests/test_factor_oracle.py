"""CPU tests of the spin-factorised operator restatement (oracle/spin_factor.py).

- The per-spin tables multiply back to every pair's F and S of the full
  oracle evaluator (ow_fs), in both arithmetic modes and for m = 0 and m > 0
  reference pairs, with the earlier parent as bra (ordered_parent_pair).
- The factorised y = (F + lambda S) x equals the dense operator of the
  reference's full-M path (upper triangle evaluated, mirrored) when every
  reference pair has m = 0.
- Component numbering: one id per distinct excitation string of a parent,
  shared by exactly the determinants carrying it.
"""
import numpy as np
import pytest

import oracle as O
import spin_factor as OF
from nwick import basis, synth


def _system(orth):
    sysm = synth.make_system(10, 3, 2, 3, seed=7, orth_refs=orth, orthonormal_ao=False, n_aux=20)
    rng = np.random.default_rng(1)
    fa = rng.normal(size=(10, 10)); fa = fa + fa.T
    fb = rng.normal(size=(10, 10)); fb = fb + fb.T
    slab = O.Slab(sysm)
    slab.update_fock(sysm, fa, fb)
    dets = basis.excited_basis(3, 10, 3, 2, per_ref=14, seed=5, include_refs=True)
    return sysm, slab, dets


@pytest.mark.parametrize("orth", [(), (2,)])
@pytest.mark.parametrize("mode", [0, 1])
def test_tables_reproduce_pair_elements(orth, mode):
    sysm, slab, dets = _system(orth)
    fac = OF.SpinFactorisation(dets, sysm.nref)
    tabs = OF.tables(slab, dets, fac, mode)
    if orth:
        assert any(m != (0, 0) for m in slab.m.values())
    worst = 0.0
    for q in range(sysm.nref):
        for p in range(sysm.nref):
            F, S = OF.block(fac, tabs, q, p, 0.0)
            for r, i in enumerate(fac.members[q]):
                for c, j in enumerate(fac.members[p]):
                    bra, ket = (dets[i], dets[j]) if fac.target_left(q, p) else (dets[j], dets[i])
                    f, s = slab.fs(bra, ket, mode)
                    scale = max(1.0, abs(f))
                    worst = max(worst, abs(F[r, c] - f) / scale, abs(S[r, c] - s))
    assert worst < 1e-12, worst


def test_apply_equals_full_operator_m0():
    sysm, slab, dets = _system(())
    assert all(m == (0, 0) for m in slab.m.values())
    fac = OF.SpinFactorisation(dets, sysm.nref)
    tabs = OF.tables(slab, dets, fac, 0)
    F, S = O.fs_matrix(slab, dets, 0)
    rng = np.random.default_rng(3)
    x = rng.normal(size=len(dets))
    lam = -1.37
    y, ys = OF.apply(fac, tabs, lam, x)
    yr = (F + lam * S) @ x
    assert np.max(np.abs(y - yr)) < 1e-10 * max(1.0, np.max(np.abs(yr)))
    assert np.max(np.abs(ys - S @ x)) < 1e-12


def test_component_numbering():
    sysm, slab, dets = _system(())
    fac = OF.SpinFactorisation(dets, sysm.nref)
    for p in range(sysm.nref):
        mem = fac.members[p]
        for spin, ids in ((0, fac.aid), (1, fac.bid)):
            strings = {}
            for i in mem:
                d = dets[i]
                key = (d.ha, d.pa) if spin == 0 else (d.hb, d.pb)
                strings.setdefault(key, set()).add(int(ids[i]))
            assert all(len(v) == 1 for v in strings.values())
            assert len({next(iter(v)) for v in strings.values()}) == len(strings) == fac.ncomp(p, spin)
        # ranks ascend with the component id
        ranks = [bin(dets[r].ha).count("1") for r in fac.reps[p][0]]
        assert ranks == sorted(ranks)
