"""Contraction ranks above four per spin (triple and quadruple excitations)
on the GPU's run-time-rank evaluators, against the oracle (whose cofactors
and minors come from explicit LU determinants, as the reference's generic
paths; pinned for these ranks against brute force and the generalised
Slater-Condon route in tests/test_oracle.py), at 1e-10 per element: H, S and
F, S matrices and batched requests in both arithmetic modes, m = 0 and m = 1
reference pairs, per-spin overlap factors and the PT2 matvec."""
import numpy as np
import pytest

import nwick
import oracle as O
from nwick import basis, synth

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def rank_case():
    n, na, nb = 12, 4, 4
    sysm = synth.make_system(n, na, nb, 3, seed=123, orth_refs=(2,), orth_count=1, orthonormal_ao=False, n_aux=24)
    counts = {"a": 2, "b": 2, "aa": 3, "ab": 3, "aaa": 4, "aab": 4, "abb": 3, "bbb": 3, "aaaa": 2, "aabb": 2,
              "abbb": 2}
    dets = basis.excited_basis(3, n, na, nb, per_ref=0, seed=123, include_refs=True, counts=counts)
    print(f"rank case: n={n} nocc={na}+{nb} ndet={len(dets)} (ranks up to 4 per determinant)", flush=True)
    eng = nwick.Engine.from_system(sysm)
    meta = eng.build()
    assert meta[(0, 2)].ma == 1
    slab = O.Slab(sysm)
    rng = np.random.default_rng(1)
    fa = rng.normal(size=(n, n)); fa = 0.5 * (fa + fa.T)
    fb = rng.normal(size=(n, n)); fb = 0.5 * (fb + fb.T)
    eng.update_fock(fa, fb)
    slab.update_fock(sysm, fa, fb)
    return sysm, dets, eng, slab


@pytest.mark.parametrize("mode", [0, 1])
def test_hs_matrix_high_rank(rank_case, mode):
    sysm, dets, eng, slab = rank_case
    rec = basis.to_records(dets)
    H, S = eng.hs_matrix(rec, mode=mode)
    Ho, So = O.hs_matrix(slab, dets, mode=mode)
    la = np.array([[a.rank_a + b.rank_a for b in dets] for a in dets])
    assert la.max() == 8
    dh = float(np.max(np.abs(H - Ho)))
    print(f"mode {mode}: max|dH| {dh:.3e} over {len(dets) ** 2} elements, max|H| {np.max(np.abs(Ho)):.2f}", flush=True)
    assert np.max(np.abs(S - So)) < TOL and dh < TOL * max(1.0, np.max(np.abs(Ho)))


@pytest.mark.parametrize("mode", [0, 1])
def test_fs_matrix_and_matvec_high_rank(rank_case, mode):
    sysm, dets, eng, slab = rank_case
    rec = basis.to_records(dets)
    F, S = eng.fs_matrix(rec, mode=mode)
    Fo, So = O.fs_matrix(slab, dets, mode=mode)
    assert np.max(np.abs(S - So)) < TOL and np.max(np.abs(F - Fo)) < TOL * max(1.0, np.max(np.abs(Fo)))
    x = np.random.default_rng(3).normal(size=len(dets))
    y = eng.matvec_m(rec, -1.5, x, mode=mode)
    assert np.max(np.abs(y - (Fo + 1.5 * So) @ x)) < 1e-9


def test_batched_and_spin_overlaps_high_rank(rank_case):
    sysm, dets, eng, slab = rank_case
    rec = basis.to_records(dets)
    rng = np.random.default_rng(5)
    req = rng.integers(0, len(dets), size=(3000, 2)).astype(np.uint64)
    sa = eng.s_same_batched(rec, req, 0)
    sb = eng.s_same_batched(rec, req, 1)
    for mode in (0, 1):
        h, s = eng.hs_batched(rec, req, mode=mode)
        f, s2 = eng.fs_batched(rec, req, mode=mode)
        for q in range(0, len(req), 4):
            a, b = dets[int(req[q, 0])], dets[int(req[q, 1])]
            if a.parent == b.parent and bin(a.oa ^ b.oa).count("1") + bin(a.ob ^ b.ob).count("1") > 4:
                assert h[q] == 0.0 and s[q] == 0.0
            else:
                ho, so = slab.hs(a, b, mode)
                assert abs(h[q] - ho) < TOL * max(1.0, abs(ho)) and abs(s[q] - so) < TOL
            fo, so2 = slab.fs(a, b, mode)
            assert abs(f[q] - fo) < TOL * max(1.0, abs(fo)) and abs(s2[q] - so2) < TOL
            if mode == 0:
                assert abs(sa[q] - slab.fs_spin(a, b, 0, 0)[0]) < TOL
                assert abs(sb[q] - slab.fs_spin(a, b, 1, 0)[0]) < TOL


@pytest.mark.parametrize("mode", [0, 1])
def test_pt2_with_triples_uses_pairwise_matvec(mode):
    """A PT2 candidate space with triple excitations: the factor tables are
    refused (a triple against a double is contraction rank 5 on one spin),
    the driver takes the pairwise matvec, and E_PT2 equals a dense solve
    formed from the oracle's elements."""
    from nwick.pt2 import PT2
    from pt2_dense import dense_pt2
    n, na, nb = 8, 3, 3
    sysm = synth.make_system(n, na, nb, 2, seed=41, orth_refs=(1,), orth_count=1, orthonormal_ao=False, n_aux=16)
    refs = basis.excited_basis(2, n, na, nb, per_ref=0, include_refs=True, counts={})
    counts = {"a": 5, "b": 5, "aa": 5, "ab": 8, "bb": 5, "aaa": 3, "aab": 3, "abb": 2}
    cands = basis.excited_basis(2, n, na, nb, per_ref=0, seed=41, include_refs=False, counts=counts)
    print(f"pt2 triples: mode {mode} ncur={len(refs)} ncand={len(cands)}", flush=True)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    pt = PT2(eng, sysm, refs, cands, mode=mode, factor_tables="device")
    pt.setup()
    assert pt.operator == "pairwise"
    r = pt.solve(res_tol=1e-11, restart=len(cands) + 10, max_iter=4 * len(cands))
    dn = dense_pt2(sysm, refs, cands, mode=mode)
    print(f"pt2 triples: gpu {r['ept2']:.12f} dense {dn['ept2']:.12f} converged {r['converged']}", flush=True)
    assert r["converged"]
    assert abs(r["e_noci"] - dn["e_noci"]) < 1e-10
    assert abs(r["ept2"] - dn["ept2"]) < 1e-8
