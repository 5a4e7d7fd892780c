"""CPU tests of the oracle: pinned against mathematics that fixes the answer.

- Jacobi SVD vs numpy/LAPACK singular values and reconstruction.
- Excitation phases: C restatement vs the Python mirror, and (through every
  excitation class) the brute-force determinant expansion.
- Wick H,S vs the independent generalised Slater-Condon route (the
  reference's own *_agree check, tests/noci.rs:303-320) for m = 0 pairs.
- Wick H,S (corrected mode) vs a brute-force second-quantised expansion in
  the full determinant space, for m = 0, 1 and 2 reference pairs.
- The reference-arithmetic mode differs from the corrected one for m > 0
  exactly by the closed form of its v0 double count (la = lb = 0 case).
- Reference-pair overlap equals det(C_x,occ^T S C_w,occ) per spin.
- Hermiticity of H and S.
- Slab block sizes follow assign_offsets (58 nmo^2 + 3 nmo^4 for m = 0).
"""
import numpy as np
import pytest

import oracle as O
from brute import BruteForce
from nwick import basis, synth


def _cols(mask):
    return [i for i in range(128) if mask >> i & 1]


def test_svd_matches_lapack():
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 10, 17):
        A = rng.normal(size=(n, n))
        if n > 3:
            A[:, 2] = A[:, 0] * 0.5 - A[:, 1]  # exactly rank deficient
        U, s, V = O.svd(A)
        ref = np.linalg.svd(A, compute_uv=False)
        assert np.allclose(s, ref, atol=1e-12)
        assert np.allclose(U @ np.diag(s) @ V.T, A, atol=1e-12)
        assert np.allclose(U.T @ U, np.eye(n), atol=1e-12)
        assert np.allclose(V.T @ V, np.eye(n), atol=1e-12)


def test_excitation_phase_c_vs_python():
    rng = np.random.default_rng(1)
    for _ in range(500):
        n = 20
        occ = int(rng.integers(0, 1 << n))
        occ_list = [i for i in range(n) if occ >> i & 1]
        vir_list = [i for i in range(n) if not occ >> i & 1]
        k = int(rng.integers(0, min(len(occ_list), len(vir_list), 4) + 1))
        h = basis._mask(rng.choice(occ_list, k, replace=False)) if k else 0
        p = basis._mask(rng.choice(vir_list, k, replace=False)) if k else 0
        assert O.excitation_phase(occ, h, p) == basis.excitation_phase_bits(occ, h, p)


@pytest.mark.parametrize("orthonormal", [True, False])
def test_wick_vs_generalised_slater_condon_m0(orthonormal):
    sysm = synth.make_system(10, 3, 2, 3, seed=11, orthonormal_ao=orthonormal, n_aux=20)
    slab = O.Slab(sysm, tol=1e-8)
    assert all(v == (0, 0) for v in slab.m.values())
    dets = basis.excited_basis(3, 10, 3, 2, per_ref=None, include_refs=True)
    ea = sysm.eri_asym()
    rng = np.random.default_rng(2)
    for i, j in rng.choice(len(dets), size=(600, 2)):
        hw, sw = slab.hs(dets[i], dets[j], 0)
        hn, sn = O.hs_naive(sysm, dets[i], dets[j], eri_asym=ea)
        assert abs(hw - hn) < 1e-10 and abs(sw - sn) < 1e-10


@pytest.mark.parametrize("seed,orth_count,orthonormal", [(3, 1, False), (4, 2, True), (5, 1, True)])
def test_wick_vs_bruteforce(seed, orth_count, orthonormal):
    sysm = synth.make_system(7, 3, 2, 2, seed=seed, orth_refs=(1,), orth_count=orth_count,
                             orthonormal_ao=orthonormal, n_aux=10)
    bf = BruteForce(sysm.S, sysm.h, sysm.eri, 3, 2)
    slab = O.Slab(sysm, tol=1e-8)
    assert slab.m[(0, 1)] == (orth_count, 0) and slab.m[(1, 0)] == (orth_count, 0)
    assert slab.m[(0, 0)] == (0, 0) and slab.m[(1, 1)] == (0, 0)
    dets = basis.excited_basis(2, 7, 3, 2, per_ref=None, include_refs=True)
    ea = sysm.eri_asym()
    rng = np.random.default_rng(seed)
    vec = {}
    for i, j in rng.choice(len(dets), size=(500, 2)):
        for k in (i, j):
            if k not in vec:
                d = dets[k]
                vec[k] = bf.vector(*sysm.refs[d.parent], _cols(d.oa), _cols(d.ob))
        hb, sb = bf.hs(vec[i], vec[j])
        h1, s1 = slab.hs(dets[i], dets[j], 1)
        hn, sn = O.hs_naive(sysm, dets[i], dets[j], eri_asym=ea)
        assert abs(h1 - hb) < 1e-10 and abs(s1 - sb) < 1e-10
        assert abs(hn - hb) < 1e-10 and abs(sn - sb) < 1e-10
        if slab.m[(dets[i].parent, dets[j].parent)] == (0, 0):
            h0, s0 = slab.hs(dets[i], dets[j], 0)
            assert h0 == h1 and s0 == s1


def test_reference_mode_v0_double_count_closed_form():
    """For reference determinants on an m_alpha = 1 pair, the reference
    arithmetic adds v0[1] for both (0,1) and (1,0) assignments while v0[1]
    already holds both: H_ref - H_exact = pref * S_beta * v0[1] / 2."""
    sysm = synth.make_system(9, 3, 3, 2, seed=7, orth_refs=(1,), orth_count=1, n_aux=20)
    slab = O.Slab(sysm, tol=1e-8)
    dets = basis.excited_basis(2, 9, 3, 3, per_ref=0, include_refs=True, counts={})
    r0, r1 = dets[0], dets[1]
    assert slab.m[(0, 1)] == (1, 0)
    h0, _ = slab.hs(r0, r1, 0)
    h1, _ = slab.hs(r0, r1, 1)
    _, _, meta = slab.pair_view(0, 1)
    pref = meta[1] * meta[0] * meta[12] * meta[11]
    assert abs((h0 - h1) - pref * 1.0 * meta[9] / 2) < 1e-12
    assert abs(h0 - h1) > 1e-4


def test_reference_pair_overlap_closed_form():
    sysm = synth.make_system(12, 4, 3, 3, seed=8, orthonormal_ao=False, n_aux=20)
    slab = O.Slab(sysm, tol=1e-8)
    refs = basis.excited_basis(3, 12, 4, 3, per_ref=0, include_refs=True, counts={})
    for a in refs:
        for b in refs:
            _, s = slab.hs(a, b, 0)
            ca, cb = sysm.refs[a.parent]
            ga, gb = sysm.refs[b.parent]
            ref = (np.linalg.det(ca[:, :4].T @ sysm.S @ ga[:, :4]) * np.linalg.det(cb[:, :3].T @ sysm.S @ gb[:, :3]))
            assert abs(s - ref) < 1e-12


def test_hermiticity():
    sysm = synth.make_system(9, 2, 2, 3, seed=9, orth_refs=(2,), orth_count=1, n_aux=20)
    slab = O.Slab(sysm, tol=1e-8)
    dets = basis.excited_basis(3, 9, 2, 2, per_ref=12, seed=1)
    for mode in (0, 1):
        for a in dets[::3]:
            for b in dets[1::4]:
                hab, sab = slab.hs(a, b, mode)
                hba, sba = slab.hs(b, a, mode)
                assert abs(sab - sba) < 1e-12
                if mode == 1 or slab.m[(a.parent, b.parent)] == (0, 0):
                    assert abs(hab - hba) < 1e-10


def test_slab_block_sizes():
    nmo = 7
    off, size = O.pair_offsets(0, 0, nmo)
    assert size == 58 * nmo ** 2 + 3 * nmo ** 4
    off, size = O.pair_offsets(1, 0, nmo)
    # alpha: slots with branch sum <= 1 (3), beta: 1, II: (a0+aj<=1) x (b0+bj<=0) = 3, no hcol0
    assert size == 56 * nmo ** 2 + 7 * nmo ** 4
    off, size = O.pair_offsets(2, 2, nmo)
    # 7 J slots per spin (branch sum <= 2) and all 16 II blocks
    assert size == 56 * nmo ** 2 + 30 * nmo ** 4


def test_generation_order_matches_reference_enumeration():
    """Random excitation subsets come out in generate_excited_basis order."""
    full = basis.all_excitations(8, 3, 2, orders=(1, 2))
    pos = {e: i for i, e in enumerate(full)}
    rng = np.random.default_rng(3)
    sub = basis.random_excitations(rng, 8, 3, 2, {"a": 3, "b": 3, "aa": 5, "bb": 1, "ab": 6})
    idx = [pos[e] for e in sub]
    assert idx == sorted(idx)


def test_fock_elements_vs_generalised_slater_condon():
    """Generalised-Fock and overlap elements (corrected mode) against the
    one-body Loewdin route on the excited determinants, incl. m = 1 pairs."""
    for orth in ((), (2,)):
        sysm = synth.make_system(10, 3, 2, 3, seed=4, orth_refs=orth, orthonormal_ao=False, n_aux=20)
        rng = np.random.default_rng(0)
        fa = rng.normal(size=(10, 10)); fa = fa + fa.T
        fb = rng.normal(size=(10, 10)); fb = fb + fb.T
        slab = O.Slab(sysm)
        slab.update_fock(sysm, fa, fb)
        dets = basis.excited_basis(3, 10, 3, 2, per_ref=15, seed=2, include_refs=True)
        for i in range(0, len(dets), 3):
            for j in range(0, len(dets), 2):
                f, s = slab.fs(dets[i], dets[j], 1)
                fn, sn = O.fs_naive(sysm, dets[i], dets[j], fa, fb)
                assert abs(f - fn) < 1e-10 and abs(s - sn) < 1e-10
                if slab.m[(dets[i].parent, dets[j].parent)] == (0, 0):
                    assert slab.fs(dets[i], dets[j], 0) == (f, s)


@pytest.mark.parametrize("orth_count", [0, 1])
def test_high_rank_vs_bruteforce(orth_count):
    """Triple excitations (contraction ranks up to 6 per spin): the corrected
    element against brute force and the generalised Slater-Condon route on
    m = 0 and m = 1 reference pairs, and the reference-arithmetic element
    against the numpy restatement of the Rust text, which handles any rank."""
    import m_readings as R
    sysm = synth.make_system(7, 3, 3, 2, seed=17 + orth_count, orth_refs=(1,) if orth_count else (),
                             orth_count=max(orth_count, 1), orthonormal_ao=False, n_aux=14, enuc=0.8)
    bf = BruteForce(sysm.S, sysm.h, sysm.eri, 3, 3)
    slab = O.Slab(sysm, tol=1e-8)
    assert slab.m[(0, 1)] == (orth_count, 0)
    dets = basis.excited_basis(2, 7, 3, 3, per_ref=0, seed=4, include_refs=True,
                               counts={"aaa": 3, "aab": 3, "abb": 2, "bbb": 2, "aa": 2, "b": 1})
    ea = sysm.eri_asym()
    vec = {}
    seen_l = set()
    for i in range(len(dets)):
        for j in range(len(dets)):
            a, b = dets[i], dets[j]
            if a.parent != 0 or b.parent != 1:
                continue
            for k, d in ((i, a), (j, b)):
                if k not in vec:
                    vec[k] = bf.vector(*sysm.refs[d.parent], _cols(d.oa), _cols(d.ob))
            hb, sb = bf.hs(vec[i], vec[j])
            h1, s1 = slab.hs(a, b, 1)
            assert abs(h1 - (hb + sysm.enuc * sb)) < 1e-10 and abs(s1 - sb) < 1e-10
            hn, sn = O.hs_naive(sysm, a, b, eri_asym=ea)
            assert abs(hn - (hb + sysm.enuc * sb)) < 1e-10
            seen_l.add(max(a.rank_a + b.rank_a, a.rank_b + b.rank_b))
            if orth_count:
                h0, s0 = slab.hs(a, b, 0)
                hr, sr = R.element(slab.pair_view(0, 1), 7, a, b, O.RD_ALL, sysm.enuc, slab.tol)
                assert abs(h0 - hr) < 1e-10 * max(1.0, abs(hr)) and abs(s0 - sr) < 1e-12
    assert max(seen_l) == 6
