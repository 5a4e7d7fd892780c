"""Rank-4 blocks on demand from integral factors, through the C ABI.

A context given Cholesky-type factors (the synthetic systems' own L_K, with
(ab|cd) = sum_K L_K[a,b] L_K[c,d]) builds the J and II blocks of m = 0 pairs
only where each evaluation call's determinants read them.  Every result must
equal the dense-transform engine's and the oracle's: 1e-10 Ha per element
(FP64), as the dense path.  Exports complete the tensors first.
"""
import numpy as np
import pytest

import nwick
import oracle as O
from nwick import basis, synth

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _pair(sysm, prec=nwick.PREC_FP64, factors=True):
    eng = nwick.Engine.from_system(sysm)
    if prec != nwick.PREC_FP64:
        eng.set_precision(prec)
    if factors:
        dev = eng.set_eri_factors(sysm.chol)
        assert dev < 1e-12
    return eng


@pytest.fixture(scope="module")
def small():
    sysm = synth.make_system(14, 3, 2, 3, seed=77, orth_refs=(2,), orth_count=1, orthonormal_ao=False, n_aux=30)
    dets = basis.excited_basis(3, 14, 3, 2, per_ref=40, seed=77, include_refs=True)
    print(f"small: n=14 nocc=3+2 nref=3 naux=30 ndet={len(dets)}", flush=True)
    dense = nwick.Engine.from_system(sysm)
    meta = dense.build()
    assert meta[(0, 2)].ma == 1 and meta[(0, 1)].ma == 0
    return sysm, dets, dense


@pytest.mark.parametrize("mode", [0, 1])
def test_factored_equals_dense_and_oracle(small, mode):
    sysm, dets, dense = small
    eng = _pair(sysm)
    eng.build()
    rec = basis.to_records(dets)
    H, S = eng.hs_matrix(rec, mode=mode)
    Hd, Sd = dense.hs_matrix(rec, mode=mode)
    Ho, So = O.hs_matrix(O.Slab(sysm, tol=1e-8), dets, mode=mode)
    print(f"mode {mode}: factored-dense {np.max(np.abs(H - Hd)):.2e}, factored-oracle {np.max(np.abs(H - Ho)):.2e}",
          flush=True)
    assert np.max(np.abs(S - Sd)) < TOL and np.max(np.abs(H - Hd)) < TOL
    assert np.max(np.abs(S - So)) < TOL and np.max(np.abs(H - Ho)) < TOL
    # every pair (m = 0 and the m_alpha = 1 pair (0, 2), with exchange products)
    # was built on demand, and only in part
    for x, w in [(0, 1), (0, 2)]:
        ja, jb, ii = eng.rank4_coverage(x, w)
        assert 0.0 < ja < 1.0 and 0.0 < jb < 1.0 and 0.0 < ii < 1.0
    eng.close()


def test_coverage_grows_with_later_calls(small):
    """A call on a subset builds only its blocks; a later call on the full
    list adds the rest it reads and equals the dense result; a repeat builds
    nothing."""
    sysm, dets, dense = small
    eng = _pair(sysm)
    eng.build()
    sub = [d for d in dets if d.parent != 1][:30]
    Hs, Ss = eng.hs_matrix(basis.to_records(sub))
    Hsd, Ssd = dense.hs_matrix(basis.to_records(sub))
    assert np.max(np.abs(Hs - Hsd)) < TOL
    c1 = eng.rank4_coverage(0, 0)
    rec = basis.to_records(dets)
    H, S = eng.hs_matrix(rec)
    assert eng.timings()["transform_ms"] > 0.0
    c2 = eng.rank4_coverage(0, 0)
    assert all(b >= a for a, b in zip(c1, c2)) and sum(c2) > sum(c1)
    Hd, Sd = dense.hs_matrix(rec)
    assert np.max(np.abs(H - Hd)) < TOL and np.max(np.abs(S - Sd)) < TOL
    H2, S2 = eng.hs_matrix(rec)
    assert eng.timings()["transform_ms"] == 0.0
    assert np.array_equal(H2, H)
    eng.close()


def test_batched_project_and_streamed(small):
    sysm, dets, dense = small
    rec = basis.to_records(dets)
    rng = np.random.default_rng(9)
    req = rng.integers(0, len(dets), size=(4000, 2)).astype(np.uint64)
    eng = _pair(sysm)
    eng.build()
    for mode in (0, 1):
        h, s = eng.hs_batched(rec, req, mode=mode)
        hd, sd = dense.hs_batched(rec, req, mode=mode)
        assert np.max(np.abs(h - hd)) < TOL and np.max(np.abs(s - sd)) < TOL
    eng.close()
    eng = _pair(sysm)
    eng.build()
    ncur = 25
    coef = rng.normal(size=ncur)
    v, s = eng.project_hs(rec, ncur, coef)
    vd, sd = dense.project_hs(rec, ncur, coef)
    assert np.max(np.abs(v - vd)) < TOL and np.max(np.abs(s - sd)) < TOL
    eng.close()
    eng = _pair(sysm)
    eng.build()
    eng.hs_matrix(rec)
    whole = eng.rank4_stats()
    # the block products by launch class partition the rectangle, single and
    # exchange stages, in time and in flops
    ks = eng.rank4_kernel_stats()
    kms = ks["ms_f64_large"] + ks["ms_f64_narrow"] + ks["ms_f32_large"] + ks["ms_f32_narrow"]
    kfl = ks["flops_f64_large"] + ks["flops_f64_narrow"] + ks["flops_f32_large"] + ks["flops_f32_narrow"]
    assert kms == pytest.approx(whole["ms_rect"] + whole["ms_single"] + whole["ms_exchange"], rel=1e-9)
    assert kfl == pytest.approx(whole["flops_rect"] + whole["flops_single"] + whole["flops_exchange"], rel=1e-12)
    assert ks["flops_f64_large"] > 0.0 and ks["flops_f32_large"] == 0.0  # an FP64 context
    eng.close()
    eng = _pair(sysm)
    H, S, ng = eng.hs_matrix_streamed(rec, slab_bytes=5 << 19)
    assert ng > 1
    Hd, Sd = dense.hs_matrix(rec)
    assert np.max(np.abs(H - Hd)) < TOL and np.max(np.abs(S - Sd)) < TOL
    # the stage records of a streamed call add up over its groups: the same
    # blocks as one call over every pair
    st = eng.rank4_stats()
    for k in ("flops_rect", "flops_single", "flops_exchange", "flops_factor", "single_blocks"):
        assert st[k] == pytest.approx(whole[k], rel=1e-12), k
    assert st["ms_total"] > 0.0
    eng.close()


def test_export_completes_every_block(small):
    sysm, dets, dense = small
    eng = _pair(sysm)
    eng.build()
    eng.hs_matrix(basis.to_records(dets[:20]))
    assert eng.rank4_coverage(0, 1)[0] < 1.0
    for x, w in [(0, 1), (1, 1), (0, 2)]:
        g, off, meta = eng.pair_slab(x, w)
        gd, offd, metad = dense.pair_slab(x, w)
        assert np.array_equal(off, offd)
        assert np.max(np.abs(g - gd)) < 1e-11
    assert eng.rank4_coverage(0, 1) == (1.0, 1.0, 1.0)
    eng.close()


@pytest.mark.parametrize("prec", [nwick.PREC_MIXED, nwick.PREC_BF16X3])
def test_mixed_precision_factored(small, prec):
    """FP32 products (NW_PREC_MIXED: FP32 MFMA; NW_PREC_BF16X3: BF16 MFMA on
    three-way split operands) of FP32-rounded factors for pairs above the
    sigma floor: the error against FP64 is of the size of the dense reduced-
    precision transforms' of the same precision, at the default floor and with
    every pair forced to FP32 (floor 0)."""
    sysm, dets, dense = small
    rec = basis.to_records(dets)
    Hd, Sd = dense.hs_matrix(rec)
    scale = np.max(np.abs(Hd))
    for floor in (0.1, 0.0):
        errs = {}
        for factors in (False, True):
            eng = _pair(sysm, prec=prec, factors=factors)
            eng.set_mixed_floor(floor)
            meta = eng.build()
            n32 = sum(m.fp32 for m in meta.values())
            H, S = eng.hs_matrix(rec)
            assert np.max(np.abs(S - Sd)) < TOL
            errs[factors] = np.max(np.abs(H - Hd)) / scale
            eng.close()
        print(f"prec {prec} floor {floor}: {n32} FP32 pairs; max rel |dH| dense {errs[False]:.2e}, "
              f"factored {errs[True]:.2e}", flush=True)
        assert n32 > 0
        assert errs[True] < max(10 * errs[False], 1e-9)


def test_factors_that_do_not_match_are_refused(small):
    sysm, dets, dense = small
    eng = nwick.Engine.from_system(sysm)
    bad = sysm.chol.copy()
    bad[3, 2, 5] += 1e-3
    bad[3, 5, 2] += 1e-3
    with pytest.raises(nwick.NwickError) as e:
        eng.set_eri_factors(bad)
    assert e.value.code == 1
    asym = sysm.chol.copy()
    asym[0, 1, 2] += 1e-6
    with pytest.raises(nwick.NwickError):
        eng.set_eri_factors(asym)
    assert eng.set_eri_factors(None) == 0.0
    eng.close()


def test_noalloc_after_blocks_exist(small):
    torch = pytest.importorskip("torch")
    sysm, dets, dense = small
    eng = _pair(sysm)
    eng.build()
    rec = basis.to_records(dets)
    req = np.random.default_rng(5).integers(0, len(dets), size=(2000, 2)).astype(np.uint64)
    rbytes = torch.from_numpy(rec.view(np.uint8).copy()).cuda()
    dreq = torch.from_numpy(req.view(np.int64).copy()).cuda()
    h = torch.empty(len(req), dtype=torch.float64, device="cuda")
    s = torch.empty_like(h)
    with pytest.raises(nwick.NwickError) as e:
        eng.hs_batched_device(rbytes, dreq, h, s, noalloc=True)
    assert e.value.code == 6
    eng.hs_batched(rec, req)
    eng.hs_batched_device(rbytes, dreq, h, s, noalloc=True)
    hd, sd = dense.hs_batched(rec, req)
    assert np.max(np.abs(h.cpu().numpy() - hd)) < TOL and np.max(np.abs(s.cpu().numpy() - sd)) < TOL
    eng.close()


@pytest.fixture(scope="module")
def n100():
    sysm = synth.make_system(100, 10, 10, 2, seed=2024, n_aux=300, orth_refs=(1,), orth_count=1)
    counts = {"a": 4, "b": 4, "aa": 10, "bb": 10, "ab": 11}
    dets = basis.excited_basis(2, 100, 10, 10, per_ref=39, seed=2024, include_refs=True, counts=counts)
    print(f"n_bas=100: 10+10 e, 2 refs, naux=300, ndet={len(dets)}", flush=True)
    return sysm, dets


@pytest.mark.parametrize("prec", [nwick.PREC_FP64, nwick.PREC_MIXED, nwick.PREC_BF16X3])
def test_n100_factored_vs_dense(n100, prec):
    """The headline shape (n_bas = 100, 10+10 electrons, 300 factors): FP64
    factored = dense to 1e-10; FP32 factored products (FP32 MFMA, or BF16 MFMA
    on split operands) against the FP64 dense build as the FP32 transforms are
    (relative 1e-6)."""
    sysm, dets = n100
    rec = basis.to_records(dets)
    dense = nwick.Engine.from_system(sysm)
    dense.build()
    Hd, Sd = dense.hs_matrix(rec)
    dense.close()
    eng = _pair(sysm, prec=prec)
    eng.build()
    H, S = eng.hs_matrix(rec)
    dh = float(np.max(np.abs(H - Hd)))
    print(f"n_bas=100 prec {prec}: max|dH| = {dh:.3e}, blocks {eng.timings()['transform_ms']:.2f} ms, "
          f"coverage (0,0) {eng.rank4_coverage(0, 0)}", flush=True)
    assert np.max(np.abs(S - Sd)) < TOL
    assert dh < (TOL if prec == nwick.PREC_FP64 else 1e-6 * np.max(np.abs(Hd)))
    eng.close()
