"""GPU parity of the HIP engine against the CPU oracle, through the C ABI.

Tolerance: 1e-10 Ha per element (H) and 1e-10 for S (FP64 path); the zero
classification m per ordered reference pair must match exactly.
"""
import numpy as np
import pytest

import nwick
import oracle as O
from nwick import basis, synth

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _case(n, na, nb, nref, seed, orth_refs=(), orth_count=1, orthonormal=False, per_ref=30):
    sysm = synth.make_system(n, na, nb, nref, seed=seed, orth_refs=orth_refs, orth_count=orth_count,
                             orthonormal_ao=orthonormal, n_aux=2 * n)
    dets = basis.excited_basis(nref, n, na, nb, per_ref=per_ref, seed=seed, include_refs=True)
    return sysm, dets


@pytest.fixture(scope="module")
def case_m1():
    sysm, dets = _case(12, 3, 2, 3, seed=21, orth_refs=(2,), orth_count=1)
    print(f"case_m1: n=12 nocc=3+2 nref=3 ndet={len(dets)}", flush=True)
    eng = nwick.Engine.from_system(sysm)
    meta = eng.build(tol=1e-8)
    slab = O.Slab(sysm, tol=1e-8)
    return sysm, dets, eng, meta, slab


def test_zero_classification_and_reduced_overlap(case_m1):
    sysm, dets, eng, meta, slab = case_m1
    for (x, w), pm in meta.items():
        assert (pm.ma, pm.mb) == slab.m[(x, w)]
        om = slab.meta[(x, w)]
        assert abs(pm.phase_a * pm.tilde_s_a - om[1] * om[0]) < 1e-12
        assert abs(pm.phase_b * pm.tilde_s_b - om[12] * om[11]) < 1e-12
    assert meta[(0, 2)].ma == 1 and meta[(2, 0)].ma == 1 and meta[(0, 0)].ma == 0


def test_slab_intermediates_m0_pairs(case_m1):
    """X, Y, F, V, J, II, hcol0 blocks agree with the oracle's slab for m = 0 pairs
    (these are invariant to the SVD gauge)."""
    sysm, dets, eng, meta, slab = case_m1
    for (x, w), pm in meta.items():
        if (pm.ma, pm.mb) != (0, 0):
            continue
        g, off, gmeta = eng.pair_slab(x, w)
        sub, ooff, ometa = slab.pair_view(x, w)
        assert np.array_equal(off, ooff)
        size = int(off[94])
        assert np.max(np.abs(g[:size] - sub[:size])) < 1e-10
        assert np.max(np.abs(gmeta - ometa)) < 1e-10


@pytest.mark.parametrize("mode", [0, 1])
def test_hs_matrix_vs_oracle(case_m1, mode):
    sysm, dets, eng, meta, slab = case_m1
    rec = basis.to_records(dets)
    H, S = eng.hs_matrix(rec, mode=mode)
    Ho, So = O.hs_matrix(slab, dets, mode=mode)
    assert np.max(np.abs(S - So)) < TOL
    assert np.max(np.abs(H - Ho)) < TOL


def test_hs_batched_vs_oracle(case_m1):
    sysm, dets, eng, meta, slab = case_m1
    rec = basis.to_records(dets)
    rng = np.random.default_rng(4)
    req = rng.integers(0, len(dets), size=(3000, 2)).astype(np.uint64)
    for mode in (0, 1):
        h, s = eng.hs_batched(rec, req, mode=mode)
        for q in range(0, len(req), 7):
            a, b = int(req[q, 0]), int(req[q, 1])
            A, B = dets[a], dets[b]
            if A.parent == B.parent and bin(A.oa ^ B.oa).count("1") + bin(A.ob ^ B.ob).count("1") > 4:
                assert h[q] == 0.0 and s[q] == 0.0
                continue
            ho, so = slab.hs(A, B, mode)
            assert abs(h[q] - ho) < TOL and abs(s[q] - so) < TOL


@pytest.mark.parametrize("orthonormal", [True, False])
def test_m2_pairs_both_modes(orthonormal):
    sysm, dets = _case(13, 4, 3, 3, seed=31, orth_refs=(1,), orth_count=2, orthonormal=orthonormal, per_ref=18)
    print(f"case_m2: n=13 nocc=4+3 orthonormal={orthonormal} ndet={len(dets)}", flush=True)
    eng = nwick.Engine.from_system(sysm)
    meta = eng.build()
    slab = O.Slab(sysm)
    assert meta[(0, 1)].ma == 2
    rec = basis.to_records(dets)
    for mode in (0, 1):
        H, S = eng.hs_matrix(rec, mode=mode)
        Ho, So = O.hs_matrix(slab, dets, mode=mode)
        assert np.max(np.abs(S - So)) < TOL
        assert np.max(np.abs(H - Ho)) < TOL


def test_pair_mask_and_unbuilt_pair_error():
    sysm, dets = _case(10, 2, 2, 3, seed=41, per_ref=8)
    eng = nwick.Engine.from_system(sysm)
    mask = np.triu(np.ones((3, 3), dtype=np.uint8))
    eng.build(pair_mask=mask)
    rec = basis.to_records(dets)
    H, S = eng.hs_matrix(rec)  # only pairs with parent_a <= parent_b are touched
    slab = O.Slab(sysm)
    Ho, So = O.hs_matrix(slab, dets)
    assert np.max(np.abs(H - Ho)) < TOL
    with pytest.raises(nwick.NwickError):
        eng.hs_batched(rec, np.array([[len(dets) - 1, 0]], dtype=np.uint64))


def test_batched_request_errors_and_rank_limit():
    """Batched requests are binned on the GPU: an index past the determinant
    list and a determinant with an excitation rank above four on a spin (a
    quintuple) are errors rather than unwritten outputs, in the batched and
    full-matrix paths; a pair of contraction rank five on one spin (an alpha
    triple against an alpha double) is evaluated by the run-time-rank
    kernels and matches the oracle.  Valid requests after an error still
    evaluate, and match the full matrix."""
    sysm, dets = _case(12, 5, 3, 2, seed=61, per_ref=6)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    rec = basis.to_records(dets)
    with pytest.raises(nwick.NwickError):
        eng.hs_batched(rec, np.array([[0, len(dets)]], dtype=np.uint64))
    oa, ob = (1 << 5) - 1, (1 << 3) - 1
    triple = basis.Det(1, oa, ob, (0, 1, 2), (5, 6, 7))
    double = basis.Det(0, oa, ob, (0, 1), (8, 9))
    quint = basis.Det(0, oa, ob, (0, 1, 2, 3, 4), (5, 6, 7, 8, 9))
    n = len(dets)
    rec3 = basis.to_records(dets + [triple, double])
    h, s = eng.hs_batched(rec3, np.array([[n + 1, n]], dtype=np.uint64))
    slab = O.Slab(sysm)
    for mode in (0, 1):
        h, s = eng.hs_batched(rec3, np.array([[n + 1, n]], dtype=np.uint64), mode=mode)
        ho, so = slab.hs(double, triple, mode)
        assert abs(h[0] - ho) < TOL and abs(s[0] - so) < TOL
    rec5 = basis.to_records(dets + [quint])
    with pytest.raises(nwick.NwickError):
        eng.hs_batched(rec5, np.array([[n, 0]], dtype=np.uint64))
    with pytest.raises(nwick.NwickError):
        eng.hs_matrix(rec5)
    req = np.array([[a, b] for a in range(n) for b in range(n)], dtype=np.uint64)
    h, s = eng.hs_batched(rec, req)
    H, S = eng.hs_matrix(rec)
    assert np.max(np.abs(h.reshape(n, n) - H)) < 1e-10 and np.max(np.abs(s.reshape(n, n) - S)) < 1e-12


@pytest.mark.parametrize("prec", [nwick.PREC_MIXED, nwick.PREC_BF16X3])
def test_mixed_precision_build_close_to_fp64(prec):
    """Mixed build (FP32 MFMA, or BF16 MFMA on three-way split operands):
    elements within 1e-6 of FP64 and the lowest NOCI energy within 1 uHa;
    FP64 quantities (S, m) unchanged."""
    sysm, dets = _case(16, 3, 3, 3, seed=51, orth_refs=(2,), orth_count=1, orthonormal=False, per_ref=40)
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    H64, S64 = eng.hs_matrix(rec, mode=1)
    eng.set_precision(prec)
    meta = eng.build()
    assert meta[(0, 2)].ma == 1
    Hmx, Smx = eng.hs_matrix(rec, mode=1)
    assert np.max(np.abs(Smx - S64)) < 1e-14
    assert np.max(np.abs(Hmx - H64)) < 1e-6
    import scipy.linalg
    e64 = scipy.linalg.eigh(H64, S64, eigvals_only=True)[0]
    emx = scipy.linalg.eigh(Hmx, Smx, eigvals_only=True)[0]
    print(f"precision {prec}: max|dH|={np.max(np.abs(Hmx - H64)):.3e} dE={emx - e64:.3e}", flush=True)
    assert abs(emx - e64) < 1e-6


def test_mixed_floor_selects_pairs():
    """The sigma floor decides per reference pair: floor 0 puts every pair on
    FP32 transforms, a floor above 1 keeps every pair FP64 (then the mixed
    build equals the FP64 build bit for bit), and the default keeps exactly the
    pairs whose smallest nonzero singular value is below 0.1 in FP64."""
    sysm, dets = _case(16, 3, 3, 4, seed=52, orthonormal=False, per_ref=24)
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    H64, S64 = eng.hs_matrix(rec, mode=0)
    eng.set_precision(nwick.PREC_MIXED)
    eng.set_mixed_floor(2.0)
    meta = eng.build()
    assert all(m.fp32 == 0 for m in meta.values())
    Hk, Sk = eng.hs_matrix(rec, mode=0)
    assert np.array_equal(Hk, H64) and np.array_equal(Sk, S64)
    eng.set_mixed_floor(0.0)
    meta = eng.build()
    assert all(m.fp32 == 1 for m in meta.values())
    eng.set_mixed_floor(0.1)
    meta = eng.build()
    for (x, w), m in meta.items():
        sv = []
        for sp in (0, 1):
            A = sysm.refs[x][sp][:, :3].T @ sysm.S @ sysm.refs[w][sp][:, :3]
            s = np.linalg.svd(A, compute_uv=False)
            sv += [v for v in s if v > 1e-8]
        assert abs(m.sigma_min - min(sv)) < 1e-10
        assert m.fp32 == (1 if min(sv) >= 0.1 else 0)


@pytest.mark.parametrize("n", [11, 14])
def test_transform_shapes_odd_and_two_mod_four(n):
    """n_bas odd (generic / non-pipelined transform) and n_bas = 2 mod 4
    (pipelined transform with a zero-padded K tile, FP64 and the float2 FP32
    form) against the oracle."""
    sysm, dets = _case(n, 3, 2, 3, seed=90 + n, orthonormal=False, per_ref=20)
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    H, S = eng.hs_matrix(rec, mode=1)
    Ho, So = O.hs_matrix(O.Slab(sysm, tol=1e-8), dets, mode=1)
    assert np.max(np.abs(H - Ho)) < TOL and np.max(np.abs(S - So)) < TOL
    if n % 2 == 0:
        # well-conditioned references (small rotations: every singular value
        # near 1) so FP32 rounding is not amplified and the FP32 kernel's
        # indexing is what the comparison tests
        sysm = synth.make_system(n, 3, 2, 3, seed=90 + n, kappa_sigma=0.05, n_aux=2 * n)
        dets = basis.excited_basis(3, n, 3, 2, per_ref=20, seed=90 + n, include_refs=True)
        rec = basis.to_records(dets)
        eng = nwick.Engine.from_system(sysm)
        eng.build()
        H, S = eng.hs_matrix(rec, mode=1)
        eng.set_mixed_floor(0.0)
        err = {}
        for prec in (nwick.PREC_MIXED, nwick.PREC_BF16X3):
            eng.set_precision(prec)
            meta = eng.build()
            assert all(m.fp32 == 1 for m in meta.values())
            Hm, Sm = eng.hs_matrix(rec, mode=1)
            err[prec] = float(np.max(np.abs(Hm - H)))
            print(f"n={n} precision {prec}: max|dH|={err[prec]:.3e}", flush=True)
            assert np.array_equal(Sm, S)
            assert err[prec] < 1e-5 * max(1.0, np.max(np.abs(H)))
        # the split BF16 products carry FP32-level error
        assert err[nwick.PREC_BF16X3] < 8.0 * err[nwick.PREC_MIXED] + 1e-12


@pytest.mark.parametrize("n", [16, 50, 100])
def test_bf16x3_three_launch_transform(n, monkeypatch):
    """The bf16x3 transform as three launches (NW_TRANSFORM_PASSES=3: half
    transform of each (p,q) slice, two leading-index passes) against the
    default four quarter passes and against FP64: n = 16 (one slice tile),
    50 (n = 2 mod 4: float2 slice loads, ragged K and column tiles), 100
    (seven slice tiles, the n_bas = 100 shape)."""
    sysm = synth.make_system(n, 3, 2, 2, seed=300 + n, kappa_sigma=0.05, n_aux=2 * n)
    dets = basis.excited_basis(2, n, 3, 2, per_ref=10, seed=300 + n, include_refs=True)
    rec = basis.to_records(dets)
    print(f"three-launch transform: n={n} ndet={len(dets)}", flush=True)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    H64, S64 = eng.hs_matrix(rec, mode=1)
    g64, off64, _ = eng.pair_slab(0, 1)
    eng.set_mixed_floor(0.0)
    eng.set_precision(nwick.PREC_BF16X3)
    monkeypatch.delenv("NW_TRANSFORM_PASSES", raising=False)
    meta = eng.build()
    assert all(m.fp32 == 1 for m in meta.values())
    H4, S4 = eng.hs_matrix(rec, mode=1)
    # the rank-4 blocks are kept in FP32 on the device; the exported slab is
    # widened back to FP64 and agrees with the FP64 build's block
    g4, off4, _ = eng.pair_slab(0, 1)
    assert np.array_equal(off4, off64)
    size = int(off64[94])
    assert np.max(np.abs(g4[:size] - g64[:size])) < 1e-5 * max(1.0, float(np.max(np.abs(g64[:size]))))
    monkeypatch.setenv("NW_TRANSFORM_PASSES", "3")
    eng.build()
    H3, S3 = eng.hs_matrix(rec, mode=1)
    scale = max(1.0, float(np.max(np.abs(H64))))
    e4 = float(np.max(np.abs(H4 - H64)))
    e3 = float(np.max(np.abs(H3 - H64)))
    d = float(np.max(np.abs(H3 - H4)))
    print(f"n={n}: max|H4-H64|={e4:.3e} max|H3-H64|={e3:.3e} max|H3-H4|={d:.3e}", flush=True)
    assert np.array_equal(S3, S64) and np.array_equal(S4, S64)
    assert e3 < 1e-5 * scale
    assert e3 < 4.0 * e4 + 1e-12
    assert d < 4.0 * e4 + 1e-12


def test_large_basis_mfma_contractions():
    """n_bas = 36: the co-density x integral-matrix products (J - K, J) take
    the FP64 MFMA path (K = n^2 >= 256, N K >= 2^20); H, S against the oracle,
    including an m = 1 reference pair."""
    sysm, dets = _case(36, 3, 2, 2, seed=36, orth_refs=(1,), orth_count=1, per_ref=12)
    print(f"large basis: n=36 ndet={len(dets)}", flush=True)
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    meta = eng.build(tol=1e-8)
    assert meta[(0, 1)].ma == 1
    H, S = eng.hs_matrix(rec, mode=1)
    Ho, So = O.hs_matrix(O.Slab(sysm, tol=1e-8), dets, mode=1)
    print(f"large basis: max|dH|={np.max(np.abs(H - Ho)):.3e} max|dS|={np.max(np.abs(S - So)):.3e}", flush=True)
    assert np.max(np.abs(S - So)) < TOL
    assert np.max(np.abs(H - Ho)) < TOL * max(1.0, np.max(np.abs(Ho)))
